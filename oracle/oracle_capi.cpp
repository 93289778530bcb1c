// ORACLE — test infrastructure only.  C entry points over the FP64 restatement
// so that the pytest suite (ctypes) can drive it with the same inputs it gives
// the HIP library.
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

#include "oracle.hpp"

using namespace oracle;

namespace {
thread_local std::string g_err;

template <class F> int guard(F&& f) {
    try {
        f();
        return 0;
    } catch (std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

Change make_change(int everything, int ngroups, const int* group_index, const int* all, const int* internal,
                   const int* offsets, const int* indices) {
    Change c;
    c.everything = everything != 0;
    for (int g = 0; g < ngroups; ++g) {
        GroupChange gc;
        gc.group_index = group_index[g];
        gc.all = all[g] != 0;
        gc.internal = internal[g] != 0;
        for (int k = offsets[g]; k < offsets[g + 1]; ++k) gc.relative_atom_indices.push_back(indices[k]);
        c.groups.push_back(gc);
    }
    return c;
}

// Apply new coordinates for the listed absolute particle indices to the trial space.
void stage(System& s, int m, const int* pidx, const double* xyz) {
    for (int k = 0; k < m; ++k) {
        Particle& p = s.trial_space.particles.at(static_cast<size_t>(pidx[k]));
        p.pos = {xyz[3 * k], xyz[3 * k + 1], xyz[3 * k + 2]};
    }
}
} // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

void* orc_create(const char* json) {
    try {
        return new System(json);
    } catch (std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void orc_destroy(void* h) { delete static_cast<System*>(h); }

int orc_set_particles(void* h, int n, const double* xyz, const double* q, const int* id) {
    return guard([&] { static_cast<System*>(h)->set_particles(n, xyz, q, id); });
}

int orc_set_groups(void* h, int ng, const int* begin, const int* size, const int* molid, const double* cm) {
    return guard([&] { static_cast<System*>(h)->set_groups(ng, begin, size, molid, cm); });
}

int orc_init(void* h, double* initial_energy) {
    return guard([&] {
        auto* s = static_cast<System*>(h);
        s->init();
        if (initial_energy) *initial_energy = s->initial_energy;
    });
}

int orc_num_terms(void* h) { return static_cast<System*>(h)->num_terms(); }

int orc_term_name(void* h, int i, char* buf, int len) {
    return guard([&] {
        auto names = static_cast<System*>(h)->term_names();
        std::strncpy(buf, names.at(static_cast<size_t>(i)).c_str(), static_cast<size_t>(len - 1));
        buf[len - 1] = 0;
    });
}

int orc_total_energy(void* h, int trial, double* terms, double* total) {
    return guard([&] {
        std::vector<double> t;
        const double e = static_cast<System*>(h)->total_energy(trial != 0, &t);
        for (size_t i = 0; i < t.size(); ++i) terms[i] = t[i];
        if (total) *total = e;
    });
}

// Energies of the affected pairs of a Change in the trial state (new
// coordinates given for the moved particles) and in the accepted state, per term.
// Leaves the trial state as it was (the staged coordinates are reverted).
int orc_energy_change(void* h, int everything, int ngroups, const int* gi, const int* all, const int* internal,
                      const int* off, const int* idx, int m, const int* pidx, const double* xyz, double* e_new,
                      double* e_old) {
    return guard([&] {
        auto* s = static_cast<System*>(h);
        Change c = make_change(everything, ngroups, gi, all, internal, off, idx);
        stage(*s, m, pidx, xyz);
        std::vector<double> en, eo;
        s->energy_both(c, en, eo);
        for (size_t i = 0; i < en.size(); ++i) {
            e_new[i] = en[i];
            e_old[i] = eo[i];
        }
        // revert trial state (as a rejected move would)
        for (int k = 0; k < m; ++k) s->trial_space.particles[static_cast<size_t>(pidx[k])] =
                                        s->accepted_space.particles[static_cast<size_t>(pidx[k])];
        s->trial_pot->sync_from(*s->accepted_pot, c);
    });
}

// Metropolis driver for single-particle translations (montecarlo.cpp:139-187,
// move.cpp:225-293, core.cpp:249-261) with its own std::mt19937 proposal source.
// Records each proposal so that another implementation can replay it.
int orc_run_translate(void* h, long nmoves, unsigned seed, int molid, double default_dp, int* rec_group,
                      int* rec_index, double* rec_xyz, double* rec_u, double* rec_du, unsigned char* rec_acc,
                      double* rec_enew, double* rec_eold) {
    return guard([&] {
        auto* s = static_cast<System*>(h);
        std::mt19937 engine(seed);
        std::uniform_real_distribution<double> unif(0.0, 1.0);
        auto rnd = [&] { return unif(engine); };
        std::vector<int> candidates;
        for (size_t g = 0; g < s->accepted_space.groups.size(); ++g)
            if (s->accepted_space.groups[g].molid == molid) candidates.push_back(static_cast<int>(g));
        if (candidates.empty()) throw std::runtime_error("no groups of the requested molecule");
        const int nterm = s->num_terms();
        std::vector<double> en, eo;
        for (long n = 0; n < nmoves; ++n) {
            const int gidx = candidates[std::uniform_int_distribution<size_t>(0, candidates.size() - 1)(engine)];
            Group& g = s->trial_space.groups[static_cast<size_t>(gidx)];
            if (g.size == 0) throw std::runtime_error("empty group");
            const int rel = static_cast<int>(std::uniform_int_distribution<int>(0, g.size - 1)(engine));
            Particle& p = s->trial_space.particles[static_cast<size_t>(g.begin + rel)];
            const AtomType& at = s->atoms.at(static_cast<size_t>(p.id));
            const double dp = at.has_dp ? at.dp : default_dp;
            Vec3 dir;
            double r2;
            do {
                dir = {rnd() - 0.5, rnd() - 0.5, rnd() - 0.5};
                r2 = dir.squared_norm();
            } while (r2 > 0.25);
            dir = dir * (1.0 / std::sqrt(r2));
            p.pos = p.pos + dir * dp * rnd();
            s->trial_space.geometry.boundary(p.pos);
            const double u = rnd();
            Change c;
            GroupChange gc;
            gc.group_index = gidx;
            gc.internal = true;
            gc.relative_atom_indices = {rel};
            c.groups.push_back(gc);
            bool acc = false;
            double du = 0;
            if (rec_xyz) {
                rec_xyz[3 * n] = p.pos.x;
                rec_xyz[3 * n + 1] = p.pos.y;
                rec_xyz[3 * n + 2] = p.pos.z;
            }
            s->metropolis_move(c, u, &acc, &du, &en, &eo);
            if (rec_group) rec_group[n] = gidx;
            if (rec_index) rec_index[n] = rel;
            if (rec_u) rec_u[n] = u;
            if (rec_du) rec_du[n] = du;
            if (rec_acc) rec_acc[n] = acc ? 1 : 0;
            for (int t = 0; t < nterm; ++t) {
                if (rec_enew) rec_enew[n * nterm + t] = en[static_cast<size_t>(t)];
                if (rec_eold) rec_eold[n * nterm + t] = eo[static_cast<size_t>(t)];
            }
        }
    });
}

// Metropolis run with the device-resident proposal scheme: move t of replica r
// draws Philox4x32-10 blocks at counters (t_lo, t_hi, r, 0) and (t_lo, t_hi, r, 1)
// with key (seed_lo, seed_hi): particle = (x0 * n) >> 32 over all particles, unit
// vector from z = 2u-1 and phi = 2 pi u', length dp*U[0,1), a 53-bit acceptance
// deviate; acceptance u <= exp(-beta du).  Records per move: particle, proposed
// position, du (kT at the configured temperature) and the decision.
int orc_run_philox_translate(void* h, long nmoves, unsigned long long seed, unsigned replica,
                             unsigned long long move0, double beta, int* rec_index, double* rec_xyz, double* rec_du,
                             unsigned char* rec_acc) {
    return guard([&] {
        auto* s = static_cast<System*>(h);
        const int n = static_cast<int>(s->trial_space.particles.size());
        std::vector<int> group_of(static_cast<size_t>(n)), rel_of(static_cast<size_t>(n));
        for (size_t g = 0; g < s->trial_space.groups.size(); ++g) {
            const Group& G = s->trial_space.groups[g];
            for (int k = 0; k < G.size; ++k) {
                group_of[static_cast<size_t>(G.begin + k)] = static_cast<int>(g);
                rel_of[static_cast<size_t>(G.begin + k)] = k;
            }
        }
        auto u01 = [](uint32_t x) { return (static_cast<double>(x) + 0.5) * 2.3283064365386963e-10; };
        for (long t = 0; t < nmoves; ++t) {
            const unsigned long long mv = move0 + static_cast<unsigned long long>(t);
            const auto a = philox4x32_10({static_cast<uint32_t>(mv), static_cast<uint32_t>(mv >> 32), replica, 0u},
                                         {static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32)});
            const auto b = philox4x32_10({static_cast<uint32_t>(mv), static_cast<uint32_t>(mv >> 32), replica, 1u},
                                         {static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32)});
            const int i = static_cast<int>((static_cast<uint64_t>(a[0]) * static_cast<uint64_t>(n)) >> 32);
            const double z = 2.0 * u01(a[1]) - 1.0;
            const double phi = 6.283185307179586 * u01(a[2]);
            const double sz = std::sqrt(1.0 - z * z);
            const Vec3 dir{sz * std::cos(phi), sz * std::sin(phi), z};
            const double len_u = static_cast<double>(a[3]) * 2.3283064365386963e-10;
            const double acc_u = (static_cast<double>(b[0] >> 5) * 67108864.0 + static_cast<double>(b[1] >> 6)) *
                                 1.1102230246251565e-16;
            Particle& p = s->trial_space.particles[static_cast<size_t>(i)];
            const double dp = s->atoms.at(static_cast<size_t>(p.id)).dp;
            const double scale = dp * len_u;
            p.pos = {p.pos.x + dir.x * scale, p.pos.y + dir.y * scale, p.pos.z + dir.z * scale};
            s->trial_space.geometry.boundary(p.pos);
            if (rec_xyz) {
                rec_xyz[3 * t] = p.pos.x;
                rec_xyz[3 * t + 1] = p.pos.y;
                rec_xyz[3 * t + 2] = p.pos.z;
            }
            Change c;
            GroupChange gc;
            gc.group_index = group_of[static_cast<size_t>(i)];
            gc.internal = true;
            gc.relative_atom_indices = {rel_of[static_cast<size_t>(i)]};
            c.groups.push_back(gc);
            // performMove with beta-scaled acceptance
            s->trial_pot->update_state(s->trial_space, &s->accepted_space, c);
            const double e_new = s->trial_pot->energy(s->trial_space, &s->accepted_space, c);
            const double e_old = s->accepted_pot->energy(s->accepted_space, nullptr, c);
            double du;
            if (std::isnan(e_old) && !std::isnan(e_new)) du = -infty;
            else if (std::isnan(e_new)) du = infty;
            else if (e_new > 0.0 && std::isinf(e_new)) du = infty;
            else {
                du = e_new - e_old;
                if (std::isnan(du)) du = 0.0;
            }
            const double bdu = beta * du;
            bool acc;
            if (std::isinf(bdu) && bdu < 0) acc = true;
            else if (-bdu > max_exp_argument) acc = true;
            else acc = acc_u <= std::exp(-bdu);
            if (acc) {
                s->accepted_space.particles[static_cast<size_t>(i)] = p;
                s->accepted_pot->sync_from(*s->trial_pot, c);
                s->stats.accepted++;
                s->stats.sum_du += du;
            } else {
                p = s->accepted_space.particles[static_cast<size_t>(i)];
                s->trial_pot->sync_from(*s->accepted_pot, c);
            }
            s->stats.moves++;
            if (rec_index) rec_index[t] = i;
            if (rec_du) rec_du[t] = du;
            if (rec_acc) rec_acc[t] = acc ? 1 : 0;
        }
    });
}

// performMove (montecarlo.cpp:139-187) on a Change whose new coordinates are in the
// trial space, with acceptance u <= exp(-beta du); syncs the changed groups.
static void philox_metropolis(System& s, const Change& c, double acc_u, double beta, double* du_out, bool* acc_out) {
    s.trial_pot->update_state(s.trial_space, &s.accepted_space, c);
    const double e_new = s.trial_pot->energy(s.trial_space, &s.accepted_space, c);
    const double e_old = s.accepted_pot->energy(s.accepted_space, nullptr, c);
    double du;
    if (std::isnan(e_old) && !std::isnan(e_new)) du = -infty;
    else if (std::isnan(e_new)) du = infty;
    else if (e_new > 0.0 && std::isinf(e_new)) du = infty;
    else {
        du = e_new - e_old;
        if (std::isnan(du)) du = 0.0;
    }
    const double bdu = beta * du;
    bool acc;
    if (std::isinf(bdu) && bdu < 0) acc = true;
    else if (-bdu > max_exp_argument) acc = true;
    else acc = acc_u <= std::exp(-bdu);
    auto sync_space = [&](Space& dst, const Space& src) { // space.cpp:199-239
        for (auto& gc : c.groups) {
            const Group& gs = src.groups[static_cast<size_t>(gc.group_index)];
            Group& gd = dst.groups[static_cast<size_t>(gc.group_index)];
            if (gc.all) {
                for (int i = 0; i < gs.size; ++i) dst.particles[gs.begin + i] = src.particles[gs.begin + i];
            } else {
                for (int i : gc.relative_atom_indices) dst.particles[gs.begin + i] = src.particles[gs.begin + i];
            }
            gd.mass_center = gs.mass_center;
        }
    };
    if (acc) {
        sync_space(s.accepted_space, s.trial_space);
        s.accepted_pot->sync_from(*s.trial_pot, c);
        s.stats.accepted++;
        s.stats.sum_du += du;
    } else {
        sync_space(s.trial_space, s.accepted_space);
        s.trial_pot->sync_from(*s.accepted_pot, c);
    }
    s.stats.moves++;
    *du_out = du;
    *acc_out = acc;
}

// Metropolis run over a mix of move kinds with the device-resident proposal scheme.
// Move t of replica r draws three Philox4x32-10 blocks a, b, c at counters
// (t_lo, t_hi, r, 0|1|2) with key (seed_lo, seed_hi):
//   kind        = first k with (b2 * sum(weight)) >> 32 < cumulative weight of k
//   acceptance  = 53-bit deviate from b0, b1; u <= exp(-beta du)
//   kind 0, translate one atom of a molecule type (AtomicTranslateRotate,
//     move.cpp:225-345): group among the groups of that type (c3, only if several),
//     atom (a0 * size) >> 32, unit vector from z = 2u(a1)-1 and phi = 2 pi u(a2),
//     length dp * a3/2^32 (the atom's dp if it has one, else the move's dp); a
//     molecular group's mass centre is recomputed with PBC around the old one.
//   kind 1, translate and rotate a molecule (TranslateRotate, move.cpp:1618-1689):
//     group (a0 * count) >> 32; displacement as in kind 0 with the move's dp, then a
//     rotation by dprot * (u(c2) - 1/2) about the unit vector from c0, c1 through the
//     mass centre (Geometry::rotate, geometry.h:613-626); Change all=true.
//   kind 2, pivot (PivotMove, chainmove.cpp:216-268 and 76-104): chain (a0 * count)
//     >> 32, harmonic bond (c3 * nbonds) >> 32, the atoms after the bond's second
//     atom if u(b3) > 1/2 (origin = second atom) else the atoms before its first atom
//     (origin = first atom); rotation by dprot * (u(c2) - 1/2) about the bond axis in
//     the frame translated by the origin; mass centre recomputed; internal=true.
//     (The reference also translates the unmoved chain atoms back and forth, which
//     changes them at rounding level; the restatement leaves them untouched.)
// A move that changes nothing (empty Change) is counted as accepted with du = 0.
int orc_run_philox_moves(void* h, long nmoves, unsigned long long seed, unsigned replica, unsigned long long move0,
                         double beta, int nkinds, const int* kind, const int* molid, const unsigned* weight,
                         const double* dp, const double* dprot, int* rec_kind, int* rec_group, long* rec_off,
                         long rec_cap, int* rec_index, double* rec_xyz, double* rec_cm, double* rec_du,
                         unsigned char* rec_acc) {
    return guard([&] {
        auto* s = static_cast<System*>(h);
        if (nkinds < 1) throw std::runtime_error("no move kinds");
        std::vector<unsigned long long> cum(static_cast<size_t>(nkinds));
        unsigned long long total = 0;
        std::vector<std::vector<int>> groups_of(static_cast<size_t>(nkinds));
        for (int k = 0; k < nkinds; ++k) {
            total += weight[k];
            cum[static_cast<size_t>(k)] = total;
            for (size_t g = 0; g < s->accepted_space.groups.size(); ++g)
                if (s->accepted_space.groups[g].molid == molid[k]) groups_of[static_cast<size_t>(k)].push_back(static_cast<int>(g));
            if (groups_of[static_cast<size_t>(k)].empty()) throw std::runtime_error("move kind without groups");
            if (kind[k] < 0 || kind[k] > 2) throw std::runtime_error("unknown move kind");
        }
        if (total == 0 || total > 0xffffffffull) throw std::runtime_error("move weights must sum to 1..2^32-1");
        auto u01 = [](uint32_t x) { return (static_cast<double>(x) + 0.5) * 2.3283064365386963e-10; };
        auto pick = [](uint32_t x, size_t n) { return static_cast<size_t>((static_cast<uint64_t>(x) * n) >> 32); };
        auto unit = [&](uint32_t zw, uint32_t pw) {
            const double z = 2.0 * u01(zw) - 1.0;
            const double phi = 6.283185307179586 * u01(pw);
            const double sz = std::sqrt(1.0 - z * z);
            return Vec3{sz * std::cos(phi), sz * std::sin(phi), z};
        };
        Space& T = s->trial_space;
        const Cuboid& geo = T.geometry;
        long off = 0;
        if (rec_off) rec_off[0] = 0;
        for (long t = 0; t < nmoves; ++t) {
            const unsigned long long mv = move0 + static_cast<unsigned long long>(t);
            const std::array<uint32_t, 2> key{static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32)};
            const auto a = philox4x32_10({static_cast<uint32_t>(mv), static_cast<uint32_t>(mv >> 32), replica, 0u}, key);
            const auto b = philox4x32_10({static_cast<uint32_t>(mv), static_cast<uint32_t>(mv >> 32), replica, 1u}, key);
            const auto c = philox4x32_10({static_cast<uint32_t>(mv), static_cast<uint32_t>(mv >> 32), replica, 2u}, key);
            const uint64_t sel = (static_cast<uint64_t>(b[2]) * total) >> 32;
            int k = 0;
            while (sel >= cum[static_cast<size_t>(k)]) ++k;
            const double acc_u = (static_cast<double>(b[0] >> 5) * 67108864.0 + static_cast<double>(b[1] >> 6)) *
                                 1.1102230246251565e-16;
            const auto& G = groups_of[static_cast<size_t>(k)];
            Change ch;
            GroupChange gc;
            int gidx = -1;
            std::vector<int> moved; // absolute indices, ascending
            if (kind[k] == 0) {
                gidx = G.size() == 1 ? G[0] : G[pick(c[3], G.size())];
                Group& g = T.groups[static_cast<size_t>(gidx)];
                if (g.size > 0) {
                    const int rel = static_cast<int>(pick(a[0], static_cast<size_t>(g.size)));
                    Particle& p = T.particles[static_cast<size_t>(g.begin + rel)];
                    const AtomType& at = s->atoms.at(static_cast<size_t>(p.id));
                    const double step = at.has_dp ? at.dp : dp[k];
                    if (step > 0.0) {
                        const Vec3 dir = unit(a[1], a[2]);
                        const double scale = step * (static_cast<double>(a[3]) * 2.3283064365386963e-10);
                        p.pos = {p.pos.x + dir.x * scale, p.pos.y + dir.y * scale, p.pos.z + dir.z * scale};
                        geo.boundary(p.pos);
                        if (!g.atomic) g.mass_center = mass_center(T, g, s->atoms, true, g.mass_center * -1.0);
                        gc.group_index = gidx;
                        gc.internal = true;
                        gc.relative_atom_indices = {rel};
                        ch.groups.push_back(gc);
                        moved.push_back(g.begin + rel);
                    }
                }
            } else if (kind[k] == 1) {
                gidx = G[pick(a[0], G.size())];
                Group& g = T.groups[static_cast<size_t>(gidx)];
                bool changed = false;
                if (dp[k] > 0.0 && g.size > 0) { // Group::translate (group.cpp:115-123)
                    const Vec3 dir = unit(a[1], a[2]);
                    const double scale = dp[k] * (static_cast<double>(a[3]) * 2.3283064365386963e-10);
                    const Vec3 d{dir.x * scale, dir.y * scale, dir.z * scale};
                    g.mass_center = g.mass_center + d;
                    geo.boundary(g.mass_center);
                    for (int i = 0; i < g.size; ++i) {
                        Particle& p = T.particles[static_cast<size_t>(g.begin + i)];
                        p.pos = p.pos + d;
                        geo.boundary(p.pos);
                    }
                    changed = true;
                }
                if (dprot[k] > std::numeric_limits<double>::epsilon() && g.size > 0) {
                    const Vec3 axis = unit(c[0], c[1]);
                    const double angle = dprot[k] * (u01(c[2]) - 0.5);
                    const auto R = rotation_matrix(axis, angle);
                    const Vec3 shift = g.mass_center * -1.0;
                    for (int i = 0; i < g.size; ++i) {
                        Particle& p = T.particles[static_cast<size_t>(g.begin + i)];
                        p.pos = p.pos + shift;
                        geo.boundary(p.pos);
                        p.pos = rotate(R, p.pos) - shift;
                        geo.boundary(p.pos);
                    }
                    changed = true;
                }
                if (changed) {
                    gc.group_index = gidx;
                    gc.all = true;
                    gc.internal = false;
                    ch.groups.push_back(gc);
                    for (int i = 0; i < g.size; ++i) moved.push_back(g.begin + i);
                }
            } else {
                gidx = G[pick(a[0], G.size())];
                Group& g = T.groups[static_cast<size_t>(gidx)];
                const auto& bonds = s->molecules.at(static_cast<size_t>(g.molid)).bonds;
                if (g.size > 2 && !bonds.empty() && std::fabs(dprot[k]) > 1e-9) {
                    const HarmonicBond& bd = bonds[pick(c[3], bonds.size())];
                    int atom0 = g.begin + bd.i, atom1 = g.begin + bd.j;
                    std::vector<int> segment;
                    if (u01(b[3]) > 0.5) {
                        for (int i = atom1 + 1; i < g.begin + g.size; ++i) segment.push_back(i);
                        std::swap(atom0, atom1);
                    } else {
                        for (int i = g.begin; i < atom0; ++i) segment.push_back(i);
                    }
                    if (!segment.empty()) {
                        const double angle = dprot[k] * (u01(c[2]) - 0.5);
                        const Vec3 shift = T.particles[static_cast<size_t>(atom0)].pos;
                        auto shifted = [&](int i) {
                            Vec3 v = T.particles[static_cast<size_t>(i)].pos - shift;
                            geo.boundary(v);
                            return v;
                        };
                        const Vec3 origin = shifted(atom0);
                        const Vec3 v = geo.vdist(origin, shifted(atom1));
                        const double norm = std::sqrt(v.squared_norm());
                        const auto R = rotation_matrix(Vec3{v.x / norm, v.y / norm, v.z / norm}, angle);
                        std::vector<Vec3> frame(static_cast<size_t>(g.size));
                        for (int i = 0; i < g.size; ++i) frame[static_cast<size_t>(i)] = shifted(g.begin + i);
                        for (int i : segment)
                            frame[static_cast<size_t>(i - g.begin)] = rotate(R, frame[static_cast<size_t>(i - g.begin)] - origin) + origin;
                        double wsum = 0.0;
                        Vec3 cm{0, 0, 0};
                        for (int i = 0; i < g.size; ++i) {
                            const double w = s->atoms.at(static_cast<size_t>(T.particles[static_cast<size_t>(g.begin + i)].id)).mw;
                            cm = cm + frame[static_cast<size_t>(i)] * w;
                            wsum += w;
                        }
                        cm = Vec3{cm.x / wsum, cm.y / wsum, cm.z / wsum} + shift;
                        geo.boundary(cm);
                        g.mass_center = cm;
                        for (int i : segment) {
                            Vec3 p = frame[static_cast<size_t>(i - g.begin)] + shift;
                            geo.boundary(p);
                            T.particles[static_cast<size_t>(i)].pos = p;
                        }
                        gc.group_index = gidx;
                        gc.all = false;
                        gc.internal = true;
                        for (int i : segment) gc.relative_atom_indices.push_back(i - g.begin);
                        ch.groups.push_back(gc);
                        moved = segment;
                    }
                }
            }
            // record the proposal (moved atoms' trial positions, trial mass centre)
            if (rec_off) {
                if (off + static_cast<long>(moved.size()) > rec_cap) throw std::runtime_error("record capacity exceeded");
                for (size_t q = 0; q < moved.size(); ++q) {
                    const Vec3& p = T.particles[static_cast<size_t>(moved[q])].pos;
                    if (rec_index) rec_index[off + static_cast<long>(q)] = moved[q];
                    if (rec_xyz) {
                        rec_xyz[3 * (off + static_cast<long>(q))] = p.x;
                        rec_xyz[3 * (off + static_cast<long>(q)) + 1] = p.y;
                        rec_xyz[3 * (off + static_cast<long>(q)) + 2] = p.z;
                    }
                }
                off += static_cast<long>(moved.size());
                rec_off[t + 1] = off;
            }
            if (rec_cm) {
                const Vec3 cm = gidx >= 0 ? T.groups[static_cast<size_t>(gidx)].mass_center : Vec3{0, 0, 0};
                rec_cm[3 * t] = cm.x;
                rec_cm[3 * t + 1] = cm.y;
                rec_cm[3 * t + 2] = cm.z;
            }
            double du = 0.0;
            bool acc = true;
            if (!ch.groups.empty()) philox_metropolis(*s, ch, acc_u, beta, &du, &acc);
            else {
                s->stats.moves++;
                s->stats.accepted++;
            }
            if (rec_kind) rec_kind[t] = k;
            if (rec_group) rec_group[t] = gidx;
            if (rec_du) rec_du[t] = du;
            if (rec_acc) rec_acc[t] = acc ? 1 : 0;
        }
    });
}

// Replay a recorded single-particle proposal stream.
int orc_replay_translate(void* h, long nmoves, const int* group, const int* index, const double* xyz,
                         const double* u, double* out_du, unsigned char* out_acc, double* out_enew,
                         double* out_eold) {
    return guard([&] {
        auto* s = static_cast<System*>(h);
        const int nterm = s->num_terms();
        std::vector<double> en, eo;
        for (long n = 0; n < nmoves; ++n) {
            Group& g = s->trial_space.groups.at(static_cast<size_t>(group[n]));
            Particle& p = s->trial_space.particles.at(static_cast<size_t>(g.begin + index[n]));
            p.pos = {xyz[3 * n], xyz[3 * n + 1], xyz[3 * n + 2]};
            Change c;
            GroupChange gc;
            gc.group_index = group[n];
            gc.internal = true;
            gc.relative_atom_indices = {index[n]};
            c.groups.push_back(gc);
            bool acc = false;
            double du = 0;
            s->metropolis_move(c, u[n], &acc, &du, &en, &eo);
            if (out_du) out_du[n] = du;
            if (out_acc) out_acc[n] = acc ? 1 : 0;
            for (int t = 0; t < nterm; ++t) {
                if (out_enew) out_enew[n * nterm + t] = en[static_cast<size_t>(t)];
                if (out_eold) out_eold[n * nterm + t] = eo[static_cast<size_t>(t)];
            }
        }
    });
}

// Exclusion query on a parsed molecule type (MoleculeData::isPairExcluded).
int orc_molecule_excluded(void* h, int molid, int i, int j, int* excluded) {
    return guard([&] {
        auto* s = static_cast<System*>(h);
        *excluded = s->molecules.at(static_cast<size_t>(molid)).is_pair_excluded(i, j) ? 1 : 0;
    });
}

int orc_group_mass_center(void* h, int trial, int group, double* cm) {
    return guard([&] {
        auto* s = static_cast<System*>(h);
        const Space& sp = trial ? s->trial_space : s->accepted_space;
        const Vec3& c = sp.groups.at(static_cast<size_t>(group)).mass_center;
        cm[0] = c.x;
        cm[1] = c.y;
        cm[2] = c.z;
    });
}

int orc_get_positions(void* h, int trial, double* xyz) {
    return guard([&] {
        auto* s = static_cast<System*>(h);
        const Space& sp = trial ? s->trial_space : s->accepted_space;
        for (size_t i = 0; i < sp.particles.size(); ++i) {
            xyz[3 * i] = sp.particles[i].pos.x;
            xyz[3 * i + 1] = sp.particles[i].pos.y;
            xyz[3 * i + 2] = sp.particles[i].pos.z;
        }
    });
}

int orc_mc_stats(void* h, long* moves, long* accepted, double* sum_du) {
    return guard([&] {
        auto* s = static_cast<System*>(h);
        *moves = s->stats.moves;
        *accepted = s->stats.accepted;
        *sum_du = s->stats.sum_du;
    });
}

// ---- known-answer-test helpers

// Ewald policy pieces on an explicit particle set (energy.cpp:249-317).
int orc_ewald_parts(const char* coulomb_json, double temperature, double box, int n, const double* xyz,
                    const double* q, int* num_k, double* self, double* surface, double* reciprocal) {
    return guard([&] {
        Units u;
        u.temperature = temperature;
        EwaldTerm t(jsonlite::parse(coulomb_json), u);
        Space spc;
        spc.geometry.set({box, box, box});
        for (int i = 0; i < n; ++i) {
            Particle p;
            p.pos = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
            p.charge = q[i];
            spc.particles.push_back(p);
        }
        Group g;
        g.begin = 0;
        g.size = n;
        spc.groups.push_back(g);
        EwaldTerm::update_box(t.data, spc.geometry.len);
        EwaldTerm::full_update(t.data, spc);
        Change c;
        c.everything = true;
        *num_k = t.data.num_kvectors();
        *self = EwaldTerm::self_energy(t.data, c, spc);
        *surface = EwaldTerm::surface_energy(t.data, c, spc);
        *reciprocal = EwaldTerm::reciprocal_energy(t.data);
    });
}

// k-vector table (n as integers would lose nothing; kx,ky,kz,Ak are returned).
int orc_ewald_kvectors(const char* coulomb_json, double temperature, double box, int cap, double* out, int* num_k) {
    return guard([&] {
        Units u;
        u.temperature = temperature;
        EwaldTerm t(jsonlite::parse(coulomb_json), u);
        EwaldTerm::update_box(t.data, {box, box, box});
        *num_k = t.data.num_kvectors();
        for (int k = 0; k < std::min(cap, *num_k); ++k) {
            out[4 * k] = t.data.kvec[k].x;
            out[4 * k + 1] = t.data.kvec[k].y;
            out[4 * k + 2] = t.data.kvec[k].z;
            out[4 * k + 3] = t.data.Ak[k];
        }
    });
}

// Andrea spline of the tabulate.h doctest function f(x) = x sin(x)/2 + 2, or of the
// short-range function S(q) of a `coulomb` block (which = 1).
int orc_spline_test(int which, const char* coulomb_json, double utol, double ftol, double xmin, double xmax,
                    int cap, double* knots, double* coeff, int* nknots, int neval, const double* xs, double* ys) {
    return guard([&] {
        AndreaSpline sp;
        sp.set_tolerance(utol, ftol);
        SplineTable t;
        if (which == 0) {
            t = sp.generate([](double x) { return 0.5 * x * std::sin(x) + 2; }, xmin, xmax);
        } else {
            Units u;
            PairTerm pt = make_coulomb_term(jsonlite::parse(coulomb_json), u);
            t = pt.S;
        }
        *nknots = static_cast<int>(t.knots.size());
        for (int i = 0; i < std::min(cap, *nknots); ++i) knots[i] = t.knots[i];
        for (int i = 0; i < std::min(cap * 6, static_cast<int>(t.coeff.size())); ++i) coeff[i] = t.coeff[i];
        for (int i = 0; i < neval; ++i) ys[i] = AndreaSpline::eval(t, xs[i]);
    });
}

// One pair energy through the configured nonbonded term (types and charges given).
int orc_pair_energy(void* h, int type_a, double qa, int type_b, double qb, double r2, double* out) {
    return guard([&] {
        auto* s = static_cast<System*>(h);
        for (auto& t : s->accepted_pot->terms) {
            if (auto* nb = dynamic_cast<NonbondedTerm*>(t.get())) {
                Particle a, b;
                a.id = type_a;
                a.charge = qa;
                b.id = type_b;
                b.charge = qb;
                *out = nb->pot(a, b, r2);
                return;
            }
        }
        throw std::runtime_error("no nonbonded term");
    });
}

// Exact (functor) pair energy of the nonbonded term, bypassing a spline table.
int orc_pair_energy_exact(void* h, int type_a, double qa, int type_b, double qb, double r2, double* out) {
    return guard([&] {
        auto* s = static_cast<System*>(h);
        for (auto& t : s->accepted_pot->terms) {
            if (auto* nb = dynamic_cast<NonbondedTerm*>(t.get())) {
                Particle a, b;
                a.id = type_a;
                a.charge = qa;
                b.id = type_b;
                b.charge = qb;
                *out = nb->pot.exact(a, b, r2);
                return;
            }
        }
        throw std::runtime_error("no nonbonded term");
    });
}

// The nonbonded_splined table of a type pair: knots (ascending r^2), 6 coefficients per
// interval, rmin^2, rmax^2 and the hard-sphere flag.
int orc_splined_table(void* h, int type_a, int type_b, int cap, double* knots, double* coeff, int* num_knots,
                      double* rmin2, double* rmax2, int* hardsphere) {
    return guard([&] {
        auto* s = static_cast<System*>(h);
        for (auto& t : s->accepted_pot->terms) {
            if (auto* nb = dynamic_cast<NonbondedTerm*>(t.get())) {
                if (!nb->pot.splined) throw std::runtime_error("the nonbonded term is not splined");
                const int nt = nb->pot.ntypes;
                if (type_a < 0 || type_b < 0 || type_a >= nt || type_b >= nt) throw std::runtime_error("bad type");
                const SplinedPair& p = nb->pot.tables[static_cast<size_t>(type_a * nt + type_b)];
                const int n = static_cast<int>(p.table.knots.size());
                *num_knots = n;
                *rmin2 = p.rmin2;
                *rmax2 = p.rmax2;
                *hardsphere = p.hardsphere ? 1 : 0;
                if (n > cap) throw std::runtime_error("knot buffer too small");
                std::copy(p.table.knots.begin(), p.table.knots.end(), knots);
                std::copy(p.table.coeff.begin(), p.table.coeff.end(), coeff);
                return;
            }
        }
        throw std::runtime_error("no nonbonded term");
    });
}

double orc_bjerrum_length(double temperature, double epsr) {
    Units u;
    u.temperature = temperature;
    return u.bjerrum_length(epsr);
}

void orc_philox(const unsigned* ctr, const unsigned* key, unsigned* out) {
    auto r = philox4x32_10({ctr[0], ctr[1], ctr[2], ctr[3]}, {key[0], key[1]});
    for (int i = 0; i < 4; ++i) out[i] = r[static_cast<size_t>(i)];
}

} // extern "C"
