// C ABI of the engine (declared in include/faunus_amd.h).
#include <cmath>
#include <cstring>
#include <exception>
#include <memory>
#include <string>
#include <vector>

#include "../../include/faunus_amd.h"
#include "engine.hpp"

struct fmc_engine {
    std::unique_ptr<fmc::Engine> impl;
};

namespace {
thread_local std::string g_error;

template <class F> int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return 1;
    } catch (...) {
        g_error = "unknown error";
        return 1;
    }
}

fmc::Engine& eng(fmc_engine* e) {
    if (!e || !e->impl) throw std::runtime_error("null engine handle");
    return *e->impl;
}

struct ChangeLists {
    std::vector<int> groups;
    std::vector<char> all, internal;
    std::vector<std::vector<int>> rel;
};

ChangeLists unpack(const fmc_change* c) {
    if (!c) throw std::runtime_error("null change");
    if (c->everything || c->volume_change || c->matter_change)
        throw std::runtime_error("everything/volume/matter changes go through fmc_total_energy or are unsupported");
    ChangeLists L;
    for (int g = 0; g < c->num_groups; ++g) {
        L.groups.push_back(c->group_index[g]);
        L.all.push_back(c->all ? static_cast<char>(c->all[g] != 0) : 0);
        L.internal.push_back(c->internal ? static_cast<char>(c->internal[g] != 0) : 0);
        std::vector<int> r;
        if (c->offsets)
            for (int k = c->offsets[g]; k < c->offsets[g + 1]; ++k) r.push_back(c->indices[k]);
        L.rel.push_back(r);
    }
    return L;
}

double sum_terms(const double* t, int n) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += t[i];
    return s;
}

// MetropolisMonteCarlo::getEnergyChange (montecarlo.cpp:193-209)
double energy_change(double new_energy, double old_energy) {
    if (std::isnan(old_energy) && !std::isnan(new_energy)) return -INFINITY;
    if (std::isnan(new_energy)) return INFINITY;
    if (new_energy > 0.0 && std::isinf(new_energy)) return INFINITY;
    const double du = new_energy - old_energy;
    if (std::isnan(du)) return 0.0;
    return du;
}

// MetropolisMonteCarlo::metropolisCriterion (montecarlo.cpp:17-34); the uniform
// deviate is supplied by the caller.
bool metropolis(double du, double u) {
    if (std::isnan(du)) throw std::runtime_error("Metropolis error: energy cannot be NaN");
    if (std::isinf(du) && du < 0.0) return true;
    if (-du > 709.782712893384) return true;
    return u <= std::exp(-du);
}
} // namespace

extern "C" {

const char* fmc_last_error(void) { return g_error.c_str(); }

int fmc_device_count(int* count) {
    return guarded([&] {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
        *count = n;
    });
}

int fmc_create(const char* config_json, fmc_engine** out) {
    return guarded([&] {
        if (!config_json || !out) throw std::runtime_error("null argument");
        auto* h = new fmc_engine;
        try {
            h->impl = std::make_unique<fmc::Engine>(config_json);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int fmc_destroy(fmc_engine* engine) {
    return guarded([&] { delete engine; });
}

int fmc_set_particles(fmc_engine* e, int n, const double* xyz, const double* charge, const int* atom_id) {
    return guarded([&] { eng(e).set_particles(n, xyz, charge, atom_id); });
}

int fmc_set_groups(fmc_engine* e, int ng, const int* begin, const int* size, const int* molid, const double* cm) {
    return guarded([&] { eng(e).set_groups(ng, begin, size, molid, cm); });
}

int fmc_parse_state(const char* state_json, int* n, int* num_groups, double* box, double* xyz, double* charge,
                    int* atom_id, int* begin, int* size, int* molid, double* mass_center) {
    return guarded([&] {
        if (!state_json) throw std::runtime_error("null state text");
        const fmc::StateCfg st = fmc::StateCfg::from_json(state_json);
        const int np = static_cast<int>(st.charge.size()), ng = static_cast<int>(st.size.size());
        if (n) *n = np;
        if (num_groups) *num_groups = ng;
        if (box) std::copy(st.box, st.box + 3, box);
        if (xyz) std::copy(st.xyz.begin(), st.xyz.end(), xyz);
        if (charge) std::copy(st.charge.begin(), st.charge.end(), charge);
        if (atom_id) std::copy(st.atom_id.begin(), st.atom_id.end(), atom_id);
        if (begin) std::copy(st.begin.begin(), st.begin.end(), begin);
        if (size) std::copy(st.size.begin(), st.size.end(), size);
        if (molid) std::copy(st.molid.begin(), st.molid.end(), molid);
        if (mass_center) std::copy(st.mass_center.begin(), st.mass_center.end(), mass_center);
    });
}

int fmc_load_state(fmc_engine* e, const char* state_json) {
    return guarded([&] {
        if (!state_json) throw std::runtime_error("null state text");
        const fmc::StateCfg st = fmc::StateCfg::from_json(state_json);
        const double* box = eng(e).model().box;
        for (int d = 0; d < 3; ++d)
            if (std::fabs(st.box[d] - box[d]) > 1e-12 * std::fabs(box[d]))
                throw std::runtime_error("state: box differs from the input's geometry");
        eng(e).set_particles(static_cast<int>(st.charge.size()), st.xyz.data(), st.charge.data(), st.atom_id.data());
        eng(e).set_groups(static_cast<int>(st.size.size()), st.begin.data(), st.size.data(), st.molid.data(),
                          st.mass_center.data());
    });
}

int fmc_init(fmc_engine* e, double* initial_energy) {
    return guarded([&] {
        eng(e).init();
        if (initial_energy) {
            std::vector<double> t(static_cast<size_t>(eng(e).num_terms()));
            eng(e).total_energy(t.data());
            *initial_energy = sum_terms(t.data(), static_cast<int>(t.size()));
        }
    });
}

int fmc_num_terms(fmc_engine* e, int* count) {
    return guarded([&] { *count = eng(e).num_terms(); });
}

int fmc_term_name(fmc_engine* e, int index, char* buffer, int length) {
    return guarded([&] {
        const auto& names = eng(e).model().term_names;
        if (index < 0 || index >= static_cast<int>(names.size())) throw std::runtime_error("term index out of range");
        if (length <= 0) throw std::runtime_error("buffer too small");
        std::strncpy(buffer, names[static_cast<size_t>(index)].c_str(), static_cast<size_t>(length - 1));
        buffer[length - 1] = 0;
    });
}

int fmc_stage_trial(fmc_engine* e, int count, const int* particle_index, const double* xyz) {
    return guarded([&] { eng(e).stage_trial(count, particle_index, xyz); });
}

int fmc_stage_mass_center(fmc_engine* e, int group, const double* cm) {
    return guarded([&] {
        if (!cm) throw std::runtime_error("null mass centre");
        eng(e).stage_mass_center(group, cm);
    });
}

int fmc_get_mass_center(fmc_engine* e, int group, double* cm) {
    return guarded([&] {
        const auto& c = eng(e).mass_centers();
        if (group < 0 || 3 * static_cast<size_t>(group) + 2 >= c.size()) throw std::runtime_error("group out of range");
        for (int d = 0; d < 3; ++d) cm[d] = c[3 * static_cast<size_t>(group) + static_cast<size_t>(d)];
    });
}

int fmc_delta_u(fmc_engine* e, const fmc_change* change, double* new_terms, double* old_terms) {
    return guarded([&] {
        ChangeLists L = unpack(change);
        eng(e).delta(L.groups, L.all, L.internal, L.rel, new_terms, old_terms);
    });
}

int fmc_energy(fmc_engine* e, const fmc_change* change, int state, double* terms, double* total) {
    return guarded([&] {
        fmc::Engine& E = eng(e);
        const int nt = E.num_terms();
        if (change && change->everything) {
            if (state != FMC_STATE_ACCEPTED) throw std::runtime_error("Change::everything is evaluated on the accepted state");
            E.total_energy(terms);
        } else {
            ChangeLists L = unpack(change);
            std::vector<double> en(static_cast<size_t>(nt)), eo(static_cast<size_t>(nt));
            E.delta(L.groups, L.all, L.internal, L.rel, en.data(), eo.data());
            const std::vector<double>& src = (state == FMC_STATE_TRIAL) ? en : eo;
            std::copy(src.begin(), src.end(), terms);
        }
        if (total) *total = sum_terms(terms, nt);
    });
}

int fmc_sync(fmc_engine* e, int accept) {
    return guarded([&] { eng(e).sync(accept != 0); });
}

int fmc_total_energy(fmc_engine* e, double* terms, double* total) {
    return guarded([&] {
        eng(e).total_energy(terms);
        if (total) *total = sum_terms(terms, eng(e).num_terms());
    });
}

int fmc_check_q(fmc_engine* e, double* max_abs_dq, double* rel_reciprocal) {
    return guarded([&] { eng(e).check_q(max_abs_dq, rel_reciprocal); });
}

int fmc_get_positions(fmc_engine* e, double* xyz) {
    return guarded([&] { eng(e).get_positions(xyz); });
}

int fmc_replay_translate(fmc_engine* e, long count, const int* group, const int* index, const double* xyz,
                         const double* uniform, double* du_out, unsigned char* accepted, double* new_terms,
                         double* old_terms) {
    return guarded([&] {
        fmc::Engine& E = eng(e);
        const int nt = E.num_terms();
        std::vector<double> en(static_cast<size_t>(nt)), eo(static_cast<size_t>(nt));
        const auto& gb = E.group_begin();
        std::vector<int> groups(1);
        std::vector<char> all(1, 0), internal(1, 1);
        std::vector<std::vector<int>> rel(1, std::vector<int>(1));
        for (long n = 0; n < count; ++n) {
            const int g = group[n];
            if (g < 0 || g >= static_cast<int>(gb.size())) throw std::runtime_error("replay: group out of range");
            const int abs_index = gb[static_cast<size_t>(g)] + index[n];
            E.stage_trial(1, &abs_index, xyz + 3 * n);
            groups[0] = g;
            rel[0][0] = index[n];
            E.delta(groups, all, internal, rel, en.data(), eo.data());
            const double du = energy_change(sum_terms(en.data(), nt), sum_terms(eo.data(), nt));
            const bool acc = metropolis(du, uniform[n]);
            E.sync(acc);
            if (du_out) du_out[n] = du;
            if (accepted) accepted[n] = acc ? 1 : 0;
            for (int t = 0; t < nt; ++t) {
                if (new_terms) new_terms[n * nt + t] = en[static_cast<size_t>(t)];
                if (old_terms) old_terms[n * nt + t] = eo[static_cast<size_t>(t)];
            }
        }
    });
}

int fmc_replay_moves(fmc_engine* e, long count, const int* group, const long* offsets, const int* index,
                     const double* xyz, const double* mass_center, const unsigned char* whole, const double* uniform,
                     double* du_out, unsigned char* accepted) {
    return guarded([&] {
        fmc::Engine& E = eng(e);
        const int nt = E.num_terms();
        std::vector<double> en(static_cast<size_t>(nt)), eo(static_cast<size_t>(nt));
        const auto& gb = E.group_begin();
        const auto& gsz = E.group_size();
        std::vector<int> groups(1);
        std::vector<char> all(1, 0), internal(1, 1);
        std::vector<std::vector<int>> rel(1);
        for (long n = 0; n < count; ++n) {
            const long o0 = offsets[n], o1 = offsets[n + 1];
            if (o1 < o0) throw std::runtime_error("replay: offsets must not decrease");
            if (o1 == o0) { // a move that changed nothing
                if (du_out) du_out[n] = 0.0;
                if (accepted) accepted[n] = 1;
                continue;
            }
            const int g = group[n];
            if (g < 0 || g >= static_cast<int>(gb.size())) throw std::runtime_error("replay: group out of range");
            E.stage_trial(static_cast<int>(o1 - o0), index + o0, xyz + 3 * o0);
            if (mass_center) E.stage_mass_center(g, mass_center + 3 * n);
            groups[0] = g;
            all[0] = whole && whole[n] ? 1 : 0;
            internal[0] = all[0] ? 0 : 1;
            rel[0].clear();
            if (!all[0])
                for (long k = o0; k < o1; ++k) {
                    const int r = index[k] - gb[static_cast<size_t>(g)];
                    if (r < 0 || r >= gsz[static_cast<size_t>(g)]) throw std::runtime_error("replay: atom outside its group");
                    rel[0].push_back(r);
                }
            E.delta(groups, all, internal, rel, en.data(), eo.data());
            const double du = energy_change(sum_terms(en.data(), nt), sum_terms(eo.data(), nt));
            const bool acc = metropolis(du, uniform[n]);
            E.sync(acc);
            if (du_out) du_out[n] = du;
            if (accepted) accepted[n] = acc ? 1 : 0;
        }
    });
}

int fmc_run_translate(fmc_engine* e, long count, const int* group, const int* index, const double* displacement,
                      const double* uniform, double* du_out, unsigned char* accepted) {
    return guarded([&] {
        fmc::Engine& E = eng(e);
        const int nt = E.num_terms();
        std::vector<double> en(static_cast<size_t>(nt)), eo(static_cast<size_t>(nt));
        const auto& gb = E.group_begin();
        const auto& box = E.model().box;
        std::vector<int> groups(1);
        std::vector<char> all(1, 0), internal(1, 1);
        std::vector<std::vector<int>> rel(1, std::vector<int>(1));
        std::vector<double> xyz(3 * static_cast<size_t>(E.num_particles()));
        E.get_positions(xyz.data());
        for (long n = 0; n < count; ++n) {
            const int g = group[n];
            if (g < 0 || g >= static_cast<int>(gb.size())) throw std::runtime_error("run: group out of range");
            const int i = gb[static_cast<size_t>(g)] + index[n];
            // AtomicTranslateRotate::translateParticle (move.cpp:225-240) and
            // Cuboid::boundary (geometry.cpp:91-104)
            double p[3];
            for (int d = 0; d < 3; ++d) {
                double x = xyz[3 * static_cast<size_t>(i) + d] + displacement[3 * n + d];
                if (std::fabs(x) > 0.5 * box[d]) {
                    const double s = x * (1.0 / box[d]);
                    x -= box[d] * static_cast<int>(s > 0.0 ? s + 0.5 : s - 0.5);
                }
                p[d] = x;
            }
            E.stage_trial(1, &i, p);
            groups[0] = g;
            rel[0][0] = index[n];
            E.delta(groups, all, internal, rel, en.data(), eo.data());
            const double du = energy_change(sum_terms(en.data(), nt), sum_terms(eo.data(), nt));
            const bool acc = metropolis(du, uniform[n]);
            E.sync(acc);
            if (acc)
                for (int d = 0; d < 3; ++d) xyz[3 * static_cast<size_t>(i) + d] = p[d];
            if (du_out) du_out[n] = du;
            if (accepted) accepted[n] = acc ? 1 : 0;
        }
    });
}

int fmc_copy_probe(long long bytes, int repeats, double* ms_per_launch) {
    return guarded([&] { *ms_per_launch = fmc::copy_probe(bytes, repeats); });
}

int fmc_set_timing(fmc_engine* e, int enable) {
    return guarded([&] { eng(e).set_timing(enable != 0); });
}

int fmc_get_timing(fmc_engine* e, double* total_ms, long* launches) {
    return guarded([&] { eng(e).timing(total_ms, launches); });
}

int fmc_coulomb_spline(fmc_engine* e, int capacity, double* knots, double* coeff, int* num_knots) {
    return guarded([&] {
        const auto& S = eng(e).model().coulomb.S;
        *num_knots = static_cast<int>(S.knots.size());
        for (int i = 0; i < capacity && i < *num_knots; ++i) knots[i] = S.knots[static_cast<size_t>(i)];
        for (int i = 0; i < 6 * capacity && i < static_cast<int>(S.coeff.size()); ++i) coeff[i] = S.coeff[static_cast<size_t>(i)];
    });
}

int fmc_splined_table(fmc_engine* e, int type_a, int type_b, int capacity, double* knots, double* coeff,
                      int* num_knots, double* rmin2, double* rmax2, int* hardsphere) {
    return guarded([&] {
        const fmc::Model& m = eng(e).model();
        if (!m.splined) throw std::runtime_error("the nonbonded term is not nonbonded_splined");
        if (type_a < 0 || type_b < 0 || type_a >= m.ntypes || type_b >= m.ntypes) throw std::runtime_error("bad atom type");
        const fmc::SplinedPairCfg& s = m.spl[static_cast<size_t>(type_a * m.ntypes + type_b)];
        *num_knots = static_cast<int>(s.table.knots.size());
        *rmin2 = s.rmin2;
        *rmax2 = s.rmax2;
        *hardsphere = s.hardsphere ? 1 : 0;
        for (int i = 0; i < capacity && i < *num_knots; ++i) knots[i] = s.table.knots[static_cast<size_t>(i)];
        for (int i = 0; i < 6 * capacity && i < static_cast<int>(s.table.coeff.size()); ++i)
            coeff[i] = s.table.coeff[static_cast<size_t>(i)];
    });
}

int fmc_num_kvectors(fmc_engine* e, int* count) {
    return guarded([&] { *count = eng(e).num_kvectors(); });
}

} // extern "C"

// ---------------------------------------------------------------- batches
#include "mc_batch.hpp"

struct fmc_batch {
    std::unique_ptr<fmc::Batch> impl;
};

namespace {
fmc::Batch& bat(fmc_batch* b) {
    if (!b || !b->impl) throw std::runtime_error("null batch handle");
    return *b->impl;
}
} // namespace

extern "C" {

int fmc_batch_set_cell_divisions(fmc_batch* b, int divisions) {
    return guarded([&] { bat(b).set_cell_divisions(divisions); });
}

int fmc_batch_set_cell_capacity(fmc_batch* b, int capacity) {
    return guarded([&] { bat(b).set_cell_capacity(capacity); });
}

int fmc_batch_get_replica_state(fmc_batch* b, int replica, double* xyz, double* mass_center, unsigned long long* moves,
                                unsigned long long* accepted) {
    return guarded([&] {
        if (!xyz) throw std::runtime_error("get_replica_state: positions buffer is null");
        bat(b).get_replica_state(replica, xyz, mass_center, moves, accepted);
    });
}

int fmc_batch_set_replica_state(fmc_batch* b, int replica, const double* xyz, const double* mass_center,
                                unsigned long long moves, unsigned long long accepted) {
    return guarded([&] {
        if (!xyz) throw std::runtime_error("set_replica_state: positions are null");
        bat(b).set_replica_state(replica, xyz, mass_center, moves, accepted);
    });
}

int fmc_batch_create(int fp32_pairs, fmc_batch** out) {
    return guarded([&] {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n < 1) throw std::runtime_error("no HIP device visible");
        auto* b = new fmc_batch;
        b->impl = std::make_unique<fmc::Batch>(fp32_pairs);
        *out = b;
    });
}

int fmc_batch_destroy(fmc_batch* b) {
    return guarded([&] { delete b; });
}

int fmc_batch_add_replica(fmc_batch* b, const char* config_json, int n, const double* xyz, const double* charge,
                          const int* atom_id, int num_groups, const int* begin, const int* size, const int* molid,
                          double beta, unsigned long long seed, int ladder, int* replica) {
    return guarded([&] {
        const int r = bat(b).add_replica(config_json, n, xyz, charge, atom_id, num_groups, begin, size, molid, beta,
                                         seed, ladder);
        if (replica) *replica = r;
    });
}

int fmc_batch_add_replica_moves(fmc_batch* b, const char* config_json, int n, const double* xyz, const double* charge,
                                const int* atom_id, int num_groups, const int* begin, const int* size, const int* molid,
                                double beta, unsigned long long seed, int ladder, int num_kinds, const int* kind,
                                const int* kind_molid, const unsigned* weight, const double* dp, const double* dprot,
                                int* replica) {
    return guarded([&] {
        if (num_kinds < 1 || !kind || !kind_molid || !weight || !dp || !dprot) throw std::runtime_error("empty move set");
        std::vector<fmc::MoveKindSpec> kinds(static_cast<size_t>(num_kinds));
        for (int k = 0; k < num_kinds; ++k) {
            kinds[static_cast<size_t>(k)].kind = kind[k];
            kinds[static_cast<size_t>(k)].molid = kind_molid[k];
            kinds[static_cast<size_t>(k)].weight = weight[k];
            kinds[static_cast<size_t>(k)].dp = dp[k];
            kinds[static_cast<size_t>(k)].dprot = dprot[k];
        }
        const int r = bat(b).add_replica(config_json, n, xyz, charge, atom_id, num_groups, begin, size, molid, beta,
                                         seed, ladder, &kinds);
        if (replica) *replica = r;
    });
}

int fmc_batch_run(fmc_batch* b, const long long* moves, int trace_capacity) {
    return guarded([&] {
        std::vector<long long> m(moves, moves + bat(b).size());
        bat(b).run(m, trace_capacity > 0, trace_capacity);
    });
}

int fmc_batch_run_split(fmc_batch* b, long long moves, int groups, int stride, int trace_capacity) {
    return guarded([&] { bat(b).run_split(moves, groups, stride, trace_capacity > 0, trace_capacity); });
}

int fmc_pair_probe(fmc_engine* e, int type_a, int type_b, double qa, double qb, int n, const double* r2, double* u64,
                   double* u32) {
    return guarded([&] { fmc::pair_probe(eng(e), type_a, type_b, qa, qb, n, r2, u64, u32); });
}

int fmc_batch_capture_interval(fmc_batch* b, const long long* moves, unsigned long long seed, unsigned long long round0) {
    return guarded([&] {
        if (!moves) throw std::runtime_error("capture_interval: moves are null");
        bat(b).capture_interval(std::vector<long long>(moves, moves + bat(b).size()), seed, round0);
    });
}

int fmc_batch_run_ladders(fmc_batch* b, const long long* moves, const int* intervals, unsigned long long seed,
                          unsigned long long round0) {
    return guarded([&] {
        if (!moves || !intervals) throw std::runtime_error("run_ladders: moves or intervals are null");
        const int n = bat(b).size();
        bat(b).run_ladders(std::vector<long long>(moves, moves + n), std::vector<int>(intervals, intervals + n), seed,
                           round0);
    });
}

int fmc_batch_replay(fmc_batch* b, int count) {
    return guarded([&] { bat(b).replay(count); });
}

int fmc_batch_swap_counts(fmc_batch* b, long long* attempted, long long* accepted) {
    return guarded([&] { bat(b).swap_counts(attempted, accepted); });
}

int fmc_batch_autotune(fmc_batch* b, long long trial_moves, int capacity, int* num_candidates, int* values,
                       double* us_per_move, int* chosen) {
    return guarded([&] {
        const auto r = bat(b).autotune(trial_moves);
        *num_candidates = static_cast<int>(r.size());
        for (int i = 0; i < capacity && i < static_cast<int>(r.size()); ++i) {
            values[i] = r[static_cast<size_t>(i)].value;
            us_per_move[i] = r[static_cast<size_t>(i)].us_per_move;
        }
        *chosen = bat(b).tuned();
    });
}

int fmc_batch_run_split_static(fmc_batch* b, long long moves, int groups, int stride) {
    return guarded([&] { bat(b).run_split_static(moves, groups, stride); });
}

int fmc_batch_run_pipe(fmc_batch* b, long long moves, int groups, int depth, int trace_capacity) {
    return guarded([&] { bat(b).run_pipe(moves, groups, depth, trace_capacity > 0, trace_capacity); });
}

int fmc_batch_split_error(fmc_batch* b, int* error) {
    return guarded([&] { *error = bat(b).split_error(); });
}

int fmc_batch_sync(fmc_batch* b, double* kernel_ms) {
    return guarded([&] {
        bat(b).sync();
        if (kernel_ms) *kernel_ms = bat(b).last_kernel_ms();
    });
}

int fmc_batch_set_timing(fmc_batch* b, int enable) {
    return guarded([&] { bat(b).set_timing(enable != 0); });
}

int fmc_batch_swap(fmc_batch* b, unsigned long long seed, unsigned long long round, long long* attempted,
                   long long* accepted) {
    return guarded([&] {
        bat(b).swap_ladders(seed, round);
        if (attempted) *attempted = bat(b).swaps_attempted();
        if (accepted) *accepted = bat(b).swaps_accepted();
    });
}

int fmc_batch_replica_info(fmc_batch* b, int replica, double* energy, double* beta, unsigned long long* moves,
                           unsigned long long* accepted, int* error) {
    return guarded([&] {
        if (replica < 0 || replica >= bat(b).size()) throw std::runtime_error("replica out of range");
        const fmc::ReplicaState s = bat(b).state(replica);
        if (energy) *energy = s.energy;
        if (beta) *beta = s.beta;
        if (moves) *moves = s.moves;
        if (accepted) *accepted = s.accepted;
        if (error) *error = s.error;
    });
}

int fmc_batch_sample_energy(fmc_batch* b, long long every) {
    return guarded([&] { bat(b).sample_energy(every); });
}

int fmc_batch_energy_stats(fmc_batch* b, int replica, unsigned long long* samples, double* sum, double* sum_sq) {
    return guarded([&] {
        if (replica < 0 || replica >= bat(b).size()) throw std::runtime_error("replica out of range");
        const fmc::ReplicaState s = bat(b).state(replica);
        if (samples) *samples = s.e_samples;
        if (sum) *sum = s.e_sum;
        if (sum_sq) *sum_sq = s.e_sum2;
    });
}

int fmc_batch_rdf(fmc_batch* b, int replica, int type_a, int type_b, double dr, int nbins, unsigned long long* counts) {
    return guarded([&] {
        if (!counts) throw std::runtime_error("rdf: null counts");
        bat(b).rdf(replica, type_a, type_b, dr, nbins, counts);
    });
}

int fmc_batch_positions(fmc_batch* b, int replica, double* xyz) {
    return guarded([&] {
        if (replica < 0 || replica >= bat(b).size()) throw std::runtime_error("replica out of range");
        bat(b).positions(replica, xyz);
    });
}

int fmc_batch_trace(fmc_batch* b, int replica, int count, int* index, unsigned char* accepted, double* du,
                    double* xyz) {
    return guarded([&] {
        if (replica < 0 || replica >= bat(b).size()) throw std::runtime_error("replica out of range");
        auto t = bat(b).trace(replica, count);
        for (int k = 0; k < count; ++k) {
            const auto& e = t[static_cast<size_t>(k)];
            if (index) index[k] = e.index;
            if (accepted) accepted[k] = static_cast<unsigned char>(e.accepted);
            if (du) du[k] = e.du;
            if (xyz) {
                xyz[3 * k] = e.x;
                xyz[3 * k + 1] = e.y;
                xyz[3 * k + 2] = e.z;
            }
        }
    });
}

int fmc_batch_total_energy(fmc_batch* b, int replica, double* terms, double* total) {
    return guarded([&] {
        if (replica < 0 || replica >= bat(b).size()) throw std::runtime_error("replica out of range");
        bat(b).sync();
        fmc::Engine& e = bat(b).engine(replica);
        e.refresh_host_positions();
        e.total_energy(terms);
        if (total) *total = sum_terms(terms, e.num_terms());
    });
}

int fmc_batch_check_q(fmc_batch* b, int replica, double* max_abs_dq, double* rel_reciprocal) {
    return guarded([&] {
        if (replica < 0 || replica >= bat(b).size()) throw std::runtime_error("replica out of range");
        bat(b).sync();
        bat(b).engine(replica).check_q(max_abs_dq, rel_reciprocal);
    });
}

} // extern "C"

// ---------------------------------------------------------------- serving
#include "serve.hpp"

struct fmc_server {
    std::unique_ptr<fmc::Server> impl;
};

namespace {
fmc::Server& srv(fmc_server* s) {
    if (!s || !s->impl) throw std::runtime_error("null server handle");
    return *s->impl;
}
} // namespace

extern "C" {

int fmc_server_create(int fp32_pairs, long long hbm_budget_bytes, int max_slots, fmc_server** out) {
    return guarded([&] {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n < 1) throw std::runtime_error("no HIP device visible");
        auto* s = new fmc_server;
        s->impl = std::make_unique<fmc::Server>(fp32_pairs, hbm_budget_bytes, max_slots);
        *out = s;
    });
}

int fmc_server_destroy(fmc_server* s) {
    return guarded([&] { delete s; });
}

int fmc_server_set_split(fmc_server* s, int atoms_per_workgroup, int max_workgroups) {
    return guarded([&] { srv(s).set_split(atoms_per_workgroup, max_workgroups); });
}

int fmc_server_submit(fmc_server* s, const char* config_json, int n, const double* xyz, const double* charge,
                      const int* atom_id, int num_groups, const int* begin, const int* size, const int* molid,
                      long long moves, unsigned long long seed, int* job) {
    return guarded([&] {
        const int id = srv(s).submit(config_json, n, xyz, charge, atom_id, num_groups, begin, size, molid, moves, seed);
        if (job) *job = id;
    });
}

int fmc_server_submit_moves(fmc_server* s, const char* config_json, int n, const double* xyz, const double* charge,
                            const int* atom_id, int num_groups, const int* begin, const int* size, const int* molid,
                            long long moves, unsigned long long seed, int num_kinds, const int* kind,
                            const int* kind_molid, const unsigned* weight, const double* dp, const double* dprot,
                            int* job) {
    return guarded([&] {
        if (num_kinds < 1 || !kind || !kind_molid || !weight || !dp || !dprot) throw std::runtime_error("empty move set");
        std::vector<fmc::MoveKindSpec> kinds(static_cast<size_t>(num_kinds));
        for (int k = 0; k < num_kinds; ++k)
            kinds[static_cast<size_t>(k)] = fmc::MoveKindSpec{kind[k], kind_molid[k], weight[k], dp[k], dprot[k]};
        const int id = srv(s).submit(config_json, n, xyz, charge, atom_id, num_groups, begin, size, molid, moves, seed,
                                     &kinds);
        if (job) *job = id;
    });
}

int fmc_server_step(fmc_server* s, long long quantum, int* resident, int* queued) {
    return guarded([&] {
        srv(s).step(quantum);
        if (resident) *resident = srv(s).resident();
        if (queued) *queued = srv(s).queued();
    });
}

int fmc_server_job(fmc_server* s, int job, int* state, double* energy, unsigned long long* moves,
                   unsigned long long* accepted, double* t_submit, double* t_admit, double* t_done) {
    return guarded([&] {
        const fmc::JobRecord& j = srv(s).job(job);
        if (state) *state = j.state;
        if (energy) *energy = j.energy;
        if (moves) *moves = j.moves_done;
        if (accepted) *accepted = j.accepted;
        if (t_submit) *t_submit = j.t_submit;
        if (t_admit) *t_admit = j.t_admit;
        if (t_done) *t_done = j.t_done;
    });
}

int fmc_server_job_positions(fmc_server* s, int job, double* xyz) {
    return guarded([&] {
        const fmc::JobRecord& j = srv(s).job(job);
        if (j.state != 2) throw std::runtime_error("job not finished");
        std::copy(j.final_xyz.begin(), j.final_xyz.end(), xyz);
    });
}

int fmc_server_status(fmc_server* s, double* seconds, long long* moves_done, long long* bytes_in_use) {
    return guarded([&] {
        if (seconds) *seconds = srv(s).now();
        if (moves_done) *moves_done = srv(s).moves_done_total();
        if (bytes_in_use) *bytes_in_use = static_cast<long long>(srv(s).footprint_in_use());
    });
}

} // extern "C"

extern "C" int fmc_server_job_error(fmc_server* s, int job, char* buffer, int length) {
    return guarded([&] {
        const fmc::JobRecord& j = srv(s).job(job);
        if (!buffer || length <= 0) throw std::runtime_error("buffer too small");
        std::strncpy(buffer, j.error.c_str(), static_cast<size_t>(length - 1));
        buffer[length - 1] = 0;
    });
}
