// ORACLE — test infrastructure only (see oracle.hpp).  FP64 restatement of the
// reference's nonbonded + Ewald energy path and its Metropolis driver.
#include "oracle.hpp"

#include <algorithm>
#include <memory>
#include <numeric>
#include <stdexcept>

namespace oracle {

using jsonlite::Value;

// ---------------------------------------------------------------- Andrea spline
// tabulate.h:223-307 (generate), 113-141 (SetUBuffer), 148-182 (CheckUBuffer).
SplineTable AndreaSpline::generate(const std::function<double(double)>& f, double xmin, double xmax) const {
    const double rmin = std::sqrt(xmin);
    const double rmax = std::sqrt(xmax);
    SplineTable t;
    t.xmin = rmin * rmin;
    t.xmax = rmax * rmax;
    double r_floor = rmin;
    double r_upper = rmax;
    double z_upper = rmax * rmax;
    bool repulsive = false;
    std::vector<double> knots_desc{z_upper};
    std::vector<std::array<double, 6>> packets_desc;

    auto quintic = [&](double zlo, double zhi, double u0l, double u1l, double u2l, double u0h,
                       double u1h, double u2h) -> std::array<double, 6> {
        if (std::fabs(u0l) < 1e-9 && std::fabs(u1l) < 1e-9) return {0, 0, 0, 0, 0, 0};
        const double h1 = zhi - zlo, h2 = h1 * h1, h3 = h2 * h1;
        const double c0 = u0l, c1 = u1l, c2 = u2l * 0.5;
        const double a = 6 * (u0h - c0 - c1 * h1 - c2 * h2) / h3;
        const double b = 2 * (u1h - c1 - 2 * c2 * h1) / h2;
        const double c = (u2h - 2 * c2) / h1;
        const double c3 = (10 * a - 12 * b + 3 * c) / 6;
        const double c4 = (-15 * a + 21 * b - 6 * c) / (6 * h1);
        const double c5 = (2 * a - 3 * b + c) / (2 * h2);
        return {c0, c1, c2, c3, c4, c5};
    };
    // returns {within tolerance, repulsive part found}
    auto verify = [&](const std::array<double, 6>& c, double rlo, double rhi) -> std::pair<bool, bool> {
        const int ncheck = 11;
        const double step = (rhi - rlo) / (ncheck - 1);
        bool rep = false;
        for (int i = 0; i < ncheck; i++) {
            const double r1 = rlo + step * static_cast<double>(i);
            const double x = r1 * r1;
            const double u0 = f(x);
            const double u1 = d1(f, x);
            const double dz = x - rlo * rlo;
            const double usum = c[0] + dz * (c[1] + dz * (c[2] + dz * (c[3] + dz * (c[4] + dz * c[5]))));
            const double fsum = c[1] + dz * (2 * c[2] + dz * (3 * c[3] + dz * (4 * c[4] + dz * (5 * c[5]))));
            if (std::fabs(usum - u0) > utol) return {false, false};
            if (ftol != -1 && std::fabs(fsum - u1) > ftol) return {false, false};
            if (umaxtol != -1 && std::fabs(usum) > umaxtol) rep = true;
            if (fmaxtol != -1 && std::fabs(usum) > fmaxtol) rep = true;
        }
        return {true, rep};
    };

    int i = 0;
    for (; i < max_grid; i++) {
        double r_lower = r_upper;
        double z_lower = 0;
        std::array<double, 6> packet{};
        double width = r_upper - rmin;
        int j = 0;
        for (; j < max_shrink; j++) {
            z_upper = r_upper * r_upper;
            r_lower = r_upper - width;
            if (r_floor > r_lower) r_lower = r_floor;
            z_lower = r_lower * r_lower;
            packet = quintic(z_lower, z_upper, f(z_lower), d1(f, z_lower), d2(f, z_lower), f(z_upper),
                             d1(f, z_upper), d2(f, z_upper));
            auto [ok, rep] = verify(packet, r_lower, r_upper);
            repulsive = rep;
            if (ok) {
                r_upper = r_lower;
                break;
            }
            width *= shrink;
        }
        if (j >= max_shrink) throw std::runtime_error("Andrea spline: try to increase utol/ftol");
        knots_desc.push_back(z_lower);
        packets_desc.push_back(packet);
        if (repulsive) {
            r_floor = r_lower;
            t.xmin = r_lower * r_lower;
        }
        if (r_lower <= r_floor || repulsive) break;
    }
    if (i >= max_grid) throw std::runtime_error("Andrea spline: try to increase utol/ftol");
    t.knots.assign(knots_desc.rbegin(), knots_desc.rend());
    for (auto it = packets_desc.rbegin(); it != packets_desc.rend(); ++it)
        t.coeff.insert(t.coeff.end(), it->begin(), it->end());
    return t;
}

// tabulate.h:184-202 — interval by lower_bound, Horner from the top coefficient.
double AndreaSpline::eval(const SplineTable& t, double x) {
    auto it = std::lower_bound(t.knots.begin(), t.knots.end(), x);
    long pos = static_cast<long>(it - t.knots.begin()) - 1;
    const long last = static_cast<long>(t.knots.size()) - 2;
    if (pos < 0) pos = 0;   // the reference leaves x <= first knot undefined
    if (pos > last) pos = last;
    const size_t p6 = 6 * static_cast<size_t>(pos);
    const double dz = x - t.knots[static_cast<size_t>(pos)];
    double sum = 0;
    for (size_t k = 5; k > 0; k--) sum = dz * (sum + t.coeff[p6 + k]);
    return sum + t.coeff[p6];
}

// ---------------------------------------------------------------- pair terms
double PairTerm::short_range(double q) const {
    switch (scheme) {
    case Scheme::Plain: return 1.0;
    case Scheme::Ewald: return std::erfc(alpha * rc * q);
    case Scheme::Fanourgakis: {
        const double q2 = q * q, q4 = q2 * q2, q5 = q4 * q, q6 = q5 * q, q7 = q6 * q;
        return 1.0 - 1.75 * q + 5.25 * q5 - 7.0 * q6 + 2.5 * q7;
    }
    }
    return 1.0;
}

double PairTerm::operator()(const Particle& a, const Particle& b, double r2) const {
    const size_t ij = static_cast<size_t>(a.id) * ntypes + b.id;
    switch (kind) {
    case PotKind::LennardJones: { // potentials.h:42-49
        double x = sigma2[ij] / r2;
        x = x * x * x;
        return eps4[ij] * (x * x - x);
    }
    case PotKind::WCA: { // potentials.h:151-160
        double x = sigma2[ij];
        if (r2 > x * 1.2599210498948732) return 0;
        x = x / r2;
        x = x * x * x;
        return eps4[ij] * (x * x - x + 0.25);
    }
    case PotKind::HardSphere: // potentials.h:203-207
        return r2 < sigma2[ij] ? infty : 0.0;
    case PotKind::CoulombPlainPM: // potentials.h:472-476
        return lB * a.charge * b.charge / std::sqrt(r2);
    case PotKind::CoulombGalore: { // potentials.h:591-598 (+ CoulombGalore ion_ion_energy)
        const double r = std::sqrt(r2) + std::numeric_limits<double>::epsilon();
        if (scheme == Scheme::Plain) return lB * (a.charge * b.charge / r);
        if (r < rc) {
            const double q = r * rc_inv;
            const double s = splined ? AndreaSpline::eval(S, q) : short_range(q);
            return lB * (a.charge * b.charge / r * s);
        }
        return 0.0;
    }
    }
    return 0.0;
}

static double mix(const std::string& rule, bool is_sigma, double a, double b) {
    // potentials.cpp:18-47 / potentials_base.h:104-115
    if (rule == "arithmetic") return 0.5 * (a + b);
    if (rule == "geometric") return std::sqrt(a * b);
    if (rule == "LB" || rule == "lorentz_berthelot") return is_sigma ? 0.5 * (a + b) : std::sqrt(a * b);
    if (rule == "undefined") return std::numeric_limits<double>::quiet_NaN();
    throw std::runtime_error("unknown combination rule " + rule);
}

static std::vector<double> pair_matrix(const std::vector<AtomType>& atoms, const std::string& rule, bool is_sigma,
                                       const std::function<double(const AtomType&)>& get,
                                       const std::function<double(double)>& modifier) {
    // potentials.cpp:49-71 (diagonal takes the type's own value)
    const size_t n = atoms.size();
    std::vector<double> m(n * n);
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j)
            m[i * n + j] = (i == j) ? modifier(get(atoms[i])) : modifier(mix(rule, is_sigma, get(atoms[i]), get(atoms[j])));
    return m;
}

static int atom_index(const std::vector<AtomType>& atoms, const std::string& name) {
    for (size_t i = 0; i < atoms.size(); ++i)
        if (atoms[i].name == name) return static_cast<int>(i);
    throw std::runtime_error("unknown atom '" + name + "'");
}

static void apply_custom_pairs(const Value& j, const std::vector<AtomType>& atoms, std::vector<double>& m,
                               const std::string& key, const std::function<double(double)>& modifier) {
    if (!j.contains("custom")) return;
    const Value& c = j.at("custom");
    std::vector<std::pair<std::string, const Value*>> entries;
    if (c.is_array()) {
        for (auto& e : c.items()) {
            auto& kv = jsonlite::single_item(e);
            entries.emplace_back(kv.first, &kv.second);
        }
    } else {
        for (auto& kv : c.members()) entries.emplace_back(kv.first, &kv.second);
    }
    const size_t n = atoms.size();
    for (auto& [names, val] : entries) {
        auto sp = names.find(' ');
        int a = atom_index(atoms, names.substr(0, sp));
        int b = atom_index(atoms, names.substr(sp + 1));
        if (val->contains(key)) m[a * n + b] = m[b * n + a] = modifier(val->at(key).as_number());
    }
}

PairTerm make_coulomb_term(const Value& j, const Units& u) {
    PairTerm t;
    t.kind = PotKind::CoulombGalore;
    t.lB = u.bjerrum_length(j.at("epsr").as_number());
    const std::string type = j.at("type").as_string();
    const double utol = j.value("utol", 0.005 / t.lB); // potentials.cpp:1632
    t.splined = j.value("splined", true);
    if (type == "plain") {
        if (j.contains("cutoff")) throw std::runtime_error("unexpected cutoff for plain: it's *always* infinity");
        t.scheme = Scheme::Plain;
        t.rc = infty;
        t.self_prefactor = 0.0;
        return t;
    }
    if (type == "ewald") {
        t.scheme = Scheme::Ewald;
        t.alpha = j.at("alpha").as_number();
        if (j.value("kappa", 0.0) != 0.0 || j.contains("debyelength") || j.contains("molarity"))
            throw std::runtime_error("screened (Yukawa) Ewald is not supported");
    } else if (type == "fanourgakis") {
        t.scheme = Scheme::Fanourgakis;
    } else {
        throw std::runtime_error("unsupported coulomb type '" + type + "'");
    }
    t.rc = j.at("cutoff").as_number();
    t.rc_inv = 1.0 / t.rc;
    // Self energy per particle = lB * c * q^2 / Rc.  Ewald: c = -alpha Rc/sqrt(pi), which
    // equals PolicyIonIon::selfEnergy's -alpha q^2/sqrt(pi) (energy.cpp:510-517, pinned
    // by the doctest at 275-279).  Fanourgakis: c = -1, inferred from the golden
    // examples/bulk initial energy (0.1 % agreement); parity unpinned.
    const double c_self = (t.scheme == Scheme::Ewald) ? -t.alpha * t.rc / std::sqrt(pi) : -1.0;
    t.self_prefactor = t.lB * c_self / t.rc;
    if (t.splined) {
        AndreaSpline sp;
        sp.set_tolerance(utol);
        PairTerm exact = t;
        exact.splined = false;
        t.S = sp.generate([&](double q) { return exact.short_range(q); }, 0.0, 1.0);
    }
    return t;
}

static PairTerm make_mixed_term(PotKind kind, const Value& j, const std::vector<AtomType>& atoms, const Units& u) {
    PairTerm t;
    t.kind = kind;
    t.ntypes = static_cast<int>(atoms.size());
    std::string rule = j.value("mixing", kind == PotKind::HardSphere ? "arithmetic" : "LB");
    if (rule != "arithmetic" && rule != "geometric" && rule != "LB" && rule != "lorentz_berthelot" && rule != "undefined")
        throw std::runtime_error("unknown combination rule " + rule);
    auto sq = [](double x) { return x * x; };
    auto sigma = [](const AtomType& a) { return a.sigma; };
    t.sigma2 = pair_matrix(atoms, rule, true, sigma, sq);
    apply_custom_pairs(j, atoms, t.sigma2, "sigma", sq);
    if (kind != PotKind::HardSphere) {
        auto four_eps = [&u](double x) { return 4 * x; };
        auto eps = [&u](const AtomType& a) { return u.kJmol(a.eps); };
        t.eps4 = pair_matrix(atoms, rule, false, eps, four_eps);
        apply_custom_pairs(j, atoms, t.eps4, "eps", [&u](double x) { return 4 * u.kJmol(x); });
    }
    // MixerPairPotentialBase::from_json (potentials.cpp:285-310) reads no cutoff: a file
    // with one would run untruncated LJ in the reference, so it is refused here
    if (kind == PotKind::LennardJones && j.contains("cutoff"))
        throw std::runtime_error("lennardjones has no cutoff in the reference; use nonbonded_splined with rmax");
    return t;
}

// Parse one potential record {"name": {...}} (potentials.cpp:1202-1310).
static PairTerm make_term(const std::string& name, const Value& cfg, const std::vector<AtomType>& atoms,
                          const Units& u) {
    PairTerm t;
    if (name == "lennardjones") t = make_mixed_term(PotKind::LennardJones, cfg, atoms, u);
    else if (name == "wca") t = make_mixed_term(PotKind::WCA, cfg, atoms, u);
    else if (name == "hardsphere") t = make_mixed_term(PotKind::HardSphere, cfg, atoms, u);
    else if (name == "coulomb") t = make_coulomb_term(cfg, u);
    else throw std::runtime_error("potential '" + name + "': unknown or unsupported potential");
    t.ntypes = static_cast<int>(atoms.size());
    return t;
}

// ---------------------------------------------------------------- SplinedPotential
void build_splined_tables(PairPotential& pp, const Value& js, const std::vector<AtomType>& atoms) {
    AndreaSpline spline;
    spline.set_tolerance(js.value("utol", 1e-3), js.value("ftol", 1e-2));
    const bool hardsphere = js.value("hardsphere", false);
    const double energy_at_rmin = js.value("u_at_rmin", 20.0);
    const double energy_at_rmax = js.value("u_at_rmax", 1e-6);
    const double dr = 1e-2;
    const long max_iterations = 1000000;
    const int nt = pp.ntypes;
    pp.splined = true;
    pp.tables.assign(static_cast<size_t>(nt * nt), SplinedPair{});
    for (int i = 0; i < nt; ++i) {
        for (int j = 0; j <= i; ++j) {
            if (atoms[static_cast<size_t>(i)].implicit || atoms[static_cast<size_t>(j)].implicit) continue;
            // the pair's particles carry the atom types' charges (Particle(AtomData))
            Particle a, b;
            a.id = i;
            a.charge = atoms[static_cast<size_t>(i)].charge;
            b.id = j;
            b.charge = atoms[static_cast<size_t>(j)].charge;
            auto u = [&](double r2) { return pp.exact(a, b, r2); };
            double rmin = 0.5 * (atoms[static_cast<size_t>(i)].sigma + atoms[static_cast<size_t>(j)].sigma);
            double rmax = js.value("rmax", rmin * 10);
            if (js.contains("cutoff_g2g")) {
                const Value& c = js.at("cutoff_g2g");
                if (c.is_number()) rmax = c.as_number();
                else if (c.is_object()) rmax = c.at("default").as_number();
            }
            long iterations = 0;
            while (rmin >= dr) { // findLowerDistance
                if (iterations++ == max_iterations)
                    throw std::runtime_error("Pair potential spline error: cannot determine minimum distance");
                const double e = std::fabs(u(rmin * rmin));
                if (e > energy_at_rmin * 1.1) rmin += dr;
                else if (e < energy_at_rmin / 1.1) rmin -= dr;
                else break;
            }
            iterations = 0;
            while (rmax >= dr) { // findUpperDistance
                if (iterations++ == max_iterations)
                    throw std::runtime_error("Pair potential spline error: cannot determine maximum distance");
                if (std::fabs(u(rmax * rmax)) > energy_at_rmax) rmax += dr;
                else break;
            }
            SplinedPair s; // createKnots
            s.table = spline.generate(u, rmin * rmin, rmax * rmax);
            s.rmin2 = s.table.xmin;
            s.rmax2 = s.table.xmax;
            s.hardsphere = hardsphere;
            if (AndreaSpline::eval(s.table, s.rmin2 + dr) < 0) s.hardsphere = false; // attractive below rmin
            pp.tables[static_cast<size_t>(i * nt + j)] = s;
            pp.tables[static_cast<size_t>(j * nt + i)] = s;
        }
    }
}

// ---------------------------------------------------------------- Ewald
// energy.cpp:28-59
EwaldTerm::EwaldTerm(const Value& j, const Units& u) {
    name = "ewald";
    data.alpha = j.at("alpha").as_number();
    data.r_cutoff = j.at("cutoff").as_number();
    data.spherical_sum = j.value("spherical_sum", true);
    data.lB = u.bjerrum_length(j.at("epsr").as_number());
    data.surface_dielectric_constant = j.value("epss", 0.0);
    data.const_inf = (data.surface_dielectric_constant < 1) ? 0 : 1;
    data.kappa = j.value("kappa", 0.0);
    data.kappa2 = data.kappa * data.kappa;
    data.n_cutoff = j.contains("kcutoff") ? j.at("kcutoff").as_number() : j.at("ncutoff").as_number();
    std::string scheme = j.value("ewaldscheme", "PBC");
    if (j.value("ipbc", false)) scheme = "IPBC";
    if (scheme == "PBC") data.ipbc = false;
    else if (scheme == "IPBC") data.ipbc = true;
    else throw std::runtime_error("ewaldscheme '" + scheme + "' not supported (PBC or IPBC)");
}

void EwaldTerm::update_box(EwaldData& d, const Vec3& box) {
    d.box = box;
    const int nc = static_cast<int>(std::ceil(d.n_cutoff));
    const double maxL = std::max(box.x, std::max(box.y, box.z));
    d.check_k2_zero = 0.1 * std::pow(2 * pi / maxL, 2);
    d.kvec.clear();
    d.Ak.clear();
    const int total = (2 * nc + 1) * (2 * nc + 1) * (2 * nc + 1) - 1;
    if (total == 0) {
        d.kvec.push_back({1, 0, 0});
        d.Ak.push_back(0);
    } else {
        const double nc2 = d.n_cutoff * d.n_cutoff;
        const int start = d.ipbc ? 0 : 1;
        for (int nx = 0; nx <= nc; nx++) {
            const double dnx2 = double(nx * nx);
            const double xfactor = (nx > 0) ? 2.0 : 1.0;
            for (int ny = -nc * start; ny <= nc; ny++) {
                const double dny2 = double(ny * ny);
                const double yfactor = (d.ipbc && ny > 0) ? 2.0 : 1.0;
                for (int nz = -nc * start; nz <= nc; nz++) {
                    double factor = xfactor * yfactor;
                    if (d.ipbc && nz > 0) factor *= 2;
                    Vec3 kv{2 * pi * (double(nx) / box.x), 2 * pi * (double(ny) / box.y), 2 * pi * (double(nz) / box.z)};
                    const double k2 = kv.squared_norm() + d.kappa2;
                    if (k2 < d.check_k2_zero) continue;
                    if (d.spherical_sum) {
                        const double dnz2 = double(nz * nz);
                        if ((dnx2 + dny2 + dnz2) / nc2 > 1) continue;
                    }
                    d.kvec.push_back(kv);
                    d.Ak.push_back(factor * std::exp(-k2 / (4 * d.alpha * d.alpha)) / k2);
                }
            }
        }
    }
    d.Qre.assign(d.kvec.size(), 0.0);
    d.Qim.assign(d.kvec.size(), 0.0);
}

void EwaldTerm::full_update(EwaldData& d, const Space& spc) {
    for (size_t k = 0; k < d.kvec.size(); k++) {
        const Vec3& q = d.kvec[k];
        double re = 0, im = 0;
        for (const auto& g : spc.groups) {
            for (int i = g.begin; i < g.begin + g.size; ++i) {
                const Particle& p = spc.particles[i];
                if (d.ipbc) {
                    re += std::cos(q.x * p.pos.x) * std::cos(q.y * p.pos.y) * std::cos(q.z * p.pos.z) * p.charge;
                } else {
                    const double qr = q.dot(p.pos);
                    re += p.charge * std::cos(qr);
                    im += p.charge * std::sin(qr);
                }
            }
        }
        d.Qre[k] = re;
        d.Qim[k] = im;
    }
}

static std::vector<int> change_indices(const GroupChange& gc, int size_new, int size_old) {
    if (gc.all) {
        std::vector<int> v(static_cast<size_t>(std::max(size_new, size_old)));
        std::iota(v.begin(), v.end(), 0);
        return v;
    }
    return gc.relative_atom_indices;
}

void EwaldTerm::partial_update(EwaldData& d, const Change& c, const Space& spc, const Space& old) {
    for (size_t k = 0; k < d.kvec.size(); k++) {
        const Vec3& q = d.kvec[k];
        double re = d.Qre[k], im = d.Qim[k];
        for (const auto& gc : c.groups) {
            const Group& gn = spc.groups.at(gc.group_index);
            const Group& go = old.groups.at(gc.group_index);
            const auto idx = (d.ipbc) ? gc.relative_atom_indices : change_indices(gc, gn.size, go.size);
            for (int i : idx) {
                if (i < gn.size) {
                    const Particle& p = spc.particles[gn.begin + i];
                    if (d.ipbc) re += std::cos(q.x * p.pos.x) * std::cos(q.y * p.pos.y) * std::cos(q.z * p.pos.z) * p.charge;
                    else {
                        const double qr = q.dot(p.pos);
                        re += p.charge * std::cos(qr);
                        im += p.charge * std::sin(qr);
                    }
                }
                if (i < go.size) {
                    const Particle& p = old.particles[go.begin + i];
                    if (d.ipbc) re -= std::cos(q.x * p.pos.x) * std::cos(q.y * p.pos.y) * std::cos(q.z * p.pos.z) * p.charge;
                    else {
                        const double qr = q.dot(p.pos);
                        re -= p.charge * std::cos(qr);
                        im -= p.charge * std::sin(qr);
                    }
                }
            }
        }
        d.Qre[k] = re;
        d.Qim[k] = im;
    }
}

double EwaldTerm::reciprocal_energy(const EwaldData& d) {
    double e = 0;
    for (size_t k = 0; k < d.Qre.size(); k++) e += d.Ak[k] * (d.Qre[k] * d.Qre[k] + d.Qim[k] * d.Qim[k]);
    return 2 * pi * e * d.lB / (d.box.x * d.box.y * d.box.z);
}

double EwaldTerm::surface_energy(const EwaldData& d, const Change& c, const Space& spc) {
    if (d.const_inf < 0.5 || c.empty()) return 0.0;
    const double volume = d.box.x * d.box.y * d.box.z;
    Vec3 qr{0, 0, 0};
    for (const auto& g : spc.groups)
        for (int i = g.begin; i < g.begin + g.size; ++i) qr = qr + spc.particles[i].pos * spc.particles[i].charge;
    return d.const_inf * 2.0 * pi / ((2.0 * d.surface_dielectric_constant + 1.0) * volume) * qr.squared_norm() * d.lB;
}

double EwaldTerm::self_energy(const EwaldData& d, const Change& c, const Space& spc) {
    double q2 = 0, qt = 0;
    if (c.everything && !c.volume_change) {
        for (const auto& g : spc.groups)
            for (int i = g.begin; i < g.begin + g.size; ++i) {
                q2 += spc.particles[i].charge * spc.particles[i].charge;
                qt += spc.particles[i].charge;
            }
    }
    double Vcc = -pi / 2.0 / d.alpha / d.alpha / (d.box.x * d.box.y * d.box.z) * qt * qt;
    const double beta = d.kappa / (2.0 * d.alpha);
    if (beta > 1e-6) Vcc *= (1.0 - std::exp(-beta * beta)) / beta / beta;
    return (-d.alpha * q2 / std::sqrt(pi) * (std::exp(-beta * beta) + std::sqrt(pi) * beta * std::erf(beta)) + Vcc) * d.lB;
}

void EwaldTerm::init(const Space& spc) { // energy.cpp:554-558
    update_box(data, spc.geometry.len);
    full_update(data, spc);
}

void EwaldTerm::update_state(const Space& spc, const Space* old_spc, const Change& c) { // 564-575
    if (c) {
        if (!c.groups.empty() && old_spc && !c.everything && !c.volume_change) partial_update(data, c, spc, *old_spc);
        else {
            update_box(data, spc.geometry.len);
            full_update(data, spc);
        }
    }
}

double EwaldTerm::energy(const Space& spc, const Space*, const Change& c) { // 581-587
    if (c) return surface_energy(data, c, spc) + reciprocal_energy(data);
    return 0.0;
}

void EwaldTerm::sync_from(const Term& other, const Change& c) { // 640-658
    const auto& o = dynamic_cast<const EwaldTerm&>(other);
    if (c.everything || c.volume_change) data = o.data;
    else {
        data.Qre = o.data.Qre;
        data.Qim = o.data.Qim;
    }
}

// ---------------------------------------------------------------- self energy
// externalpotential.cpp:94-126 with the CoulombGalore monopole self-energy functor.
double SelfEnergyTerm::energy(const Space& spc, const Space*, const Change& c) {
    double e = 0.0;
    if (c.volume_change || c.everything || c.matter_change) {
        for (const auto& g : spc.groups) {
            double ge = 0;
            for (int i = g.begin; i < g.begin + g.size; ++i) ge += prefactor * spc.particles[i].charge * spc.particles[i].charge;
            e += ge;
            if (!std::isfinite(e)) break;
        }
        return e;
    }
    for (const auto& gc : c.groups) {
        const Group& g = spc.groups.at(gc.group_index);
        if (gc.all) {
            double ge = 0;
            for (int i = g.begin; i < g.begin + g.size; ++i) ge += prefactor * spc.particles[i].charge * spc.particles[i].charge;
            e += ge;
        } else {
            for (int i : gc.relative_atom_indices) {
                const double q = spc.particles[g.begin + i].charge;
                e += prefactor * q * q;
            }
        }
        if (!std::isfinite(e)) break;
    }
    return e;
}

// ---------------------------------------------------------------- Nonbonded
// GroupCutoff::cut, energy.h:761-768.
bool NonbondedTerm::cut(const Space& spc, const Group& g1, const Group& g2) const {
    if (g1.atomic || g2.atomic || g2g_cutoff2.empty()) return false;
    const size_t nmol = molecules->size();
    return spc.geometry.sqdist(g1.mass_center, g2.mass_center) >=
           g2g_cutoff2[static_cast<size_t>(g1.molid) * nmol + static_cast<size_t>(g2.molid)];
}

// from_json(GroupCutoff) (energy.cpp:1897-1933): a number is one cutoff for every pair of
// molecule types; an object holds "default" (no cutoff if absent) and "A B" pairs of
// molecule names.
std::vector<double> parse_g2g(const Value& c, const std::vector<MoleculeType>& mols) {
    const size_t nmol = mols.size();
    const double big = std::numeric_limits<double>::max();
    auto squared = [&](double rc) { return rc >= std::sqrt(big) ? big : rc * rc; };
    std::vector<double> t(nmol * nmol, big);
    if (c.is_number()) {
        std::fill(t.begin(), t.end(), squared(c.as_number()));
        return t;
    }
    if (!c.is_object()) throw std::runtime_error("cutoff_g2g: a number or an object");
    if (c.contains("default")) std::fill(t.begin(), t.end(), squared(c.at("default").as_number()));
    auto find = [&](const std::string& name) {
        for (size_t m = 0; m < nmol; ++m)
            if (mols[m].name == name) return m;
        throw std::runtime_error("cutoff_g2g: unknown molecule '" + name + "'");
    };
    for (auto& [key, v] : c.members()) {
        if (key == "default") continue;
        const auto sp = key.find(' ');
        if (sp == std::string::npos || key.find(' ', sp + 1) != std::string::npos)
            throw std::runtime_error("cutoff_g2g: invalid molecules names '" + key + "'");
        const size_t a = find(key.substr(0, sp)), b = find(key.substr(sp + 1));
        t[a * nmol + b] = t[b * nmol + a] = squared(v.as_number());
    }
    return t;
}

double NonbondedTerm::energy(const Space& spc, const Space*, const Change& change) {
    // InstantEnergyAccumulator: serial sum in enumeration order (energy.h:557-597).
    double acc = 0.0;
    auto p2p = [&](const Particle& a, const Particle& b) { acc += pair(spc, a, b); };
    const auto& groups = spc.groups;
    auto P = [&](const Group& g, int i) -> const Particle& { return spc.particles[g.begin + i]; };
    auto excluded = [&](const Group& g, int i, int j) {
        return !g.atomic && (*molecules)[g.molid].is_pair_excluded(i, j);
    };
    auto is_rigid = [&](const Group& g) { return (*molecules)[g.molid].rigid; };

    auto group2group = [&](const Group& g1, const Group& g2) { // 972-982
        if (!cut(spc, g1, g2))
            for (int i = 0; i < g1.size; ++i)
                for (int j = 0; j < g2.size; ++j) p2p(P(g1, i), P(g2, j));
    };
    auto group2group_idx = [&](const Group& g1, const Group& g2, const std::vector<int>& idx1) { // 1015-1027
        if (!cut(spc, g1, g2))
            for (int i : idx1)
                for (int j = 0; j < g2.size; ++j) p2p(P(g1, i), P(g2, j));
    };
    auto group_internal_all = [&](const Group& g) { // 858-873
        if (!is_rigid(g))
            for (int i = 0; i < g.size - 1; ++i)
                for (int j = i + 1; j < g.size; ++j)
                    if (g.atomic || !excluded(g, i, j)) p2p(P(g, i), P(g, j));
    };
    auto group_internal_one = [&](const Group& g, int index) { // 885-914
        if (is_rigid(g)) return;
        for (int i = 0; i < index; ++i)
            if (g.atomic || !excluded(g, index, i)) p2p(P(g, index), P(g, i));
        for (int i = index + 1; i < g.size; ++i)
            if (g.atomic || !excluded(g, index, i)) p2p(P(g, index), P(g, i));
    };
    auto group_internal_idx = [&](const Group& g, const std::vector<int>& index) { // 931-961
        if (is_rigid(g)) return;
        if (index.size() == 1) return group_internal_one(g, index[0]);
        std::vector<char> in(static_cast<size_t>(g.size), 0);
        for (int i : index) in[static_cast<size_t>(i)] = 1;
        std::vector<int> complement;
        for (int i = 0; i < g.size; ++i)
            if (!in[static_cast<size_t>(i)]) complement.push_back(i);
        for (int i : index)
            for (int j : complement)
                if (!excluded(g, i, j)) p2p(P(g, i), P(g, j));
        for (size_t a = 0; a < index.size(); ++a)
            for (size_t b = a + 1; b < index.size(); ++b)
                if (!excluded(g, index[a], index[b])) p2p(P(g, index[a]), P(g, index[b]));
    };
    auto group2all_one = [&](const Group& g, int index) { // 1182-1195
        const Particle& p = P(g, index);
        for (const auto& other : groups) {
            if (&other != &g && !cut(spc, other, g))
                for (int j = 0; j < other.size; ++j) p2p(p, P(other, j));
        }
    };
    auto group2all_idx = [&](const Group& g, const std::vector<int>& index) { // 1211-1224
        if (index.size() == 1) return group2all_one(g, index[0]);
        for (const auto& other : groups)
            if (&other != &g) group2group_idx(g, other, index);
    };
    auto group2all = [&](const Group& g) { // 1155-1163
        for (const auto& other : groups)
            if (&other != &g) group2group(g, other);
    };

    if (change.everything) { // all(), 1290-1299
        for (size_t a = 0; a < groups.size(); ++a) {
            group_internal_all(groups[a]);
            for (size_t b = a + 1; b < groups.size(); ++b) group2group(groups[a], groups[b]);
        }
    } else if (change.volume_change) { // 1317-1326
        for (size_t a = 0; a < groups.size(); ++a) {
            const Group& g = groups[a];
            if (g.atomic || (*molecules)[g.molid].compressible) group_internal_all(g);
            for (size_t b = a + 1; b < groups.size(); ++b) group2group(groups[a], groups[b]);
        }
    } else if (!change.matter_change) {
        if (change.groups.size() == 1) { // accumulateGroup, 1347-1377
            const auto& cd = change.groups[0];
            const Group& g = groups.at(cd.group_index);
            if (cd.relative_atom_indices.size() == 1) {
                group2all_one(g, cd.relative_atom_indices[0]);
                if (cd.internal) group_internal_one(g, cd.relative_atom_indices[0]);
            } else if (cd.relative_atom_indices.empty()) {
                group2all(g);
                if (cd.internal) group_internal_all(g);
            } else {
                group2all_idx(g, cd.relative_atom_indices);
                if (cd.internal) group_internal_idx(g, cd.relative_atom_indices);
            }
        } else { // groups2all, 1270-1279 (+ groups2self 1247-1261)
            std::vector<int> moved;
            for (auto& gc : change.groups) moved.push_back(gc.group_index);
            for (size_t a = 0; a < moved.size(); ++a)
                for (size_t b = a + 1; b < moved.size(); ++b) group2group(groups[moved[a]], groups[moved[b]]);
            std::vector<char> is_moved(groups.size(), 0);
            for (int m : moved) is_moved[static_cast<size_t>(m)] = 1;
            for (int m : moved)
                for (size_t o = 0; o < groups.size(); ++o)
                    if (!is_moved[o]) group2group(groups[m], groups[o]);
        }
    } else {
        throw std::runtime_error("matter change (speciation) is not supported by this path");
    }
    return acc;
}

// ---------------------------------------------------------------- Hamiltonian
double Hamiltonian::energy(const Space& spc, const Space* old_spc, const Change& c) { // energy.cpp:1227-1241
    latest.clear();
    for (auto& t : terms) {
        const double e = t->energy(spc, old_spc, c);
        latest.push_back(e);
        if (e >= maximum_allowed_energy || std::isnan(e)) break;
    }
    return std::accumulate(latest.begin(), latest.end(), 0.0);
}

// ---------------------------------------------------------------- System
System::System(const std::string& config_json) {
    Value cfg = jsonlite::parse(config_json);
    units.temperature = cfg.value("temperature", 298.15);
    for (auto& rec : cfg.at("atomlist").items()) { // atomdata.cpp:140-190
        auto& kv = jsonlite::single_item(rec);
        AtomType a;
        a.name = kv.first;
        const Value& v = kv.second;
        a.charge = v.value("q", 0.0);
        a.sigma = v.value("sigma", 0.0);
        if (std::fabs(a.sigma) < 1e-20) a.sigma = 2.0 * v.value("r", 0.0);
        a.eps = v.value("eps", 0.0);
        a.has_dp = v.contains("dp");
        a.dp = v.value("dp", 0.0);
        a.dprot = v.value("dprot", 0.0);
        a.mw = v.value("mw", 1.0);
        a.implicit = v.value("implicit", false);
        atoms.push_back(a);
    }
    if (cfg.contains("moleculelist")) {
        for (auto& rec : cfg.at("moleculelist").items()) {
            auto& kv = jsonlite::single_item(rec);
            MoleculeType m;
            m.name = kv.first;
            m.atomic = kv.second.value("atomic", false);
            m.rigid = kv.second.value("rigid", false);
            m.compressible = kv.second.value("compressible", false);
            if (kv.second.contains("atoms"))
                for (auto& a : kv.second.at("atoms").items()) m.atoms.push_back(atom_index(atoms, a.as_string()));
            if (kv.second.contains("bondlist")) // bonds.cpp:23-60
                for (auto& rec : kv.second.at("bondlist").items()) {
                    auto& [btype, bp] = jsonlite::single_item(rec);
                    if (btype != "harmonic") throw std::runtime_error("bond type '" + btype + "' is not supported");
                    HarmonicBond b;
                    const Value& idx = bp.at("index");
                    if (idx.size() != 2) throw std::runtime_error("harmonic bond needs exactly 2 indices");
                    b.i = static_cast<int>(idx.at(0).as_number());
                    b.j = static_cast<int>(idx.at(1).as_number());
                    b.half_k = units.kJmol(bp.at("k").as_number()) / 2.0; // bonds.cpp:102-104
                    b.req = bp.at("req").as_number();
                    m.bonds.push_back(b);
                }
            // excluded_neighbours: atom pairs joined by a path of at most d bonds
            // (NeighboursGenerator, molecule.cpp:280-294); exclusionlist: explicit pairs.
            const int d = static_cast<int>(kv.second.value("excluded_neighbours", 0.0));
            if (d > 0) {
                int size = 0;
                for (auto& b : m.bonds) size = std::max(size, std::max(b.i, b.j) + 1);
                std::vector<std::vector<int>> adj(static_cast<size_t>(size));
                for (auto& b : m.bonds) {
                    adj[b.i].push_back(b.j);
                    adj[b.j].push_back(b.i);
                }
                for (int s0 = 0; s0 < size; ++s0) {
                    std::vector<int> dist(static_cast<size_t>(size), -1);
                    std::vector<int> frontier{s0};
                    dist[s0] = 0;
                    for (int step = 1; step <= d && !frontier.empty(); ++step) {
                        std::vector<int> next;
                        for (int a : frontier)
                            for (int b : adj[a])
                                if (dist[b] < 0) {
                                    dist[b] = step;
                                    next.push_back(b);
                                }
                        frontier = next;
                    }
                    for (int t = s0 + 1; t < size; ++t)
                        if (dist[t] > 0) m.excluded.emplace_back(s0, t);
                }
            }
            if (kv.second.contains("exclusionlist"))
                for (auto& p : kv.second.at("exclusionlist").items()) {
                    int i = static_cast<int>(p.at(0).as_number()), j = static_cast<int>(p.at(1).as_number());
                    m.excluded.emplace_back(std::min(i, j), std::max(i, j));
                }
            std::sort(m.excluded.begin(), m.excluded.end());
            m.excluded.erase(std::unique(m.excluded.begin(), m.excluded.end()), m.excluded.end());
            molecules.push_back(m);
        }
    } else {
        MoleculeType m;
        m.name = "atomic";
        m.atomic = true;
        molecules.push_back(m);
    }
    const Value& geo = cfg.at("geometry");
    if (geo.value("type", "cuboid") != "cuboid") throw std::runtime_error("only cuboid geometry is supported");
    const Value& L = geo.at("length");
    Vec3 len = L.is_array() ? Vec3{L.at(0).as_number(), L.at(1).as_number(), L.at(2).as_number()}
                            : Vec3{L.as_number(), L.as_number(), L.as_number()};
    accepted_space.geometry.set(len);
    trial_space.geometry.set(len);
    energy_json = cfg.at("energy");
    accepted_pot = build_hamiltonian();
    trial_pot = build_hamiltonian();
}

std::unique_ptr<Hamiltonian> System::build_hamiltonian() {
    auto h = std::make_unique<Hamiltonian>();
    h->units = units;
    h->atoms = atoms;
    h->molecules = molecules;
    const int nt = static_cast<int>(atoms.size());
    for (auto& rec : energy_json.items()) { // energy.cpp:1171-1211
        auto& [key, value] = jsonlite::single_item(rec);
        if (key == "maxenergy") {
            h->maximum_allowed_energy = value.as_number();
            continue;
        }
        if (key == "bonded") { // Hamiltonian::createEnergy (energy.cpp:1328), intra-molecular bonds only
            if (value.contains("bondlist") && value.at("bondlist").size() > 0)
                throw std::runtime_error("inter-molecular bonds are not supported by this path");
            auto b = std::make_unique<BondedTerm>();
            b->name = "bonded";
            b->molecules = &h->molecules;
            h->terms.push_back(std::move(b));
            continue;
        }
        auto nb = std::make_unique<NonbondedTerm>();
        nb->name = "nonbonded";
        nb->molecules = &h->molecules;
        PairPotential& pp = nb->pot;
        pp.ntypes = nt;
        pp.umatrix.assign(static_cast<size_t>(nt * nt), {});
        const Value* coulomb_json = nullptr;
        if (key == "nonbonded" || key == "nonbonded_exact" || key == "nonbonded_splined") {
            // FunctorPotential (aniso form); SplinedPotential takes sqdist (energy.cpp:1301-1305)
            pp.anisotropic_form = key != "nonbonded_splined";
            auto add_list = [&](const Value& list) {
                std::vector<int> idx;
                for (auto& r : list.items()) {
                    if (!r.is_object() || r.size() != 1) continue;
                    auto& [pname, pcfg] = jsonlite::single_item(r);
                    pp.terms.push_back(make_term(pname, pcfg, atoms, units));
                    if (pname == "coulomb" && !pp.has_self_energy) {
                        pp.has_self_energy = true;
                        pp.self_prefactor = pp.terms.back().self_prefactor;
                    }
                    idx.push_back(static_cast<int>(pp.terms.size()) - 1);
                }
                return idx;
            };
            auto def = add_list(value.at("default"));
            for (auto& cell : pp.umatrix) cell = def;
            for (auto& [pk, pv] : value.members()) {
                auto sp = pk.find(' ');
                if (sp == std::string::npos || pk == "default") continue;
                int a = atom_index(atoms, pk.substr(0, sp)), b = atom_index(atoms, pk.substr(sp + 1));
                auto lst = add_list(pv);
                pp.umatrix[a * nt + b] = lst;
                pp.umatrix[b * nt + a] = lst;
            }
            for (auto& r : value.at("default").items())
                if (r.is_object() && r.size() == 1 && r.contains("coulomb")) { coulomb_json = &r.at("coulomb"); break; }
            if (key == "nonbonded_splined") build_splined_tables(pp, value, atoms);
        } else if (key == "nonbonded_coulomblj" || key == "nonbonded_coulombwca" || key == "nonbonded_pm" ||
                   key == "nonbonded_coulombhs" || key == "nonbonded_pmwca") { // CombinedPairPotential
            pp.anisotropic_form = false;
            const bool galore = (key == "nonbonded_coulomblj" || key == "nonbonded_coulombwca");
            PairTerm first;
            if (galore) first = make_coulomb_term(value.at("coulomb"), units);
            else {
                first.kind = PotKind::CoulombPlainPM;
                first.lB = units.bjerrum_length(value.at("coulomb").at("epsr").as_number());
            }
            first.ntypes = nt;
            PairTerm second;
            if (key == "nonbonded_coulomblj") second = make_term("lennardjones", value.at("lennardjones"), atoms, units);
            else if (key == "nonbonded_coulombwca" || key == "nonbonded_pmwca") second = make_term("wca", value.at("wca"), atoms, units);
            else second = make_term("hardsphere", value.contains("hardsphere") ? value.at("hardsphere") : Value::object(), atoms, units);
            pp.terms = {first, second};
            for (auto& cell : pp.umatrix) cell = {0, 1};
            if (galore) {
                pp.has_self_energy = true;
                pp.self_prefactor = first.self_prefactor;
                coulomb_json = &value.at("coulomb");
            }
        } else {
            throw std::runtime_error("energy term '" + key + "' is not supported by this path");
        }
        if (value.contains("cutoff_g2g")) nb->g2g_cutoff2 = parse_g2g(value.at("cutoff_g2g"), h->molecules);
        if (pp.has_self_energy) { // PairEnergy::addPairPotentialSelfEnergy, energy.h:462-468
            auto self = std::make_unique<SelfEnergyTerm>();
            self->name = "selfenergy";
            self->prefactor = pp.self_prefactor;
            h->terms.push_back(std::move(self));
        }
        h->terms.push_back(std::move(nb));
        if (coulomb_json && coulomb_json->value("type", "") == "ewald") // addEwald, energy.cpp:1134-1160
            h->terms.push_back(std::make_unique<EwaldTerm>(*coulomb_json, units));
    }
    return h;
}

std::vector<std::string> System::term_names() const {
    std::vector<std::string> v;
    for (auto& t : accepted_pot->terms) v.push_back(t->name);
    return v;
}

void System::set_particles(int n, const double* xyz, const double* charge, const int* atom_id) {
    for (Space* s : {&accepted_space, &trial_space}) {
        s->particles.resize(static_cast<size_t>(n));
        for (int i = 0; i < n; ++i) {
            s->particles[i].pos = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
            s->particles[i].charge = charge[i];
            s->particles[i].id = atom_id[i];
        }
    }
}

void System::set_groups(int ngroups, const int* begin, const int* size, const int* molid, const double* mass_center) {
    for (Space* s : {&accepted_space, &trial_space}) {
        s->groups.resize(static_cast<size_t>(ngroups));
        for (int g = 0; g < ngroups; ++g) {
            Group& G = s->groups[g];
            G.begin = begin[g];
            G.size = size[g];
            G.molid = molid[g];
            const MoleculeType& m = molecules.at(static_cast<size_t>(molid[g]));
            G.atomic = m.atomic;
            G.rigid = m.rigid;
            G.compressible = m.compressible;
            if (mass_center) G.mass_center = {mass_center[3 * g], mass_center[3 * g + 1], mass_center[3 * g + 2]};
            else if (!G.atomic && G.size > 0) // PBC-safe centre around the first atom (Group::updateMassCenter)
                G.mass_center = oracle::mass_center(*s, G, atoms, true, s->particles[static_cast<size_t>(G.begin)].pos * -1.0);
        }
    }
}

void System::init() { // montecarlo.cpp:43-71
    Change everything;
    everything.everything = true;
    accepted_pot->init(accepted_space);
    initial_energy = accepted_pot->energy(accepted_space, nullptr, everything);
    trial_space = accepted_space;
    trial_pot->init(trial_space);
    const double trial_energy = trial_pot->energy(trial_space, nullptr, everything);
    if (std::isfinite(initial_energy) && std::isfinite(trial_energy) &&
        std::fabs((initial_energy - trial_energy) / initial_energy) > 1e-6)
        throw std::runtime_error("error aligning energies");
    stats = MCStats{};
}

double System::total_energy(bool trial, std::vector<double>* terms_out) {
    Change everything;
    everything.everything = true;
    Hamiltonian& h = pot(trial);
    const double e = h.energy(trial ? trial_space : accepted_space, nullptr, everything);
    if (terms_out) {
        *terms_out = h.latest;
        terms_out->resize(h.terms.size(), 0.0);
    }
    return e;
}

void System::energy_both(const Change& c, std::vector<double>& e_new, std::vector<double>& e_old) {
    trial_pot->update_state(trial_space, &accepted_space, c);
    trial_pot->energy(trial_space, &accepted_space, c);
    e_new = trial_pot->latest;
    e_new.resize(trial_pot->terms.size(), 0.0);
    accepted_pot->energy(accepted_space, nullptr, c);
    e_old = accepted_pot->latest;
    e_old.resize(accepted_pot->terms.size(), 0.0);
}

static double energy_change(double new_energy, double old_energy) { // montecarlo.cpp:193-209
    if (std::isnan(old_energy) && !std::isnan(new_energy)) return -infty;
    if (std::isnan(new_energy)) return infty;
    if (new_energy > 0.0 && std::isinf(new_energy)) return infty;
    const double du = new_energy - old_energy;
    if (std::isnan(du)) return 0.0;
    return du;
}

static bool metropolis(double du, double u) { // montecarlo.cpp:17-34 (u drawn by the caller)
    if (std::isnan(du)) throw std::runtime_error("Metropolis error: energy cannot be NaN");
    if (std::isinf(du) && du < 0.0) return true;
    if (-du > max_exp_argument) return true;
    return u <= std::exp(-du);
}

// performMove (montecarlo.cpp:139-187) for a Change whose new coordinates have
// already been written into the trial space.
double System::metropolis_move(const Change& c, double uniform, bool* accepted_out, double* du_out,
                               std::vector<double>* e_new, std::vector<double>* e_old) {
    trial_pot->update_state(trial_space, &accepted_space, c);
    const double new_energy = trial_pot->energy(trial_space, &accepted_space, c);
    if (e_new) { *e_new = trial_pot->latest; e_new->resize(trial_pot->terms.size(), 0.0); }
    const double old_energy = accepted_pot->energy(accepted_space, nullptr, c);
    if (e_old) { *e_old = accepted_pot->latest; e_old->resize(accepted_pot->terms.size(), 0.0); }
    double du = energy_change(new_energy, old_energy);
    const bool acc = metropolis(du, uniform);
    auto sync_space = [&](Space& dst, const Space& src) { // space.cpp:199-239 (changed groups only)
        for (auto& gc : c.groups) {
            const Group& gs = src.groups[gc.group_index];
            Group& gd = dst.groups[gc.group_index];
            if (gc.all || gc.relative_atom_indices.empty()) {
                for (int i = 0; i < gs.size; ++i) dst.particles[gs.begin + i] = src.particles[gs.begin + i];
            } else {
                for (int i : gc.relative_atom_indices) dst.particles[gs.begin + i] = src.particles[gs.begin + i];
            }
            gd.mass_center = gs.mass_center;
            gd.size = gs.size;
        }
    };
    if (acc) {
        sync_space(accepted_space, trial_space);
        accepted_pot->sync_from(*trial_pot, c);
        stats.accepted++;
    } else {
        sync_space(trial_space, accepted_space);
        trial_pot->sync_from(*accepted_pot, c);
        du = 0.0;
    }
    stats.moves++;
    stats.sum_du += du;
    if (accepted_out) *accepted_out = acc;
    if (du_out) *du_out = acc ? du : energy_change(new_energy, old_energy);
    return du;
}

// ---------------------------------------------------------------- Bonded
double BondedTerm::energy(const Space& spc, const Space*, const Change& change) { // energy.cpp:1047-1088
    double energy = 0.0;
    if (!change) return energy;
    auto group_bonds = [&](const Group& g) -> const std::vector<HarmonicBond>& {
        return (*molecules)[static_cast<size_t>(g.molid)].bonds;
    };
    if (change.everything || change.volume_change) {
        for (const Group& g : spc.groups) {
            if (g.size == 0) continue;
            double e = 0.0;
            for (const auto& b : group_bonds(g)) e += bond(spc, g, b);
            energy += e;
        }
        return energy;
    }
    for (const auto& gc : change.groups) { // internalGroupEnergy, 1068-1088
        const Group& g = spc.groups.at(static_cast<size_t>(gc.group_index));
        if (!gc.internal || g.size == 0) continue;
        double e = 0.0;
        if (gc.all) {
            for (const auto& b : group_bonds(g)) e += bond(spc, g, b);
        } else { // sumEnergy (energy.h:360-376): bonds with any index among the moved atoms
            const auto& idx = gc.relative_atom_indices;
            for (const auto& b : group_bonds(g))
                if (std::binary_search(idx.begin(), idx.end(), b.i) || std::binary_search(idx.begin(), idx.end(), b.j))
                    e += bond(spc, g, b);
        }
        energy += e;
    }
    return energy;
}

Vec3 mass_center(const Space& spc, const Group& g, const std::vector<AtomType>& atoms, bool pbc, const Vec3& shift) {
    double wsum = 0.0;
    Vec3 c{0, 0, 0};
    for (int k = 0; k < g.size; ++k) {
        const Particle& p = spc.particles[static_cast<size_t>(g.begin + k)];
        Vec3 t = p.pos + shift;
        if (pbc) spc.geometry.boundary(t);
        const double w = atoms.at(static_cast<size_t>(p.id)).mw;
        c = c + t * w;
        wsum += w;
    }
    if (std::fabs(wsum) <= std::numeric_limits<double>::epsilon()) return {0, 0, 0};
    c = Vec3{c.x / wsum, c.y / wsum, c.z / wsum} - shift;
    if (pbc) spc.geometry.boundary(c);
    return c;
}

std::array<double, 9> rotation_matrix(const Vec3& u, double angle) {
    const double s = std::sin(angle), c = std::cos(angle);
    const Vec3 sa = u * s, ca = u * (1.0 - c);
    return {ca.x * u.x + c, ca.x * u.y - sa.z, ca.x * u.z + sa.y,
            ca.x * u.y + sa.z, ca.y * u.y + c, ca.y * u.z - sa.x,
            ca.x * u.z - sa.y, ca.y * u.z + sa.x, ca.z * u.z + c};
}

// ---------------------------------------------------------------- Philox
std::array<uint32_t, 4> philox4x32_10(std::array<uint32_t, 4> c, std::array<uint32_t, 2> k) {
    constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = static_cast<uint64_t>(M0) * c[0];
        const uint64_t p1 = static_cast<uint64_t>(M1) * c[2];
        const uint32_t hi0 = static_cast<uint32_t>(p0 >> 32), lo0 = static_cast<uint32_t>(p0);
        const uint32_t hi1 = static_cast<uint32_t>(p1 >> 32), lo1 = static_cast<uint32_t>(p1);
        c = {hi1 ^ c[1] ^ k[0], lo1, hi0 ^ c[3] ^ k[1], lo0};
        k[0] += W0;
        k[1] += W1;
    }
    return c;
}

} // namespace oracle
