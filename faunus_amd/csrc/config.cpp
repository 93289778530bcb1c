// Parsing of the reference's input JSON into the flat model the GPU engine uses.
#include "config.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <stdexcept>

namespace fmc {

using jsonlite::Value;

// ------------------------------------------------------------------ spline
// Knot placement follows tabulate.h:223-307: intervals are grown downwards from the
// upper end in sqrt(x); each is the quintic Hermite interpolant of f, f', f'' (central
// differences, step 1e-4) and is accepted when 11 checkpoints are within utol.
Spline Spline::build(const std::function<double(double)>& f, double xmin, double xmax, double utol, double ftol) {
    const double h = 1e-4;
    auto df = [&](double x) { return (f(x + 0.5 * h) - f(x - 0.5 * h)) / h; };
    auto d2f = [&](double x) { return (df(x + 0.5 * h) - df(x - 0.5 * h)) / h; };
    const double r_min = std::sqrt(xmin), r_max = std::sqrt(xmax);
    double r_hi = r_max;
    std::vector<double> xs{r_max * r_max};
    std::vector<std::array<double, 6>> cs;
    for (int iter = 0; iter < 1200; ++iter) {
        double width = r_hi - r_min;
        double r_lo = r_hi;
        std::array<double, 6> c{};
        bool accepted = false;
        for (int shrink = 0; shrink < 100; ++shrink) {
            const double x_hi = r_hi * r_hi;
            r_lo = std::max(r_hi - width, r_min);
            const double x_lo = r_lo * r_lo;
            const double f0 = f(x_lo), f1 = df(x_lo), f2 = d2f(x_lo);
            if (std::fabs(f0) < 1e-9 && std::fabs(f1) < 1e-9) {
                c = {0, 0, 0, 0, 0, 0};
            } else {
                const double g0 = f(x_hi), g1 = df(x_hi), g2 = d2f(x_hi);
                const double d = x_hi - x_lo, dd = d * d, ddd = dd * d;
                const double c2 = 0.5 * f2;
                const double A = 6 * (g0 - f0 - f1 * d - c2 * dd) / ddd;
                const double B = 2 * (g1 - f1 - 2 * c2 * d) / dd;
                const double Cc = (g2 - 2 * c2) / d;
                c = {f0, f1, c2, (10 * A - 12 * B + 3 * Cc) / 6, (-15 * A + 21 * B - 6 * Cc) / (6 * d),
                     (2 * A - 3 * B + Cc) / (2 * dd)};
            }
            bool ok = true;
            const double step = (r_hi - r_lo) / 10.0;
            for (int k = 0; k < 11 && ok; ++k) {
                const double r = r_lo + step * k;
                const double x = r * r;
                const double z = x - r_lo * r_lo;
                const double u = c[0] + z * (c[1] + z * (c[2] + z * (c[3] + z * (c[4] + z * c[5]))));
                if (std::fabs(u - f(x)) > utol) ok = false;
                if (ok && ftol != -1.0) {
                    const double du = c[1] + z * (2 * c[2] + z * (3 * c[3] + z * (4 * c[4] + z * 5 * c[5])));
                    if (std::fabs(du - df(x)) > ftol) ok = false;
                }
            }
            if (ok) {
                accepted = true;
                break;
            }
            width *= 0.9;
        }
        if (!accepted) throw std::runtime_error("spline: tolerance cannot be met");
        xs.push_back(r_lo * r_lo);
        cs.push_back(c);
        r_hi = r_lo;
        if (r_lo <= r_min) break;
    }
    Spline s;
    s.knots.assign(xs.rbegin(), xs.rend());
    for (auto it = cs.rbegin(); it != cs.rend(); ++it) s.coeff.insert(s.coeff.end(), it->begin(), it->end());
    return s;
}

double Spline::eval(double x) const {
    long pos = static_cast<long>(std::lower_bound(knots.begin(), knots.end(), x) - knots.begin()) - 1;
    pos = std::max(0L, std::min(pos, static_cast<long>(knots.size()) - 2));
    const double* c = &coeff[6 * static_cast<size_t>(pos)];
    const double z = x - knots[static_cast<size_t>(pos)];
    double s = 0;
    for (int k = 5; k > 0; --k) s = z * (s + c[k]);
    return s + c[0];
}

double Model::pair_exact(int a, double qa, int b, double qb, double r2) const {
    const PairParam& p = pairs[static_cast<size_t>(a * ntypes + b)];
    double u = 0.0;
    if (p.lj_e4 != 0.0) {
        double x = p.lj_s2 / r2;
        x = x * x * x;
        u = u + p.lj_e4 * (x * x - x);
    }
    if (p.wca_e4 != 0.0 && !(r2 > p.wca_s2 * 1.2599210498948732)) {
        double x = p.wca_s2 / r2;
        x = x * x * x;
        u = u + p.wca_e4 * (x * x - x + 0.25);
    }
    if (r2 < p.hs_s2) u = u + kInf;
    if (p.coulomb != 0.0) {
        const CoulombModel& c = coulomb;
        double e = 0.0;
        if (c.scheme == kCoulPM) e = c.lB * qa * qb / std::sqrt(r2);
        else {
            const double r = std::sqrt(r2) + 2.220446049250313e-16;
            if (c.scheme == kCoulPlain) e = c.lB * (qa * qb / r);
            else if (r < c.rc) {
                const double q = r * c.rc_inv;
                const double s = c.splined ? c.S.eval(q) : c.short_range_exact(q);
                e = c.lB * (qa * qb / r * s);
            }
        }
        u = u + e;
    }
    return u;
}

double Model::pair_energy(int a, double qa, int b, double qb, double r2) const {
    if (!splined) return pair_exact(a, qa, b, qb, r2);
    const SplinedPairCfg& s = spl[static_cast<size_t>(a * ntypes + b)];
    if (r2 >= s.rmax2) return 0.0;
    if (r2 > s.rmin2) return s.table.eval(r2);
    if (s.hardsphere) return kInf;
    return pair_exact(a, qa, b, qb, r2);
}

double CoulombModel::short_range_exact(double q) const {
    if (scheme == kCoulEwald) return std::erfc(alpha * rc * q);
    if (scheme == kCoulFanourgakis) {
        const double q5 = q * q * q * q * q;
        return 1.0 - 1.75 * q + 5.25 * q5 - 7.0 * q5 * q + 2.5 * q5 * q * q;
    }
    return 1.0;
}

// ------------------------------------------------------------------ parsing
namespace {

int type_index(const Model& m, const std::string& name) {
    for (size_t i = 0; i < m.atoms.size(); ++i)
        if (m.atoms[i].name == name) return static_cast<int>(i);
    throw std::runtime_error("unknown atom type '" + name + "'");
}

double combine(const std::string& rule, bool sigma, double a, double b) {
    if (rule == "arithmetic") return 0.5 * (a + b);
    if (rule == "geometric") return std::sqrt(a * b);
    if (rule == "LB" || rule == "lorentz_berthelot") return sigma ? 0.5 * (a + b) : std::sqrt(a * b);
    if (rule == "undefined") return std::numeric_limits<double>::quiet_NaN();
    throw std::runtime_error("unknown combination rule " + rule);
}

// Per-pair sigma^2 and 4*epsilon (in kT) with the potential's mixing rule and its
// optional "custom" pair overrides (potentials.cpp:49-93, 672-689, 956-962).
void mixed_matrices(const Model& m, const Value& j, bool with_eps, const std::string& default_rule,
                    std::vector<double>& s2, std::vector<double>& e4) {
    const std::string rule = j.value("mixing", default_rule.c_str());
    const int n = m.ntypes;
    s2.assign(static_cast<size_t>(n * n), 0.0);
    e4.assign(static_cast<size_t>(n * n), 0.0);
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) {
            const auto& A = m.atoms[static_cast<size_t>(a)];
            const auto& B = m.atoms[static_cast<size_t>(b)];
            const double s = (a == b) ? A.sigma : combine(rule, true, A.sigma, B.sigma);
            s2[static_cast<size_t>(a * n + b)] = s * s;
            if (with_eps) {
                const double ea = m.units.kJmol(A.eps), eb = m.units.kJmol(B.eps);
                e4[static_cast<size_t>(a * n + b)] = 4 * ((a == b) ? ea : combine(rule, false, ea, eb));
            }
        }
    if (!j.contains("custom")) return;
    const Value& c = j.at("custom");
    auto apply = [&](const std::string& names, const Value& v) {
        const auto sp = names.find(' ');
        const int a = type_index(m, names.substr(0, sp)), b = type_index(m, names.substr(sp + 1));
        if (v.contains("sigma")) {
            const double s = v.at("sigma").as_number();
            s2[static_cast<size_t>(a * n + b)] = s2[static_cast<size_t>(b * n + a)] = s * s;
        }
        if (with_eps && v.contains("eps")) {
            const double e = 4 * m.units.kJmol(v.at("eps").as_number());
            e4[static_cast<size_t>(a * n + b)] = e4[static_cast<size_t>(b * n + a)] = e;
        }
    };
    if (c.is_array())
        for (auto& e : c.items()) {
            auto& kv = jsonlite::single_item(e);
            apply(kv.first, kv.second);
        }
    else
        for (auto& kv : c.members()) apply(kv.first, kv.second);
}

CoulombModel parse_coulomb(const Model& m, const Value& j, bool pm) {
    CoulombModel c;
    c.lB = m.units.bjerrum(j.at("epsr").as_number());
    if (pm) { // pairpotential::Coulomb (potentials.cpp:452-460): epsr only
        c.scheme = kCoulPM;
        return c;
    }
    const std::string type = j.at("type").as_string();
    c.splined = j.value("splined", true);
    if (type == "plain") {
        if (j.contains("cutoff")) throw std::runtime_error("unexpected cutoff for plain: it's *always* infinity");
        c.scheme = kCoulPlain;
        return c;
    }
    if (type == "ewald") {
        c.scheme = kCoulEwald;
        c.alpha = j.at("alpha").as_number();
        if (j.value("kappa", 0.0) != 0.0 || j.contains("debyelength") || j.contains("molarity"))
            throw std::runtime_error("screened (Yukawa) Ewald is not supported by the GPU engine");
    } else if (type == "fanourgakis") {
        c.scheme = kCoulFanourgakis;
    } else {
        throw std::runtime_error("coulomb type '" + type + "' is not supported by the GPU engine");
    }
    c.rc = j.at("cutoff").as_number();
    c.rc_inv = 1.0 / c.rc;
    const double cs = (c.scheme == kCoulEwald) ? -c.alpha * c.rc / std::sqrt(kPi) : -1.0;
    c.self_prefactor = c.lB * cs / c.rc;
    if (c.splined) {
        CoulombModel exact = c;
        c.S = Spline::build([exact](double q) { return exact.short_range_exact(q); }, 0.0, 1.0,
                            j.value("utol", 0.005 / c.lB));
    }
    return c;
}

} // namespace

static void build_splined(Model& m, const Value& js);

// cutoff_g2g (from_json(GroupCutoff), energy.cpp:1897-1933): a number is one cutoff for
// every pair of molecule types; an object holds "default" (no cutoff if absent) and
// "A B" entries naming two molecule types.
static std::vector<double> parse_g2g(const Value& c, const std::vector<MoleculeCfg>& mols) {
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
        for (size_t k = 0; k < nmol; ++k)
            if (mols[k].name == name) return k;
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


Model Model::from_json(const std::string& text) {
    Model m;
    const Value cfg = jsonlite::parse(text);
    m.units.temperature = cfg.value("temperature", 298.15);
    for (auto& rec : cfg.at("atomlist").items()) { // atomdata.cpp:140-190
        auto& kv = jsonlite::single_item(rec);
        AtomTypeCfg a;
        a.name = kv.first;
        a.charge = kv.second.value("q", 0.0);
        a.sigma = kv.second.value("sigma", 0.0);
        if (std::fabs(a.sigma) < 1e-20) a.sigma = 2.0 * kv.second.value("r", 0.0);
        a.eps = kv.second.value("eps", 0.0);
        a.has_dp = kv.second.contains("dp");
        a.dp = kv.second.value("dp", 0.0);
        a.dprot = kv.second.value("dprot", 0.0);
        a.mw = kv.second.value("mw", 1.0);
        m.atoms.push_back(a);
    }
    m.ntypes = static_cast<int>(m.atoms.size());
    if (m.ntypes > 255) throw std::runtime_error("at most 255 atom types are supported");
    if (cfg.contains("moleculelist")) {
        for (auto& rec : cfg.at("moleculelist").items()) {
            auto& kv = jsonlite::single_item(rec);
            MoleculeCfg mc;
            mc.name = kv.first;
            mc.atomic = kv.second.value("atomic", false);
            mc.rigid = kv.second.value("rigid", false);
            mc.compressible = kv.second.value("compressible", false);
            if (kv.second.contains("atoms"))
                for (auto& a : kv.second.at("atoms").items()) mc.atoms.push_back(type_index(m, a.as_string()));
            if (kv.second.contains("bondlist")) // bonds.cpp:23-60, molecule.cpp:508-517
                for (auto& brec : kv.second.at("bondlist").items()) {
                    auto& [btype, bp] = jsonlite::single_item(brec);
                    if (btype != "harmonic") throw std::runtime_error("bond type '" + btype + "' is not supported by the GPU engine");
                    const Value& idx = bp.at("index");
                    if (idx.size() != 2) throw std::runtime_error("harmonic bond needs exactly 2 indices");
                    BondCfg b;
                    b.i = static_cast<int>(idx.at(0).as_number());
                    b.j = static_cast<int>(idx.at(1).as_number());
                    if (b.i < 0 || b.j < 0 || b.i == b.j) throw std::runtime_error("bonded index out of range");
                    b.half_k = m.units.kJmol(bp.at("k").as_number()) / 2.0;
                    b.req = bp.at("req").as_number();
                    mc.bonds.push_back(b);
                }
            // exclusions (molecule.cpp:533-547): pairs within excluded_neighbours bonds
            // of each other (shortest path in the bond graph), plus exclusionlist
            const int dmax = static_cast<int>(kv.second.value("excluded_neighbours", 0.0));
            if (dmax > 0) {
                int size = 0;
                for (auto& b : mc.bonds) size = std::max(size, std::max(b.i, b.j) + 1);
                std::vector<std::vector<int>> adj(static_cast<size_t>(size));
                for (auto& b : mc.bonds) {
                    adj[static_cast<size_t>(b.i)].push_back(b.j);
                    adj[static_cast<size_t>(b.j)].push_back(b.i);
                }
                for (int s0 = 0; s0 < size; ++s0) {
                    std::vector<int> dist(static_cast<size_t>(size), -1), frontier{s0};
                    dist[static_cast<size_t>(s0)] = 0;
                    for (int d = 1; d <= dmax && !frontier.empty(); ++d) {
                        std::vector<int> next;
                        for (int a : frontier)
                            for (int b : adj[static_cast<size_t>(a)])
                                if (dist[static_cast<size_t>(b)] < 0) {
                                    dist[static_cast<size_t>(b)] = d;
                                    next.push_back(b);
                                }
                        frontier.swap(next);
                    }
                    for (int t = s0 + 1; t < size; ++t)
                        if (dist[static_cast<size_t>(t)] > 0) mc.excluded.emplace_back(s0, t);
                }
            }
            if (kv.second.contains("exclusionlist"))
                for (auto& p : kv.second.at("exclusionlist").items()) {
                    int a = static_cast<int>(p.at(0).as_number()), b = static_cast<int>(p.at(1).as_number());
                    mc.excluded.emplace_back(std::min(a, b), std::max(a, b));
                }
            std::sort(mc.excluded.begin(), mc.excluded.end());
            mc.excluded.erase(std::unique(mc.excluded.begin(), mc.excluded.end()), mc.excluded.end());
            m.molecules.push_back(mc);
        }
    } else {
        MoleculeCfg mc;
        mc.name = "atomic";
        mc.atomic = true;
        m.molecules.push_back(mc);
    }
    const Value& geo = cfg.at("geometry");
    if (geo.value("type", "cuboid") != "cuboid") throw std::runtime_error("only cuboid geometry is supported");
    const Value& L = geo.at("length");
    for (int d = 0; d < 3; ++d) m.box[d] = L.is_array() ? L.at(static_cast<size_t>(d)).as_number() : L.as_number();

    const int n = m.ntypes;
    m.pairs.assign(static_cast<size_t>(n * n), PairParam{});
    int nonbonded_terms = 0;
    const Value* coulomb_json = nullptr;
    for (auto& rec : cfg.at("energy").items()) { // energy.cpp:1171-1211
        auto& [key, value] = jsonlite::single_item(rec);
        if (key == "maxenergy") {
            m.maximum_allowed_energy = value.as_number();
            continue;
        }
        if (key == "bonded") { // Hamiltonian::createEnergy (energy.cpp:1328): intra-molecular bonds
            if (value.contains("bondlist") && value.at("bondlist").size() > 0)
                throw std::runtime_error("inter-molecular bonds are not supported by the GPU engine");
            if (m.bonded_term >= 0) throw std::runtime_error("only one bonded term is supported");
            m.bonded_term = static_cast<int>(m.term_names.size());
            m.term_names.push_back("bonded");
            continue;
        }
        if (++nonbonded_terms > 1) throw std::runtime_error("only one nonbonded energy term is supported");
        // per type pair: which additive terms are present
        auto add_term = [&](const std::string& pname, const Value& pcfg, const std::vector<std::pair<int, int>>& cells) {
            if (pname == "lennardjones" || pname == "wca") {
                std::vector<double> s2, e4;
                mixed_matrices(m, pcfg, true, "LB", s2, e4);
                for (auto [a, b] : cells) {
                    PairParam& p = m.pairs[static_cast<size_t>(a * n + b)];
                    if (pname == "lennardjones") {
                        if (p.lj_e4 != 0 || p.lj_s2 != 0) throw std::runtime_error("duplicate lennardjones for a pair");
                        // MixerPairPotentialBase::from_json (potentials.cpp:285-310) reads no
                        // cutoff: a file with one runs untruncated LJ in the reference
                        if (pcfg.contains("cutoff"))
                            throw std::runtime_error("lennardjones has no cutoff in the reference; use nonbonded_splined "
                                                     "with rmax");
                        p.lj_s2 = s2[static_cast<size_t>(a * n + b)];
                        p.lj_e4 = e4[static_cast<size_t>(a * n + b)];
                    } else {
                        if (p.wca_e4 != 0 || p.wca_s2 != 0) throw std::runtime_error("duplicate wca for a pair");
                        p.wca_s2 = s2[static_cast<size_t>(a * n + b)];
                        p.wca_e4 = e4[static_cast<size_t>(a * n + b)];
                    }
                }
            } else if (pname == "hardsphere") {
                std::vector<double> s2, e4;
                mixed_matrices(m, pcfg, false, "arithmetic", s2, e4);
                for (auto [a, b] : cells) m.pairs[static_cast<size_t>(a * n + b)].hs_s2 = s2[static_cast<size_t>(a * n + b)];
            } else if (pname == "coulomb") {
                CoulombModel c = parse_coulomb(m, pcfg, false);
                if (m.coulomb.scheme != kCoulNone &&
                    (m.coulomb.scheme != c.scheme || m.coulomb.lB != c.lB || m.coulomb.rc != c.rc || m.coulomb.alpha != c.alpha))
                    throw std::runtime_error("all pairs must share one coulomb model");
                if (m.coulomb.scheme == kCoulNone) m.coulomb = c;
                for (auto [a, b] : cells) m.pairs[static_cast<size_t>(a * n + b)].coulomb = 1.0;
            } else {
                throw std::runtime_error("potential '" + pname + "' is not supported by the GPU engine");
            }
        };
        std::vector<std::pair<int, int>> all_cells;
        for (int a = 0; a < n; ++a)
            for (int b = 0; b < n; ++b) all_cells.emplace_back(a, b);
        if (key == "nonbonded" || key == "nonbonded_exact" || key == "nonbonded_splined") {
            m.anisotropic_form = key != "nonbonded_splined";
            m.splined = key == "nonbonded_splined";
            // cells overridden by "A B" lists do not take the default list
            std::vector<char> overridden(static_cast<size_t>(n * n), 0);
            for (auto& [pk, pv] : value.members()) {
                const auto sp = pk.find(' ');
                if (pk == "default" || sp == std::string::npos) continue;
                const int a = type_index(m, pk.substr(0, sp)), b = type_index(m, pk.substr(sp + 1));
                overridden[static_cast<size_t>(a * n + b)] = overridden[static_cast<size_t>(b * n + a)] = 1;
                std::vector<std::pair<int, int>> cells{{a, b}};
                if (a != b) cells.emplace_back(b, a);
                for (auto& r : pv.items()) {
                    if (!r.is_object() || r.size() != 1) continue;
                    auto& [pname, pcfg] = jsonlite::single_item(r);
                    add_term(pname, pcfg, cells);
                }
            }
            std::vector<std::pair<int, int>> def_cells;
            for (auto c : all_cells)
                if (!overridden[static_cast<size_t>(c.first * n + c.second)]) def_cells.push_back(c);
            for (auto& r : value.at("default").items()) {
                if (!r.is_object() || r.size() != 1) continue;
                auto& [pname, pcfg] = jsonlite::single_item(r);
                add_term(pname, pcfg, def_cells);
                if (pname == "coulomb" && !coulomb_json) coulomb_json = &pcfg;
            }
        } else if (key == "nonbonded_coulomblj" || key == "nonbonded_coulombwca") {
            m.anisotropic_form = false;
            add_term("coulomb", value.at("coulomb"), all_cells);
            add_term(key == "nonbonded_coulomblj" ? "lennardjones" : "wca",
                     value.at(key == "nonbonded_coulomblj" ? "lennardjones" : "wca"), all_cells);
            coulomb_json = &value.at("coulomb");
        } else if (key == "nonbonded_pm" || key == "nonbonded_coulombhs" || key == "nonbonded_pmwca") {
            m.anisotropic_form = false;
            m.coulomb = parse_coulomb(m, value.at("coulomb"), true);
            for (auto& p : m.pairs) p.coulomb = 1.0;
            if (key == "nonbonded_pmwca") add_term("wca", value.at("wca"), all_cells);
            else add_term("hardsphere", value.contains("hardsphere") ? value.at("hardsphere") : Value::object(), all_cells);
        } else {
            throw std::runtime_error("energy term '" + key + "' is not supported by the GPU engine");
        }
        if (value.contains("cutoff_g2g")) m.g2g2 = parse_g2g(value.at("cutoff_g2g"), m.molecules);
        const bool galore = m.coulomb.scheme == kCoulPlain || m.coulomb.scheme == kCoulEwald ||
                            m.coulomb.scheme == kCoulFanourgakis;
        if (galore) {
            m.has_self_term = true;
            m.term_names.push_back("selfenergy");
        }
        m.term_names.push_back("nonbonded");
        if (coulomb_json && coulomb_json->value("type", "") == "ewald") {
            const Value& j = *coulomb_json;
            EwaldModel& e = m.ewald; // energy.cpp:28-59
            e.present = true;
            e.alpha = j.at("alpha").as_number();
            e.rc = j.at("cutoff").as_number();
            e.n_cutoff = j.contains("kcutoff") ? j.at("kcutoff").as_number() : j.at("ncutoff").as_number();
            e.spherical_sum = j.value("spherical_sum", true);
            e.lB = m.units.bjerrum(j.at("epsr").as_number());
            e.epss = j.value("epss", 0.0);
            e.const_inf = (e.epss < 1) ? 0 : 1;
            const std::string scheme = j.value("ewaldscheme", "PBC");
            if (scheme != "PBC" || j.value("ipbc", false))
                throw std::runtime_error("ewaldscheme '" + scheme + "' is not supported by the GPU engine (PBC only)");
            if (j.value("kappa", 0.0) != 0.0) throw std::runtime_error("Yukawa-Ewald is not supported by the GPU engine");
            m.term_names.push_back("ewald");
        }
    }
    if (nonbonded_terms == 0) throw std::runtime_error("energy: no nonbonded term");
    // range beyond which every pair term vanishes (used to decide on the cell list)
    double range = 0.0;
    const bool coul_any = std::any_of(m.pairs.begin(), m.pairs.end(), [](const PairParam& p) { return p.coulomb != 0; });
    if (coul_any) range = std::max(range, m.coulomb.rc);
    for (const auto& p : m.pairs) {
        if (p.lj_e4 != 0) range = kInf;
        if (p.wca_e4 != 0) range = std::max(range, std::sqrt(p.wca_s2 * 1.2599210498948732));
        if (p.hs_s2 != 0) range = std::max(range, std::sqrt(p.hs_s2));
    }
    if (m.splined) {
        for (auto& rec : cfg.at("energy").items()) {
            auto& [key, value] = jsonlite::single_item(rec);
            if (key == "nonbonded_splined") build_splined(m, value);
        }
        range = 0.0; // every table is zero from its rmax on
        for (const auto& s : m.spl) range = std::max(range, std::sqrt(s.rmax2));
    }
    m.pair_range = range;
    return m;
}

// SplinedPotential::from_json (potentials.cpp:1513-1550): for every type pair, rmin from
// half the summed sigmas and rmax from "rmax" (default 10 rmin) or "cutoff_g2g", moved in
// steps of 0.01 A until |u(rmin)| is within 10 % of u_at_rmin (findLowerDistance,
// 1452-1477) and |u(rmax)| <= u_at_rmax (findUpperDistance, 1489-1511); then an Andrea
// table of u(r^2) with utol / ftol (createKnots, 1559-1600).  The pair's particles carry
// the atom types' charges.
static void build_splined(Model& m, const Value& js) {
    const double utol = js.value("utol", 1e-3), ftol = js.value("ftol", 1e-2);
    const bool hardsphere = js.value("hardsphere", false);
    const double energy_at_rmin = js.value("u_at_rmin", 20.0), energy_at_rmax = js.value("u_at_rmax", 1e-6);
    const double dr = 1e-2;
    const long max_iterations = 1000000;
    const int n = m.ntypes;
    m.spl.assign(static_cast<size_t>(n * n), SplinedPairCfg{});
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) {
            const double qi = m.atoms[static_cast<size_t>(i)].charge, qj = m.atoms[static_cast<size_t>(j)].charge;
            auto u = [&](double r2) { return m.pair_exact(i, qi, j, qj, r2); };
            double rmin = 0.5 * (m.atoms[static_cast<size_t>(i)].sigma + m.atoms[static_cast<size_t>(j)].sigma);
            double rmax = js.value("rmax", rmin * 10);
            if (js.contains("cutoff_g2g")) {
                const Value& c = js.at("cutoff_g2g");
                if (c.is_number()) rmax = c.as_number();
                else if (c.is_object()) rmax = c.at("default").as_number();
            }
            long it = 0;
            while (rmin >= dr) {
                if (it++ == max_iterations) throw std::runtime_error("pair potential spline: cannot determine rmin");
                const double e = std::fabs(u(rmin * rmin));
                if (e > energy_at_rmin * 1.1) rmin += dr;
                else if (e < energy_at_rmin / 1.1) rmin -= dr;
                else break;
            }
            it = 0;
            while (rmax >= dr) {
                if (it++ == max_iterations) throw std::runtime_error("pair potential spline: cannot determine rmax");
                if (std::fabs(u(rmax * rmax)) > energy_at_rmax) rmax += dr;
                else break;
            }
            SplinedPairCfg s;
            const double r0 = std::sqrt(rmin * rmin), r1 = std::sqrt(rmax * rmax);
            s.table = Spline::build(u, rmin * rmin, rmax * rmax, utol, ftol);
            s.rmin2 = r0 * r0;
            s.rmax2 = r1 * r1;
            s.hardsphere = hardsphere && !(s.table.eval(s.rmin2 + dr) < 0);
            m.spl[static_cast<size_t>(i * n + j)] = s;
            m.spl[static_cast<size_t>(j * n + i)] = s;
        }
}

StateCfg StateCfg::from_json(const std::string& text) {
    const jsonlite::Value j = jsonlite::parse(text);
    StateCfg st;
    const jsonlite::Value& geo = j.at("geometry");
    if (geo.value("type", "cuboid") != "cuboid") throw std::runtime_error("state: only cuboid geometry is supported");
    const jsonlite::Value& len = geo.at("length");
    for (int d = 0; d < 3; ++d) st.box[d] = len.is_array() ? len.at(static_cast<size_t>(d)).as_number() : len.as_number();
    if (!j.contains("groups")) throw std::runtime_error("state: 'groups' missing (insertmolecules is not supported)");
    for (const auto& p : j.at("particles").items()) {
        const jsonlite::Value& pos = p.at("pos");
        if (pos.size() != 3) throw std::runtime_error("state: particle position needs three coordinates");
        for (int d = 0; d < 3; ++d) st.xyz.push_back(pos.at(static_cast<size_t>(d)).as_number());
        st.charge.push_back(p.at("q").as_number());
        st.atom_id.push_back(static_cast<int>(p.at("id").as_number()));
    }
    int begin = 0;
    for (const auto& g : j.at("groups").items()) {
        const int size = static_cast<int>(g.at("size").as_number());
        const int capacity = static_cast<int>(g.value("capacity", static_cast<double>(size)));
        if (size < 0 || capacity != size)
            throw std::runtime_error("state: groups with inactive particles (capacity > size) are not supported");
        st.begin.push_back(begin);
        st.size.push_back(size);
        st.molid.push_back(static_cast<int>(g.at("id").as_number()));
        const jsonlite::Value& cm = g.at("cm");
        for (int d = 0; d < 3; ++d) st.mass_center.push_back(cm.at(static_cast<size_t>(d)).as_number());
        begin += capacity;
    }
    if (begin != static_cast<int>(st.charge.size()))
        throw std::runtime_error("state: the groups do not cover the particles exactly (load error)");
    return st;
}

} // namespace fmc
