// Host-side model of the reference's input for the nonbonded + Ewald path:
// atom types, molecule types, the `energy:` list and the cuboid geometry, parsed
// from the reference's JSON (the YAML input after yason.py conversion).
//
// Supported energy terms (energy.cpp:1280-1327): nonbonded / nonbonded_exact
// (FunctorPotential with "default" and "A B" lists), nonbonded_splined
// (SplinedPotential: the same lists tabulated per type pair), nonbonded_coulomblj,
// nonbonded_coulombwca, nonbonded_pm / nonbonded_coulombhs, nonbonded_pmwca, plus
// `maxenergy` and `bonded` (intra-molecular harmonic bonds, energy.cpp:974-1124).  Pair potentials: lennardjones, wca, hardsphere, coulomb
// {type: plain|ewald|fanourgakis}.  Ewald reciprocal space is added when the
// coulomb type is "ewald" (addEwald, energy.cpp:1134-1160).
#pragma once
#include <cstdint>
#include <functional>
#include <limits>
#include <string>
#include <vector>

#include "json_lite.hpp"

namespace fmc {

constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double kInf = std::numeric_limits<double>::infinity();

// units.h:14-41 (same constants as the reference).
struct UnitSystem {
    double temperature = 298.15;
    double kT() const { return temperature * 1.380658e-23; }
    double bjerrum(double epsr) const {
        return 1.602177e-19 * 1.602177e-19 / (4 * kPi * 8.85419e-12 * epsr * 1e-10 * kT());
    }
    double kJmol(double v) const { return v / kT() / 6.022137e23 * 1e3; }
};

// Quintic spline with adaptive knots (the reference's Andrea tabulator,
// tabulate.h:95-308), used for the Coulomb short-range function S(q).
struct Spline {
    std::vector<double> knots; // ascending
    std::vector<double> coeff; // 6 per interval
    double eval(double x) const;
    static Spline build(const std::function<double(double)>& f, double xmin, double xmax, double utol,
                        double ftol = -1.0);
};

enum CoulombScheme : int { kCoulNone = 0, kCoulPlain = 1, kCoulPM = 2, kCoulEwald = 3, kCoulFanourgakis = 4 };

struct CoulombModel {
    int scheme = kCoulNone;
    double lB = 0.0;
    double rc = kInf, rc_inv = 0.0, alpha = 0.0;
    bool splined = true;
    Spline S;
    double self_prefactor = 0.0; // per-particle self energy = self_prefactor * q^2 (kT)
    double short_range_exact(double q) const;
};

// Per type pair parameters of the additive pair form u = LJ + WCA + HS + Coulomb.
// Terms absent for a pair have zero coefficients.
struct PairParam {
    double lj_s2 = 0, lj_e4 = 0;
    double wca_s2 = 0, wca_e4 = 0;
    double hs_s2 = 0;
    double coulomb = 0; // 1 if the pair carries the Coulomb term
};

// SplinedPotential's data for one type pair (potentials.h:770-824): Andrea table of the
// exact pair energy in r^2 on [rmin^2, rmax^2], zero beyond rmax, exact (or +inf with
// `hardsphere`) below rmin.
struct SplinedPairCfg {
    Spline table;
    double rmin2 = 0, rmax2 = 0;
    bool hardsphere = false;
};

struct EwaldModel {
    bool present = false;
    double alpha = 0, rc = 0, n_cutoff = 0, lB = 0, epss = 0, const_inf = 0;
    bool spherical_sum = true;
};

struct AtomTypeCfg {
    std::string name;
    double charge = 0, sigma = 0, eps = 0, dp = 0, dprot = 0, mw = 1;
    bool has_dp = false;
};

// HarmonicBond (bonds.cpp:100-105): indices relative to the molecule, k in kT/A^2 / 2.
struct BondCfg {
    int i = 0, j = 0;
    double half_k = 0, req = 0;
};

struct MoleculeCfg {
    std::string name;
    bool atomic = false, rigid = false, compressible = false;
    std::vector<int> atoms;                     // atom types listed by "atoms"
    std::vector<BondCfg> bonds;                 // "bondlist" (harmonic)
    std::vector<std::pair<int, int>> excluded;  // excluded_neighbours + exclusionlist, sorted (i<j)
};

struct Model {
    UnitSystem units;
    double box[3] = {0, 0, 0};
    std::vector<AtomTypeCfg> atoms;
    std::vector<MoleculeCfg> molecules;
    int ntypes = 0;
    std::vector<PairParam> pairs; // ntypes*ntypes
    CoulombModel coulomb;
    EwaldModel ewald;
    bool has_self_term = false;
    // GroupCutoff (energy.h:746-768; from_json, energy.cpp:1897-1933): squared cutoff of
    // the mass-centre distance per pair of molecule types (nmol x nmol, row major); empty
    // when the nonbonded term has no cutoff_g2g
    std::vector<double> g2g2;
    bool has_g2g() const { return !g2g2.empty(); }
    double maximum_allowed_energy = kInf;
    bool anisotropic_form = true;
    std::vector<std::string> term_names; // Hamiltonian order: energy list, with [selfenergy] before nonbonded, [ewald] after
    int bonded_term = -1;                // index of "bonded" in term_names, -1 if absent
    double pair_range = kInf; // distance beyond which every pair term vanishes
    bool splined = false;               // nonbonded_splined
    std::vector<SplinedPairCfg> spl;    // ntypes*ntypes (symmetric) when splined

    // Exact pair energy u = LJ + WCA + HS + Coulomb of types a, b with charges qa, qb.
    double pair_exact(int a, double qa, int b, double qb, double r2) const;
    // The pair energy the nonbonded term uses (splined when the term is).
    double pair_energy(int a, double qa, int b, double qb, double r2) const;

    static Model from_json(const std::string& text);
};

// A reference state file (the Space JSON of Space::to_json / from_json, space.cpp:449-500,
// with Group to_json / from_json, group.cpp:231-253): cuboid geometry, particles in order
// (pos, q, atom id) and groups laid end to end (molecule id, size, mass centre).
struct StateCfg {
    double box[3] = {0, 0, 0};
    std::vector<double> xyz, charge;
    std::vector<int> atom_id;
    std::vector<int> begin, size, molid;
    std::vector<double> mass_center; // 3 per group
    static StateCfg from_json(const std::string& text);
};

} // namespace fmc
