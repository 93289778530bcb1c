// GPU energy engine for one Metropolis system: both MC states (accepted + trial
// delta) resident in HBM, the nonbonded pair ΔU and the Ewald reciprocal ΔU
// computed by hand-written gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "config.hpp"

namespace fmc {

constexpr int kMaxInlineMoved = 4;     // moved particles passed in the kernel arguments
constexpr int kMaxMoved = 4096;        // moved particles per Change
constexpr int kMaxKnots = 64;          // Coulomb S(q) spline knots
constexpr int kSplineBuckets = 1024;   // uniform buckets of q in [0, 1) indexing the knots
constexpr int kMaxTypes = 16;          // atom types handled by the device kernels
constexpr int kPhaseTableMoved = 8;    // moved particles with LDS phase tables (Ewald)
constexpr int kBlock = 256;
constexpr int kSplMaxKnots = 24;   // knots of one nonbonded_splined table (utol 1e-3 gives 7-10)
constexpr int kSplMaxTables = 10;  // distinct type pairs with a table (four atom types)
constexpr int kSplBuckets = 32;    // uniform buckets in r^2 indexing a table's knots

// Device copy of one type pair's parameters (64 bytes).
struct DevPair {
    double lj_s2, lj_e4, wca_s2, wca_e4, hs_s2, coul;
    int spl; // the pair's nonbonded_splined table (0 when the term is not splined)
    int pad;
    double pad2;
};

// One nonbonded_splined table (SplinedPotential's KnotData, potentials.h:773-781): the
// Andrea spline of the exact pair energy in r^2 on [rmin2, rmax2]; bucket[b] is the first
// knot >= x0 + b / inv_w, where the interval search starts.
struct DevSplineTable {
    double rmin2, rmax2, x0, inv_w;
    int n, steps, hs, pad;
    double knots[kSplMaxKnots];
    double coeff[6 * kSplMaxKnots];
    unsigned char bucket[kSplBuckets + 8];
};

// Device copy of the nonbonded model's shared data: the Coulomb scheme, its constants
// and spline of S(q), and the nonbonded_splined tables (nspl > 0 when the term is splined).
struct alignas(16) DevCoulomb {
    int scheme;   // CoulombScheme
    int splined;
    int nknots;
    int spline_steps; // forward steps from a bucket's start that always reach the interval
    double lB, rc, rc_inv, alpha;
    double knots[kMaxKnots];
    double coeff[6 * kMaxKnots];
    // bucket[b]: index of the first knot >= b / kSplineBuckets (start of the interval search)
    unsigned short bucket[kSplineBuckets + 1];
    int nspl;
    DevSplineTable spl[kSplMaxTables];
};

// One moved particle: absolute index, type/group, new and old coordinates (x,y,z,q).
struct MovedParticle {
    int idx, type, gid, pad;
    double4 newp;
    double4 oldp;
};

// Per-move kernel arguments (fits in the dispatch packet's kernarg segment).
struct MoveArgs {
    int m;                // moved particles
    int change_group;     // group index of a single-group change, -1 for several groups
    int internal;         // same-group pairs are part of the change (Change::GroupChange::internal)
    int pad;
    const MovedParticle* ext; // used when m > kMaxInlineMoved (device memory)
    MovedParticle inl[kMaxInlineMoved];
};

// Results of one evaluation, written by the last block into pinned host memory.
struct DeltaResult {
    double pair_new, pair_old, rec_new, rec_old, bond_new, bond_old;
    unsigned long long seq;
    unsigned long long pad;
};

// One bond as seen from one of its atoms (per-particle lists) and one bond of the
// flat list used for the total energy.
struct DevBond {
    int other, pad;
    double half_k, req;
};
struct DevBondPair {
    int i, j;
    double half_k, req;
};

struct CellList {
    bool enabled = false;
    int nc[3] = {0, 0, 0};
    int cap = 0;
    int radius = 1;         // stencil half-width in cells: cells >= pair range / radius
    double inv_cell[3] = {0, 0, 0};
    int* count = nullptr;   // [ncell]
    int* members = nullptr; // [ncell * cap]
    int* cell_of = nullptr; // [N]
    int* overflow = nullptr;
};

class Engine {
  public:
    explicit Engine(const std::string& config_json);
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    const Model& model() const { return model_; }
    int num_particles() const { return n_; }
    int num_kvectors() const { return nk_; }
    int num_terms() const { return static_cast<int>(model_.term_names.size()); }

    void set_particles(int n, const double* xyz, const double* charge, const int* atom_id);
    void set_groups(int ng, const int* begin, const int* size, const int* molid, const double* mass_center);
    void init(); // k-vectors, A_k, full Q(k), cell list (Ewald::init + MetropolisMonteCarlo::init)

    // Trial coordinates of the particles a move displaced (absolute indices).
    void stage_trial(int m, const int* idx, const double* xyz);
    // Trial mass centre of a moved molecular group (needed with a group cutoff).
    void stage_mass_center(int group, const double* cm);
    const std::vector<double>& mass_centers() const { return g_cm_; }
    // the range the cell list was built for (pair potentials, or cutoff_g2g over rigid molecules)
    double cell_range() const { return cell_range_; }
    double interaction_range() const;
    // Energies of the affected pairs/terms of `change` in both states, per term in
    // Hamiltonian order.  Writes the trial Q(k) (Ewald::updateState).
    void delta(const std::vector<int>& groups, const std::vector<char>& all, const std::vector<char>& internal,
               const std::vector<std::vector<int>>& rel, double* e_new, double* e_old);
    // accepted.sync(trial) (accept) or trial.sync(accepted) (reject)
    void sync(bool accept);
    // Change::everything energies of the accepted state, per term.
    void total_energy(double* terms);
    // Full recompute of Q(k) from positions compared with the running Q(k):
    // max |Q_run - Q_full| and relative difference of the reciprocal energy.
    void check_q(double* max_abs_dq, double* rel_rec);
    void get_positions(double* xyz) const;
    hipStream_t stream() const { return stream_; }
    // timing of the fused per-move kernel (HIP events on the engine's stream)
    void set_timing(bool on);
    void timing(double* total_ms, long* launches);
    long launches() const { return launches_; }
    bool cell_list_enabled() const { return cells_.enabled; }
    // Cells of at least pair range / d (d = 1 or 2) with a (2d+1)^3 stencil; before init().
    // Only the general device kernel walks d = 2; the other paths need d = 1.
    void set_cell_divisions(int d);
    // Cell divisions of an initialised engine, applied by the next reset_positions.
    void regrid(int d);
    int cell_radius() const { return cells_.radius; }
    // Members per cell of the cell list built by init(): 0 chooses max(2 x the fullest
    // cell, 4 x the mean + 32); a fixed capacity must hold the fullest initial cell.
    void set_cell_capacity(int cap);
    int cell_capacity() const { return cells_.cap; }

    // Pair-parameter and spline tables (as uploaded).
    const std::vector<DevPair>& dev_pairs() const { return h_pairs_; }
    const DevCoulomb& dev_coulomb() const { return h_coul_; }

    // Raw device state for the device-resident MC drivers.
    double4* d_pos() { return d_pos_; }
    const uint32_t* d_meta() const { return d_meta_; }
    double2* d_q_acc() { return d_q_acc_; }
    const int* d_kn() const { return d_kn_; }
    const double* d_ak() const { return d_ak_; }
    const DevPair* d_pairs() const { return d_pairs_; }
    const DevCoulomb* d_coul() const { return d_coul_; }
    CellList& cells() { return cells_; }
    int* d_group_flags() { return d_gflags_; }
    const double* d_g2g_cutoffs() { return d_g2g_; }
    double* d_mass_centers() { return d_gcm_; }
    const int* d_bond_offsets() const { return d_bond_off_; }
    const DevBond* d_bonds() const { return d_bond_; }
    int num_bonds() const { return nbonds_; }
    const int* d_exclusion_offsets() const { return d_excl_off_; }
    const int* d_exclusions() const { return d_excl_; }
    const std::vector<int>& group_molid() const { return g_molid_; }
    double self_energy_of(int idx) const;
    const std::vector<int>& group_of() const { return h_gid_; }
    const std::vector<int>& group_begin() const { return g_begin_; }
    const std::vector<int>& group_size() const { return g_size_; }
    const std::vector<double>& host_charges() const { return h_q_; }
    const std::vector<int>& host_types() const { return h_type_; }
    void refresh_host_positions();
    // Replace the accepted positions (3n) of an initialised engine and rebuild what
    // depends on them: Q(k) (full update), the cell list (automatic capacity), the
    // dipole, and the molecular mass centres (from `mass_center`, 3 per group, or
    // recomputed when null).  MetropolisMonteCarlo::restore + init (montecarlo.cpp:118-137).
    void reset_positions(const double* xyz, const double* mass_center);

  private:
    Model model_;
    int n_ = 0, nk_ = 0;
    hipStream_t stream_ = nullptr;
    int device_ = 0;
    int num_cu_ = 0;
    // particles (accepted state)
    double4* d_pos_ = nullptr;
    uint32_t* d_meta_ = nullptr; // type | group << 8
    std::vector<double> h_xyz_, h_q_;
    std::vector<int> h_type_, h_gid_;
    // groups
    std::vector<int> g_begin_, g_size_, g_molid_;
    std::vector<double> g_cm_;
    double cell_range_ = 0.0;
    bool any_molecular_ = false;
    int* d_gflags_ = nullptr;    // bit0 molecular, bit1 rigid, molecule type from bit 8
    double* d_g2g_ = nullptr;    // cutoff_g2g: squared cutoff per pair of molecule types
    double* d_gcm_ = nullptr;    // accepted mass centres, 3 per group
    int staged_cm_group_ = -1, pending_cm_group_ = -1;
    double staged_cm_[3] = {0, 0, 0}, pending_cm_[3] = {0, 0, 0};
    // bonds (intra-molecular harmonic), per particle and as one list
    int* d_bond_off_ = nullptr;
    DevBond* d_bond_ = nullptr;
    DevBondPair* d_bond_all_ = nullptr;
    int nbonds_ = 0;
    int* d_excl_off_ = nullptr; // excluded partners per particle (null: no exclusions)
    int* d_excl_ = nullptr;
    // potentials
    std::vector<DevPair> h_pairs_;
    DevCoulomb h_coul_{};
    DevPair* d_pairs_ = nullptr;
    DevCoulomb* d_coul_ = nullptr;
    // Ewald
    int* d_kn_ = nullptr;      // packed (nx, ny, nz) as signed bytes
    double* d_ak_ = nullptr;
    double2* d_q_acc_ = nullptr;
    double2* d_q_trial_ = nullptr;
    double dipole_[3] = {0, 0, 0};
    double dipole_trial_[3] = {0, 0, 0};
    // reduction scratch
    double2* d_partials_ = nullptr;
    unsigned* d_ticket_ = nullptr;
    int max_blocks_ = 0;
    DeltaResult* h_result_ = nullptr; // pinned
    MovedParticle* d_moved_ = nullptr;
    std::vector<MovedParticle> staged_;
    // cell list
    CellList cells_;
    int cell_divisions_ = 1;
    int cell_capacity_ = 0;
    // pending change (between delta and sync)
    bool pending_ = false;
    std::vector<MovedParticle> pending_moved_;
    unsigned long long seq_ = 0;
    // timing
    bool timing_ = false;
    std::vector<hipEvent_t> ev_pool_;
    std::vector<std::pair<int, int>> ev_pairs_;
    long launches_ = 0;

    void upload_potentials();
    void build_kvectors();
    void full_q(double2* out);
    void build_cell_list();
    void launch_delta(const MoveArgs& args, int pair_blocks, int ewald_blocks, bool bond_block);
    void build_group_tables();
    // PBC-safe mass centre of group g around its first atom (Geometry::massCenter)
    void group_mass_center_host(int g, double* cm) const;
    double reciprocal_energy_host(const std::vector<double2>& q) const;
    double surface_energy(const double* dipole) const;
};

void check_hip(hipError_t e, const char* what);
// Time (ms per launch) of a streaming copy of `bytes` bytes read + written.
double copy_probe(long long bytes, int repeats);

} // namespace fmc
