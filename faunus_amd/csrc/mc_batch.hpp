// Device-resident Metropolis: every replica is advanced by one workgroup of a
// persistent kernel (proposal, pair dU, Ewald dU, acceptance test and commit all
// on the GPU).  Replicas of a temperature ladder share one parameter set at the
// reference temperature and differ in beta = T_ref / T; replica exchange permutes
// the beta labels (configurations never move).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "engine.hpp"

namespace fmc {

constexpr int kGeneralMaxMoved = 1024; // atoms moved together by one general-kernel move

// One kind of trial move of the general move set (the reference's MoveCollection
// entries, move.cpp:513-525, picked with probability weight / total weight).
struct MoveKindDev {
    int kind;         // 0 translate one atom, 1 translate+rotate a molecule, 2 pivot
    int group_first;  // groups of the target molecule type: kind_groups[group_first ...]
    int group_count;
    int bond_first;   // kind 2: the molecule's harmonic bonds, pivot_bonds[bond_first ...]
    int bond_count;
    int pad;
    unsigned long long cum; // cumulative weight
    double dp, dprot;
};

// Host description of one move kind (see MoveKindDev).
struct MoveKindSpec {
    int kind = 0, molid = 0;
    unsigned weight = 1;
    double dp = 0.0, dprot = 0.0;
};

struct FixPos; // mc_common.hpp

// Per-replica device state (one per workgroup).
struct ReplicaState {
    int n, nk, ncut, ntypes;
    int cells, nc0, nc1, nc2;
    int cap, kpt, replica_id, ladder;
    double L, half, inv;
    double inv_cell;
    double rec_prefactor; // 2 pi lB / V
    double4* pos;
    FixPos* pos32;        // FP32-pair form: fixed-point coordinates and FP32 charge, or null
    const uint32_t* meta;
    double2* q;
    const int* kn;
    const double* gk;     // A_k / (nx>0 ? 2 : 1) as a function of |n|^2 (cubic box)
    int* cell_count;
    int* cell_members;
    int* cell_of;
    const DevPair* pairs;
    const DevCoulomb* coul;
    const double* dp_type;
    unsigned long long seed;
    // dynamic
    double beta;
    double energy;      // running total energy (kT at T_ref)
    unsigned long long moves;
    unsigned long long accepted;
    double sum_du;
    int error;
    int terms;          // pair terms present in the pair table (kTermLJ | kTermWCA | kTermHS)
    // general move set (mc_general_kernel); nkinds == 0 for the translate-only kernel
    int nkinds, max_moved;
    unsigned long long total_weight;
    const MoveKindDev* kinds;
    const int* kind_groups;
    const int2* pivot_bonds;  // (i, j) relative to the molecule
    const int* gbegin;
    const int* gsize;
    const int* gflags;        // bit0 molecular, bit1 rigid, molecule type from bit 8
    double* gcm;              // mass centres (translate+rotate targets; every molecular group with cutoff_g2g)
    const double* dp_atom;    // per type: the atom's dp, or -1 when the move's dp applies
    const int* bond_off;      // per-particle bond lists (bonded term), or null
    const DevBond* bonds;
    const int* excl_off;      // per-particle excluded partners, or null
    const int* excl;
    const float4* stab;       // FP32 forms: cubic table of S(q) (kSTable intervals), or null
    int parts;                // workgroups sharing this replica in a split general launch (1 otherwise)
    int nmol;                 // molecule types (rows of g2g2)
    const double* g2g2;       // cutoff_g2g: squared cutoff per pair of molecule types, or null
    const double* mw;         // per type: the atom mass (mass centres of changed molecular groups)
    void* gslots;             // split general launch: 2 x parts exchange slots
    // energy samples (SystemEnergy, analysis.h:547-575): after every move whose index + 1 is a
    // multiple of sample_every (0: off), the running energy is added to e_sum and its square to e_sum2
    long long sample_every;
    unsigned long long e_samples;
    double e_sum, e_sum2;
    double pair_range;        // distance beyond which every pair term vanishes (cell-list forms)
    double dip[3];            // surface term: the replica's dipole sum q r (positions as stored)
    double surf;              // surface term prefactor const_inf 2 pi lB / ((2 eps_s + 1) V), 0 without it
    int cell_radius;          // stencil half-width in cells (1: 27 cells, 2: 125 cells)
    int threads;              // workgroup size of the translate-only kernel chosen by the tuner (0: the default rule)
};

// Free-running temperature ladders in one persistent launch (Batch::run_ladders): replica
// r runs intervals[r] swap intervals of its moves[r] moves; after each, the replicas of
// a ladder meet at a barrier in HBM and the last to arrive swaps the ladder's rungs, so
// ladders of different sizes run their intervals at their own pace.  The kernel takes it
// by a device pointer (null: one interval, no barrier, the ordinary launch).
struct LadderRun {
    const int* intervals;     // per replica
    const int* ladder_of;     // per replica: its ladder's index into begin/size, -1 for none
    const int* begin;         // per ladder: first entry in rung_slot
    const int* size;          // per ladder: rungs
    int* rung_slot;           // replica index per rung, ordered by decreasing beta
    unsigned* arrive;         // per ladder: arrivals so far (zeroed before the launch)
    unsigned* gen;            // per ladder: intervals whose swap is done (zeroed before the launch)
    unsigned long long* counters; // swaps attempted, accepted
    unsigned long long seed, round0;
    long long spin_limit;     // polls of a waiting replica before it stops with error 2
};

// One recorded move of the trace (tests).
struct TraceEntry {
    int index;
    int accepted;
    double du;     // energy change (kT at T_ref) before the reject reset
    double x, y, z; // proposed position
};

// Device state of one replica (engine already initialised); device buffers it makes
// are appended to `owned`, which the caller frees.
ReplicaState make_replica_state(Engine& e, bool fp32, unsigned long long seed, int stream_id, double beta, int ladder,
                                std::vector<void*>& owned, const std::vector<MoveKindSpec>* kinds = nullptr);
// Pair energies at the squared distances r2[n] of types ta, tb with charges qa, qb in the
// two device forms: FP64 (u64) and FP32 (u32, the device-resident kernels' form).
void pair_probe(Engine& e, int ta, int tb, double qa, double qb, int n, const double* r2, double* u64, double* u32);
// Workgroup size of the translate-only kernel for a replica: 512 threads (two waves per
// SIMD) without a cell list and with K <= 4096, else 256.
int advance_threads(const ReplicaState& s);
// Persistent launches advancing `count` replicas (states already on the device), one
// per run of replicas with the same workgroup size.  `ladders` (a device pointer, null
// for one interval without swaps) makes it one cooperative launch of free-running
// ladders.
void launch_advance(const ReplicaState* h_states, int count, ReplicaState* d_states, const long long* d_moves, bool f32,
                    TraceEntry* trace, int trace_cap, hipStream_t stream, const LadderRun* ladders = nullptr);
// Same for replicas with a general move set (molecular moves, bonds).  Replicas with
// parts > 1 are shared by that many workgroups: the launch is then cooperative, its
// grid is the sum of the parts and block b advances part map[b].y of replica map[b].x
// (d_blockmap, as laid out by general_blockmap, on the device; map_blocks its length,
// 0 for one block per part).
void launch_general(const ReplicaState* h_states, int count, ReplicaState* d_states, const long long* d_moves, bool f32,
                    TraceEntry* trace, int trace_cap, hipStream_t stream, const int2* d_blockmap = nullptr,
                    int map_blocks = 0);
// Share a replica among `parts` workgroups of later general launches (1 undoes it);
// allocates its exchange slots into `owned`.
void set_replica_parts(ReplicaState& s, int parts, std::vector<void*>& owned);
// Host block map of a general launch over these replicas; returns the grid size.  With
// stride > 1 the parts sit on every stride-th block (idle blocks between them), so that
// with stride = the XCD count the parts of a replica share one XCD and its L2.
int general_blockmap(const ReplicaState* h_states, int count, std::vector<int2>& map, int stride = 1);
// Workgroups of the split general kernel that fit on the device at once.
int general_coresident_blocks(bool f32);

// One candidate launch shape of the tile autotuner and its measured kernel time per
// replica-move (microseconds).
struct TuneCandidate {
    int value; // workgroup threads (translate-only replicas) or cell divisions (move sets)
    double us_per_move;
};

// A set of replicas advanced together on one stream.
class Batch {
  public:
    explicit Batch(int precision_fp32);
    // cells of at least pair range / d for the replicas added afterwards (d = 1 or 2; the
    // finer grid is walked by the general kernel only)
    void set_cell_divisions(int d);
    // members per cell of the cell lists of the replicas added afterwards (0: automatic)
    void set_cell_capacity(int cap);
    ~Batch();
    int add_replica(const std::string& config_json, int n, const double* xyz, const double* charge, const int* ids,
                    int ngroups, const int* gbegin, const int* gsize, const int* gmolid, double beta,
                    unsigned long long seed, int ladder, const std::vector<MoveKindSpec>* kinds = nullptr);
    void finalize();
    // advance every replica by `moves[r]` moves (one launch); trace optional
    void run(const std::vector<long long>& moves, bool trace, int trace_cap);
    // one replica spread over `groups` workgroups (every `stride`-th block of the grid)
    void run_split(long long moves, int groups, int stride, bool trace, int trace_cap);
    // run_split of a translate-only replica that allocates nothing: the exchange slots live
    // in a static per-GPU workspace (at most 64 workgroups); the batch must hold its
    // device state already
    void run_split_static(long long moves, int groups, int stride);
    int split_error();
    // one replica over `groups` workgroups with up to `depth` moves in flight
    void run_pipe(long long moves, int groups, int depth, bool trace, int trace_cap);
    void swap_ladders(unsigned long long seed, unsigned long long round);
    // One replica-exchange interval captured as a HIP graph: the persistent launch
    // advancing replica r by moves[r] and one round of ladder swaps (device round counter
    // starting at round0); replay(count) launches it count times on the batch's stream.
    void capture_interval(const std::vector<long long>& moves, unsigned long long seed, unsigned long long round0);
    // intervals[r] swap intervals of moves[r] moves for every replica in one cooperative
    // launch, each ladder swapping after its own intervals (round round0 + interval)
    void run_ladders(const std::vector<long long>& moves, const std::vector<int>& intervals, unsigned long long seed,
                     unsigned long long round0);
    void replay(int count);
    // swaps attempted and accepted so far (host rounds and captured rounds)
    void swap_counts(long long* attempted, long long* accepted);
    // time the batch's candidate launch shapes on `trial` moves per replica (restored
    // afterwards) and keep the fastest; empty when the batch has no choice to make
    std::vector<TuneCandidate> autotune(long long trial);
    int tuned() const { return tuned_; }
    // energy samples every `every` moves (0: off) for all replicas, counters cleared
    void sample_energy(long long every);
    // pair-distance histogram of types a and b (a == b: pairs i < j) of replica r's current
    // configuration, bins [k dr, (k+1) dr), k < nbins, added to counts
    void rdf(int r, int type_a, int type_b, double dr, int nbins, unsigned long long* counts);
    void sync();
    int size() const { return static_cast<int>(engines_.size()); }
    ReplicaState state(int r);
    void positions(int r, double* xyz);
    // Checkpoint of replica r: positions (3n), molecular mass centres (3 per group; may
    // be null), Philox move counter and accepted count.
    void get_replica_state(int r, double* xyz, double* mass_center, unsigned long long* moves,
                           unsigned long long* accepted);
    // Resume replica r from a checkpoint: positions, mass centres (null: recomputed),
    // move counter and accepted count; Q(k), the cell list and the running energy are
    // rebuilt and the replica's error is cleared.
    void set_replica_state(int r, const double* xyz, const double* mass_center, unsigned long long moves,
                           unsigned long long accepted);
    std::vector<TraceEntry> trace(int r, int count);
    Engine& engine(int r) { return *engines_[static_cast<size_t>(r)]; }
    hipStream_t stream() const { return stream_; }
    void set_timing(bool on) { timing_ = on; }
    double last_kernel_ms() const { return last_ms_; }
    int threads_for(int r) const;
    void set_threads(int threads) { threads_override_ = threads; }
    long long swaps_attempted() const { return swaps_att_; }
    long long swaps_accepted() const { return swaps_acc_; }

  private:
    hipGraphExec_t graph_ = nullptr;
    std::vector<void*> graph_owned_; // ladders, rung order and counters of the captured interval
    unsigned long long* graph_counters_ = nullptr; // attempted, accepted, next round
    int tuned_ = 0;
    int fp32_ = 0;
    int general_ = -1; // -1 unset, 0 translate-only replicas, 1 general move sets
    int cell_divisions_ = 1;
    int cell_capacity_ = 0;
    std::vector<std::unique_ptr<Engine>> engines_;
    std::vector<ReplicaState> h_states_;
    std::vector<void*> owned_; // device buffers made for the replicas
    ReplicaState* d_states_ = nullptr;
    long long* d_moves_ = nullptr;
    TraceEntry* d_trace_ = nullptr;
    int trace_cap_ = 0;
    hipStream_t stream_ = nullptr;
    bool finalized_ = false;
    bool timing_ = false;
    double last_ms_ = 0;
    int threads_override_ = 0;
    long long swaps_att_ = 0, swaps_acc_ = 0;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    bool ev_pending_ = false; // both events recorded around a launch not yet read by sync()
    void time_begin();
    void time_end();
    void* split_slots_ = nullptr;
    void* split_priv_ = nullptr; // private position copies of the split kernel's workgroups 1..G-1
    size_t split_priv_bytes_ = 0;
    void* pipe_slots_ = nullptr;
    int* split_err_ = nullptr;
    int2* d_map_ = nullptr; // block map of a split general launch
    int map_cap_ = 0;
    std::vector<int2> h_map_;
    long long h_split_moves_ = 0;
    void run_split_general(long long moves, int groups, int stride, bool trace, int trace_cap);
};

} // namespace fmc
