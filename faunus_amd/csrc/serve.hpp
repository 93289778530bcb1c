// Serving loop for Metropolis MC jobs with continuous batching: jobs are admitted
// while their HBM footprint fits the budget and workgroup slots are free, every
// scheduling round advances all resident jobs by one quantum of moves in a single
// persistent launch (one workgroup per job), and finished jobs leave the batch so
// queued jobs take their slots in the next round.  A job's Markov chain depends only
// on its own Philox stream (keyed by its seed, with replica counter 0), so its result
// is the same alone and in any mix.  Large jobs with a general move set can be shared
// by several workgroups (the split form of the general kernel); their number of
// workgroups is a function of the job alone (its particle and k-vector counts), never
// of the mix, so the equality holds for them too.
#pragma once
#include <chrono>
#include <deque>
#include <memory>
#include <string>
#include <vector>

#include "mc_batch.hpp"

namespace fmc {

struct JobRecord {
    int id = 0;
    int state = 0; // 0 queued, 1 resident, 2 done, 3 failed
    long long moves_total = 0;
    unsigned long long seed = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    size_t footprint = 0;
    // inputs kept until admission
    std::string config;
    std::vector<double> xyz, charge;
    std::vector<int> ids, gbegin, gsize, gmolid;
    std::vector<MoveKindSpec> kinds; // empty: single-particle translations
    // results
    double energy = 0;
    unsigned long long moves_done = 0, accepted = 0;
    std::vector<double> final_xyz;
    std::string error;
};

class Server {
  public:
    Server(int fp32_pairs, long long hbm_budget_bytes, int max_slots);
    ~Server();
    int submit(const std::string& config_json, int n, const double* xyz, const double* charge, const int* ids,
               int ngroups, const int* gbegin, const int* gsize, const int* gmolid, long long moves,
               unsigned long long seed, const std::vector<MoveKindSpec>* kinds = nullptr);
    // One scheduling round: admit, advance every resident job by up to `quantum` moves
    // (scaled by per-move cost so rounds take similar time per job), retire finished jobs.
    void step(long long quantum);
    // Jobs with a general move set and no cell list get ceil(n / atoms_per_part)
    // workgroups (at most max_parts, at least enough for 16 k-vectors per thread);
    // atoms_per_part 0 keeps one workgroup per job.
    void set_split(int atoms_per_part, int max_parts);
    int queued() const { return static_cast<int>(queue_.size()); }
    int resident() const { return static_cast<int>(active_.size()); }
    const JobRecord& job(int id) const { return *jobs_.at(static_cast<size_t>(id)); }
    double now() const;
    long long moves_done_total() const { return moves_total_; }
    size_t footprint_in_use() const { return used_; }

  private:
    struct Slot {
        int job;
        bool general;
        int parts;
        std::unique_ptr<Engine> engine;
        ReplicaState state;
        std::vector<void*> owned;
        double cost;
    };
    int fp32_;
    long long budget_;
    int max_slots_;
    std::vector<std::unique_ptr<JobRecord>> jobs_;
    std::deque<int> queue_;
    std::vector<Slot> active_;
    size_t used_ = 0;
    long long moves_total_ = 0;
    hipStream_t stream_ = nullptr;
    ReplicaState* d_states_ = nullptr;
    long long* d_moves_ = nullptr;
    int capacity_ = 0;
    int split_atoms_ = 0, max_parts_ = 16;
    int coresident_ = 0;            // workgroups of the split general kernel that fit at once
    int2* d_map_ = nullptr;         // block maps of the split launches
    int map_cap_ = 0;
    std::vector<int2> h_map_;
    std::chrono::steady_clock::time_point t0_;

    void admit();
    int parts_for(const ReplicaState& s) const;
    void retire(size_t k);
    static size_t estimate_footprint(const Model& m, int n);
};

} // namespace fmc
