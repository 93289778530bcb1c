// Serving loop with continuous batching (see serve.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "serve.hpp"

namespace fmc {

Server::Server(int fp32_pairs, long long hbm_budget_bytes, int max_slots)
    : fp32_(fp32_pairs), budget_(hbm_budget_bytes), max_slots_(max_slots) {
    if (budget_ <= 0) throw std::runtime_error("server: HBM budget must be positive");
    if (max_slots_ < 1) throw std::runtime_error("server: at least one slot");
    check_hip(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    t0_ = std::chrono::steady_clock::now();
}

Server::~Server() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (auto& s : active_)
        for (void* p : s.owned) (void)hipFree(p);
    (void)hipFree(d_states_);
    (void)hipFree(d_moves_);
    (void)hipFree(d_map_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void Server::set_split(int atoms_per_part, int max_parts) {
    if (atoms_per_part < 0) throw std::runtime_error("server: atoms per workgroup must be >= 0");
    if (max_parts < 1 || max_parts > 64) throw std::runtime_error("server: 1 to 64 workgroups per job");
    if (!active_.empty() || !queue_.empty()) throw std::runtime_error("server: set the split policy before submitting jobs");
    split_atoms_ = atoms_per_part;
    max_parts_ = max_parts;
}

// Workgroups of a job: from its size alone, so that its chain does not depend on the mix.
int Server::parts_for(const ReplicaState& s) const {
    if (s.nkinds == 0 || s.cells) return 1;
    int parts = 1;
    if (split_atoms_ > 0) parts = std::min(max_parts_, (s.n + split_atoms_ - 1) / split_atoms_);
    const int kmax_parts = (s.nk + 16 * 256 - 1) / (16 * 256); // at most 16 k-vectors per thread
    return std::max(std::max(parts, kmax_parts), 1);
}

double Server::now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

// Device bytes an engine holds for a system of n particles (positions, FP32 mirror,
// metadata, k-vectors, A_k, both Q(k) buffers, cell list, reduction and move buffers).
size_t Server::estimate_footprint(const Model& m, int n) {
    size_t k = 0;
    if (m.ewald.present) {
        const double nc = m.ewald.n_cutoff;
        k = static_cast<size_t>(4.0 / 3.0 * kPi * nc * nc * nc / 2.0 + 4 * nc * nc + 8);
    }
    size_t cells = 0;
    if (std::isfinite(m.pair_range) && m.pair_range > 0) {
        const double nc = std::floor(m.box[0] / m.pair_range);
        if (nc >= 3) cells = static_cast<size_t>(nc * nc * nc) * (4 + 4 * 48) + 4 * static_cast<size_t>(n);
    }
    const size_t per_particle = 32 + 4 + 16;
    return per_particle * static_cast<size_t>(n) + k * (4 + 8 + 32) + cells + 4096 * (16 + 80) + 16 * 1024;
}

int Server::submit(const std::string& config_json, int n, const double* xyz, const double* charge, const int* ids,
                   int ngroups, const int* gbegin, const int* gsize, const int* gmolid, long long moves,
                   unsigned long long seed, const std::vector<MoveKindSpec>* kinds) {
    if (n <= 0 || moves < 0) throw std::runtime_error("server: bad job");
    const Model m = Model::from_json(config_json); // validates the input before queueing
    auto j = std::make_unique<JobRecord>();
    j->id = static_cast<int>(jobs_.size());
    j->moves_total = moves;
    j->seed = seed;
    j->t_submit = now();
    j->footprint = estimate_footprint(m, n);
    if (static_cast<long long>(j->footprint) > budget_) throw std::runtime_error("server: job larger than the HBM budget");
    j->config = config_json;
    j->xyz.assign(xyz, xyz + 3 * static_cast<size_t>(n));
    j->charge.assign(charge, charge + n);
    j->ids.assign(ids, ids + n);
    j->gbegin.assign(gbegin, gbegin + ngroups);
    j->gsize.assign(gsize, gsize + ngroups);
    j->gmolid.assign(gmolid, gmolid + ngroups);
    if (kinds) {
        if (kinds->empty()) throw std::runtime_error("server: empty move set");
        j->kinds = *kinds;
    }
    const int id = j->id;
    jobs_.push_back(std::move(j));
    queue_.push_back(id);
    return id;
}

void Server::admit() {
    while (!queue_.empty() && static_cast<int>(active_.size()) < max_slots_) {
        JobRecord& j = *jobs_[static_cast<size_t>(queue_.front())];
        if (static_cast<long long>(used_ + j.footprint) > budget_) break; // first in, first out
        queue_.pop_front();
        Slot s;
        s.job = j.id;
        try {
            s.engine = std::make_unique<Engine>(j.config);
            s.engine->set_particles(static_cast<int>(j.charge.size()), j.xyz.data(), j.charge.data(), j.ids.data());
            s.engine->set_groups(static_cast<int>(j.gbegin.size()), j.gbegin.data(), j.gsize.data(), j.gmolid.data(),
                                 nullptr);
            s.engine->init();
            // the stream is keyed by the job seed alone, so a job runs the same chain in any mix
            s.general = !j.kinds.empty();
            s.state = make_replica_state(*s.engine, fp32_ != 0, j.seed, 0, 1.0, -1, s.owned,
                                         s.general ? &j.kinds : nullptr);
            s.state.kpt = (s.state.nk + 255) / 256;
            s.parts = s.general ? parts_for(s.state) : 1;
            if (s.parts > 1) {
                if (coresident_ == 0) coresident_ = general_coresident_blocks(fp32_ != 0);
                if (s.parts > coresident_) throw std::runtime_error("job needs more workgroups than fit on the device");
                set_replica_parts(s.state, s.parts, s.owned);
            } else if (s.state.kpt > (s.general ? 16 : 68)) {
                throw std::runtime_error("too many k-vectors for the device MC");
            }
        } catch (const std::exception& e) {
            for (void* p : s.owned) (void)hipFree(p);
            j.state = 3;
            j.error = e.what();
            j.t_done = now();
            continue;
        }
        // per-move time scale: the pair and Ewald work of one workgroup, plus the
        // exchange between workgroups of a split job
        s.cost = (static_cast<double>(s.state.n) + 2.0 * s.state.nk) / s.parts + (s.parts > 1 ? 1500.0 : 0.0);
        j.state = 1;
        j.t_admit = now();
        j.config.clear();
        j.xyz.clear();
        j.charge.clear();
        j.ids.clear();
        used_ += j.footprint;
        active_.push_back(std::move(s));
    }
}

void Server::retire(size_t k) {
    Slot& s = active_[k];
    JobRecord& j = *jobs_[static_cast<size_t>(s.job)];
    j.energy = s.state.energy;
    j.moves_done = s.state.moves;
    j.accepted = s.state.accepted;
    j.final_xyz.resize(3 * static_cast<size_t>(s.engine->num_particles()));
    s.engine->refresh_host_positions();
    s.engine->get_positions(j.final_xyz.data());
    j.state = s.state.error ? 3 : 2;
    if (s.state.error == 2) j.error = "device error (a workgroup of the split job did not arrive)";
    else if (s.state.error) j.error = "device error (cell capacity)";
    j.t_done = now();
    for (void* p : s.owned) (void)hipFree(p);
    used_ -= j.footprint;
    active_.erase(active_.begin() + static_cast<long>(k));
}

void Server::step(long long quantum) {
    admit();
    if (active_.empty()) return;
    const int count = static_cast<int>(active_.size());
    if (count > capacity_) {
        (void)hipFree(d_states_);
        (void)hipFree(d_moves_);
        capacity_ = std::max(count, 2 * capacity_);
        check_hip(hipMalloc(&d_states_, sizeof(ReplicaState) * capacity_), "hipMalloc states");
        check_hip(hipMalloc(&d_moves_, sizeof(long long) * capacity_), "hipMalloc moves");
    }
    // jobs with and without a cell list, and with and without a general move set, run
    // in separate launches
    std::vector<ReplicaState> states(static_cast<size_t>(count));
    std::vector<long long> moves(static_cast<size_t>(count));
    std::vector<int> order;
    // translate-only jobs also split by workgroup size (one launch each)
    auto launch_class = [&](const Slot& s) {
        return (s.state.cells != 0 ? 1 : 0) + (s.general ? 2 : 0) + (!s.general && advance_threads(s.state) == 512 ? 4 : 0);
    };
    // split jobs first within their class, so that they lead the cooperative launches
    for (int pass = 0; pass < 8; ++pass)
        for (int split = 1; split >= 0; --split)
            for (int k = 0; k < count; ++k) {
                const Slot& s = active_[static_cast<size_t>(k)];
                if (launch_class(s) == pass && (s.parts > 1 ? 1 : 0) == split) order.push_back(k);
            }
    for (int r = 0; r < count; ++r) {
        const Slot& s = active_[static_cast<size_t>(order[static_cast<size_t>(r)])];
        const JobRecord& j = *jobs_[static_cast<size_t>(s.job)];
        const long long remaining = j.moves_total - static_cast<long long>(s.state.moves);
        const long long share = std::max(1LL, static_cast<long long>(static_cast<double>(quantum) * 1000.0 / s.cost));
        states[static_cast<size_t>(r)] = s.state;
        moves[static_cast<size_t>(r)] = std::max(0LL, std::min(remaining, share));
    }
    check_hip(hipMemcpyAsync(d_states_, states.data(), sizeof(ReplicaState) * count, hipMemcpyHostToDevice, stream_),
              "copy states");
    check_hip(hipMemcpyAsync(d_moves_, moves.data(), sizeof(long long) * count, hipMemcpyHostToDevice, stream_),
              "copy moves");
    // launches: [first, last) replicas, with their block map at map_off (-1: no split)
    struct Launch {
        int first, last, cls, map_off;
    };
    std::vector<Launch> launches;
    h_map_.clear();
    int first = 0;
    while (first < count) {
        const int cls = launch_class(active_[static_cast<size_t>(order[static_cast<size_t>(first)])]);
        int last = first;
        while (last < count && launch_class(active_[static_cast<size_t>(order[static_cast<size_t>(last)])]) == cls) ++last;
        int a = first;
        while (a < last) {
            // a cooperative launch takes split jobs (and then single ones) while the
            // grid stays co-resident; the rest go in one ordinary launch
            int b = a;
            int map_off = -1;
            if (states[static_cast<size_t>(a)].parts > 1) {
                int blocks = 0;
                while (b < last) {
                    const int p = std::max(1, states[static_cast<size_t>(b)].parts);
                    if (blocks + p > coresident_) break;
                    blocks += p;
                    ++b;
                }
                std::vector<int2> map;
                general_blockmap(states.data() + a, b - a, map);
                map_off = static_cast<int>(h_map_.size());
                h_map_.insert(h_map_.end(), map.begin(), map.end());
            } else {
                b = last;
            }
            launches.push_back(Launch{a, b, cls, map_off});
            a = b;
        }
        first = last;
    }
    if (!h_map_.empty()) {
        if (static_cast<int>(h_map_.size()) > map_cap_) {
            (void)hipFree(d_map_);
            d_map_ = nullptr;
            map_cap_ = std::max(static_cast<int>(h_map_.size()), 2 * map_cap_);
            check_hip(hipMalloc(&d_map_, sizeof(int2) * static_cast<size_t>(map_cap_)), "hipMalloc block map");
        }
        check_hip(hipMemcpyAsync(d_map_, h_map_.data(), sizeof(int2) * h_map_.size(), hipMemcpyHostToDevice, stream_),
                  "copy block map");
    }
    for (const Launch& l : launches) {
        if (l.cls == 2 || l.cls == 3)
            launch_general(states.data() + l.first, l.last - l.first, d_states_ + l.first, d_moves_ + l.first,
                           fp32_ != 0, nullptr, 0, stream_, l.map_off >= 0 ? d_map_ + l.map_off : nullptr);
        else
            launch_advance(states.data() + l.first, l.last - l.first, d_states_ + l.first, d_moves_ + l.first,
                           fp32_ != 0, nullptr, 0, stream_);
    }
    check_hip(hipMemcpyAsync(states.data(), d_states_, sizeof(ReplicaState) * count, hipMemcpyDeviceToHost, stream_),
              "copy states back");
    check_hip(hipStreamSynchronize(stream_), "server step");
    for (int r = 0; r < count; ++r) {
        Slot& s = active_[static_cast<size_t>(order[static_cast<size_t>(r)])];
        moves_total_ += static_cast<long long>(states[static_cast<size_t>(r)].moves - s.state.moves);
        s.state = states[static_cast<size_t>(r)];
    }
    for (size_t k = active_.size(); k-- > 0;) {
        const Slot& s = active_[k];
        const JobRecord& j = *jobs_[static_cast<size_t>(s.job)];
        if (static_cast<long long>(s.state.moves) >= j.moves_total || s.state.error) retire(k);
    }
}

} // namespace fmc
