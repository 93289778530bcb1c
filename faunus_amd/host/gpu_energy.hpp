// C++ host side of the drop-in boundary: EnergyTerm-shaped adapters over the C ABI
// (include/faunus_amd.h).  They mirror the reference's Energy::EnergyTerm interface
// (src/externalpotential.h:21-43) — energy(Change), updateState(Change),
// sync(EnergyTerm*, Change), init(), state, name — so that the reference's
// Nonbonded and Ewald terms (src/energy.h:1502-1598, 266-285) can be replaced by
// GpuNonbonded / GpuEwald objects sharing one GPU engine, and its
// MetropolisMonteCarlo::performMove (src/montecarlo.cpp:139-187) keeps its shape.
//
// Error convention: the C ABI returns a status; these classes throw
// std::runtime_error with fmc_last_error() (the reference throws on configuration and
// sync errors).  +inf / NaN energies pass through as values.
#pragma once
#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/faunus_amd.h"

namespace faunus_amd {

inline void check(int status) {
    if (status != 0) throw std::runtime_error(fmc_last_error());
}

// Faunus::Change (src/space.h:29-75), the fields this path reads.
struct Change {
    struct GroupChange {
        int group_index = 0;
        bool all = false;
        bool internal = false;
        std::vector<int> relative_atom_indices;
    };
    bool everything = false, volume_change = false, matter_change = false;
    std::vector<GroupChange> groups;
    explicit operator bool() const { return everything || volume_change || matter_change || !groups.empty(); }
};

// Holds the C view of a Change for the duration of a call.
class ChangeView {
  public:
    explicit ChangeView(const Change& c) {
        for (const auto& g : c.groups) {
            index_.push_back(g.group_index);
            all_.push_back(g.all ? 1 : 0);
            internal_.push_back(g.internal ? 1 : 0);
            offsets_.push_back(static_cast<int>(indices_.size()));
            indices_.insert(indices_.end(), g.relative_atom_indices.begin(), g.relative_atom_indices.end());
        }
        offsets_.push_back(static_cast<int>(indices_.size()));
        if (indices_.empty()) indices_.push_back(0);
        view_.everything = c.everything ? 1 : 0;
        view_.volume_change = c.volume_change ? 1 : 0;
        view_.matter_change = c.matter_change ? 1 : 0;
        view_.num_groups = static_cast<int>(index_.size());
        view_.group_index = index_.data();
        view_.all = all_.data();
        view_.internal = internal_.data();
        view_.offsets = offsets_.data();
        view_.indices = indices_.data();
    }
    const fmc_change* get() const { return &view_; }

  private:
    std::vector<int> index_, offsets_, indices_;
    std::vector<unsigned char> all_, internal_;
    fmc_change view_{};
};

// One GPU engine shared by the trial and accepted Hamiltonians of an MC run: it holds
// both states (accepted configuration + the pending trial delta) and caches the
// last fused evaluation so that the trial and accepted energy() calls of one move
// cost one launch.
class Engine {
  public:
    explicit Engine(const std::string& config_json) {
        fmc_engine* e = nullptr;
        check(fmc_create(config_json.c_str(), &e));
        handle_.reset(e);
        int n = 0;
        check(fmc_num_terms(e, &n));
        for (int i = 0; i < n; ++i) {
            char buf[64];
            check(fmc_term_name(e, i, buf, sizeof(buf)));
            names_.emplace_back(buf);
        }
    }
    fmc_engine* get() const { return handle_.get(); }
    const std::vector<std::string>& term_names() const { return names_; }
    int term_index(const std::string& name) const {
        for (size_t i = 0; i < names_.size(); ++i)
            if (names_[i] == name) return static_cast<int>(i);
        return -1;
    }
    // Energies of both states for `change` (one launch per move generation).
    void evaluate(const Change& change, unsigned long generation) {
        if (generation == cached_generation_) return;
        ChangeView v(change);
        new_terms_.assign(names_.size(), 0.0);
        old_terms_.assign(names_.size(), 0.0);
        check(fmc_delta_u(get(), v.get(), new_terms_.data(), old_terms_.data()));
        cached_generation_ = generation;
    }
    double term(bool trial, int index) const { return trial ? new_terms_[static_cast<size_t>(index)] : old_terms_[static_cast<size_t>(index)]; }
    void invalidate() { cached_generation_ = ~0ul; }
    // State::sync of the move `generation` (montecarlo.cpp:170-176): accept copies the trial
    // state into the accepted one, reject restores the trial state.  Every term of both
    // Hamiltonians calls it; the engine, which holds both states, commits once per move.
    void commit(bool accept, unsigned long generation) {
        if (generation == committed_generation_) {
            if (accept != committed_accept_) throw std::runtime_error("sync error: terms disagree on the decision");
            return;
        }
        check(fmc_sync(get(), accept ? 1 : 0));
        committed_generation_ = generation;
        committed_accept_ = accept;
    }

  private:
    struct Deleter {
        void operator()(fmc_engine* e) const { fmc_destroy(e); }
    };
    std::unique_ptr<fmc_engine, Deleter> handle_;
    std::vector<std::string> names_;
    std::vector<double> new_terms_, old_terms_;
    unsigned long cached_generation_ = ~0ul;
    unsigned long committed_generation_ = ~0ul;
    bool committed_accept_ = false;
};

// Energy::EnergyTerm shape (src/externalpotential.h:21-43).
class EnergyTerm {
  public:
    enum class MonteCarloState { ACCEPTED, TRIAL, NONE };
    MonteCarloState state = MonteCarloState::NONE;
    std::string name;
    virtual ~EnergyTerm() = default;
    virtual double energy(const Change& change) = 0;
    virtual void updateState(const Change&) {}
    virtual void sync(EnergyTerm*, const Change&) {}
    virtual void init() {}
};

// The GPU term `term_name` ("selfenergy", "nonbonded", "ewald" or "bonded") of one state.
class GpuTerm : public EnergyTerm {
  public:
    GpuTerm(std::shared_ptr<Engine> engine, std::string term_name, const unsigned long* generation)
        : engine_(std::move(engine)), index_(engine_->term_index(term_name)), generation_(generation) {
        if (index_ < 0) throw std::runtime_error("engine has no term " + term_name);
        name = std::move(term_name);
    }
    // Nonbonded::energy (src/energy.h:1561-1577) / Ewald::energy (src/energy.cpp:581-587)
    double energy(const Change& change) override {
        if (!change) return 0.0;
        if (change.everything) {
            std::vector<double> t(engine_->term_names().size());
            double total = 0;
            check(fmc_total_energy(engine_->get(), t.data(), &total));
            return t[static_cast<size_t>(index_)];
        }
        engine_->evaluate(change, *generation_);
        return engine_->term(state == MonteCarloState::TRIAL, index_);
    }
    // EnergyTerm::sync(other, change): the accepted term syncing from the trial one means
    // the move was accepted, the trial term syncing from the accepted one that it was
    // rejected (MetropolisMonteCarlo::performMove, src/montecarlo.cpp:170-176).
    void sync(EnergyTerm* other, const Change&) override {
        auto* o = dynamic_cast<GpuTerm*>(other);
        if (!o || o->name != name || o->engine_ != engine_) throw std::runtime_error("sync error");
        if (state == other->state) throw std::runtime_error("sync error: both terms in one state");
        engine_->commit(state == MonteCarloState::ACCEPTED, *generation_);
    }

  protected:
    std::shared_ptr<Engine> engine_;
    int index_;
    const unsigned long* generation_;
};

// Nonbonded<PairEnergy, GroupPairing> (src/energy.h:1502-1598): the affected-pair energy of
// one state; updateState has nothing to do (the pair sum reads the Space).
class GpuNonbonded : public GpuTerm {
  public:
    GpuNonbonded(std::shared_ptr<Engine> engine, const unsigned long* generation)
        : GpuTerm(std::move(engine), "nonbonded", generation) {}
};

// Ewald (src/energy.h:266-285; src/energy.cpp:539-662): updateState applies the partial
// update of the trial structure factors (energy.cpp:564-575); the engine fuses it with
// both states' energies in one launch, which this call starts.
class GpuEwald : public GpuTerm {
  public:
    GpuEwald(std::shared_ptr<Engine> engine, const unsigned long* generation)
        : GpuTerm(std::move(engine), "ewald", generation) {}
    void updateState(const Change& change) override {
        if (change && !change.everything && state == MonteCarloState::TRIAL) engine_->evaluate(change, *generation_);
    }
};

// ParticleSelfEnergy (src/externalpotential.h:194-198) and Bonded (src/energy.cpp:974-1124).
class GpuSelfEnergy : public GpuTerm {
  public:
    GpuSelfEnergy(std::shared_ptr<Engine> engine, const unsigned long* generation)
        : GpuTerm(std::move(engine), "selfenergy", generation) {}
};
class GpuBonded : public GpuTerm {
  public:
    GpuBonded(std::shared_ptr<Engine> engine, const unsigned long* generation)
        : GpuTerm(std::move(engine), "bonded", generation) {}
};

// Hamiltonian (src/energy.h:1954-1977) over the GPU terms of one state, in the engine's
// term order: energy with the ordered sum and early break (src/energy.cpp:1227-1241),
// updateState (1248-1252) and sync (1254-1268).
class GpuHamiltonian {
  public:
    GpuHamiltonian(std::shared_ptr<Engine> engine, EnergyTerm::MonteCarloState state, const unsigned long* generation,
                   double maximum_allowed_energy = INFINITY)
        : max_energy_(maximum_allowed_energy) {
        for (const auto& n : engine->term_names()) {
            if (n == "nonbonded") terms_.push_back(std::make_unique<GpuNonbonded>(engine, generation));
            else if (n == "ewald") terms_.push_back(std::make_unique<GpuEwald>(engine, generation));
            else if (n == "selfenergy") terms_.push_back(std::make_unique<GpuSelfEnergy>(engine, generation));
            else if (n == "bonded") terms_.push_back(std::make_unique<GpuBonded>(engine, generation));
            else throw std::runtime_error("unknown engine term " + n);
            terms_.back()->state = state;
        }
    }
    double energy(const Change& change) {
        latest_.clear();
        double sum = 0.0;
        for (auto& t : terms_) {
            const double e = t->energy(change);
            latest_.push_back(e);
            sum += e;
            if (e >= max_energy_ || std::isnan(e)) break;
        }
        return sum;
    }
    void updateState(const Change& change) {
        for (auto& t : terms_) t->updateState(change);
    }
    void sync(GpuHamiltonian& other, const Change& change) {
        if (other.terms_.size() != terms_.size()) throw std::runtime_error("hamiltonian mismatch");
        latest_ = other.latest_;
        for (size_t i = 0; i < terms_.size(); ++i) terms_[i]->sync(other.terms_[i].get(), change);
    }
    size_t size() const { return terms_.size(); }
    EnergyTerm& operator[](size_t i) { return *terms_[i]; }
    const std::vector<double>& latest_energies() const { return latest_; }

  private:
    std::vector<std::unique_ptr<GpuTerm>> terms_;
    std::vector<double> latest_;
    double max_energy_;
};

// MetropolisMonteCarlo::performMove (src/montecarlo.cpp:139-187) over the GPU
// Hamiltonians, for moves that displace particles: the caller provides the Change,
// the moved particles' trial coordinates and the acceptance deviate.
class Metropolis {
  public:
    explicit Metropolis(const std::string& config_json)
        : engine_(std::make_shared<Engine>(config_json)),
          trial_(engine_, EnergyTerm::MonteCarloState::TRIAL, &generation_),
          accepted_(engine_, EnergyTerm::MonteCarloState::ACCEPTED, &generation_) {}
    Engine& engine() { return *engine_; }

    double init(int n, const double* xyz, const double* charge, const int* atom_id, const std::vector<int>& group_begin,
                const std::vector<int>& group_size, const std::vector<int>& group_molid) {
        check(fmc_set_particles(engine_->get(), n, xyz, charge, atom_id));
        check(fmc_set_groups(engine_->get(), static_cast<int>(group_begin.size()), group_begin.data(),
                             group_size.data(), group_molid.data(), nullptr));
        double e = 0;
        check(fmc_init(engine_->get(), &e));
        initial_energy_ = e;
        return e;
    }

    // returns true if accepted; du is the energy change before the reject reset
    bool perform(const Change& change, const std::vector<int>& moved, const std::vector<double>& xyz, double uniform,
                 double* du_out = nullptr) {
        ++generation_;
        // Move::move changed the trial Space
        check(fmc_stage_trial(engine_->get(), static_cast<int>(moved.size()), moved.data(), xyz.data()));
        trial_.updateState(change);
        const double new_energy = trial_.energy(change);
        const double old_energy = accepted_.energy(change);
        const double du = energy_change(new_energy, old_energy);
        const bool accept = metropolis(du, uniform);
        if (accept) accepted_.sync(trial_, change); // state->sync(*trial_state, change)
        else trial_.sync(accepted_, change);        // trial_state->sync(*state, change)
        if (accept) sum_of_energy_changes_ += du;
        if (du_out) *du_out = du;
        return accept;
    }

    // MetropolisMonteCarlo::relativeEnergyDrift (src/montecarlo.cpp:85-99)
    double relative_energy_drift() {
        Change everything;
        everything.everything = true;
        const double e = accepted_.energy(everything);
        const double du = e - initial_energy_;
        if (!std::isfinite(du)) return NAN;
        if (std::fabs(du) <= 2.220446049250313e-16) return 0.0;
        return (e - (initial_energy_ + sum_of_energy_changes_)) /
               (std::fabs(initial_energy_) > 2.220446049250313e-16 ? initial_energy_ : e);
    }

    static double energy_change(double new_energy, double old_energy) { // montecarlo.cpp:193-209
        if (std::isnan(old_energy) && !std::isnan(new_energy)) return -INFINITY;
        if (std::isnan(new_energy)) return INFINITY;
        if (new_energy > 0.0 && std::isinf(new_energy)) return INFINITY;
        const double du = new_energy - old_energy;
        return std::isnan(du) ? 0.0 : du;
    }
    static bool metropolis(double du, double u) { // montecarlo.cpp:17-34
        if (std::isnan(du)) throw std::runtime_error("Metropolis error: energy cannot be NaN");
        if (std::isinf(du) && du < 0.0) return true;
        if (-du > 709.782712893384) return true;
        return u <= std::exp(-du);
    }

  private:
    unsigned long generation_ = 0;
    std::shared_ptr<Engine> engine_;
    GpuHamiltonian trial_, accepted_;
    double initial_energy_ = 0, sum_of_energy_changes_ = 0;
};

} // namespace faunus_amd
