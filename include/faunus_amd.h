/*
 * faunus_amd — C ABI of the MI355X energy engine for Faunus's Metropolis MC.
 *
 * The engine replaces the reference's nonbonded and Ewald energy terms behind
 * the Energy::EnergyTerm / Hamiltonian surface (reference: src/externalpotential.h:21-43,
 * src/energy.h:1954-1977).  Plain pointers and sizes only; every call returns 0 on
 * success and a non-zero status on error, with the message in fmc_last_error()
 * (thread-local).  No exception crosses this boundary.  +inf / NaN energies are
 * returned as values, exactly as the reference's terms return them.
 *
 * Energies are in kT, lengths in angstrom, charges in e.  Term energies are
 * reported in Hamiltonian order: the `energy` list order, with [selfenergy] just
 * before nonbonded and [ewald] just after it (src/energy.cpp:1171-1211; self energy
 * is added by PairEnergy::addPairPotentialSelfEnergy, src/energy.h:462-468), and
 * `bonded` (intra-molecular harmonic bonds, src/energy.cpp:974-1124) where listed.
 */
#ifndef FAUNUS_AMD_H
#define FAUNUS_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fmc_engine fmc_engine;

#define FMC_STATE_ACCEPTED 0
#define FMC_STATE_TRIAL 1

/*
 * Flat form of Faunus::Change (src/space.h:29-75): flags plus a list of group
 * changes; relative_atom_indices of group g are indices[offsets[g] .. offsets[g+1]).
 */
typedef struct fmc_change {
    int everything;
    int volume_change;
    int matter_change;
    int num_groups;
    const int* group_index;         /* Change::GroupChange::group_index */
    const unsigned char* all;       /* Change::GroupChange::all */
    const unsigned char* internal;  /* Change::GroupChange::internal */
    const int* offsets;             /* num_groups + 1 */
    const int* indices;             /* relative_atom_indices, sorted per group */
} fmc_change;

/* Thread-local message of the last failed call. */
const char* fmc_last_error(void);

/* Number of visible HIP devices (0 means the engine cannot run). */
int fmc_device_count(int* count);

/*
 * Build an engine from the reference's input JSON: {"temperature", "geometry",
 * "atomlist", "moleculelist", "energy": [...]} — the same keys the reference parses in
 * Hamiltonian::Hamiltonian (src/energy.cpp:1171-1211), Hamiltonian::createEnergy
 * (1280-1327), EwaldData(json) (28-59) and NewCoulombGalore::from_json
 * (src/potentials.cpp:1628-1699).  Replaces constructing the Nonbonded and Ewald
 * EnergyTerms.  Unsupported keys or terms are an error.
 */
int fmc_create(const char* config_json, fmc_engine** out);
int fmc_destroy(fmc_engine* engine);

/* Space::particles (src/particle.h:220-227): positions (3n), charges, atom ids. */
int fmc_set_particles(fmc_engine* engine, int n, const double* xyz, const double* charge, const int* atom_id);

/*
 * A reference state file: the Space JSON written by Space::to_json and read by
 * MetropolisMonteCarlo::restore (src/montecarlo.cpp:118-137; src/space.cpp:449-500;
 * src/group.cpp:231-253).  Reads the cuboid box, the particles (pos, q, id) and the
 * groups (molecule id, size, mass centre, laid end to end).  fmc_parse_state fills
 * whichever arrays are non-null (call it first with nulls for n and num_groups): box[3],
 * xyz[3n], charge[n], atom_id[n], begin/size/molid[num_groups], mass_center[3 num_groups].
 * fmc_load_state checks the box against the input's geometry and sets the engine's
 * particles and groups (fmc_init follows).  Groups with inactive particles
 * (capacity > size) and states without groups are errors, as is any mismatch between
 * the group sizes and the particle count (the reference's "load error").  The random
 * generator states in the file are not used: proposals come from the caller or Philox.
 */
int fmc_parse_state(const char* state_json, int* n, int* num_groups, double* box, double* xyz, double* charge,
                    int* atom_id, int* begin, int* size, int* molid, double* mass_center);
int fmc_load_state(fmc_engine* engine, const char* state_json);

/* Space::groups (src/group.h:177-194): contiguous ranges in particle order. */
int fmc_set_groups(fmc_engine* engine, int num_groups, const int* begin, const int* size, const int* molid,
                   const double* mass_center);

/*
 * MetropolisMonteCarlo::init (src/montecarlo.cpp:43-71) / Ewald::init
 * (src/energy.cpp:554-558): k-vectors, A_k, full Q(k); both states aligned.
 * Writes the initial total energy (Change::everything) if initial_energy != NULL.
 */
int fmc_init(fmc_engine* engine, double* initial_energy);

int fmc_num_terms(fmc_engine* engine, int* count);
int fmc_term_name(fmc_engine* engine, int index, char* buffer, int length);

/*
 * Trial coordinates of the particles a move displaced (absolute indices, 3 doubles
 * each) — the part of the trial Space that Move::move (src/move.cpp:65-76) mutated.
 */
int fmc_stage_trial(fmc_engine* engine, int count, const int* particle_index, const double* xyz);

/*
 * Trial mass centre of the moved molecular group (Group::mass_center of the trial
 * Space after Group::translate/rotate, src/group.cpp:110-123, or the recomputed centre
 * of AtomicTranslateRotate::translateParticle, src/move.cpp:225-240, and
 * ChainRotationMove::rotate_segment, src/chainmove.cpp:76-104).  Needed when
 * cutoff_g2g is set (GroupCutoff::cut, src/energy.h:761-768); kept on accept.
 */
int fmc_stage_mass_center(fmc_engine* engine, int group, const double* cm);
/* Accepted-state mass centre of a group (molecular groups; atomic groups report 0). */
int fmc_get_mass_center(fmc_engine* engine, int group, double* cm);

/*
 * Ewald::updateState (src/energy.cpp:564-575) followed by Hamiltonian::energy(Change)
 * (src/energy.cpp:1227-1241) on both states in one launch: per-term energies of the
 * affected pairs/terms in the trial state (new_terms) and accepted state (old_terms).
 * This is the fused form of the two energy() calls in
 * MetropolisMonteCarlo::performMove (src/montecarlo.cpp:153-157).
 */
int fmc_delta_u(fmc_engine* engine, const fmc_change* change, double* new_terms, double* old_terms);

/* Hamiltonian::energy(Change) of one state (FMC_STATE_TRIAL or FMC_STATE_ACCEPTED). */
int fmc_energy(fmc_engine* engine, const fmc_change* change, int state, double* terms, double* total);

/*
 * State::sync after the Metropolis decision (src/montecarlo.cpp:170-176,
 * src/energy.cpp:640-658, src/space.cpp:199-239): accept != 0 copies the trial state
 * into the accepted state, accept == 0 restores the trial state.
 */
int fmc_sync(fmc_engine* engine, int accept);

/* Hamiltonian::energy with Change::everything on the accepted state, per term. */
int fmc_total_energy(fmc_engine* engine, double* terms, double* total);

/*
 * Drift check: rebuild Q(k) from positions (Ewald full update, src/energy.cpp:191-206)
 * and compare it with the running Q(k): max |dQ| and relative reciprocal-energy difference.
 */
int fmc_check_q(fmc_engine* engine, double* max_abs_dq, double* rel_reciprocal);

/* Accepted-state positions (3n doubles). */
int fmc_get_positions(fmc_engine* engine, double* xyz);

/*
 * The reference's Metropolis step (MetropolisMonteCarlo::performMove,
 * src/montecarlo.cpp:139-187 with getEnergyChange 193-209 and metropolisCriterion
 * 17-34) for a stream of single-particle translations, one GPU evaluation per move.
 * Inputs per move: group index, index within the group, new position (3), uniform
 * deviate of the acceptance test.  Outputs per move: du (new-old, before the
 * reject reset), accepted flag and per-term energies of both states (may be NULL).
 */
int fmc_replay_translate(fmc_engine* engine, long count, const int* group, const int* index, const double* xyz,
                         const double* uniform, double* du, unsigned char* accepted, double* new_terms,
                         double* old_terms);

/*
 * The same Metropolis step for a stream of multi-atom moves: move n changes the
 * atoms index[offsets[n] .. offsets[n+1]) (absolute, ascending, all in group[n]) to
 * the coordinates xyz (3 per atom); whole[n] != 0 marks a rigid-body change of the
 * whole group (Change all=true, internal=false, as TranslateRotate reports it,
 * src/move.cpp:1676-1680), otherwise the change is internal with relative indices (as
 * AtomicTranslateRotate and PivotMove report it, src/move.cpp:311-313,
 * src/chainmove.cpp:107-121).  mass_center (3 per move, may be NULL) gives the trial
 * mass centre of the moved group.  An empty move counts as accepted with du = 0.
 */
int fmc_replay_moves(fmc_engine* engine, long count, const int* group, const long* offsets, const int* index,
                     const double* xyz, const double* mass_center, const unsigned char* whole, const double* uniform,
                     double* du, unsigned char* accepted);

/*
 * Same Metropolis step for translations given as displacements from the current
 * accepted position: new = boundary(old + displacement) as in
 * AtomicTranslateRotate::translateParticle (src/move.cpp:225-240) and
 * Cuboid::boundary (src/geometry.cpp:91-104).  One GPU evaluation per move.
 */
int fmc_run_translate(fmc_engine* engine, long count, const int* group, const int* index, const double* displacement,
                      const double* uniform, double* du, unsigned char* accepted);

/* Kernel timing of the per-move launch (HIP events on the engine's stream). */
int fmc_set_timing(fmc_engine* engine, int enable);
int fmc_get_timing(fmc_engine* engine, double* total_ms, long* launches);

/*
 * Device-resident Metropolis over many replicas (one workgroup per replica, one
 * persistent launch per call).  Each replica is a system as in fmc_create plus
 * positions and groups; proposals are single-particle translations drawn on the
 * GPU from a per-replica Philox4x32-10 stream (the semantics of
 * AtomicTranslateRotate::_move, src/move.cpp:267-293, and performMove,
 * src/montecarlo.cpp:139-187).  beta = T_ref/T scales every energy (all terms are
 * energies in kT at T_ref); replicas with the same ladder id >= 0 form a
 * parallel-tempering ladder whose temperature labels fmc_batch_swap exchanges
 * (ParallelTempering, src/move.cpp:844-965).  fp32_pairs != 0 evaluates pair terms in
 * FP32 with FP64 accumulation; Q(k), dU and the acceptance test stay FP64.
 * No reference counterpart as one call: it replaces the loop of performMove calls.
 */
typedef struct fmc_batch fmc_batch;
int fmc_batch_create(int fp32_pairs, fmc_batch** out);
/*
 * Cell-list grid of the replicas added afterwards: cells of at least pair range / d with
 * a (2d+1)^3 stencil (d = 1: 27 cells, the default; d = 2: 125 smaller cells, about 2.4x
 * fewer candidates).  d = 2 is walked by the general move-set kernel only; translate-only
 * replicas and the per-move path refuse it.  No reference counterpart (the reference
 * loops over all pairs, src/energy.h:1155-1224).
 */
int fmc_batch_set_cell_divisions(fmc_batch* batch, int divisions);
/*
 * Members per cell of the cell lists of the replicas added afterwards (0, the default:
 * max(2 x the fullest initial cell, 4 x the mean + 32)).  A fixed capacity must hold the
 * fullest initial cell.  A move that would overfill a cell is committed and the replica
 * stops after it with error 1 (fmc_batch_replica_info); its cell list is rebuilt when its
 * state is set again (fmc_batch_set_replica_state).  No reference counterpart (the
 * reference has no cell list on this path).
 */
int fmc_batch_set_cell_capacity(fmc_batch* batch, int capacity);
/*
 * Checkpoint and resume of one replica, the device-resident counterpart of SaveState
 * (src/analysis.cpp:645-680: positions, group mass centres and the move random-number
 * state) and MetropolisMonteCarlo::restore followed by init (src/montecarlo.cpp:118-137).
 * The random state is the replica's Philox move counter.  get: positions (3n), mass
 * centres (3 per group; may be NULL), move counter and accepted count.  set: the same
 * (NULL mass centres are recomputed from the positions); Q(k), the cell list (automatic
 * capacity) and the running energy (Change::everything) are rebuilt, the error flag is
 * cleared.  A chain run M moves, checkpointed, resumed and run M more is the chain of
 * 2M moves.
 */
int fmc_batch_get_replica_state(fmc_batch* batch, int replica, double* xyz, double* mass_center,
                                unsigned long long* moves, unsigned long long* accepted);
int fmc_batch_set_replica_state(fmc_batch* batch, int replica, const double* xyz, const double* mass_center,
                                unsigned long long moves, unsigned long long accepted);
int fmc_batch_destroy(fmc_batch* batch);
int fmc_batch_add_replica(fmc_batch* batch, const char* config_json, int n, const double* xyz, const double* charge,
                          const int* atom_id, int num_groups, const int* begin, const int* size, const int* molid,
                          double beta, unsigned long long seed, int ladder, int* replica);
/* Advance replica r by moves[r] trial moves; trace_capacity > 0 records that many moves per replica. */
/*
 * Replica with a general move set: num_kinds kinds (at most 8), each one of
 *   0 translate one atom of molecule type kind_molid (AtomicTranslateRotate,
 *     src/move.cpp:225-345; the atom's dp, else dp[k]),
 *   1 translate by dp[k] and rotate by dprot[k] a molecule of that type
 *     (TranslateRotate, src/move.cpp:1618-1689),
 *   2 pivot a chain of that type about a random harmonic bond by dprot[k]
 *     (PivotMove, src/chainmove.cpp:216-268),
 * picked with probability weight[k] / sum(weight) as MoveCollection::sample does
 * (src/move.cpp:513-525).  Proposals come from the replica's Philox stream (counters
 * (move, replica, 0|1|2)); the Hamiltonian may hold a bonded term.  A batch holds
 * either replicas added this way or translate-only replicas; cutoff_g2g must be unset
 * and a molecule type moved by kind 1 takes no kind 0 or 2 moves.
 */
int fmc_batch_add_replica_moves(fmc_batch* batch, const char* config_json, int n, const double* xyz,
                                const double* charge, const int* atom_id, int num_groups, const int* begin,
                                const int* size, const int* molid, double beta, unsigned long long seed, int ladder,
                                int num_kinds, const int* kind, const int* kind_molid, const unsigned* weight,
                                const double* dp, const double* dprot, int* replica);
int fmc_batch_run(fmc_batch* batch, const long long* moves, int trace_capacity);
/*
 * Advance a batch of ONE replica by `moves` moves with `groups` cooperating
 * workgroups (workgroup 0 walks the pair stencils and commits, the others own slices of
 * the k-vectors; every `stride`-th block of the grid takes part, so with stride = number
 * of XCDs the participants share one XCD).  The workgroups exchange per-move partial
 * sums through device memory, each publishing its share of the next move right after
 * the previous decision; a wait that
 * exceeds its spin bound stops the run and sets the error read by
 * fmc_batch_split_error instead of hanging.  A replica with a general move set
 * (fmc_batch_add_replica_moves, no cell list, 1..64 groups) runs the split form of the
 * general kernel: each workgroup also takes a slice of the pair sums.
 */
int fmc_batch_run_split(fmc_batch* batch, long long moves, int groups, int stride, int trace_capacity);
/*
 * fmc_batch_run_split for a translate-only replica without allocating device memory.
 * No counterpart in the reference: for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The exchange
 * slots and the time-out flag live in a static workspace of 16.3 KB that each GPU keeps
 * for the life of the process (slots for at most 64 workgroups: more is an error); it is
 * cleared by a fill kernel on the batch's stream before every launch.  The batch must
 * already hold its device state (after any earlier fmc_batch_run* or
 * fmc_batch_replica_info call), and no trace is recorded.  Calls that overlap on one GPU
 * share the workspace and race.  The chain equals fmc_batch_run_split's at the same
 * workgroup count, bit for bit; without the private position copies that
 * fmc_batch_run_split allocates, workgroup 0 also sends, with each move's pair change,
 * the particle of the move after next, so it runs a little slower.  Errors read through fmc_batch_split_error / fmc_batch_replica_info.
 */
int fmc_batch_run_split_static(fmc_batch* batch, long long moves, int groups, int stride);
/*
 * As fmc_batch_run_split, with up to `depth` (1..4) moves evaluated ahead of their
 * predecessors' decisions: each move's partial sums carry the Ewald cross terms with
 * the pending moves' dQ(k) and the pair terms with each pending particle at its old
 * and new position, so every decision is exactly the one-move-at-a-time Metropolis
 * decision while the exchange between workgroups overlaps the next moves' work.
 */
int fmc_batch_run_pipe(fmc_batch* batch, long long moves, int groups, int depth, int trace_capacity);
int fmc_batch_split_error(fmc_batch* batch, int* error);
/* Wait for the batch; kernel time of the last run (ms, HIP events) if timing is on. */
int fmc_batch_sync(fmc_batch* batch, double* kernel_ms);
int fmc_batch_set_timing(fmc_batch* batch, int enable);
/* One odd/even round of neighbour exchanges on every ladder. */
int fmc_batch_swap(fmc_batch* batch, unsigned long long seed, unsigned long long round, long long* attempted,
                   long long* accepted);
/*
 * One replica-exchange interval as a HIP graph (ParallelTempering every interval,
 * src/move.cpp:844-965): the persistent launch advancing replica r by moves[r] moves,
 * then one odd/even round of neighbour swaps on every ladder, the round taken from a
 * device counter that starts at round0, so that replays alternate the pairings.  The
 * ladders' rung order and the swap counters stay on the device.  fmc_batch_replay
 * launches the captured interval `count` times (asynchronously; fmc_batch_sync waits);
 * fmc_batch_swap_counts gives the swaps attempted and accepted so far.  Replicas shared
 * by several workgroups are not captured.  No reference counterpart as one call.
 */
int fmc_batch_capture_interval(fmc_batch* batch, const long long* moves, unsigned long long seed,
                               unsigned long long round0);
int fmc_batch_replay(fmc_batch* batch, int count);
int fmc_batch_swap_counts(fmc_batch* batch, long long* attempted, long long* accepted);
/*
 * Free-running ladders (ParallelTempering every interval, src/move.cpp:844-965): one
 * cooperative launch in which replica r runs intervals[r] intervals of moves[r] moves;
 * after each interval the rungs of a ladder meet at a barrier in device memory and the
 * last to arrive runs that ladder's odd/even round of swaps (round round0 + interval,
 * the rule and deviates of fmc_batch_capture_interval).  A ladder waits only for its own
 * rungs, so ladders of different sizes run their intervals at their own pace; the rungs
 * of one ladder must share the interval count.  Translate-only replicas of one workgroup
 * size, all resident at once (an error otherwise).  Blocks until done; the swaps count
 * in fmc_batch_swap_counts.  No reference counterpart as one call.
 */
int fmc_batch_run_ladders(fmc_batch* batch, const long long* moves, const int* intervals, unsigned long long seed,
                          unsigned long long round0);
/*
 * Tile autotuner: times each candidate launch shape of the batch on trial_moves moves
 * per replica (from checkpoints restored afterwards) and keeps the fastest for the whole
 * batch.  Candidates: 256 or 512 threads per workgroup for translate-only replicas
 * (512 when every replica has K <= 4096); cells of at least the pair range or half of it
 * (cell divisions 1 or 2) for move-set replicas with a cell list.  Writes up to
 * `capacity` candidates with their kernel time per replica-move (microseconds) and the
 * chosen value; num_candidates 0 means the batch has no choice.  No reference counterpart.
 */
int fmc_batch_autotune(fmc_batch* batch, long long trial_moves, int capacity, int* num_candidates, int* values,
                       double* us_per_move, int* chosen);
int fmc_batch_replica_info(fmc_batch* batch, int replica, double* energy, double* beta, unsigned long long* moves,
                           unsigned long long* accepted, int* error);
int fmc_batch_positions(fmc_batch* batch, int replica, double* xyz);
int fmc_batch_trace(fmc_batch* batch, int replica, int count, int* index, unsigned char* accepted, double* du,
                    double* xyz);
/* Change::everything energy of a replica recomputed from its positions, per term. */
int fmc_batch_total_energy(fmc_batch* batch, int replica, double* terms, double* total);
/*
 * On-device observables.
 * fmc_batch_sample_energy: from now on every replica adds its running energy (kT at
 * T_ref) and its square to device accumulators after each move whose index + 1 is a
 * multiple of `every` (0: off), skipping non-finite energies, as SystemEnergy::_sample
 * (src/analysis.cpp:339-347) does at its sampling interval; the accumulators are cleared.
 * The one-workgroup, split and general kernels sample; fmc_batch_run_pipe refuses to run
 * while sampling is on.
 * fmc_batch_energy_stats reads a replica's sample count, sum and sum of squares.
 * fmc_batch_rdf adds the pair-distance histogram of atom types a and b of a replica's
 * current configuration (a == b: pairs i < j) to counts[nbins], bin floor(r / dr), as
 * AtomRDF::_sample (src/analysis.cpp:1556-1600) fills its Equidistant2DTable
 * (src/aux/equidistant_table.h:32-39); normalisation to g(r) stays with the caller.
 */
int fmc_batch_sample_energy(fmc_batch* batch, long long every);
int fmc_batch_energy_stats(fmc_batch* batch, int replica, unsigned long long* samples, double* sum, double* sum_sq);
int fmc_batch_rdf(fmc_batch* batch, int replica, int type_a, int type_b, double dr, int nbins,
                  unsigned long long* counts);
int fmc_batch_check_q(fmc_batch* batch, int replica, double* max_abs_dq, double* rel_reciprocal);

/*
 * Serving loop with continuous batching.  Each job is one translational Metropolis run
 * (the loop of MetropolisMonteCarlo::move, src/montecarlo.cpp:87-106, over a system of
 * the form fmc_create accepts) of `moves` moves with Philox proposals keyed by `seed`;
 * the server has no single reference counterpart: it replaces many independent
 * `faunus` processes.  Jobs are admitted first-in first-out while their estimated HBM
 * footprint fits `hbm_budget_bytes` and fewer than `max_slots` jobs are resident;
 * fmc_server_step advances every resident job by up to `quantum`*1000/(n + 2K) moves
 * (n particles, K k-vectors) in one persistent launch per cell-list kind and retires
 * finished jobs.  A job's result does not depend on which other jobs share its launches.
 * Job states: 0 queued, 1 resident, 2 done, 3 failed (reason from fmc_server_job_error).
 */
typedef struct fmc_server fmc_server;
int fmc_server_create(int fp32_pairs, long long hbm_budget_bytes, int max_slots, fmc_server** out);
int fmc_server_destroy(fmc_server* server);
/*
 * Split policy (before the first submit): a job with a general move set and no cell
 * list is shared by ceil(n / atoms_per_workgroup) workgroups, at most max_workgroups
 * (1..64) and at least enough to hold its k-vectors at 16 per thread; 0 atoms keeps
 * one workgroup per job.  The workgroups of a job split its k-vectors and pair sums
 * and exchange partial sums every move in a cooperative launch.  The count depends on
 * the job alone, so a job's chain stays the same in any mix (it differs from the
 * one-workgroup chain only by the summation order of the partial sums).
 */
int fmc_server_set_split(fmc_server* server, int atoms_per_workgroup, int max_workgroups);
int fmc_server_submit(fmc_server* server, const char* config_json, int n, const double* xyz, const double* charge,
                      const int* atom_id, int num_groups, const int* begin, const int* size, const int* molid,
                      long long moves, unsigned long long seed, int* job);
/* A job with a general move set (same move kinds as fmc_batch_add_replica_moves). */
int fmc_server_submit_moves(fmc_server* server, const char* config_json, int n, const double* xyz,
                            const double* charge, const int* atom_id, int num_groups, const int* begin,
                            const int* size, const int* molid, long long moves, unsigned long long seed,
                            int num_kinds, const int* kind, const int* kind_molid, const unsigned* weight,
                            const double* dp, const double* dprot, int* job);
int fmc_server_step(fmc_server* server, long long quantum, int* resident, int* queued);
/* Times are seconds since the server was created (host steady clock). */
int fmc_server_job(fmc_server* server, int job, int* state, double* energy, unsigned long long* moves,
                   unsigned long long* accepted, double* t_submit, double* t_admit, double* t_done);
int fmc_server_job_positions(fmc_server* server, int job, double* xyz);
/* Reason a job failed (empty string otherwise). */
int fmc_server_job_error(fmc_server* server, int job, char* buffer, int length);
int fmc_server_status(fmc_server* server, double* seconds, long long* moves_done, long long* bytes_in_use);

/*
 * Streaming copy of `bytes` bytes (read once, written once), `repeats` launches: the
 * known-size reference for scaling HBM counter readings.  No reference counterpart.
 */
int fmc_copy_probe(long long bytes, int repeats, double* ms_per_launch);

/* Andrea spline of the Coulomb short-range function built by the engine (for tests). */
int fmc_coulomb_spline(fmc_engine* engine, int capacity, double* knots, double* coeff, int* num_knots);
int fmc_num_kvectors(fmc_engine* engine, int* count);
/*
 * nonbonded_splined table of a type pair built by the engine (SplinedPotential's
 * KnotData, src/potentials.h:773-781; for tests): knots in r^2 (ascending), 6
 * coefficients per interval, rmin^2, rmax^2 and the hard-sphere flag.
 */
int fmc_splined_table(fmc_engine* engine, int type_a, int type_b, int capacity, double* knots, double* coeff,
                      int* num_knots, double* rmin2, double* rmax2, int* hardsphere);
/*
 * The nonbonded term's pair energy (PairEnergy::potential, src/energy.h:428-438) of
 * atom types a, b with charges qa, qb at n squared distances, evaluated on the GPU by
 * the device functions of the FP64 forms (u64) and of the FP32-pair forms (u32).  For
 * tests of the device pair arithmetic against the oracle.
 */
int fmc_pair_probe(fmc_engine* engine, int type_a, int type_b, double qa, double qb, int n, const double* r2,
                   double* u64, double* u32);

#ifdef __cplusplus
}
#endif

#endif /* FAUNUS_AMD_H */
