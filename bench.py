"""Benchmark of the Metropolis energy path on one MI355X.

Headline workload (N=1 line): N = 20 000 Na/Cl ions (q=+-1, LJ sigma 3.3/4.4 A,
eps 0.5/0.4 kJ/mol, Lorentz-Berthelot), cubic L = 255 A, Ewald alpha=0.25 A^-1,
Rc=12 A, kmax=20 (K=17 328), the pair energy splined and cut at 11.99 A
(nonbonded_splined, the reference's way to cut LJ), single-particle translations
dp = 2 A, FP64, device-resident Metropolis (one persistent launch per step).  A step
is STEP_MOVES trial moves; the metric is trial moves per second.

Also measured and reported beside it: the plain FP64 one-move-per-launch path on
the same system, the FP32-pair form, the 256-replica 2:1-salt mix (8 compositions x
32-rung temperature ladder: cost-balanced launches, swap intervals of 100 sweeps
replayed from a HIP graph, and free-running ladders of 100-sweep intervals in one
cooperative launch, all after the tile autotuner) with the plain path on the
same per-replica move counts, 64 rigid-water replicas, serving under Poisson arrivals
to 1000 sweeps per job, and CPU baselines (FP64 oracle: one core for the single chain,
one replica or job per core for the batched and serving forms).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the headline kernel (STEP_MOVES moves), from rocprofv3 FETCH_SIZE
# and WRITE_SIZE runs scaled by the copy-probe calibration
# (profiles/mc_split_kernel__nacl20k_kmax20.txt)
HEADLINE_TRAFFIC_BYTES = 378.3e6
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
STEP_MOVES = 20000
PLAIN_MOVES = 2000


T_START = time.perf_counter()


def log(*a):
    print("[%6.1f s]" % (time.perf_counter() - T_START), *a, file=sys.stderr, flush=True)


def cpu_workers():
    """Host cores used by the multi-core CPU baselines: the process's CPU share, at most
    16 (one GPU's share of the host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_replicas_baseline(work, seconds):
    """One independent FP64 oracle chain per host core, all cores at once (the reference's
    recommended way to run replicas, docs/_docs/running.md:131-138): work(w, seconds)
    runs worker w's chain for about `seconds` and returns (moves, elapsed).  The oracle
    library releases the GIL, so threads run the chains in parallel."""
    from concurrent.futures import ThreadPoolExecutor
    cores = cpu_workers()
    t0 = time.perf_counter()
    with ThreadPoolExecutor(cores) as ex:
        res = list(ex.map(lambda w: work(w, seconds), range(cores)))
    wall = time.perf_counter() - t0
    moves = sum(m for m, _ in res)
    return dict(value=moves / wall, cores=cores, moves=moves, wall_s=wall,
                per_core=[m / t for m, t in res])


def nacl_system(seed=1):
    from tests import systems
    return systems.nacl_lj(n_pairs=10000, box=255.0, seed=seed)


def ewald_flops_per_move(K, m=1):
    """Algorithmic flops of the Ewald dU of one move: 2mK(5+2) + 8K plus 2mK sincos
    counted as 20 flops each."""
    return 2 * m * K * 7 + 8 * K + 2 * m * K * 20


def pair_bytes_cells(n, box, cutoff):
    nc = int(box // cutoff)
    return 27.0 * n / nc ** 3 * 36 + 2 * 27 * 8


def check_replicas(b, count, what):
    """Every replica of a batch ended its launches without an error (a full cell, a
    workgroup that did not arrive)."""
    bad = [(r, b.info(r)["error"]) for r in range(count) if b.info(r)["error"]]
    if bad:
        raise RuntimeError(f"{what}: replicas stopped with an error: {bad[:8]}")


def fp32_error(sysd, moves, seed, kinds=None):
    """The FP32-pair form against the FP64 form on the same Philox stream (one replica,
    one workgroup): the largest |du32 - du64| / max(|du64|, |E|) (E the system's total
    energy) over the moves up to and including the first differing decision, and that
    move.  The FP64 form is the one the parity tests hold to the FP64 oracle."""
    import faunus_amd as F
    tr, e0 = {}, None
    for fp32 in (False, True):
        b = F.Batch(fp32_pairs=fp32)
        b.add(sysd, seed=seed, moves=kinds)
        if not fp32:
            e0 = b.info(0)["energy"]
        b.run([moves], trace=moves)
        b.sync()
        check_replicas(b, 1, "fp32 error replica")
        tr[fp32] = b.trace(0, moves)
        b.close()
    d64, d32 = np.asarray(tr[False]["du"], float), np.asarray(tr[True]["du"], float)
    mism = np.nonzero(np.asarray(tr[False]["accepted"]) != np.asarray(tr[True]["accepted"]))[0]
    upto = int(mism[0]) + 1 if len(mism) else moves
    a, b = d64[:upto], d32[:upto]
    fin = np.isfinite(a) & np.isfinite(b)
    if not np.array_equal(np.isinf(a), np.isinf(b)):
        raise RuntimeError("FP32 and FP64 forms disagree on an infinite energy change")
    rel = np.abs(b[fin] - a[fin]) / np.maximum(np.abs(a[fin]), abs(e0))
    return dict(max_rel_err_vs_fp64=float(rel.max()) if len(rel) else 0.0, moves=moves,
                compared=int(upto), first_decision_mismatch=int(mism[0]) if len(mism) else None,
                energy_scale=abs(e0))


def bench_plain_single(sysd, moves, seed=7):
    import faunus_amd as F
    cfg, xyz, q, ids, g = sysd
    eng = F.Engine(cfg)
    eng.set_particles(xyz, q, ids)
    eng.set_groups(g["begin"], g["size"], g["molid"])
    eng.init()
    rng = np.random.default_rng(seed)
    n = len(q)
    dp = cfg["atomlist"][0][next(iter(cfg["atomlist"][0]))]["dp"]
    idx = rng.integers(0, n, size=moves).astype(np.int32)
    v = rng.normal(size=(moves, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    disp = v * (dp * rng.random(moves))[:, None]
    u = rng.random(moves)
    grp = np.zeros(moves, np.int32)
    w = min(200, moves // 5)
    eng.run_translate(grp[:w], idx[:w], disp[:w], u[:w])
    eng.set_timing(True)
    t0 = time.perf_counter()
    du, acc = eng.run_translate(grp[w:], idx[w:], disp[w:], u[w:])
    dt = time.perf_counter() - t0
    kern_ms, launches = eng.timing()
    eng.close()
    return dict(rate=(moves - w) / dt, acceptance=float(acc.mean()), kernel_us=kern_ms * 1e3 / max(1, launches))


def bench_device_single(sysd, args, fp32, steps=None):
    import faunus_amd as F
    b = F.Batch(fp32_pairs=fp32)
    steps = steps or args.steps
    b.add(sysd, seed=2024)
    b.set_timing(True)
    for _ in range(min(args.warmup, 1)):
        b.run([STEP_MOVES])
    b.sync()
    kernel_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        b.run([STEP_MOVES])
        kernel_ms += b.sync()
    dt = time.perf_counter() - t0
    check_replicas(b, 1, "device-resident replica")
    info = b.info(0)
    tot, _ = b.total_energy(0)
    dq, drec = b.check_q(0)
    b.close()
    moves = steps * STEP_MOVES
    return dict(rate=moves / dt, ms_per_step=dt * 1e3 / steps, kernel_ms=kernel_ms, moves=moves,
                acceptance=info["accepted"] / max(1, info["moves"]),
                drift=(info["energy"] - tot) / abs(tot), q_drift=dq)


def bench_split(sysd, args, groups, stride, steps=None):
    """One replica over `groups` cooperating workgroups (k-vector slices)."""
    import faunus_amd as F
    steps = steps or args.steps
    b = F.Batch()
    b.add(sysd, seed=2024)
    b.set_timing(True)
    b.run_split(STEP_MOVES, groups=groups, stride=stride)
    b.sync()
    kernel_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        b.run_split(STEP_MOVES, groups=groups, stride=stride)
        kernel_ms += b.sync()
    dt = time.perf_counter() - t0
    err = b.split_error()
    info = b.info(0)
    tot, _ = b.total_energy(0)
    b.close()
    if err:
        raise RuntimeError("split run stopped with error %d (1: full cell, 2: synchronisation time-out)" % err)
    moves = steps * STEP_MOVES
    return dict(rate=moves / dt, ms_per_step=dt * 1e3 / steps, kernel_ms=kernel_ms, moves=moves,
                acceptance=info["accepted"] / max(1, info["moves"]), drift=(info["energy"] - tot) / abs(tot),
                groups=groups, stride=stride)


def bench_general_split(n_beads=500, n_total=43000, moves=4000, parts=(1, 2, 4, 8, 16, 32, 64), fp32=True):
    """One polyelectrolyte replica (the largest serving job: 500-bead chain with
    counterions and salt, N=43k, Ewald kmax=8, ion and bead translations and pivots)
    advanced by the split general kernel over 1..64 workgroups: kernel time per move."""
    import faunus_amd as F
    from tests import systems
    cfg, xyz, q, ids, g, kinds = systems.polyelectrolyte(n_beads, n_total, seed=5)
    out = {}
    for G in parts:
        b = F.Batch(fp32_pairs=fp32)
        b.add((cfg, xyz, q, ids, g), seed=77, moves=kinds)
        b.set_timing(True)
        b.run_split(200, groups=G)
        b.sync()
        b.run_split(moves, groups=G)
        ms = b.sync()
        err = b.split_error()
        info = b.info(0)
        b.close()
        if err:
            raise RuntimeError("split general run reported an error")
        out[G] = dict(us_per_move=ms * 1e3 / moves, acceptance=info["accepted"] / max(1, info["moves"]))
        log("general split N=%d parts=%d: %.2f us/move" % (len(q), G, out[G]["us_per_move"]))
    return dict(n=len(q), fp32=fp32, moves=moves, per_parts=out)


def bench_pipe(sysd, args, groups, depth, steps=None, fp32=False):
    """One replica over `groups` workgroups with up to `depth` moves in flight."""
    import faunus_amd as F
    steps = steps or args.steps
    b = F.Batch(fp32_pairs=fp32)
    b.add(sysd, seed=2024)
    b.set_timing(True)
    b.run_pipe(STEP_MOVES, groups=groups, depth=depth)
    b.sync()
    kernel_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        b.run_pipe(STEP_MOVES, groups=groups, depth=depth)
        kernel_ms += b.sync()
    dt = time.perf_counter() - t0
    err = b.split_error()
    info = b.info(0)
    tot, _ = b.total_energy(0)
    dq, _ = b.check_q(0)
    b.close()
    if err:
        raise RuntimeError("pipelined run reported a synchronisation time-out")
    moves = steps * STEP_MOVES
    return dict(rate=moves / dt, ms_per_step=dt * 1e3 / steps, kernel_ms=kernel_ms, moves=moves,
                acceptance=info["accepted"] / max(1, info["moves"]), drift=(info["energy"] - tot) / abs(tot),
                q_drift=dq, groups=groups, depth=depth)


def salt_mix(n_comp=8, rungs=32, seed=1):
    """8 compositions x 32 rungs: N ~ U[600, 9000] (multiple of 3), c ~ logU[0.01, 1] M."""
    from tests import systems
    rng = np.random.default_rng(seed)
    comps = []
    for c in range(n_comp):
        n = int(rng.integers(200, 3001)) * 3
        conc = float(10 ** rng.uniform(-2, 0))
        comps.append((n, conc, systems.salt21(n=n, conc=conc, seed=100 + c)))
    temps = np.linspace(280.0, 360.0, rungs)
    return comps, temps


def batch_of_mix(comps, temps, fp32):
    import faunus_amd as F
    b = F.Batch(fp32_pairs=fp32)
    tref = 298.15
    sizes = []
    for ci, (n, conc, sysd) in enumerate(comps):
        for ti, T in enumerate(temps):
            b.add(sysd, beta=tref / T, seed=10_000 + ci * 100 + ti, ladder=ci)
            sizes.append(n)
    return b, np.array(sizes)


def mix_flops(sizes, K=1152, fp32=True):
    """Algorithmic flops of one trial move of each replica: pair terms 2 (N-1) x 27 (minimum
    image and r^2 12, hard sphere 1, Ewald real space with the S(q) spline 14) and the
    Ewald dU 2K x 7 + 8K plus 2K sincos at 20 flops each."""
    pair = 2.0 * (sizes - 1) * 27.0
    ewald = ewald_flops_per_move(K) * np.ones_like(pair)
    return pair, ewald


def batch_roofline(pair_flops, ewald_flops, kernel_s, fp32):
    pf, ef = float(pair_flops), float(ewald_flops)
    out = {"bound": "valu", "unit": "TFLOP/s", "kernel_s": kernel_s,
           "pair_tflops": pf / kernel_s / 1e12, "pair_precision": "fp32" if fp32 else "fp64",
           "pair_peak": FP32_PEAK_TFLOPS if fp32 else FP64_PEAK_TFLOPS,
           "ewald_tflops_fp64": ef / kernel_s / 1e12, "ewald_peak_fp64": FP64_PEAK_TFLOPS}
    out["frac"] = out["pair_tflops"] / out["pair_peak"] + out["ewald_tflops_fp64"] / FP64_PEAK_TFLOPS
    out["note"] = ("algorithmic flops per replica-move (pairs 2(N-1) x 27, Ewald 2K x 7 + 8K + 2K sincos x 20) summed "
                   "over the launch / kernel time (HIP events); frac adds the two units' shares of their peaks")
    return out


def bench_batch(comps, temps, fp32, args, budget_moves, tune=True):
    """Cost-balanced form: one persistent launch per step gives replica r moves inversely
    proportional to its per-move cost (N + 2K), so every workgroup works for about the same
    time, then one round of ladder swaps; the tile autotuner runs first."""
    b, sizes = batch_of_mix(comps, temps, fp32)
    cost = sizes + 2 * 1152.0
    moves = np.maximum(1, np.round(budget_moves * cost.mean() / cost)).astype(np.int64)
    b.set_timing(True)
    b.run(moves)
    b.sync()
    tuned = b.autotune(300) if tune else (0, {})
    kernel_ms = 0.0
    t0 = time.perf_counter()
    for step in range(args.steps):
        b.run(moves)
        kernel_ms += b.sync()
        b.swap(seed=4242, rnd=step)
    dt = time.perf_counter() - t0
    check_replicas(b, len(sizes), "batched salt replicas")
    total_moves = int(moves.sum()) * args.steps
    drifts = []
    for r in (0, len(sizes) // 2, len(sizes) - 1):
        tot, _ = b.total_energy(r)
        drifts.append(abs(b.info(r)["energy"] - tot) / abs(tot))
    sweeps = [round(float(moves[ci * len(temps)] * args.steps / n), 1) for ci, (n, _, _) in enumerate(comps)]
    swaps = b.swap_counts()
    b.close()
    pf, ef = mix_flops(sizes, fp32=fp32)
    roof = batch_roofline((pf * moves).sum() * args.steps, (ef * moves).sum() * args.steps, kernel_ms / 1e3, fp32)
    return dict(rate=total_moves / dt, kernel_ms=kernel_ms, seconds=dt, moves=total_moves, swaps=swaps,
                max_drift=float(max(drifts)), sweeps_per_composition=sweeps, moves_per_replica=moves,
                tuned_threads=tuned[0], tuner_us_per_move=tuned[1], roofline=roof)


def bench_batch_intervals(comps, temps, fp32, intervals=1, sweeps=100):
    """The configuration's own schedule: every replica advances `sweeps` sweeps (N moves
    each), then one round of ladder swaps, the interval captured once as a HIP graph
    (persistent launch + swap kernel with a device round counter) and replayed.  The
    interval lasts as long as the slowest (largest) replica's sweeps."""
    b, sizes = batch_of_mix(comps, temps, fp32)
    b.set_timing(True)
    b.run(np.full(len(sizes), 100, np.int64))
    b.sync()
    tuned = b.autotune(300)
    moves = (sweeps * sizes).astype(np.int64)
    b.capture_interval(moves, seed=4242, round0=0)
    t0 = time.perf_counter()
    b.replay(intervals)
    kernel_ms = b.sync()
    dt = time.perf_counter() - t0
    check_replicas(b, len(sizes), "batched salt replicas (graph intervals)")
    swaps = b.swap_counts()
    b.close()
    total = int(moves.sum()) * intervals
    pf, ef = mix_flops(sizes, fp32=fp32)
    roof = batch_roofline((pf * moves).sum() * intervals, (ef * moves).sum() * intervals, kernel_ms / 1e3, fp32)
    return dict(rate=total / dt, seconds=dt, kernel_ms=kernel_ms, moves=total, intervals=intervals,
                sweeps_per_interval=sweeps, swaps=swaps, tuned_threads=tuned[0], tuner_us_per_move=tuned[1],
                roofline=roof)


def bench_batch_free(comps, temps, fp32, sweeps=100, big_intervals=1):
    """The configuration's schedule with free-running ladders: every replica advances
    `sweeps` sweeps per interval and each composition's ladder swaps after each of its own
    intervals at its own barrier, all in one cooperative launch.  A ladder of cost c
    (interval moves 100 N x per-move cost N + 2K) runs about big_intervals x c_max / c
    intervals, so every workgroup works for about as long as the largest composition's
    intervals and none waits for another composition."""
    b, sizes = batch_of_mix(comps, temps, fp32)
    b.set_timing(True)
    b.run(np.full(len(sizes), 100, np.int64))
    b.sync()
    tuned = b.autotune(300)
    moves = (sweeps * sizes).astype(np.int64)
    cost = moves * (sizes + 2 * 1152.0)
    intervals = np.maximum(1, np.round(big_intervals * cost.max() / cost)).astype(np.int32)
    t0 = time.perf_counter()
    b.run_ladders(moves, intervals, seed=4242, round0=0)
    kernel_ms = b.sync()
    dt = time.perf_counter() - t0
    check_replicas(b, len(sizes), "batched salt replicas (free-running ladders)")
    swaps = b.swap_counts()
    b.close()
    total = int((moves * intervals).sum())
    pf, ef = mix_flops(sizes, fp32=fp32)
    roof = batch_roofline((pf * moves * intervals).sum(), (ef * moves * intervals).sum(), kernel_ms / 1e3, fp32)
    per_comp = [int(intervals[ci * len(temps)]) for ci in range(len(comps))]
    return dict(rate=total / dt, seconds=dt, kernel_ms=kernel_ms, moves=total, sweeps_per_interval=sweeps,
                intervals_per_composition=per_comp, composition_sizes=[n for n, _, _ in comps], swaps=swaps,
                tuned_threads=tuned[0], tuner_us_per_move=tuned[1], roofline=roof)


def bench_plain_mix(comps, moves_per_replica, temps, per_replica_moves=300):
    """Plain FP64 one-move-per-launch path on the same mix and the same per-replica move
    counts: the per-move time of each composition (one engine at T_ref; the per-move cost
    does not depend on T) is measured on a sample and the batched launch's moves are
    priced with it."""
    per_move = []
    for n, conc, sysd in comps:
        r = bench_plain_single(sysd, per_replica_moves + 60)
        per_move.append(1.0 / r["rate"])
    per_move = np.repeat(np.array(per_move), len(temps))
    return float(moves_per_replica.sum() / (moves_per_replica * per_move).sum())


def cpu_mix_baseline(comps, seconds=12.0):
    """One FP64 oracle replica per host core, the compositions dealt round-robin, each a
    bounded sample of single-particle translations (Philox proposal stream)."""
    from oracle import oracle as O

    def work(w, secs):
        n, conc, (cfg, xyz, q, ids, g) = comps[w % len(comps)]
        s = O.System(cfg, fast=True)
        s.set_particles(xyz, q, ids)
        s.set_groups(g["begin"], g["size"], g["molid"])
        s.init()
        t0 = time.perf_counter()
        s.run_philox_translate(20, seed=w, replica=w)
        per = (time.perf_counter() - t0) / 20
        nm = max(50, int(secs / per))
        t0 = time.perf_counter()
        s.run_philox_translate(nm, seed=w, replica=w, move0=20)
        return nm, time.perf_counter() - t0

    out = cpu_replicas_baseline(work, seconds)
    out.update(unit="replica-moves/s", kind="port",
               sample=f"{out['moves']} single-particle translations over {out['cores']} concurrent FP64 oracle "
                      f"replicas of the 8 compositions (one per core, {out['wall_s']:.1f} s)")
    return out


def cpu_baseline(seconds=15.0):
    """FP64 oracle (the reference's Release flags), one core, single-particle
    translations on the N = 20 000 system; bounded sample of moves."""
    from oracle import oracle as O
    cfg, xyz, q, ids, g = nacl_system()
    s = O.System(cfg, fast=True)
    s.set_particles(xyz, q, ids)
    s.set_groups(g["begin"], g["size"], g["molid"])
    s.init()
    t0 = time.perf_counter()
    s.run_translate(20, seed=3)
    per = (time.perf_counter() - t0) / 20
    nm = max(50, int(seconds / per))
    t0 = time.perf_counter()
    s.run_translate(nm, seed=5)
    dt = time.perf_counter() - t0
    return dict(value=nm / dt, unit="trial moves/s", cores=1, kind="port",
                sample=f"{nm} single-particle translations of the N=20000 system (FP64 oracle, all pairs as in "
                       f"the reference, no cell list, 1 thread, {dt:.1f} s)")


HEADLINE_GROUPS = 16
# consecutive blocks (spread over the XCDs); placing the 16 workgroups on one XCD (stride
# 8) measured 6.6-6.9 us per move on one box but 7.4-8.4 on others, against a steady 7.1-7.5
HEADLINE_STRIDE = 1


def water_system(n_mol=4096, seed=100):
    """4096 rigid waters written as the reference's water example writes its energy
    (examples/water/ewald.yml: nonbonded_coulomblj, LJ truncated by cutoff_g2g = 10 A on
    the molecules' mass centres) with alpha = 0.28, Rc = 10 A, kmax = 9 (K = 1661)."""
    from tests import systems
    return systems.water_g2g(n_mol=n_mol, seed=seed, alpha=0.28, cutoff=10.0, kmax=9, g2g=10.0)


def water_systems(n_mol=4096, comps=4):
    return [water_system(n_mol, seed=100 + c) for c in range(comps)]


def bench_water(args, fp32=True, replicas=64, n_mol=4096, moves_per_step=3000, steps=None):
    """64 rigid-water replicas (4096 molecules, L=49.68 A, Ewald alpha=0.28, Rc=10,
    kmax=9 (K=1661), LJ truncated by cutoff_g2g = 10 A as in the reference's water
    example, cell list) on a 280-360 K ladder with molecular translate+rotate moves (dp
    0.5 A, dprot 0.5 rad), one persistent launch per step and one round of neighbour
    swaps after it; the tile autotuner picks the cell grid (cells of the interaction
    range or half of it) first."""
    import faunus_amd as F
    steps = steps or args.steps
    comps = water_systems(n_mol)
    temps = np.linspace(280.0, 360.0, replicas)
    b = F.Batch(fp32_pairs=fp32, cell_divisions=args.water_cell_divisions)
    for r in range(replicas):
        cfg, xyz, q, ids, g, kinds = comps[r % len(comps)]
        b.add((cfg, xyz, q, ids, g), beta=298.15 / temps[r], seed=9000 + r, ladder=0, moves=kinds)
    b.set_timing(True)
    b.run([moves_per_step] * replicas)
    b.sync()
    tuned = b.autotune(400)
    kernel_ms = 0.0
    t0 = time.perf_counter()
    for step in range(steps):
        b.run([moves_per_step] * replicas)
        kernel_ms += b.sync()
        b.swap(seed=77, rnd=step)
    dt = time.perf_counter() - t0
    check_replicas(b, replicas, "water replicas")
    total = replicas * moves_per_step * steps
    drift = []
    for r in (0, replicas - 1):
        tot, _ = b.total_energy(r)
        drift.append(abs(b.info(r)["energy"] - tot) / abs(tot))
    acc = np.mean([b.info(r)["accepted"] / b.info(r)["moves"] for r in range(replicas)])
    swaps = b.swap_counts()
    b.close()
    # algorithmic flops per move: the three moved atoms against the three atoms of every
    # molecule whose centre lies within the group cutoff (the pairs the reference
    # evaluates), in both states, 27 flops each, and the Ewald dU of three moved atoms
    L = comps[0][0]["geometry"]["length"]
    g2g = 10.0
    n_mol_in = n_mol * (4.0 / 3.0 * math.pi * g2g ** 3) / L ** 3
    pf = 2 * 3 * 3 * n_mol_in * 27.0 * total
    ef = ewald_flops_per_move(1661, m=3) * total
    return dict(rate=total / dt, kernel_ms=kernel_ms, seconds=dt, moves=total, swaps=swaps,
                max_drift=float(max(drift)), acceptance=float(acc), tuned_cell_divisions=tuned[0],
                tuner_us_per_move=tuned[1], roofline=batch_roofline(pf, ef, kernel_ms / 1e3, fp32))


def cpu_water_baseline(n_mol=4096, seconds=12.0):
    """One FP64 oracle water replica per host core (translate+rotate moves, Philox stream)."""
    from oracle import oracle as O
    comps = water_systems(n_mol)

    def work(w, secs):
        cfg, xyz, q, ids, g, kinds = comps[w % len(comps)]
        s = O.System(cfg, fast=True)
        s.set_particles(xyz, q, ids)
        s.set_groups(g["begin"], g["size"], g["molid"])
        s.init()
        t0 = time.perf_counter()
        s.run_philox_moves(20, seed=w, kinds=kinds, replica=w)
        per = (time.perf_counter() - t0) / 20
        nm = max(50, int(secs / per))
        t0 = time.perf_counter()
        s.run_philox_moves(nm, seed=w, kinds=kinds, replica=w, move0=20)
        return nm, time.perf_counter() - t0

    out = cpu_replicas_baseline(work, seconds)
    out.update(unit="replica-moves/s", kind="port",
               sample=f"{out['moves']} molecular translate+rotate moves over {out['cores']} concurrent FP64 oracle "
                      f"replicas of 4096 waters (one per core, {out['wall_s']:.1f} s)")
    return out


def bench_plain_water(n_mol=4096, nmoves=400, seed=5):
    """Plain FP64 path on the same water system: one fused launch and one host round
    trip per molecular move (fmc_replay_moves).  Proposals (translate dp 0.5 A and
    rotate 0.5 rad about the mass centre) are drawn from the initial configuration."""
    import faunus_amd as F
    from tests import systems
    cfg, xyz, q, ids, g, kinds = water_system(n_mol, seed=100)
    e = F.Engine(cfg)
    e.set_particles(xyz, q, ids)
    e.set_groups(g["begin"], g["size"], g["molid"])
    e.init()
    L = cfg["geometry"]["length"]
    rng = np.random.default_rng(seed)
    grp = rng.integers(0, n_mol, nmoves).astype(np.int32)
    idx = np.concatenate([3 * w + np.arange(3) for w in grp]).astype(np.int32)
    new = np.zeros((3 * nmoves, 3))
    cms = np.zeros((nmoves, 3))
    for t, w in enumerate(grp):
        cm = e.mass_center(int(w))
        d = rng.normal(size=3)
        d *= 0.5 * rng.random() / np.linalg.norm(d)
        axis = rng.normal(size=3)
        axis /= np.linalg.norm(axis)
        ang = 0.5 * (rng.random() - 0.5)
        K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
        R = np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K
        p = xyz[3 * w:3 * w + 3] - cm
        p -= L * np.round(p / L)
        p = p @ R.T + cm + d
        new[3 * t:3 * t + 3] = p - L * np.round(p / L)
        c2 = cm + d
        cms[t] = c2 - L * np.round(c2 / L)
    off = np.arange(0, 3 * nmoves + 1, 3, dtype=np.int64)
    u = rng.random(nmoves)
    e.replay_moves(grp[:20], off[:21], idx[:60], new[:60], u[:20], whole=np.ones(20, np.uint8), mass_center=cms[:20])
    e.set_timing(True)
    t0 = time.perf_counter()
    e.replay_moves(grp[20:], off[20:] - 60, idx[60:], new[60:], u[20:], whole=np.ones(nmoves - 20, np.uint8),
                   mass_center=cms[20:])
    dt = time.perf_counter() - t0
    ms, launches = e.timing()
    return dict(rate=(nmoves - 20) / dt, kernel_us=ms * 1e3 / max(1, launches))


def serve_stream(n_jobs, seed=1, rate=50.0):
    """Poisson arrivals (rate jobs/s) of polyelectrolyte jobs: chain length ~ U[50, 500],
    total N ~ logU[1k, 50k]; the job id seeds the configuration and the Philox stream."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_jobs))
    out = []
    for j in range(n_jobs):
        nb = int(rng.integers(50, 501))
        nt = max(int(round(10 ** rng.uniform(3.0, math.log10(50000.0)))), 2 * nb + 2)
        out.append((float(arrivals[j]), nb, nt, j))
    return out


def bounded_stream(n_jobs, max_n, seed=1, rate=50.0):
    """The first n_jobs jobs of the Poisson stream with N <= max_n, at their own arrival
    times (shifted so that the first arrives at 0)."""
    picked = [s for s in serve_stream(40 * n_jobs, seed, rate) if s[2] <= max_n][:n_jobs]
    t0 = picked[0][0]
    return [(a - t0, nb, nt, j) for a, nb, nt, j in picked]


def bench_serve(stream, sweeps, fp32=True, quantum=5000, split_atoms=2048, max_parts=16):
    """Continuous batching under a Poisson stream of polyelectrolyte jobs (chain of
    50-500 beads with harmonic bonds, counterions and 1:1 salt, 5 mM monomer, hard
    spheres + Ewald alpha=5/L, kmax=8; ion and bead translations and pivots): submit each
    job at its arrival time, run scheduling rounds while anything is queued or resident,
    and report each job's time from arrival to completion of `sweeps` sweeps (N moves
    per sweep).  Jobs above `split_atoms` particles are shared by ceil(N / split_atoms)
    workgroups (at most `max_parts`)."""
    import faunus_amd as F
    from tests import systems
    n_jobs = len(stream)
    jobs = [systems.polyelectrolyte(nb, nt, seed=1000 + j) for (_, nb, nt, j) in stream]
    log("serve: %d jobs built, %d particles in total" % (n_jobs, sum(len(j[2]) for j in jobs)))
    srv = F.Server(fp32_pairs=fp32, hbm_budget_bytes=200 << 30, max_slots=512, split_atoms=split_atoms,
                   max_parts=max_parts)
    t0 = srv.status()["seconds"]
    ids = {}
    nxt = 0
    rounds = 0
    peak = 0
    last_report = 0.0
    while True:
        now = srv.status()["seconds"] - t0
        if now - last_report > 5.0:
            st = srv.status()
            log("serve: %.1f s, %d submitted, %d rounds, %.3e moves done" % (now, nxt, rounds, st["moves_done"]))
            last_report = now
        while nxt < n_jobs and stream[nxt][0] <= now:
            cfg, xyz, q, idl, g, kinds = jobs[nxt]
            ids[nxt] = srv.submit((cfg, xyz, q, idl, g), sweeps * len(q), seed=1000 + stream[nxt][3], kinds=kinds)
            nxt += 1
        r, q = srv.step(quantum)
        rounds += 1
        peak = max(peak, r)
        if r == 0 and q == 0:
            if nxt >= n_jobs:
                break
            time.sleep(max(0.0, stream[nxt][0] - (srv.status()["seconds"] - t0)))
    lat, moves = [], 0
    t_end = 0.0
    for j, jid in ids.items():
        info = srv.job(jid)
        if info["state"] != F.Server.DONE:
            raise RuntimeError(f"serving job {j} failed: {info}")
        done = info["t_done"] - t0
        lat.append(done - stream[j][0])
        moves += info["moves"]
        t_end = max(t_end, done)
    srv.close()
    lat = np.array(lat)
    ns = np.array([len(sysd[2]) for sysd in jobs])
    wall = t_end - stream[0][0]
    span = stream[-1][0] - stream[0][0]
    offered = (moves / span) if span > 0 else float("inf")
    return dict(jobs=n_jobs, sweeps_per_job=sweeps, split_atoms=split_atoms, max_parts=max_parts,
                p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)), moves_per_s=moves / wall,
                wall_s=wall, rounds=rounds, peak_resident=peak, n_min=int(ns.min()), n_median=int(np.median(ns)),
                n_max=int(ns.max()), offered_moves_per_s=offered, rho=offered / (moves / wall))


def cpu_serve_baseline(stream, sweeps, seconds=1.0):
    """The same jobs run one per host core, first come first served in arrival order: each
    job's single-core FP64 oracle rate is measured on a bounded sample (about `seconds`
    each, all cores busy) and its time to `sweeps` sweeps is priced with it; the FCFS
    schedule over the cores gives p50/p99 from arrival to completion."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle as O
    from tests import systems

    def rate_of(k):
        _, nb, nt, j = stream[k]
        cfg, xyz, q, ids, g, kinds = systems.polyelectrolyte(nb, nt, seed=1000 + j)
        s = O.System(cfg, fast=True)
        s.set_particles(xyz, q, ids)
        s.set_groups(g["begin"], g["size"], g["molid"])
        s.init()
        t0 = time.perf_counter()
        s.run_philox_moves(20, seed=j, kinds=kinds)
        per = (time.perf_counter() - t0) / 20
        nm = max(40, int(seconds / per))
        t0 = time.perf_counter()
        s.run_philox_moves(nm, seed=j, kinds=kinds, move0=20)
        return nm / (time.perf_counter() - t0), len(q)

    cores = cpu_workers()
    with ThreadPoolExecutor(cores) as ex:
        rates = list(ex.map(rate_of, range(len(stream))))
    free = [0.0] * cores
    lat = []
    for k, (arr, _, _, _) in enumerate(stream):
        rate, n = rates[k]
        c = int(np.argmin(free))
        start = max(arr, free[c])
        free[c] = start + sweeps * n / rate
        lat.append(free[c] - arr)
    lat = np.array(lat)
    return dict(p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)), cores=cores, kind="port",
                unit="s (arrival to 1000 sweeps)",
                sample=f"each job's FP64 oracle rate measured for about {seconds:.1f} s on one core ({len(stream)} jobs, "
                       f"{cores} cores at once), its {sweeps} sweeps priced with it and scheduled first come first "
                       f"served over {cores} cores")


def fp32_errors(headline_sysd):
    """fp32_error on each benchmarked configuration: the headline system, every
    composition of the salt mix, a 4096-water replica and two serving jobs (the
    stream's first and the largest polyelectrolyte)."""
    from tests import systems
    out = {"nacl_20k": fp32_error(headline_sysd, 2000, seed=31)}
    comps, _ = salt_mix()
    mix = [fp32_error(sysd, 1500, seed=41 + c) for c, (_, _, sysd) in enumerate(comps)]
    worst = max(mix, key=lambda r: r["max_rel_err_vs_fp64"])
    out["salt_mix"] = dict(worst, per_composition=[r["max_rel_err_vs_fp64"] for r in mix])
    cfg, xyz, q, ids, g, kinds = water_system(4096, seed=100)
    out["water_4096"] = fp32_error((cfg, xyz, q, ids, g), 800, seed=51, kinds=kinds)
    for name, (nb, nt) in (("polyelectrolyte_job_first", serve_stream(1)[0][1:3]), ("polyelectrolyte_43k", (500, 43000))):
        cfg, xyz, q, ids, g, kinds = systems.polyelectrolyte(nb, nt, seed=1000)
        out[name] = dict(fp32_error((cfg, xyz, q, ids, g), 800, seed=61, kinds=kinds), n=len(q))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--no-batch", action="store_true", help="skip the 256-replica mix")
    ap.add_argument("--only", choices=["device", "plain", "batch", "split", "pipe", "headline", "copy", "serve",
                                       "gsplit", "water", "intervals", "fp32err", "batch32", "water32", "free"],
                    default=None,
                    help="run one leg and print its results (batch32 / water32: the FP32 batched forms alone, "
                         "for profiling)")
    ap.add_argument("--serve-jobs", type=int, default=24,
                    help="serving: the first jobs of the stream with N <= --serve-max-n, to 1000 sweeps")
    ap.add_argument("--serve-max-n", type=int, default=3000)
    ap.add_argument("--serve-sweeps", type=int, default=1000)
    ap.add_argument("--serve-split-atoms", type=int, default=512,
                    help="particles per workgroup of a serving job (0: one workgroup per job)")
    ap.add_argument("--serve-max-parts", type=int, default=16)
    ap.add_argument("--water-cell-divisions", type=int, default=2,
                    help="water replicas: starting cell grid before the autotuner (1: 27-cell stencils, 2: 125)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): the engine has no multi-GPU path; "
              "independent replicas and serving jobs would be spread over GPUs by the caller, one process each",
              file=sys.stderr)
        sys.exit(2)
    import faunus_amd as F
    if F.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.only == "copy":
        nbytes = 1 << 28
        ms = F.copy_probe(nbytes, repeats=10)
        print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "ms_per_launch": ms,
                          "GBps_read_plus_write": 2 * nbytes / ms / 1e6}))
        return
    sysd = nacl_system()
    K = 17328
    n = 20000
    out = {}
    single_steps = max(1, min(args.steps, 5))
    if args.only in (None, "headline"):
        out["split64"] = bench_split(sysd, args, HEADLINE_GROUPS, HEADLINE_STRIDE)
        log("split FP64 (G=%d): %.0f moves/s, drift %.2e" % (HEADLINE_GROUPS, out["split64"]["rate"],
                                                              out["split64"]["drift"]))
    if args.only in (None, "plain"):
        out["plain"] = bench_plain_single(sysd, PLAIN_MOVES)
        log("plain FP64 one move per launch: %.0f moves/s, kernel %.1f us" % (out["plain"]["rate"], out["plain"]["kernel_us"]))
    if args.only in (None, "device"):
        out["device64"] = bench_device_single(sysd, args, fp32=False, steps=single_steps)
        log("device-resident FP64, one workgroup: %.0f moves/s (drift %.2e)" % (out["device64"]["rate"],
                                                                                 out["device64"]["drift"]))
        out["device32"] = bench_device_single(sysd, args, fp32=True, steps=single_steps)
        log("device-resident FP32-pair, one workgroup: %.0f moves/s" % out["device32"]["rate"])
    if args.only == "pipe":
        for groups, depth in [(16, 1), (16, 2), (16, 4), (32, 2)]:
            r = bench_pipe(sysd, args, groups, depth, steps=3)
            out[f"pipe_{groups}_{depth}"] = r
            log("pipe G=%d D=%d: %.0f moves/s (kernel %.2f us/move)" % (
                groups, depth, r["rate"], r["kernel_ms"] * 1e3 / r["moves"]))
    if args.only == "split":
        for groups, stride in [(12, 1), (16, 1), (24, 1), (32, 1), (48, 1), (64, 1), (16, 8)]:
            r = bench_split(sysd, args, groups, stride, steps=3)
            out[f"split_{groups}_{stride}"] = r
            log("split G=%d stride=%d: %.0f moves/s (kernel %.2f us/move)" % (
                groups, stride, r["rate"], r["kernel_ms"] * 1e3 / r["moves"]))
    if args.only in (None, "batch", "intervals") and not args.no_batch:
        comps, temps = salt_mix()
        if args.only in (None, "batch"):
            budget = 4000
            out["batch32"] = bench_batch(comps, temps, True, args, budget)
            log("batched FP32-pair/FP64-Q, cost-balanced: %.3e replica-moves/s (tuned %s)" % (
                out["batch32"]["rate"], out["batch32"]["tuner_us_per_move"]))
            out["batch64"] = bench_batch(comps, temps, False, args, budget)
            log("batched FP64, cost-balanced: %.3e replica-moves/s" % out["batch64"]["rate"])
            out["plain_mix"] = bench_plain_mix(comps, out["batch32"]["moves_per_replica"], temps)
            log("plain FP64 one move per launch, same per-replica moves: %.0f replica-moves/s" % out["plain_mix"])
        out["intervals32"] = bench_batch_intervals(comps, temps, True)
        log("batched FP32-pair/FP64-Q, 100-sweep swap intervals from a HIP graph: %.3e replica-moves/s (%.1f s)" % (
            out["intervals32"]["rate"], out["intervals32"]["seconds"]))
        out["free32"] = bench_batch_free(comps, temps, True)
        log("batched FP32-pair/FP64-Q, free-running ladders of 100-sweep intervals: %.3e replica-moves/s "
            "(%.1f s, intervals per composition %s)" % (out["free32"]["rate"], out["free32"]["seconds"],
                                                        out["free32"]["intervals_per_composition"]))
        if "plain_mix" not in out:
            out["plain_mix"] = bench_plain_mix(comps, np.array([100 * n for n, _, _ in comps for _ in temps]), temps)
        out["plain_mix_intervals"] = bench_plain_mix(comps, np.array([100 * n for n, _, _ in comps for _ in temps]), temps)
        if not args.no_cpu:
            out["cpu_mix"] = cpu_mix_baseline(comps)
            log("CPU baseline, one oracle replica per core: %.0f replica-moves/s on %d cores" % (
                out["cpu_mix"]["value"], out["cpu_mix"]["cores"]))
    if args.only in (None, "water"):
        out["water32"] = bench_water(args, fp32=True)
        log("64 water replicas, FP32 pairs: %.3e replica-moves/s (acceptance %.2f, cells %s)" % (
            out["water32"]["rate"], out["water32"]["acceptance"], out["water32"]["tuner_us_per_move"]))
        out["water64"] = bench_water(args, fp32=False, steps=single_steps)
        log("64 water replicas, FP64: %.3e replica-moves/s" % out["water64"]["rate"])
        out["water_plain"] = bench_plain_water()
        log("plain FP64 per-move path on water: %.0f moves/s (kernel %.1f us)" % (
            out["water_plain"]["rate"], out["water_plain"]["kernel_us"]))
        if not args.no_cpu:
            out["cpu_water"] = cpu_water_baseline()
            log("CPU baseline, one oracle water replica per core: %.0f replica-moves/s on %d cores" % (
                out["cpu_water"]["value"], out["cpu_water"]["cores"]))
    if args.only in (None, "gsplit"):
        out["gsplit"] = bench_general_split()
    if args.only in (None, "serve"):
        stream = bounded_stream(args.serve_jobs, args.serve_max_n)
        out["serve"] = bench_serve(stream, args.serve_sweeps, split_atoms=args.serve_split_atoms,
                                   max_parts=args.serve_max_parts)
        log("serving %d jobs (N <= %d) to %d sweeps: p50 %.2f s, p99 %.2f s, %.3e moves/s, rho %.3f" % (
            out["serve"]["jobs"], args.serve_max_n, args.serve_sweeps, out["serve"]["p50_s"], out["serve"]["p99_s"],
            out["serve"]["moves_per_s"], out["serve"]["rho"]))
        full = serve_stream(32)
        out["serve_full"] = bench_serve(full, 2, split_atoms=2048, max_parts=16)
        log("serving the full stream's first 32 jobs to 2 sweeps: p50 %.2f s, p99 %.2f s, %.3e moves/s" % (
            out["serve_full"]["p50_s"], out["serve_full"]["p99_s"], out["serve_full"]["moves_per_s"]))
        if not args.no_cpu:
            out["cpu_serve"] = cpu_serve_baseline(stream, args.serve_sweeps)
            log("CPU baseline serving, one job per core: p50 %.1f s, p99 %.1f s" % (
                out["cpu_serve"]["p50_s"], out["cpu_serve"]["p99_s"]))
    if args.only == "batch32":
        comps, temps = salt_mix()
        out["batch32"] = bench_batch(comps, temps, True, args, 4000)
        log("batched FP32-pair/FP64-Q, cost-balanced: %.3e replica-moves/s" % out["batch32"]["rate"])
    if args.only == "free":
        comps, temps = salt_mix()
        out["free32"] = bench_batch_free(comps, temps, True)
        log("free-running ladders: %.3e replica-moves/s (%.1f s, intervals %s)" % (
            out["free32"]["rate"], out["free32"]["seconds"], out["free32"]["intervals_per_composition"]))
    if args.only == "water32":
        out["water32"] = bench_water(args, fp32=True)
        log("64 water replicas, FP32 pairs: %.3e replica-moves/s" % out["water32"]["rate"])
    if args.only in (None, "fp32err"):
        out["fp32err"] = fp32_errors(sysd)
        log("FP32-pair form vs FP64, max relative dU error per config: %s" % {
            k: "%.2e" % v["max_rel_err_vs_fp64"] for k, v in out["fp32err"].items()})
    if args.only is not None:
        print(json.dumps({k: v for k, v in out.items()}, default=lambda x: x.tolist() if hasattr(x, "tolist") else float(x)))
        return
    cpu = None if args.no_cpu else cpu_baseline()
    d = out["split64"]
    kernel_s_per_move = d["kernel_ms"] / 1e3 / d["moves"]
    flops = ewald_flops_per_move(K)
    algo_bytes = 64.0 * K + pair_bytes_cells(n, 255.0, 11.99)
    achieved_gbs = algo_bytes / kernel_s_per_move / 1e9
    achieved_tf = flops / kernel_s_per_move / 1e12
    line = {
        "metric": "trial moves/s at N=20k Ewald",
        "value": d["rate"],
        "unit": "trial moves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": d["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (random sequential addition, seed 1; device Philox proposals)",
        "config": {"workload": "N=20000 Na/Cl LJ(LB)+Ewald(alpha 0.25, Rc 12, kmax 20, K=17328), pair energy "
                               "nonbonded_splined (rmax 11.99 A, utol 1e-3), L=255 A, translate dp=2 A, "
                               "device-resident Metropolis, %d cooperating workgroups of 512 threads (workgroup 0: "
                               "pair walk and commit; the others: k-vector slices)" % HEADLINE_GROUPS,
                   "step_moves": STEP_MOVES, "acceptance": d["acceptance"], "energy_drift": d["drift"]},
        "roofline": {"bound": "hbm", "achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved_gbs / HBM_PEAK_GBS, "traffic": HEADLINE_TRAFFIC_BYTES,
                     "traffic_unit": "bytes per launch (%d moves)" % STEP_MOVES, "kernel": "mc_split_kernel",
                     "algorithmic_bytes_per_move": algo_bytes, "kernel_us_per_move": kernel_s_per_move * 1e6,
                     "fp64_tflops": achieved_tf, "fp64_peak_tflops": FP64_PEAK_TFLOPS,
                     "algorithmic_flops_per_move": flops,
                     "limiter": "per-move latency chain of workgroup 0 (one Metropolis chain: every workgroup "
                                "publishes its share of the next move right after each decision, so the exchange "
                                "and the k-vector slices of workgroups 1-15 run in the shadow of workgroup 0's "
                                "stencil walk of the next move and its commit)",
                     "note": "one launch runs STEP_MOVES sequential moves; achieved = algorithmic bytes per move "
                             "/ kernel time per move (HIP events); Q(k) stays in registers, so the measured HBM "
                             "traffic (about 19 KB per move, mostly the exchange's polls) is far below the "
                             "algorithmic bytes; neither HBM nor the FP64 VALU binds: the move time is a chain of "
                             "dependent latencies"},
        "cpu_baseline": cpu,
        "forms": {
            "plain_fp64_one_move_per_launch": {"moves_per_s": out["plain"]["rate"],
                                               "kernel_us": out["plain"]["kernel_us"]},
            "device_fp64_split_%d_workgroups" % HEADLINE_GROUPS: {"moves_per_s": d["rate"]},
            "device_fp64_one_workgroup": {"moves_per_s": out["device64"]["rate"]},
            "device_fp32_pairs_one_workgroup": {"moves_per_s": out["device32"]["rate"]},
        },
    }
    if "intervals32" in out:
        iv = out["intervals32"]
        bal = out.get("batch32")
        line["forms"]["batched_256_replicas"] = {
            "fp32_pairs_fp64_q_replica_moves_per_s": iv["rate"],
            "schedule": "every replica 100 sweeps, then one round of ladder swaps; the interval captured as a HIP "
                        "graph and replayed; the interval ends with the largest replica's sweeps",
            "intervals_timed": iv["intervals"], "interval_seconds": iv["seconds"] / iv["intervals"],
            "plain_fp64_replica_moves_per_s_same_moves": out["plain_mix_intervals"],
            "speedup_fp32_batched_vs_plain": iv["rate"] / out["plain_mix_intervals"],
            "tuned_threads": iv["tuned_threads"], "tuner_us_per_replica_move": iv["tuner_us_per_move"],
            "swaps_attempted_accepted": iv["swaps"], "roofline": iv["roofline"]}
        if "free32" in out:
            fr = out["free32"]
            line["forms"]["batched_256_replicas_free_running_ladders"] = {
                "fp32_pairs_fp64_q_replica_moves_per_s": fr["rate"],
                "schedule": "every replica 100 sweeps per interval; each composition's ladder swaps after each of "
                            "its own intervals at a barrier of its own, all ladders in one cooperative launch; a "
                            "composition runs intervals in inverse proportion to its interval cost",
                "intervals_per_composition": fr["intervals_per_composition"],
                "composition_sizes": fr["composition_sizes"], "seconds": fr["seconds"],
                "plain_fp64_replica_moves_per_s_same_moves": out["plain_mix_intervals"],
                "speedup_fp32_batched_vs_plain": fr["rate"] / out["plain_mix_intervals"],
                "tuned_threads": fr["tuned_threads"], "swaps_attempted_accepted": fr["swaps"],
                "roofline": fr["roofline"]}
        if bal:
            line["forms"]["batched_256_replicas_cost_balanced"] = {
                "fp32_pairs_fp64_q_replica_moves_per_s": bal["rate"],
                "fp64_replica_moves_per_s": out["batch64"]["rate"],
                "schedule": "moves per replica inversely proportional to its per-move cost, one round of ladder "
                            "swaps per launch",
                "plain_fp64_replica_moves_per_s_same_moves": out["plain_mix"],
                "speedup_fp32_batched_vs_plain": bal["rate"] / out["plain_mix"],
                "tuned_threads": bal["tuned_threads"], "tuner_us_per_replica_move": bal["tuner_us_per_move"],
                "swaps_attempted_accepted": bal["swaps"], "sweeps_per_composition": bal["sweeps_per_composition"],
                "max_energy_drift": bal["max_drift"], "roofline": bal["roofline"],
                "fp64_roofline": out["batch64"]["roofline"]}
        if "cpu_mix" in out:
            line["forms"]["batched_256_replicas"]["cpu_baseline"] = out["cpu_mix"]
    if "water32" in out:
        line["forms"]["batched_64_water_replicas"] = {
            "fp32_pairs_fp64_q_replica_moves_per_s": out["water32"]["rate"],
            "fp64_replica_moves_per_s": out["water64"]["rate"],
            "plain_fp64_moves_per_s": out["water_plain"]["rate"],
            "speedup_fp32_batched_vs_plain": out["water32"]["rate"] / out["water_plain"]["rate"],
            "acceptance": out["water32"]["acceptance"], "swaps_attempted_accepted": out["water32"]["swaps"],
            "max_energy_drift": out["water32"]["max_drift"],
            "tuned_cell_divisions": out["water32"]["tuned_cell_divisions"],
            "tuner_us_per_replica_move": out["water32"]["tuner_us_per_move"],
            "roofline": out["water32"]["roofline"], "fp64_roofline": out["water64"]["roofline"]}
        if "cpu_water" in out:
            line["forms"]["batched_64_water_replicas"]["cpu_baseline"] = out["cpu_water"]
    if "serve" in out:
        sv = dict(out["serve"])
        sv["full_stream_sample_2_sweeps"] = out["serve_full"]
        mean_n = float(np.mean([nt for _, _, nt, _ in serve_stream(2000)]))
        sv["full_stream_offered_moves_per_s"] = 50.0 * 1000 * mean_n
        sv["full_stream_rho"] = sv["full_stream_offered_moves_per_s"] / out["serve_full"]["moves_per_s"]
        sv["note"] = ("p50/p99 measured from arrival to the 1000th sweep on the bounded job set; the full stream "
                      "(50 jobs/s, N to 50k) offers rho >> 1 times the served rate, so its steady-state p99 is unbounded")
        if "cpu_serve" in out:
            sv["cpu_baseline"] = out["cpu_serve"]
        line["forms"]["serving_poisson_50_per_s"] = sv
    if "gsplit" in out:
        line["forms"]["general_split_polyelectrolyte_43k"] = out["gsplit"]
    if "fp32err" in out:
        fe = out["fp32err"]
        line["fp32_error_per_config"] = dict(
            fe, note="FP32-pair/FP64-Q(k) form against the FP64 form (held to the FP64 oracle within 1e-10 by the "
                     "parity tests) on the same Philox stream: max |du32 - du64| / max(|du64|, |E_total|) up to the "
                     "first differing decision")
        for form, key in (("batched_256_replicas", "salt_mix"), ("batched_64_water_replicas", "water_4096"),
                          ("serving_poisson_50_per_s", "polyelectrolyte_43k")):
            if form in line["forms"]:
                line["forms"][form]["max_rel_err_vs_fp64"] = fe[key]["max_rel_err_vs_fp64"]
        line["forms"]["device_fp32_pairs_one_workgroup"]["max_rel_err_vs_fp64"] = fe["nacl_20k"]["max_rel_err_vs_fp64"]
    print(json.dumps(line, default=lambda x: x.tolist() if hasattr(x, "tolist") else float(x)), flush=True)


if __name__ == "__main__":
    main()
