"""Benchmark of the Metropolis energy path on one MI355X.

Headline workload (N=1 line): N = 20 000 Na/Cl ions (q=+-1, LJ sigma 3.3/4.4 A,
eps 0.5/0.4 kJ/mol, Lorentz-Berthelot, LJ truncated at 12 A), cubic L = 255 A,
Ewald alpha=0.25 A^-1, Rc=12 A, kmax=20 (K=17 328), single-particle translations
dp = 2 A, FP64, device-resident Metropolis (one persistent launch per step).
A step is STEP_MOVES trial moves; the metric is trial moves per second.

Also measured and reported beside it: the plain FP64 one-move-per-launch path on
the same system, the FP32-pair form, the 256-replica 2:1-salt mix (8 compositions
x 32-rung temperature ladder, FP32 pairs / FP64 Q(k), and FP64) with the plain path
on the same mix, and a CPU baseline (FP64 oracle, one core, bounded sample).
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
HEADLINE_TRAFFIC_BYTES = 232.3e6
FP64_PEAK_TFLOPS = 78.6
STEP_MOVES = 20000
PLAIN_MOVES = 2000


def log(*a):
    print(*a, file=sys.stderr, flush=True)


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


def bench_device_single(sysd, args, fp32):
    import faunus_amd as F
    b = F.Batch(fp32_pairs=fp32)
    b.add(sysd, seed=2024)
    b.set_timing(True)
    for _ in range(args.warmup):
        b.run([STEP_MOVES])
    b.sync()
    kernel_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.run([STEP_MOVES])
        kernel_ms += b.sync()
    dt = time.perf_counter() - t0
    check_replicas(b, 1, "device-resident replica")
    info = b.info(0)
    tot, _ = b.total_energy(0)
    dq, drec = b.check_q(0)
    b.close()
    moves = args.steps * STEP_MOVES
    return dict(rate=moves / dt, ms_per_step=dt * 1e3 / args.steps, kernel_ms=kernel_ms, moves=moves,
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


def bench_batch(comps, temps, fp32, args, budget_moves):
    import faunus_amd as F
    b = F.Batch(fp32_pairs=fp32)
    tref = 298.15
    cost = []
    for ci, (n, conc, sysd) in enumerate(comps):
        for ti, T in enumerate(temps):
            b.add(sysd, beta=tref / T, seed=10_000 + ci * 100 + ti, ladder=ci)
            cost.append(n + 1152.0)
    cost = np.array(cost)
    # equal wall time per replica per launch: moves inversely proportional to the per-move cost
    moves = np.maximum(1, np.round(budget_moves * cost.mean() / cost)).astype(np.int64)
    b.set_timing(True)
    b.run(moves)
    b.sync()
    sweeps = np.zeros(len(comps))
    kernel_ms = 0.0
    swaps = (0, 0)
    t0 = time.perf_counter()
    for step in range(args.steps):
        b.run(moves)
        kernel_ms += b.sync()
        swaps = b.swap(seed=4242, rnd=step)
    dt = time.perf_counter() - t0
    check_replicas(b, len(cost), "batched salt replicas")
    total_moves = int(moves.sum()) * args.steps
    # per-move drift check on a few replicas
    drifts = []
    for r in (0, len(cost) // 2, len(cost) - 1):
        tot, _ = b.total_energy(r)
        drifts.append(abs(b.info(r)["energy"] - tot) / abs(tot))
    for ci, (n, _, _) in enumerate(comps):
        sweeps[ci] = moves[ci * len(temps)] * args.steps / n
    b.close()
    return dict(rate=total_moves / dt, kernel_ms=kernel_ms, seconds=dt, moves=total_moves, swaps=swaps,
                max_drift=float(max(drifts)), sweeps_per_composition=[round(float(s), 1) for s in sweeps])


def bench_plain_mix(comps, per_replica_moves=300):
    """Plain FP64 one-move-per-launch path on replicas of the same mix (one engine per
    composition at T_ref; the per-move cost does not depend on T)."""
    total, t = 0, 0.0
    for n, conc, sysd in comps:
        r = bench_plain_single(sysd, per_replica_moves + 60)
        total += per_replica_moves
        t += per_replica_moves / r["rate"]
    return total / t


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


def bench_water(args, fp32=True, replicas=64, n_mol=4096, moves_per_step=3000):
    """64 rigid-water replicas (4096 molecules, L=49.68 A, Ewald alpha=0.28, Rc=10,
    kmax=9, LJ on O truncated at 10 A, cell list with cells of at least Rc /
    args.water_cell_divisions) on a 280-360 K ladder with molecular translate+rotate
    moves (dp 0.5 A, dprot 0.5 rad), one persistent launch per step and one round of
    neighbour swaps after it."""
    import faunus_amd as F
    from tests import systems
    comps = [systems.water(n_mol=n_mol, seed=100 + c) for c in range(4)]
    temps = np.linspace(280.0, 360.0, replicas)
    b = F.Batch(fp32_pairs=fp32, cell_divisions=args.water_cell_divisions)
    for r in range(replicas):
        cfg, xyz, q, ids, g, kinds = comps[r % len(comps)]
        b.add((cfg, xyz, q, ids, g), beta=298.15 / temps[r], seed=9000 + r, ladder=0, moves=kinds)
    b.set_timing(True)
    b.run([moves_per_step] * replicas)
    b.sync()
    kernel_ms = 0.0
    swaps = (0, 0)
    t0 = time.perf_counter()
    for step in range(args.steps):
        b.run([moves_per_step] * replicas)
        kernel_ms += b.sync()
        swaps = b.swap(seed=77, rnd=step)
    dt = time.perf_counter() - t0
    check_replicas(b, replicas, "water replicas")
    total = replicas * moves_per_step * args.steps
    drift = []
    for r in (0, replicas - 1):
        tot, _ = b.total_energy(r)
        drift.append(abs(b.info(r)["energy"] - tot) / abs(tot))
    acc = np.mean([b.info(r)["accepted"] / b.info(r)["moves"] for r in range(replicas)])
    b.close()
    return dict(rate=total / dt, kernel_ms=kernel_ms, seconds=dt, moves=total, swaps=swaps,
                max_drift=float(max(drift)), acceptance=float(acc))


def bench_plain_water(n_mol=4096, nmoves=400, seed=5):
    """Plain FP64 path on the same water system: one fused launch and one host round
    trip per molecular move (fmc_replay_moves).  Proposals (translate dp 0.5 A and
    rotate 0.5 rad about the mass centre) are drawn from the initial configuration."""
    import faunus_amd as F
    from tests import systems
    cfg, xyz, q, ids, g, kinds = systems.water(n_mol=n_mol, seed=100)
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
        out.append((float(arrivals[j]), nb, nt))
    return out


def bench_serve(n_jobs=48, sweeps=4, fp32=True, quantum=5000, seed=1, split_atoms=2048, max_parts=16):
    """Continuous batching under a Poisson stream of polyelectrolyte jobs (chain of
    50-500 beads with harmonic bonds, counterions and 1:1 salt, N log-uniform in
    1k-50k, 5 mM monomer, hard spheres + Ewald alpha=5/L, kmax=8; ion and bead
    translations and pivots): submit each job at its arrival time, run scheduling
    rounds while anything is queued or resident, and report the per-job time from
    arrival to completion of `sweeps` sweeps (N moves per sweep).  Jobs above
    `split_atoms` particles are shared by ceil(N / split_atoms) workgroups (at most
    `max_parts`)."""
    import faunus_amd as F
    from tests import systems
    stream = serve_stream(n_jobs, seed)
    jobs = [systems.polyelectrolyte(nb, nt, seed=1000 + j) for j, (_, nb, nt) in enumerate(stream)]
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
            ids[nxt] = srv.submit((cfg, xyz, q, idl, g), sweeps * len(q), seed=1000 + nxt, kinds=kinds)
            nxt += 1
        r, q = srv.step(quantum)
        rounds += 1
        peak = max(peak, r)
        if r == 0 and q == 0:
            if nxt >= n_jobs:
                break
            time.sleep(max(0.0, stream[nxt][0] - (srv.status()["seconds"] - t0)))
    lat, per_sweep, moves = [], [], 0
    t_end = 0.0
    for j, jid in ids.items():
        info = srv.job(jid)
        if info["state"] != F.Server.DONE:
            raise RuntimeError(f"serving job {j} failed: {info}")
        done = info["t_done"] - t0
        lat.append(done - stream[j][0])
        per_sweep.append((info["t_done"] - info["t_admit"]) / sweeps)
        moves += info["moves"]
        t_end = max(t_end, done)
    srv.close()
    lat = np.array(lat)
    ns = np.array([len(sysd[2]) for sysd in jobs])
    return dict(jobs=n_jobs, sweeps_per_job=sweeps, split_atoms=split_atoms, max_parts=max_parts, p50_s=float(np.percentile(lat, 50)),
                p99_s=float(np.percentile(lat, 99)), moves_per_s=moves / (t_end - stream[0][0]),
                wall_s=t_end - stream[0][0], rounds=rounds, peak_resident=peak,
                n_min=int(ns.min()), n_median=int(np.median(ns)), n_max=int(ns.max()),
                projected_p50_s_to_1000_sweeps=float(np.percentile(per_sweep, 50) * 1000),
                projected_p99_s_to_1000_sweeps=float(np.percentile(per_sweep, 99) * 1000))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the 256-replica mix")
    ap.add_argument("--only", choices=["device", "plain", "batch", "split", "pipe", "headline", "copy", "serve",
                                       "gsplit",
                                       "water"], default=None)
    ap.add_argument("--serve-jobs", type=int, default=32)
    ap.add_argument("--serve-sweeps", type=int, default=2)
    ap.add_argument("--serve-split-atoms", type=int, default=2048,
                    help="particles per workgroup of a serving job (0: one workgroup per job)")
    ap.add_argument("--serve-max-parts", type=int, default=16)
    ap.add_argument("--water-cell-divisions", type=int, default=2,
                    help="water replicas: cells of at least Rc / d (1: 27-cell stencils, 2: 125)")
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
    if args.only in (None, "headline"):
        out["split64"] = bench_split(sysd, args, HEADLINE_GROUPS, HEADLINE_STRIDE)
        log("split FP64 (G=%d): %.0f moves/s, drift %.2e" % (HEADLINE_GROUPS, out["split64"]["rate"],
                                                              out["split64"]["drift"]))
    if args.only in (None, "plain"):
        out["plain"] = bench_plain_single(sysd, PLAIN_MOVES)
        log("plain FP64 one move per launch: %.0f moves/s, kernel %.1f us" % (out["plain"]["rate"], out["plain"]["kernel_us"]))
    if args.only in (None, "device"):
        out["device64"] = bench_device_single(sysd, args, fp32=False)
        log("device-resident FP64, one workgroup: %.0f moves/s (drift %.2e)" % (out["device64"]["rate"],
                                                                                 out["device64"]["drift"]))
        out["device32"] = bench_device_single(sysd, args, fp32=True)
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
    if args.only in (None, "batch") and not args.no_batch:
        comps, temps = salt_mix()
        budget = 4000
        out["batch32"] = bench_batch(comps, temps, True, args, budget)
        log("batched FP32-pair/FP64-Q: %.3e replica-moves/s" % out["batch32"]["rate"])
        out["batch64"] = bench_batch(comps, temps, False, args, budget)
        log("batched FP64: %.3e replica-moves/s" % out["batch64"]["rate"])
        out["plain_mix"] = bench_plain_mix(comps)
        log("plain FP64 one move per launch on the mix: %.0f replica-moves/s" % out["plain_mix"])
    if args.only in (None, "water"):
        out["water32"] = bench_water(args, fp32=True)
        log("64 water replicas, FP32 pairs: %.3e replica-moves/s (acceptance %.2f)" % (
            out["water32"]["rate"], out["water32"]["acceptance"]))
        out["water64"] = bench_water(args, fp32=False)
        log("64 water replicas, FP64: %.3e replica-moves/s" % out["water64"]["rate"])
        out["water_plain"] = bench_plain_water()
        log("plain FP64 per-move path on water: %.0f moves/s (kernel %.1f us)" % (
            out["water_plain"]["rate"], out["water_plain"]["kernel_us"]))
    if args.only in (None, "gsplit"):
        out["gsplit"] = bench_general_split()
    if args.only in (None, "serve"):
        out["serve"] = bench_serve(n_jobs=args.serve_jobs, sweeps=args.serve_sweeps,
                                   split_atoms=args.serve_split_atoms, max_parts=args.serve_max_parts)
        log("serving: p50 %.2f s, p99 %.2f s to %d sweeps, %.3e moves/s" % (
            out["serve"]["p50_s"], out["serve"]["p99_s"], out["serve"]["sweeps_per_job"], out["serve"]["moves_per_s"]))
    if args.only is not None:
        print(json.dumps({k: v for k, v in out.items()}, default=float))
        return
    cpu = None if args.no_cpu else cpu_baseline()
    d = out["split64"]
    kernel_s_per_move = d["kernel_ms"] / 1e3 / d["moves"]
    flops = ewald_flops_per_move(K)
    algo_bytes = 64.0 * K + pair_bytes_cells(n, 255.0, 12.0)
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
        "config": {"workload": "N=20000 Na/Cl LJ(LB, cut 12 A)+Ewald(alpha 0.25, Rc 12, kmax 20, K=17328), "
                               "L=255 A, translate dp=2 A, device-resident Metropolis, %d cooperating "
                               "workgroups of 512 threads (k-vector slices)" % HEADLINE_GROUPS,
                   "step_moves": STEP_MOVES, "acceptance": d["acceptance"], "energy_drift": d["drift"]},
        "roofline": {"bound": "hbm", "achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved_gbs / HBM_PEAK_GBS, "traffic": HEADLINE_TRAFFIC_BYTES,
                     "traffic_unit": "bytes per launch (%d moves)" % STEP_MOVES, "kernel": "mc_split_kernel",
                     "algorithmic_bytes_per_move": algo_bytes, "kernel_us_per_move": kernel_s_per_move * 1e6,
                     "fp64_tflops": achieved_tf, "fp64_peak_tflops": FP64_PEAK_TFLOPS,
                     "algorithmic_flops_per_move": flops,
                     "note": "one launch runs STEP_MOVES sequential moves; achieved = algorithmic bytes per move "
                             "/ kernel time per move (HIP events); Q(k) stays in registers, so the measured HBM "
                             "traffic (12 KB per move) is far below the algorithmic bytes (1.11 MB per move); "
                             "the move time is a chain of latencies"},
        "cpu_baseline": cpu,
        "forms": {
            "plain_fp64_one_move_per_launch": {"moves_per_s": out["plain"]["rate"],
                                               "kernel_us": out["plain"]["kernel_us"]},
            "device_fp64_split_%d_workgroups" % HEADLINE_GROUPS: {"moves_per_s": d["rate"]},
            "device_fp64_one_workgroup": {"moves_per_s": out["device64"]["rate"]},
            "device_fp32_pairs_one_workgroup": {"moves_per_s": out["device32"]["rate"]},
        },
    }
    if "batch32" in out:
        line["forms"]["batched_256_replicas"] = {
            "fp32_pairs_fp64_q_replica_moves_per_s": out["batch32"]["rate"],
            "fp64_replica_moves_per_s": out["batch64"]["rate"],
            "plain_fp64_replica_moves_per_s": out["plain_mix"],
            "speedup_fp32_batched_vs_plain": out["batch32"]["rate"] / out["plain_mix"],
            "swaps_attempted_accepted": out["batch32"]["swaps"],
            "sweeps_per_composition": out["batch32"]["sweeps_per_composition"],
            "max_energy_drift": out["batch32"]["max_drift"]}
    if "water32" in out:
        line["forms"]["batched_64_water_replicas"] = {
            "fp32_pairs_fp64_q_replica_moves_per_s": out["water32"]["rate"],
            "fp64_replica_moves_per_s": out["water64"]["rate"],
            "plain_fp64_moves_per_s": out["water_plain"]["rate"],
            "speedup_fp32_batched_vs_plain": out["water32"]["rate"] / out["water_plain"]["rate"],
            "acceptance": out["water32"]["acceptance"], "swaps_attempted_accepted": out["water32"]["swaps"],
            "max_energy_drift": out["water32"]["max_drift"], "cell_divisions": args.water_cell_divisions}
    if "serve" in out:
        line["forms"]["serving_poisson_50_per_s"] = out["serve"]
    if "gsplit" in out:
        line["forms"]["general_split_polyelectrolyte_43k"] = out["gsplit"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
