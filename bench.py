"""Benchmark of the Metropolis energy path on one MI355X.

Workload (default): N = 20 000 Na/Cl ions (q=+-1, LJ sigma 3.3/4.4 A, eps 0.5/0.4
kJ/mol, Lorentz-Berthelot, LJ truncated at 12 A), cubic L = 255 A, Ewald alpha=0.25,
Rc=12 A, kmax=20 (K=17 328), single-particle translations dp = 2 A.  A step is
STEP_MOVES trial moves; the metric is trial moves per second.

Prints one JSON line (rank 0) with the roofline of the dominant kernel (HIP
events on the engine's stream) and a CPU baseline (the FP64 oracle, one core, on a
bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
STEP_MOVES = 1000


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def nacl_system(seed=1):
    from tests import systems
    return systems.nacl_lj(n_pairs=10000, box=255.0, seed=seed)


def translate_proposals(n_moves, n_particles, dp, rng):
    """Random particle, random unit vector (uniform on the sphere) times dp*U(0,1)."""
    idx = rng.integers(0, n_particles, size=n_moves).astype(np.int32)
    v = rng.normal(size=(n_moves, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    disp = v * (dp * rng.random(n_moves))[:, None]
    u = rng.random(n_moves)
    return idx, disp, u


def cell_candidates(eng_box, n, cutoff):
    nc = int(eng_box // cutoff)
    return 27.0 * n / nc ** 3


def bench_plain(args):
    import faunus_amd as F
    cfg, xyz, q, ids, g = nacl_system()
    eng = F.Engine(cfg)
    eng.set_particles(xyz, q, ids)
    eng.set_groups(g["begin"], g["size"], g["molid"])
    eng.init()
    K = eng.num_kvectors
    n = len(q)
    rng = np.random.default_rng(7)
    total = (args.warmup + args.steps) * STEP_MOVES
    idx, disp, u = translate_proposals(total, n, 2.0, rng)
    grp = np.zeros(total, np.int32)
    # warmup
    w = args.warmup * STEP_MOVES
    if w:
        eng.run_translate(grp[:w], idx[:w], disp[:w], u[:w])
    eng.set_timing(True)
    t0 = time.perf_counter()
    du, acc = eng.run_translate(grp[w:], idx[w:], disp[w:], u[w:])
    t1 = time.perf_counter()
    kern_ms, launches = eng.timing()
    eng.set_timing(False)
    moves = args.steps * STEP_MOVES
    rate = moves / (t1 - t0)
    avg_kernel_s = kern_ms / 1e3 / max(launches, 1)
    n_cand = cell_candidates(255.0, n, 12.0)
    bytes_pair = n_cand * 36 + 2 * 27 * 8
    bytes_ewald = 64.0 * K
    algo_bytes = bytes_pair + bytes_ewald
    achieved = algo_bytes / avg_kernel_s / 1e9
    dq, drec = eng.check_q()
    log(f"plain: {rate:.0f} moves/s, acceptance {acc.mean():.3f}, kernel {avg_kernel_s*1e6:.2f} us/launch, "
        f"{achieved:.1f} GB/s algorithmic, Q drift {dq:.2e}")
    return dict(rate=rate, ms_per_step=(t1 - t0) * 1e3 / args.steps, acceptance=float(acc.mean()),
                kernel_us=avg_kernel_s * 1e6, achieved=achieved, algo_bytes=algo_bytes, K=K, n=n,
                q_drift=dq)


def cpu_baseline(seconds=15.0):
    """FP64 oracle (reference Release flags), 1 core, single-particle translations on
    the same N = 20 000 system; bounded sample of moves."""
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
                sample=f"{nm} single-particle translations of the N=20000 system "
                       f"(oracle FP64 restatement, all-pairs as in the reference, 1 thread, {dt:.1f} s)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU layout is described "
              "in DESIGN.md", file=sys.stderr)
        sys.exit(2)
    import faunus_amd as F
    if F.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    res = bench_plain(args)
    cpu = None if args.no_cpu else cpu_baseline()
    line = {
        "metric": "trial moves/s at N=20k Ewald",
        "value": res["rate"],
        "unit": "trial moves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (random sequential addition, seed 1)",
        "config": {"workload": "N=20000 Na/Cl LJ(cut 12)+Ewald(alpha 0.25, Rc 12, kmax 20, K=%d), L=255, "
                               "translate dp=2, plain FP64 one move per launch" % res["K"],
                   "step_moves": STEP_MOVES, "acceptance": res["acceptance"]},
        "roofline": {"bound": "hbm", "achieved": res["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": res["achieved"] / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "delta_kernel", "kernel_us": res["kernel_us"],
                     "algorithmic_bytes_per_launch": res["algo_bytes"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
