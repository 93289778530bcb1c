"""Shared helpers for the GPU parity tests: build the same system in the oracle and
in the HIP engine, and compare per-move energies with a stated tolerance."""
import sys

import numpy as np

from oracle import oracle as O

# FP64 path: |gpu - cpu| <= REL * max(|cpu|, E_scale) with E_scale >= 1 kT, the
# absolute floor needed because the reference forms dU as a difference of
# whole-system totals (montecarlo.cpp:153-157).
REL_FP64 = 1e-10

# The FP32-pair form's running energy sums the FP32 rounding of every accepted dU (each
# within about 1e-8 |E| on these systems, far inside the 1e-6 per-move bar), so after a
# few thousand moves a full FP64 recompute may differ from it by up to this fraction of
# |E|.  The chain itself (decisions, dU) is held to the per-move bar.
DRIFT_FP32 = 1e-5


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def build_pair(system, engine_cls):
    cfg, xyz, q, ids, g = system
    orc = O.System(cfg)
    orc.set_particles(xyz, q, ids)
    orc.set_groups(g["begin"], g["size"], g["molid"])
    e_cpu = orc.init()
    eng = engine_cls(cfg)
    eng.set_particles(xyz, q, ids)
    eng.set_groups(g["begin"], g["size"], g["molid"])
    e_gpu = eng.init()
    return orc, eng, e_cpu, e_gpu


def close(a, b, rel=REL_FP64, floor=1.0):
    """|a - b| <= rel * max(|b|, floor) where both are finite; equal infinities pass."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    both_inf = np.isinf(a) & np.isinf(b) & (np.sign(a) == np.sign(b))
    fin = np.isfinite(a) & np.isfinite(b)
    af, bf = np.where(fin, a, 0.0), np.where(fin, b, 0.0)
    scale = np.maximum(np.abs(bf), floor)
    return both_inf | (fin & (np.abs(af - bf) <= rel * scale))


def acceptance_deviate(seed, replica, move):
    """The 53-bit uniform deviate of the acceptance test of a device-resident move:
    Philox4x32-10 block (move_lo, move_hi, replica, 1) under key (seed_lo, seed_hi)."""
    b = O.philox([move & 0xFFFFFFFF, move >> 32, replica, 1], [seed & 0xFFFFFFFF, seed >> 32])
    return ((b[0] >> 5) * 67108864.0 + (b[1] >> 6)) * 1.1102230246251565e-16


def metropolis_accepts(du, u, beta=1.0):
    """metropolisCriterion (montecarlo.cpp:17-34) with the beta-scaled energy change."""
    x = beta * du
    if np.isnan(x):
        return False
    if np.isinf(x) and x < 0:
        return True
    if -x > 709.782712893384:
        return True
    with np.errstate(over="ignore"):
        return bool(u <= np.exp(-x))


def compare_chain(tr, rec, e_scale, rel, seed, replica=0, move0=0, beta=1.0):
    """Compare a device trace with the oracle's run of the same Philox stream.

    Every du up to and including the first move whose decision differs must agree
    within rel * max(|du|, e_scale) (equal infinities pass), and that first differing
    decision must be a near-tie: each side's decision is the Metropolis outcome of its
    own du with the move's acceptance deviate, so the two du straddle the threshold
    while lying within the tolerance of each other.  Returns a report dict."""
    n = len(rec["du"])
    mism = np.nonzero(np.asarray(tr["accepted"]) != np.asarray(rec["accepted"]))[0]
    first = int(mism[0]) if len(mism) else n
    upto = min(n, first + 1)
    if "index" in rec and len(rec["index"]) == n:
        assert np.array_equal(tr["index"][:upto], rec["index"][:upto]), "proposed particles differ"
    dg, dc = np.asarray(tr["du"][:upto], float), np.asarray(rec["du"][:upto], float)
    assert np.array_equal(np.isinf(dg), np.isinf(dc)), "infinite du differ"
    fin = np.isfinite(dc) & np.isfinite(dg)
    err = np.abs(dg[fin] - dc[fin]) / np.maximum(np.abs(dc[fin]), e_scale)
    tie = None
    if first < n:
        u = acceptance_deviate(seed, replica, move0 + first)
        own_g = metropolis_accepts(dg[first], u, beta) == bool(tr["accepted"][first])
        own_c = metropolis_accepts(dc[first], u, beta) == bool(rec["accepted"][first])
        tie = dict(move=first, u=u, du_gpu=float(dg[first]), du_oracle=float(dc[first]))
        assert own_g and own_c, ("the first differing decision is not a near-tie", tie)
        if np.isfinite(dc[first]) and np.isfinite(dg[first]):
            assert abs(dg[first] - dc[first]) <= rel * max(abs(dc[first]), e_scale), ("tie beyond tolerance", tie)
    return dict(first_mismatch=first, n=n, max_rel=float(err.max()) if len(err) else 0.0,
                accepted=int(np.asarray(rec["accepted"][:first]).sum()), mismatches=len(mism), tie=tie)


def compare_replay(rec, out, nterm, rel=REL_FP64):
    """Compare a GPU replay with the oracle's recorded run; returns a report dict."""
    en_c, eo_c = rec["e_new"], rec["e_old"]
    en_g, eo_g = out["e_new"], out["e_old"]
    scale = np.maximum(np.abs(en_c).sum(1), np.abs(eo_c).sum(1))
    scale = np.where(np.isfinite(scale), scale, 1.0)
    ok_terms = close(en_g, en_c, rel, floor=1.0) & close(eo_g, eo_c, rel, floor=1.0)
    du_c, du_g = rec["du"], out["du"]
    fin = np.isfinite(du_c) & np.isfinite(du_g)
    diff = np.abs(np.where(fin, du_g, 0.0) - np.where(fin, du_c, 0.0))
    ok_du = np.where(fin, diff <= rel * np.maximum(np.abs(np.where(fin, du_c, 0.0)), np.maximum(scale, 1.0)),
                     du_c == du_g)
    mism = np.nonzero(rec["accepted"] != out["accepted"])[0]
    # a decision may differ only at a near-tie of u and exp(-du)
    ties = [int(i) for i in mism if np.isfinite(du_c[i]) and abs(rec["u"][i] - np.exp(-du_c[i])) < 1e-9]
    with np.errstate(invalid="ignore", divide="ignore"):
        rel_err = np.where(fin & (scale > 0), diff / np.maximum(np.abs(np.where(fin, du_c, 0.0)), scale), 0.0)
    return dict(term_fail=int((~ok_terms).sum()), du_fail=int((~ok_du).sum()), decision_mismatch=len(mism),
                near_ties=ties, max_rel_du=float(np.max(rel_err)) if len(rel_err) else 0.0,
                accepted=int(out["accepted"].sum()))
