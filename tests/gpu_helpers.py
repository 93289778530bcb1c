"""Shared helpers for the GPU parity tests: build the same system in the oracle and
in the HIP engine, and compare per-move energies with a stated tolerance."""
import sys

import numpy as np

from oracle import oracle as O

# FP64 path: |gpu - cpu| <= REL * max(|cpu|, E_scale) with E_scale >= 1 kT, the
# absolute floor needed because the reference forms dU as a difference of
# whole-system totals (montecarlo.cpp:153-157).
REL_FP64 = 1e-10


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
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    both_inf = np.isinf(a) & np.isinf(b) & (np.sign(a) == np.sign(b))
    scale = np.maximum(np.maximum(np.abs(b), floor), 0)
    ok = both_inf | (np.abs(a - b) <= rel * scale)
    return ok


def compare_replay(rec, out, nterm, rel=REL_FP64):
    """Compare a GPU replay with the oracle's recorded run; returns a report dict."""
    en_c, eo_c = rec["e_new"], rec["e_old"]
    en_g, eo_g = out["e_new"], out["e_old"]
    scale = np.maximum(np.abs(en_c).sum(1), np.abs(eo_c).sum(1))
    scale = np.where(np.isfinite(scale), scale, 1.0)
    ok_terms = close(en_g, en_c, rel, floor=1.0) & close(eo_g, eo_c, rel, floor=1.0)
    du_c, du_g = rec["du"], out["du"]
    fin = np.isfinite(du_c) & np.isfinite(du_g)
    ok_du = np.where(fin, np.abs(du_g - du_c) <= rel * np.maximum(np.abs(du_c), np.maximum(scale, 1.0)),
                     du_c == du_g)
    mism = np.nonzero(rec["accepted"] != out["accepted"])[0]
    # a decision may differ only at a near-tie of u and exp(-du)
    ties = [int(i) for i in mism if np.isfinite(du_c[i]) and abs(rec["u"][i] - np.exp(-du_c[i])) < 1e-9]
    with np.errstate(invalid="ignore"):
        rel_err = np.where(fin & (scale > 0), np.abs(du_g - du_c) / np.maximum(np.abs(du_c), scale), 0.0)
    return dict(term_fail=int((~ok_terms).sum()), du_fail=int((~ok_du).sum()), decision_mismatch=len(mism),
                near_ties=ties, max_rel_du=float(np.max(rel_err)) if len(rel_err) else 0.0,
                accepted=int(out["accepted"].sum()))
