"""Failure handling and checkpoint/resume of device-resident replicas.

Cell-list overflow: with the cell capacity fixed at the fullest initial cell, the
first accepted move into a full cell is predicted on the host from the oracle's chain
(same Philox stream, same cell arithmetic).  The replica must stop right after that
move with error 1 -- the one-workgroup kernel and the split kernel, whose other
workgroups stop at the next exchange -- with the oracle's decisions, positions and
move count up to there, and every launch must end normally.  Resuming the replica
from its own checkpoint rebuilds the cell list and continues the oracle's chain.

Checkpoint/resume: a chain of 2M moves equals M moves, a checkpoint (positions, mass
centres, Philox move counter), a resume in a new batch and M more moves, bit for bit
(SaveState, analysis.cpp:645-680; MetropolisMonteCarlo::restore, montecarlo.cpp:118-137)."""
import numpy as np
import pytest

import faunus_amd as F
from oracle import oracle as O
from tests import systems
from tests.gpu_helpers import compare_chain, log, DRIFT_FP32

pytestmark = pytest.mark.gpu


def oracle_of(cfg, xyz, q, ids, g):
    s = O.System(cfg)
    s.set_particles(xyz, q, ids)
    s.set_groups(g["begin"], g["size"], g["molid"])
    return s, s.init()


def cell_index(p, box, nc):
    """The device's cell of a position: floor((x + L/2) * (nc / L)) mod nc per axis."""
    c = np.floor((np.asarray(p) + 0.5 * box) * (nc / box)).astype(np.int64) % nc
    return int((c[2] * nc + c[1]) * nc + c[0])


def predict_overflow(xyz0, box, nc, cap, rec):
    """First move whose accepted particle enters a cell already holding `cap` members."""
    cell_of = [cell_index(p, box, nc) for p in xyz0]
    count = np.bincount(cell_of, minlength=nc ** 3)
    for t in range(len(rec["du"])):
        if not rec["accepted"][t]:
            continue
        i = int(rec["index"][t])
        cnew = cell_index(rec["xyz"][t], box, nc)
        if cnew == cell_of[i]:
            continue
        if count[cnew] >= cap:
            return t
        count[cell_of[i]] -= 1
        count[cnew] += 1
        cell_of[i] = cnew
    return None


def small_cell_system():
    # L = 40 A and a 12 A pair range: a 3x3x3 grid
    return systems.nacl_lj(n_pairs=300, box=40.0, seed=3, kmax=6, dp=4.0)


@pytest.mark.parametrize("split", [False, True])
def test_cell_overflow_stops_the_replica_and_resume_continues(split):
    sysd = small_cell_system()
    cfg, xyz, q, ids, g = sysd
    box, nc = 40.0, 3
    cap = int(np.bincount([cell_index(p, box, nc) for p in xyz], minlength=nc ** 3).max())
    nm, seed = 3000, 123
    orc, e0 = oracle_of(*sysd)
    rec = orc.run_philox_translate(nm, seed=seed, replica=0)
    t_star = predict_overflow(xyz, box, nc, cap, rec)
    assert t_star is not None and t_star + 1 < nm, "the chain must overflow a full cell"
    b = F.Batch(cell_capacity=cap)
    b.add(sysd, seed=seed)
    print("cell overflow split=%s cap=%d predicted move %d" % (split, cap, t_star), flush=True)
    if split:
        b.run_split(nm, groups=8, stride=1, trace=nm)
    else:
        b.run([nm], trace=nm)
    b.sync()
    info = b.info(0)
    log("overflow", info, "split error", b.split_error())
    assert info["error"] == 1 and info["moves"] == t_star + 1, (info, t_star)
    assert b.split_error() == 1
    done = t_star + 1
    sub = {k: v[:done] for k, v in rec.items()}
    rep = compare_chain({k: v[:done] for k, v in b.trace(0, done).items()}, sub, max(1.0, abs(e0)), 1e-10, seed=seed)
    assert rep["first_mismatch"] == done and rep["max_rel"] <= 1e-10, rep
    orc2, _ = oracle_of(*sysd)
    orc2.run_philox_translate(done, seed=seed, replica=0)
    np.testing.assert_allclose(b.positions(0), orc2.positions(), rtol=0, atol=1e-9)
    # a stopped replica refuses further moves until its state is set again
    b.run([100])
    b.sync()
    assert b.info(0)["moves"] == done
    # resume from its own checkpoint: the cell list is rebuilt with the automatic capacity
    st = b.get_state(0)
    assert st["moves"] == done
    b.set_state(0, st)
    rest = nm - done
    if split:
        b.run_split(rest, groups=8, stride=1, trace=rest)
    else:
        b.run([rest], trace=rest)
    b.sync()
    info = b.info(0)
    assert info["error"] == 0 and info["moves"] == nm, info
    rec2 = orc2.run_philox_translate(rest, seed=seed, replica=0, move0=done)
    rep2 = compare_chain(b.trace(0, rest), rec2, max(1.0, abs(e0)), 1e-10, seed=seed, move0=done)
    log("resumed after overflow", rep2)
    assert rep2["first_mismatch"] == rest and rep2["max_rel"] <= 1e-10, rep2
    np.testing.assert_allclose(b.positions(0), orc2.positions(), rtol=0, atol=1e-9)
    tot, _ = b.total_energy(0)
    assert abs(info["energy"] - tot) <= 1e-9 * abs(tot)


def checkpoint_roundtrip(sysd, kinds, m, seed, fp32=False, split_groups=0):
    def make():
        b = F.Batch(fp32_pairs=fp32)
        b.add(sysd, seed=seed, moves=kinds)
        return b

    def run(b, n):
        if split_groups:
            b.run_split(n, groups=split_groups)
        else:
            b.run([n])
        b.sync()

    whole = make()
    run(whole, 2 * m)
    first = make()
    run(first, m)
    st = first.get_state(0)
    first.close()
    resumed = make()
    resumed.set_state(0, st)
    run(resumed, m)
    a, r = whole.info(0), resumed.info(0)
    log("checkpoint round trip", a, r)
    assert r["moves"] == a["moves"] == 2 * m and r["accepted"] == a["accepted"]
    assert np.array_equal(resumed.positions(0), whole.positions(0))
    # the resumed running energy starts from a full recompute; the FP32-pair form's running
    # sum carries its own rounding
    assert abs(r["energy"] - a["energy"]) <= (DRIFT_FP32 if fp32 else 1e-9) * max(1.0, abs(a["energy"]))
    assert r["error"] == 0 and a["error"] == 0
    whole.close()
    resumed.close()


def test_checkpoint_resume_translate_replica():
    checkpoint_roundtrip(systems.salt21(n=600, conc=0.1, seed=4, kmax=6), None, 1500, seed=17)


def test_checkpoint_resume_headline_split_kernel():
    sysd = systems.nacl_lj(n_pairs=1000, box=255.0 * (2000 / 20000) ** (1 / 3), seed=2, kmax=10)
    checkpoint_roundtrip(sysd, None, 1500, seed=18, split_groups=8)


@pytest.mark.parametrize("fp32", [False, True])
def test_checkpoint_resume_water_replica(fp32):
    cfg, xyz, q, ids, g, kinds = systems.water(n_mol=1000, seed=3, kmax=6, cutoff=10.0)
    checkpoint_roundtrip((cfg, xyz, q, ids, g), kinds, 800, seed=19, fp32=fp32)


def test_checkpoint_resume_polyelectrolyte_split():
    cfg, xyz, q, ids, g, kinds = systems.polyelectrolyte(40, 400, seed=7, kmax=6)
    checkpoint_roundtrip((cfg, xyz, q, ids, g), kinds, 1500, seed=20, split_groups=4)
