"""Device-resident Metropolis (one workgroup per replica) against the oracle running
the same Philox proposal stream.  FP64: every du within 1e-10 * max(|du|, E_scale)
and identical decisions (a decision may differ only at a logged near-tie); FP32
pair terms: du within 1e-6 * max(|du|, E_scale) up to the first near-tie."""
import numpy as np
import pytest

import faunus_amd as F
from oracle import oracle as O
from tests import systems
from tests.gpu_helpers import compare_chain, log, DRIFT_FP32

pytestmark = pytest.mark.gpu


def oracle_of(system):
    cfg, xyz, q, ids, g = system
    s = O.System(cfg)
    s.set_particles(xyz, q, ids)
    s.set_groups(g["begin"], g["size"], g["molid"])
    e0 = s.init()
    return s, e0


def compare_traces(tr, rec, e_scale, rel, seed, replica=0, move0=0, beta=1.0):
    return compare_chain(tr, rec, e_scale, rel, seed, replica=replica, move0=move0, beta=beta)


@pytest.mark.parametrize("fp32", [False, True])
def test_device_mc_primitive_model(fp32):
    sysd = systems.primitive_model(n_pairs=100, box=118.0, seed=1)
    orc, e0 = oracle_of(sysd)
    b = F.Batch(fp32_pairs=fp32)
    b.add(sysd, beta=1.0, seed=12345)
    nm = 4000
    b.run([nm], trace=nm)
    b.sync()
    tr = b.trace(0, nm)
    rec = orc.run_philox_translate(nm, seed=12345, replica=0)
    e_scale = max(1.0, abs(e0))
    rel = 1e-6 if fp32 else 1e-10
    rep = compare_traces(tr, rec, e_scale, rel, seed=12345)
    log("device MC primitive model fp32=%s" % fp32, rep)
    if not fp32:
        assert rep["first_mismatch"] == nm, rep
        assert rep["max_rel"] <= 1e-10
        np.testing.assert_allclose(b.positions(0), orc.positions(), rtol=0, atol=1e-9)
    else:
        assert rep["max_rel"] <= 1e-6, rep
        assert rep["first_mismatch"] > nm // 4, rep
    info = b.info(0)
    assert info["moves"] == nm and info["error"] == 0
    # running energy equals a full recompute; Q(k) equals the full update
    tot, terms = b.total_energy(0)
    log("running", info["energy"], "recomputed", tot)
    # FP32 pair terms accumulate ~1e-7-relative errors per move into the running sum
    drift_tol = DRIFT_FP32 if fp32 else 1e-9
    assert abs(info["energy"] - tot) <= drift_tol * max(1.0, abs(tot))
    dq, drec = b.check_q(0)
    assert dq < 1e-9 and abs(drec) < 1e-11


def test_device_mc_cell_list_nacl():
    box = 255.0 * (2000 / 20000) ** (1 / 3)
    sysd = systems.nacl_lj(n_pairs=1000, box=box, seed=2, kmax=10)
    orc, e0 = oracle_of(sysd)
    b = F.Batch()
    b.add(sysd, seed=99)
    nm = 3000
    b.run([nm], trace=nm)
    b.sync()
    tr = b.trace(0, nm)
    rec = orc.run_philox_translate(nm, seed=99, replica=0)
    rep = compare_traces(tr, rec, max(1.0, abs(e0)), 1e-10, seed=99)
    log("device MC cell list", rep)
    assert rep["first_mismatch"] == nm and rep["max_rel"] <= 1e-10
    tot, _ = b.total_energy(0)
    assert abs(b.info(0)["energy"] - tot) <= 1e-9 * abs(tot)


def test_device_mc_full_size_nacl_short():
    """N=20 000, K=17 328: 68 k-vectors per thread in registers."""
    sysd = systems.nacl_lj(n_pairs=10000, box=255.0, seed=1)
    orc, e0 = oracle_of(sysd)
    b = F.Batch()
    b.add(sysd, seed=7)
    nm = 200
    b.run([nm], trace=nm)
    b.sync()
    tr = b.trace(0, nm)
    rec = orc.run_philox_translate(nm, seed=7, replica=0)
    rep = compare_traces(tr, rec, max(1.0, abs(e0)), 1e-10, seed=7)
    log("device MC full size", rep)
    assert rep["first_mismatch"] == nm and rep["max_rel"] <= 1e-10
    dq, drec = b.check_q(0)
    assert dq < 1e-9


@pytest.mark.parametrize("groups,stride", [(3, 1), (4, 1), (16, 8), (64, 1)])
def test_split_replica_matches_oracle(groups, stride):
    """One replica spread over several cooperating workgroups (k-vector slices)."""
    box = 255.0 * (2000 / 20000) ** (1 / 3)
    sysd = systems.nacl_lj(n_pairs=1000, box=box, seed=2, kmax=10)
    orc, e0 = oracle_of(sysd)
    b = F.Batch()
    b.add(sysd, seed=31)
    nm = 1500
    b.run_split(nm, groups=groups, stride=stride, trace=nm)
    b.sync()
    assert b.split_error() == 0
    tr = b.trace(0, nm)
    rec = orc.run_philox_translate(nm, seed=31, replica=0)
    rep = compare_traces(tr, rec, max(1.0, abs(e0)), 1e-10, seed=31)
    log("split", groups, stride, rep)
    assert rep["first_mismatch"] == nm and rep["max_rel"] <= 1e-10
    # a second launch continues the same chain
    b.run_split(500, groups=groups, stride=stride, trace=500)
    b.sync()
    tr2 = b.trace(0, 500)
    rec2 = orc.run_philox_translate(500, seed=31, replica=0, move0=nm)
    rep2 = compare_traces(tr2, rec2, max(1.0, abs(e0)), 1e-10, seed=31, move0=nm)
    assert rep2["first_mismatch"] == 500
    tot, _ = b.total_energy(0)
    assert abs(b.info(0)["energy"] - tot) <= 1e-9 * abs(tot)
    dq, _ = b.check_q(0)
    assert dq < 1e-9


@pytest.mark.parametrize("groups", [1, 3, 8])
def test_split_without_cell_list_matches_oracle(groups):
    """The split kernel on a system without a cell list (2:1 salt with Rc = L/2):
    workgroup 0 sums all pairs, the others the k-vector slices."""
    sysd = systems.salt21(n=600, conc=0.1, seed=4, kmax=8)
    orc, e0 = oracle_of(sysd)
    b = F.Batch()
    b.add(sysd, seed=77)
    nm = 2000
    b.run_split(nm, groups=groups, stride=1, trace=nm)
    b.sync()
    assert b.split_error() == 0
    tr = b.trace(0, nm)
    rec = orc.run_philox_translate(nm, seed=77, replica=0)
    rep = compare_traces(tr, rec, max(1.0, abs(e0)), 1e-10, seed=77)
    log("split without cells", groups, rep)
    assert rep["first_mismatch"] == nm and rep["max_rel"] <= 1e-10
    np.testing.assert_allclose(b.positions(0), orc.positions(), rtol=0, atol=1e-9)


@pytest.mark.parametrize("groups,depth,kind", [(4, 1, "cells"), (16, 2, "cells"), (16, 4, "cells"),
                                               (64, 3, "cells"), (8, 4, "pm200"), (16, 4, "salt")])
def test_pipelined_replica_matches_oracle(groups, depth, kind):
    """Moves evaluated ahead of their predecessors' decisions; exact decisions."""
    if kind == "cells":
        box = 255.0 * (2000 / 20000) ** (1 / 3)
        sysd = systems.nacl_lj(n_pairs=1000, box=box, seed=2, kmax=10)
    elif kind == "pm200":  # N=200: proposals often hit a pending particle (pipeline drains)
        sysd = systems.primitive_model(n_pairs=100, box=118.0, seed=1)
    else:  # no cell list (Rc = L/2): all-pairs branch
        sysd = systems.salt21(n=900, conc=0.05, seed=4)
    orc, e0 = oracle_of(sysd)
    b = F.Batch()
    b.add(sysd, seed=77)
    nm = 2000
    b.run_pipe(nm, groups=groups, depth=depth, trace=nm)
    b.sync()
    assert b.split_error() == 0
    tr = b.trace(0, nm)
    rec = orc.run_philox_translate(nm, seed=77, replica=0)
    rep = compare_traces(tr, rec, max(1.0, abs(e0)), 1e-10, seed=77)
    log("pipe", groups, depth, kind, rep)
    assert rep["first_mismatch"] == nm and rep["max_rel"] <= 1e-10
    b.run_pipe(700, groups=groups, depth=depth, trace=700)
    b.sync()
    rec2 = orc.run_philox_translate(700, seed=77, replica=0, move0=nm)
    rep2 = compare_traces(b.trace(0, 700), rec2, max(1.0, abs(e0)), 1e-10, seed=77, move0=nm)
    assert rep2["first_mismatch"] == 700
    tot, _ = b.total_energy(0)
    assert abs(b.info(0)["energy"] - tot) <= 1e-9 * abs(tot)
    dq, _ = b.check_q(0)
    assert dq < 1e-9
    np.testing.assert_allclose(b.positions(0), orc.positions(), rtol=0, atol=1e-9)


def test_split_full_size():
    sysd = systems.nacl_lj(n_pairs=10000, box=255.0, seed=1)
    orc, e0 = oracle_of(sysd)
    b = F.Batch()
    b.add(sysd, seed=5)
    nm = 300
    b.run_split(nm, groups=64, stride=1, trace=nm)
    b.sync()
    assert b.split_error() == 0
    tr = b.trace(0, nm)
    rec = orc.run_philox_translate(nm, seed=5, replica=0)
    rep = compare_traces(tr, rec, max(1.0, abs(e0)), 1e-10, seed=5)
    log("split full size", rep)
    assert rep["first_mismatch"] == nm and rep["max_rel"] <= 1e-10


def test_batch_of_uneven_replicas_matches_oracle_per_replica():
    """Several 2:1 salt replicas of different N and concentration in one launch."""
    specs = [(600, 0.5, 11), (900, 0.05, 12), (1203, 0.2, 13), (300, 1.0, 14)]
    b = F.Batch()
    sysds = []
    for k, (n, c, seed) in enumerate(specs):
        sysd = systems.salt21(n=n, conc=c, seed=seed)
        sysds.append(sysd)
        b.add(sysd, seed=1000 + k)
    nm = 1500
    b.run([nm] * len(specs), trace=nm)
    b.sync()
    for r, sysd in enumerate(sysds):
        orc, e0 = oracle_of(sysd)
        rec = orc.run_philox_translate(nm, seed=1000 + r, replica=r)
        tr = b.trace(r, nm)
        rep = compare_traces(tr, rec, max(1.0, abs(e0)), 1e-10, seed=1000 + r, replica=r)
        log("batch replica", r, rep)
        assert rep["first_mismatch"] == nm and rep["max_rel"] <= 1e-10


def test_temperature_ladder_swaps_permute_betas():
    sysd = systems.salt21(n=300, conc=0.1, seed=3)
    betas = [298.15 / t for t in (280.0, 300.0, 320.0, 340.0)]
    b = F.Batch()
    for k, beta in enumerate(betas):
        b.add(sysd, beta=beta, seed=50 + k, ladder=0)
    for rnd in range(6):
        b.run([600] * 4)
        att, acc = b.swap(seed=77, rnd=rnd)
    b.sync()
    got = sorted(b.info(r)["beta"] for r in range(4))
    assert np.allclose(got, sorted(betas))
    log("ladder swaps attempted", att, "accepted", acc)
    assert att == 9
    for r in range(4):
        tot, _ = b.total_energy(r)
        assert abs(b.info(r)["energy"] - tot) <= 1e-9 * max(1.0, abs(tot))


def test_split_three_cell_grid_matches_oracle():
    """A 40 A box with 12 A cells: a 3x3x3 grid, so every stencil holds every cell and
    each accepted move changes cells that the next move's early-loaded neighbours came
    from.  Those loads are then discarded and taken again; decisions stay exact."""
    sysd = systems.nacl_lj(n_pairs=300, box=40.0, seed=3, kmax=6)
    orc, e0 = oracle_of(sysd)
    b = F.Batch()
    b.add(sysd, seed=41)
    nm = 1500
    print("split three-cell grid", flush=True)
    b.run_split(nm, groups=8, stride=1, trace=nm)
    b.sync()
    assert b.split_error() == 0
    tr = b.trace(0, nm)
    rec = orc.run_philox_translate(nm, seed=41, replica=0)
    rep = compare_traces(tr, rec, max(1.0, abs(e0)), 1e-10, seed=41)
    log("split three-cell grid", rep, "accepted", int(rec["accepted"].sum()))
    assert rep["first_mismatch"] == nm and rep["max_rel"] <= 1e-10
    tot, _ = b.total_energy(0)
    assert abs(b.info(0)["energy"] - tot) <= 1e-9 * abs(tot)
    np.testing.assert_allclose(b.positions(0), orc.positions(), rtol=0, atol=1e-9)


@pytest.mark.parametrize("groups", [3, 8])
def test_split_small_system_repeats_particles(groups):
    """N = 24 on a 3x3x3 cell grid: the next move's particle is the current one about
    once in 24 moves.  After such a move is accepted, the k-vector workgroups build the
    next move's share again and workgroup 0 loads that particle again; decisions stay
    exact against the oracle, and the final positions and energy agree."""
    sysd = systems.nacl_lj(n_pairs=12, box=40.0, seed=5, kmax=6)
    orc, e0 = oracle_of(sysd)
    b = F.Batch()
    b.add(sysd, seed=43)
    nm = 3000
    print("split small system, groups", groups, flush=True)
    b.run_split(nm, groups=groups, stride=1, trace=nm)
    b.sync()
    assert b.split_error() == 0
    tr = b.trace(0, nm)
    rec = orc.run_philox_translate(nm, seed=43, replica=0)
    rep = compare_traces(tr, rec, max(1.0, abs(e0)), 1e-10, seed=43)
    idx, acc = np.asarray(rec["index"]), np.asarray(rec["accepted"]).astype(bool)
    repeats = int(np.sum((idx[1:] == idx[:-1]) & acc[:-1]))
    log("split small system", groups, rep, "accepted repeats", repeats)
    assert repeats >= 20
    assert rep["first_mismatch"] == nm and rep["max_rel"] <= 1e-10
    tot, _ = b.total_energy(0)
    assert abs(b.info(0)["energy"] - tot) <= 1e-9 * max(1.0, abs(tot))
    np.testing.assert_allclose(b.positions(0), orc.positions(), rtol=0, atol=1e-9)
    b.close()
