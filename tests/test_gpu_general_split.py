"""Split form of the general device kernel (one replica shared by several workgroups
that exchange per-move partial sums) against the oracle running the same Philox
stream: FP64 du within 1e-10 relative to max(|du|, E_scale) with identical decisions,
final positions within 1e-9 A, running energy against a full recompute and Q(k)
against a full rebuild; FP32 pair terms within 1e-6 up to the first near-tie.  Also:
a run continued across launches and across the split and one-workgroup forms, and
split serving jobs that are bitwise the same alone and in a mix."""
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


def compare(tr, rec, e_scale, seed, replica=0, move0=0, rel=1e-6):
    return compare_chain(tr, rec, e_scale, rel, seed, replica=replica, move0=move0)


def check_fp64(b, orc, rep, label):
    log(label, rep)
    info = b.info(0)
    assert info["error"] == 0, info
    assert rep["first_mismatch"] == rep["n"] and rep["max_rel"] <= 1e-10, rep
    np.testing.assert_allclose(b.positions(0), orc.positions(), rtol=0, atol=1e-9)
    tot, _ = b.total_energy(0)
    assert abs(info["energy"] - tot) <= 1e-9 * max(1.0, abs(tot))
    dq, _ = b.check_q(0)
    assert dq < 1e-9


@pytest.mark.parametrize("groups", [4, 16])
def test_split_polyelectrolyte_fp64(groups):
    cfg, xyz, q, ids, g, kinds = systems.polyelectrolyte(40, 400, seed=7, kmax=6)
    sysd = (cfg, xyz, q, ids, g)
    orc, e0 = oracle_of(*sysd)
    nm = 3000
    b = F.Batch(fp32_pairs=False)
    b.add(sysd, seed=33, moves=kinds)
    print("split polyelectrolyte fp64 groups=%d moves=%d" % (groups, nm), flush=True)
    b.run_split(nm, groups=groups, trace=nm)
    b.sync()
    assert b.split_error() == 0
    rec = orc.run_philox_moves(nm, seed=33, kinds=kinds, replica=0)
    rep = compare(b.trace(0, nm), rec, max(1.0, abs(e0)), seed=33)
    assert b.info(0)["moves"] == nm
    check_fp64(b, orc, rep, "split polyelectrolyte groups=%d" % groups)


def test_split_polyelectrolyte_with_exclusions_fp64():
    """Bonded neighbours up to two bonds apart excluded from the nonbonded pairs; pivots
    move up to 39 beads whose exclusion lists span several workgroups' slices."""
    cfg, xyz, q, ids, g, kinds = systems.polyelectrolyte(40, 400, seed=7, kmax=6, excluded_neighbours=2)
    sysd = (cfg, xyz, q, ids, g)
    orc, e0 = oracle_of(*sysd)
    nm = 3000
    b = F.Batch(fp32_pairs=False)
    b.add(sysd, seed=34, moves=kinds)
    print("split polyelectrolyte exclusions groups=8", flush=True)
    b.run_split(nm, groups=8, trace=nm)
    b.sync()
    assert b.split_error() == 0
    rec = orc.run_philox_moves(nm, seed=34, kinds=kinds, replica=0)
    rep = compare(b.trace(0, nm), rec, max(1.0, abs(e0)), seed=34)
    check_fp64(b, orc, rep, "split polyelectrolyte exclusions")


def test_split_polyelectrolyte_fp32():
    cfg, xyz, q, ids, g, kinds = systems.polyelectrolyte(40, 400, seed=7, kmax=6)
    sysd = (cfg, xyz, q, ids, g)
    orc, e0 = oracle_of(*sysd)
    nm = 3000
    b = F.Batch(fp32_pairs=True)
    b.add(sysd, seed=33, moves=kinds)
    print("split polyelectrolyte fp32 groups=8", flush=True)
    b.run_split(nm, groups=8, trace=nm)
    b.sync()
    assert b.split_error() == 0
    rec = orc.run_philox_moves(nm, seed=33, kinds=kinds, replica=0)
    rep = compare(b.trace(0, nm), rec, max(1.0, abs(e0)), seed=33)
    log("split polyelectrolyte fp32", rep)
    assert rep["max_rel"] <= 1e-6 and rep["first_mismatch"] > 500, rep
    tot, _ = b.total_energy(0)
    assert abs(b.info(0)["energy"] - tot) <= DRIFT_FP32 * max(1.0, abs(tot))


def test_split_water_rigid_moves_continued_across_forms():
    """Rigid water (translate+rotate of molecules): the first half of the run split
    over 8 workgroups in two launches, the second half on one workgroup."""
    cfg, xyz, q, ids, g, kinds = systems.water(n_mol=216, seed=3, kmax=6, rmax=None, cutoff=9.0)
    sysd = (cfg, xyz, q, ids, g)
    orc, e0 = oracle_of(*sysd)
    b = F.Batch(fp32_pairs=False)
    b.add(sysd, seed=8, moves=kinds)
    print("split water 216 groups=8", flush=True)
    b.run_split(400, groups=8)
    b.run_split(400, groups=8)
    b.sync()
    assert b.split_error() == 0
    b.run([400])
    b.sync()
    rec = orc.run_philox_moves(1200, seed=8, kinds=kinds, replica=0)
    info = b.info(0)
    log("split water continued", info, "oracle accepted", int(rec["accepted"].sum()))
    assert info["moves"] == 1200 and info["error"] == 0
    assert info["accepted"] == int(rec["accepted"].sum())
    np.testing.assert_allclose(b.positions(0), orc.positions(), rtol=0, atol=1e-9)
    tot, _ = b.total_energy(0)
    assert abs(info["energy"] - tot) <= 1e-9 * abs(tot)
    dq, _ = b.check_q(0)
    assert dq < 1e-9


def drain(srv, quantum, max_rounds=100000):
    for _ in range(max_rounds):
        r, q = srv.step(quantum)
        if r == 0 and q == 0:
            return
    raise AssertionError("server did not drain")


@pytest.mark.parametrize("fp32", [False, True])
def test_split_serving_jobs_alone_equal_mixed(fp32):
    """Chain jobs above 150 particles get ceil(n / 150) workgroups (at most 4); each
    result is bitwise the same alone and in a mix with one-workgroup chain jobs and
    translate-only jobs, and the FP64 split result ends at the oracle's positions."""
    chains = []
    for r, (nb, nt) in enumerate([(30, 300), (60, 500), (20, 120)]):
        cfg, xyz, q, ids, g, kinds = systems.polyelectrolyte(nb, nt, seed=40 + r, kmax=5)
        chains.append(("chain%d" % r, (cfg, xyz, q, ids, g), kinds, 1500 + 500 * r, 60 + r))
    alone = {}
    for name, sysd, kinds, moves, seed in chains:
        s = F.Server(fp32_pairs=fp32, split_atoms=150, max_parts=4)
        print("split serving alone", name, "fp32", fp32, flush=True)
        j = s.submit(sysd, moves, seed, kinds=kinds)
        drain(s, quantum=10**6)
        info = s.job(j)
        assert info["state"] == F.Server.DONE and info["moves"] == moves, info
        alone[name] = (info, s.positions(j))
        s.close()
        if not fp32:
            orc, _ = oracle_of(*sysd)
            rec = orc.run_philox_moves(moves, seed=seed, kinds=kinds, replica=0)
            log("split serving job", name, info["accepted"], "oracle", int(rec["accepted"].sum()))
            assert info["accepted"] == int(rec["accepted"].sum())
            np.testing.assert_allclose(alone[name][1], orc.positions(), rtol=0, atol=1e-9)
    s = F.Server(fp32_pairs=fp32, split_atoms=150, max_parts=4)
    ids = {}
    for rep in range(2):
        for name, sysd, kinds, moves, seed in (chains[::-1] if rep else chains):
            ids[(name, rep)] = s.submit(sysd, moves, seed, kinds=kinds)
    s.submit(systems.salt21(n=300, seed=3, kmax=6), 2000, 13)
    print("split serving mixed", len(ids), "chain jobs", flush=True)
    drain(s, quantum=200)
    for (name, rep), j in ids.items():
        info = s.job(j)
        a_info, a_xyz = alone[name]
        assert info["state"] == F.Server.DONE, info
        assert info["energy"] == a_info["energy"] and info["accepted"] == a_info["accepted"], (name, rep)
        assert np.array_equal(s.positions(j), a_xyz), (name, rep)
    s.close()


def test_split_jobs_beyond_the_coresident_grid_run_in_several_launches():
    """Six chain jobs of 300 particles at 5 particles per workgroup take 60 workgroups
    each, 360 together: more than fit on the device at once (one 512-thread workgroup
    per CU), so the server runs them in more than one cooperative launch per round.
    Each result is bitwise the same as the job alone (one launch of 60)."""
    jobs = []
    for r in range(6):
        cfg, xyz, q, ids, g, kinds = systems.polyelectrolyte(30, 300, seed=70 + r, kmax=5)
        jobs.append(((cfg, xyz, q, ids, g), kinds, 600, 90 + r))
    alone = []
    for sysd, kinds, moves, seed in jobs:
        s = F.Server(fp32_pairs=False, split_atoms=5, max_parts=64)
        j = s.submit(sysd, moves, seed, kinds=kinds)
        print("wide split job alone", seed, flush=True)
        drain(s, quantum=10**6)
        info = s.job(j)
        assert info["state"] == F.Server.DONE, info
        alone.append((info, s.positions(j)))
        s.close()
    s = F.Server(fp32_pairs=False, split_atoms=5, max_parts=64)
    ids = [s.submit(sysd, moves, seed, kinds=kinds) for sysd, kinds, moves, seed in jobs]
    print("wide split jobs mixed", flush=True)
    drain(s, quantum=100)
    for j, (a_info, a_xyz) in zip(ids, alone):
        info = s.job(j)
        assert info["state"] == F.Server.DONE, info
        assert info["energy"] == a_info["energy"] and info["accepted"] == a_info["accepted"]
        assert np.array_equal(s.positions(j), a_xyz)
    s.close()
