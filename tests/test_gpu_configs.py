"""Parity at the benchmark configurations' own sizes, against the oracle running the
same Philox proposal streams.

- Rigid water at 4096 molecules (L = 49.68 A, Ewald alpha 0.28, Rc 10, kmax 9, K = 1661),
  two replicas of a 280-360 K ladder in one launch, finer cell grid: FP64 decisions
  exact, du within 1e-10 * max(|du|, E_scale), positions within 1e-9 A; FP32 pairs
  within 1e-6 up to and including the first differing decision, which must be a
  near-tie (tests.gpu_helpers.compare_chain).
- Polyelectrolyte serving jobs (chain with harmonic bonds, counterions and salt, HS +
  Ewald alpha = 5/L, kmax 8) at N ~ 1.5k and ~ 10k through the server with the split
  policy on: FP64 accepted count and final positions equal the oracle's.
- N = 20 000 Na/Cl (K = 17 328) in the FP32-pair form, and the headline split kernel
  at its own setting (16 workgroups).

Each case logs its maximum relative du error against the FP64 oracle."""
import numpy as np
import pytest

import faunus_amd as F
from oracle import oracle as O
from tests import systems
from tests.gpu_helpers import compare_chain, log, DRIFT_FP32

pytestmark = pytest.mark.gpu

TREF = 298.15


def oracle_of(cfg, xyz, q, ids, g):
    s = O.System(cfg)
    s.set_particles(xyz, q, ids)
    s.set_groups(g["begin"], g["size"], g["molid"])
    return s, s.init()


@pytest.mark.parametrize("fp32", [False, True])
def test_water_4096_ladder_pair(fp32):
    """Two rungs (280 K and 360 K) of the water ladder in one launch, 400 moves each."""
    cfg, xyz, q, ids, g, kinds = systems.water(n_mol=4096, seed=11)
    assert len(q) == 12288
    temps = (280.0, 360.0)
    nm = 400
    b = F.Batch(fp32_pairs=fp32, cell_divisions=2)
    for r, T in enumerate(temps):
        b.add((cfg, xyz, q, ids, g), beta=TREF / T, seed=600 + r, ladder=0, moves=kinds)
    print("water 4096 ladder pair fp32=%s" % fp32, flush=True)
    b.run([nm, nm], trace=nm)
    b.sync()
    for r, T in enumerate(temps):
        info = b.info(r)
        assert info["error"] == 0 and info["moves"] == nm, info
        orc, e0 = oracle_of(cfg, xyz, q, ids, g)
        rec = orc.run_philox_moves(nm, seed=600 + r, kinds=kinds, replica=r, beta=TREF / T)
        rel = 1e-6 if fp32 else 1e-10
        rep = compare_chain(b.trace(r, nm), rec, max(1.0, abs(e0)), rel, seed=600 + r, replica=r)
        log("water 4096 T=%g fp32=%s" % (T, fp32), rep, "K", b.check_q(r))
        assert rep["max_rel"] <= rel, rep
        if not fp32:
            assert rep["first_mismatch"] == nm, rep
            np.testing.assert_allclose(b.positions(r), orc.positions(), rtol=0, atol=1e-9)
        tot, _ = b.total_energy(r)
        assert abs(info["energy"] - tot) <= (DRIFT_FP32 if fp32 else 1e-9) * abs(tot)
        dq, _ = b.check_q(r)
        assert dq < 1e-9


@pytest.mark.parametrize("n_beads,n_total,moves", [(100, 1500, 6000), (400, 10000, 3000)])
def test_serving_polyelectrolyte_jobs_fp64(n_beads, n_total, moves):
    """Serving jobs of the Poisson stream's kind (translations of ions and beads, pivots)
    through the server with jobs above 2048 particles split over workgroups."""
    cfg, xyz, q, ids, g, kinds = systems.polyelectrolyte(n_beads, n_total, seed=21)
    sysd = (cfg, xyz, q, ids, g)
    assert abs(cfg["energy"][1]["nonbonded"]["default"][1]["coulomb"]["alpha"] * cfg["geometry"]["length"] - 5.0) < 1e-12
    srv = F.Server(fp32_pairs=False, hbm_budget_bytes=8 << 30, max_slots=8, split_atoms=2048, max_parts=16)
    print("serving polyelectrolyte job N=%d moves=%d" % (len(q), moves), flush=True)
    j = srv.submit(sysd, moves, seed=77, kinds=kinds)
    while True:
        r, qd = srv.step(4000)
        if r == 0 and qd == 0:
            break
    info = srv.job(j)
    assert info["state"] == F.Server.DONE, info
    orc, e0 = oracle_of(*sysd)
    rec = orc.run_philox_moves(moves, seed=77, kinds=kinds, replica=0)
    log("serving job N=%d" % len(q), "gpu accepted", info["accepted"], "oracle", int(rec["accepted"].sum()))
    assert info["moves"] == moves
    assert info["accepted"] == int(rec["accepted"].sum())
    np.testing.assert_allclose(srv.positions(j), orc.positions(), rtol=0, atol=1e-9)
    srv.close()


def test_nacl_20k_fp32_pairs():
    """The FP32-pair / FP64-Q(k) form at N = 20 000, K = 17 328."""
    sysd = systems.nacl_lj(n_pairs=10000, box=255.0, seed=1)
    orc, e0 = oracle_of(*sysd)
    nm = 600
    b = F.Batch(fp32_pairs=True)
    b.add(sysd, seed=9)
    print("nacl 20k fp32", flush=True)
    b.run([nm], trace=nm)
    b.sync()
    rec = orc.run_philox_translate(nm, seed=9, replica=0)
    rep = compare_chain(b.trace(0, nm), rec, max(1.0, abs(e0)), 1e-6, seed=9)
    log("nacl 20k fp32", rep)
    assert rep["max_rel"] <= 1e-6, rep
    tot, _ = b.total_energy(0)
    assert abs(b.info(0)["energy"] - tot) <= DRIFT_FP32 * abs(tot)


def test_headline_split_16_workgroups_full_size():
    """The headline setting: N = 20 000, 16 cooperating workgroups of the split kernel."""
    sysd = systems.nacl_lj(n_pairs=10000, box=255.0, seed=1)
    orc, e0 = oracle_of(*sysd)
    nm = 1000
    b = F.Batch()
    b.add(sysd, seed=2024)
    print("headline split G=16 N=20000", flush=True)
    b.run_split(nm, groups=16, stride=1, trace=nm)
    b.sync()
    assert b.split_error() == 0
    rec = orc.run_philox_translate(nm, seed=2024, replica=0)
    rep = compare_chain(b.trace(0, nm), rec, max(1.0, abs(e0)), 1e-10, seed=2024)
    log("headline split G=16", rep)
    assert rep["first_mismatch"] == nm and rep["max_rel"] <= 1e-10, rep
    np.testing.assert_allclose(b.positions(0), orc.positions(), rtol=0, atol=1e-9)
    tot, _ = b.total_energy(0)
    assert abs(b.info(0)["energy"] - tot) <= 1e-9 * abs(tot)
    dq, _ = b.check_q(0)
    assert dq < 1e-9
