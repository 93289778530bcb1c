"""nonbonded_splined (SplinedPotential, potentials.h:770-824, potentials.cpp:1513-1600)
on the GPU against the oracle:

- the engine builds the same tables as the oracle (knots, coefficients, rmin^2,
  rmax^2, hard-sphere flag) for the ion and water inputs;
- per-move replay (the reference's proposals) and the device-resident batched form
  (FP64 and FP32 distances, same Philox stream), with two ions placed inside rmin so
  that the exact energy below rmin is taken, and with `hardsphere`;
- per-term energies and du within 1e-10 * max(|x|, E_scale), identical decisions
  (FP32: 1e-6 up to the first differing decision, which must be a near-tie)."""
import numpy as np
import pytest

import faunus_amd as F
from oracle import oracle as O
from tests import systems
from tests.gpu_helpers import build_pair, compare_chain, compare_replay, log, DRIFT_FP32

pytestmark = pytest.mark.gpu

EWALD12 = {"type": "ewald", "epsr": 78.7, "cutoff": 12.0, "alpha": 0.25, "ncutoff": 6}


def ion_system(hardsphere=False, close_pair=True, default=None):
    term = {"default": default or [{"lennardjones": {"mixing": "LB"}}, {"coulomb": dict(EWALD12)}],
            "rmax": 11.99, "u_at_rmax": 0.01}
    if hardsphere:
        term["hardsphere"] = True
    cfg, xyz, q, ids, g = systems.ions([{"nonbonded_splined": term}], n_pairs=120, box=42.0, seed=6, dp=3.0)
    xyz = xyz.copy()
    if close_pair:
        # Na 0 and Cl 1 at 2.5 A, inside rmin (about 2.9 A): the exact energy applies
        d = xyz[1] - xyz[0]
        xyz[1] = xyz[0] + 2.5 * d / np.linalg.norm(d)
        xyz[1] -= 42.0 * np.round(xyz[1] / 42.0)
    return cfg, xyz, q, ids, g


def test_engine_tables_equal_the_oracle_tables():
    for cfg in (ion_system()[0], systems.water(n_mol=64, seed=2, kmax=5, cutoff=9.0)[0],
                systems.nacl_lj(n_pairs=100, box=60.0, seed=1, kmax=6)[0]):
        orc = O.System(cfg)
        eng = F.Engine(cfg)
        nt = len(cfg["atomlist"])
        for a in range(nt):
            for b in range(nt):
                to, te = orc.splined_table(a, b), eng.splined_table(a, b)
                assert len(to["knots"]) == len(te["knots"]), (a, b)
                np.testing.assert_allclose(te["knots"], to["knots"], rtol=1e-14, atol=0)
                np.testing.assert_allclose(te["coeff"], to["coeff"], rtol=1e-9, atol=1e-12)
                assert te["rmin2"] == to["rmin2"] and te["rmax2"] == to["rmax2"]
                assert te["hardsphere"] == to["hardsphere"]
        eng.close()


@pytest.mark.parametrize("hardsphere", [False, True])
def test_per_move_replay(hardsphere):
    sysd = ion_system(hardsphere=hardsphere)
    orc, eng, e_cpu, e_gpu = build_pair(sysd, F.Engine)
    assert abs(e_gpu - e_cpu) <= 1e-10 * max(1.0, abs(e_cpu)) or (np.isinf(e_cpu) and e_cpu == e_gpu), (e_gpu, e_cpu)
    rec = orc.run_translate(3000, seed=7)
    out = eng.replay_translate(rec["group"], rec["index"], rec["xyz"], rec["u"])
    rep = compare_replay(rec, out, orc.num_terms)
    log("splined per-move hardsphere=%s" % hardsphere, rep)
    assert rep["term_fail"] == 0 and rep["du_fail"] == 0, rep
    assert rep["decision_mismatch"] == 0, rep
    np.testing.assert_array_equal(eng.positions(), orc.positions())


@pytest.mark.parametrize("fp32", [False, True])
def test_device_resident(fp32):
    sysd = ion_system()
    cfg, xyz, q, ids, g = sysd
    orc = O.System(cfg)
    orc.set_particles(xyz, q, ids)
    orc.set_groups(g["begin"], g["size"], g["molid"])
    e0 = orc.init()
    b = F.Batch(fp32_pairs=fp32)
    b.add(sysd, seed=71)
    nm = 3000
    b.run([nm], trace=nm)
    b.sync()
    rec = orc.run_philox_translate(nm, seed=71, replica=0)
    rel = 1e-6 if fp32 else 1e-10
    rep = compare_chain(b.trace(0, nm), rec, max(1.0, abs(e0)), rel, seed=71)
    log("splined device fp32=%s" % fp32, rep)
    assert rep["max_rel"] <= rel, rep
    if not fp32:
        assert rep["first_mismatch"] == nm, rep
        np.testing.assert_allclose(b.positions(0), orc.positions(), rtol=0, atol=1e-9)
    tot, _ = b.total_energy(0)
    assert abs(b.info(0)["energy"] - tot) <= (DRIFT_FP32 if fp32 else 1e-9) * max(1.0, abs(tot))


def test_split_kernel():
    sysd = ion_system()
    cfg, xyz, q, ids, g = sysd
    orc = O.System(cfg)
    orc.set_particles(xyz, q, ids)
    orc.set_groups(g["begin"], g["size"], g["molid"])
    e0 = orc.init()
    b = F.Batch()
    b.add(sysd, seed=72)
    nm = 2000
    b.run_split(nm, groups=4, stride=1, trace=nm)
    b.sync()
    assert b.split_error() == 0
    rec = orc.run_philox_translate(nm, seed=72, replica=0)
    rep = compare_chain(b.trace(0, nm), rec, max(1.0, abs(e0)), 1e-10, seed=72)
    log("splined split", rep)
    assert rep["first_mismatch"] == nm and rep["max_rel"] <= 1e-10, rep
    np.testing.assert_allclose(b.positions(0), orc.positions(), rtol=0, atol=1e-9)
