"""GPU parity for the pair potentials, mixing rules and Hamiltonian policies that the
configuration parser accepts beyond the benchmark systems, on the per-move path
(fmc_replay_translate, the reference's proposal scheme) and, where the device-resident
kernels take them, on the batched form (FP64 and FP32 pairs, same Philox stream):

- nonbonded_coulombwca (WCA + Ewald), nonbonded_pm (plain Coulomb + hard spheres),
  nonbonded_pmwca (plain Coulomb + WCA): potentials.h:151-160, 203-207, 472-476;
  energy.cpp:1280-1327;
- Lennard-Jones with arithmetic and geometric mixing and a custom pair
  (potentials.cpp:18-60, 672-689);
- the Ewald surface term epss = 1 (energy.cpp:466-482), per-move path only;
- the maxenergy early break of Hamiltonian::energy (energy.cpp:1227-1241), per-move
  path only (the device-resident forms refuse it);
- a NaN pair energy in the old state (two LJ particles at one point), which
  getEnergyChange turns into du = -inf (montecarlo.cpp:193-209).

Bar: per-term energies and du within 1e-10 * max(|x|, E_scale), identical decisions."""
import json

import numpy as np
import pytest

import faunus_amd as F
from oracle import oracle as O
from tests import systems
from tests.gpu_helpers import build_pair, compare_chain, compare_replay, log

pytestmark = pytest.mark.gpu

EWALD = {"type": "ewald", "epsr": 78.7, "cutoff": 14.0, "alpha": 0.22, "ncutoff": 6}

VARIANTS = {
    "coulombwca": [{"nonbonded_coulombwca": {"coulomb": EWALD, "wca": {"mixing": "LB"}}}],
    "pm": [{"nonbonded_pm": {"coulomb": {"epsr": 78.7}}}],
    "pmwca": [{"nonbonded_pmwca": {"coulomb": {"epsr": 78.7}, "wca": {"mixing": "LB"}}}],
    "lj_arithmetic": [{"nonbonded": {"default": [{"lennardjones": {"mixing": "arithmetic"}}, {"coulomb": EWALD}]}}],
    "lj_geometric": [{"nonbonded": {"default": [{"lennardjones": {"mixing": "geometric"}}, {"coulomb": EWALD}]}}],
    "lj_custom": [{"nonbonded": {"default": [
        {"lennardjones": {"mixing": "LB", "custom": {"Na Cl": {"sigma": 3.1, "eps": 1.5}}}}, {"coulomb": EWALD}]}}],
    "surface_epss1": [{"nonbonded": {"default": [{"hardsphere": {}}, {"coulomb": dict(EWALD, epss=1.0)}]}}],
    "maxenergy": [{"nonbonded": {"default": [{"hardsphere": {}}, {"coulomb": EWALD}]}}, {"maxenergy": 0.0}],
}
DEVICE_OK = {"coulombwca", "pm", "pmwca", "lj_arithmetic", "lj_geometric", "lj_custom", "surface_epss1"}


def system(name, n_pairs=120, box=42.0, seed=5):
    return systems.ions(json.loads(json.dumps(VARIANTS[name])), n_pairs=n_pairs, box=box, seed=seed, dp=3.0)


@pytest.mark.parametrize("name", sorted(VARIANTS))
def test_per_move_replay(name):
    sysd = system(name)
    orc, eng, e_cpu, e_gpu = build_pair(sysd, F.Engine)
    assert eng.term_names() == orc.term_names()
    assert abs(e_gpu - e_cpu) <= 1e-10 * max(1.0, abs(e_cpu)), (e_gpu, e_cpu)
    rec = orc.run_translate(2500, seed=3)
    out = eng.replay_translate(rec["group"], rec["index"], rec["xyz"], rec["u"])
    rep = compare_replay(rec, out, orc.num_terms)
    log("per-move", name, rep)
    assert rep["term_fail"] == 0 and rep["du_fail"] == 0, rep
    assert rep["decision_mismatch"] == 0, rep
    np.testing.assert_array_equal(eng.positions(), orc.positions())
    if name == "maxenergy":
        # the break is taken: some states skip the Ewald term
        skipped = (rec["e_new"][:, -1] == 0.0) & (rec["e_new"][:, -2] >= 0.0)
        assert skipped.sum() > 50, int(skipped.sum())


@pytest.mark.parametrize("name", sorted(DEVICE_OK))
@pytest.mark.parametrize("fp32", [False, True])
def test_device_resident(name, fp32):
    sysd = system(name)
    cfg, xyz, q, ids, g = sysd
    orc = O.System(cfg)
    orc.set_particles(xyz, q, ids)
    orc.set_groups(g["begin"], g["size"], g["molid"])
    e0 = orc.init()
    b = F.Batch(fp32_pairs=fp32)
    b.add(sysd, seed=61)
    nm = 2500
    b.run([nm], trace=nm)
    b.sync()
    rec = orc.run_philox_translate(nm, seed=61, replica=0)
    rel = 1e-6 if fp32 else 1e-10
    rep = compare_chain(b.trace(0, nm), rec, max(1.0, abs(e0)), rel, seed=61)
    log("device", name, "fp32" if fp32 else "fp64", rep)
    assert rep["max_rel"] <= rel, rep
    if not fp32:
        assert rep["first_mismatch"] == nm, rep
        np.testing.assert_allclose(b.positions(0), orc.positions(), rtol=0, atol=1e-9)
    assert b.info(0)["error"] == 0


def test_device_resident_refuses_maxenergy_and_pipelined_surface():
    b = F.Batch()
    with pytest.raises(RuntimeError, match="maxenergy"):
        b.add(system("maxenergy"), seed=1)
    b.close()
    b = F.Batch()
    b.add(system("surface_epss1"), seed=1)
    with pytest.raises(RuntimeError, match="surface"):
        b.run_pipe(10, groups=4, depth=2)
    b.close()


def test_surface_term_in_the_split_and_general_kernels():
    """The surface term (epss = 1) with a running dipole on the split kernel (G = 4) and
    on the general kernel (a polyelectrolyte with ion and bead translations and pivots,
    one workgroup and split over four): FP64 decisions and du against the oracle."""
    sysd = system("surface_epss1")
    cfg, xyz, q, ids, g = sysd
    orc = O.System(cfg)
    orc.set_particles(xyz, q, ids)
    orc.set_groups(g["begin"], g["size"], g["molid"])
    e0 = orc.init()
    b = F.Batch()
    b.add(sysd, seed=62)
    nm = 2000
    b.run_split(nm, groups=4, trace=nm)
    b.sync()
    rec = orc.run_philox_translate(nm, seed=62, replica=0)
    rep = compare_chain(b.trace(0, nm), rec, max(1.0, abs(e0)), 1e-10, seed=62)
    log("surface split", rep)
    assert rep["first_mismatch"] == nm and rep["max_rel"] <= 1e-10, rep
    np.testing.assert_allclose(b.positions(0), orc.positions(), rtol=0, atol=1e-9)
    b.close()
    cfg, xyz, q, ids, g, kinds = systems.polyelectrolyte(30, 300, seed=9, kmax=6)
    cfg["energy"][1]["nonbonded"]["default"][1]["coulomb"]["epss"] = 1.0
    for groups in (1, 4):
        orc = O.System(cfg)
        orc.set_particles(xyz, q, ids)
        orc.set_groups(g["begin"], g["size"], g["molid"])
        e0 = orc.init()
        b = F.Batch()
        b.add((cfg, xyz, q, ids, g), seed=63, moves=kinds)
        nm = 2000
        if groups > 1:
            b.run_split(nm, groups=groups, trace=nm)
        else:
            b.run([nm], trace=nm)
        b.sync()
        rec = orc.run_philox_moves(nm, seed=63, kinds=kinds, replica=0)
        rep = compare_chain(b.trace(0, nm), rec, max(1.0, abs(e0)), 1e-10, seed=63)
        log("surface general groups=%d" % groups, rep)
        assert rep["first_mismatch"] == nm and rep["max_rel"] <= 1e-10, rep
        np.testing.assert_allclose(b.positions(0), orc.positions(), rtol=0, atol=1e-9)
        tot, _ = b.total_energy(0)
        assert abs(b.info(0)["energy"] - tot) <= 1e-9 * max(1.0, abs(tot))
        b.close()


def test_nan_old_state_gives_minus_infinite_du():
    """Two LJ particles at one point: the old state's pair energy is NaN (inf - inf), the
    Hamiltonian stops at it, and moving one of them away has du = -inf and is accepted."""
    energy = [{"nonbonded": {"default": [{"lennardjones": {"mixing": "LB"}}, {"coulomb": EWALD}]}}]
    cfg, xyz, q, ids, g = systems.ions(energy, n_pairs=40, box=40.0, seed=8)
    xyz = xyz.copy()
    xyz[1] = xyz[0]
    sysd = (cfg, xyz, q, ids, g)
    orc, eng, e_cpu, e_gpu = build_pair(sysd, F.Engine)
    assert np.isnan(e_cpu) and np.isnan(e_gpu)
    grp = np.zeros(3, np.int32)
    idx = np.array([1, 5, 1], np.int32)
    new = np.array([xyz[0] + [6.0, 0.0, 0.0], xyz[5] + [0.5, 0.0, 0.0], xyz[0] + [6.5, 0.0, 0.0]])
    new -= 40.0 * np.round(new / 40.0)
    u = np.array([0.5, 0.5, 0.5])
    rc = orc.replay_translate(grp, idx, new, u)
    out = eng.replay_translate(grp, idx, new, u)
    log("nan old state", rc["du"], out["du"], rc["accepted"], out["accepted"])
    assert rc["du"][0] == -np.inf and out["du"][0] == -np.inf
    assert rc["accepted"][0] == 1 and out["accepted"][0] == 1
    rep = compare_replay(rc, out, orc.num_terms)
    assert rep["du_fail"] == 0 and rep["decision_mismatch"] == 0, rep
