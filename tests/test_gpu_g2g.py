"""The group-to-group cutoff (cutoff_g2g, GroupCutoff::cut, energy.h:761-768) on the
device-resident general kernel against the oracle running the same Philox stream:

- rigid water written as the reference's water example writes it (nonbonded_coulomblj,
  LJ truncated only by cutoff_g2g on the molecules' mass centres), without a cell list
  (216 molecules, L = 18.6 A) and with one (4096 molecules, L = 49.7 A: the cells cover
  the cutoff plus twice the largest atom-to-centre distance of the rigid molecules, on
  both cell grids);
- flexible chains with counterions, where atom translations and pivots move a chain's
  mass centre (recomputed on the device from the trial positions, as translateParticle
  and rotate_segment do), on one workgroup and split over four.

FP64: identical decisions, du within 1e-10 * max(|du|, |E|), final positions within
1e-9 A; FP32 pairs: within 1e-6 up to the first differing decision, which must be a
near-tie."""
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


def check(b, orc, e0, rec, nm, seed, fp32, what, telescopes=True):
    """telescopes: the running energy equals a full recompute.  It does not when a
    group is changed in part: the move's dU holds the moved atoms' pairs only, while the
    changed mass centre can cut or uncut the group's other pairs too (GroupPairing with
    GroupCutoff, energy.h:885-1027); the running energy then follows the oracle's own
    running energy (the start energy plus the accepted dU)."""
    rep = compare_chain(b.trace(0, nm), rec, max(1.0, abs(e0)), 1e-6 if fp32 else 1e-10, seed=seed)
    log(what, rep)
    info = b.info(0)
    assert info["error"] == 0 and info["moves"] == nm
    if telescopes:
        tot, _ = b.total_energy(0)
    else:
        tot = e0 + float(np.where(np.asarray(rec["accepted"]) == 1, rec["du"], 0.0).sum())
    if fp32:
        assert rep["max_rel"] <= 1e-6, rep
        assert abs(info["energy"] - tot) <= DRIFT_FP32 * max(1.0, abs(tot))
    else:
        assert rep["first_mismatch"] == nm and rep["max_rel"] <= 1e-10, rep
        np.testing.assert_allclose(b.positions(0), orc.positions(), rtol=0, atol=1e-9)
        assert abs(info["energy"] - tot) <= 1e-9 * max(1.0, abs(tot))
    return rep


@pytest.mark.parametrize("n_mol,divisions,fp32", [(216, 1, False), (216, 1, True), (4096, 1, False), (4096, 2, False),
                                                  (4096, 1, True)])
def test_water_cutoff_g2g(n_mol, divisions, fp32):
    cfg, xyz, q, ids, g, kinds = systems.water_g2g(n_mol=n_mol, seed=4)
    nm, seed = (1200 if n_mol < 1000 else 600), 21
    orc, e0 = oracle_of(cfg, xyz, q, ids, g)
    b = F.Batch(fp32_pairs=fp32, cell_divisions=divisions)
    b.add((cfg, xyz, q, ids, g), seed=seed, moves=kinds)
    print("water g2g n_mol=%d divisions=%d fp32=%s" % (n_mol, divisions, fp32), flush=True)
    # the tuner times the cell grids only where the replica has a cell list; it restores
    # the chain afterwards
    chosen, cands = b.autotune(20)
    assert bool(cands) == (n_mol == 4096), cands
    b.run([nm], trace=nm)
    b.sync()
    rec = orc.run_philox_moves(nm, seed=seed, kinds=kinds, replica=0)
    check(b, orc, e0, rec, nm, seed, fp32, "water g2g n_mol=%d cells=%s" % (n_mol, cands))
    b.close()


@pytest.mark.parametrize("fp32,groups", [(False, 1), (True, 1), (False, 4)])
def test_chains_cutoff_g2g(fp32, groups):
    cfg, xyz, q, ids, g, kinds = systems.chains_g2g(seed=3)
    nm, seed = 3000, 22
    orc, e0 = oracle_of(cfg, xyz, q, ids, g)
    b = F.Batch(fp32_pairs=fp32)
    b.add((cfg, xyz, q, ids, g), seed=seed, moves=kinds)
    print("chains g2g fp32=%s groups=%d" % (fp32, groups), flush=True)
    if groups > 1:
        b.run_split(nm, groups=groups, trace=nm)
    else:
        b.run([nm], trace=nm)
    b.sync()
    assert b.split_error() == 0
    rec = orc.run_philox_moves(nm, seed=seed, kinds=kinds, replica=0)
    check(b, orc, e0, rec, nm, seed, fp32, "chains g2g fp32=%s groups=%d" % (fp32, groups), telescopes=False)
    # the chains' mass centres followed the atom moves and pivots
    st = b.get_state(0)
    for c in range(len(g["begin"]) - 1):
        np.testing.assert_allclose(st["mass_center"][c], orc.mass_center(c), rtol=0, atol=1e-9)
    b.close()


def test_cutoff_g2g_pairs_of_molecule_types():
    """{default, "A B"} entries: the oracle and the engine's total energies agree, and a
    pair entry changes the energy (the chains' own cutoff against the default)."""
    cfg, xyz, q, ids, g, kinds = systems.chains_g2g(seed=3)
    e = {}
    for name, c in [("default", {"default": 25.0}), ("pair", {"default": 25.0, "chain chain": 12.0}),
                    ("number", 12.0)]:
        cfg["energy"][1]["nonbonded"]["cutoff_g2g"] = c
        orc, e0 = oracle_of(cfg, xyz, q, ids, g)
        eng = F.Engine(cfg)
        eng.set_particles(xyz, q, ids)
        eng.set_groups(g["begin"], g["size"], g["molid"])
        e1 = eng.init()
        eng.close()
        assert abs(e1 - e0) <= 1e-10 * max(1.0, abs(e0)), (name, e1, e0)
        e[name] = e0
    assert e["pair"] == e["number"] and e["pair"] != e["default"], e
