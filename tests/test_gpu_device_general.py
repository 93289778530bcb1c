"""Device-resident Metropolis with general move sets (atom translations inside
molecules, rigid translate+rotate, pivot) against the oracle running the same Philox
proposal stream: FP64 du within 1e-10 relative to max(|du|, E_scale) with identical
decisions, final positions, running energy against a full recompute and Q(k) against
a full rebuild; FP32 pair terms within 1e-6 up to the first near-tie."""
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


def run(sysd, kinds, nm, seed, fp32=False, cell_divisions=1):
    cfg, xyz, q, ids, g = sysd
    orc, e0 = oracle_of(cfg, xyz, q, ids, g)
    b = F.Batch(fp32_pairs=fp32, cell_divisions=cell_divisions)
    b.add(sysd, seed=seed, moves=kinds)
    b.run([nm], trace=nm)
    b.sync()
    tr = b.trace(0, nm)
    rec = orc.run_philox_moves(nm, seed=seed, kinds=kinds, replica=0)
    return b, orc, e0, compare(tr, rec, max(1.0, abs(e0)), seed=seed)


@pytest.mark.parametrize("fp32,excl", [(False, 0), (True, 0), (False, 1)])
def test_polyelectrolyte_general_moves(fp32, excl):
    cfg, xyz, q, ids, g, kinds = systems.polyelectrolyte(40, 400, seed=7, kmax=6, excluded_neighbours=excl)
    b, orc, e0, rep = run((cfg, xyz, q, ids, g), kinds, 3000, seed=33, fp32=fp32)
    log("device general polyelectrolyte fp32=%s" % fp32, rep)
    info = b.info(0)
    assert info["moves"] == 3000 and info["error"] == 0
    tot, _ = b.total_energy(0)
    if not fp32:
        assert rep["first_mismatch"] == rep["n"] and rep["max_rel"] <= 1e-10, rep
        np.testing.assert_allclose(b.positions(0), orc.positions(), rtol=0, atol=1e-9)
        assert abs(info["energy"] - tot) <= 1e-9 * max(1.0, abs(tot))
    else:
        assert rep["max_rel"] <= 1e-6 and rep["first_mismatch"] > 500, rep
        assert abs(info["energy"] - tot) <= DRIFT_FP32 * max(1.0, abs(tot))
    dq, _ = b.check_q(0)
    assert dq < 1e-9


@pytest.mark.parametrize("n_mol,cells,rc", [(216, False, 9.0), (1000, True, 10.0), (1000, True, 8.0)])
def test_water_general_moves(n_mol, cells, rc):
    """Cell list at Rc = 10 A (cells of L/3 = 10.35 A: one stencil per moved atom) and at
    Rc = 8 A (Rc + the O-H distance fits in a cell: one stencil per state for the whole
    molecule)."""
    cfg, xyz, q, ids, g, kinds = systems.water(n_mol=n_mol, seed=3, kmax=6, rmax="auto" if cells else None,
                                                cutoff=rc)
    if cells:
        box = cfg["geometry"]["length"]
        cell = box / np.floor(box / rc)
        assert (rc + 1.0 <= cell) == (rc < 9.0)  # which stencil form the case exercises
    b, orc, e0, rep = run((cfg, xyz, q, ids, g), kinds, 1500, seed=8)
    log("device general water", n_mol, rc, rep)
    assert rep["first_mismatch"] == rep["n"] and rep["max_rel"] <= 1e-10, rep
    np.testing.assert_allclose(b.positions(0), orc.positions(), rtol=0, atol=1e-9)
    tot, _ = b.total_energy(0)
    assert abs(b.info(0)["energy"] - tot) <= 1e-9 * abs(tot)


@pytest.mark.parametrize("rc,molecule_stencil", [(9.0, True), (10.0, False)])
def test_water_on_the_finer_cell_grid(rc, molecule_stencil):
    """Cells of at least Rc / 2 and 125-cell stencils (L = 31.05 A, 6 cells of 5.18 A):
    at Rc = 9 A, Rc plus the O-H distance fits two cells and one stencil per state serves
    all three atoms; at Rc = 10 A each atom walks its own stencil, one (atom, state,
    cell) combination per lane."""
    cfg, xyz, q, ids, g, kinds = systems.water(n_mol=1000, seed=3, kmax=6, cutoff=rc)
    box = cfg["geometry"]["length"]
    cell = box / np.floor(2 * box / rc)
    assert (rc + 1.0 <= 2 * cell) == molecule_stencil
    b, orc, e0, rep = run((cfg, xyz, q, ids, g), kinds, 1500, seed=8, cell_divisions=2)
    log("device general water, finer cells", rc, rep)
    assert rep["first_mismatch"] == rep["n"] and rep["max_rel"] <= 1e-10, rep
    np.testing.assert_allclose(b.positions(0), orc.positions(), rtol=0, atol=1e-9)
    tot, _ = b.total_energy(0)
    assert abs(b.info(0)["energy"] - tot) <= 1e-9 * abs(tot)


def test_finer_cell_grid_is_refused_by_the_translate_kernel():
    b = F.Batch(cell_divisions=2)
    with pytest.raises(RuntimeError, match="general"):
        b.add(systems.nacl_lj(n_pairs=800, box=255.0 * (1600 / 20000) ** (1 / 3), seed=2, kmax=9), seed=1)


def test_batch_of_uneven_polyelectrolytes_and_ladder():
    """Replicas of different chain length and size in one launch, two of them a
    temperature ladder; each follows its own oracle stream."""
    specs = [(30, 300, 5), (50, 500, 6), (40, 400, 7)]
    b = F.Batch()
    sysds = []
    for r, (nb, nt, seed) in enumerate(specs):
        cfg, xyz, q, ids, g, kinds = systems.polyelectrolyte(nb, nt, seed=seed, kmax=5)
        sysds.append(((cfg, xyz, q, ids, g), kinds))
        b.add((cfg, xyz, q, ids, g), seed=500 + r, moves=kinds)
    nm = 1200
    b.run([nm] * len(specs), trace=nm)
    b.sync()
    for r, (sysd, kinds) in enumerate(sysds):
        orc, e0 = oracle_of(*sysd)
        rec = orc.run_philox_moves(nm, seed=500 + r, kinds=kinds, replica=r)
        rep = compare(b.trace(r, nm), rec, max(1.0, abs(e0)), seed=500 + r, replica=r)
        log("general batch replica", r, rep)
        assert rep["first_mismatch"] == nm and rep["max_rel"] <= 1e-10, rep
