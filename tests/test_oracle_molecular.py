"""Oracle pins for molecular systems: the harmonic bond and exclusion known answers
the reference's own doctests hold, and invariants of the translate+rotate and pivot
proposals (rigid-body distances, bond lengths, mass centres) together with a
per-move check of the Change-based energy against full recomputation, which fails if
any affected pair or bond is dropped or double counted."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests import systems

KT_PER_KJMOL = 1e3 / (1.380658e-23 * 298.15 * 6.022137e23)


def _oracle(cfg, xyz, q, ids, g):
    s = O.System(cfg)
    s.set_particles(xyz, q, ids)
    s.set_groups(g["begin"], g["size"], g["molid"])
    e0 = s.init()
    return s, e0


def test_harmonic_bond_known_answer():
    # bonds.cpp:686-690: {"harmonic": {"index":[1,2], "k":10, "req":2}} on p_60deg_4a -> 10 kJ/mol / 2 * 4
    cfg = {"temperature": 298.15, "geometry": {"type": "cuboid", "length": 100.0},
           "atomlist": [{"A": {"q": 0.0, "sigma": 0.0}}],
           "moleculelist": [{"m": {"atoms": ["A", "A", "A"],
                                   "bondlist": [{"harmonic": {"index": [1, 2], "k": 10.0, "req": 2.0}}]}}],
           "energy": [{"bonded": {}}]}
    xyz = np.array([[1.0, 1.0 + math.sqrt(3), 4.0], [1.0, 1.0, 1.0], [1.0, 5.0, 1.0]])
    s, e0 = _oracle(cfg, xyz, np.zeros(3), np.zeros(3, np.int32), dict(begin=[0], size=[3], molid=[0]))
    assert e0 == pytest.approx(10.0 / 2 * 4 * KT_PER_KJMOL, rel=1e-12)


def _mol_cfg(n, bonds, **extra):
    m = {"atoms": ["A"] * n, "bondlist": [{"harmonic": {"index": [i, j], "k": 1.0, "req": 1.0}} for i, j in bonds]}
    m.update(extra)
    return {"temperature": 298.15, "geometry": {"type": "cuboid", "length": 100.0},
            "atomlist": [{"A": {"q": 0.0, "sigma": 0.0}}], "moleculelist": [{"m": m}], "energy": [{"bonded": {}}]}


def _excluded_pairs(cfg, n):
    s = O.System(cfg)
    return {(i, j) for i in range(n) for j in range(i + 1, n) if s.excluded(0, i, j)}


def test_excluded_neighbours_linear_chain():
    # molecule.cpp:301-327: decamer, distance 3 -> (n-1)+(n-2)+(n-3) pairs, all (i, i+d)
    n = 10
    pairs = _excluded_pairs(_mol_cfg(n, [(i, i + 1) for i in range(n - 1)], excluded_neighbours=3), n)
    assert len(pairs) == 9 + 8 + 7
    assert pairs == {(i, i + d) for d in (1, 2, 3) for i in range(n - d)}


def test_excluded_neighbours_cycle():
    # molecule.cpp:329-350: cyclic hexamer, distance 3 -> mer + mer + mer/2 pairs
    n = 6
    pairs = _excluded_pairs(_mol_cfg(n, [(i, (i + 1) % n) for i in range(n)], excluded_neighbours=3), n)
    assert len(pairs) == 6 + 6 + 3


def test_exclusions_peptide_ring():
    # molecule.cpp:556-575: 8-mer chain + bond (0,7), excluded_neighbours 2, exclusionlist [[3,6]]
    n = 8
    cfg = _mol_cfg(n, [(i, i + 1) for i in range(n - 1)] + [(0, 7)], excluded_neighbours=2, exclusionlist=[[3, 6]])
    s = O.System(cfg)
    assert s.excluded(0, 0, 1) and s.excluded(0, 0, 2) and not s.excluded(0, 0, 3)
    assert s.excluded(0, 7, 5) and not s.excluded(0, 7, 4)
    assert s.excluded(0, 3, 6) and s.excluded(0, 6, 3) and not s.excluded(0, 3, 7)
    assert s.excluded(0, 7, 1)


def _per_move_consistency(sysd, kinds, nmoves, seed):
    """Run one move at a time; every du must equal the change of the full energy."""
    cfg, xyz, q, ids, g = sysd
    s, e0 = _oracle(cfg, xyz, q, ids, g)
    e_prev = e0
    worst = 0.0
    n_acc = 0
    for t in range(nmoves):
        rec = s.run_philox_moves(1, seed=seed, kinds=kinds, move0=t)
        e_now, _ = s.total_energy()
        if rec["accepted"][0]:
            n_acc += 1
            ref = e_now - e_prev
            worst = max(worst, abs(rec["du"][0] - ref) / max(1.0, abs(e_now)))
        else:
            assert e_now == pytest.approx(e_prev, rel=1e-12, abs=1e-12)
        e_prev = e_now
    return s, worst, n_acc


def test_pivot_and_bead_moves_match_full_recompute():
    cfg, xyz, q, ids, g, kinds = systems.polyelectrolyte(30, 150, seed=11, kmax=5)
    s, worst, n_acc = _per_move_consistency((cfg, xyz, q, ids, g), kinds, 300, seed=3)
    assert n_acc > 30
    assert worst < 1e-10, worst


def test_transrot_water_matches_full_recompute():
    cfg, xyz, q, ids, g, kinds = systems.water(n_mol=64, seed=2, kmax=5, rmax=None)
    s, worst, n_acc = _per_move_consistency((cfg, xyz, q, ids, g), kinds, 200, seed=4)
    assert n_acc > 20
    assert worst < 1e-10, worst


def _min_image(d, box):
    return d - box * np.round(d / box)


def test_transrot_is_rigid_and_tracks_mass_center():
    cfg, xyz, q, ids, g, kinds = systems.water(n_mol=27, seed=3, kmax=4, rmax=None)
    box = cfg["geometry"]["length"]
    s, e0 = _oracle(cfg, xyz, q, ids, g)
    rec = s.run_philox_moves(400, seed=9, kinds=kinds)
    assert rec["accepted"].sum() > 50
    final = s.positions()
    mw = np.array([15.999, 1.008, 1.008])
    for w in range(27):
        a, b = xyz[3 * w:3 * w + 3], final[3 * w:3 * w + 3]
        da = _min_image(a[:, None] - a[None], box)
        db = _min_image(b[:, None] - b[None], box)
        np.testing.assert_allclose(np.linalg.norm(da, axis=-1), np.linalg.norm(db, axis=-1), atol=1e-10)
        rel = _min_image(b - b[0], box) + b[0]
        cm = (rel * mw[:, None]).sum(0) / mw.sum()
        cm = _min_image(cm, box)
        np.testing.assert_allclose(_min_image(s.mass_center(w) - cm, box), 0.0, atol=1e-9)


def test_pivot_preserves_bond_lengths():
    cfg, xyz, q, ids, g, kinds = systems.polyelectrolyte(40, 200, seed=5, kmax=4)
    box = cfg["geometry"]["length"]
    pivot_only = [k for k in kinds if k["kind"] == 2]
    s, e0 = _oracle(cfg, xyz, q, ids, g)
    rec = s.run_philox_moves(300, seed=2, kinds=pivot_only)
    assert rec["accepted"].sum() > 20
    chain0, chain1 = xyz[:40], s.positions()[:40]
    l0 = np.linalg.norm(_min_image(np.diff(chain0, axis=0), box), axis=1)
    l1 = np.linalg.norm(_min_image(np.diff(chain1, axis=0), box), axis=1)
    np.testing.assert_allclose(l0, l1, atol=1e-9)
    # a pivot keeps every bond length, so the bonded term never changes
    _, t0 = s.total_energy()
    assert abs(t0[0] - e0_bonded(cfg, xyz, q, ids, g)) < 1e-9


def e0_bonded(cfg, xyz, q, ids, g):
    s, _ = _oracle(cfg, xyz, q, ids, g)
    _, t = s.total_energy()
    return t[0]
