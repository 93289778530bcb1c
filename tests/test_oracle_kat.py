"""Pin the FP64 oracle against the reference's own known answers, golden example
outputs and closed-form physics.  CPU only."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests import systems

EWALD_KAT = {"epsr": 1.0, "alpha": 0.894427190999916, "epss": 1.0, "ncutoff": 11.0,
             "spherical_sum": True, "cutoff": 5.0}


def approx(a, b, rel=1e-9, abs_=0.0):
    return abs(a - b) <= max(rel * abs(b), abs_)


# --- Ewald policy doctests, energy.cpp:74-99 and 249-317 -----------------------
def test_ewald_kvector_counts_and_parts_pbc():
    lB = O.bjerrum_length(298.15, 1.0)
    nk, self_e, surf, rec = O.ewald_parts(EWALD_KAT, 298.15, 10.0, [[0, 0, 0], [1, 0, 0]], [1.0, -1.0])
    assert nk == 2975
    assert approx(self_e, -1.0092530088080642 * lB)
    assert approx(surf, 0.0020943951023931952 * lB)
    assert approx(rec, 0.21303063979675319 * lB)


def test_ewald_kvector_counts_and_parts_ipbc():
    lB = O.bjerrum_length(298.15, 1.0)
    cj = dict(EWALD_KAT, ewaldscheme="IPBC")
    nk, self_e, surf, rec = O.ewald_parts(cj, 298.15, 10.0, [[0, 0, 0], [1, 0, 0]], [1.0, -1.0])
    assert nk == 846
    assert approx(rec, 0.0865107467 * lB, rel=1e-9)


@pytest.mark.parametrize("kmax,count", [(7, 783), (8, 1152), (9, 1661), (20, 17328)])
def test_kvector_counts_of_benchmark_configs(kmax, count):
    k = O.ewald_kvectors({"epsr": 1, "alpha": 0.2, "ncutoff": kmax, "cutoff": 10}, 298.15, 100.0)
    assert len(k) == count
    n = np.rint(k[:, :3] * 100.0 / (2 * math.pi)).astype(int)
    assert (n[:, 0] >= 0).all() and ((n ** 2).sum(1) <= kmax * kmax).all()


def _nacl_two_ions(splined=False):
    cfg = {"temperature": 298.15, "geometry": {"type": "cuboid", "length": 10},
           "atomlist": [{"Na": {"sigma": 3.8, "eps": 0.1, "q": 1.0}}, {"Cl": {"sigma": 4.0, "eps": 0.05, "q": -1.0}}],
           "moleculelist": [{"salt": {"atomic": True, "atoms": ["Na", "Cl"]}}],
           "energy": [{"nonbonded": {"default": [{"coulomb": dict(EWALD_KAT, type="ewald", splined=splined)}]}}]}
    s = O.System(cfg)
    s.set_particles([[0, 0, 0], [1, 0, 0]], [1.0, -1.0], [0, 1])
    s.set_groups([0], [2], [0])
    s.init()
    return s


def test_ewald_energy_and_partial_update_doctest():
    # energy.cpp:665-762: E = (surface + reciprocal); moving ion 0 to (0.1,0.1,0.1)
    # gives E = 103.7300260099 kT and dE = -16.8380445846 kT, via partial update with
    # relative_atom_indices={0} (and with all=true).
    lB = O.bjerrum_length(298.15, 1.0)
    for change in ({"groups": [{"group_index": 0, "relative_atom_indices": [0], "internal": True}]},
                   {"groups": [{"group_index": 0, "all": True, "internal": True}]}):
        s = _nacl_two_ions()
        names = s.term_names()
        assert names == ["selfenergy", "nonbonded", "ewald"]
        _, terms = s.total_energy()
        assert approx(terms[2], (0.0020943951023931952 + 0.21303063979675319) * lB)
        idx = [0] if change["groups"][0].get("relative_atom_indices") else [0, 1]
        xyz = [[0.1, 0.1, 0.1]] if len(idx) == 1 else [[0.1, 0.1, 0.1], [1, 0, 0]]
        en, eo = s.energy_change(change, idx, xyz)
        assert approx(en[2], 103.7300260099, rel=1e-9)
        assert approx(en[2] - eo[2], -16.8380445846, rel=1e-9)


# --- Andrea spline doctest, tabulate.h:313-365 --------------------------------
def test_andrea_spline_doctest():
    knots, coeff, ys = O.spline_test(0, utol=2e-6, ftol=1e-4, xmin=0, xmax=10, xs=[1e-9, 5.0, 10.0])
    assert len(knots) == 19 and len(coeff) == 108
    assert knots[0] == pytest.approx(0.0) and knots[-1] == pytest.approx(10.0)
    assert knots[1] == pytest.approx(0.212991, rel=1e-5)
    assert knots[2] == pytest.approx(0.782554, rel=1e-5)
    assert coeff[0] == pytest.approx(2.0) and coeff[1] == pytest.approx(0.0, abs=1e-12)
    assert coeff[2] == pytest.approx(0.5, rel=1e-5)
    assert coeff[-1] == pytest.approx(-0.0441931, rel=1e-5)
    f = lambda x: 0.5 * x * math.sin(x) + 2
    for x, y in zip([1e-9, 5.0, 10.0], ys):
        assert y == pytest.approx(f(x), rel=1e-5)


def test_ewald_short_range_spline_within_utol():
    # S(q)=erfc(alpha*Rc*q) splined on q in [0,1] with utol = 0.005/lB (potentials.cpp:1632)
    cj = {"type": "ewald", "epsr": 78.7, "cutoff": 30.0, "alpha": 0.2, "ncutoff": 7}
    lB = O.bjerrum_length(298.15, 78.7)
    q = np.linspace(1e-4, 1.0 - 1e-9, 4001)
    knots, coeff, ys = O.spline_test(1, coulomb=cj, xs=q)
    exact = np.array([math.erfc(6.0 * x) for x in q])
    assert np.max(np.abs(ys - exact)) <= 0.005 / lB * 1.0001
    assert 3 <= len(knots) < 200


# --- Philox4x32-10 known answers (Random123 kat_vectors) -----------------------
def test_philox_known_answers():
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert O.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


# --- pair-potential doctests, potentials.cpp -------------------------------------
def test_bjerrum_length():
    assert O.bjerrum_length(298.15, 80.0) == pytest.approx(7.0056973292, rel=1e-10)  # potentials.cpp:660


def _lj_system(mixing, custom=None):
    lj = {"mixing": mixing}
    if custom:
        lj["custom"] = custom
    cfg = {"temperature": 298.15, "geometry": {"type": "cuboid", "length": 100},
           "atomlist": [{"A": {"sigma": 2.0, "eps": 0.9}}, {"B": {"sigma": 8.0, "eps": 0.1}},
                        {"C": {"sigma": 5.0, "eps": 1.1}}],
           "energy": [{"nonbonded": {"default": [{"lennardjones": lj}]}}]}
    return O.System(cfg)


def test_lennard_jones_mixing_doctest():
    # potentials.cpp:711-780 (d = 0.9 nm)
    kJ = 1e3 / (298.15 * 1.380658e-23 * 6.022137e23)
    d = 9.0
    lj = lambda s, e: 4 * e * kJ * ((s / d) ** 12 - (s / d) ** 6)
    s = _lj_system("LB")
    assert s.pair_energy(0, 0, 0, 0, d * d) == pytest.approx(lj(2.0, 0.9), rel=1e-12)
    assert s.pair_energy(0, 0, 1, 0, d * d) == pytest.approx(lj(5.0, 0.3), rel=1e-12)
    s = _lj_system("geometric")
    assert s.pair_energy(0, 0, 1, 0, d * d) == pytest.approx(lj(4.0, 0.3), rel=1e-12)
    s = _lj_system("LB", custom=[{"A B": {"eps": 0.5, "sigma": 8}}])
    assert s.pair_energy(0, 0, 1, 0, d * d) == pytest.approx(lj(8.0, 0.5), rel=1e-12)
    assert s.pair_energy(0, 0, 0, 0, d * d) == pytest.approx(lj(2.0, 0.9), rel=1e-12)
    with pytest.raises(RuntimeError):
        _lj_system("unknown")


def test_hardsphere_doctest():
    # potentials.cpp:979-1013
    def hs(custom=None):
        h = {"mixing": "arithmetic"}
        if custom:
            h["custom"] = custom
        cfg = {"temperature": 298.15, "geometry": {"type": "cuboid", "length": 100},
               "atomlist": [{"A": {"sigma": 2}}, {"B": {"sigma": 8}}],
               "energy": [{"nonbonded": {"default": [{"hardsphere": h}]}}]}
        return O.System(cfg)
    s = hs()
    assert s.pair_energy(0, 0, 0, 0, 2.01 ** 2) == 0.0
    assert s.pair_energy(0, 0, 0, 0, 1.99 ** 2) == math.inf
    assert s.pair_energy(0, 0, 1, 0, 5.01 ** 2) == 0.0
    assert s.pair_energy(0, 0, 1, 0, 4.99 ** 2) == math.inf
    s = hs([{"A B": {"sigma": 6}}])
    assert s.pair_energy(0, 0, 1, 0, 6.01 ** 2) == 0.0
    assert s.pair_energy(0, 0, 1, 0, 5.99 ** 2) == math.inf


def test_functor_potential_composition_doctest():
    # potentials.cpp:1333-1368: default [coulomb plain], "A B" [coulomb, wca], "C C" [hardsphere]
    cfg = {"temperature": 298.15, "geometry": {"type": "cuboid", "length": 100},
           "atomlist": [{"A": {"q": 1.0, "r": 1.1, "eps": 0.1}}, {"B": {"q": -1.0, "r": 2.0, "eps": 0.05}},
                        {"C": {"r": 1.0}}],
           "energy": [{"nonbonded": {
               "default": [{"coulomb": {"epsr": 80.0, "type": "plain"}}],
               "A B": [{"coulomb": {"epsr": 80.0, "type": "plain"}}, {"wca": {"mixing": "LB"}}],
               "C C": [{"hardsphere": {}}]}}]}
    s = O.System(cfg)
    lB = O.bjerrum_length(298.15, 80.0)
    eps = np.finfo(float).eps
    r = 2.0
    coul = lambda qa, qb: lB * qa * qb / (r + eps)
    kJ = 1e3 / (298.15 * 1.380658e-23 * 6.022137e23)
    sig = 0.5 * (2.2 + 4.0)
    epsab = math.sqrt(0.1 * 0.05) * kJ
    wca = 4 * epsab * ((sig / r) ** 12 - (sig / r) ** 6 + 0.25) if r * r <= 2 ** (1 / 3) * sig * sig else 0.0
    assert s.pair_energy(0, 1.0, 0, 1.0, r * r) == pytest.approx(coul(1, 1), rel=1e-12)
    assert s.pair_energy(0, 1.0, 1, -1.0, r * r) == pytest.approx(coul(1, -1) + wca, rel=1e-12)
    assert s.pair_energy(2, 0.0, 2, 0.0, (2.02) ** 2) == 0.0
    assert s.pair_energy(2, 0.0, 2, 0.0, (1.98) ** 2) == math.inf


# --- whole-configuration golden energies (example state files) -----------------
def test_minimal_example_initial_energy():
    d = np.load(systems_path("minimal.npz"))
    cfg = {"temperature": 300, "geometry": {"type": "cuboid", "length": 50},
           "atomlist": [{"Na": {"q": 1.0, "eps": 0.15, "sigma": 4.0, "dp": 50}},
                        {"Cl": {"q": -1.0, "eps": 0.20, "sigma": 4.0, "dp": 50}}],
           "moleculelist": [{"salt": {"atoms": ["Na", "Cl"], "atomic": True}}],
           "energy": [{"nonbonded": {"default": [{"lennardjones": {"mixing": "LB"}},
                                                 {"coulomb": {"type": "plain", "epsr": 80}}]}}]}
    s = O.System(cfg)
    s.set_particles(d["xyz"], d["q"], d["ids"])
    s.set_groups([0], [len(d["q"])], [0])
    e0 = s.init()
    # examples/minimal/minimal.out.json:10 prints 5 significant digits
    assert round(e0, 3) == float(d["init_energy"])


def test_bulk_example_initial_energy_within_reference_tolerance():
    # examples/bulk/bulk.out.json:28; the reference checks this run with
    # jsoncompare --tol 0.05 (examples/CMakeLists.txt:242-248).  The Fanourgakis
    # self energy inside CoulombGalore is unpinned, so only that tolerance applies.
    d = np.load(systems_path("bulk.npz"))
    cfg = {"temperature": 1100, "geometry": {"type": "cuboid", "length": 42.5},
           "atomlist": [{"Na": {"q": 1.0, "sigma": 3.33, "eps": 0.01158968}},
                        {"Cl": {"q": -1.0, "sigma": 4.4, "eps": 0.4184}}],
           "moleculelist": [{"salt": {"atoms": ["Na", "Cl"], "atomic": True}}],
           "energy": [{"nonbonded_coulomblj": {"lennardjones": {"mixing": "LB"},
                                               "coulomb": {"type": "fanourgakis", "epsr": 1, "cutoff": 14}}}]}
    s = O.System(cfg)
    s.set_particles(d["xyz"], d["q"], d["ids"])
    s.set_groups([0], [len(d["q"])], [0])
    e0 = s.init()
    assert abs(e0 - float(d["init_energy"])) < 0.05 * abs(float(d["init_energy"]))


def systems_path(name):
    import os
    return os.path.join(os.path.dirname(__file__), "golden", name)


# --- closed forms -----------------------------------------------------------------
def test_madelung_constant_rocksalt():
    """Real + reciprocal + self Ewald energy of a rock-salt lattice equals the
    textbook Madelung energy -M lB/d per ion pair (M = 1.747564594633)."""
    d = 2.8
    n = 4  # ions per edge, box = n*d
    box = n * d
    pos, q, ids = [], [], []
    for i in range(n):
        for j in range(n):
            for k in range(n):
                pos.append([i * d - box / 2 + 0.25, j * d - box / 2 + 0.25, k * d - box / 2 + 0.25])
                s = 1.0 if (i + j + k) % 2 == 0 else -1.0
                q.append(s)
                ids.append(0 if s > 0 else 1)
    rc = box / 2
    alpha = 3.6 / rc
    cfg = {"temperature": 298.15, "geometry": {"type": "cuboid", "length": box},
           "atomlist": [{"Na": {"q": 1.0}}, {"Cl": {"q": -1.0}}],
           "energy": [{"nonbonded": {"default": [{"coulomb": {
               "type": "ewald", "epsr": 80.0, "cutoff": rc, "alpha": alpha, "ncutoff": 12, "splined": False}}]}}]}
    s = O.System(cfg)
    s.set_particles(np.array(pos), np.array(q), np.array(ids, np.int32))
    s.set_groups([0], [len(q)], [0])
    e = s.init()
    lB = O.bjerrum_length(298.15, 80.0)
    madelung = -1.747564594633 * lB / d * (len(q) / 2)
    assert e == pytest.approx(madelung, rel=2e-6)
    # a dropped self term or a wrong reciprocal prefactor fails this by >1 %
    _, terms = s.total_energy()
    assert abs(terms[0]) > 0.05 * abs(e)


# --- invariants of the restated MC path ----------------------------------------------
def test_delta_u_equals_difference_of_full_energies_and_partial_equals_full_q():
    cfg, xyz, q, ids, g = systems.primitive_model(n_pairs=20, box=40.0, seed=3, cutoff=20.0, kmax=5)
    s = O.System(cfg)
    s.set_particles(xyz, q, ids)
    s.set_groups(g["begin"], g["size"], g["molid"])
    s.init()
    rng = np.random.default_rng(7)
    for _ in range(5):
        i = int(rng.integers(0, len(q)))
        new = xyz[i] + rng.normal(size=3)
        new -= 40.0 * np.round(new / 40.0)
        ch = {"groups": [{"group_index": 0, "relative_atom_indices": [i], "internal": True}]}
        en, eo = s.energy_change(ch, [i], [new])
        # full energies of both configurations, fresh systems (full Q recompute)
        full = []
        for X in (xyz, np.where(np.arange(len(q))[:, None] == i, new, xyz)):
            t = O.System(cfg)
            t.set_particles(X, q, ids)
            t.set_groups(g["begin"], g["size"], g["molid"])
            t.init()
            full.append(t.total_energy()[1])
        if np.isfinite(en[1]) and np.isfinite(full[1][1]):
            assert en[1] - eo[1] == pytest.approx(full[1][1] - full[0][1], rel=1e-9, abs=1e-9)
        assert en[2] == pytest.approx(full[1][2], rel=1e-11)   # partial Q update == full Q
        assert eo[2] == pytest.approx(full[0][2], rel=1e-11)
        # reverse move gives the negated change
        if np.isfinite(en[1]):
            t = O.System(cfg)
            X = xyz.copy()
            X[i] = new
            t.set_particles(X, q, ids)
            t.set_groups(g["begin"], g["size"], g["molid"])
            t.init()
            rn, ro = t.energy_change(ch, [i], [xyz[i]])
            assert (rn - ro)[1:].sum() == pytest.approx(-((en - eo)[1:].sum()), rel=1e-9, abs=1e-9)


def test_mc_drift_is_zero():
    """Sum of accepted du equals E_final - E_initial (montecarlo.cpp:85-99)."""
    cfg, xyz, q, ids, g = systems.primitive_model(n_pairs=30, box=50.0, seed=5, cutoff=20.0, kmax=5, dp=5.0)
    s = O.System(cfg)
    s.set_particles(xyz, q, ids)
    s.set_groups(g["begin"], g["size"], g["molid"])
    e0 = s.init()
    rec = s.run_translate(3000, seed=11)
    moves, acc, sum_du = s.stats()
    assert moves == 3000 and 0 < acc < 3000
    t = O.System(cfg)
    t.set_particles(s.positions(), q, ids)
    t.set_groups(g["begin"], g["size"], g["molid"])
    e1 = t.init()
    assert (e1 - (e0 + sum_du)) / abs(e0) == pytest.approx(0.0, abs=1e-9)
    # replaying the recorded stream reproduces every du and decision
    r = O.System(cfg)
    r.set_particles(xyz, q, ids)
    r.set_groups(g["begin"], g["size"], g["molid"])
    r.init()
    out = r.replay_translate(rec["group"], rec["index"], rec["xyz"], rec["u"])
    assert (out["accepted"] == rec["accepted"]).all()
    assert np.array_equal(out["du"], rec["du"])
