"""The oracle's nonbonded_splined term (SplinedPotential, potentials.h:770-824 and
potentials.cpp:1452-1600) pinned by what the reference's algorithm fixes:

- the spline follows the exact pair energy to the tolerance over the whole splined
  interval (the Andrea tabulator itself is pinned by its doctest in
  test_oracle_kat.py), so evaluating in r instead of r^2, dropping a term of the
  functor or a wrong knot order fails;
- zero from rmax on, the exact energy at and below rmin, +inf below rmin with
  `hardsphere` for a repulsive pair and the exact energy for an attractive one;
- rmin is where |u| first lies within 10 % of u_at_rmin on the 0.01 A walk from
  half the summed sigmas, and |u(rmax)| <= u_at_rmax with rmax from "rmax" (or
  "cutoff_g2g") unless pushed out;
- tables are symmetric in the type pair and use the atom types' charges;
- a move's du is the difference of full O(N^2) splined energies."""
import json

import numpy as np
import pytest

from oracle import oracle as O
from tests import systems

EWALD12 = {"type": "ewald", "epsr": 78.7, "cutoff": 12.0, "alpha": 0.25, "ncutoff": 6}


def ions_cfg(extra, default=None, box=40.0):
    default = default or [{"lennardjones": {"mixing": "LB"}}, {"coulomb": dict(EWALD12)}]
    term = {"default": default}
    term.update(extra)
    return systems.ions([{"nonbonded_splined": term}], n_pairs=60, box=box, seed=2)


def make(extra, default=None):
    cfg, xyz, q, ids, g = ions_cfg(extra, default)
    s = O.System(cfg)
    s.set_particles(xyz, q, ids)
    s.set_groups(g["begin"], g["size"], g["molid"])
    return s, (cfg, xyz, q, ids, g)


CHARGES = (1.0, -1.0)


@pytest.mark.parametrize("utol", [1e-3, 1e-4])
def test_spline_follows_the_exact_energy(utol):
    s, _ = make(dict(rmax=11.99, u_at_rmax=0.01, utol=utol))
    for a, b in [(0, 0), (0, 1), (1, 1)]:
        t = s.splined_table(a, b)
        assert t["knots"][0] == pytest.approx(t["rmin2"]) and t["knots"][-1] == pytest.approx(t["rmax2"])
        assert np.all(np.diff(t["knots"]) > 0)
        r = np.linspace(np.sqrt(t["rmin2"]) * 1.0001, np.sqrt(t["rmax2"]) * 0.9999, 3000)
        qa, qb = CHARGES[a], CHARGES[b]
        spl = np.array([s.pair_energy(a, qa, b, qb, x * x) for x in r])
        ex = np.array([s.pair_energy_exact(a, qa, b, qb, x * x) for x in r])
        # the tabulator checks 11 points per interval against utol; between them the
        # quintic stays within a small multiple
        assert np.max(np.abs(spl - ex)) <= 2 * utol, (a, b, np.max(np.abs(spl - ex)))
        assert np.max(np.abs(spl - ex)) > 1e-9  # it is a spline, not the exact function


def test_boundaries_and_search_rules():
    s, _ = make(dict(rmax=11.99, u_at_rmax=0.01))
    sig = (3.3, 4.4)
    for a, b in [(0, 0), (0, 1), (1, 1)]:
        t = s.splined_table(a, b)
        qa, qb = CHARGES[a], CHARGES[b]
        rmin, rmax = np.sqrt(t["rmin2"]), np.sqrt(t["rmax2"])
        assert rmax == pytest.approx(11.99, abs=1e-12)
        assert s.pair_energy(a, qa, b, qb, t["rmax2"]) == 0.0
        assert s.pair_energy(a, qa, b, qb, 200.0) == 0.0
        # at and below rmin: the exact energy
        for r2 in (t["rmin2"], 0.9 * t["rmin2"], 0.5 * t["rmin2"]):
            assert s.pair_energy(a, qa, b, qb, r2) == s.pair_energy_exact(a, qa, b, qb, r2)
        # findLowerDistance: |u(rmin)| within 10 % of u_at_rmin = 20, reached on the
        # 0.01 A grid from half the summed sigmas
        u_rmin = abs(s.pair_energy_exact(a, qa, b, qb, rmin * rmin))
        assert 20 / 1.1 <= u_rmin <= 20 * 1.1, (a, b, rmin, u_rmin)
        steps = (rmin - 0.5 * (sig[a] + sig[b])) / 0.01
        assert abs(steps - round(steps)) < 1e-6
        # one step further out the energy is already below the band
        assert abs(s.pair_energy_exact(a, qa, b, qb, (rmin + 0.01) ** 2)) < 20 * 1.1
        assert abs(s.pair_energy_exact(a, qa, b, qb, rmax * rmax)) <= 0.01
        # symmetric tables
        u = s.splined_table(b, a)
        assert np.array_equal(u["knots"], t["knots"]) and np.array_equal(u["coeff"], t["coeff"])


def test_rmax_is_pushed_out_until_u_at_rmax():
    """With the default u_at_rmax = 1e-6 the LJ tail at 10 A is larger, so rmax walks out
    in 0.01 A steps until |u| <= 1e-6 (findUpperDistance)."""
    s, _ = make(dict(rmax=10.0), default=[{"lennardjones": {"mixing": "LB"}}])
    for a, b in [(0, 0), (1, 1)]:
        t = s.splined_table(a, b)
        rmax = np.sqrt(t["rmax2"])
        assert rmax > 10.5
        assert abs(s.pair_energy_exact(a, 0, b, 0, rmax * rmax)) <= 1e-6
        assert abs(s.pair_energy_exact(a, 0, b, 0, (rmax - 0.01) ** 2)) > 1e-6
        assert abs((rmax - 10.0) / 0.01 - round((rmax - 10.0) / 0.01)) < 1e-6


def test_cutoff_g2g_sets_rmax():
    s, _ = make(dict(cutoff_g2g=9.0, u_at_rmax=0.05))
    assert np.sqrt(s.splined_table(0, 1)["rmax2"]) == pytest.approx(9.0, abs=1e-12)


def test_hardsphere_below_rmin_only_for_repulsive_pairs():
    # Coulomb only: like charges repel at rmin (u = +20 kT), unlike ones attract (-20 kT)
    s, _ = make(dict(rmax=11.99, u_at_rmax=0.01, hardsphere=True), default=[{"coulomb": dict(EWALD12)}])
    like = s.splined_table(0, 0)
    unlike = s.splined_table(0, 1)
    assert like["hardsphere"] and not unlike["hardsphere"]
    assert s.pair_energy(0, 1.0, 0, 1.0, 0.9 * like["rmin2"]) == np.inf
    r2 = 0.9 * unlike["rmin2"]
    assert s.pair_energy(0, 1.0, 1, -1.0, r2) == s.pair_energy_exact(0, 1.0, 1, -1.0, r2)


def test_tables_use_the_atom_type_charges():
    """A particle whose charge differs from its type's gets the type's table above rmin
    (the tables are built from Particle(AtomData)) and its own charge below rmin."""
    s, _ = make(dict(rmax=11.99, u_at_rmax=0.01))
    t = s.splined_table(0, 1)
    r2 = 0.5 * (t["rmin2"] + t["rmax2"])
    assert s.pair_energy(0, 0.5, 1, -1.0, r2) == s.pair_energy(0, 1.0, 1, -1.0, r2)
    r2 = 0.8 * t["rmin2"]
    assert s.pair_energy(0, 0.5, 1, -1.0, r2) == s.pair_energy_exact(0, 0.5, 1, -1.0, r2)


def test_the_spline_fails_across_the_coulomb_cutoff():
    """Tabulating across the real-space cutoff (rmax = Rc) meets the jump of the Coulomb
    term at Rc and the tabulator gives up, as the reference's does."""
    with pytest.raises(RuntimeError, match="Andrea"):
        make(dict(rmax=12.0, u_at_rmax=0.01))


def test_move_du_is_the_difference_of_full_energies():
    s, (cfg, xyz, q, ids, g) = make(dict(rmax=11.99, u_at_rmax=0.01))
    s.init()
    rec = s.run_translate(200, seed=4)
    acc = rec["accepted"].astype(bool)
    e_end = s.total_energy()[0]
    ref = O.System(cfg)
    ref.set_particles(xyz, q, ids)
    ref.set_groups(g["begin"], g["size"], g["molid"])
    e0 = ref.init()
    assert e_end - e0 == pytest.approx(rec["du"][acc].sum(), rel=1e-9, abs=1e-9)


def test_lennardjones_cutoff_is_refused():
    cfg, xyz, q, ids, g = systems.ions([{"nonbonded": {"default": [
        {"lennardjones": {"mixing": "LB", "cutoff": 12.0}}, {"coulomb": dict(EWALD12)}]}}], n_pairs=10)
    with pytest.raises(RuntimeError, match="no cutoff"):
        O.System(cfg)
