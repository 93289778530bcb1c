"""Molecular Changes through the per-move GPU path (the drop-in EnergyTerm form):
rigid-molecule translate+rotate (Change all=true, internal=false), single bead
translations inside a chain and pivot moves (internal=true with relative indices),
with the harmonic bonded term and an optional group cutoff.  The oracle runs the
Philox proposal stream and records every proposal; the GPU engine replays it and
must give the same du (1e-10 relative to max(|du|, E_scale)) and decisions."""
import math

import numpy as np
import pytest

import faunus_amd as F
from oracle import oracle as O
from tests import systems
from tests.gpu_helpers import log

pytestmark = pytest.mark.gpu


def accept_deviate(seed, replica, t):
    b = O.philox([t & 0xffffffff, t >> 32, replica, 1], [seed & 0xffffffff, seed >> 32])
    return ((b[0] >> 5) * 67108864.0 + (b[1] >> 6)) * 1.1102230246251565e-16


def energy_change(new, old):
    if math.isnan(old) and not math.isnan(new):
        return -math.inf
    if math.isnan(new) or (new > 0 and math.isinf(new)):
        return math.inf
    du = new - old
    return 0.0 if math.isnan(du) else du


def build(sysd):
    cfg, xyz, q, ids, g = sysd
    orc = O.System(cfg)
    orc.set_particles(xyz, q, ids)
    orc.set_groups(g["begin"], g["size"], g["molid"])
    e0 = orc.init()
    eng = F.Engine(cfg)
    eng.set_particles(xyz, q, ids)
    eng.set_groups(g["begin"], g["size"], g["molid"])
    e0g = eng.init()
    return orc, eng, e0, e0g


def replay(eng, rec, kinds, begin, molecular, seed, nmoves):
    du_out = np.zeros(nmoves)
    acc_out = np.zeros(nmoves, np.uint8)
    for t in range(nmoves):
        o0, o1 = int(rec["offsets"][t]), int(rec["offsets"][t + 1])
        if o1 == o0:
            acc_out[t] = 1
            continue
        k = kinds[rec["kind"][t]]
        g = int(rec["group"][t])
        idx = rec["index"][o0:o1]
        eng.stage_trial(idx, rec["xyz"][o0:o1])
        if molecular[g]:
            eng.stage_mass_center(g, rec["cm"][t])
        if k["kind"] == 1:
            change = {"groups": [{"group_index": g, "all": True, "internal": False}]}
        else:
            change = {"groups": [{"group_index": g, "internal": True,
                                  "relative_atom_indices": [int(i) - begin[g] for i in idx]}]}
        en, eo = eng.delta_u(change)
        du = energy_change(float(np.sum(en)), float(np.sum(eo)))
        u = accept_deviate(seed, 0, t)
        acc = du == -math.inf or -du > 709.782712893384 or u <= math.exp(-du)
        eng.sync(acc)
        du_out[t] = du
        acc_out[t] = acc
    return du_out, acc_out


def compare(rec, du, acc, e_scale):
    mism = np.nonzero(acc != rec["accepted"])[0]
    first = int(mism[0]) if len(mism) else len(du)
    fin = np.isfinite(rec["du"][:first])
    err = np.abs(du[:first][fin] - rec["du"][:first][fin]) / np.maximum(np.abs(rec["du"][:first][fin]), e_scale)
    return dict(first_mismatch=first, n=len(du), max_rel=float(err.max()) if len(err) else 0.0,
                accepted=int(acc[:first].sum()))


def run_case(sysd, kinds, nmoves, seed):
    cfg, xyz, q, ids, g = sysd
    orc, eng, e0, e0g = build(sysd)
    assert e0g == pytest.approx(e0, rel=1e-10, abs=1e-10)
    rec = orc.run_philox_moves(nmoves, seed=seed, kinds=kinds)
    molecular = []
    for m in g["molid"]:
        mol = list(cfg["moleculelist"][m].values())[0]
        molecular.append(not mol.get("atomic", False))
    du, acc = replay(eng, rec, kinds, list(g["begin"]), molecular, seed, nmoves)
    rep = compare(rec, du, acc, max(1.0, abs(e0)))
    tot, terms = eng.total_energy()
    otot, oterms = orc.total_energy()
    rep["final_rel"] = abs(tot - otot) / max(1.0, abs(otot))
    return rep, eng, orc, rec


@pytest.mark.parametrize("excl", [0, 2])
def test_polyelectrolyte_pivot_and_translations(excl):
    cfg, xyz, q, ids, g, kinds = systems.polyelectrolyte(40, 300, seed=7, kmax=6, excluded_neighbours=excl)
    rep, eng, orc, rec = run_case((cfg, xyz, q, ids, g), kinds, 1500, seed=21)
    log("polyelectrolyte", rep, "kinds", np.bincount(rec["kind"]))
    assert rep["first_mismatch"] == rep["n"], rep
    assert rep["max_rel"] <= 1e-10 and rep["final_rel"] <= 1e-10, rep
    np.testing.assert_allclose(eng.positions(), orc.positions(), rtol=0, atol=1e-9)
    dq, _ = eng.check_q()
    assert dq < 1e-9


def test_rigid_water_transrot():
    cfg, xyz, q, ids, g, kinds = systems.water(n_mol=216, seed=3, kmax=6, rmax=None, cutoff=9.0)
    rep, eng, orc, rec = run_case((cfg, xyz, q, ids, g), kinds, 800, seed=5)
    log("water transrot", rep)
    assert rep["first_mismatch"] == rep["n"], rep
    assert rep["max_rel"] <= 1e-10 and rep["final_rel"] <= 1e-10, rep
    np.testing.assert_allclose(eng.positions(), orc.positions(), rtol=0, atol=1e-9)


def test_water_with_cell_list_and_group_cutoff():
    """nonbonded_splined with cutoff_g2g = 11 A: the tables end at 11 A (the spline takes
    rmax from cutoff_g2g) inside the 11.5 A Coulomb cutoff, so the per-move path walks a
    cell list and cuts molecule pairs on their mass centres."""
    cfg, xyz, q, ids, g, kinds = systems.water(n_mol=2000, seed=4, kmax=7, cutoff=11.5)
    term = cfg["energy"][0]["nonbonded_splined"]
    del term["rmax"]
    term["cutoff_g2g"] = 11.0
    rep, eng, orc, rec = run_case((cfg, xyz, q, ids, g), kinds, 500, seed=6)
    log("water cells + g2g", rep)
    assert rep["first_mismatch"] == rep["n"], rep
    assert rep["max_rel"] <= 1e-10 and rep["final_rel"] <= 1e-10, rep
    # accepted mass centres follow the oracle's
    for grp in (0, 1, 500, 999):
        np.testing.assert_allclose(eng.mass_center(grp), orc.mass_center(grp), rtol=0, atol=1e-9)


def test_group_cutoff_needs_staged_mass_center():
    cfg, xyz, q, ids, g, kinds = systems.water(n_mol=64, seed=5, kmax=4, rmax=None, cutoff=6.0)
    cfg["energy"][0]["nonbonded"]["cutoff_g2g"] = 8.0
    eng = F.Engine(cfg)
    eng.set_particles(xyz, q, ids)
    eng.set_groups(g["begin"], g["size"], g["molid"])
    eng.init()
    eng.stage_trial([0, 1, 2], xyz[0:3] + 0.1)
    with pytest.raises(RuntimeError, match="mass centre"):
        eng.delta_u({"groups": [{"group_index": 0, "all": True}]})
