"""The oracle's group-to-group cutoff (GroupCutoff::cut, energy.h:761-768; from_json,
energy.cpp:1897-1933) against a brute-force sum: the nonbonded term of rigid water
equals the sum of the pair energies of every pair of atoms in different molecules whose
mass centres are closer than the cutoff (minimum image), for a number, a {default}
object and a molecule-pair entry; a cutoff larger than the box equals no cutoff, and an
unknown molecule name is refused."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import systems


def nonbonded(cfg, xyz, q, ids, g):
    s = O.System(cfg)
    s.set_particles(xyz, q, ids)
    s.set_groups(g["begin"], g["size"], g["molid"])
    s.init()
    tot, terms = s.total_energy()
    return s, terms[s.term_names().index("nonbonded")]


def brute(s, cfg, xyz, q, ids, g, cut):
    box = cfg["geometry"]["length"]
    cms = np.array([s.mass_center(k) for k in range(len(g["begin"]))])
    grp = np.repeat(np.arange(len(g["begin"])), g["size"])
    total = 0.0
    n = len(q)
    for i in range(n):
        for j in range(i + 1, n):
            if grp[i] == grp[j]:
                continue  # rigid molecules: no internal pairs
            dc = cms[grp[i]] - cms[grp[j]]
            dc -= box * np.round(dc / box)
            if dc @ dc >= cut * cut:
                continue
            d = xyz[i] - xyz[j]
            d -= box * np.round(d / box)
            total += s.pair_energy(int(ids[i]), float(q[i]), int(ids[j]), float(q[j]), float(d @ d))
    return total


@pytest.mark.parametrize("form", ["number", "default", "pair"])
def test_g2g_cut_matches_brute_force(form):
    cut = 4.0
    cfg, xyz, q, ids, g, _ = systems.water_g2g(n_mol=27, seed=5, kmax=3, cutoff=4.5, alpha=0.3)
    value = {"number": cut, "default": {"default": cut}, "pair": {"water water": cut}}[form]
    cfg["energy"][0]["nonbonded_coulomblj"]["cutoff_g2g"] = value
    s, e = nonbonded(cfg, xyz, q, ids, g)
    ref = brute(s, cfg, xyz, q, ids, g, cut)
    assert abs(e - ref) <= 1e-12 * max(1.0, abs(ref)), (e, ref)
    # some group pairs are cut at this distance, so the cut changed the energy
    big = dict(cfg)
    big["energy"] = [{"nonbonded_coulomblj": dict(cfg["energy"][0]["nonbonded_coulomblj"], cutoff_g2g=1e6)}]
    _, e_all = nonbonded(big, xyz, q, ids, g)
    assert abs(e_all - e) > 1e-3


def test_g2g_beyond_the_box_equals_no_cutoff():
    cfg, xyz, q, ids, g, _ = systems.water_g2g(n_mol=27, seed=5, kmax=3, cutoff=4.5, alpha=0.3)
    _, with_cut = nonbonded(cfg, xyz, q, ids, g)  # 10 A: beyond every centre distance of a 9.3 A box
    del cfg["energy"][0]["nonbonded_coulomblj"]["cutoff_g2g"]
    _, without = nonbonded(cfg, xyz, q, ids, g)
    assert with_cut == without


def test_g2g_unknown_molecule_is_refused():
    cfg, xyz, q, ids, g, _ = systems.water_g2g(n_mol=27, seed=5, kmax=3, cutoff=4.5, alpha=0.3)
    cfg["energy"][0]["nonbonded_coulomblj"]["cutoff_g2g"] = {"water ice": 5.0}
    with pytest.raises(RuntimeError, match="unknown molecule"):
        O.System(cfg)
