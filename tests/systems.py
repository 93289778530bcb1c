"""Synthetic systems of the benchmark configurations, built identically for the
oracle and the HIP library (same seeds, same random sequential addition)."""
import math

import numpy as np

AVOGADRO = 6.022137e23


def rsa_positions(n, box, sigma, rng, max_tries=2_000_000):
    """Random sequential addition of n non-overlapping spheres in a cubic periodic box
    centred at the origin; ``sigma`` is the per-particle diameter array."""
    pos = np.zeros((n, 3))
    ncell = max(1, int(box // max(sigma.max(), 1e-9)))
    cell = box / ncell
    grid = {}
    placed = 0
    tries = 0
    while placed < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("RSA failed; box too dense")
        p = (rng.random(3) - 0.5) * box
        c = tuple(((p + box / 2) // cell).astype(int) % ncell)
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    key = ((c[0] + dx) % ncell, (c[1] + dy) % ncell, (c[2] + dz) % ncell)
                    for j in grid.get(key, ()):
                        d = p - pos[j]
                        d -= box * np.round(d / box)
                        s = 0.5 * (sigma[placed] + sigma[j])
                        if d @ d < s * s:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if ok:
            pos[placed] = p
            grid.setdefault(c, []).append(placed)
            placed += 1
    return pos


def primitive_model(n_pairs=100, box=118.0, seed=1, alpha=0.2, cutoff=30.0, kmax=7, epsr=78.7, splined=True,
                    dp=10.0):
    """1:1 primitive-model electrolyte: hard spheres sigma=4 A + Ewald."""
    cfg = {
        "temperature": 298.15,
        "geometry": {"type": "cuboid", "length": box},
        "atomlist": [{"Na": {"q": 1.0, "sigma": 4.0, "dp": dp}}, {"Cl": {"q": -1.0, "sigma": 4.0, "dp": dp}}],
        "moleculelist": [{"salt": {"atoms": ["Na", "Cl"], "atomic": True}}],
        "energy": [{"nonbonded": {"default": [
            {"hardsphere": {}},
            {"coulomb": {"type": "ewald", "epsr": epsr, "cutoff": cutoff, "alpha": alpha, "ncutoff": kmax,
                         "splined": splined}}]}}],
    }
    rng = np.random.default_rng(seed)
    n = 2 * n_pairs
    ids = np.array([0, 1] * n_pairs, np.int32)
    q = np.where(ids == 0, 1.0, -1.0)
    xyz = rsa_positions(n, box, np.full(n, 4.0), rng)
    return cfg, xyz, q, ids, dict(begin=[0], size=[n], molid=[0])


def nacl_lj(n_pairs=10000, box=255.0, seed=1, alpha=0.25, cutoff=12.0, kmax=20, lj_cutoff=12.0, splined=True,
            dp=2.0):
    """Na/Cl Lennard-Jones (Lorentz-Berthelot) ions + Ewald, LJ truncated at lj_cutoff."""
    lj = {"mixing": "LB"}
    if lj_cutoff is not None:
        lj["cutoff"] = lj_cutoff
    cfg = {
        "temperature": 298.15,
        "geometry": {"type": "cuboid", "length": box},
        "atomlist": [{"Na": {"q": 1.0, "sigma": 3.3, "eps": 0.5, "dp": dp}},
                     {"Cl": {"q": -1.0, "sigma": 4.4, "eps": 0.4, "dp": dp}}],
        "moleculelist": [{"salt": {"atoms": ["Na", "Cl"], "atomic": True}}],
        "energy": [{"nonbonded": {"default": [
            {"lennardjones": lj},
            {"coulomb": {"type": "ewald", "epsr": 78.7, "cutoff": cutoff, "alpha": alpha, "ncutoff": kmax,
                         "splined": splined}}]}}],
    }
    rng = np.random.default_rng(seed)
    n = 2 * n_pairs
    ids = np.array([0, 1] * n_pairs, np.int32)
    q = np.where(ids == 0, 1.0, -1.0)
    sig = np.where(ids == 0, 3.3, 4.4)
    xyz = rsa_positions(n, box, sig * 0.9, rng)
    return cfg, xyz, q, ids, dict(begin=[0], size=[n], molid=[0])


def salt21_box(n, conc_molar):
    """Box length (A) of a 2:1 salt with n ions (n/3 formula units) at conc_molar."""
    return ((n / 3) / (conc_molar * AVOGADRO * 1e-27)) ** (1.0 / 3.0)


def salt21(n=600, conc=0.1, seed=1, temperature=298.15, kmax=8, splined=True):
    """2:1 salt (q=+2 and q=-1, sigma=4 A) hard spheres + Ewald alpha=5/L, Rc=L/2;
    n is the number of ions and must be a multiple of 3."""
    if n % 3:
        raise ValueError("salt21: n must be a multiple of 3 (one M and two X per formula unit)")
    box = salt21_box(n, conc)
    cfg = {
        "temperature": temperature,
        "geometry": {"type": "cuboid", "length": box},
        "atomlist": [{"M": {"q": 2.0, "sigma": 4.0, "dp": 0.2 * box}}, {"X": {"q": -1.0, "sigma": 4.0, "dp": 0.2 * box}}],
        "moleculelist": [{"salt": {"atoms": ["M", "X"], "atomic": True}}],
        "energy": [{"nonbonded": {"default": [
            {"hardsphere": {}},
            {"coulomb": {"type": "ewald", "epsr": 78.7, "cutoff": box / 2, "alpha": 5.0 / box, "ncutoff": kmax,
                         "splined": splined}}]}}],
    }
    rng = np.random.default_rng(seed)
    ids = np.array([0, 1, 1] * (n // 3), np.int32)
    q = np.where(ids == 0, 2.0, -1.0)
    xyz = rsa_positions(n, box, np.full(n, 4.0), rng)
    return cfg, xyz, q, ids, dict(begin=[0], size=[n], molid=[0])
