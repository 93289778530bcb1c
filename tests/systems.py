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


def pair_cut(default, rmax):
    """The nonbonded term of a pair list cut at rmax: nonbonded_splined, whose tables are
    zero from rmax on (SplinedPotential, the reference's only way to truncate LJ),
    with u_at_rmax (0.1 kT) high enough that rmax stays where it is put rather than
    walking out onto the Coulomb cutoff; rmax None: the exact nonbonded term."""
    if rmax is None:
        return {"nonbonded": {"default": default}}
    return {"nonbonded_splined": {"default": default, "rmax": rmax, "u_at_rmax": 0.1}}


def nacl_lj(n_pairs=10000, box=255.0, seed=1, alpha=0.25, cutoff=12.0, kmax=20, rmax=11.99, splined=True,
            dp=2.0):
    """Na/Cl Lennard-Jones (Lorentz-Berthelot) ions + Ewald (real-space cutoff `cutoff`),
    the pair energy splined and cut at rmax (nonbonded_splined; 0.01 A inside the
    Coulomb cutoff, where the tabulated energy is continuous)."""
    cfg = {
        "temperature": 298.15,
        "geometry": {"type": "cuboid", "length": box},
        "atomlist": [{"Na": {"q": 1.0, "sigma": 3.3, "eps": 0.5, "dp": dp}},
                     {"Cl": {"q": -1.0, "sigma": 4.4, "eps": 0.4, "dp": dp}}],
        "moleculelist": [{"salt": {"atoms": ["Na", "Cl"], "atomic": True}}],
        "energy": [pair_cut([
            {"lennardjones": {"mixing": "LB"}},
            {"coulomb": {"type": "ewald", "epsr": 78.7, "cutoff": cutoff, "alpha": alpha, "ncutoff": kmax,
                         "splined": splined}}], rmax)],
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


def lattice_positions(n, box, sigma, rng):
    """n non-overlapping spheres of diameter <= sigma: a random subset of the sites of
    a simple cubic lattice (spacing a = box/m >= sigma) with per-site jitter of at
    most (a - sigma)/2 per coordinate, so any two centres stay >= sigma apart."""
    m = int(math.ceil(n ** (1.0 / 3.0)))
    if box / m < sigma:
        raise RuntimeError("lattice too dense for the requested sigma")
    a = box / m
    sites = rng.choice(m ** 3, size=n, replace=False)
    ijk = np.stack([sites % m, (sites // m) % m, sites // (m * m)], axis=1).astype(np.float64)
    jit = (a - sigma) / 2 * (2 * rng.random((n, 3)) - 1)
    return (ijk + 0.5) * a - box / 2 + jit


def ions(energy, n_pairs=100, box=60.0, seed=1, dp=5.0, temperature=298.15):
    """Na/Cl-like ions (q=+-1, sigma 3.3/4.4 A, eps 0.5/0.4 kJ/mol) in a cubic box with
    the given `energy` list (reference input form), random sequential addition start."""
    cfg = {
        "temperature": temperature,
        "geometry": {"type": "cuboid", "length": box},
        "atomlist": [{"Na": {"q": 1.0, "sigma": 3.3, "eps": 0.5, "dp": dp}},
                     {"Cl": {"q": -1.0, "sigma": 4.4, "eps": 0.4, "dp": dp}}],
        "moleculelist": [{"salt": {"atoms": ["Na", "Cl"], "atomic": True}}],
        "energy": energy,
    }
    rng = np.random.default_rng(seed)
    n = 2 * n_pairs
    ids = np.array([0, 1] * n_pairs, np.int32)
    q = np.where(ids == 0, 1.0, -1.0)
    sig = np.where(ids == 0, 3.3, 4.4)
    xyz = rsa_positions(n, box, sig, rng)
    return cfg, xyz, q, ids, dict(begin=[0], size=[n], molid=[0])


def water(n_mol=4096, box=None, seed=1, alpha=0.28, cutoff=10.0, kmax=9, rmax="auto", dp=0.5, dprot=0.5,
          splined=True):
    """Rigid 3-site water (O-H 1.0 A, H-O-H 109.47 deg, q_O=-0.8476, q_H=+0.4238, LJ
    on O sigma=3.166 A eps=0.650 kJ/mol; H sigma 2 A as in the reference's water
    examples, with eps 0), molecules on a cubic lattice with random orientations; Ewald
    alpha=0.28, Rc=10, kmax=9; the pair energy splined and cut at rmax (default 0.01 A
    inside Rc; None: the exact nonbonded term).  The move set is molecular
    translate+rotate with dp and dprot."""
    if box is None:
        box = 49.68 * (n_mol / 4096) ** (1.0 / 3.0)
    if rmax == "auto":
        rmax = cutoff - 0.01
    cfg = {
        "temperature": 298.15,
        "geometry": {"type": "cuboid", "length": box},
        "atomlist": [{"OW": {"q": -0.8476, "sigma": 3.166, "eps": 0.650, "mw": 15.999}},
                     {"HW": {"q": 0.4238, "sigma": 2.0, "eps": 0.0, "mw": 1.008}}],
        "moleculelist": [{"water": {"atoms": ["OW", "HW", "HW"], "rigid": True}}],
        "energy": [pair_cut([
            {"lennardjones": {"mixing": "LB"}},
            {"coulomb": {"type": "ewald", "epsr": 1.0, "cutoff": cutoff, "alpha": alpha, "ncutoff": kmax,
                         "splined": splined}}], rmax)],
    }
    rng = np.random.default_rng(seed)
    theta = math.radians(109.47)
    local = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [math.cos(theta), math.sin(theta), 0.0]])
    m = int(math.ceil(n_mol ** (1.0 / 3.0)))
    a = box / m
    sites = rng.choice(m ** 3, size=n_mol, replace=False)
    ijk = np.stack([sites % m, (sites // m) % m, sites // (m * m)], axis=1).astype(np.float64)
    centres = (ijk + 0.5) * a - box / 2
    xyz = np.zeros((3 * n_mol, 3))
    for w in range(n_mol):
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        pos = centres[w] + local @ q.T
        pos -= box * np.round(pos / box)
        xyz[3 * w:3 * w + 3] = pos
    ids = np.tile(np.array([0, 1, 1], np.int32), n_mol)
    q = np.where(ids == 0, -0.8476, 0.4238)
    groups = dict(begin=list(range(0, 3 * n_mol, 3)), size=[3] * n_mol, molid=[0] * n_mol)
    moves = [dict(kind=1, molid=0, weight=1, dp=dp, dprot=dprot)]
    return cfg, xyz, q, ids, groups, moves


def polyelectrolyte(n_beads, n_total, seed, kmax=8, dp=10.0, bead_dp=2.0, dprot=3.0, monomer_molar=0.005,
                    bond_k=0.5, bond_req=4.0, excluded_neighbours=0):
    """Bead-spring polyelectrolyte (n_beads, q=-1, sigma=4 A, harmonic bonds k=0.5
    kJ/mol/A^2, req=4 A) with n_beads counterions (q=+1) and 1:1 salt up to n_total
    particles; hard spheres + Ewald alpha=5/L, Rc=L/2, kmax=8; cubic box at the given
    monomer concentration; excluded_neighbours > 0 removes the nonbonded pairs of
    beads within that many bonds.  Groups: the chain (molecular), then one atomic
    group of ions.  The move set translates ions (dp) and beads (bead_dp) and pivots the chain
    (dprot), weighted by the number of ions, beads and bonds."""
    n_salt = max(0, (n_total - 2 * n_beads) // 2)
    n_ions = n_beads + 2 * n_salt
    box = (n_beads / (monomer_molar * AVOGADRO * 1e-27)) ** (1.0 / 3.0)
    bonds = [{"harmonic": {"index": [i, i + 1], "k": bond_k, "req": bond_req}} for i in range(n_beads - 1)]
    cfg = {
        "temperature": 298.15,
        "geometry": {"type": "cuboid", "length": box},
        "atomlist": [{"P": {"q": -1.0, "sigma": 4.0, "dp": bead_dp}}, {"C": {"q": 1.0, "sigma": 4.0, "dp": dp}},
                     {"A": {"q": -1.0, "sigma": 4.0, "dp": dp}}],
        "moleculelist": [{"chain": {"atoms": ["P"] * n_beads, "bondlist": bonds,
                                    "excluded_neighbours": excluded_neighbours}},
                         {"ions": {"atoms": ["C", "A"], "atomic": True}}],
        "energy": [{"bonded": {}},
                   {"nonbonded": {"default": [
                       {"hardsphere": {}},
                       {"coulomb": {"type": "ewald", "epsr": 78.7, "cutoff": box / 2, "alpha": 5.0 / box,
                                    "ncutoff": kmax}}]}}],
    }
    rng = np.random.default_rng(seed)
    # chain: random walk with step 4.2 A and consecutive bonds at least 75 deg from reversing
    chain = np.zeros((n_beads, 3))
    step = bond_req * 1.05
    prev = None
    i = 1
    tries = 0
    while i < n_beads:
        tries += 1
        if tries > 200 * n_beads:
            raise RuntimeError("chain growth failed")
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        if prev is not None and v @ prev < -0.25:
            continue
        p = chain[i - 1] + step * v
        d = chain[:i - 1] - p
        if i > 1 and (np.einsum("ij,ij->i", d, d) < 4.4 ** 2).any():
            continue
        chain[i] = p
        prev = v
        i += 1
    chain -= chain.mean(axis=0)
    chain -= box * np.round(chain / box)
    # ions on a jittered lattice, away from every bead
    m = int(math.ceil((2 * n_ions) ** (1.0 / 3.0)))
    a = box / m
    if a < 4.0:
        raise RuntimeError("too dense")
    sites = rng.permutation(m ** 3)
    cand = (np.stack([sites % m, (sites // m) % m, sites // (m * m)], axis=1).astype(np.float64) + 0.5) * a - box / 2
    cand += (a - 4.0) / 2 * (2 * rng.random(cand.shape) - 1)
    ok = np.ones(len(cand), bool)
    for b in chain:  # drop candidates closer than 4.4 A to any bead
        d = cand - b
        d -= box * np.round(d / box)
        ok &= np.einsum("ij,ij->i", d, d) >= 4.4 ** 2
    ions = cand[ok][:n_ions]
    if len(ions) < n_ions:
        raise RuntimeError("ion placement failed")
    xyz = np.concatenate([chain, ions])
    ids = np.concatenate([np.zeros(n_beads, np.int32), np.ones(n_beads + n_salt, np.int32),
                          np.full(n_salt, 2, np.int32)])
    q = np.where(ids == 1, 1.0, -1.0)
    groups = dict(begin=[0, n_beads], size=[n_beads, n_ions], molid=[0, 1])
    moves = [dict(kind=0, molid=1, weight=n_ions, dp=dp), dict(kind=0, molid=0, weight=n_beads, dp=bead_dp),
             dict(kind=2, molid=0, weight=max(1, n_beads - 1), dprot=dprot)]
    return cfg, xyz, q, ids, groups, moves


def water_g2g(n_mol=1000, box=None, seed=1, alpha=0.2, cutoff=9.0, kmax=7, g2g=10.0, dp=0.5, dprot=0.5):
    """Rigid water written as the reference's water example writes it
    (examples/water/ewald.yml): nonbonded_coulomblj with LB-mixed LJ that only the
    group-to-group cutoff truncates (cutoff_g2g on the O-H-H mass centres), Ewald
    real space cut at `cutoff`; geometry, charges and moves as water()."""
    cfg, xyz, q, ids, groups, moves = water(n_mol=n_mol, box=box, seed=seed, alpha=alpha, cutoff=cutoff, kmax=kmax,
                                            dp=dp, dprot=dprot)
    cfg["energy"] = [{"nonbonded_coulomblj": {
        "lennardjones": {"mixing": "LB"},
        "coulomb": {"type": "ewald", "epsr": 1.0, "cutoff": cutoff, "alpha": alpha, "ncutoff": kmax},
        "cutoff_g2g": g2g}}]
    return cfg, xyz, q, ids, groups, moves


def chains_g2g(n_chains=6, n_beads=12, box=70.0, seed=1, kmax=6, g2g=25.0, dp=6.0, bead_dp=2.0, dprot=2.0,
               bond_k=0.5, bond_req=4.0):
    """Several bead-spring chains (q=-1, sigma=4 A, harmonic bonds) with their counterions
    (q=+1), hard spheres + Ewald (alpha=5/L, Rc=L/2), and a group-to-group cutoff between
    the chains (cutoff_g2g: {default: g2g}); the counterions form one atomic group, which
    the cutoff never cuts.  Moves: ion and bead translations and pivots, weighted by the
    number of ions, beads and bonds."""
    bonds = [{"harmonic": {"index": [i, i + 1], "k": bond_k, "req": bond_req}} for i in range(n_beads - 1)]
    cfg = {
        "temperature": 298.15,
        "geometry": {"type": "cuboid", "length": box},
        "atomlist": [{"P": {"q": -1.0, "sigma": 4.0, "dp": bead_dp, "mw": 2.0}}, {"C": {"q": 1.0, "sigma": 4.0, "dp": dp}}],
        "moleculelist": [{"chain": {"atoms": ["P"] * n_beads, "bondlist": bonds}},
                         {"ions": {"atoms": ["C"], "atomic": True}}],
        "energy": [{"bonded": {}},
                   {"nonbonded": {"default": [
                       {"hardsphere": {}},
                       {"coulomb": {"type": "ewald", "epsr": 78.7, "cutoff": box / 2, "alpha": 5.0 / box,
                                    "ncutoff": kmax}}],
                       "cutoff_g2g": {"default": g2g}}}],
    }
    rng = np.random.default_rng(seed)
    placed = []
    step = bond_req * 1.05
    for c in range(n_chains):
        for _ in range(1000):
            start = (rng.random(3) - 0.5) * box
            chain = [start]
            ok = True
            for i in range(1, n_beads):
                for _ in range(200):
                    v = rng.normal(size=3)
                    v /= np.linalg.norm(v)
                    p = chain[-1] + step * v
                    pw = p - box * np.round(p / box)
                    near = [x for x in placed + chain[:-1]]
                    if near:
                        d = np.array(near) - pw
                        d -= box * np.round(d / box)
                        if (np.einsum("ij,ij->i", d, d) < 4.4 ** 2).any():
                            continue
                    chain.append(p)
                    break
                else:
                    ok = False
                    break
            if ok:
                placed.extend([p - box * np.round(p / box) for p in chain])
                break
        else:
            raise RuntimeError("chain placement failed")
    beads = np.array(placed)
    n_ions = n_chains * n_beads
    ions = []
    while len(ions) < n_ions:
        p = (rng.random(3) - 0.5) * box
        d = np.concatenate([beads, np.array(ions).reshape(-1, 3)]) - p
        d -= box * np.round(d / box)
        if (np.einsum("ij,ij->i", d, d) >= 4.4 ** 2).all():
            ions.append(p)
    xyz = np.concatenate([beads, np.array(ions)])
    ids = np.concatenate([np.zeros(n_ions, np.int32), np.ones(n_ions, np.int32)])
    q = np.where(ids == 1, 1.0, -1.0)
    begin = [c * n_beads for c in range(n_chains)] + [n_chains * n_beads]
    groups = dict(begin=begin, size=[n_beads] * n_chains + [n_ions], molid=[0] * n_chains + [1])
    moves = [dict(kind=0, molid=1, weight=n_ions, dp=dp), dict(kind=0, molid=0, weight=n_ions, dp=bead_dp),
             dict(kind=2, molid=0, weight=n_chains * (n_beads - 1), dprot=dprot)]
    return cfg, xyz, q, ids, groups, moves
