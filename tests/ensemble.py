"""Ensemble averages for the statistical parity checks: the reference's example
systems (configuration files and golden averages from examples/minimal and
examples/bulk, extracted into tests/golden/*.npz) and a pair distribution function."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def minimal_system():
    """examples/minimal/minimal.yml: 20 Na + 20 Cl, LJ (LB) + plain Coulomb (epsr 80),
    L = 50 A, T = 300 K, dp = 50 A; starting state examples/minimal/minimal.state.json."""
    d = np.load(os.path.join(HERE, "golden", "minimal.npz"))
    cfg = {"temperature": 300, "geometry": {"type": "cuboid", "length": 50},
           "atomlist": [{"Na": {"q": 1.0, "eps": 0.15, "sigma": 4.0, "dp": 50}},
                        {"Cl": {"q": -1.0, "eps": 0.20, "sigma": 4.0, "dp": 50}}],
           "moleculelist": [{"salt": {"atoms": ["Na", "Cl"], "atomic": True}}],
           "energy": [{"nonbonded": {"default": [{"lennardjones": {"mixing": "LB"}},
                                                 {"coulomb": {"type": "plain", "epsr": 80}}]}}]}
    n = len(d["q"])
    return (cfg, d["xyz"], d["q"], d["ids"], dict(begin=[0], size=[n], molid=[0])), float(d["mean_energy"])


def bulk_system():
    """examples/bulk/bulk.yml: 1152 Na + 1152 Cl, LJ (LB) + Fanourgakis Coulomb
    (epsr 1, cutoff 14), L = 42.5 A, T = 1100 K, dp 1.0 (Na) / 0.5 (Cl)."""
    d = np.load(os.path.join(HERE, "golden", "bulk.npz"))
    cfg = {"temperature": 1100, "geometry": {"type": "cuboid", "length": 42.5},
           "atomlist": [{"Na": {"q": 1.0, "sigma": 3.33, "eps": 0.01158968, "dp": 1.0}},
                        {"Cl": {"q": -1.0, "sigma": 4.4, "eps": 0.4184, "dp": 0.5}}],
           "moleculelist": [{"salt": {"atoms": ["Na", "Cl"], "atomic": True}}],
           "energy": [{"nonbonded_coulomblj": {"lennardjones": {"mixing": "LB"},
                                               "coulomb": {"type": "fanourgakis", "epsr": 1, "cutoff": 14}}}]}
    n = len(d["q"])
    return (cfg, d["xyz"], d["q"], d["ids"], dict(begin=[0], size=[n], molid=[0])), float(d["mean_energy"])


def rdf_counts(xyz, ids, box, a, b, rmax, dr):
    """Histogram of a-b pair distances (minimum image) up to rmax."""
    pa, pb = xyz[ids == a], xyz[ids == b]
    d = pa[:, None, :] - pb[None, :, :]
    d -= box * np.round(d / box)
    r = np.sqrt((d * d).sum(-1)).ravel()
    return np.histogram(r[r < rmax], bins=int(round(rmax / dr)), range=(0.0, rmax))[0].astype(np.float64)


def block_stats(samples, blocks=20):
    """Mean and standard error from block averages (rows are consecutive samples)."""
    s = np.asarray(samples, np.float64)
    nb = len(s) // blocks
    bm = s[:nb * blocks].reshape(blocks, nb, *s.shape[1:]).mean(axis=1)
    return bm.mean(axis=0), bm.std(axis=0, ddof=1) / np.sqrt(blocks)
