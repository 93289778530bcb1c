"""The device pair functions against the oracle's pair energy, pair by pair: the FP64
form (pair_energy) within 1e-11 * max(|u|, 1e-3 kT), and the FP32-pair form at the same
distance within 1e-6 * max(|u|, 8 kT) -- a few units of FP32 rounding of the largest
term of a pair (terms that cancel near the LJ minimum included).  Pairs inside the hot
radius (LJ or WCA repulsion above 8 kT) are evaluated again in FP64, so deep overlaps
(|u| > 1000 kT) must meet the FP64 bound.  Every type pair of each potential form the
parser takes, on a grid of distances from well inside contact to beyond the pair
range."""
import json

import numpy as np
import pytest

import faunus_amd as F
from oracle import oracle as O
from tests import systems
from tests.gpu_helpers import log
from tests.test_gpu_potentials import VARIANTS

pytestmark = pytest.mark.gpu

HOT = 8.0  # kT: floor of the FP32 bound

CASES = dict(VARIANTS)
CASES["splined_nacl"] = systems.nacl_lj(n_pairs=10, box=60.0, seed=1, kmax=6)[0]["energy"]
CASES["primitive_model"] = systems.primitive_model(n_pairs=10, box=60.0)[0]["energy"]


def config(name):
    if name == "water_splined":
        return systems.water(n_mol=8, seed=1, kmax=4)[0]
    return systems.ions(json.loads(json.dumps(CASES[name])), n_pairs=10, box=60.0, seed=1)[0]


@pytest.mark.parametrize("name", sorted(CASES) + ["water_splined"])
def test_pair_functions_match_the_oracle(name):
    cfg = config(name)
    orc = O.System(cfg)
    eng = F.Engine(cfg)
    atoms = [next(iter(a.values())) for a in cfg["atomlist"]]
    worst64 = worst32 = 0.0
    for a in range(len(atoms)):
        for b in range(len(atoms)):
            qa, qb = atoms[a]["q"], atoms[b]["q"]
            r = np.concatenate([np.linspace(0.6, 6.0, 800), np.linspace(6.0, 16.0, 400)])
            r2 = r * r
            u64, u32 = eng.pair_probe(a, b, qa, qb, r2)
            uo = np.array([orc.pair_energy(a, qa, b, qb, float(x)) for x in r2])
            inf = np.isinf(uo)
            assert np.array_equal(inf, np.isinf(u64)) and np.array_equal(inf, np.isinf(u32)), (name, a, b)
            fin = ~inf
            e64 = np.abs(u64[fin] - uo[fin]) / np.maximum(np.abs(uo[fin]), 1e-3)
            deep = np.abs(uo[fin]) > 1e3
            # FP32 bound, and the FP64 bound (in units of the FP32 one) on deep overlaps
            e32 = np.maximum(np.abs(u32[fin] - uo[fin]) / (1e-6 * np.maximum(np.abs(uo[fin]), HOT)),
                             np.where(deep, np.abs(u32[fin] - uo[fin]) / (1e-11 * np.maximum(np.abs(uo[fin]), 1e3)), 0.0))
            k = int(np.argmax(e32)) if len(e32) else 0
            log(name, a, b, "fp64 max rel %.2e" % (e64.max() if len(e64) else 0),
                "fp32 worst/bound %.3f at r=%.4f (u %.9g vs %.9g)" % (e32.max() if len(e32) else 0,
                                                                     r[fin][k] if len(e32) else 0,
                                                                     u32[fin][k] if len(e32) else 0,
                                                                     uo[fin][k] if len(e32) else 0))
            worst64 = max(worst64, float(e64.max()) if len(e64) else 0.0)
            worst32 = max(worst32, float(e32.max()) if len(e32) else 0.0)
    eng.close()
    assert worst64 <= 1e-11, (name, worst64)
    assert worst32 <= 1.0, (name, worst32)
