"""Ensemble averages of the device-resident Metropolis forms: mean energy and
acceptance on the reference's examples (golden values from examples/minimal and
examples/bulk within the reference's own jsoncompare tolerances), and the FP32-pair
form's Na-Cl g(r) and mean energy against the FP64 oracle's within statistical
error (independent chains)."""
import numpy as np
import pytest

import faunus_amd as F
from oracle import oracle as O
from tests import systems
from tests.ensemble import block_stats, bulk_system, minimal_system, rdf_counts
from tests.gpu_helpers import log, DRIFT_FP32

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fp32", [False, True])
def test_minimal_mean_energy_and_acceptance(fp32):
    sysd, golden = minimal_system()
    b = F.Batch(fp32_pairs=fp32)
    b.add(sysd, seed=21)
    energies = []
    for _ in range(800):
        b.run([4000])
        b.sync()
        energies.append(b.info(0)["energy"])
    mean, sem = block_stats(energies)
    info = b.info(0)
    acc = info["accepted"] / info["moves"]
    log("minimal fp32=%s mean %.3f +- %.3f (golden %.3f), acceptance %.3f" % (fp32, mean, sem, golden, acc))
    assert abs(mean - golden) <= 0.02 * abs(golden)
    assert abs(acc - 0.651) < 0.02
    tot, _ = b.total_energy(0)
    assert abs(info["energy"] - tot) <= (DRIFT_FP32 if fp32 else 1e-9) * abs(tot)


def test_bulk_mean_energy_and_acceptance():
    sysd, golden = bulk_system()
    b = F.Batch()
    b.add(sysd, seed=5)
    energies = []
    for _ in range(5):  # bulk.yml: 250 steps of 2304 moves, energy every 50 steps
        b.run([50 * 2304])
        b.sync()
        energies.append(b.info(0)["energy"])
    info = b.info(0)
    acc = info["accepted"] / info["moves"]
    log("bulk mean %.1f (golden %.1f), acceptance %.3f (reference 0.393)" % (np.mean(energies), golden, acc))
    assert abs(np.mean(energies) - golden) <= 0.05 * abs(golden)
    assert abs(acc - 0.393) < 0.03


def test_fp32_form_rdf_and_energy_match_fp64_oracle():
    sysd = systems.primitive_model(n_pairs=100, box=118.0, seed=1)
    cfg, xyz, q, ids, g = sysd
    box = cfg["geometry"]["length"]
    nsamp, stride = 400, 2000
    rmax, dr = 30.0, 1.0
    # GPU, FP32 pairs
    b = F.Batch(fp32_pairs=True)
    b.add(sysd, seed=101)
    b.run([20000])
    b.sync()
    eg, hg = [], []
    for _ in range(nsamp):
        b.run([stride])
        b.sync()
        eg.append(b.info(0)["energy"])
        hg.append(rdf_counts(b.positions(0), ids, box, 0, 1, rmax, dr))
    # oracle, FP64, independent stream
    s = O.System(cfg, fast=True)
    s.set_particles(xyz, q, ids)
    s.set_groups(g["begin"], g["size"], g["molid"])
    s.init()
    s.run_philox_translate(20000, seed=202, replica=0, move0=0)
    eo, ho = [], []
    for k in range(nsamp):
        s.run_philox_translate(stride, seed=202, replica=0, move0=20000 + k * stride)
        eo.append(s.total_energy()[0])
        ho.append(rdf_counts(s.positions(), ids, box, 0, 1, rmax, dr))
    mg, sg = block_stats(eg)
    mo, so = block_stats(eo)
    hmg, hsg = block_stats(hg)
    hmo, hso = block_stats(ho)
    z_e = abs(mg - mo) / np.hypot(sg, so)
    sig = np.hypot(hsg, hso)
    mask = (hmg + hmo) > 20  # bins with counts
    z_h = np.abs(hmg - hmo)[mask] / np.maximum(sig[mask], 1e-9)
    log("fp32 vs fp64 oracle: energy %.2f +- %.2f vs %.2f +- %.2f (z %.2f); g(r) max z %.2f over %d bins" % (
        mg, sg, mo, so, z_e, z_h.max(), mask.sum()))
    assert z_e < 4.0
    assert z_h.max() < 4.5
