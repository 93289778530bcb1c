"""On-device observables: the device energy accumulators equal the running energies
read on the host at the same move indices (sum exactly, sum of squares to 1e-12), for
the translate-only, general, split general and split translate kernels; the device
pair-distance histogram equals a host histogram of the read-back positions computed
with the same arithmetic (exact integer counts), for unlike and like atom types."""
import numpy as np
import pytest

import faunus_amd as F
from tests import systems
from tests.gpu_helpers import log

pytestmark = pytest.mark.gpu


def host_rdf(xyz, ids, box, a, b, dr, nbins):
    """AtomRDF sampling restated with numpy: unlike types all (a, b) pairs, like types
    i < j; minimum image by +-L, r from x^2 + y^2 + z^2 rounded term by term, bin
    floor(r * (1 / dr))."""
    ia = np.nonzero(ids == a)[0]
    ib = np.nonzero(ids == b)[0]
    half = box / 2.0
    counts = np.zeros(nbins, np.uint64)
    inv_dr = 1.0 / dr
    for i in ia:
        js = ib[ib > i] if a == b else ib
        if len(js) == 0:
            continue
        d = xyz[i][None, :] - xyz[js]
        d = np.where(d > half, d - box, np.where(d < -half, d + box, d))
        r = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
        k = np.floor(r * inv_dr)
        k = k[k < nbins].astype(np.int64)
        counts += np.bincount(k, minlength=nbins).astype(np.uint64)
    return counts


def sampled_run(b, host, chunks, every, split=None):
    """Run `chunks` chunks of `every` moves, appending the energy after each to `host`;
    return the device sample count, sum and sum of squares, and the host's, summed in
    the same order."""
    for _ in range(chunks):
        if split is None:
            b.run([every])
        else:
            b.run_split(every, groups=split)
        b.sync()
        host.append(b.info(0)["energy"])
    st = b.energy_stats(0)
    s1 = 0.0
    s2 = 0.0
    for e in host:
        s1 += e
        s2 += e * e
    return st["samples"], st["sum"], st["sum_sq"], len(host), s1, s2


@pytest.mark.parametrize("form", ["translate", "general", "general_split", "translate_split"])
def test_device_energy_samples_equal_host_reads(form):
    if form.startswith("translate"):
        sysd = systems.salt21(n=300, seed=3, kmax=6)
        kinds = None
    else:
        cfg, xyz, q, ids, g, kinds = systems.polyelectrolyte(30, 300, seed=40, kmax=5)
        sysd = (cfg, xyz, q, ids, g)
    b = F.Batch(fp32_pairs=False)
    if kinds is None:
        b.add(sysd, seed=5)
    else:
        b.add(sysd, seed=5, moves=kinds)
    b.run([100])  # moves before sampling starts do not count
    b.sync()
    every = 37
    b.sample_energy(every)
    # sample at absolute move indices: align the chunks with the sampling grid
    first = every - (100 % every)
    if form.endswith("split"):
        b.run_split(first, groups=4)
    else:
        b.run([first])
    b.sync()
    host = [b.info(0)["energy"]]
    print("energy samples", form, flush=True)
    n, s1, s2, nh, h1, h2 = sampled_run(b, host, 40, every, split=4 if form.endswith("split") else None)
    log("energy samples", form, n, nh, s1, h1)
    assert n == nh
    assert s1 == h1
    assert abs(s2 - h2) <= 1e-12 * abs(h2)


@pytest.mark.parametrize("a,b_,fp32", [(0, 1, False), (0, 0, False), (1, 1, True)])
def test_device_rdf_equals_host_histogram(a, b_, fp32):
    sysd = systems.primitive_model(n_pairs=100, box=118.0, seed=1)
    cfg, xyz, q, ids, g = sysd
    box = cfg["geometry"]["length"]
    b = F.Batch(fp32_pairs=fp32)
    b.add(sysd, seed=9)
    dr, nbins = 0.5, 100
    total = np.zeros(nbins, np.uint64)
    host = np.zeros(nbins, np.uint64)
    print("rdf", a, b_, fp32, flush=True)
    for _ in range(5):
        b.run([3000])
        b.sync()
        b.rdf(0, a, b_, dr, nbins, total)
        host += host_rdf(b.positions(0), np.asarray(ids), box, a, b_, dr, nbins)
    log("rdf", a, b_, int(total.sum()), int(host.sum()))
    assert total.sum() > 0
    assert np.array_equal(total, host)


def test_rdf_on_a_large_replica_counts_every_pair():
    """N = 20 000 with a cell list: the unlike-type histogram out to beyond the box
    diagonal holds every Na-Cl pair exactly once."""
    cfg, xyz, q, ids, g = systems.nacl_lj(n_pairs=10000, box=255.0, seed=1, kmax=20)
    b = F.Batch()
    b.add((cfg, xyz, q, ids, g), seed=3)
    nbins = 4500  # 0.1 A bins to 450 A > sqrt(3) * 255 / 2
    print("rdf large", flush=True)
    counts = b.rdf(0, 0, 1, 0.1, nbins)
    ids = np.asarray(ids)
    assert int(counts.sum()) == int((ids == 0).sum()) * int((ids == 1).sum())
