"""Serving loop with continuous batching: a job's result is bitwise the same run
alone, in a mix of other jobs, and under another scheduling quantum; FP64 jobs end
at the oracle's positions for the same Philox stream; admission respects the HBM
budget in first-in first-out order."""
import numpy as np
import pytest

import faunus_amd as F
from oracle import oracle as O
from tests import systems
from tests.gpu_helpers import log

pytestmark = pytest.mark.gpu


def drain(srv, quantum, max_rounds=100000):
    peak = 0
    for _ in range(max_rounds):
        r, q = srv.step(quantum)
        peak = max(peak, r)
        if r == 0 and q == 0:
            return peak
    raise AssertionError("server did not drain")


def small_box(n_pairs):
    return 255.0 * (2 * n_pairs / 20000) ** (1 / 3)


def job_systems():
    return [
        ("primitive", systems.primitive_model(n_pairs=100, box=118.0, seed=1), 3000, 11),
        ("nacl_cells", systems.nacl_lj(n_pairs=800, box=small_box(800), seed=2, kmax=9), 2000, 12),
        ("salt21", systems.salt21(n=300, seed=3, kmax=6), 2500, 13),
        ("primitive_b", systems.primitive_model(n_pairs=60, box=100.0, seed=4), 4000, 14),
        # ~3.7k k-vectors: needs the wider per-thread Q(k) instance of the kernel
        ("salt21_k12", systems.salt21(n=198, seed=5, kmax=12), 1500, 15),
    ]


@pytest.mark.parametrize("fp32", [False, True])
def test_alone_equals_mixed(fp32):
    jobs = job_systems()
    # each job alone, large quantum
    alone = {}
    for name, sysd, moves, seed in jobs:
        s = F.Server(fp32_pairs=fp32)
        j = s.submit(sysd, moves, seed)
        drain(s, quantum=10**6)
        info = s.job(j)
        assert info["state"] == F.Server.DONE and info["moves"] == moves, info
        alone[name] = (info, s.positions(j))
        s.close()
    # all jobs mixed in one server, several copies, small quantum (many rounds)
    s = F.Server(fp32_pairs=fp32)
    ids = {}
    for rep in range(2):
        for name, sysd, moves, seed in jobs[::-1] if rep else jobs:
            ids[(name, rep)] = s.submit(sysd, moves, seed)
    peak = drain(s, quantum=300)
    log("mixed peak resident", peak, "status", s.status())
    assert peak == len(ids)
    for (name, rep), j in ids.items():
        info = s.job(j)
        a_info, a_xyz = alone[name]
        assert info["state"] == F.Server.DONE
        assert info["moves"] == a_info["moves"] and info["accepted"] == a_info["accepted"], (name, info, a_info)
        assert info["energy"] == a_info["energy"], (name, info["energy"], a_info["energy"])
        assert np.array_equal(s.positions(j), a_xyz), name
    s.close()


def test_fp64_job_matches_oracle():
    for name, sysd, moves, seed in job_systems()[:3]:
        cfg, xyz, q, ids, g = sysd
        orc = O.System(cfg)
        orc.set_particles(xyz, q, ids)
        orc.set_groups(g["begin"], g["size"], g["molid"])
        orc.init()
        rec = orc.run_philox_translate(moves, seed=seed, replica=0)
        s = F.Server(fp32_pairs=False)
        j = s.submit(sysd, moves, seed)
        drain(s, quantum=500)
        info = s.job(j)
        log("serve vs oracle", name, info, "oracle accepted", int(rec["accepted"].sum()))
        assert info["state"] == F.Server.DONE
        assert info["accepted"] == int(rec["accepted"].sum())
        np.testing.assert_allclose(s.positions(j), orc.positions(), rtol=0, atol=1e-9)
        s.close()


def test_hbm_budget_admission_fifo():
    jobs = job_systems()
    # footprint of each job alone
    sizes = []
    for _, sysd, moves, seed in jobs:
        probe = F.Server(fp32_pairs=True, hbm_budget_bytes=1 << 40)
        probe.submit(sysd, moves, seed)
        probe.step(1)
        sizes.append(probe.status()["bytes_in_use"])
        probe.close()
    budget = int(max(sizes) * 1.2)
    s = F.Server(fp32_pairs=True, hbm_budget_bytes=budget, max_slots=64)
    ids = [s.submit(sysd, moves, seed) for _, sysd, moves, seed in jobs * 2]
    tiny = F.Server(fp32_pairs=True, hbm_budget_bytes=min(sizes) // 2)
    with pytest.raises(RuntimeError, match="budget"):
        tiny.submit(jobs[0][1], 10, 1)
    tiny.close()
    used_max, peak = 0, 0
    for _ in range(100000):
        r, q = s.step(400)
        peak = max(peak, r)
        used_max = max(used_max, s.status()["bytes_in_use"])
        if r == 0 and q == 0:
            break
    log("budget", budget, "sizes", sizes, "peak resident", peak, "max bytes", used_max)
    assert used_max <= budget and peak < len(ids)
    admits = [s.job(j)["t_admit"] for j in ids]
    assert admits == sorted(admits), admits
    assert all(s.job(j)["state"] == F.Server.DONE for j in ids)
    s.close()


def test_job_too_many_kvectors_fails_cleanly():
    s = F.Server(fp32_pairs=True)
    bad = s.submit(systems.nacl_lj(n_pairs=200, box=60.0, seed=5, kmax=23), 100, 1)
    good = s.submit(systems.primitive_model(n_pairs=40, box=90.0, seed=6), 500, 2)
    drain(s, quantum=10**6)
    assert s.job(bad)["state"] == F.Server.FAILED
    assert s.job(good)["state"] == F.Server.DONE
    s.close()


def chain_jobs():
    out = []
    for r, (nb, nt) in enumerate([(30, 300), (60, 500), (45, 350)]):
        cfg, xyz, q, ids, g, kinds = systems.polyelectrolyte(nb, nt, seed=40 + r, kmax=5)
        out.append(("chain%d" % r, (cfg, xyz, q, ids, g), kinds, 2000 + 500 * r, 60 + r))
    return out


def test_chain_jobs_alone_equal_mixed_and_match_oracle():
    """Polyelectrolyte jobs (ion and bead translations, pivots, bonds) share launches
    with translate-only jobs; each result is bitwise the same as alone, and the FP64
    result ends at the oracle's positions for the same stream."""
    chains = chain_jobs()
    alone = {}
    for name, sysd, kinds, moves, seed in chains:
        s = F.Server(fp32_pairs=False)
        j = s.submit(sysd, moves, seed, kinds=kinds)
        drain(s, quantum=10**6)
        info = s.job(j)
        assert info["state"] == F.Server.DONE and info["moves"] == moves, info
        alone[name] = (info, s.positions(j))
        s.close()
        cfg, xyz, q, ids, g = sysd
        orc = O.System(cfg)
        orc.set_particles(xyz, q, ids)
        orc.set_groups(g["begin"], g["size"], g["molid"])
        orc.init()
        rec = orc.run_philox_moves(moves, seed=seed, kinds=kinds, replica=0)
        log("chain job", name, info["accepted"], "oracle", int(rec["accepted"].sum()))
        assert info["accepted"] == int(rec["accepted"].sum())
        np.testing.assert_allclose(alone[name][1], orc.positions(), rtol=0, atol=1e-9)
    s = F.Server(fp32_pairs=False)
    ids = {}
    atomic = job_systems()[:2]
    for name, sysd, kinds, moves, seed in chains:
        ids[name] = s.submit(sysd, moves, seed, kinds=kinds)
    for name, sysd, moves, seed in atomic:
        s.submit(sysd, moves, seed)
    drain(s, quantum=200)
    for name, j in ids.items():
        info = s.job(j)
        a_info, a_xyz = alone[name]
        assert info["state"] == F.Server.DONE
        assert info["energy"] == a_info["energy"] and info["accepted"] == a_info["accepted"], name
        assert np.array_equal(s.positions(j), a_xyz), name
    s.close()
