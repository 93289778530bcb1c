"""The replica-exchange interval captured as a HIP graph and the tile autotuner.

- Replaying the captured interval (persistent launch + one round of ladder swaps with
  the round taken from a device counter) equals the eager sequence of the same launch
  and host-driven swap rounds, replay by replay, with the program's own eager work and
  synchronisation between the replays: positions bit for bit, temperature labels,
  running energies and swap counts (ParallelTempering, move.cpp:844-965).
- Free-running ladders (one cooperative launch, each ladder swapping at its own barrier
  after each of its own intervals) equal the captured interval replayed as often, and
  with unequal interval counts each ladder equals the eager sequence cut at its own
  count: a ladder's chain depends on nothing but its rungs.
- The autotuner times the candidate launch shapes, keeps one, and leaves every chain
  where it was: tuning and then running gives the chain of running without tuning."""
import numpy as np
import pytest

import faunus_amd as F
from tests import systems
from tests.gpu_helpers import log

pytestmark = pytest.mark.gpu

TREF = 298.15


def ladder_batch(fp32=False):
    b = F.Batch(fp32_pairs=fp32)
    temps = (280.0, 300.0, 330.0, 360.0)
    for c, (n, conc) in enumerate([(300, 0.1), (450, 0.3)]):
        sysd = systems.salt21(n=n, conc=conc, seed=20 + c, kmax=6)
        for k, T in enumerate(temps):
            b.add(sysd, beta=TREF / T, seed=300 + 10 * c + k, ladder=c)
    return b


@pytest.mark.parametrize("fp32", [False, True])
def test_captured_interval_equals_eager_rounds(fp32):
    g, e = ladder_batch(fp32), ladder_batch(fp32)
    moves = np.array([3 * 300] * 4 + [3 * 450] * 4, np.int64)  # three sweeps per rung
    g.capture_interval(moves, seed=77, round0=0)
    for rnd in range(6):
        print("replay", rnd, flush=True)
        g.replay(1)
        g.sync()
        e.run(moves)
        e.swap(seed=77, rnd=rnd)
        e.sync()
        for r in range(8):
            a, b = g.info(r), e.info(r)
            assert a["beta"] == b["beta"] and a["moves"] == b["moves"] and a["accepted"] == b["accepted"], (rnd, r, a, b)
            assert a["energy"] == b["energy"], (rnd, r)
            assert np.array_equal(g.positions(r), e.positions(r))
    ga, gb = g.swap_counts()
    ea, eb = e.swap_counts()
    log("swaps graph", (ga, gb), "eager", (ea, eb))
    assert (ga, gb) == (ea, eb) and ga == 6 * 3
    assert sorted(g.info(r)["beta"] for r in range(4)) == sorted(TREF / T for T in (280.0, 300.0, 330.0, 360.0))
    g.close()
    e.close()


def same_replica(a, b, ra, rb=None):
    rb = ra if rb is None else rb
    x, y = a.info(ra), b.info(rb)
    assert x["beta"] == y["beta"] and x["moves"] == y["moves"] and x["accepted"] == y["accepted"], (ra, x, y)
    assert x["energy"] == y["energy"] and x["error"] == y["error"] == 0, (ra, x, y)
    assert np.array_equal(a.positions(ra), b.positions(rb)), ra


@pytest.mark.parametrize("fp32", [False, True])
def test_free_running_ladders_equal_captured_replays(fp32):
    f, g = ladder_batch(fp32), ladder_batch(fp32)
    moves = np.array([3 * 300] * 4 + [3 * 450] * 4, np.int64)
    g.capture_interval(moves, seed=91, round0=5)
    g.replay(6)
    g.sync()
    f.run_ladders(moves, 6, seed=91, round0=5)
    for r in range(8):
        same_replica(f, g, r)
    fa, fb = f.swap_counts()
    ga, gb = g.swap_counts()
    log("swaps free-running", (fa, fb), "graph", (ga, gb))
    assert (fa, fb) == (ga, gb) and fa == 6 * 3
    f.close()
    g.close()


@pytest.mark.parametrize("fp32", [False, True])
def test_free_running_ladders_with_unequal_intervals(fp32):
    """Ladder 0 runs 8 intervals, ladder 1 three, a replica outside any ladder five; the
    eager batch runs every replica 8 rounds and is read at each one's own count."""
    def batch():
        b = ladder_batch(fp32)
        b.add(systems.salt21(n=600, conc=0.2, seed=29, kmax=6), beta=1.0, seed=399)
        return b

    f, e = batch(), batch()
    moves = np.array([3 * 300] * 4 + [3 * 450] * 4 + [2 * 600], np.int64)
    intervals = np.array([8] * 4 + [3] * 4 + [5], np.int32)
    f.run_ladders(moves, intervals, seed=13, round0=0)
    seen = {}
    for rnd in range(8):
        e.run(moves)
        e.swap(seed=13, rnd=rnd)
        e.sync()
        for r in range(9):
            if intervals[r] == rnd + 1:
                seen[r] = (e.info(r), e.positions(r))
    for r in range(9):
        x, (y, pos) = f.info(r), seen[r]
        assert x["beta"] == y["beta"] and x["moves"] == y["moves"] == intervals[r] * moves[r], (r, x, y)
        assert x["accepted"] == y["accepted"] and x["energy"] == y["energy"] and x["error"] == 0, (r, x, y)
        assert np.array_equal(f.positions(r), pos), r
    att, acc = f.swap_counts()
    log("swaps", att, acc)
    assert att == 4 * 2 + 4 * 1 + 2 * 2 + 1 and 0 <= acc <= att
    # the rungs of one ladder must share the interval count
    with pytest.raises(RuntimeError, match="interval count"):
        f.run_ladders(moves, [1, 2, 1, 1, 1, 1, 1, 1, 1], seed=13)
    f.close()
    e.close()


def test_autotune_keeps_the_chains():
    def batch():
        b = F.Batch()
        for k, (n, conc) in enumerate([(600, 0.1), (900, 0.5), (300, 0.05)]):
            b.add(systems.salt21(n=n, conc=conc, seed=40 + k, kmax=8), seed=500 + k)
        return b

    tuned, plain = batch(), batch()
    tuned.run([200, 200, 200])
    plain.run([200, 200, 200])
    chosen, cands = tuned.autotune(300)
    log("autotune", chosen, cands)
    assert set(cands) == {256, 512} and chosen in cands
    assert chosen == min(cands, key=cands.get)
    for r in range(3):
        assert tuned.info(r)["moves"] == 200
    tuned.run([1500] * 3)
    plain.run([1500] * 3)
    for r in range(3):
        a, b = tuned.info(r), plain.info(r)
        assert a["moves"] == b["moves"] == 1700 and a["accepted"] == b["accepted"], (a, b)
        assert np.array_equal(tuned.positions(r), plain.positions(r))
        assert abs(a["energy"] - b["energy"]) <= 1e-9 * abs(b["energy"])


def test_autotune_picks_the_cell_grid_of_move_set_replicas():
    cfg, xyz, q, ids, g, kinds = systems.water(n_mol=1000, seed=3, kmax=6, cutoff=9.0)
    b = F.Batch()
    b.add((cfg, xyz, q, ids, g), seed=7, moves=kinds)
    chosen, cands = b.autotune(200)
    log("autotune water", chosen, cands)
    assert set(cands) == {1, 2} and chosen == min(cands, key=cands.get)
    b.run([500])
    info = b.info(0)
    assert info["moves"] == 500 and info["error"] == 0
    tot, _ = b.total_energy(0)
    assert abs(info["energy"] - tot) <= 1e-9 * abs(tot)
