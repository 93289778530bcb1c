"""fmc_batch_run_split_static -- the split kernel with its exchange slots in a static
per-GPU workspace, for callers that cannot allocate -- against fmc_batch_run_split on
identical batches: the same chain (positions bit for bit, accepted count, running
energy), on two calls in a row, the second with a smaller system and fewer moves, so
that the workspace left by the first call must be cleared for the second to run."""
import ctypes as C

import numpy as np
import pytest

import faunus_amd as F
from tests import systems
from tests.gpu_helpers import log

pytestmark = pytest.mark.gpu


def static_run(batch, moves, groups, stride=1):
    rc = F.lib().fmc_batch_run_split_static(batch.h, C.c_longlong(moves), C.c_int(groups), C.c_int(stride))
    if rc != 0:
        raise RuntimeError(F.lib().fmc_last_error().decode())


def compare(sysd, moves, groups, seed):
    ref = F.Batch()
    ref.add(sysd, seed=seed)
    ref.run_split(moves, groups=groups)
    ref.sync()
    new = F.Batch()
    new.add(sysd, seed=seed)
    new.info(0)  # the batch takes its device state before the call that allocates nothing
    static_run(new, moves, groups)
    new.sync()
    a, b = ref.info(0), new.info(0)
    log("split", a, "static", b)
    assert b["error"] == 0 and new.split_error() == 0
    assert b["moves"] == a["moves"] == moves and b["accepted"] == a["accepted"]
    assert b["energy"] == a["energy"]
    assert np.array_equal(new.positions(0), ref.positions(0))
    ref.close()
    new.close()


def test_static_workspace_matches_run_split_on_two_calls():
    large = systems.nacl_lj(n_pairs=1000, box=255.0 * (2000 / 20000) ** (1 / 3), seed=2, kmax=10)
    small = systems.nacl_lj(n_pairs=300, box=40.0, seed=3, kmax=6)
    compare(large, 1500, 8, seed=91)
    compare(small, 700, 4, seed=92)


@pytest.mark.parametrize("groups", [3, 8])
def test_static_workspace_matches_run_split_when_particles_repeat(groups):
    """N = 24: the move after next often moves the particle of the pending move, so the
    particle message of the call that allocates nothing is corrected by that move's
    decision about once in 24 moves."""
    tiny = systems.nacl_lj(n_pairs=12, box=40.0, seed=5, kmax=6)
    compare(tiny, 3000, groups, seed=94)


def test_static_workspace_refuses_what_it_cannot_hold():
    sysd = systems.nacl_lj(n_pairs=300, box=40.0, seed=3, kmax=6)
    b = F.Batch()
    b.add(sysd, seed=1)
    with pytest.raises(RuntimeError, match="device state"):
        static_run(b, 10, 4)
    b.info(0)
    with pytest.raises(RuntimeError, match="at most 64 workgroups"):
        static_run(b, 10, 65)
    b.close()
