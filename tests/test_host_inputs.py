"""Host-side input validation that needs no GPU: particle and group arrays of
inconsistent lengths are rejected before they reach the C ABI, and the test
systems refuse sizes they cannot build consistently."""
import numpy as np
import pytest

import faunus_amd as F
from tests import systems


def test_particle_arrays_must_agree():
    xyz = np.zeros((10, 3))
    with pytest.raises(ValueError):
        F._particles(xyz, np.zeros(9), np.zeros(10, np.int32))
    with pytest.raises(ValueError):
        F._particles(xyz, np.zeros(10), np.zeros(11, np.int32))
    a, b, c = F._particles(xyz, np.zeros(10), np.zeros(10, np.int32))
    assert a.flags.c_contiguous and len(b) == 10 and c.dtype == np.int32


def test_group_arrays_must_agree():
    with pytest.raises(ValueError):
        F._groups([0, 5], [5], [0, 0])
    with pytest.raises(ValueError):
        F._groups([0], [5], [0], mass_center=np.zeros(4))
    gb, gs, gm, cm = F._groups([0, 5], [5, 5], [0, 0])
    assert gb.dtype == np.int32 and cm is None


def test_salt21_sizes_are_consistent():
    with pytest.raises(ValueError):
        systems.salt21(n=200)
    cfg, xyz, q, ids, g = systems.salt21(n=198, seed=5, kmax=6)
    assert len(xyz) == len(q) == len(ids) == sum(g["size"]) == 198
