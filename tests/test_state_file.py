"""Reference state files (Space JSON, src/space.cpp:449-500, src/group.cpp:231-253):
the C-ABI parser reads examples/minimal/minimal.state.json (a data fixture) into the
same arrays as the independent numpy extraction in tests/golden/minimal.npz, bit for
bit, and rejects the forms the reference rejects or this build does not take (no CPU
compute beyond parsing; the library is loaded, no GPU is touched)."""
import json
import os

import numpy as np
import pytest

import faunus_amd as F

HERE = os.path.dirname(os.path.abspath(__file__))


def state_text():
    with open(os.path.join(HERE, "golden", "minimal.state.json")) as f:
        return f.read()


def test_minimal_state_parses_to_the_fixture_arrays():
    st = F.parse_state(state_text())
    d = np.load(os.path.join(HERE, "golden", "minimal.npz"))
    assert np.array_equal(st["xyz"], d["xyz"])
    assert np.array_equal(st["q"], d["q"])
    assert np.array_equal(st["ids"], d["ids"])
    assert np.array_equal(st["groups"]["size"], d["group_sizes"])
    assert np.array_equal(st["groups"]["begin"], [0])
    assert np.array_equal(st["box"], d["length"])


def test_groups_are_laid_end_to_end():
    s = json.loads(state_text())
    parts = s["particles"]
    s["groups"] = [dict(s["groups"][0], size=15, id=0), dict(s["groups"][0], size=25, id=0, cm=[1.0, 2.0, 3.0])]
    st = F.parse_state(s)
    assert list(st["groups"]["begin"]) == [0, 15] and list(st["groups"]["size"]) == [15, 25]
    assert np.array_equal(st["mass_center"][1], [1.0, 2.0, 3.0])
    assert len(st["q"]) == len(parts)


@pytest.mark.parametrize("edit,msg", [
    (lambda s: s["groups"][0].update(size=39), "cover"),
    (lambda s: s["groups"][0].update(capacity=41), "capacity"),
    (lambda s: s.pop("groups"), "groups"),
    (lambda s: s["geometry"].update(type="sphere"), "cuboid"),
    (lambda s: s["particles"][3].pop("pos"), "pos"),
])
def test_malformed_states_are_rejected(edit, msg):
    s = json.loads(state_text())
    edit(s)
    with pytest.raises(RuntimeError, match=msg):
        F.parse_state(s)
