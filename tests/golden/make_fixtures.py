"""Extract compact fixtures from data files held by the reference's own tests.

Reads (as text/JSON data only) the example state files and golden outputs under
the reference checkout and writes small ``.npz`` fixtures next to this script:

* ``minimal.npz``  examples/minimal/minimal.state.json (40 ions) + golden
  ``systemenergy.init`` from examples/minimal/minimal.out.json:10
* ``bulk.npz``     examples/bulk/bulk.state.json (2304 ions) + golden
  ``systemenergy.init`` from examples/bulk/bulk.out.json:28
* ``minimal.state.json``  examples/minimal/minimal.state.json as data (the input of the
  state-file loader tests)

Usage: python tests/golden/make_fixtures.py /path/to/reference
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def extract(ref, name):
    with open(os.path.join(ref, "examples", name, f"{name}.state.json")) as f:
        state = json.load(f)
    with open(os.path.join(ref, "examples", name, f"{name}.out.json")) as f:
        out = json.load(f)
    init = None
    for a in out["analysis"]:
        if "systemenergy" in a:
            init = a["systemenergy"]["init"]
            mean = a["systemenergy"].get("mean")
    parts = state["particles"]
    xyz = np.array([p["pos"] for p in parts], dtype=np.float64)
    q = np.array([p["q"] for p in parts], dtype=np.float64)
    ids = np.array([p["id"] for p in parts], dtype=np.int32)
    groups = state["groups"]
    sizes = np.array([g["size"] for g in groups], np.int32)
    atomic = np.array([g.get("atomic", False) for g in groups], np.int8)
    length = np.array(state["geometry"]["length"], np.float64)
    np.savez(os.path.join(HERE, f"{name}.npz"), xyz=xyz, q=q, ids=ids, group_sizes=sizes,
             group_atomic=atomic, length=length, init_energy=np.float64(init),
             mean_energy=np.float64(mean if mean is not None else np.nan))
    print(name, len(q), init, mean)


def copy_state(ref, name):
    with open(os.path.join(ref, "examples", name, f"{name}.state.json")) as f:
        state = json.load(f)
    with open(os.path.join(HERE, f"{name}.state.json"), "w") as f:
        json.dump(state, f, indent=1)


if __name__ == "__main__":
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    extract(ref, "minimal")
    extract(ref, "bulk")
    copy_state(ref, "minimal")
