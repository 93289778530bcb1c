"""Statistical pin of the restated Metropolis chain (device proposal scheme) on the
reference's own example: the mean energy of examples/minimal agrees with the golden
systemenergy.mean (examples/minimal/minimal.out.json) within the reference's own
jsoncompare tolerance of 2 % (examples/CMakeLists.txt:77-81)."""
import numpy as np

from oracle import oracle as O
from tests.ensemble import block_stats, minimal_system


def test_minimal_mean_energy_matches_reference():
    sysd, golden = minimal_system()
    cfg, xyz, q, ids, g = sysd
    s = O.System(cfg, fast=True)
    s.set_particles(xyz, q, ids)
    s.set_groups(g["begin"], g["size"], g["molid"])
    s.init()
    energies = []
    # the reference samples every 100 steps of 40 moves (minimal.yml)
    for k in range(800):
        s.run_philox_translate(4000, seed=11, replica=0, move0=k * 4000)
        energies.append(s.total_energy()[0])
    mean, sem = block_stats(energies)
    assert abs(mean - golden) <= 0.02 * abs(golden), (mean, sem, golden)
    moves, acc, _ = s.stats()
    assert abs(acc / moves - 0.651) < 0.02  # minimal.out.json: transrot acceptance 0.651
