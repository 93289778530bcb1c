"""Parity of the HIP engine (through the C ABI) with the FP64 oracle on recorded
proposal streams.  Tolerance: |gpu - cpu| <= 1e-10 * max(|cpu|, E_scale), E_scale
= per-move sum of |term energies| (>= 1 kT); accept/reject identical except at
logged near-ties |u - exp(-du)| < 1e-9."""
import numpy as np
import pytest

import faunus_amd as F
from oracle import oracle as O
from tests import systems
from tests.gpu_helpers import build_pair, close, compare_replay, log

pytestmark = pytest.mark.gpu


def test_engine_spline_matches_oracle_spline():
    cfg = systems.primitive_model(n_pairs=10, box=60.0)[0]
    eng = F.Engine(cfg)
    knots, coeff = eng.coulomb_spline()
    cj = cfg["energy"][0]["nonbonded"]["default"][1]["coulomb"]
    k2, c2, _ = O.spline_test(1, coulomb=cj)
    assert len(knots) == len(k2)
    np.testing.assert_allclose(knots, k2, rtol=0, atol=1e-15)
    np.testing.assert_allclose(coeff, c2, rtol=1e-12, atol=1e-14)


def _check_init(orc, eng, e_cpu, e_gpu):
    assert eng.term_names() == orc.term_names()
    _, tc = orc.total_energy()
    _, tg = eng.total_energy()
    log("init terms cpu", tc, "gpu", tg)
    assert close(tg, tc, floor=1.0).all(), (tg, tc)
    assert abs(e_gpu - e_cpu) <= 1e-10 * max(abs(e_cpu), np.abs(tc).sum())


@pytest.mark.parametrize("splined", [True, False])
def test_primitive_model_replay(splined):
    """1:1 primitive model, N=200, L=118 A, HS + Ewald (alpha 0.2, Rc 30, kmax 7)."""
    sysd = systems.primitive_model(n_pairs=100, box=118.0, seed=1, splined=splined)
    orc, eng, e_cpu, e_gpu = build_pair(sysd, F.Engine)
    _check_init(orc, eng, e_cpu, e_gpu)
    log("primitive model K =", eng.num_kvectors)
    assert eng.num_kvectors == 783
    rec = orc.run_translate(3000, seed=1)
    xyz0 = sysd[1]
    out = eng.replay_translate(rec["group"], rec["index"], rec["xyz"], rec["u"])
    rep = compare_replay(rec, out, orc.num_terms)
    log("primitive model replay", rep)
    assert rep["term_fail"] == 0 and rep["du_fail"] == 0
    assert rep["decision_mismatch"] == len(rep["near_ties"]) == 0
    # running Q(k) equals the full recompute; total energies agree after the run
    dq, drec = eng.check_q()
    log("Q drift", dq, drec)
    assert dq < 1e-10 and abs(drec) < 1e-12
    _, tc = orc.total_energy()
    _, tg = eng.total_energy()
    assert close(tg, tc, floor=1.0).all(), (tg, tc)
    np.testing.assert_array_equal(eng.positions(), orc.positions())
    assert not np.array_equal(orc.positions(), xyz0)


def test_nacl_lj_cell_list_replay_small():
    """Na/Cl LJ (cut 12 A) + Ewald (Rc 12), N=2000 at ~1 M: exercises the cell list."""
    box = 255.0 * (2000 / 20000) ** (1 / 3)
    sysd = systems.nacl_lj(n_pairs=1000, box=box, seed=2, kmax=10)
    orc, eng, e_cpu, e_gpu = build_pair(sysd, F.Engine)
    _check_init(orc, eng, e_cpu, e_gpu)
    rec = orc.run_translate(3000, seed=2)
    out = eng.replay_translate(rec["group"], rec["index"], rec["xyz"], rec["u"])
    rep = compare_replay(rec, out, orc.num_terms)
    log("nacl small replay", rep)
    assert rep["term_fail"] == 0 and rep["du_fail"] == 0
    assert rep["decision_mismatch"] == 0
    dq, drec = eng.check_q()
    assert dq < 1e-10 and abs(drec) < 1e-12


def test_nacl_lj_full_size_replay():
    """The N=20 000, L=255 A, kmax=20 (K=17 328) system: per-move parity on 300 moves."""
    sysd = systems.nacl_lj(n_pairs=10000, box=255.0, seed=1)
    orc, eng, e_cpu, e_gpu = build_pair(sysd, F.Engine)
    assert eng.num_kvectors == 17328
    _check_init(orc, eng, e_cpu, e_gpu)
    rec = orc.run_translate(300, seed=1)
    out = eng.replay_translate(rec["group"], rec["index"], rec["xyz"], rec["u"])
    rep = compare_replay(rec, out, orc.num_terms)
    log("nacl full replay", rep)
    assert rep["term_fail"] == 0 and rep["du_fail"] == 0
    assert rep["decision_mismatch"] == 0
    assert rep["accepted"] > 0


def test_salt21_replay():
    """2:1 salt replica (q=+2/-1, HS + Ewald alpha=5/L, Rc=L/2, kmax 8, K=1152)."""
    sysd = systems.salt21(n=900, conc=0.05, seed=4)
    orc, eng, e_cpu, e_gpu = build_pair(sysd, F.Engine)
    assert eng.num_kvectors == 1152
    _check_init(orc, eng, e_cpu, e_gpu)
    rec = orc.run_translate(2000, seed=4)
    out = eng.replay_translate(rec["group"], rec["index"], rec["xyz"], rec["u"])
    rep = compare_replay(rec, out, orc.num_terms)
    log("salt21 replay", rep)
    assert rep["term_fail"] == 0 and rep["du_fail"] == 0 and rep["decision_mismatch"] == 0


def test_energy_of_each_state_and_reject_restores():
    sysd = systems.primitive_model(n_pairs=40, box=60.0, seed=9, cutoff=25.0, kmax=6)
    orc, eng, _, _ = build_pair(sysd, F.Engine)
    xyz = sysd[1]
    rng = np.random.default_rng(3)
    for _ in range(20):
        i = int(rng.integers(0, len(xyz)))
        new = xyz[i] + rng.normal(scale=2.0, size=3)
        new -= 60.0 * np.round(new / 60.0)
        ch = {"groups": [{"group_index": 0, "relative_atom_indices": [i], "internal": True}]}
        en_c, eo_c = orc.energy_change(ch, [i], [new])
        eng.stage_trial([i], [new])
        tot_t, en_g = eng.energy(ch, F.STATE_TRIAL)
        eng.sync(False)
        eng.stage_trial([i], [new])
        tot_a, eo_g = eng.energy(ch, F.STATE_ACCEPTED)
        eng.sync(False)
        assert close(en_g, en_c).all() and close(eo_g, eo_c).all()
    # rejected moves leave the accepted state untouched
    np.testing.assert_array_equal(eng.positions(), xyz)
    _, tc = orc.total_energy()
    _, tg = eng.total_energy()
    assert close(tg, tc).all()


def test_multi_particle_change_in_one_group():
    """A Change with several relative_atom_indices in one atomic group (internal=true):
    moved x static pairs plus moved x moved pairs (energy.h:931-961)."""
    sysd = systems.primitive_model(n_pairs=30, box=60.0, seed=5, cutoff=25.0, kmax=5)
    orc, eng, _, _ = build_pair(sysd, F.Engine)
    xyz = sysd[1]
    rng = np.random.default_rng(1)
    for m in (2, 3, 5, 9):
        idx = np.sort(rng.choice(len(xyz), size=m, replace=False))
        new = xyz[idx] + rng.normal(scale=1.5, size=(m, 3))
        new -= 60.0 * np.round(new / 60.0)
        ch = {"groups": [{"group_index": 0, "relative_atom_indices": [int(i) for i in idx], "internal": True}]}
        en_c, eo_c = orc.energy_change(ch, idx, new)
        eng.stage_trial(idx, new)
        en_g, eo_g = eng.delta_u(ch)
        eng.sync(False)
        log("multi", m, en_c, en_g)
        assert close(en_g, en_c).all() and close(eo_g, eo_c).all()


def test_unsupported_input_fails_loudly():
    bad = systems.primitive_model(n_pairs=5, box=40.0)[0]
    bad = dict(bad)
    bad["energy"] = [{"nonbonded": {"default": [{"hertz": {}}]}}]
    with pytest.raises(RuntimeError):
        F.Engine(bad)


def test_groups_past_the_last_particle_are_rejected():
    """A group list that claims more particles than were set is an error raised
    before any per-particle table is written (host-side bounds check)."""
    cfg, xyz, q, ids, g = systems.primitive_model(n_pairs=10, box=60.0, seed=3)
    e = F.Engine(cfg)
    e.set_particles(xyz, q, ids)
    with pytest.raises(RuntimeError, match="past the last particle"):
        e.set_groups([0], [len(q) + 2], [0])
    with pytest.raises(RuntimeError, match="past the last particle"):
        e.set_groups([0, 10], [10, len(q)], [0, 0])
    e.set_groups([0], [len(q)], [0])
    assert np.isfinite(e.init())
