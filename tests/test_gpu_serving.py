"""Serving pool: continuous-batching minimization with resident state.

The pool advances every resident system by the reference's per-cycle launch sequence
(src/MolecularMechanics/hpc_minimization.cu:262-405), so a system's trajectory must not depend on
what else is resident or on how its cycles are split across calls: results are compared bit for
bit with the system minimized alone and with the batch per-step minimizer.
"""
import numpy as np
import pytest

from stormm_amd import serving, synthetic
from stormm_amd import synthesis as S

pytestmark = pytest.mark.gpu


def conformers(natom, seed, count, perturb=0.3):
    topo, x = synthetic.branched_molecule(natom, seed=seed, gb=True)
    rng = np.random.default_rng(seed)
    return topo, [x + rng.normal(0.0, perturb, x.shape) for _ in range(count)]


def run_until_done(pool, slots, tick):
    for _ in range(10000):
        steps, done, rms, nrg = pool.status(slots)
        if done.all():
            return steps, rms, nrg
        pool.minimize(tick)
    raise AssertionError("pool did not finish")


@pytest.mark.parametrize("prec", ["single", "double"])
def test_trajectory_is_the_same_alone_and_in_a_changing_mix(prec):
    topo, confs = conformers(60, 11, 3)
    kw = dict(precision=prec, total_cycles=40, sd_cycles=10, conv_rms=0.0)
    alone = serving.MinimizationPool(**kw)
    a_slots = alone.admit_conformers(topo, confs)
    a_steps, _, a_nrg = run_until_done(alone, a_slots, 7)
    a_x = [alone.coordinates(s, topo.natom, fixed=True) for s in a_slots]
    assert (a_steps == 40).all()

    mix = serving.MinimizationPool(**kw)
    o1_t, o1_c = conformers(120, 12, 5)
    o2_t, o2_c = conformers(33, 13, 9)
    o1 = mix.admit_conformers(o1_t, o1_c)
    mix.minimize(3)
    slots = mix.admit_conformers(topo, confs)
    mix.minimize(5)
    o2 = mix.admit_conformers(o2_t, o2_c)
    mix.minimize(11)
    mix.release(o1)
    mix.minimize(2)
    o3_t, o3_c = conformers(200, 14, 2)
    mix.admit_conformers(o3_t, o3_c)
    m_steps, _, m_nrg = run_until_done(mix, slots, 13)
    assert np.array_equal(m_steps, a_steps)
    assert np.array_equal(m_nrg, a_nrg)
    for s, ax in zip(slots, a_x):
        assert np.array_equal(mix.coordinates(s, topo.natom, fixed=True), ax)
    del o2


@pytest.mark.parametrize("prec", ["single", "double"])
def test_pool_matches_the_batch_per_step_minimizer(prec):
    topos, coords = synthetic.config2(nsys=6, seed=21)
    cycles = 25
    bits = (32, 23, 26) if prec == "single" else (32, 40, 32)
    ag = S.AtomGraphSynthesis(topos)
    ps = S.PhaseSpaceSynthesis(ag, bits[0], bits[1])
    sc = S.ScoreCard(ag.system_count, bits[2])
    lm = S.LineMinimization(ag.system_count, 0.01)
    ps.upload(coords)
    S.launch_minimization(prec, ag, ps, S.minimize_controls(cycles, 10, 0.01, 0.0), sc, lm,
                          persistent=False)
    xf, _ = ps.fixed()
    pool = serving.MinimizationPool(precision=prec, total_cycles=cycles, sd_cycles=10,
                                    conv_rms=0.0)
    slots = pool.admit(topos, coords)
    steps, _, nrg = run_until_done(pool, slots, 4)
    assert (steps == cycles).all()
    for s, t, ref in zip(slots, topos, xf):
        assert np.array_equal(pool.coordinates(s, t.natom, fixed=True), ref)
    assert np.array_equal(nrg, sc.potential_energies())


def test_budget_admission_and_release():
    topo, confs = conformers(100, 31, 16)
    # Room for roughly 40 systems of 100 atoms (128 padded) plus templates.
    budget = int((40 * 128 * 172 + 64 * 400) / 0.8) + 1
    pool = serving.MinimizationPool(budget_bytes=budget, max_systems=64, total_cycles=5)
    cap0 = pool.capacity()
    assert cap0["capacity_atoms"] >= 40 * 128
    a = pool.admit_conformers(topo, confs)
    b = pool.admit_conformers(topo, confs)
    assert a is not None and b is not None
    assert pool.admit_conformers(topo, confs) is None  # 48 x 128 atoms exceed the budget
    pool.minimize(2)
    pool.release(a)
    c = pool.admit_conformers(topo, confs)
    assert c is not None
    pool.minimize(2)
    pool.release(b)
    pool.release(c)
    cap = pool.capacity()
    assert cap["free_atoms"] == cap0["free_atoms"] and cap["free_slots"] == cap0["free_slots"]
    with pytest.raises(Exception):
        pool.release(c)


def test_systems_stop_at_the_rms_threshold_or_the_cycle_limit():
    topo, confs = conformers(40, 41, 6, perturb=0.05)
    pool = serving.MinimizationPool(total_cycles=400, sd_cycles=50, conv_rms=1.0)
    slots = pool.admit_conformers(topo, confs)
    steps, rms, nrg = run_until_done(pool, slots, 25)
    for st_, r in zip(steps, rms):
        assert st_ == 400 or r <= 1.0
    assert (steps < 400).any()
    # Stopped systems stay put.
    x0 = [pool.coordinates(s, topo.natom, fixed=True) for s in slots]
    pool.minimize(10)
    for s, x in zip(slots, x0):
        assert np.array_equal(pool.coordinates(s, topo.natom, fixed=True), x)


def test_serving_loop_completes_every_request_in_both_modes():
    reqs = serving.serving_workload(seconds=0.05, rate=200, molecules=8, max_atoms=96,
                                    max_conformers=8)
    assert len(reqs) > 0
    for mode in ("continuous", "static"):
        pool = serving.MinimizationPool(total_cycles=30, conv_rms=0.1)
        res = serving.serve(pool, [serving.Request(r.rid, r.arrival, r.topo, r.coords)
                                   for r in reqs], mode=mode, tick_cycles=5, realtime=False)
        assert res["requests"] == len(reqs)
        assert res["latency_p99_s"] >= res["latency_p50_s"] > 0.0


def test_served_results_equal_plain_batched_launches():
    reqs = serving.serving_workload(seconds=0.03, rate=200, molecules=6, max_atoms=80,
                                    max_conformers=6)
    served = [serving.Request(r.rid, r.arrival, r.topo, r.coords) for r in reqs]
    plain = [serving.Request(r.rid, r.arrival, r.topo, r.coords) for r in reqs]
    pool = serving.MinimizationPool(total_cycles=40, sd_cycles=10, conv_rms=0.5)
    serving.serve(pool, served, "continuous", tick_cycles=7, realtime=False)
    serving.serve_per_request(plain, total_cycles=40, sd_cycles=10, conv_rms=0.5,
                              realtime=False)
    for a, b in zip(sorted(served, key=lambda r: r.rid), sorted(plain, key=lambda r: r.rid)):
        assert np.array_equal(a.steps, b.steps)
        assert np.array_equal(a.energies, b.energies)


def test_template_budget_rejection_rolls_back_and_unused_templates_are_dropped():
    """A molecule whose work template does not fit next to a resident one is refused with nothing
    kept (the resident topology is still admitted afterwards); once the resident systems leave,
    their unused template is dropped and the new molecule fits.  Its trajectory is the same as in
    a pool of its own (the dropped template's arrays were packed away underneath it)."""
    kw = dict(precision="double", total_cycles=30, sd_cycles=5, conv_rms=0.0, max_systems=4)
    a_topo, a_confs = conformers(150, 21, 3)
    b_topo, b_confs = conformers(170, 22, 2)
    probe_a = serving.MinimizationPool(**kw)
    probe_a.admit_conformers(a_topo, a_confs[:1])
    t_a = probe_a.capacity()["template_bytes"]
    probe_b = serving.MinimizationPool(**kw)
    b_alone = probe_b.admit_conformers(b_topo, b_confs)
    t_b = probe_b.capacity()["template_bytes"]
    assert t_a > 0 and t_b > 0
    # Template share of the budget (one fifth): room for either template, not for both.
    budget = int(1.5 * max(t_a, t_b) / 0.2)
    assert 1.5 * max(t_a, t_b) < t_a + t_b
    pool = serving.MinimizationPool(budget_bytes=budget, **kw)
    a_slots = pool.admit_conformers(a_topo, a_confs[:2])
    assert a_slots is not None
    before = pool.capacity()
    assert before["template_bytes"] == t_a
    assert pool.admit_conformers(b_topo, b_confs) is None
    assert pool.capacity() == before
    more = pool.admit_conformers(a_topo, a_confs[2:])
    assert more is not None and pool.capacity()["template_bytes"] == t_a
    pool.minimize(3)
    pool.release(a_slots + more)
    b_slots = pool.admit_conformers(b_topo, b_confs)
    assert b_slots is not None
    assert pool.capacity()["template_bytes"] == t_b
    run_until_done(pool, b_slots, 7)
    run_until_done(probe_b, b_alone, 7)
    assert np.array_equal(pool.status(b_slots)[3], probe_b.status(b_alone)[3])
    for s, r in zip(b_slots, b_alone):
        assert np.array_equal(pool.coordinates(s, b_topo.natom, fixed=True),
                              probe_b.coordinates(r, b_topo.natom, fixed=True))
