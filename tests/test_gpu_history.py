"""ScoreCard energy history on the device: the time series that ScoreCard::commit appends to
(src/Potential/scorecard.cpp:366-, hpc_scorecard.cu:170-215), reportHistory / reportAverageStates
/ reportVarianceOfStates (scorecard.cpp:869-1134), and the snapshots a minimization commits every
ntpr cycles (src/MolecularMechanics/hpc_minimization.cu:292-300), checked the way
test/MolecularMechanics/test_hpc_minimization.cu:914-945 checks them (back-to-back runs give the
same history) and against independent quantities: a sample taken at cycle c holds the energies
at the start of that cycle, which are the final energies of a minimization of c cycles."""
import numpy as np
import pytest

import oracle
from stormm_amd import _lib, synthetic
from stormm_amd import synthesis as S

pytestmark = pytest.mark.gpu

BITS = {"single": (32, 23, 26), "double": (32, 40, 32)}


def evaluate_into(sc, ag, ps, prec):
    sc.initialize()
    S.launch_mm_evaluation(prec, ag, None, ps, sc, False, True)


def test_commit_appends_rows_and_grows_the_series():
    topos, coords = synthetic.config1(nsys=6, seed=31)
    ag = S.AtomGraphSynthesis(topos)
    ps = S.PhaseSpaceSynthesis(ag)
    sc = S.ScoreCard(ag.system_count, 26)
    ps.upload(coords)
    assert sc.sample_count() == 0 and sc.sample_capacity() == 16
    rows = []
    for k in range(21):
        # Different energies per sample: move one atom of system 2 a little further each time.
        moved = [np.array(c, dtype=np.float64) for c in coords]
        moved[2][0, 0] += 0.01 * k
        ps.upload(moved)
        evaluate_into(sc, ag, ps, "single")
        sc.commit()
        sc.set_last_time_step(100 + k)
        rows.append(sc.accumulators().copy())
    # The 17th commit found the series full and doubled it, keeping what it held.
    assert sc.sample_count() == 21 and sc.sample_capacity() == 32
    series, steps = sc.series()
    assert np.array_equal(steps, 100 + np.arange(21))
    for k in range(21):
        pe = rows[k][:, :13].sum(axis=1)
        assert np.array_equal(series[k][:, :13], rows[k][:, :13]), k
        assert np.array_equal(series[k][:, 28], pe) and np.array_equal(series[k][:, 29], pe)
    # reportHistory of one system, and over systems.
    h = sc.report_history(2)
    assert np.allclose(h, [r[2, :13].sum() * 2.0 ** -26 for r in rows], rtol=0, atol=1e-12)
    assert len(np.unique(h)) > 10
    mean, std = sc.report_history_over_systems()
    v = np.array([r[:, :13].sum(axis=1) for r in rows]) * 2.0 ** -26
    assert np.allclose(mean, v.mean(axis=1)) and np.allclose(std, v.std(axis=1, ddof=1))
    # Averages and variances from the running sums.
    avg = sc.report_average_states()
    var = sc.report_variance_of_states()
    assert np.allclose(avg[:, 28], v.mean(axis=0), rtol=1e-12)
    # System 2 is the one whose energy moves (the others' variances are rounding noise of the
    # running sums, as in the reference's formula).
    assert np.isclose(var[2, 28], v[:, 2].std(ddof=1), rtol=1e-6)
    # A partial download and a reset.
    part, st = sc.series(5, 3)
    assert np.array_equal(part, series[5:8]) and np.array_equal(st, [105, 106, 107])
    sc.reset_sample_count(4)
    assert sc.sample_count() == 4
    with pytest.raises(_lib.StormmError):
        sc.series(3, 2)
    with pytest.raises(_lib.StormmError):
        sc.reset_sample_count(5)


def minimize(topos, coords, prec, cycles, mode, ntpr, rms=0.0, sc=None, sd=10):
    gpos, frc, nrg = BITS[prec]
    ag = S.AtomGraphSynthesis(topos)
    ps = S.PhaseSpaceSynthesis(ag, gpos, frc)
    sc = sc or S.ScoreCard(ag.system_count, nrg)
    lm = S.LineMinimization(ag.system_count, 0.01)
    ps.upload(coords)
    mc = S.minimize_controls(cycles, sd, 0.01, rms, clash_cycles=0, ntpr=ntpr)
    S.launch_minimization(prec, ag, ps, mc, sc, lm, persistent=mode)
    return ag, sc


@pytest.mark.parametrize("prec", ["single", "double"])
def test_minimization_history_is_the_energy_at_the_start_of_each_sampled_cycle(prec):
    topos, coords = synthetic.config2(nsys=12, seed=33)
    cycles, ntpr = 40, 10
    _, sc = minimize(topos, coords, prec, cycles, 0, ntpr)
    series, steps = sc.series()
    assert series.shape[0] == 4 and np.array_equal(steps, [0, 10, 20, 30])
    # Sample j: the final energies of a minimization of j * ntpr cycles (its last evaluation
    # is at the coordinates cycle j * ntpr starts from).
    for j in range(4):
        _, ref = minimize(topos, coords, prec, j * ntpr, 0, 0)
        acc = ref.accumulators()
        assert np.array_equal(series[j][:, :13], acc[:, :13]), j
        assert np.array_equal(series[j][:, 28], acc[:, :13].sum(axis=1)), j
    # Energies fall along the history.
    pe = series[:, :, 28]
    assert (pe[-1] < pe[0]).all()
    if prec == "double":
        # The oracle's energy trace holds the energy at the start of every cycle (FP64 device
        # trajectories track it to ~1e-10 A over 30 cycles).
        gpos, frc, nrg = BITS[prec]
        for s, t in enumerate(topos):
            fixed = np.rint(np.asarray(coords[s]) * 2.0 ** gpos).astype(np.int64)
            _, trace = oracle.minimize(t, fixed, gpos, frc, nrg, cycles, 10, 0.01, dp_mode=True)
            assert np.abs(pe[:, s] * 2.0 ** -nrg - trace[0:cycles:ntpr]).max() < 1e-6, s
    # A second run appends to the same ScoreCard; back-to-back runs record the same history
    # (test_hpc_minimization.cu:924-945).
    minimize(topos, coords, prec, cycles, 0, ntpr, sc=sc)
    both, steps2 = sc.series()
    assert both.shape[0] == 8 and np.array_equal(both[:4], both[4:])
    assert np.array_equal(steps2, [0, 10, 20, 30] * 2)


@pytest.mark.parametrize("prec", ["single", "double"])
def test_every_minimization_mode_records_the_same_history(prec):
    topos, coords = synthetic.config2(nsys=10, seed=34)
    ref = None
    for mode in (0, 2, 1, 3, 4):
        _, sc = minimize(topos, coords, prec, 25, mode, 4)
        series, steps = sc.series()
        assert np.array_equal(steps, [0, 4, 8, 12, 16, 20, 24])
        if ref is None:
            ref = series
        assert np.array_equal(series, ref), mode
        run, sq, n = sc.history()
        assert n == 7
        v = series.astype(np.float64) * 2.0 ** -BITS[prec][2]
        assert np.allclose(run, v.sum(axis=0), rtol=1e-12, atol=1e-9)


def test_systems_that_stop_early_repeat_their_last_sample():
    topos, coords = synthetic.config2(nsys=16, seed=35)
    cycles, ntpr = 120, 10
    ag, sc = minimize(topos, coords, "single", cycles, 1, ntpr, rms=1.0)
    run, _ = ag.minimization_status()
    assert (run < cycles).any(), "no system stopped early"
    series, _ = sc.series()
    assert series.shape[0] == 12
    for s in range(ag.system_count):
        last = int(run[s]) // ntpr
        for j in range(last + 1, 12):
            assert np.array_equal(series[j][s], series[last][s]), (s, j)
    # The same run in graph-replay mode records the same history.
    _, sc2 = minimize(topos, coords, "single", cycles, 2, ntpr, rms=1.0)
    assert np.array_equal(sc2.series()[0], series)


@pytest.mark.parametrize("mode", [0, 1])
def test_kit_minimization_writes_the_callers_series(mode):
    """The kit form writes the caller's ScoreCardWriter time series from its sampled_step_count
    on, and refuses a series without room for the run's samples."""
    from test_gpu_kits import CallerState, DevArray, initial_values, split_words

    topos, coords = synthetic.config2(nsys=5, seed=36)
    gpos, frc, nrg = BITS["single"]
    cycles, ntpr = 30, 10
    _, ref = minimize(topos, coords, "single", cycles, mode, ntpr, sd=5)
    ag = S.AtomGraphSynthesis(topos)
    lo, hi = split_words(initial_values(topos, coords, gpos, frc), False)
    st = CallerState(ag, gpos, frc, nrg, lo, hi)
    nsys = ag.system_count
    series = DevArray(np.full((5, nsys, 32), -7, dtype=np.int64))
    run = DevArray(np.zeros((nsys, 32)))
    sq = DevArray(np.zeros((nsys, 32)))
    st.scw.running_accumulators = run.ptr.value
    st.scw.squared_accumulators = sq.ptr.value
    st.scw.time_series = series.ptr.value
    st.scw.sample_capacity = 4
    st.scw.sampled_step_count = 2
    mc = S.minimize_controls(cycles, 5, 0.01, 0.0, clash_cycles=0, ntpr=ntpr)
    L = _lib.load()
    import ctypes
    args = (_lib.PREC_SINGLE, ag.handle, ctypes.byref(st.psw), ctypes.byref(mc),
            ctypes.byref(st.scw), ctypes.byref(st.lmw), mode, None)
    with pytest.raises(_lib.StormmError, match="time series holds 4"):
        _lib.check(L.stormm_launch_minimization_kit(*args))
    st.scw.sample_capacity = 5
    _lib.check(L.stormm_launch_minimization_kit(*args))
    got = series.get()
    want, _ = ref.series()
    assert np.array_equal(got[2:], want)
    assert (got[:2] == -7).all()
    ref_run, ref_sq, _ = ref.history()
    assert np.allclose(run.get(), ref_run, rtol=1e-12) and np.allclose(sq.get(), ref_sq,
                                                                        rtol=1e-12)


def test_pool_slot_history_equals_the_batch_history():
    """The serving pool's per-slot history (one sample every ntpr cycles from admission) holds
    the batch minimizer's samples for the same systems, whatever the pool's call pattern."""
    from stormm_amd import serving

    topos, coords = synthetic.config2(nsys=6, seed=37)
    cycles, ntpr = 25, 5
    _, sc = minimize(topos, coords, "single", cycles, 0, ntpr)
    ref, _ = sc.series()
    pool = serving.MinimizationPool(precision="single", total_cycles=cycles, sd_cycles=10,
                                    conv_rms=0.0, clash_cycles=0, ntpr=ntpr)
    other_t, other_c = synthetic.config2(nsys=3, seed=38)
    pool.admit(other_t, other_c)
    pool.minimize(3)
    slots = pool.admit(topos, coords)
    for tick in (4, 1, 7, 13, 9):
        pool.minimize(tick)
    steps, done, _, _ = pool.status(slots)
    assert (steps == cycles).all() and done.all()
    for s, slot in enumerate(slots):
        nrg, rows = pool.history(slot)
        assert rows.shape == (5, 32)
        assert np.array_equal(rows, ref[:, s, :]), s
        assert np.array_equal(nrg, ref[:, s, 28] * 2.0 ** -26)


def test_pool_history_of_systems_stopped_by_the_rms_force():
    from stormm_amd import serving

    topos, coords = synthetic.config2(nsys=8, seed=39)
    cycles, ntpr = 200, 10
    pool = serving.MinimizationPool(precision="single", total_cycles=cycles, sd_cycles=50,
                                    conv_rms=1.0, clash_cycles=0, ntpr=ntpr)
    slots = pool.admit(topos, coords)
    for _ in range(100):
        if pool.minimize(20) == 0:
            break
    steps, done, _, nrg = pool.status(slots)
    assert done.all() and (steps < cycles).any()
    for slot, st_, e in zip(slots, steps, nrg):
        hist, _ = pool.history(slot)
        want = -(-int(st_) // ntpr) + (1 if (st_ % ntpr == 0 and st_ < cycles) else 0)
        assert len(hist) == want, (slot, st_, len(hist))
        assert hist[-1] < hist[0] and e < hist[0]


def test_noalloc_minimization_writes_the_callers_series():
    """The allocation-free entry point records the same history as the kit entry point in mode
    1, including the samples that systems stopped by the RMS force repeat, and refuses a series
    without room."""
    from test_gpu_kits import CallerState, DevArray
    import ctypes

    L = _lib.load()
    gpos, frc, nrg = BITS["single"]
    topos, coords = synthetic.config2(nsys=9, seed=40)
    ag = S.AtomGraphSynthesis(topos)
    fixed = [np.rint(np.asarray(c) * 2.0 ** gpos).astype(np.int64) for c in coords]
    cycles, ntpr = 120, 10
    mc = S.minimize_controls(cycles, 10, 0.01, 1.0, clash_cycles=0, ntpr=ntpr)
    series = {}
    for name in ("kit", "noalloc"):
        st = CallerState(ag, gpos, frc, nrg, fixed)
        rows = DevArray(np.zeros((12, ag.system_count, 32), dtype=np.int64))
        st.scw.time_series = rows.ptr.value
        st.scw.sample_capacity = 11
        args = (_lib.PREC_SINGLE, ag.handle, ctypes.byref(st.psw), ctypes.byref(mc),
                ctypes.byref(st.scw), ctypes.byref(st.lmw))
        call = ((lambda: L.stormm_launch_minimization_kit(*args, 1, None)) if name == "kit"
                else (lambda: L.stormm_launch_minimization_kit_noalloc(*args, None)))
        with pytest.raises(_lib.StormmError, match="time series holds 11"):
            _lib.check(call())
        st.scw.sample_capacity = 12
        _lib.check(call())
        _lib.check(L.stormm_device_synchronize())
        series[name] = rows.get()
    assert np.array_equal(series["kit"], series["noalloc"])
    assert (series["kit"][:, :, 28] != 0).all()
