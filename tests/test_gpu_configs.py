"""The benchmark workloads at their full sizes on the device.

* 1024 heterogeneous systems, N ~ U[24, 256], OBC-II GB: every system against the FP64 oracle in
  both precision models (bars of test_gpu_eval.py: FP32 energies within 1e-6 of the energy scale
  S = sum_k |E_k| and forces within 1e-5 relative RMS; FP64 within 1e-8 and 1e-8).
* 16384 systems, N ~ U[32, 96]: both precision models over the whole batch, bit-identical
  reruns, finite results, and 64 systems spread over the batch against the oracle.
* 4096 systems, N ~ U[16, 128], 500 cycles (50 steepest descent, 25 clash-damped) in FP32: the
  whole batch in one graph-captured launch equals one launch per size class bit for bit, and
  every system's energy falls; 64 systems in FP64 over 60 cycles (conjugate gradient from cycle
  50) track the oracle's minimizer to 2e-5 A and 1e-5 of the energy.  Conjugate gradient
  amplifies the last-bit differences of the device's double-precision forces about tenfold per
  ten cycles (tests/test_gpu_minimize_reference.py), so after ten conjugate-gradient cycles the
  measured worst cases are 2.3e-6 A and 8.4e-7 of the energy; a wrong formula or constant shows
  at 1e-3 or more.
* The served stream (N ~ logU[16, 512], 1-64 conformers per request, 300 cycles or RMS force
  <= 0.1): every request of a trace completes, and requests served in the mix end exactly where
  they end when served alone.
"""
import numpy as np
import pytest

import oracle
from stormm_amd import serving, synthetic
from stormm_amd import synthesis as S

pytestmark = pytest.mark.gpu

BITS = {"single": (32, 23, 26), "double": (32, 40, 32)}
BARS = {"single": (1e-6, 1e-5), "double": (1e-8, 1e-8)}


def evaluate(topos, coords, prec, shape=None):
    gpos, frc, nrg = BITS[prec]
    ag = S.AtomGraphSynthesis(topos, shape)
    ps = S.PhaseSpaceSynthesis(ag, gpos, frc)
    sc = S.ScoreCard(ag.system_count, nrg)
    ps.upload(coords)
    ps.initialize_forces()
    sc.initialize()
    S.launch_mm_evaluation(prec, ag, None, ps, sc, True, True)
    assert ag.status() == 0
    xf, ff = ps.fixed()
    return dict(E=sc.accumulators()[:, :13].copy(), X=ps.coordinates(), F=ps.forces(), Ffix=ff)


def worst_errors(topos, res, prec, pick):
    nrg_bits = BITS[prec][2]
    e_bar, f_bar = BARS[prec]
    worst_e, worst_f = 0.0, 0.0
    for s in pick:
        eo_d, _, fo = oracle.eval_mm(topos[s], res["X"][s], True, nrg_bits)
        eo = np.array([eo_d[k] for k in oracle.SLOT_NAMES])
        eg = res["E"][s].astype(np.float64) * 2.0 ** -nrg_bits
        scale = np.abs(eo).sum()
        e_err = max(np.abs(eg - eo).max(), abs(eg.sum() - eo.sum())) / scale
        f_err = float(np.sqrt(((res["F"][s] - fo) ** 2).sum() / (fo ** 2).sum()))
        assert e_err <= e_bar and f_err <= f_bar, (s, topos[s].natom, e_err, f_err)
        worst_e, worst_f = max(worst_e, e_err), max(worst_f, f_err)
    return worst_e, worst_f


def spread(natoms, count):
    order = np.argsort(np.asarray(natoms), kind="stable")
    return [int(order[i]) for i in np.linspace(0, len(order) - 1, count).round().astype(int)]


@pytest.mark.parametrize("prec", ["single", "double"])
def test_whole_1024_system_batch_matches_the_oracle(prec):
    topos, coords = synthetic.config1(nsys=1024, seed=1)
    res = evaluate(topos, coords, prec)
    we, wf = worst_errors(topos, res, prec, range(len(topos)))
    print(f"[{prec}] 1024 systems: worst energy error / S {we:.2e}, worst force rel. RMS {wf:.2e}")


@pytest.mark.parametrize("prec", ["single", "double"])
def test_16384_system_batch(prec):
    topos, coords = synthetic.config4(nsys=16384, seed=4)
    a = evaluate(topos, coords, prec)
    b = evaluate(topos, coords, prec)
    assert np.array_equal(a["E"], b["E"])
    assert all(np.array_equal(x, y) for x, y in zip(a["Ffix"], b["Ffix"]))
    assert np.isfinite(a["E"].astype(np.float64)).all()
    assert all(np.isfinite(f).all() for f in a["F"])
    we, wf = worst_errors(topos, a, prec, spread([t.natom for t in topos], 64))
    print(f"[{prec}] 16384 systems, 64 checked: worst energy error / S {we:.2e}, "
          f"worst force rel. RMS {wf:.2e}")


def minimize(topos, coords, prec, cycles, sd, mode=1, clash=25):
    gpos, frc, nrg = BITS[prec]
    ag = S.AtomGraphSynthesis(topos)
    ps = S.PhaseSpaceSynthesis(ag, gpos, frc)
    sc = S.ScoreCard(ag.system_count, nrg)
    lm = S.LineMinimization(ag.system_count, 0.01)
    ps.upload(coords)
    mc = S.minimize_controls(cycles, sd, 0.01, 0.0, clash_cycles=clash)
    S.launch_minimization(prec, ag, ps, mc, sc, lm, persistent=mode)
    xf, _ = ps.fixed()
    return xf, sc.potential_energies()


def start_energies(topos, coords, prec):
    gpos, frc, nrg = BITS[prec]
    ag = S.AtomGraphSynthesis(topos)
    ps = S.PhaseSpaceSynthesis(ag, gpos, frc)
    sc = S.ScoreCard(ag.system_count, nrg)
    ps.upload(coords)
    sc.initialize()
    S.launch_mm_evaluation(prec, ag, None, ps, sc, False, True)
    return sc.potential_energies()


def test_4096_system_minimization_whole_batch_equals_size_classes():
    topos, coords = synthetic.config2(nsys=4096, seed=2)
    e0 = start_energies(topos, coords, "single")
    xw, ew = minimize(topos, coords, "single", 500, 50)
    assert np.isfinite(ew).all()
    assert (ew < e0).all(), int((ew >= e0).sum())
    edges = [16, 44, 72, 100, 129]
    for lo, hi in zip(edges[:-1], edges[1:]):
        idx = [i for i, t in enumerate(topos) if lo <= t.natom < hi]
        xc, ec = minimize([topos[i] for i in idx], [coords[i] for i in idx], "single", 500, 50)
        assert np.array_equal(ec, ew[idx]), (lo, hi)
        assert all(np.array_equal(xc[k], xw[i]) for k, i in enumerate(idx)), (lo, hi)
    print(f"4096 systems, 500 cycles: mean energy {e0.mean():.2f} -> {ew.mean():.2f} kcal/mol")


def test_4096_system_minimization_sample_tracks_the_oracle_in_double_precision():
    topos, coords = synthetic.config2(nsys=4096, seed=2)
    pick = spread([t.natom for t in topos], 64)
    sub_t, sub_c = [topos[i] for i in pick], [coords[i] for i in pick]
    cycles, sd, clash = 60, 50, 25
    gpos, frc, nrg = BITS["double"]
    xf, e = minimize(sub_t, sub_c, "double", cycles, sd, clash=clash)
    worst, worst_e = 0.0, 0.0
    for s, t in enumerate(sub_t):
        fixed = np.rint(np.asarray(sub_c[s]) * 2.0 ** gpos).astype(np.int64)
        out, trace = oracle.minimize(t, fixed, gpos, frc, nrg, cycles, sd, 0.01, dp_mode=True,
                                     clash_cycles=clash)
        drift = float(np.abs(xf[s] - out).max()) * 2.0 ** -gpos
        rel = abs(e[s] - trace[-1]) / abs(trace[-1])
        worst, worst_e = max(worst, drift), max(worst_e, rel)
    print(f"64 systems, {cycles} cycles in FP64: largest coordinate drift from the oracle "
          f"{worst:.2e} A, largest relative energy difference {worst_e:.2e}")
    assert worst < 2e-5 and worst_e < 1e-5


def test_served_trace_over_the_full_size_range():
    reqs = serving.serving_workload(0.08, 200.0, seed=5)
    sizes = [r.topo.natom for r in reqs]
    confs = [len(r.coords) for r in reqs]
    print(f"{len(reqs)} requests, atoms {min(sizes)}..{max(sizes)}, conformers "
          f"{min(confs)}..{max(confs)}")
    fresh = [serving.Request(r.rid, r.arrival, r.topo, r.coords) for r in reqs]
    pool = serving.MinimizationPool(budget_bytes=256 << 20, total_cycles=300, conv_rms=0.1)
    res = serving.serve(pool, fresh, "continuous", 10, realtime=False)
    assert res["requests"] == len(reqs)
    for r in fresh:
        assert (r.steps <= 300).all() and np.isfinite(r.energies).all()
    # The largest, the smallest and the most-conformer requests again, each alone.
    by_size = sorted(fresh, key=lambda r: r.topo.natom)
    again = {by_size[0].rid, by_size[-1].rid, max(fresh, key=lambda r: len(r.coords)).rid}
    for r in fresh:
        if r.rid not in again:
            continue
        alone = serving.MinimizationPool(budget_bytes=256 << 20, total_cycles=300, conv_rms=0.1)
        one = serving.Request(r.rid, 0.0, r.topo, r.coords)
        serving.serve(alone, [one], "continuous", 10, realtime=False)
        assert np.array_equal(one.steps, r.steps), r.rid
        assert np.array_equal(one.energies, r.energies), r.rid
