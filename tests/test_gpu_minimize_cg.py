"""GPU minimizer with the conjugate-gradient update active (steepest-descent cycles fewer than
the cycles run), against the oracle's GPU-semantics minimizer and against textbook linear
conjugate gradient on a quadratic bowl (the pins of tests/test_oracle_minimizer.py, now run on
the device).

Tolerances.  Double precision (40 force bits, 32 energy bits): every mode follows the oracle's
trajectory to 1e-6 A in the coordinates and 1e-8 relative in the final energy over 40 cycles of
which 35 are conjugate gradient; the device's reciprocal and square root differ from the CPU's
in the last bits, and the fixed-precision rounding absorbs that over these runs.  Single
precision (23 force bits): the forces differ from the oracle's by ~1e-6 relative, so the
trajectories separate slowly; over 20 cycles (15 conjugate gradient) the final energies agree
within 1e-4 relative and the coordinates within 2e-3 A.
"""
import numpy as np
import pytest

import oracle
from stormm_amd import synthetic
from stormm_amd import synthesis as S
from test_oracle_minimizer import CG_CASE, bowl, textbook_cg

pytestmark = pytest.mark.gpu

MODES = {"double": (32, 40, 32), "single": (32, 23, 26)}
MODE_NAMES = {0: "per-step launches", 1: "fused cycles, whole-run graph", 2: "graph replay"}


def gpu_minimize(topos, coords, prec, cycles, persistent, sd, dx0=0.01, clash=None,
                 restraints=None):
    gpos, frc, nrg = MODES[prec]
    ag = S.AtomGraphSynthesis(topos, restraints=restraints)
    ps = S.PhaseSpaceSynthesis(ag, gpos, frc)
    sc = S.ScoreCard(ag.system_count, nrg)
    lm = S.LineMinimization(ag.system_count, dx0)
    ps.upload(coords)
    mc = S.minimize_controls(cycles, sd, dx0, 0.0, clash_cycles=clash)
    S.launch_minimization(prec, ag, ps, mc, sc, lm, persistent=persistent)
    xf, _ = ps.fixed()
    return xf, sc.potential_energies(), lm.values()


@pytest.mark.parametrize("persistent", [0, 2])
def test_bowl_conjugate_gradient_on_the_device(persistent):
    """Two curvatures: the device's steepest-descent step lands on textbook step 1 and its first
    conjugate-gradient step reaches the minimum, as the oracle's does (the persistent kernel
    takes no restraints)."""
    ks = CG_CASE["ks"]
    topo, x0, targets, rst = bowl(ks, CG_CASE["seed"])
    expect = textbook_cg(ks, x0, targets, 2)
    gpos, frc, nrg = MODES["double"]
    energies = []
    for cycles in (3, 4, 5):
        _, e, _ = gpu_minimize([topo], [x0], "double", cycles, persistent, 2,
                               dx0=CG_CASE["dx0"], clash=0, restraints=[rst])
        energies.append(float(e[0]))
    print(f"[{MODE_NAMES[persistent]}] energies after 3, 4, 5 cycles: {energies}; "
          f"textbook {expect}")
    assert energies[0] == pytest.approx(expect[0], rel=1e-8)
    assert energies[1] == pytest.approx(expect[1], rel=1e-8)
    assert expect[1] > 0.1 * expect[0] and energies[2] < 1e-8 * expect[0]
    fixed = np.rint(x0 * 2.0 ** gpos).astype(np.int64)
    _, trace = oracle.minimize(topo, fixed, gpos, frc, nrg, 5, 2, CG_CASE["dx0"], dp_mode=True,
                               restraints=rst)
    assert energies[2] == pytest.approx(trace[-1], abs=1e-8 * expect[0])


@pytest.mark.parametrize("persistent", [1, 0, 2])
def test_double_precision_conjugate_gradient_tracks_the_oracle(persistent):
    topos, coords = synthetic.config2(nsys=6, seed=41)
    gpos, frc, nrg = MODES["double"]
    cycles, sd, clash = 40, 5, 8
    xf, e, lm = gpu_minimize(topos, coords, "double", cycles, persistent, sd, clash=clash)
    worst = 0.0
    for s, t in enumerate(topos):
        fixed = np.rint(np.asarray(coords[s]) * 2.0 ** gpos).astype(np.int64)
        out, trace, rec = oracle.minimize(t, fixed, gpos, frc, nrg, cycles, sd, 0.01,
                                          dp_mode=True, clash_cycles=clash, record=True)
        assert np.any(np.abs(rec[sd:, 8]) > 1e-3), "conjugate gradient never engaged"
        drift = np.abs(xf[s] - out).max() * 2.0 ** -gpos
        worst = max(worst, drift)
        assert drift < 1e-6, (s, drift)
        assert e[s] == pytest.approx(trace[-1], rel=1e-8, abs=1e-6)
        assert lm["l_move"][s] == pytest.approx(rec[-1, 7], rel=1e-8)
    print(f"[{MODE_NAMES[persistent]}] largest coordinate drift from the oracle {worst:.2e} A")


@pytest.mark.parametrize("persistent", [1, 0, 2])
def test_single_precision_conjugate_gradient_tracks_the_oracle(persistent):
    topos, coords = synthetic.config2(nsys=6, seed=42)
    gpos, frc, nrg = MODES["single"]
    cycles, sd = 20, 5
    xf, e, _ = gpu_minimize(topos, coords, "single", cycles, persistent, sd, clash=0)
    worst_x, worst_e, worst_here = 0.0, 0.0, 0.0
    for s, t in enumerate(topos):
        fixed = np.rint(np.asarray(coords[s]) * 2.0 ** gpos).astype(np.int64)
        out, trace = oracle.minimize(t, fixed, gpos, frc, nrg, cycles, sd, 0.01, dp_mode=False)
        # Energy scale S = sum_k |E_k| at the device's end point: the total alone can nearly
        # cancel, so errors are stated relative to S, as for single evaluations.
        eo, _, _ = oracle.eval_mm(t, xf[s] * 2.0 ** -gpos, False, nrg)
        scale = sum(abs(eo[k]) for k in oracle.SLOT_NAMES)
        worst_x = max(worst_x, np.abs(xf[s] - out).max() * 2.0 ** -gpos)
        worst_e = max(worst_e, abs(e[s] - trace[-1]) / scale)
        # The reported energy is the FP32 evaluation at the end point (single-evaluation bar).
        worst_here = max(worst_here, abs(e[s] - oracle.total_potential(eo)) / scale)
        assert e[s] < trace[0]
    print(f"[{MODE_NAMES[persistent]}] FP32 vs oracle after {cycles} cycles: coordinates "
          f"{worst_x:.2e} A, end energy {worst_e:.2e} of S (oracle trajectory), "
          f"{worst_here:.2e} of S (oracle at the device's end point)")
    assert worst_x < 2e-3 and worst_e < 1e-5 and worst_here < 1e-6
