"""The device minimizer on the molecules of the reference's GPU minimization test
(test/MolecularMechanics/test_hpc_minimization.cu:1042-1104) with its double-precision set-up:
40 position / 40 force / 32 energy bits (38 position bits for the fused-cycle mode, which takes
no overflow words), 10 steepest-descent cycles, then conjugate gradient, no clash damping.

Against the oracle (the GPU-semantics restatement pinned by tests/test_oracle_minimizer.py and
tests/test_oracle_minimizer_reference.py).  The device's double-precision forces differ from the
oracle's in the last bits (hardware reciprocal and square-root estimates refined by
Newton-Raphson steps), and conjugate gradient amplifies such differences geometrically: measured
on the device, the largest coordinate difference is ~1e-11 A after 5 cycles, ~1e-9 A after 20,
~1e-7 A after 40 and 1e-7 .. 7e-5 A after 100 cycles depending on the molecule (symmetry_L1 the
most sensitive), growing smoothly, with no jump at any cycle.  So the trajectories are compared
tightly over 20 cycles (10 of them conjugate gradient), where any difference in a formula,
constant or rounding would show at 1e-6 A or more, and over 100 cycles to the bounds that growth
allows.  Trp-cage is compared over its first 20 cycles.
"""
import numpy as np
import pytest

import oracle
from conftest import golden_system
from stormm_amd import synthesis as S

pytestmark = pytest.mark.gpu

SMALL = ["bromobenzene_iso", "symmetry_L1", "stereo_L1"]
MODE_NAMES = {0: "per-step launches", 1: "fused cycles, whole-run graph", 2: "graph replay"}


def run_device(names, mode, cycles, gpos):
    systems = [golden_system(n) for n in names]
    topos = [t for t, _ in systems]
    coords = [np.asarray(x) for _, x in systems]
    ag = S.AtomGraphSynthesis(topos)
    ps = S.PhaseSpaceSynthesis(ag, gpos, 40)
    sc = S.ScoreCard(ag.system_count, 32)
    lm = S.LineMinimization(ag.system_count, 0.01)
    ps.upload(coords)
    S.launch_minimization("double", ag, ps, S.minimize_controls(cycles, 10, 0.01, 0.0,
                                                               clash_cycles=0),
                          sc, lm, persistent=mode)
    xf, _ = ps.fixed_values()
    return topos, coords, xf, sc.potential_energies()


@pytest.mark.parametrize("mode", [0, 2, 1])
def test_reference_molecules_track_the_oracle(mode):
    gpos = 38 if mode == 1 else 40
    names = SMALL + ["trpcage"]
    worst = {}
    for cycles, subset, x_bar, e_bar in ((20, SMALL + ["trpcage"], 1e-7, 1e-6),
                                         (100, SMALL, 5e-4, 2e-3)):
        topos, coords, xf, e = run_device(subset, mode, cycles, gpos)
        for s, name in enumerate(subset):
            fixed = np.rint(coords[s] * 2.0 ** gpos).astype(np.int64)
            out, trace, rec = oracle.minimize(topos[s], fixed, gpos, 40, 32, cycles, 10, 0.01,
                                              dp_mode=True, record=True)
            assert np.any(np.abs(rec[10:, 8]) > 1e-3), "conjugate gradient never engaged"
            drift = np.abs(np.asarray(xf[s], dtype=float) - out.astype(float)).max() * \
                2.0 ** -gpos
            worst[(name, cycles)] = drift
            assert drift < x_bar, (name, cycles, drift)
            # Energies in kcal/mol; the end points lie below the start.
            assert abs(e[s] - trace[-1]) < e_bar, (name, cycles, e[s], trace[-1])
            assert e[s] < trace[0]
    print(f"[{MODE_NAMES[mode]}, {gpos} position bits] largest coordinate drift from the "
          f"oracle (A): " + ", ".join(f"{k[0]}@{k[1]} {v:.1e}" for k, v in worst.items()))
    assert {k[0] for k in worst} == set(names)
