"""The device minimizer on the molecules of the reference's GPU minimization test
(test/MolecularMechanics/test_hpc_minimization.cu:1042-1104) with its double-precision set-up:
40 position / 40 force / 32 energy bits (38 position bits for the persistent kernel, which takes
no overflow words), 100 cycles of which 10 steepest descent, no clash damping.

Against the oracle (the GPU-semantics restatement pinned by tests/test_oracle_minimizer.py and
tests/test_oracle_minimizer_reference.py): the small molecules' trajectories are insensitive to
last-bit differences, so every mode ends within 1e-6 A of the oracle's coordinates and 1e-8
relative of its energy after all 100 cycles (90 of them conjugate gradient).  Trp-cage is
sensitive (the reference's own end point moves by 3e-3 kcal/mol between batches), so it is
compared over its first 30 cycles, to the same bars.
"""
import numpy as np
import pytest

import oracle
from conftest import golden_system
from stormm_amd import synthesis as S

pytestmark = pytest.mark.gpu

SMALL = ["bromobenzene_iso", "symmetry_L1", "stereo_L1"]
MODE_NAMES = {0: "per-step launches", 1: "persistent", 2: "graph replay"}


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
    for cycles, subset in ((100, SMALL), (30, ["trpcage"])):
        topos, coords, xf, e = run_device(subset, mode, cycles, gpos)
        for s, name in enumerate(subset):
            fixed = np.rint(coords[s] * 2.0 ** gpos).astype(np.int64)
            out, trace, rec = oracle.minimize(topos[s], fixed, gpos, 40, 32, cycles, 10, 0.01,
                                              dp_mode=True, record=True)
            assert np.any(np.abs(rec[10:, 8]) > 1e-3), "conjugate gradient never engaged"
            drift = np.abs(np.asarray(xf[s], dtype=float) - out.astype(float)).max() * \
                2.0 ** -gpos
            worst[name] = drift
            assert drift < 1e-6, (name, drift)
            assert e[s] == pytest.approx(trace[-1], rel=1e-8, abs=1e-7), name
    print(f"[{MODE_NAMES[mode]}, {gpos} position bits] largest coordinate drift from the "
          f"oracle per molecule (A): " + ", ".join(f"{k} {v:.1e}" for k, v in worst.items()))
    assert set(worst) == set(names)
