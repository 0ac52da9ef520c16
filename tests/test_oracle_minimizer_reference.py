"""The oracle's minimizer against the final energies of the reference's own GPU minimizations
(test/MolecularMechanics/min_energy.m, written by test_hpc_minimization.cu:840-846).

What those numbers can and cannot pin.  Each is the ScoreCard energy after the last energy-only
evaluation of a fixed number of cycles (FP64: 100 cycles, 10 steepest descent, 40 / 40 / 32
bits; FP32: 500 cycles, 50 steepest descent, 28 / 24 / 32 bits).
* A trajectory's end point after a fixed cycle count is a property of that implementation's
  rounding, not of the algorithm alone: the reference itself ends Trp-cage at -534.84286089,
  -534.83985375 and -534.8423845 kcal/mol in three batches that differ only in what else shares
  the launch (its energy partial sums round differently), and in the oracle a change of one bit in
  the energy or force scale moves the 100-cycle Trp-cage end point by tens of kcal/mol.
* For the small molecules the oracle's 100-cycle FP64 end points are stable under such
  perturbations (to 1e-8 kcal/mol for bromobenzene) but do not equal the reference's
  (bromobenzene 2.6601925 against 2.6709043): the reference's FP64 run stops short of the basin
  minimum that its own FP32 run, and the oracle, reach.  What differs in the reference's double-
  precision path could not be established by reading it, so the FP64 100-cycle values are not
  used as pins.
* The basin minima are path independent.  The reference's FP32 500-cycle run ends at or just
  above the minimum of the basin it descends into; the oracle minimized to convergence must find
  a minimum at most that high, and for bromobenzene, which the reference converges fully, equal
  to it within the snapshot margin (1e-5).  Alanine dipeptide carries CMAP terms, which this path
  does not evaluate, so it is not among the cases.

These complement tests/test_oracle_minimizer.py, which pins the per-cycle semantics (trial
points, cubic move, Polak-Ribiere conjugation, direction rounding) on closed forms.
"""
import dataclasses

import numpy as np
import pytest

import oracle
from conftest import golden_system

# Convergence runs: FP64, 40 / 40 / 32 bits, 10 steepest-descent cycles.
CONVERGE_CYCLES = 1500
# How far below the reference's FP32 500-cycle end point the converged minimum may lie: the
# reference converges bromobenzene (margin of its snapshot), the other two are still descending
# slowly at 500 cycles (the oracle's own 500-cycle FP32-mode runs stop 2.8e-3 and 4.4e-3 above
# the minimum as well).
BELOW_REF = {"bromobenzene_iso": 1.0e-5, "symmetry_L1": 2.0e-4, "stereo_L1": 5.0e-3}


def fixed_coords(xyz, bits):
    return np.rint(np.asarray(xyz) * 2.0 ** bits).astype(np.int64)


def converged_minimum(name):
    topo, xyz = golden_system(name)
    out, trace, rec = oracle.minimize(topo, fixed_coords(xyz, 40), 40, 40, 32, CONVERGE_CYCLES,
                                      sd_cycles=10, dx0=0.01, dp_mode=True, record=True)
    return topo, out, rec


@pytest.mark.parametrize("name", ["bromobenzene_iso", "symmetry_L1", "stereo_L1"])
def test_converged_minimum_matches_the_reference_end_state(known_answers, name):
    ref = known_answers["minimization"][name]["f"]
    topo, out, rec = converged_minimum(name)
    e_min = rec[-1, 3]
    # Converged: the last cycles' trial points agree to 1e-7 kcal/mol.
    assert np.ptp(rec[-5:, :4]) < 1e-7, rec[-5:, :4]
    assert e_min <= ref + 1e-7, (e_min, ref)
    assert ref - e_min <= BELOW_REF[name], (e_min, ref)
    # The end state is a minimum of the energy the oracle evaluates: the RMS force is small
    # against the starting structure's.
    X = out.astype(np.float64) * 2.0 ** -40
    _, _, f = oracle.eval_mm(topo, X, True, 32)
    _, xyz0 = golden_system(name)
    _, _, f0 = oracle.eval_mm(topo, np.asarray(xyz0), True, 32)
    assert np.sqrt((f ** 2).mean()) < 1e-3 * np.sqrt((f0 ** 2).mean())


def test_single_precision_bits_reach_the_converged_bromobenzene_energy(known_answers):
    """The reference's FP32 set-up (28 / 24 / 32 bits, 500 cycles, 50 steepest descent) on the
    oracle reaches the reference's FP32 end point within the snapshot margin."""
    topo, xyz = golden_system("bromobenzene_iso")
    ka = known_answers["minimization"]
    _, _, rec = oracle.minimize(topo, fixed_coords(xyz, 28), 28, 24, 32, 500, sd_cycles=50,
                                dx0=0.01, dp_mode=False, record=True)
    assert rec[-1, 3] == pytest.approx(ka["bromobenzene_iso"]["f"], abs=ka["margin"])


def test_planted_energy_mistake_moves_the_minimum(known_answers):
    """A planted mistake in the minimizer alone still converges to the same minimum (the check
    above pins the energy surface and convergence, not the per-cycle steps), while the reference
    end state is missed if the surface is changed: scaling the 1:4 Lennard-Jones term by 1.05 in
    the oracle's topology moves bromobenzene's minimum by far more than the margin."""
    topo, xyz = golden_system("bromobenzene_iso")
    ref = known_answers["minimization"]["bromobenzene_iso"]["f"]
    oracle.plant_mistake("no_conjugation")
    try:
        _, _, rec = oracle.minimize(topo, fixed_coords(xyz, 40), 40, 40, 32, CONVERGE_CYCLES,
                                    sd_cycles=10, dx0=0.01, dp_mode=True, record=True)
    finally:
        oracle.plant_mistake(None)
    assert rec[-1, 3] == pytest.approx(ref, abs=1e-5)
    wrong = dataclasses.replace(topo, dihe_scnb=np.asarray(topo.dihe_scnb) / 1.05)
    _, _, rec2 = oracle.minimize(wrong, fixed_coords(xyz, 40), 40, 40, 32, CONVERGE_CYCLES,
                                 sd_cycles=10, dx0=0.01, dp_mode=True, record=True)
    assert abs(rec2[-1, 3] - ref) > 1e-3


def test_reference_trp_cage_end_points_depend_on_the_batch(known_answers):
    """The fact that rules the FP64 100-cycle end points out as pins, in the reference's own
    data: the same Trp-cage minimization ends at three energies in three batches, 3e-3 kcal/mol
    apart, 300 times the snapshot margin."""
    t = known_answers["minimization"]["trpcage"]
    ends = [t["d"]] + t["d_other_batches"]
    assert max(ends) - min(ends) > 100 * known_answers["minimization"]["margin"]
