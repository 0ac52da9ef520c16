"""Pins of the oracle's GPU-semantics minimizer (conjugate gradient + line search of
src/Math/conjugate_gradient.cui, src/Math/hpc_reduction.cu:39-321 and
src/MolecularMechanics/line_movement.cui) that do not rely on the oracle's own formulas.

The energy surface is a quadratic bowl: atoms held to target points by harmonic positional
restraints (E = k |x - t|^2 inside the well, src/Restraints/restraint_util.tpp) with no other
terms, so every quantity the minimizer produces has a closed form:

* the trial energies of one line search follow from the bowl and the reference's move factors
  (1.04 / 0.8 after the first trial, x 1.03 / x 0.85 after the second, line_movement.cui:107-128);
* the cubic fit through four points of a parabola is exact, so a line search that brackets the
  minimum lands on it, as textbook steepest descent / linear conjugate gradient with exact line
  searches do;
* with two distinct curvatures, linear conjugate gradient reaches the minimum in two exact line
  searches, which pins the Polak-Ribiere mixing of the previous direction;
* the Polak-Ribiere gamma = (f - f_prev) . f / |f_prev|^2 vanishes when a rejected line search
  leaves the forces unchanged (Fletcher-Reeves would give 1);
* the single-precision model truncates the normalized direction to the fixed-point grid
  (conjugate_gradient.cui:232-237, an implicit double-to-integer conversion) and the
  double-precision model rounds it to nearest (doubleToInt95, accumulation.cui:121-133).

Each check is also run with a deliberate mistake planted in the oracle (oracle.plant_mistake) to
show that it fails on that mistake.

The reference's final GPU energies in test/MolecularMechanics/min_energy.m are not used as a pin:
test_minimization_trajectories_are_chaotic shows why.
"""
import numpy as np
import pytest

import oracle
from conftest import golden_system
from stormm_amd import restraints as R
from test_oracle_math import make_topology

GPOS, FRC, NRG = 40, 40, 32


def bowl(ks, seed, offset=0.6):
    """Atoms without charges or Lennard-Jones terms, each held to a random target point by a
    harmonic positional restraint of stiffness k (E = k r^2 for r < 10 A)."""
    n = len(ks)
    rng = np.random.default_rng(seed)
    topo = make_topology(n, q=np.zeros(n), eps=np.zeros(n))
    targets = rng.uniform(-3.0, 3.0, (n, 3))
    x0 = targets + rng.normal(0.0, offset, (n, 3))
    rst = R.concatenate([R.apply_positional_restraint(i, targets[i],
                                                      R.FlatBottomPlan.harmonic(k, 0.0))
                         for i, k in enumerate(ks)])
    return topo, x0, targets, rst


def textbook_cg(ks, x0, targets, steps):
    """Linear conjugate gradient with exact line searches on E = sum k |x - t|^2 (Hestenes and
    Stiefel 1952): the energies after each step."""
    kk = np.repeat(np.asarray(ks, dtype=np.float64), 3)
    x = (np.asarray(x0) - np.asarray(targets)).reshape(-1)
    g = 2.0 * kk * x
    d = -g
    out = [float(np.sum(kk * x * x))]
    for _ in range(steps):
        alpha = (g @ g) / (d @ (2.0 * kk * d))
        x = x + alpha * d
        gn = 2.0 * kk * x
        d = -gn + ((gn @ gn) / (g @ g)) * d
        g = gn
        out.append(float(np.sum(kk * x * x)))
    return out


def run(topo, x0, rst, cycles, dx0, sd_cycles=2, dp_mode=True, frc_bits=FRC):
    fixed = np.rint(np.asarray(x0) * 2.0 ** GPOS).astype(np.int64)
    return oracle.minimize(topo, fixed, GPOS, frc_bits, NRG, cycles, sd_cycles=sd_cycles,
                           dx0=dx0, dp_mode=dp_mode, restraints=rst, record=True)


@pytest.fixture
def planted():
    yield oracle.plant_mistake
    oracle.plant_mistake(None)


# Two curvatures (k = 1 and 8); with this start and dx0 = 4 A the first three line searches
# overshoot and are rejected, the fourth (steepest descent) and fifth (conjugate gradient) bracket
# their minima.
CG_CASE = dict(ks=[1.0] * 3 + [8.0] * 3, seed=3, dx0=4.0)


def cg_bowl_violations():
    ks = CG_CASE["ks"]
    topo, x0, targets, rst = bowl(ks, CG_CASE["seed"])
    expect = textbook_cg(ks, x0, targets, 2)
    _, trace, rec = run(topo, x0, rst, 6, CG_CASE["dx0"])
    e_start = rec[:, 0]
    scale = expect[0]
    bad = []
    # Cycles 0-2 overshoot: every trial point is above the start, the particles go back.
    if not np.allclose(e_start[:4], expect[0], rtol=0, atol=1e-9 * scale):
        bad.append("rejected line searches moved the system")
    # After a rejected search the forces are those of the same point: Polak-Ribiere gamma ~ 0.
    if np.any(np.abs(rec[1:4, 8]) > 1e-9):
        bad.append(f"gamma after rejected searches {rec[1:4, 8]}")
    # Cycle 3 is steepest descent with an exact line search: textbook step 1.
    if abs(e_start[4] - expect[1]) > 1e-9 * scale:
        bad.append(f"steepest-descent step {e_start[4]} vs {expect[1]}")
    # Cycle 4 mixes the previous direction: two curvatures, so the minimum is reached.
    if not (expect[1] > 0.1 * scale and trace[-1] < 1e-9 * scale):
        bad.append(f"conjugate-gradient step leaves E = {trace[-1]}")
    return bad


def test_conjugate_gradient_reaches_a_two_curvature_minimum_in_two_line_searches():
    assert cg_bowl_violations() == []


@pytest.mark.parametrize("mistake", ["fletcher_reeves", "no_conjugation", "gamma_sign"])
def test_conjugate_gradient_pin_fails_on_planted_mistakes(planted, mistake):
    planted(mistake)
    assert cg_bowl_violations() != []


# Isotropic bowl: the steepest-descent direction points straight at the minimum, a distance D
# away, so the energy after moving s along it is k (D - s)^2.
ISO_K = 2.0


def trial_point_violations():
    bad = []
    topo, x0, targets, rst = bowl([ISO_K] * 5, seed=8)
    dist = float(np.sqrt(np.sum((x0 - targets) ** 2)))
    up1, down1, up2, down2 = 1.04, 0.8, 1.03, 0.85   # line_movement.cui:107-128
    for ratio in (0.1, 0.8, 2.5):
        step = ratio * dist
        _, _, rec = run(topo, x0, rst, 1, step)
        e = lambda s: ISO_K * (dist - s) ** 2
        e0, e1 = e(0.0), e(step)
        mfb = up1 if e1 < e0 else down1
        e2 = e(step * (1.0 + mfb))
        mfc = mfb * (up2 if e2 < e1 else down2)
        e3 = e(step * (1.0 + mfb + mfc))
        want = [e0, e1, e2, e3, mfb, mfc]
        got = rec[0, :6]
        if not np.allclose(got, want, rtol=1e-9, atol=1e-9):
            bad.append((ratio, got, want))
        # The cubic fit: the parabola's minimum at D / step units of the move length.
        xmin = dist / step
        span = 1.0 + mfb + mfc
        if e0 < min(e1, e2, e3):
            want_move, want_len = -span, 0.8 * step
        elif xmin < span:
            want_move, want_len = xmin - span, (1.05 if xmin > 0.6 else 0.95) * step
        else:
            want_move, want_len = 0.0, 1.05 * step
        if not (np.isclose(rec[0, 6], want_move, rtol=1e-5, atol=1e-9)
                and np.isclose(rec[0, 7], want_len, rtol=1e-12)):
            bad.append((ratio, rec[0, 6:8], (want_move, want_len)))
    return bad


def test_line_search_trial_points_and_cubic_move_on_a_parabola():
    assert trial_point_violations() == []


@pytest.mark.parametrize("mistake", ["first_growth", "first_shrink", "second_growth"])
def test_line_search_pin_fails_on_planted_mistakes(planted, mistake):
    planted(mistake)
    assert trial_point_violations() != []


def normalization_violations():
    """First trial energy on a one-atom bowl with a coarse 6-bit force grid, where rounding the
    normalized direction up or down moves the atom measurably differently."""
    bad = []
    topo, x0, targets, rst = bowl([ISO_K], seed=21, offset=0.7)
    fbits, step = 6, 0.2
    force = np.rint(2.0 * ISO_K * (targets - x0) * 2.0 ** fbits)
    unit = force * (2.0 ** fbits / np.sqrt(np.sum(force * force)))
    for dp_mode, rounding in ((False, np.trunc), (True, np.rint)):
        direction = rounding(unit)
        moved = x0 + np.rint(step * 2.0 ** (GPOS - fbits) * direction) * 2.0 ** -GPOS
        want = ISO_K * float(np.sum((moved - targets) ** 2))
        _, _, rec = run(topo, x0, rst, 1, step, dp_mode=dp_mode, frc_bits=fbits)
        if abs(rec[0, 1] - want) > 1e-9:
            bad.append((dp_mode, rec[0, 1], want))
    return bad


def test_direction_rounding_follows_each_precision_model():
    assert normalization_violations() == []


def test_direction_rounding_pin_fails_on_planted_mistake(planted):
    planted("normalization_rounding")
    assert normalization_violations() != []


def test_minimization_trajectories_are_chaotic():
    """Why test/MolecularMechanics/min_energy.m cannot pin this minimizer at its 1e-5 kcal/mol
    margin (test_hpc_minimization.cu:836-837): moving each start coordinate of Trp-cage by at
    most one unit of the 2^-40 A grid changes the energy after 100 FP64 cycles (10 of steepest
    descent, 40 / 40 / 32 bits, no clash damping, the snapshot's set-up at
    test_hpc_minimization.cu:400-411,1024-1026) by more than a tenth of a kcal/mol.  A match at
    1e-5 would need the reference GPU's own floating-point results bit for bit.  (The snapshot's
    alanine dipeptide also carries a CMAP term, outside this path.)"""
    topo, xyz = golden_system("trpcage")
    handle = oracle.SystemHandle(topo)
    base = np.rint(np.asarray(xyz) * 2.0 ** GPOS).astype(np.int64)
    finals = []
    for seed in range(4):
        start = base + np.random.default_rng(seed).integers(-1, 2, base.shape) * (seed > 0)
        _, trace = oracle.minimize(topo, start, GPOS, FRC, NRG, 100, sd_cycles=10, dx0=0.01,
                                   dp_mode=True, handle=handle)
        finals.append(trace[-1])
    assert np.ptp(finals) > 0.1
    assert all(f < trace[0] - 50.0 for f in finals)
