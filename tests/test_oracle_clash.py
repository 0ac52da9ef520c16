"""Pins of the oracle's clash damping (src/Potential/soft_core_potentials.tpp:32-55,145-188 as
used by nonbonded_potential.tpp:108-172 and valence_potential.tpp:1337-1360) and of its
minimizer schedule (src/MolecularMechanics/hpc_minimization.cu:269-281).

The softened cores are built to meet the plain potentials with the same value and slope at the
clash boundary, so value and slope continuity there catch a wrong constant term or a wrong
curvature; the forces are checked against central differences inside the cores (a wrong force
expression or sign fails); outside the cores and with zero parameters the damped evaluation must
give the plain potentials; and the core shapes are checked to be polynomials in (r + 1) of the
stated degree with no lower-order terms, which a dropped or extra term would break.
"""
import math

import numpy as np
import pytest

import oracle
from stormm_amd import synthetic
from stormm_amd.topology import AMBER_COULOMB

from test_oracle_math import make_topology


def pair_energy(t, r, d, ratio, key):
    xyz = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]])
    nrg, _, f = oracle.eval_mm(t, xyz, True, clash_distance=d, clash_ratio=ratio)
    return nrg[key], f


def two_atoms(q=(0.4, -0.25), sigma=(3.1, 2.7), eps=(0.15, 0.05)):
    return make_topology(2, q=list(q), sigma=list(sigma), eps=list(eps))


@pytest.mark.parametrize("d", [0.2, 0.8, 1.7])
def test_coulomb_core_meets_the_plain_potential_with_equal_slope(d):
    t = two_atoms(sigma=(0.5, 0.5), eps=(1e-12, 1e-12))
    qq = AMBER_COULOMB * 0.4 * -0.25
    h = 1.0e-9
    e_in, _ = pair_energy(t, d - h, d, 0.0, "ELECTROSTATIC")
    e_out, _ = pair_energy(t, d + h, d, 0.0, "ELECTROSTATIC")
    # No jump at the boundary beyond what the slope gives over 2h.
    assert abs(e_in - e_out) <= 2.0 * h * abs(qq) / d ** 2 * 1.01 + 1e-10
    h = 1.0e-6
    s_in = (pair_energy(t, d - h, d, 0.0, "ELECTROSTATIC")[0] -
            pair_energy(t, d - 3 * h, d, 0.0, "ELECTROSTATIC")[0]) / (2 * h)
    s_out = (pair_energy(t, d + 3 * h, d, 0.0, "ELECTROSTATIC")[0] -
             pair_energy(t, d + h, d, 0.0, "ELECTROSTATIC")[0]) / (2 * h)
    assert s_in == pytest.approx(s_out, rel=1e-4)
    assert s_out == pytest.approx(-qq / d ** 2, rel=1e-4)


@pytest.mark.parametrize("ratio", [0.5, 0.8])
def test_lj_core_meets_the_plain_potential_with_equal_slope(ratio):
    t = two_atoms(q=(0.0, 0.0))
    sigma = 0.5 * (3.1 + 2.7)
    rl = ratio * sigma
    e = lambda r: pair_energy(t, r, 0.0, ratio, "VDW")[0]
    h = 1.0e-9
    slope = abs(e(rl + 2e-6) - e(rl + 1e-6)) / 1e-6
    assert abs(e(rl - h) - e(rl + h)) <= 2.0 * h * slope * 1.01 + 1e-9
    h = 1.0e-6
    s_in = (e(rl - h) - e(rl - 3 * h)) / (2 * h)
    s_out = (e(rl + 3 * h) - e(rl + h)) / (2 * h)
    assert s_in == pytest.approx(s_out, rel=1e-4)


def test_core_shapes_are_pure_even_and_quartic_polynomials_in_r_plus_one():
    """Inside the cores E(r) = a (r + 1)^2 + b (Coulomb) and a (r + 1)^4 + b (LJ): fitting the
    full polynomial in u = r + 1 must give vanishing lower-order coefficients."""
    t = two_atoms(sigma=(0.5, 0.5), eps=(1e-12, 1e-12))
    u = np.linspace(0.2, 1.4, 9) + 1.0
    e = np.array([pair_energy(t, x - 1.0, 1.5, 0.0, "ELECTROSTATIC")[0] for x in u])
    c = np.polyfit(u, e, 2)
    assert abs(c[1]) < 1e-8 * abs(c[0]) * 10
    t = two_atoms(q=(0.0, 0.0))
    u = np.linspace(0.1, 1.3, 9) + 1.0
    e = np.array([pair_energy(t, x - 1.0, 0.0, 0.5, "VDW")[0] for x in u])
    c = np.polyfit(u, e, 4)
    assert np.all(np.abs(c[1:4]) < 1e-7 * abs(c[0]))
    assert c[0] != 0.0


@pytest.mark.parametrize("r", [0.1, 0.35, 1.0, 1.45])
def test_core_forces_match_central_differences(r):
    t = two_atoms()
    d, ratio = 1.2, 0.6
    h = 1.0e-6
    ep = oracle.total_potential(oracle.eval_mm(t, np.array([[0, 0, 0], [r + h, 0, 0]]), False,
                                               clash_distance=d, clash_ratio=ratio)[0])
    em = oracle.total_potential(oracle.eval_mm(t, np.array([[0, 0, 0], [r - h, 0, 0]]), False,
                                               clash_distance=d, clash_ratio=ratio)[0])
    _, f = pair_energy(t, r, d, ratio, "VDW")
    dedr = (ep - em) / (2 * h)
    assert f[1, 0] == pytest.approx(-dedr, rel=1e-6, abs=1e-6)
    assert f[0, 0] == pytest.approx(dedr, rel=1e-6, abs=1e-6)


def test_outside_the_cores_and_with_zero_parameters_the_plain_potentials_apply():
    topo, x = synthetic.branched_molecule(40, seed=5, gb=True)
    plain = oracle.eval_mm(topo, x)
    zero = oracle.eval_mm(topo, x, clash_distance=0.0, clash_ratio=0.0)
    assert plain[1] == zero[1]
    assert np.array_equal(plain[2], zero[2])
    # Every pair of a relaxed molecule lies far outside a tiny core.
    tiny = oracle.eval_mm(topo, x, clash_distance=1e-3, clash_ratio=1e-3)
    for k in ("ELECTROSTATIC", "VDW", "ELEC_ONE_FOUR", "VDW_ONE_FOUR", "BOND"):
        assert tiny[0][k] == pytest.approx(plain[0][k], rel=1e-12, abs=1e-10)
    assert np.allclose(tiny[2], plain[2], rtol=1e-10, atol=1e-9)


def test_damping_caps_a_clash():
    """Two atoms nearly on top of each other: the plain potentials explode, the damped ones stay
    at the value of the core at r = 0."""
    t = two_atoms()
    xyz = np.array([[0.0, 0.0, 0.0], [0.05, 0.0, 0.0]])
    plain = oracle.total_potential(oracle.eval_mm(t, xyz)[0])
    damped = oracle.total_potential(oracle.eval_mm(t, xyz, clash_distance=0.2,
                                                   clash_ratio=0.5)[0])
    assert plain > 1e12
    assert abs(damped) < 1e5


def test_minimizer_schedule_reduces_to_the_plain_minimizer_after_the_damped_cycles():
    """Cycles at or beyond clash_cycles evaluate the plain potentials: a run whose damping lasts
    k cycles follows the undamped run from the same start exactly when k = 0, and a damped run's
    first energy is the damped potential at the full floor and ratio."""
    topo, x = synthetic.branched_molecule(30, seed=9, gb=True)
    fixed = np.rint(x * 2.0 ** 32).astype(np.int64)
    a, ta = oracle.minimize(topo, fixed, 32, 40, 32, 12, dp_mode=True)
    b, tb = oracle.minimize(topo, fixed, 32, 40, 32, 12, dp_mode=True, clash_cycles=0)
    assert np.array_equal(a, b) and np.array_equal(ta, tb)
    c, tc = oracle.minimize(topo, fixed, 32, 40, 32, 12, dp_mode=True, clash_cycles=6,
                            clash_r0=0.2, clash_ratio=0.5)
    xd = fixed.astype(np.float64) * 2.0 ** -32
    _, acc, _ = oracle.eval_mm(topo, xd, False, nrg_bits=32, clash_distance=0.2,
                               clash_ratio=0.5)
    e0 = sum(acc[k] for k in oracle.SLOT_NAMES[:13]) * 2.0 ** -32
    assert tc[0] == pytest.approx(e0, rel=1e-12)
    assert np.all(np.isfinite(tc))
    assert tc[-1] < tc[0]
