"""Pins of the oracle that do not depend on the reference's fixtures: closed forms, finite
differences, conservation and invariance laws, branch continuity of the GB integral, and the
line-search cubic.  Each is chosen so that a dropped term, a wrong sign, a swapped index or a
transposed operand in the oracle would fail it.
"""
import math

import numpy as np
import pytest

import oracle
from stormm_amd import synthetic
from stormm_amd.topology import AMBER_COULOMB, IGB_HCT, IGB_NONE, IGB_OBC, IGB_OBC2, Topology


def make_topology(n, q=None, sigma=None, eps=None, bonds=(), bond_k=(), bond_r0=(), angles=(),
                  angle_k=(), angle_theta=(), dihedrals=(), amp=(), phase=(), freq=(),
                  scee=None, scnb=None, igb=IGB_NONE, radii=None, screen=None):
    q = np.zeros(n) if q is None else np.asarray(q, float)
    sigma = np.full(n, 3.0) if sigma is None else np.asarray(sigma, float)
    eps = np.full(n, 0.1) if eps is None else np.asarray(eps, float)
    sij = 0.5 * (sigma[:, None] + sigma[None, :])
    eij = np.sqrt(eps[:, None] * eps[None, :])
    A = (4 * eij * sij ** 12).reshape(-1)
    B = (4 * eij * sij ** 6).reshape(-1)
    nd = len(dihedrals)
    return Topology(
        natom=n, charge=q, lj_idx=np.arange(n, dtype=np.int32), n_lj_types=n, lja=A, ljb=B,
        lja14=A.copy(), ljb14=B.copy(),
        bonds=np.array(bonds, dtype=np.int32).reshape(-1, 2), bond_k=np.array(bond_k, float),
        bond_r0=np.array(bond_r0, float),
        angles=np.array(angles, dtype=np.int32).reshape(-1, 3), angle_k=np.array(angle_k, float),
        angle_theta=np.array(angle_theta, float),
        dihedrals=np.array(dihedrals, dtype=np.int32).reshape(-1, 4),
        dihe_amp=np.array(amp, float), dihe_phase=np.array(phase, float),
        dihe_freq=np.array(freq, float), dihe_improper=np.zeros(nd, dtype=np.int32),
        dihe_scee=np.full(nd, 1.2) if scee is None else np.asarray(scee, float),
        dihe_scnb=np.full(nd, 2.0) if scnb is None else np.asarray(scnb, float),
        pb_radii=np.full(n, 1.5) if radii is None else np.asarray(radii, float),
        gb_screen=np.full(n, 0.8) if screen is None else np.asarray(screen, float),
        igb=igb)


def test_two_atom_coulomb_lennard_jones_closed_form():
    r = 3.7
    q = [0.4, -0.25]
    t = make_topology(2, q=q, sigma=[3.1, 2.7], eps=[0.15, 0.05])
    xyz = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]])
    nrg, acc, f = oracle.eval_mm(t, xyz)
    s, e = 0.5 * (3.1 + 2.7), math.sqrt(0.15 * 0.05)
    e_qq = AMBER_COULOMB * q[0] * q[1] / r
    e_lj = 4 * e * ((s / r) ** 12 - (s / r) ** 6)
    assert nrg["ELECTROSTATIC"] == pytest.approx(e_qq, rel=1e-14)
    assert nrg["VDW"] == pytest.approx(e_lj, rel=1e-13)
    dedr = -e_qq / r + 4 * e * (-12 * s ** 12 / r ** 13 + 6 * s ** 6 / r ** 7)
    # Atom 1 sits at +x: its force is -dE/dr along +x; atom 0 feels the opposite.
    assert f[1, 0] == pytest.approx(-dedr, rel=1e-12)
    assert f[0, 0] == pytest.approx(dedr, rel=1e-12)
    assert np.allclose(f[:, 1:], 0.0)


def test_bonded_pair_is_excluded_from_nonbonded():
    t = make_topology(2, q=[0.5, -0.5], bonds=[(0, 1)], bond_k=[300.0], bond_r0=[1.2])
    nrg, _, _ = oracle.eval_mm(t, np.array([[0, 0, 0], [1.3, 0, 0]], float))
    assert nrg["ELECTROSTATIC"] == 0.0 and nrg["VDW"] == 0.0
    assert nrg["BOND"] == pytest.approx(300.0 * 0.1 ** 2, rel=1e-12)


@pytest.mark.parametrize("delta", [-0.07, 0.0, 0.13])
def test_bond_closed_form(delta):
    k, r0 = 317.0, 1.41
    t = make_topology(2, bonds=[(0, 1)], bond_k=[k], bond_r0=[r0])
    d = np.array([0.3, -0.5, 0.8])
    d /= np.linalg.norm(d)
    xyz = np.array([[0.2, 0.1, -0.3], [0.2, 0.1, -0.3]]) + np.array([[0, 0, 0], (r0 + delta) * d])
    nrg, _, f = oracle.eval_terms(t, xyz, ["BOND"])
    assert nrg["BOND"] == pytest.approx(k * delta ** 2, abs=1e-12)
    np.testing.assert_allclose(f[1], -2 * k * delta * d, atol=1e-10)
    np.testing.assert_allclose(f[0], 2 * k * delta * d, atol=1e-10)


def test_angle_closed_form():
    k, th0 = 55.0, math.radians(109.5)
    t = make_topology(3, angles=[(0, 1, 2)], angle_k=[k], angle_theta=[th0])
    xyz = np.array([[1.2, 0, 0], [0, 0, 0], [0, 1.4, 0]])
    nrg, _, f = oracle.eval_terms(t, xyz, ["ANGLE"])
    assert nrg["ANGLE"] == pytest.approx(k * (math.pi / 2 - th0) ** 2, rel=1e-12)
    # dE/dtheta = 2k(theta - theta0); the force on atom 0 is tangential, magnitude |dE/dtheta|/r.
    mag = 2 * k * abs(math.pi / 2 - th0) / 1.2
    assert np.linalg.norm(f[0]) == pytest.approx(mag, rel=1e-10)
    # theta = 90 deg < theta0 opens by pushing atom 0 (on +x) away from atom 2 (on +y): along -y.
    assert f[0, 1] < 0 and abs(f[0, 0]) < 1e-12
    np.testing.assert_allclose(f.sum(axis=0), 0.0, atol=1e-10)


@pytest.mark.parametrize("phi_deg", [0.0, 60.0, 180.0, -75.0])
@pytest.mark.parametrize("n,gamma", [(1, 0.0), (2, math.pi), (3, 0.0)])
def test_dihedral_closed_form(phi_deg, n, gamma):
    V = 1.37
    phi = math.radians(phi_deg)
    # Atoms placed so that the i-j-k-l torsion is exactly phi (IUPAC sign).
    xyz = np.array([[1.0, 1.0, 0.0], [1.0, 0.0, 0.0], [2.5, 0.0, 0.0],
                    [2.5, math.cos(phi), math.sin(phi)]])
    t = make_topology(4, dihedrals=[(0, 1, 2, 3)], amp=[V], phase=[gamma], freq=[n],
                      scee=[0.0], scnb=[0.0])
    nrg, _, f = oracle.eval_terms(t, xyz, ["DIHEDRAL"])
    assert nrg["PROPER_DIHEDRAL"] == pytest.approx(V * (1 + math.cos(n * phi - gamma)),
                                                   abs=1e-12)
    np.testing.assert_allclose(f.sum(axis=0), 0.0, atol=1e-10)


def _finite_difference(topo, xyz, h=1.0e-5, atoms=None):
    atoms = range(topo.natom) if atoms is None else atoms
    h_ = oracle.SystemHandle(topo)
    out = {}
    for a in atoms:
        g = np.zeros(3)
        for d in range(3):
            xp = xyz.copy()
            xm = xyz.copy()
            xp[a, d] += h
            xm[a, d] -= h
            ep = oracle.total_potential(oracle.eval_mm(topo, xp, False, handle=h_)[0])
            em = oracle.total_potential(oracle.eval_mm(topo, xm, False, handle=h_)[0])
            g[d] = -(ep - em) / (2 * h)
        out[a] = g
    return out


@pytest.mark.parametrize("igb", [IGB_NONE, IGB_HCT, IGB_OBC, IGB_OBC2])
def test_forces_match_finite_differences(igb):
    topo, xyz = synthetic.branched_molecule(28, seed=77, gb=igb != IGB_NONE)
    topo.igb = igb
    _, _, f = oracle.eval_mm(topo, xyz, True)
    fd = _finite_difference(topo, xyz, atoms=range(0, 28, 3))
    for a, g in fd.items():
        np.testing.assert_allclose(f[a], g, rtol=2e-6, atol=2e-6 * np.abs(f).max())


def test_newton_third_law_and_rigid_motion_invariance():
    topo, xyz = synthetic.branched_molecule(40, seed=5, gb=True)
    nrg, _, f = oracle.eval_mm(topo, xyz, True)
    np.testing.assert_allclose(f.sum(axis=0), 0.0, atol=1e-9 * np.abs(f).max() * topo.natom)
    torque = np.cross(xyz, f).sum(axis=0)
    np.testing.assert_allclose(torque, 0.0, atol=1e-8 * np.abs(f).max() * topo.natom)
    rng = np.random.default_rng(3)
    qr, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(qr) < 0:
        qr[:, 0] *= -1
    moved = xyz @ qr.T + np.array([3.0, -7.0, 11.0])
    nrg2, _, f2 = oracle.eval_mm(topo, moved, True)
    for k in oracle.SLOT_NAMES:
        assert nrg2[k] == pytest.approx(nrg[k], rel=1e-10, abs=1e-10)
    np.testing.assert_allclose(f2, f @ qr.T, atol=1e-8 * np.abs(f).max())


def test_gb_single_atom_self_energy_closed_form():
    """One atom: psi = 0, so R = rho and E = -q^2 C (1 - 1/eps) / (2 rho)."""
    for igb in (IGB_HCT, IGB_OBC, IGB_OBC2):
        t = make_topology(1, q=[0.7], radii=[1.7], screen=[0.8], igb=igb)
        e, acc, f, psi, radii, sdi = oracle.eval_gb(t, np.zeros((1, 3)))
        rho = 1.7 - 0.09
        assert psi[0] == 0.0 and radii[0] == pytest.approx(rho, rel=1e-15)
        assert e == pytest.approx(-0.5 * 0.49 * AMBER_COULOMB * (1 - 1 / 78.5) / rho, rel=1e-14)


def test_gb_far_pair_tends_to_screened_coulomb():
    """At large separation the GB pair term tends to -C qi qj (1 - 1/eps) / r."""
    q = [0.6, -0.45]
    t = make_topology(2, q=q, radii=[1.5, 1.8], screen=[0.8, 0.75], igb=IGB_OBC2)
    e_self = 0.0
    for i in range(2):
        ts = make_topology(1, q=[q[i]], radii=[[1.5, 1.8][i]], screen=[0.8], igb=IGB_OBC2)
        e_self += oracle.eval_gb(ts, np.zeros((1, 3)))[0]
    r = 400.0
    e = oracle.eval_gb(t, np.array([[0, 0, 0], [r, 0, 0]], float))[0]
    pair = e - e_self
    assert pair == pytest.approx(-AMBER_COULOMB * q[0] * q[1] * (1 - 1 / 78.5) / r, rel=1e-5)


@pytest.mark.parametrize("igb", [IGB_HCT, IGB_OBC2])
def test_gb_energy_continuous_across_integral_branches(igb):
    """The HCT integral's branches (far-field series, r > ri + sj, overlap, engulfed) join
    continuously; a wrong sign or swapped radius in any branch breaks this."""
    radii, screen = [1.6, 1.9], [0.85, 0.72]
    t = make_topology(2, q=[0.3, 0.2], radii=radii, screen=screen, igb=igb)
    ri = radii[0] - 0.09
    sj = screen[1] * (radii[1] - 0.09)
    for edge in (4.0 * sj, ri + sj, abs(ri - sj) + 1e-9):
        lo = oracle.eval_gb(t, np.array([[0, 0, 0], [edge - 1e-7, 0, 0]]))[0]
        hi = oracle.eval_gb(t, np.array([[0, 0, 0], [edge + 1e-7, 0, 0]]))[0]
        assert abs(hi - lo) < 5e-5 * max(1.0, abs(lo)), edge


def test_cubic_line_move_recovers_exact_cubic():
    rng = np.random.default_rng(11)
    for _ in range(50):
        coefs = rng.normal(size=4)
        mvec = np.cumsum(np.concatenate([[0.0], rng.uniform(0.5, 1.2, 3)]))
        evec = np.polyval(coefs, mvec)
        mf, nbm, got = oracle.cubic_move(evec, mvec)
        np.testing.assert_allclose(got, coefs, rtol=1e-9, atol=1e-9)


def test_cubic_line_move_decisions():
    mvec = np.array([0.0, 1.0, 2.04, 3.1])
    # Every trial point above the start: return to the start and shrink the step.
    mf, nbm, _ = oracle.cubic_move([0.0, 1.0, 2.0, 3.0], mvec)
    assert mf == pytest.approx(-mvec[3]) and nbm == 0.8
    # A cubic with its local minimum inside the bracket: E = (x - 1.5)^2 + 0.05 x^3 has
    # E'(x) = 0 at x = (-2 + sqrt(5.8)) / 0.3; move there from the last point.
    ev = [(x - 1.5) ** 2 + 0.05 * x ** 3 for x in mvec]
    xmin = (-2.0 + math.sqrt(5.8)) / 0.3
    mf, nbm, _ = oracle.cubic_move(ev, mvec)
    assert mf == pytest.approx(xmin - mvec[3], rel=1e-9) and nbm == 1.05
    # A cubic with no stationary point that keeps falling: stay at the last point, grow 1.2x.
    ev = [-x - 0.1 * x ** 3 for x in mvec]
    mf, nbm, _ = oracle.cubic_move(ev, mvec)
    assert mf == 0.0 and nbm == 1.2
    # The local minimum lies beyond the bracket (x ~ 4.67 > 3.1): stay, grow 1.05x.
    ev = [(x - 5.0) ** 2 + 0.01 * x ** 3 for x in mvec]
    mf, nbm, _ = oracle.cubic_move(ev, mvec)
    assert mf == 0.0 and nbm == 1.05


def test_minimizer_lowers_energy_and_reaches_a_local_minimum():
    """The GPU-semantics minimizer on a perturbed molecule: the energy drops, and displacing
    any coordinate of the result raises the energy (test/MolecularMechanics/
    test_minimization.cpp:53-125 applies the same local-minimum test)."""
    topo, xyz = synthetic.branched_molecule(24, seed=9, gb=True)
    xyz = xyz + np.random.default_rng(1).normal(0, 0.2, xyz.shape)
    gpos = 32
    fixed = np.rint(xyz * 2.0 ** gpos).astype(np.int64)
    out, trace = oracle.minimize(topo, fixed, gpos, 40, 32, 400, sd_cycles=50, dx0=0.01,
                                 dp_mode=True)
    assert trace[-1] < trace[0] - 1.0
    assert np.all(np.diff(trace[5:]) < 1e-6 * np.abs(trace).max() + 1e-6)
    final = out.astype(np.float64) * 2.0 ** -gpos
    e0 = oracle.total_potential(oracle.eval_mm(topo, final, False)[0])
    rng = np.random.default_rng(2)
    for _ in range(20):
        a, d = rng.integers(topo.natom), rng.integers(3)
        for step in (0.01, -0.01):
            moved = final.copy()
            moved[a, d] += step
            assert oracle.total_potential(oracle.eval_mm(topo, moved, False)[0]) > e0 - 1e-6
