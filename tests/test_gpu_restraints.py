"""GPU parity of flat-bottom restraints (src/Potential/valence_potential.tpp:1555-2160) against the
FP64 oracle, in the fused evaluation at a given controls step and inside every minimizer mode.

Tolerances: the RESTRAINT slot within 1e-8 (FP64 mode: 40 force bits, 32 energy bits) or 1e-6
(FP32 mode: 23 force bits, 26 energy bits) of the oracle relative to |E_restraint| + 1 kcal/mol;
the restraint forces (with minus without restraints) within 1e-8 / 1e-5 relative RMS of the
oracle's; in FP64 mode also every slot within 1e-8 of sum_k |E_k| and the total forces within
1e-8 relative RMS, the bars of the unrestrained parity tests.  The oracle is evaluated at exactly
the fixed-precision coordinates the device holds.
"""
import numpy as np
import pytest

import oracle
from stormm_amd import restraints as R
from stormm_amd import synthetic
from stormm_amd import synthesis as S

pytestmark = pytest.mark.gpu

MODES = {"double": dict(gpos=32, frc=40, nrg=32, e_tol=1e-8, f_tol=1e-8),
         "single": dict(gpos=32, frc=23, nrg=26, e_tol=1e-6, f_tol=1e-5)}


def _dihedral(x, i, j, k, l):
    b1, b2, b3 = x[j] - x[i], x[k] - x[j], x[l] - x[k]
    n1, n2 = np.cross(b1, b2), np.cross(b2, b3)
    return float(np.arctan2(np.linalg.norm(b2) * np.dot(b1, n2), np.dot(n1, n2)))


def _angle(x, i, j, k):
    a, b = x[i] - x[j], x[k] - x[j]
    return float(np.arccos(np.dot(a, b) / (np.linalg.norm(a) * np.linalg.norm(b))))


def restraint_set(topo, x, seed, moving=True, npos=None):
    """Restraints of every kind on one system: bimodal wells of random stiffness around the
    current geometry, some sliding between steps 5 and 25."""
    rng = np.random.default_rng(seed)
    x = np.asarray(x, dtype=np.float64)
    n = topo.natom
    recs = []
    span = (5, 25) if moving else (0, 0)

    def plans(lo, r2, r3, hi, k2, k3):
        a = R.FlatBottomPlan(k2, k3, lo, r2, r3, hi, span[0])
        shift = 0.1 * rng.uniform() if moving else 0.0
        b = R.FlatBottomPlan(1.5 * k2, 0.5 * k3, lo + shift, r2 + shift, r3 + shift, hi + shift,
                             span[1])
        return a, b

    for a in range(0, n, 3) if npos is None else range(min(npos, n)):
        r2 = 0.2 * rng.uniform()
        p, q = plans(0.0, r2, r2 + 0.3, 1.0, 1.4, 3.7)
        tgt = x[a] + rng.normal(0.0, 0.4, 3)
        recs.append(R.apply_positional_restraint(a, tgt, p, q, tgt + rng.normal(0.0, 0.2, 3)))
    for i in range(0, n - 4, 3):
        leq = float(np.linalg.norm(x[i + 3] - x[i])) + 0.2 * rng.uniform()
        recs.append(R.apply_distance_restraint(i, i + 3, *plans(leq - 0.25, leq, leq + 0.4,
                                                                leq + 0.8, 1.1, 0.9)))
    for i in range(0, n - 4, 4):
        teq = _angle(x, i, i + 2, i + 3)
        recs.append(R.apply_angle_restraint(
            i, i + 2, i + 3, *plans(max(teq - 0.4, 0.0), max(teq - 0.05, 0.0),
                                    min(teq + 0.15, 2.9), min(teq + 0.5, 3.0), 1.9, 1.8)))
    for i in range(0, n - 4, 5):
        phi = _dihedral(x, i, i + 1, i + 2, i + 3)
        recs.append(R.apply_dihedral_restraint(
            i, i + 1, i + 2, i + 3, *plans(phi - 0.75, phi - 0.15, phi + 0.10, phi + 0.55,
                                           1.1, 1.3)))
    return R.concatenate(recs)


def batch(nsys, seed, every=1, npos=None):
    """Systems with restraints on every `every`-th one; of those, every other one holds
    constant restraints and the rest sliding ones (off before step 5)."""
    topos, coords = synthetic.config2(nsys=nsys, seed=seed)
    coords = [np.asarray(c) for c in coords]
    rst = [restraint_set(t, c, seed * 100 + s, moving=(s // every) % 2 == 1, npos=npos)
           if s % every == 0 else None for s, (t, c) in enumerate(zip(topos, coords))]
    return topos, coords, rst


def sliding(rst):
    return rst is not None and (rst["final_step"] > 0).any()


def evaluate(topos, coords, rst, prec, step, launch=None):
    m = MODES[prec]
    ag = S.AtomGraphSynthesis(topos, restraints=rst)
    ps = S.PhaseSpaceSynthesis(ag, m["gpos"], m["frc"])
    sc = S.ScoreCard(ag.system_count, m["nrg"])
    ctrl = S.MolecularMechanicsControls()
    ctrl.step = step
    ps.upload(coords)
    ps.initialize_forces()
    sc.initialize()
    (launch or S.launch_mm_evaluation)(prec, ag, ctrl, ps, sc, True, True)
    assert ag.status() == 0
    return dict(E=sc.accumulators()[:, :13].copy(), X=ps.coordinates(), F=ps.forces())


@pytest.mark.parametrize("prec", ["double", "single"])
@pytest.mark.parametrize("step", [0, 12, 40])
def test_restrained_evaluation_matches_oracle(prec, step):
    """The RESTRAINT slot and the restraint forces (forces with restraints minus forces without,
    on the device and in the oracle) within the bar of the precision model, relative to the
    restraint energy and force scale; in FP64 mode every slot and the total forces too.  (The
    FP32 error of the other terms is the unrestrained kernels', tested elsewhere; these
    synthetic geometries hold clashes whose r^-12 walls amplify it.)"""
    m = MODES[prec]
    topos, coords, rst = batch(10, 31, every=2)
    res = evaluate(topos, coords, rst, prec, step)
    bare = evaluate(topos, coords, None, prec, step)
    slot = oracle.SLOT["RESTRAINT"]
    print(f"[{prec}] step {step}: {sum(len(r) for r in rst if r is not None)} restraints")
    for s, t in enumerate(topos):
        nrg, _, f = oracle.eval_mm(t, res["X"][s], True, m["nrg"], restraints=rst[s], step=step)
        _, _, f0 = oracle.eval_mm(t, res["X"][s], True, m["nrg"])
        e_rst = res["E"][s][slot] * 2.0 ** -m["nrg"]
        if rst[s] is None:
            assert res["E"][s][slot] == 0 and np.array_equal(res["F"][s], bare["F"][s])
            continue
        if step >= 5 or not sliding(rst[s]):
            assert nrg["RESTRAINT"] != 0.0
        assert abs(e_rst - nrg["RESTRAINT"]) <= m["e_tol"] * (abs(nrg["RESTRAINT"]) + 1.0), s
        df_dev = res["F"][s] - bare["F"][s]
        df_orc = f - f0
        norm = max(np.sqrt((df_orc ** 2).sum()), 1.0)
        assert np.sqrt(((df_dev - df_orc) ** 2).sum()) <= m["f_tol"] * norm, s
        if prec == "double":
            eo = np.array([nrg[k] for k in oracle.SLOT_NAMES])
            eg = res["E"][s].astype(np.float64) * 2.0 ** -m["nrg"]
            assert np.abs(eg - eo).max() <= m["e_tol"] * np.abs(eo).sum(), (s, eg - eo)
            f_err = np.sqrt(((res["F"][s] - f) ** 2).sum() / (f ** 2).sum())
            assert f_err <= m["f_tol"], (s, f_err)


@pytest.mark.parametrize("prec", ["double", "single"])
def test_restraints_leave_the_other_terms_unchanged(prec):
    topos, coords, rst = batch(6, 32)
    a = evaluate(topos, coords, rst, prec, 7)
    b = evaluate(topos, coords, None, prec, 7)
    slot = oracle.SLOT["RESTRAINT"]
    other = [k for k in range(13) if k != slot]
    assert np.array_equal(a["E"][:, other], b["E"][:, other])
    assert (b["E"][:, slot] == 0).all() and (a["E"][:, slot] != 0).all()
    # The valence launch alone carries the restraints; the non-bonded launch has none.
    v = evaluate(topos, coords, rst, prec, 7, launch=S.launch_valence)
    assert np.array_equal(v["E"][:, slot], a["E"][:, slot])
    nb = evaluate(topos, coords, rst, prec, 7, launch=S.launch_nonbonded)
    assert (nb["E"][:, slot] == 0).all()


def test_many_restraints_per_work_unit():
    """More restraints in a unit than threads in its block (every atom of 400-atom systems held
    in place twice over): each is evaluated once."""
    topos, coords = synthetic.uniform_batch(4, 380, 420, seed=33)
    rst = [R.apply_positional_restraints(np.asarray(c) + 0.3, list(range(t.natom)) * 2, 1.5,
                                         0.0, 5.0) for t, c in zip(topos, coords)]
    res = evaluate(topos, coords, rst, "double", 0)
    for s, t in enumerate(topos):
        nrg, _, _ = oracle.eval_mm(t, res["X"][s], False, 32, restraints=rst[s])
        e = res["E"][s][oracle.SLOT["RESTRAINT"]] * 2.0 ** -32
        assert e == pytest.approx(nrg["RESTRAINT"], rel=1e-10)
        assert e == pytest.approx(2 * 1.5 * 0.27 * t.natom, rel=1e-6)


def minimize(topos, coords, rst, prec, cycles, persistent, clash_cycles):
    m = MODES[prec]
    ag = S.AtomGraphSynthesis(topos, restraints=rst)
    ps = S.PhaseSpaceSynthesis(ag, m["gpos"], m["frc"])
    sc = S.ScoreCard(ag.system_count, m["nrg"])
    lm = S.LineMinimization(ag.system_count, 0.01)
    ps.upload(coords)
    mc = S.minimize_controls(cycles, 10, 0.01, 0.0, clash_cycles=clash_cycles)
    S.launch_minimization(prec, ag, ps, mc, sc, lm, persistent=persistent)
    xf, _ = ps.fixed()
    return xf, sc.potential_energies(), sc.report_instantaneous_states()


@pytest.mark.parametrize("prec", ["double", "single"])
def test_restrained_minimizer_modes_agree(prec):
    """The per-step launches and the graph replay give identical trajectories with restraints;
    the persistent minimizer refuses a restrained synthesis (its restraint path is switched
    off)."""
    topos, coords, rst = batch(8, 34, every=2)
    runs = [minimize(topos, coords, rst, prec, 30, p, 6) for p in (0, 2)]
    for x, e, _ in runs[1:]:
        assert all(np.array_equal(a, b) for a, b in zip(x, runs[0][0]))
        assert np.array_equal(e, runs[0][1])
    assert (runs[0][2][::2, oracle.SLOT["RESTRAINT"]] != 0.0).all()
    with pytest.raises(S.StormmError):
        minimize(topos, coords, rst, prec, 30, 1, 6)


def test_restrained_double_minimizer_tracks_the_oracle():
    """The restraints follow the reference minimizer's step counters: sliding restraints
    (steps 5 .. 25) over 30 cycles with 6 clash-damped cycles."""
    topos, coords, rst = batch(4, 35)
    gpos, frc, nrg = MODES["double"]["gpos"], MODES["double"]["frc"], MODES["double"]["nrg"]
    for cycles in (4, 30):
        xf, e, _ = minimize(topos, coords, rst, "double", cycles, 2, min(6, cycles))
        for s, t in enumerate(topos):
            fixed = np.rint(np.asarray(coords[s]) * 2.0 ** gpos).astype(np.int64)
            out, trace = oracle.minimize(t, fixed, gpos, frc, nrg, cycles, 10, 0.01,
                                         dp_mode=True, clash_cycles=min(6, cycles),
                                         restraints=rst[s])
            drift = np.abs(xf[s] - out).max() * 2.0 ** -gpos
            assert drift < 1e-6, (cycles, s, drift)
            assert e[s] == pytest.approx(trace[-1], rel=1e-8, abs=1e-6)


@pytest.mark.skip(reason="mode 1 refuses restrained syntheses: the persistent kernel that once "
                         "ran this case ended in a GPU memory fault whose cause was not found; "
                         "that kernel is removed, and the fused-cycle kernels that now serve "
                         "mode 1 are not run on this case until a cause is shown")
@pytest.mark.parametrize("prec", ["double", "single"])
def test_persistent_restrained_minimizer_equals_per_step(prec):
    topos, coords, rst = batch(8, 34, every=2)
    a = minimize(topos, coords, rst, prec, 30, 0, 6)
    b = minimize(topos, coords, rst, prec, 30, 1, 6)
    assert all(np.array_equal(x, y) for x, y in zip(a[0], b[0]))
    assert np.array_equal(a[1], b[1])
