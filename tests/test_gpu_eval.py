"""GPU parity of the fused energy / force evaluation against the FP64 oracle.

Tolerances (stated per precision model):
  FP64 mode (40 force bits, 32 energy bits): every energy component within 1e-8 of the oracle
    relative to the system's energy scale S = sum_k |E_k|, forces within 1e-8 relative RMS.
  FP32 mode (23 force bits, 26 energy bits): total and per-component energy within 1e-6 * S,
    forces within 1e-5 relative RMS (sqrt(sum |dF|^2) / sqrt(sum |F|^2)) per system.
The oracle is evaluated at exactly the fixed-precision coordinates the device holds.
"""
import numpy as np
import pytest

import oracle
from conftest import golden_system
from stormm_amd import synthetic
from stormm_amd import synthesis as S
from stormm_amd.topology import IGB_HCT, IGB_NONE, IGB_OBC, IGB_OBC2

pytestmark = pytest.mark.gpu

MODES = {"double": dict(gpos=32, frc=40, nrg=32, e_tol=1e-8, f_tol=1e-8),
         "single": dict(gpos=32, frc=23, nrg=26, e_tol=1e-6, f_tol=1e-5)}


def evaluate(topos, coords, prec, eval_force=True, launch=None, shape=None):
    m = MODES[prec]
    ag = S.AtomGraphSynthesis(topos, shape)
    ps = S.PhaseSpaceSynthesis(ag, m["gpos"], m["frc"])
    sc = S.ScoreCard(ag.system_count, m["nrg"])
    ps.upload(coords)
    ps.initialize_forces()
    sc.initialize()
    (launch or S.launch_mm_evaluation)(prec, ag, None, ps, sc, eval_force, True)
    assert ag.status() == 0
    xf, ff = ps.fixed()
    return dict(E=sc.accumulators()[:, :13].copy(), X=ps.coordinates(), F=ps.forces(),
                Ffix=ff, nrg_bits=m["nrg"])


def check_against_oracle(topos, res, prec, terms=("ALL",)):
    m = MODES[prec]
    worst = (0.0, 0.0)
    for s, t in enumerate(topos):
        nrg, acc, f = oracle.eval_terms(t, res["X"][s], list(terms), m["nrg"])
        eo = np.array([nrg[k] for k in oracle.SLOT_NAMES])
        eg = res["E"][s].astype(np.float64) * 2.0 ** -m["nrg"]
        scale = np.abs(eo).sum()
        comp_err = np.abs(eg - eo).max() / scale
        tot_err = abs(eg.sum() - eo.sum()) / scale
        f_err = np.sqrt(((res["F"][s] - f) ** 2).sum() / max((f ** 2).sum(), 1e-300))
        assert comp_err <= m["e_tol"] and tot_err <= m["e_tol"], (s, t.natom, comp_err, tot_err)
        assert f_err <= m["f_tol"], (s, t.natom, f_err)
        worst = (max(worst[0], tot_err), max(worst[1], f_err))
    print(f"[{prec}] worst energy error / sum|E_k| = {worst[0]:.2e}, "
          f"worst force relative RMS = {worst[1]:.2e}")
    return worst


@pytest.mark.parametrize("prec", ["double", "single"])
def test_config0_vacuum_matches_oracle(prec):
    topos, coords = synthetic.config0()
    res = evaluate(topos, coords, prec)
    check_against_oracle(topos, res, prec)
    # 32 identical copies give identical results.
    assert all(np.array_equal(res["Ffix"][0], f) for f in res["Ffix"])
    assert np.all(res["E"] == res["E"][0])


@pytest.mark.parametrize("prec", ["double", "single"])
def test_config1_sample_matches_oracle(prec):
    topos, coords = synthetic.config1(nsys=40, seed=1)
    res = evaluate(topos, coords, prec)
    check_against_oracle(topos, res, prec)


@pytest.mark.parametrize("igb", [IGB_HCT, IGB_OBC, IGB_OBC2])
@pytest.mark.parametrize("prec", ["double", "single"])
def test_gb_models_match_oracle(prec, igb):
    topos, coords = synthetic.uniform_batch(12, 20, 140, seed=31)
    for t in topos:
        t.igb = igb
    res = evaluate(topos, coords, prec)
    check_against_oracle(topos, res, prec)


@pytest.mark.parametrize("prec", ["double", "single"])
def test_reference_systems_with_stored_lj_tables(prec):
    trp, xt = golden_system("trpcage", igb=IGB_OBC2)
    ala, xa = golden_system("ala_dipeptide", igb=IGB_OBC2)
    res = evaluate([trp, ala], [xt, xa], prec)
    check_against_oracle([trp, ala], res, prec)


@pytest.mark.parametrize("prec", ["double", "single"])
def test_ragged_and_extreme_sizes(prec):
    sizes = [4, 5, 15, 16, 17, 31, 32, 33, 63, 64, 65, 255, 256, 257, 511, 512]
    topos, coords = synthetic.generate(sizes, [1000 + n for n in sizes], gb=True)
    res = evaluate(topos, coords, prec)
    check_against_oracle(topos, res, prec)


def without_terms(topo, kinds):
    """A copy of a topology with every valence term of the given kinds removed."""
    import copy
    t = copy.copy(topo)
    # Exclusions follow the bonds on both sides (the device derives them from the bonds).
    t._excl_cache = None
    if "bonds" in kinds:
        t.bonds = np.zeros((0, 2), np.int32)
        t.bond_k = t.bond_r0 = np.zeros(0)
    if "angles" in kinds:
        t.angles = np.zeros((0, 3), np.int32)
        t.angle_k = t.angle_theta = np.zeros(0)
    if "dihedrals" in kinds:
        t.dihedrals = np.zeros((0, 4), np.int32)
        t.dihe_amp = t.dihe_phase = t.dihe_freq = np.zeros(0)
        t.dihe_improper = np.zeros(0, np.int32)
        t.dihe_scee = t.dihe_scnb = np.zeros(0)
    return t


@pytest.mark.parametrize("prec", ["double", "single"])
def test_systems_lacking_a_kind_of_valence_term(prec):
    """Work units take valence terms kind by kind over their systems; systems without any
    bonds, angles or dihedrals (in every combination, next to complete ones) must be skipped
    correctly.  A system without bonds has no exclusions either, so its atoms are spread to
    twice their distances from the centroid: its former bond partners then meet as ordinary
    non-bonded pairs at 2-3 A, with forces inside the range of the fixed-precision
    accumulators (40 bits in double precision hold up to 2^23 kcal/mol/A)."""
    topos, coords = synthetic.uniform_batch(14, 12, 60, seed=37)
    drop = [(), ("dihedrals",), ("angles",), ("bonds",), ("angles", "dihedrals"),
            ("bonds", "dihedrals"), ("bonds", "angles"), ("bonds", "angles", "dihedrals")]
    cut, xyz = [], []
    for i, (t, x) in enumerate(zip(topos, coords)):
        kinds = drop[i % len(drop)]
        cut.append(without_terms(t, kinds))
        x = np.asarray(x, dtype=np.float64)
        xyz.append(x.mean(axis=0) + 2.0 * (x - x.mean(axis=0)) if "bonds" in kinds else x)
    res = evaluate(cut, xyz, prec)
    check_against_oracle(cut, res, prec)


@pytest.mark.parametrize("prec", ["double", "single"])
def test_separate_launches_equal_fused_launch(prec):
    topos, coords = synthetic.config1(nsys=24, seed=2)
    fused = evaluate(topos, coords, prec)

    def split(p, ag, ctrl, ps, sc, ef, ee):
        S.launch_nonbonded(p, ag, ctrl, ps, sc, ef, ee)
        S.launch_valence(p, ag, ctrl, ps, sc, ef, ee)

    parts = evaluate(topos, coords, prec, launch=split)
    assert np.array_equal(fused["E"], parts["E"])
    assert all(np.array_equal(a, b) for a, b in zip(fused["Ffix"], parts["Ffix"]))


def compressed(coords, factor):
    """Coordinates pulled toward each molecule's centre so that many pairs clash."""
    out = []
    for x in coords:
        x = np.asarray(x, dtype=np.float64)
        c = x.mean(axis=0)
        out.append(c + factor * (x - c))
    return out


@pytest.mark.parametrize("prec", ["double", "single"])
def test_clash_damped_evaluation_matches_oracle(prec):
    """launchNonbonded / launchValence with clash_minimum_distance and clash_ratio
    (src/Potential/hpc_nonbonded_potential.h:131-157) against the oracle's damped potentials, on
    molecules squeezed to 55 % of their size, for combining-rule and stored-table LJ."""
    topos, coords = synthetic.config1(nsys=12, seed=4)
    trp, xt = golden_system("trpcage", igb=IGB_OBC2)
    topos, coords = topos + [trp], coords + [xt]
    coords = compressed(coords, 0.55)
    m = MODES[prec]
    d, ratio = 0.2, 0.5

    def fused(p, ag, ctrl, ps, sc, ef, ee):
        S.launch_mm_evaluation(p, ag, ctrl, ps, sc, ef, ee, clash_distance=d, clash_ratio=ratio)

    def split(p, ag, ctrl, ps, sc, ef, ee):
        S.launch_nonbonded(p, ag, ctrl, ps, sc, ef, ee, clash_distance=d, clash_ratio=ratio)
        S.launch_valence(p, ag, ctrl, ps, sc, ef, ee, clash_distance=d, clash_ratio=ratio)

    res = evaluate(topos, coords, prec, launch=fused)
    parts = evaluate(topos, coords, prec, launch=split)
    assert np.array_equal(res["E"], parts["E"])
    assert all(np.array_equal(a, b) for a, b in zip(res["Ffix"], parts["Ffix"]))
    plain = evaluate(topos, coords, prec)
    worst = 0.0
    for s, t in enumerate(topos):
        nrg, _, f = oracle.eval_mm(t, res["X"][s], True, m["nrg"], clash_distance=d,
                                   clash_ratio=ratio)
        eo = np.array([nrg[k] for k in oracle.SLOT_NAMES])
        eg = res["E"][s].astype(np.float64) * 2.0 ** -m["nrg"]
        scale = np.abs(eo).sum()
        assert np.abs(eg - eo).max() <= m["e_tol"] * scale, (s, np.abs(eg - eo).max() / scale)
        rel = np.sqrt(((res["F"][s] - f) ** 2).sum() / (f ** 2).sum())
        assert rel <= m["f_tol"], (s, rel)
        worst = max(worst, rel)
        # The damping changed the clashing molecule's Lennard-Jones energy.
        assert plain["E"][s][7] != res["E"][s][7]
    print(f"[{prec}] clash-damped evaluation: worst force relative RMS {worst:.2e}")


@pytest.mark.parametrize("prec", ["double", "single"])
def test_nonbonded_and_valence_launches_match_their_terms(prec):
    topos, coords = synthetic.config1(nsys=10, seed=3)
    nb = evaluate(topos, coords, prec, launch=S.launch_nonbonded)
    va = evaluate(topos, coords, prec, launch=S.launch_valence)
    m = MODES[prec]
    for s, t in enumerate(topos):
        for res, terms, slots in [(nb, ["QQ", "LJ", "GB"], [7, 9, 11]),
                                  (va, ["BOND", "ANGLE", "DIHEDRAL", "QQ14", "LJ14"],
                                   [0, 1, 2, 3, 8, 10])]:
            nrg, _, f = oracle.eval_terms(t, res["X"][s], terms, m["nrg"])
            eo = np.array([nrg[oracle.SLOT_NAMES[k]] for k in slots])
            eg = res["E"][s][slots] * 2.0 ** -m["nrg"]
            assert np.abs(eg - eo).max() <= m["e_tol"] * np.abs(eo).sum() + 1e-9
            rel = np.sqrt(((res["F"][s] - f) ** 2).sum() / (f ** 2).sum())
            assert rel <= m["f_tol"]
        # The other slots stay untouched.
        assert np.all(nb["E"][s][[0, 1, 2, 3, 8, 10]] == 0)
        assert np.all(va["E"][s][[7, 9, 11]] == 0)


@pytest.mark.parametrize("prec", ["double", "single"])
def test_reruns_are_bit_identical(prec):
    topos, coords = synthetic.config1(nsys=32, seed=4)
    a = evaluate(topos, coords, prec)
    b = evaluate(topos, coords, prec)
    assert np.array_equal(a["E"], b["E"])
    assert all(np.array_equal(x, y) for x, y in zip(a["Ffix"], b["Ffix"]))


@pytest.mark.parametrize("prec", ["double", "single"])
def test_results_do_not_depend_on_batch_composition(prec):
    topos, coords = synthetic.config1(nsys=30, seed=6)
    mix = evaluate(topos, coords, prec)
    for s in (0, 7, 29):
        alone = evaluate([topos[s]], [coords[s]], prec)
        assert np.array_equal(alone["E"][0], mix["E"][s])
        assert np.array_equal(alone["Ffix"][0], mix["Ffix"][s])
    # Different work-unit shapes pack the systems differently: same results.
    other = evaluate(topos, coords, prec, shape={"wu_atom_capacity": 512,
                                                 "target_wu_per_cu": 1})
    assert np.array_equal(other["E"], mix["E"])
    assert all(np.array_equal(x, y) for x, y in zip(other["Ffix"], mix["Ffix"]))


@pytest.mark.parametrize("prec", ["double", "single"])
@pytest.mark.parametrize("shape", [dict(wu_atom_capacity=256, target_wu_per_cu=4, block_threads=256),
                                   dict(wu_atom_capacity=512, target_wu_per_cu=1, block_threads=1024),
                                   dict(wu_atom_capacity=384, target_wu_per_cu=2, block_threads=512)])
def test_results_do_not_depend_on_block_size(prec, shape):
    # Every accumulation is a fixed-precision integer sum or is made inside one tile task, so the
    # threads per block (and with them the task-to-wavefront assignment) cannot change a bit.
    topos, coords = synthetic.config1(nsys=48, seed=9)
    ref = evaluate(topos, coords, prec)
    other = evaluate(topos, coords, prec, shape=shape)
    assert np.array_equal(other["E"], ref["E"])
    assert all(np.array_equal(x, y) for x, y in zip(other["Ffix"], ref["Ffix"]))


@pytest.mark.parametrize("prec", ["double", "single"])
def test_energy_only_evaluation_matches_force_evaluation_energies(prec):
    topos, coords = synthetic.config1(nsys=16, seed=7)
    a = evaluate(topos, coords, prec, eval_force=True)
    b = evaluate(topos, coords, prec, eval_force=False)
    assert np.array_equal(a["E"], b["E"])
    assert all(np.all(f == 0) for f in b["Ffix"])


def test_scorecard_commit_accumulates_history():
    topos, coords = synthetic.config1(nsys=8, seed=8)
    ag = S.AtomGraphSynthesis(topos)
    ps = S.PhaseSpaceSynthesis(ag)
    sc = S.ScoreCard(ag.system_count, 26)
    ps.upload(coords)
    for _ in range(3):
        sc.initialize()
        S.launch_mm_evaluation("single", ag, None, ps, sc, False, True)
        sc.commit()
    run, sq, n = sc.history()
    inst = sc.accumulators()[:, :13] * 2.0 ** -26
    assert n == 3
    np.testing.assert_allclose(run[:, :13], 3 * inst, rtol=1e-12)
    np.testing.assert_allclose(run[:, 28], 3 * inst.sum(axis=1), rtol=1e-12)
    np.testing.assert_allclose(sq[:, :13], 3 * inst ** 2, rtol=1e-12)


def test_mismatched_scorecard_is_rejected():
    topos, coords = synthetic.config1(nsys=4, seed=9)
    ag = S.AtomGraphSynthesis(topos)
    ps = S.PhaseSpaceSynthesis(ag)
    sc = S.ScoreCard(3, 26)
    with pytest.raises(S._lib.StormmError, match="not prepared for the number of systems"):
        S.launch_mm_evaluation("single", ag, None, ps, sc, True, True)


def test_autotuner_returns_a_valid_shape_that_gives_the_same_results():
    topos, coords = synthetic.config1(nsys=64, seed=23)
    shape, ms = S.autotune(topos, "single", coords)
    assert ms > 0.0
    assert shape["block_threads"] in (256, 512, 1024)
    assert shape["wu_atom_capacity"] in (128, 256, 384, 512)
    ref = evaluate(topos, coords, "single")
    tuned = evaluate(topos, coords, "single", shape=shape)
    assert np.array_equal(tuned["E"], ref["E"])
    assert all(np.array_equal(x, y) for x, y in zip(tuned["Ffix"], ref["Ffix"]))
