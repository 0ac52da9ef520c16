"""The FP64 oracle against the reference's own known answers and force snapshots.

Known answers: test/Potential/test_generalized_born.cpp:290-292 (HCT / OBC / OBC-II GB energies),
test/Potential/test_valence_evaluation.cpp:421-429 (valence energies) and
test/Potential/test_nonbonded_evaluation.cpp:1259-1279 (non-bonded and 1:4 energies, read from the
fixed-precision ScoreCard accumulators).  Snapshots: test/Potential/{trpcage,ala_dipeptide}_*.m,
compared at the reference's snapshot tolerance of 1e-6.
"""
import numpy as np
import pytest

import oracle
from conftest import golden_system
from stormm_amd.topology import IGB_HCT, IGB_OBC, IGB_OBC2

# The GB test builds its topologies with a 2^-30 e charge grid and a 1e-10 rounding tolerance
# (test_generalized_born.cpp:229-249), which leaves the prmtop charges as read.
GB_CHARGES = dict(charge_rounding_tol=1.0e-10,
                  charge_discretization=9.31322574615478515625e-10)
SYSTEMS = [("trpcage", "trpcage", "trpcage"), ("ala_dipeptide", "ala", "ala_dipeptide")]


@pytest.mark.parametrize("igb,tag", [(IGB_HCT, "gb_hct"), (IGB_OBC, "gb_obc"),
                                     (IGB_OBC2, "gb_obc2")])
@pytest.mark.parametrize("sysname,key,_", SYSTEMS)
def test_gb_energy_known_answers(known_answers, igb, tag, sysname, key, _):
    topo, xyz = golden_system(sysname, igb=igb, **GB_CHARGES)
    e, acc, f, psi, radii, sdi = oracle.eval_gb(topo, xyz, True, 32)
    ka = known_answers[tag]
    assert abs(acc * 2.0 ** -32 - ka[key]) <= ka["margin"]
    assert abs(e - ka[key]) <= ka["margin"]


@pytest.mark.parametrize("igb,tag", [(IGB_HCT, "hct"), (IGB_OBC, "obc"), (IGB_OBC2, "obc2")])
@pytest.mark.parametrize("sysname,key,_", SYSTEMS)
def test_gb_force_snapshots(snapshots, igb, tag, sysname, key, _):
    topo, xyz = golden_system(sysname, igb=igb, **GB_CHARGES)
    f = oracle.eval_gb(topo, xyz, True, 32)[2]
    ref = snapshots[f"{key}_{tag}_gb_frc"]
    np.testing.assert_allclose(f.reshape(-1), ref, rtol=0, atol=1.0e-6)


@pytest.mark.parametrize("sysname,key,_", SYSTEMS)
def test_valence_known_answers(known_answers, sysname, key, _):
    topo, xyz = golden_system(sysname)
    nrg, acc, f = oracle.eval_mm(topo, xyz, True, 26)
    got = [nrg["BOND"], nrg["ANGLE"], nrg["PROPER_DIHEDRAL"], nrg["IMPROPER_DIHEDRAL"]]
    ka = known_answers["valence"]
    np.testing.assert_allclose(got, ka[key], rtol=0, atol=ka["margin"])


@pytest.mark.parametrize("sysname,key,_", SYSTEMS)
def test_nonbonded_known_answers(known_answers, sysname, key, _):
    topo, xyz = golden_system(sysname)
    nrg, acc, f = oracle.eval_mm(topo, xyz, True, 26)
    scale = 2.0 ** -26
    got = [acc["ELEC_ONE_FOUR"] * scale, acc["VDW_ONE_FOUR"] * scale,
           acc["ELECTROSTATIC"] * scale, acc["VDW"] * scale]
    ka = known_answers["nonbonded"]
    margin = ka["margin_trpcage"] if key == "trpcage" else ka["margin_ala"]
    np.testing.assert_allclose(got, ka[key], rtol=0, atol=margin)


TERM_SNAPSHOTS = {
    "trpcage": {"QQ14": "trpcage_qq14_forces", "LJ14": "trpcage_lj14_forces",
                "QQ": "trpcage_qq_forces", "LJ": "trpcage_lj_forces", "BOND": "trpcage_bond",
                "ANGLE": "trpcage_angl", "DIHEDRAL": "trpcage_dihe"},
    "ala_dipeptide": {"QQ14": "ala_dipeptide_qq14_forces", "LJ14": "ala_dipeptide_lj14_forces",
                      "QQ": "ala_dipeptide_qq_forces", "LJ": "ala_dipeptide_lj_forces",
                      "BOND": "ala_dipeptide_bond_forces", "ANGLE": "ala_dipeptide_angl_forces",
                      "DIHEDRAL": "ala_dipeptide_dihe_forces"},
}


@pytest.mark.parametrize("sysname", ["trpcage", "ala_dipeptide"])
@pytest.mark.parametrize("term", ["QQ14", "LJ14", "QQ", "LJ", "BOND", "ANGLE", "DIHEDRAL"])
def test_term_force_snapshots(snapshots, sysname, term):
    topo, xyz = golden_system(sysname)
    _, _, f = oracle.eval_terms(topo, xyz, [term])
    ref = snapshots[TERM_SNAPSHOTS[sysname][term]]
    np.testing.assert_allclose(f.reshape(-1), ref, rtol=0, atol=1.0e-6)


def test_charge_smoothing_is_applied_for_default_intake():
    """The default intake rounds charges to the 2^-23 grid with an integral net charge."""
    topo, _ = golden_system("trpcage")
    grid = 2.0 ** -23
    assert np.allclose(np.round(topo.charge / grid) * grid, topo.charge, rtol=0, atol=1e-15)
    assert abs(topo.charge.sum() - round(topo.charge.sum())) < 1.0e-9
