"""Build the golden fixtures that pin the oracle to the reference's own test data.

Run in a container that holds the reference tree (default /root/reference); the GPU box never
runs this.  It copies four input data files the reference's tests read (AMBER topologies and
coordinates of Trp-cage and alanine dipeptide) and extracts the per-atom force snapshots that the
reference's tests compare against from the Octave-format ``.m`` files, writing them to
``reference_snapshots.npz``.  The known-answer energies are transcribed in ``KNOWN_ANSWERS`` with
the test file and line each comes from, and written to ``known_answers.json``.

Usage: python tests/golden/make_golden.py [/path/to/reference]
"""
import json
import os
import re
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

INPUT_FILES = [
    ("test/Topology/trpcage.top", "trpcage.top"),
    ("test/Trajectory/trpcage.inpcrd", "trpcage.inpcrd"),
    ("test/Topology/ala_dipeptide.top", "ala_dipeptide.top"),
    ("test/Trajectory/ala_dipeptide.inpcrd", "ala_dipeptide.inpcrd"),
    # Three of the four small molecules of the reference's GPU minimization test
    # (test/MolecularMechanics/test_hpc_minimization.cu:1042-1051); the fourth, alanine dipeptide,
    # carries CMAP terms, which this path does not evaluate.
    ("test/Topology/bromobenzene_iso.top", "bromobenzene_iso.top"),
    ("test/Trajectory/bromobenzene_iso.inpcrd", "bromobenzene_iso.inpcrd"),
    ("test/Topology/stereo_L1.top", "stereo_L1.top"),
    ("test/Trajectory/stereo_L1.inpcrd", "stereo_L1.inpcrd"),
    ("test/Topology/symmetry_L1.top", "symmetry_L1.top"),
    ("test/Trajectory/symmetry_L1.inpcrd", "symmetry_L1.inpcrd"),
]

# (snapshot file, variable) pairs taken from the reference's test data.
SNAPSHOTS = [
    ("test/Potential/trpcage_gb_forces.m", "trpcage_hct_gb_frc"),
    ("test/Potential/trpcage_gb_forces.m", "trpcage_obc_gb_frc"),
    ("test/Potential/trpcage_gb_forces.m", "trpcage_obc2_gb_frc"),
    ("test/Potential/ala_dipeptide_gb_forces.m", "ala_hct_gb_frc"),
    ("test/Potential/ala_dipeptide_gb_forces.m", "ala_obc_gb_frc"),
    ("test/Potential/ala_dipeptide_gb_forces.m", "ala_obc2_gb_frc"),
    ("test/Potential/trpcage_nb_details.m", "trpcage_qq14_forces"),
    ("test/Potential/trpcage_nb_details.m", "trpcage_lj14_forces"),
    ("test/Potential/trpcage_nb_details.m", "trpcage_qq_forces"),
    ("test/Potential/trpcage_nb_details.m", "trpcage_lj_forces"),
    ("test/Potential/ala_dipeptide_nb_details.m", "ala_dipeptide_qq14_forces"),
    ("test/Potential/ala_dipeptide_nb_details.m", "ala_dipeptide_lj14_forces"),
    ("test/Potential/ala_dipeptide_nb_details.m", "ala_dipeptide_qq_forces"),
    ("test/Potential/ala_dipeptide_nb_details.m", "ala_dipeptide_lj_forces"),
    ("test/Potential/trpcage_details.m", "trpcage_bond"),
    ("test/Potential/trpcage_details.m", "trpcage_angl"),
    ("test/Potential/trpcage_details.m", "trpcage_dihe"),
    ("test/Potential/ala_dipeptide_details.m", "ala_dipeptide_bond_forces"),
    ("test/Potential/ala_dipeptide_details.m", "ala_dipeptide_angl_forces"),
    ("test/Potential/ala_dipeptide_details.m", "ala_dipeptide_dihe_forces"),
]

# Known-answer energies (kcal/mol) and the margins the reference's tests apply.
KNOWN_ANSWERS = {
    # test/Potential/test_generalized_born.cpp:290-292, margin 1.0e-4 (:301-351)
    "gb_hct": {"trpcage": -313.0368696, "ala": -14.7193347, "margin": 1.0e-4},
    "gb_obc": {"trpcage": -320.6017219, "ala": -16.0329696, "margin": 1.0e-4},
    "gb_obc2": {"trpcage": -298.2900710, "ala": -14.9957762, "margin": 1.0e-4},
    # test/Potential/test_valence_evaluation.cpp:421-429 (bond, angle, proper, improper)
    "valence": {"trpcage": [12.38440829, 69.48736586, -142.41607189, 0.77975874],
                "ala": [0.38340254, 0.54479669, 2.42761153, 0.00708592], "margin": 1.0e-6},
    # test/Potential/test_nonbonded_evaluation.cpp:1259-1279 (qq14, lj14, qq, lj)
    "nonbonded": {"trpcage": [1458.0998129, 62.6481811, -1816.1478854, -119.3500215],
                  "ala": [46.8072425, 3.1589264, -78.8463310, -1.2452480],
                  "margin_trpcage": 2.0e-5, "margin_ala": 1.0e-5},
    # Final energies of the reference's GPU minimizations, test/MolecularMechanics/min_energy.m
    # (written by test_hpc_minimization.cu:840-846, snapshot margin 1e-5): "d" is FP64 with 40
    # position / 40 force / 32 energy bits, 100 cycles of which 10 steepest descent; "f" is FP32
    # with 28 / 24 / 32 bits, 500 cycles of which 50 steepest descent (:1024-1029, :400-404), no
    # clash damping.  The ScoreCard holds the energy of the last energy-only evaluation (the
    # third trial point of the last cycle).  small_mol_* lines 2-592; trp_cage_* 1200-1228;
    # folded_pro_d 1184-1190 and vste_d 1334-1347 hold Trp-cage minimized in other batches.
    "minimization": {
        "bromobenzene_iso": {"d": 2.67090431, "f": 2.66019999},
        "stereo_L1": {"d": 148.14839914, "f": 144.07275783},
        "symmetry_L1": {"d": 1.13784720, "f": 0.35714461},
        "trpcage": {"d": -534.84286089, "f": -713.76852424,
                    "d_other_batches": [-534.83985375, -534.8423845]},
        "margin": 1.0e-5,
    },
}


def read_snapshot(path, var):
    with open(path, "r") as fh:
        text = fh.read()
    m = re.search(re.escape(var) + r" = \[(.*?)\];", text, re.S)
    if m is None:
        raise KeyError(f"{var} not in {path}")
    vals = np.array([float(v) for v in m.group(1).split()], dtype=np.float64)
    m2 = re.search(re.escape(var) + r" = " + re.escape(var) + r"\(1:(\d+)\);", text)
    if m2 is not None:
        vals = vals[:int(m2.group(1))]
    return vals


def main(ref_root):
    os.makedirs(os.path.join(HERE, "inputs"), exist_ok=True)
    for src, dst in INPUT_FILES:
        shutil.copyfile(os.path.join(ref_root, src), os.path.join(HERE, "inputs", dst))
    arrays = {}
    for src, var in SNAPSHOTS:
        arrays[var] = read_snapshot(os.path.join(ref_root, src), var)
    np.savez_compressed(os.path.join(HERE, "reference_snapshots.npz"), **arrays)
    with open(os.path.join(HERE, "known_answers.json"), "w") as fh:
        json.dump(KNOWN_ANSWERS, fh, indent=1, sort_keys=True)
    print(f"wrote {len(arrays)} snapshot arrays and {len(INPUT_FILES)} input files")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
