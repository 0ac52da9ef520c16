"""Host-side logic: topology intake and output, the synthetic generator, the C-ABI library's
exports and the work-unit planner (no device calls)."""
import dataclasses
import math
import os

import numpy as np
import pytest

import oracle
from conftest import golden_system
from stormm_amd import _lib, synthetic
from stormm_amd import synthesis as S
from stormm_amd.topology import (IGB_OBC2, read_inpcrd, read_prmtop, smooth_charges,
                                 write_inpcrd, write_prmtop)


def test_library_loads_and_exports_every_declared_symbol():
    lib = _lib.load()
    names = _lib.declared_symbols()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []


def test_system_transfers_reject_missing_handles_without_a_device():
    # The per-system and ranged phase-space transfers check their arguments on the host
    # before any device call.
    lib = _lib.load()
    buf = np.zeros((16, 3))
    calls = [lambda: lib.stormm_phasespace_upload_systems(None, 0, 1, buf.ctypes.data),
             lambda: lib.stormm_phasespace_download_systems(None, 0, 1, buf.ctypes.data, None),
             lambda: lib.stormm_phasespace_import_system(None, 0, buf.ctypes.data),
             lambda: lib.stormm_phasespace_export_system(None, 0, buf.ctypes.data, None)]
    for call in calls:
        assert call() != 0
        assert b"null" in lib.stormm_last_error()


def test_prmtop_round_trip_preserves_energies(tmp_path):
    topo, xyz = synthetic.branched_molecule(37, seed=3, gb=True)
    top = os.path.join(tmp_path, "mol.top")
    crd = os.path.join(tmp_path, "mol.inpcrd")
    write_prmtop(topo, top)
    write_inpcrd(xyz, crd)
    back = read_prmtop(top, igb=IGB_OBC2)
    xyz2 = read_inpcrd(crd)
    assert back.natom == topo.natom
    np.testing.assert_array_equal(back.exclusion_pairs(), topo.exclusion_pairs())
    # Charges survive the x18.2223 / /18.2223 round trip and charge smoothing to ~1e-8 e.
    np.testing.assert_allclose(back.charge, topo.charge, atol=2e-7)
    e1 = oracle.eval_mm(topo, xyz)[0]
    e2 = oracle.eval_mm(back, xyz2)[0]
    for k in oracle.SLOT_NAMES:
        assert e2[k] == pytest.approx(e1[k], rel=2e-6, abs=2e-5), k


def test_exclusions_follow_shortest_bond_paths():
    # 0-1-2-3-4 chain with a branch 2-5: 1:2, 1:3 and 1:4 pairs only.
    topo, _ = synthetic.branched_molecule(6, seed=1, gb=False)
    topo.bonds = np.array([[0, 1], [1, 2], [2, 3], [3, 4], [2, 5]], dtype=np.int32)
    topo._excl_cache = None
    pairs = {tuple(p) for p in topo.exclusion_pairs()}
    assert (0, 4) not in pairs  # four bonds apart
    for p in [(0, 1), (0, 2), (0, 3), (0, 5), (1, 4), (3, 5), (4, 5), (1, 5)]:
        assert p in pairs


def test_generated_charges_are_grid_exact_and_neutral():
    topo, _ = synthetic.branched_molecule(53, seed=12)
    grid = 2.0 ** -23
    assert np.array_equal(np.round(topo.charge / grid) * grid, topo.charge)
    assert topo.charge.sum() == 0.0
    # The reference's smoothing leaves such charges unchanged.
    np.testing.assert_array_equal(smooth_charges(topo.charge), topo.charge)


def test_generated_parameters_follow_the_workload_ranges():
    topo, xyz = synthetic.branched_molecule(120, seed=21)
    assert np.all((topo.bond_k >= 200) & (topo.bond_k <= 400))
    assert np.all((topo.bond_r0 >= 1.0) & (topo.bond_r0 <= 1.55))
    assert np.all((topo.angle_k >= 40) & (topo.angle_k <= 80))
    assert np.all(np.isin(topo.dihe_freq, [1.0, 2.0, 3.0]))
    assert np.all((topo.dihe_amp >= 0) & (topo.dihe_amp <= 2))
    assert np.all((topo.pb_radii >= 1.2) & (topo.pb_radii <= 2.0))
    assert np.all((topo.gb_screen >= 0.7) & (topo.gb_screen <= 0.9))
    assert topo.lorentz_berthelot() is not None
    # Every proper torsion carries its 1:4 pair with the 1.2 / 2.0 attenuation.
    assert np.all(topo.dihe_scee == 1.2) and np.all(topo.dihe_scnb == 2.0)
    d = np.linalg.norm(xyz[:, None] - xyz[None], axis=-1)
    excl = topo.exclusion_matrix().astype(bool) | np.eye(topo.natom, dtype=bool)
    assert d[~excl].min() > 1.5


def test_amber_tables_are_not_combining_rule():
    topo, _ = golden_system("trpcage")
    assert topo.lorentz_berthelot() is None


def test_plan_counts_pairs_tasks_and_terms():
    topos, _ = synthetic.config1(nsys=40, seed=5)
    info = S.plan_synthesis(topos)
    assert info.system_count == 40
    assert info.pair_count == sum(t.natom * (t.natom - 1) // 2 for t in topos)
    tiles = [math.ceil(t.natom / 16) for t in topos]
    assert info.task_count == sum(n * (n + 1) // 2 for n in tiles)
    assert info.valence_term_count == sum(t.valence_term_count() for t in topos)
    assert info.lj_table_systems == 0
    assert 1 <= info.work_unit_count <= 40
    assert info.max_wu_atoms <= 512


@pytest.mark.parametrize("cap", [128, 256, 512])
def test_plan_respects_work_unit_capacity(cap):
    topos, _ = synthetic.uniform_batch(60, 16, 120, seed=8)
    info = S.plan_synthesis(topos, {"wu_atom_capacity": cap, "target_wu_per_cu": 1})
    assert info.max_wu_atoms <= cap
    assert info.padded_atom_count == sum(32 * math.ceil(t.natom / 32) for t in topos)


def test_plan_rejects_oversize_systems_and_bad_models():
    big, _ = synthetic.branched_molecule(300, seed=2)
    with pytest.raises(_lib.StormmError, match="work-unit capacity"):
        S.plan_synthesis([big], {"wu_atom_capacity": 256})
    topo, _ = synthetic.branched_molecule(20, seed=2)
    topo.igb = 7
    with pytest.raises(_lib.StormmError, match="implicit solvent"):
        S.plan_synthesis([topo])
    with pytest.raises(ValueError):
        t2, _ = synthetic.branched_molecule(20, seed=3, gb=False)
        t3, _ = synthetic.branched_molecule(20, seed=4, gb=True)
        S.plan_synthesis([t2, t3])


def test_plan_rejects_fractional_dihedral_periodicity_and_bad_blocks():
    topo, _ = synthetic.branched_molecule(30, seed=2)
    assert S.plan_synthesis([topo]).valence_term_count > 0
    frac = dataclasses.replace(topo, dihe_freq=np.where(np.arange(len(topo.dihe_freq)) == 0, 2.5,
                                                        topo.dihe_freq))
    with pytest.raises(_lib.StormmError, match="periodicity"):
        S.plan_synthesis([frac])
    with pytest.raises(_lib.StormmError, match="block_threads"):
        S.plan_synthesis([topo], {"block_threads": 384})
    for bt in (256, 512, 1024):
        assert S.plan_synthesis([topo], {"block_threads": bt}).system_count == 1


def test_table_and_combining_rule_systems_plan_together():
    trp, _ = golden_system("trpcage", igb=IGB_OBC2)
    syn, _ = synthetic.branched_molecule(50, seed=6, gb=True)
    info = S.plan_synthesis([trp, syn, syn])
    assert info.lj_table_systems == 1
    assert info.work_unit_count >= 2
