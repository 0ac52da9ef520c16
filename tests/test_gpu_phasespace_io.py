"""Per-system and ranged coordinate transfers of the PhaseSpaceSynthesis (importSystem /
exportSystem and upload / download over a system range, src/Synthesis/phasespace_synthesis.cpp:
2327-2333, 2379-2385, 2472-, 3410-3436).

Transfers are integer copies of the fixed-precision arrays, so every check is bit-exact: a
synthesis whose systems were replaced one at a time or by range must hold, and evaluate to,
exactly what a synthesis uploaded whole with the same coordinates holds.
"""
import numpy as np
import pytest

from stormm_amd import synthetic
from stormm_amd import synthesis as S
from stormm_amd._lib import StormmError

pytestmark = pytest.mark.gpu


def _evaluate(ag, ps, prec="single"):
    sc = S.ScoreCard(ag.system_count, 26)
    ps.initialize_forces()
    sc.initialize()
    S.launch_mm_evaluation(prec, ag, None, ps, sc, True, True)
    assert ag.status() == 0
    return sc.accumulators()[:, :13].copy()


def test_imports_and_ranged_uploads_equal_a_whole_upload():
    topos, coords = synthetic.config1(nsys=24, seed=11)
    rng = np.random.default_rng(3)
    moved = [c + rng.normal(0.0, 0.05, c.shape) for c in coords]
    ag = S.AtomGraphSynthesis(topos)

    whole = S.PhaseSpaceSynthesis(ag, 32, 23)
    whole.upload(moved)
    e_whole = _evaluate(ag, whole)
    x_whole, f_whole = whole.fixed()

    parts = S.PhaseSpaceSynthesis(ag, 32, 23)
    parts.upload(coords)
    parts.import_system(0, moved[0])
    parts.upload_systems(1, 9, moved[1:9])
    for s in range(9, 23):
        parts.import_system(s, moved[s])
    parts.upload_systems(23, 24, moved[23:])
    e_parts = _evaluate(ag, parts)
    x_parts, f_parts = parts.fixed()

    assert np.array_equal(e_parts, e_whole)
    assert all(np.array_equal(a, b) for a, b in zip(x_parts, x_whole))
    assert all(np.array_equal(a, b) for a, b in zip(f_parts, f_whole))


def test_exports_and_ranged_downloads_match_the_whole_download():
    topos, coords = synthetic.config1(nsys=16, seed=12)
    ag = S.AtomGraphSynthesis(topos)
    ps = S.PhaseSpaceSynthesis(ag, 32, 23)
    ps.upload(coords)
    _evaluate(ag, ps)
    X, F = ps.coordinates(), ps.forces()
    for s in (0, 5, 15):
        x, f = ps.export_system(s)
        assert np.array_equal(x, X[s]) and np.array_equal(f, F[s])
    xr, fr = ps.download_systems(3, 11)
    assert len(xr) == 8
    for k, s in enumerate(range(3, 11)):
        assert np.array_equal(xr[k], X[s]) and np.array_equal(fr[k], F[s])


def test_an_import_leaves_the_other_systems_alone():
    topos, coords = synthetic.config1(nsys=8, seed=13)
    ag = S.AtomGraphSynthesis(topos)
    ps = S.PhaseSpaceSynthesis(ag, 32, 23)
    ps.upload(coords)
    before, _ = ps.fixed()
    ps.import_system(4, coords[4] + 1.0)
    after, _ = ps.fixed()
    for s in range(8):
        if s == 4:
            assert np.array_equal(after[s] - before[s], np.full_like(before[s], 1 << 32))
        else:
            assert np.array_equal(after[s], before[s])


def test_bad_ranges_are_rejected():
    topos, coords = synthetic.config1(nsys=4, seed=14)
    ag = S.AtomGraphSynthesis(topos)
    ps = S.PhaseSpaceSynthesis(ag, 32, 23)
    ps.upload(coords)
    for lo, hi in ((2, 2), (-1, 1), (3, 5)):
        with pytest.raises(ValueError):
            ps.download_systems(lo, hi)
    with pytest.raises(ValueError):
        ps.import_system(0, coords[1] if topos[1].natom != topos[0].natom
                         else np.zeros((topos[0].natom + 1, 3)))
    # The C-ABI checks ranges itself.
    L = ps._L
    buf = np.zeros((4096, 3))
    rc = L.stormm_phasespace_download_systems(ps.handle, 3, 9, buf.ctypes.data, None)
    assert rc != 0
    with pytest.raises(StormmError):
        S.check(rc)
