"""The launch entry points on caller-owned device arrays (the reference's abstract layouts:
PsSynthesisWriter, ScoreCardWriter, LinMinWriter, MMControlKit) and the 95-bit split
fixed-precision path (coordinates above 38 bits, forces above 40 bits; reference
src/Numerics/accumulation.cui:121-403, src/Synthesis/phasespace_synthesis.h:162,174).

Arrays are allocated here with hipMalloc, filled with hipMemcpy and handed over as raw pointers,
as a STORMM caller holding its own PhaseSpaceSynthesis / ScoreCard would.  The kit entry points
must give the handle entry points' results bit for bit; 40 / 40-bit double precision (the
reference's own minimization set-up, test_hpc_minimization.cu:1024-1026) must track the oracle.
"""
import ctypes

import numpy as np
import pytest

import oracle
from stormm_amd import _lib, synthetic
from stormm_amd import synthesis as S

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                   ctypes.c_int]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


class DevArray:
    """A hipMalloc'd copy of a numpy array (host to device kind 1, device to host kind 2)."""

    def __init__(self, host: np.ndarray):
        self.host = np.ascontiguousarray(host)
        self.ptr = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(self.ptr), max(self.host.nbytes, 8)) == 0
        assert hip().hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, 1) == 0

    def get(self) -> np.ndarray:
        out = np.empty_like(self.host)
        assert hip().hipDeviceSynchronize() == 0
        assert hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0
        return out

    def __del__(self):
        if self.ptr:
            hip().hipFree(self.ptr)


class CallerState:
    """A caller's own PsSynthesisWriter / ScoreCardWriter / LinMinWriter arrays for a synthesis,
    in its atom layout, starting from fixed-precision coordinates."""

    def __init__(self, ag, gpos, frc, nrg, fixed_xyz, ovrf_xyz=None, dx0=0.01):
        L = _lib.load()
        nsys = ag.system_count
        starts = np.zeros(nsys, dtype=np.int32)
        counts = np.zeros(nsys, dtype=np.int32)
        L.stormm_synthesis_atom_layout(ag.handle, starts.ctypes.data, counts.ctypes.data)
        npad = int(ag.info.padded_atom_count)
        self.starts, self.counts, self.npad = starts, counts, npad
        self.arrays = {}

        def padded(per_system, dtype):
            a = np.zeros(npad, dtype=dtype)
            for s in range(nsys):
                a[starts[s]:starts[s] + counts[s]] = per_system[s]
            return a

        for d, axis in enumerate("xyz"):
            self.arrays[f"{axis}crd"] = DevArray(padded([c[:, d] for c in fixed_xyz], np.int64))
            ov = ([o[:, d] for o in ovrf_xyz] if ovrf_xyz is not None
                  else [np.zeros(n, np.int32) for n in counts])
            self.arrays[f"{axis}crd_ovrf"] = DevArray(padded(ov, np.int32))
            for name in ("frc", "prv", "cg_temp"):
                key = f"{axis}{name}" if name != "cg_temp" else f"{axis}_cg_temp"
                self.arrays[key] = DevArray(np.zeros(npad, np.int64))
                self.arrays[key + "_ovrf"] = DevArray(np.zeros(npad, np.int32))
        self.arrays["atom_starts"] = DevArray(starts)
        self.arrays["atom_counts"] = DevArray(counts)
        w = _lib.PsWriter()
        w.system_count, w.gpos_bits, w.frc_bits = nsys, gpos, frc
        w.gpos_scale, w.inv_gpos_scale = 2.0 ** gpos, 2.0 ** -gpos
        w.frc_scale, w.inv_frc_scale = 2.0 ** frc, 2.0 ** -frc
        for name, _ in _lib.PsWriter._fields_:
            if name in self.arrays:
                setattr(w, name, self.arrays[name].ptr.value)
        self.psw = w
        self.acc = DevArray(np.zeros((nsys, 32), dtype=np.int64))
        self.scw = _lib.ScoreCardWriter(nsys, 32, nrg, self.acc.ptr.value)
        line = np.zeros((9, nsys))
        line[0] = dx0
        self.line = DevArray(line)
        self.lmw = _lib.LinMinWriter(nsys, *[self.line.ptr.value + 8 * nsys * j
                                             for j in range(9)])

    def values(self, name):
        """Exact per-system values (Python ints) of a word / overflow-word array pair."""
        w = [self.arrays[f"{a}{name}"].get() for a in "xyz"]
        o = [self.arrays[f"{a}{name}_ovrf"].get() for a in "xyz"]
        out = []
        for s in range(len(self.starts)):
            lo, n = self.starts[s], self.counts[s]
            cols = [(o[d][lo:lo + n].astype(object) << 63) + w[d][lo:lo + n].astype(object)
                    for d in range(3)]
            out.append(np.stack(cols, axis=1))
        return out


def check(rc):
    _lib.check(rc)


def split_words(values, wide):
    """Per-system exact values as (word, overflow word) arrays: above the plain limit value =
    overflow 2^63 + word with the word in [0, 2^63); below it the word holds the value."""
    lo, hi = [], []
    for v in values:
        if wide:
            lo.append(np.array([[int(x) & ((1 << 63) - 1) for x in row] for row in v],
                               dtype=np.int64))
            hi.append(np.array([[int(x) >> 63 for x in row] for row in v], dtype=np.int32))
        else:
            lo.append(np.array([[int(x) for x in row] for row in v], dtype=np.int64))
            hi.append(np.zeros((len(v), 3), dtype=np.int32))
    return lo, hi


def initial_values(topos, coords, gpos, frc):
    ag = S.AtomGraphSynthesis(topos)
    ps = S.PhaseSpaceSynthesis(ag, gpos, frc)
    ps.upload(coords)
    return ps.fixed_values()[0]


def handle_run(topos, coords, prec, gpos, frc, nrg, minimize=None):
    ag = S.AtomGraphSynthesis(topos)
    ps = S.PhaseSpaceSynthesis(ag, gpos, frc)
    sc = S.ScoreCard(ag.system_count, nrg)
    ps.upload(coords)
    if minimize is None:
        ps.initialize_forces()
        sc.initialize()
        S.launch_mm_evaluation(prec, ag, None, ps, sc, True, True)
    else:
        cycles, mode = minimize
        lm = S.LineMinimization(ag.system_count, 0.01)
        S.launch_minimization(prec, ag, ps, S.minimize_controls(cycles, 5, 0.01, 0.0),
                              sc, lm, persistent=mode)
    c, f = ps.fixed_values()
    return ag, ps, c, f, sc.accumulators()


@pytest.mark.parametrize("prec,bits", [("single", (32, 23, 26)), ("double", (32, 40, 32)),
                                       ("double", (40, 40, 32)), ("double", (60, 60, 32))])
def test_kit_evaluation_on_caller_arrays_equals_the_handle_path(prec, bits):
    gpos, frc, nrg = bits
    topos, coords = synthetic.uniform_batch(7, 24, 140, seed=71, gb=True)
    ag, ps, c_h, f_h, e_h = handle_run(topos, coords, prec, gpos, frc, nrg)
    lo, hi = split_words(c_h, gpos > 38)
    st = CallerState(ag, gpos, frc, nrg, lo, hi)
    L = _lib.load()
    check(L.stormm_check_ps_writer(ag.handle, ctypes.byref(st.psw)))
    check(L.stormm_ps_writer_initialize_forces(ctypes.byref(st.psw), st.npad, None))
    check(L.stormm_scorecard_writer_initialize(ctypes.byref(st.scw), None))
    check(L.stormm_launch_mm_evaluation_kit(_lib.PREC_DOUBLE if prec == "double"
                                            else _lib.PREC_SINGLE, ag.handle, None,
                                            ctypes.byref(st.psw), ctypes.byref(st.scw), 1, 1,
                                            0.0, 0.0, None))
    f_k = st.values("frc")
    print(f"[{prec} {bits}] {len(topos)} systems, {sum(t.natom for t in topos)} atoms")
    assert all(np.array_equal(a, b) for a, b in zip(f_k, f_h))
    assert np.array_equal(st.acc.get(), e_h)
    if frc > 40:
        assert any(np.abs(f.astype(float)).max() > 2.0 ** 63 for f in f_k), \
            "forces never used their overflow words"


@pytest.mark.parametrize("mode", [0, 2])
@pytest.mark.parametrize("bits", [(40, 40, 32), (60, 60, 32)])
def test_kit_minimization_on_caller_arrays_equals_the_handle_path(mode, bits):
    gpos, frc, nrg = bits
    topos, coords = synthetic.config2(nsys=6, seed=72)
    ag, ps, c_h, f_h, e_h = handle_run(topos, coords, "double", gpos, frc, nrg,
                                       minimize=(12, mode))
    ag2 = S.AtomGraphSynthesis(topos)
    lo, hi = split_words(initial_values(topos, coords, gpos, frc), gpos > 38)
    st = CallerState(ag2, gpos, frc, nrg, lo, hi)
    mc = S.minimize_controls(12, 5, 0.01, 0.0)
    check(_lib.load().stormm_launch_minimization_kit(
        _lib.PREC_DOUBLE, ag2.handle, ctypes.byref(st.psw), ctypes.byref(mc),
        ctypes.byref(st.scw), ctypes.byref(st.lmw), mode, None))
    c_k = st.values("crd")
    assert all(np.array_equal(a, b) for a, b in zip(c_k, c_h))
    assert np.array_equal(st.acc.get(), e_h)


def test_reference_minimization_bits_track_the_oracle():
    """FP64 with 40 coordinate and 40 force bits and 32 energy bits, 10 steepest-descent cycles
    (test_hpc_minimization.cu:400-411,1024-1026): the device trajectory follows the oracle's to
    1e-6 A over 30 cycles, every mode."""
    topos, coords = synthetic.config2(nsys=6, seed=73)
    gpos, frc, nrg = 40, 40, 32
    for mode in (0, 2):
        ag = S.AtomGraphSynthesis(topos)
        ps = S.PhaseSpaceSynthesis(ag, gpos, frc)
        sc = S.ScoreCard(ag.system_count, nrg)
        lm = S.LineMinimization(ag.system_count, 0.01)
        ps.upload(coords)
        S.launch_minimization("double", ag, ps, S.minimize_controls(30, 10, 0.01, 0.0,
                                                                   clash_cycles=0),
                              sc, lm, persistent=mode)
        xf, _ = ps.fixed_values()
        e = sc.potential_energies()
        for s, t in enumerate(topos):
            fixed = np.rint(np.asarray(coords[s]) * 2.0 ** gpos).astype(np.int64)
            out, trace = oracle.minimize(t, fixed, gpos, frc, nrg, 30, 10, 0.01, dp_mode=True)
            drift = np.abs(xf[s].astype(float) - out.astype(float)).max() * 2.0 ** -gpos
            assert drift < 1e-6, (mode, s, drift)
            assert e[s] == pytest.approx(trace[-1], rel=1e-9, abs=1e-7)


def test_wide_force_bits_store_the_forty_bit_sums_shifted():
    """Above 40 bits the work units accumulate at 40 bits and store the sums shifted to the
    stored precision: forces at 60 bits are exactly 2^20 times those at 40 bits, and the
    evaluation matches the oracle at the device coordinates."""
    topos, coords = synthetic.uniform_batch(5, 30, 120, seed=74, gb=True)
    _, ps40, c40, f40, e40 = handle_run(topos, coords, "double", 40, 40, 32)
    _, ps60, c60, f60, e60 = handle_run(topos, coords, "double", 40, 60, 32)
    assert all(np.array_equal(a * (1 << 20), b) for a, b in zip(f40, f60))
    assert np.array_equal(e40, e60)
    X = ps60.coordinates()
    F = ps60.forces()
    for s, t in enumerate(topos):
        _, _, fo = oracle.eval_mm(t, X[s], True, 32)
        err = np.sqrt(((F[s] - fo) ** 2).sum() / (fo ** 2).sum())
        assert err < 1e-8, (s, err)


def test_wide_coordinates_evaluate_like_narrow_ones():
    """Coordinates at 60 bits (with overflow words) give the same energies and forces as at 38
    bits to FP64 rounding, and the exact 95-bit values round-trip through the device."""
    topos, coords = synthetic.uniform_batch(5, 30, 120, seed=75, gb=True)
    _, ps38, c38, f38, e38 = handle_run(topos, coords, "double", 38, 40, 32)
    _, ps60, c60, f60, e60 = handle_run(topos, coords, "double", 60, 40, 32)
    for s in range(len(topos)):
        want = np.array([[int(round(v * 2.0 ** 60)) for v in row] for row in coords[s]],
                        dtype=object)
        # llround of x * 2^60 in double precision: exact powers-of-two scaling of the double.
        assert np.array_equal(c60[s], want)
    E38 = e38[:, :13].astype(float) * 2.0 ** -32
    E60 = e60[:, :13].astype(float) * 2.0 ** -32
    np.testing.assert_allclose(E60, E38, rtol=0, atol=1e-8 * np.abs(E38).sum(axis=1).max())
    for a, b in zip(f38, f60):
        fa, fb = a.astype(float), b.astype(float)
        assert np.sqrt(((fa - fb) ** 2).sum() / (fa ** 2).sum()) < 1e-8


def test_cpp_caller_with_its_own_device_arrays():
    """tests/capi/kit_boundary.cpp: a C++ program hipMallocs its arrays in the writer layouts
    and runs the kit entry points; evaluation and 12-cycle minimizations (per-step and graph) at
    32 / 23, 40 / 40 and 60 / 60 bits equal the handle entry points word for word."""
    import os
    import subprocess
    from stormm_amd import build
    exe = os.path.join(build.CAPI_BIN, "kit_boundary")
    assert os.path.exists(exe), "build the package first (python -m stormm_amd.build)"
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0 and "kit boundary ok" in out.stdout
