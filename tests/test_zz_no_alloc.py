"""The allocation-free minimization entry point (stormm_launch_minimization_kit_noalloc) against
the kit entry point it mirrors (stormm_launch_minimization_kit, mode 1), through the C-ABI on
caller-owned device arrays: every output array equal bit for bit, on two calls in a row whose
second batch is smaller.  The first call stops systems at an RMS-force threshold, so stop flags
left in the library's static state from that call would freeze systems of the second call
unless the state is cleared at the start of every call.
"""
import ctypes

import numpy as np
import pytest

from stormm_amd import _lib, synthetic
from stormm_amd import synthesis as S
from test_gpu_kits import CallerState, check

pytestmark = pytest.mark.gpu

BITS = {"single": (32, 23, 26), "double": (32, 40, 32)}


def outputs(st):
    out = {name: st.values(name) for name in ("crd", "frc", "prv", "_cg_temp")}
    out["scorecard"] = st.acc.get()
    out["line"] = st.line.get()
    return out


@pytest.mark.parametrize("prec", ["single", "double"])
def test_noalloc_minimization_equals_the_kit_entry_point(prec):
    L = _lib.load()
    gpos, frc, nrg = BITS[prec]
    code = S.precision_code(prec)
    # (systems, seed, cycles, RMS-force stop): the second call is smaller and runs every cycle.
    for nsys, seed, cycles, rms in ((14, 61, 80, 2.0), (5, 62, 40, 0.0)):
        topos, coords = synthetic.config2(nsys=nsys, seed=seed)
        ag = S.AtomGraphSynthesis(topos)
        fixed = [np.rint(np.asarray(c) * 2.0 ** gpos).astype(np.int64) for c in coords]
        ref = CallerState(ag, gpos, frc, nrg, fixed)
        new = CallerState(ag, gpos, frc, nrg, fixed)
        mc = S.minimize_controls(cycles, 10, 0.01, rms)
        print(f"[{prec}] {nsys} systems, {cycles} cycles, RMS stop {rms}", flush=True)
        check(L.stormm_launch_minimization_kit(code, ag.handle, ctypes.byref(ref.psw),
                                               ctypes.byref(mc), ctypes.byref(ref.scw),
                                               ctypes.byref(ref.lmw), 1, None))
        check(L.stormm_launch_minimization_kit_noalloc(code, ag.handle, ctypes.byref(new.psw),
                                                       ctypes.byref(mc), ctypes.byref(new.scw),
                                                       ctypes.byref(new.lmw), None))
        check(L.stormm_device_synchronize())
        a, b = outputs(ref), outputs(new)
        for name in ("crd", "frc", "prv", "_cg_temp"):
            assert all(np.array_equal(x, y) for x, y in zip(a[name], b[name])), name
        assert np.array_equal(a["scorecard"], b["scorecard"])
        assert np.array_equal(a["line"], b["line"])
        if rms > 0.0:
            # The threshold stopped some systems before the cycle limit.
            cyc = np.zeros(nsys, dtype=np.int32)
            check(L.stormm_minimization_status(ag.handle, cyc.ctypes.data, None))
            assert (cyc < cycles).any(), cyc
