"""The allocation-free reduction (lbm_reduce_populations_noalloc) against the
reduction with a caller workspace (lbm_reduce_populations), through the C ABI:
the same (sum, compensation) pairs bit for bit, on two calls in a row, the
second on a smaller input (the scratch keeps the first call's contents), for
full-value and deviation storage."""
import ctypes

import numpy as np
import pytest

from pylbm_amd import Simulation, cases
from pylbm_amd import runtime as rtm

pytestmark = pytest.mark.gpu


def _with_workspace(sim):
    rt, F = sim.rt, sim.container.F
    nb = ctypes.c_int64(0)
    rt.check(rt.lib.lbm_reduce_blocks(ctypes.byref(F.layout), ctypes.byref(nb)), "blocks")
    part = rt.malloc(8 * 2 * sim.ns * (nb.value + 1))
    w = rt.upload(np.asarray(sim.wshift, dtype=np.float64)) if sim.storage in rtm.DEVIATION_FORMATS else None
    s, c = (ctypes.c_double * sim.ns)(), (ctypes.c_double * sim.ns)()
    try:
        rt.check(rt.lib.lbm_reduce_populations(ctypes.c_void_p(F.ptr), ctypes.byref(F.layout), ctypes.c_void_p(w),
                                               ctypes.c_void_p(part), s, c, rt.stream), "reduce")
    finally:
        rt.free(part)
        if w:
            rt.free(w)
    return list(s), list(c)


def _no_alloc(sim, w):
    rt, F = sim.rt, sim.container.F
    s, c = (ctypes.c_double * sim.ns)(), (ctypes.c_double * sim.ns)()
    rt.check(rt.lib.lbm_reduce_populations_noalloc(ctypes.c_void_p(F.ptr), ctypes.byref(F.layout),
                                                   ctypes.c_void_p(w), s, c, rt.stream), "reduce noalloc")
    return list(s), list(c)


@pytest.mark.parametrize("dtype", ["float64", "float32-shifted"])
def test_noalloc_reduction_equals_workspace_reduction(hip, dtype):
    big = Simulation(cases.taylor_green_d3q19(48), dtype=dtype)
    small = Simulation(cases.karman_d2q9(64, 32, 5.0, (16.0, 16.0)), dtype=dtype)
    big.run(7)
    small.run(5)
    rt = big.rt
    wb = rt.upload(np.asarray(big.wshift, dtype=np.float64)) if big.storage in rtm.DEVIATION_FORMATS else None
    ws = rt.upload(np.asarray(small.wshift, dtype=np.float64)) if small.storage in rtm.DEVIATION_FORMATS else None
    try:
        for sim, w in ((big, wb), (small, ws)):
            got = _no_alloc(sim, w)
            want = _with_workspace(sim)
            assert np.array_equal(np.asarray(got[0]), np.asarray(want[0])), (dtype, sim.ns)
            assert np.array_equal(np.asarray(got[1]), np.asarray(want[1])), (dtype, sim.ns)
        # and again on the large input after the small one
        assert np.array_equal(np.asarray(_no_alloc(big, wb)[0]), np.asarray(_with_workspace(big)[0]))
    finally:
        for p in (wb, ws):
            if p:
                rt.free(p)
