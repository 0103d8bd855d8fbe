"""A population array beyond the 32-bit element index range on the GPU: the
D3Q19 FP32 box of 512^3 cells (2 x 10.3 GB of populations, about 2.8e9 elements
each) takes the kernels' 64-bit index variants.

* rest state: every population stays w_k exactly, so each population's FP64
  compensated sum over the interior is exactly N w_k;
* Taylor-Green vortex: mass and momentum sums conserved to FP32 rounding over
  two steps (a wrong index would move or lose populations)."""
import numpy as np
import pytest

from pylbm_amd import Simulation, cases
from pylbm_amd.storage import layout_elements

pytestmark = pytest.mark.gpu

N = 512


def _pop_sums(sim):
    import ctypes
    rt, F = sim.rt, sim.container.F
    s, c = (ctypes.c_double * sim.ns)(), (ctypes.c_double * sim.ns)()
    rt.check(rt.lib.lbm_reduce_populations_noalloc(ctypes.c_void_p(F.ptr), ctypes.byref(F.layout), None, s, c,
                                                   rt.stream), "reduce")
    return np.array(s) + np.array(c)


def test_rest_state_512_cubed(hip):
    d = cases.taylor_green_d3q19(N)
    d["init"] = {cases.RHO: 1.0, cases.QX: 0.0, cases.QY: 0.0, cases.QZ: 0.0}
    sim = Simulation(d, dtype="float32")
    assert layout_elements(sim.container.F.layout) > (1 << 31)
    w = np.asarray(sim._rest_populations(), dtype=np.float32).astype(np.float64)
    sim.run(2)
    got = _pop_sums(sim)
    assert np.array_equal(got, w * float(N) ** 3), (got, w * float(N) ** 3)


def test_taylor_green_512_cubed_conserves(hip):
    sim = Simulation(cases.taylor_green_d3q19(N), dtype="float32")
    m0 = sim.conserved_sums()
    sim.run(2)
    m1 = sim.conserved_sums()
    assert abs(m1[0] - m0[0]) <= 1e-6 * abs(m0[0]), (m0[0], m1[0])
    for c in (1, 2, 3):
        assert abs(m1[c] - m0[c]) <= 1e-6 * abs(m0[0]), (c, m0[c], m1[c])
