"""Generated HIP source: pair-parity factoring is detected for the MRT bases and
the printed transforms are exact (checked by evaluating the printed C sums in
Python on random populations)."""
import re

import numpy as np
import pytest

from pylbm_amd import cases
from pylbm_amd.generator.hip import _pairs, _row_parities, generate_source
from pylbm_amd.scheme import Scheme


@pytest.mark.parametrize("make", [lambda: cases.lid_cavity_d2q9(8), lambda: cases.taylor_green_d3q19(4),
                                  lambda: cases.d2q5q9_vectorial(8), lambda: cases.d1q3_advection(8)])
def test_parity_detected(make):
    spec = Scheme(make()).kernel_spec()
    ksym = _pairs(spec)
    assert ksym is not None
    assert _row_parities(spec.M, ksym) is not None
    assert "pair-factored=True" in generate_source(spec)


def _eval_block(src, start, end, env):
    body = src[src.index(start):]
    body = body[:body.index(end)]
    for line in body.splitlines():
        line = line.strip()
        mm = re.match(r"(?:const C |C )?(\w+) = (.*);$", line)
        if not mm:
            continue
        expr = re.sub(r"C\(([-0-9.e+]+)\)", r"(\1)", mm.group(2))
        env[mm.group(1)] = eval(expr, {}, env)
    return env


def test_printed_transforms_are_exact():
    spec = Scheme(cases.taylor_green_d3q19(4)).kernel_spec()
    src = generate_source(spec)
    M = np.array([[float(c) for c in r] for r in spec.M])
    iM = np.array([[float(c) for c in r] for r in spec.invM])
    rng = np.random.default_rng(1)
    f = rng.random(spec.ns)
    env = {"f%d" % k: f[k] for k in range(spec.ns)}
    kern = src[src.index("fused_step(const S*"):]
    env = _eval_block(kern, "const C p0", "const C q_", env)
    m = np.array([env["m%d" % i] for i in range(spec.ns)])
    assert np.allclose(m, M @ f, rtol=1e-14, atol=1e-14)
    m2 = rng.random(spec.ns)
    env2 = {"m%d" % i: m2[i] for i in range(spec.ns)}
    start = kern.index("const C r0") if "const C r0" in kern else kern.index("const C E")
    env2 = _eval_block(kern[start:], kern[start:start + 10], "G0[j] = Codec<S>::encode(", env2)
    g = np.array([env2["g%d" % k] for k in range(spec.ns)])
    assert np.allclose(g, iM @ m2, rtol=1e-13, atol=1e-14)
