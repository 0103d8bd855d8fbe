"""The C-ABI libraries load and export every entry point include/pylbm_hip.h
and include/pylbm_surrogate.h declare (no GPU needed; nothing is called)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pylbm_hip.h")
SUR_HEADER = os.path.join(ROOT, "include", "pylbm_surrogate.h")


def _declared(header=HEADER):
    txt = open(header).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\**(lbm_\w+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


def test_header_parses():
    names = _declared()
    assert "lbm_gen_step" in names and "lbm_links_apply" in names and len(names) > 25


def test_core_exports():
    from pylbm_amd.runtime import CORE_SIGNATURES, load_core_library
    lib = load_core_library()
    core = [n for n in _declared() if not n.startswith("lbm_gen_")]
    for n in core:
        assert hasattr(lib, n), n
    assert set(core) == set(CORE_SIGNATURES)


def test_surrogate_exports():
    from pylbm_amd.runtime import SURROGATE_SIGNATURES, load_core_library
    lib = load_core_library()
    names = _declared(SUR_HEADER)
    assert "lbm_sur_conv" in names and "lbm_sur_wgrad" in names
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(SURROGATE_SIGNATURES)


def test_generated_exports():
    from pylbm_amd import cases
    from pylbm_amd.generator.build import module_path
    from pylbm_amd.runtime import GEN_SIGNATURES, load_core_library
    from pylbm_amd.scheme import Scheme
    load_core_library()
    path = module_path(Scheme(cases.taylor_green_d3q19(4)).kernel_spec())
    assert os.path.exists(path), "run __graft_entry__.build() first"
    lib = ctypes.CDLL(path)
    gen = [n for n in _declared() if n.startswith("lbm_gen_")]
    for n in gen:
        assert hasattr(lib, n), n
    assert set(gen) == set(GEN_SIGNATURES)


def test_gpu_path_fails_loudly_without_device():
    """Without a GPU the runtime raises instead of falling back to the CPU."""
    import pylbm_amd.runtime as rt
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    with pytest.raises(rt.HipUnavailable):
        rt.Runtime()


def test_module_object_uses_pylbm_argument_names():
    """pylbm's call_genfunction (pylbm/symbolic.py:288-299) passes exactly the names in
    getfullargspec(func).args: the routines of the module object carry pylbm's
    argument names (pylbm/algorithm/base.py:630-667, pylbm/boundary.py:358-376)."""
    import inspect
    from pylbm_amd.module import HipModule
    expect = {
        "one_time_step": ["f", "fnew", "t", "dt"],
        "f2m": ["f", "m"],
        "m2f": ["m", "f"],
        "equilibrium": ["m"],
        "update": ["f"],
        "bounce_back": ["f", "istore", "iload0", "rhs", "ncond"],
        "anti_bounce_back": ["f", "istore", "iload0", "rhs", "ncond"],
        "neumann": ["f", "istore", "iload0", "ncond"],
        "neumannx": ["f", "istore", "iload0", "ncond"],
        "neumanny": ["f", "istore", "iload0", "ncond"],
        "neumannz": ["f", "istore", "iload0", "ncond"],
        # pylbm/boundary.py:537-566 (_get_args locals) and 606-623, 743-760 (loop symbols)
        "Bouzidi_bounce_back": ["f", "fcopy", "istore", "iload0", "iload1", "rhs", "dist", "ncond"],
        "Bouzidi_anti_bounce_back": ["f", "istore", "iload0", "iload1", "rhs", "dist", "ncond"],
    }
    for name, args in expect.items():
        assert inspect.getfullargspec(getattr(HipModule, name)).args == ["self"] + args, name


def test_layout_beyond_int32_accepted():
    """Arrays of 2^31 elements and more are accepted (64-bit element indices): a
    512^3 D3Q19 FP32 population array has about 2.8e9 elements.  lbm_reduce_blocks
    validates the layout without touching the GPU."""
    import ctypes
    from pylbm_amd import runtime as rtm
    from pylbm_amd.storage import layout_elements, make_layout
    lib = rtm.load_core_library()
    L = make_layout(19, (514, 514, 514), (1, 1, 1), rtm.LBM_F32, interleaved=True)
    assert layout_elements(L) > (1 << 31)
    nb = ctypes.c_int64(0)
    assert lib.lbm_reduce_blocks(ctypes.byref(L), ctypes.byref(nb)) == 0, lib.lbm_last_error()
    assert nb.value == 1024
    P = make_layout(19, (514, 514, 514), (1, 1, 1), rtm.LBM_F64, interleaved=False)
    assert layout_elements(P) > (1 << 31)
    assert lib.lbm_reduce_blocks(ctypes.byref(P), ctypes.byref(nb)) == 0, lib.lbm_last_error()
