"""Compile and load generated scheme modules.

Equivalent of pylbm's ``autowrap`` step (pylbm/generator/autowrap.py:52-179:
write source, run the compiler in a subprocess, import the result), with
``hipcc --offload-arch=gfx950`` instead of Cython + gcc.  Modules are cached
in-tree under ``pylbm_amd/_generated/`` keyed by a hash of the scheme's
kernel spec, so a scheme compiled once (for instance by the package build)
is loaded without recompiling.
"""
import ctypes
import os
import subprocess
import threading

from ..runtime import GEN_SIGNATURES, bind, lib_path
from .hip import generate_source, spec_hash

__all__ = ["CodeWrapError", "module_path", "compile_spec", "load_module", "HIPCC_FLAGS"]

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GEN_DIR = os.path.join(_PKG, "_generated")
CSRC = os.path.join(_PKG, "csrc")
INCLUDE = os.path.join(os.path.dirname(_PKG), "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -ffp-contract=off: every product and sum of the generated arithmetic rounds on its
# own, so the kernel variants that share it (halo / wrap reads, vector widths,
# equilibrium + m2f fused or not) give the same bits wherever they compute the
# same cell, whatever the compiler's scheduling of each variant
HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


class CodeWrapError(Exception):
    """The generated code did not compile (carries the compiler output)."""


def module_path(spec):
    return os.path.join(GEN_DIR, "lbm_gen_%s.so" % spec_hash(spec))


_LOCK = threading.Lock()
_LOADED = {}


def compile_spec(spec, force=False):
    """Write and compile the module of ``spec``; returns the .so path."""
    path = module_path(spec)
    if os.path.exists(path) and not force:
        return path
    os.makedirs(GEN_DIR, exist_ok=True)
    src = path[:-3] + ".hip"
    with open(src, "w") as fh:
        fh.write(generate_source(spec))
    tmp = path + ".tmp%d" % os.getpid()
    cmd = [HIPCC] + HIPCC_FLAGS + ["-I" + INCLUDE, "-I" + CSRC, src, "-o", tmp,
                                   "-L" + _PKG, "-lpylbm_hip", "-Wl,-rpath,$ORIGIN/.."]
    if not os.path.exists(lib_path()):
        raise CodeWrapError("core library %s missing: build the package first" % lib_path())
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        raise CodeWrapError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), proc.stdout))
    os.replace(tmp, path)
    return path


def load_module(spec):
    """ctypes handle of the compiled module for ``spec`` (compiled on a cache miss)."""
    with _LOCK:
        key = spec_hash(spec)
        if key not in _LOADED:
            path = compile_spec(spec)
            _LOADED[key] = bind(ctypes.CDLL(path), GEN_SIGNATURES)
        return _LOADED[key]
