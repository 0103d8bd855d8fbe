"""Code generation for the hip backend: sympy scheme -> HIP source -> .so."""
from .build import CodeWrapError, compile_spec, load_module  # noqa: F401
from .hip import generate_source, spec_hash  # noqa: F401
