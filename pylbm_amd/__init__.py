"""pylbm_amd -- pylbm's ``Simulation(dico)`` / ``one_time_step()`` on AMD MI355X.

A ``generator='hip'`` backend for pylbm scheme dictionaries: the scheme's
moment matrix, equilibrium and relaxation rates are printed as HIP kernels
for gfx950 (one fused pull-stream + MRT-collide kernel, link-list boundary
kernels, periodic ghost fill), compiled with hipcc and called through the C
ABI of include/pylbm_hip.h.  Names follow pylbm: ``Simulation``, ``Domain``,
``Scheme``, ``Stencil``, ``bc.BounceBack`` ..., ``Circle``, ``Sphere``,
``H5File``.
"""
from . import bc  # noqa: F401
from .domain import Domain  # noqa: F401
from .elements import (Circle, CylinderCircle, CylinderEllipse, CylinderTriangle, Ellipse,  # noqa: F401
                       Ellipsoid, Parallelepiped, Parallelogram, Sphere, Triangle)
from .hdf5 import H5File  # noqa: F401
from .monitoring import Monitor  # noqa: F401
from .scheme import Scheme  # noqa: F401
from .simulation import Simulation  # noqa: F401
from .stencil import Stencil, Velocity  # noqa: F401

__version__ = "0.1.0"
