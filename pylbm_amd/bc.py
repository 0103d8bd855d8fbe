"""Boundary-condition method classes used in the ``boundary_conditions`` dict.

Same names and meaning as pylbm/boundary.py:400-895; each class carries the
link-list rule of its reference counterpart:

* ``BounceBack``      f[k, x_out] = f[ksym, x_out + v_k] + (feq_k - feq_ksym)
* ``AntiBounceBack``  f[k, x_out] = -f[ksym, x_out + v_k] + (feq_k + feq_ksym)
* ``Neumann``         f[k, x_out] = f[k, x_out + v_k]
* ``NeumannX/Y/Z``    f[k, x_out] = f[k, x_out + v_k projected on one axis]

The device work is done by ``lbm_links_apply`` (include/pylbm_hip.h).
"""

__all__ = ["BounceBack", "AntiBounceBack", "Neumann", "NeumannX", "NeumannY", "NeumannZ",
           "BouzidiBounceBack", "BouzidiAntiBounceBack"]


class BoundaryMethodKind:
    """Rule of one link-list method: load index and sign, rhs from feq."""
    name = None
    sign = 1.0            # f_store = sign * f_load + rhs
    symmetric_load = True  # load population is ksym (bounce-back family) or k (Neumann)
    rhs_sign = None       # rhs = feq[k] + rhs_sign * feq[ksym]; None: no rhs
    shift_axis = None     # Neumann variants: only this component of v_k is applied


class BounceBack(BoundaryMethodKind):
    name = "bounce_back"
    sign = 1.0
    rhs_sign = -1.0


class AntiBounceBack(BoundaryMethodKind):
    name = "anti_bounce_back"
    sign = -1.0
    rhs_sign = 1.0


class Neumann(BoundaryMethodKind):
    name = "neumann"
    symmetric_load = False


class NeumannX(Neumann):
    name = "neumannx"
    shift_axis = 0


class NeumannY(Neumann):
    name = "neumanny"
    shift_axis = 1


class NeumannZ(Neumann):
    name = "neumannz"
    shift_axis = 2


class BouzidiBounceBack(BoundaryMethodKind):
    """Interpolated bounce-back (pylbm/boundary.py:472-623): not provided by the hip generator."""
    name = "Bouzidi_bounce_back"
    unsupported = True


class BouzidiAntiBounceBack(BoundaryMethodKind):
    """Interpolated anti-bounce-back (pylbm/boundary.py:687-760): not provided by the hip generator."""
    name = "Bouzidi_anti_bounce_back"
    unsupported = True
