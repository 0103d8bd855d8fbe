"""Boundary-condition method classes used in the ``boundary_conditions`` dict.

Same names and meaning as pylbm/boundary.py:400-895; each class carries the
link-list rule of its reference counterpart:

* ``BounceBack``      f[k, x_out] = f[ksym, x_out + v_k] + (feq_k - feq_ksym)
* ``AntiBounceBack``  f[k, x_out] = -f[ksym, x_out + v_k] + (feq_k + feq_ksym)
* ``Neumann``         f[k, x_out] = f[k, x_out + v_k]
* ``NeumannX/Y/Z``    f[k, x_out] = f[k, x_out + v_k projected on one axis]
* ``BouzidiBounceBack``      f[k, x_out] = s f[ksym, x_out + v_k] + (1 - s) f[load1] + (feq_k - feq_ksym)
* ``BouzidiAntiBounceBack``  f[k, x_out] = -s f[ksym, x_out + v_k] + (1 - s) f[load1] + (feq_k + feq_ksym)
  with, for the wall distance q of the link, s = 2q and load1 = (ksym, x_out + 2 v_k)
  when q < 0.5, s = 1/(2q) and load1 = (k, x_out + v_k) otherwise.

The device work is done by ``lbm_links_apply`` and ``lbm_links_apply_bouzidi``
(include/pylbm_hip.h).
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
    bouzidi = False       # interpolated: two loads weighted by the wall distance
    reads_copy = False    # loads come from a copy of f taken before the list runs


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
    """Interpolated bounce-back (pylbm/boundary.py:472-623); its loop reads a copy of
    f taken before it (boundary.py:545-549, 595-606)."""
    name = "Bouzidi_bounce_back"
    sign = 1.0
    rhs_sign = -1.0
    bouzidi = True
    reads_copy = True


class BouzidiAntiBounceBack(BoundaryMethodKind):
    """Interpolated anti-bounce-back (pylbm/boundary.py:687-760); its loop reads f
    itself (boundary.py:732-743)."""
    name = "Bouzidi_anti_bounce_back"
    sign = -1.0
    rhs_sign = 1.0
    bouzidi = True
