"""CPU oracle for the pylbm time-stepper -- TEST INFRASTRUCTURE ONLY.

This package is a plain restatement of the reference algorithm for the
``Simulation.one_time_step`` path, written for checking, never for shipping:

* ``stencil``   -- velocity numbering and symmetry (pylbm/stencil.py:119-373, 783-835)
* ``scheme``    -- moment matrix M, exact inverse, equilibrium, relaxation rates,
                   conserved-moment permutation (pylbm/scheme.py:136-241, 464-532, 575-607)
* ``domain``    -- dense in_or_out / distance / flag arrays with the reference's
                   box and element rules (pylbm/domain.py:252-635, pylbm/elements/*)
* ``boundary``  -- link lists istore / iload / rhs (pylbm/boundary.py:23-160, 226-305, 400-895)
* ``lbm``       -- numpy FP64 driver: initialisation, periodic ghost fill, link
                   loops, fused pull step (pylbm/simulation.py:257-420,
                   pylbm/storage.py:306-367, pylbm/algorithm/pull.py:11-59)
* ``cpu_step.c``-- C AoS restatement of the generated Cython loop, used only as the
                   timed CPU baseline (pylbm/generator/printing/cython.py:301-350)

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import anything here.  The product package ``pylbm_amd`` never
does, and fails loudly when its HIP library is missing.

Pinning (what each function is checked against) is listed per module header
and exercised in ``tests/test_oracle_*.py``.
"""
