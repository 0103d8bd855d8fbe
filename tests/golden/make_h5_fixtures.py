"""Convert the reference's end-to-end fixtures (tests/reference/*.h5 of pylbm,
written by its h5diff plugin, tests/conftest.py:221-243) into .npz files.

Each fixture holds, after ``run(dx, Tf)`` of a demo (tests/test_demo_*.py),
the coordinates ``x_0``, ``x_1``, ``x_2`` and one dataset per conserved moment
with solid cells set to 0.  pylbm's H5File stores the transpose of each field
(pylbm/hdf5.py:163-166); here it is transposed back, so ``npz[name]`` has the
shape of ``sim.m[name]`` (x first).

Run in a container where the reference is present:
    python tests/golden/make_h5_fixtures.py /root/reference/tests/reference
"""
import glob
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from h5read import read_h5  # noqa: E402


def main(src):
    out_dir = os.path.join(HERE, "h5")
    os.makedirs(out_dir, exist_ok=True)
    for path in sorted(glob.glob(os.path.join(src, "*.h5"))):
        data = read_h5(path)
        arrays = {}
        for name, arr in data.items():
            arrays[name] = arr if name.startswith("x_") else np.ascontiguousarray(arr.T)
        name = os.path.splitext(os.path.basename(path))[0]
        np.savez_compressed(os.path.join(out_dir, name + ".npz"), **arrays)
        print(name, {k: v.shape for k, v in arrays.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/tests/reference")
