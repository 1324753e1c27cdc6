"""Convert the reference's example golden files (examples/*.h5) into .npz fixtures.

Data only: the stored arrays (x_frame or u_frame, f_frame, S, t_insta) written by
the reference's own example scripts (examples/*.py, asserted in CI by
.github/workflows/historic.yml:85-94). Run once, in the build container, with an
interpreter that has h5py:

    /opt/conda/bin/python3.9 tests/golden/convert_examples_h5.py /root/reference/examples

The sha256 of every source file is stored in the fixture under "source_sha256".
"""

import hashlib
import os
import sys

import h5py
import numpy as np

NAMES = [
    "Line1d_Cuspy_Laplace",
    "Line1d_Cuspy_Laplace_LongRange",
    "Line1d_Cuspy_Laplace_Nopassing",
    "Line1d_Cuspy_Quartic",
    "Line1d_SemiSmooth_Laplace",
    "Line1d_System_Cuspy_Laplace_RandomForcing",
    "Line2d_Cuspy_Laplace",
    "Particles_Cuspy",
]


def main(src):
    out_dir = os.path.dirname(os.path.abspath(__file__))
    for name in NAMES:
        path = os.path.join(src, name + ".h5")
        with open(path, "rb") as fh:
            sha = hashlib.sha256(fh.read()).hexdigest()
        with h5py.File(path, "r") as f:
            arrays = {k: f[k][...] for k in f.keys()}
        arrays["source_sha256"] = np.array(sha)
        np.savez(os.path.join(out_dir, name + ".npz"), **arrays)
        print(name, {k: v.shape for k, v in arrays.items()}, sha[:12])


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/examples")
