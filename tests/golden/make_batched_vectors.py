"""Write tests/golden/batched_uneven_seed4.npz: the 48-mode Haar unitary of
the batched workload (workloads.batched_uneven(seed=4)) and the
double-double oracle value of each of its 4096 permanents.

Storing the unitary pins the inputs bit-for-bit (QR rounding depends on the
LAPACK build); the selections and multiplicities are integer draws of the
seeded generator and reproduce everywhere.

Run:  python tests/golden/make_batched_vectors.py
"""

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from piquasso_amd import workloads as W  # noqa: E402


def main():
    rng = np.random.default_rng(4)
    u = W.haar_unitary(48, rng)
    mats, rows, cols = W.batched_uneven(seed=4, unitary=u)
    t0 = time.time()
    dd = np.array([O.perm_dd_parts(m, r, c) for m, r, c in zip(mats, rows, cols)])
    print(f"{len(mats)} permanents in {time.time() - t0:.1f} s")
    path = os.path.join(HERE, "batched_uneven_seed4.npz")
    np.savez_compressed(path, unitary=u, dd=dd)
    print("wrote", path)


if __name__ == "__main__":
    main()
