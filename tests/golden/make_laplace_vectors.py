"""Write tests/golden/laplace_vectors.json: inputs and outputs of the
reference's ``permanent_laplace_cpp<double>`` (src/permanent_laplace.cpp:43-240),
compiled from its sources by ``make -C oracle ref`` and called through the
shim in oracle/ref_shim.cpp, next to the double-double oracle's values.

The cases are seeded random matrices with row and column multiplicities
(sum(cols) = sum(rows) + 1), some columns of multiplicity 0 (where the
reference returns the uncut product sum), rectangular shapes, a 40-mode
boson-sampling-like selection, and the degenerate inputs that return [1].

Run:  make -C oracle ref REF=/root/reference && python tests/golden/make_laplace_vectors.py
"""

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402


def cplx(a):
    a = np.asarray(a, dtype=np.complex128)
    return {"re": a.real.reshape(-1).tolist(), "im": a.imag.reshape(-1).tolist()}


def case(cid, a, r, c):
    a = np.asarray(a, dtype=np.complex128)
    ref = O.ref_laplace(a, r, c)
    dd = O.laplace_dd(a, r, c)
    return {"id": cid, "shape": list(a.shape), "matrix": cplx(a), "rows": [int(x) for x in r],
            "cols": [int(x) for x in c], "reference": cplx(ref), "dd": cplx(dd)}


def main():
    if not O.ref_available():
        raise SystemExit("build the reference first: make -C oracle ref REF=/root/reference")
    rng = np.random.default_rng(2024)
    out = []
    # random small cases with multiplicities and zero columns
    for t in range(24):
        m_r, m_c = (int(x) for x in rng.integers(1, 7, size=2))
        n = int(rng.integers(1, 9))
        a = (rng.standard_normal((m_r, m_c)) + 1j * rng.standard_normal((m_r, m_c))) / np.sqrt(2)
        r = np.bincount(rng.integers(0, m_r, n), minlength=m_r)
        c = np.bincount(rng.integers(0, m_c, n + 1), minlength=m_c)
        out.append(case(f"random_{t}", a, r, c))
    # unit multiplicities, square-plus-one-column (Clifford & Clifford step)
    for n in (5, 9, 12):
        a = (rng.standard_normal((n, n + 1)) + 1j * rng.standard_normal((n, n + 1))) / np.sqrt(2 * n)
        out.append(case(f"ones_{n}x{n + 1}", a, np.ones(n, int), np.ones(n + 1, int)))
    # boson-sampling-like: 40 modes, 12 photons in, column pattern with bunching
    z = (rng.standard_normal((40, 40)) + 1j * rng.standard_normal((40, 40))) / np.sqrt(2)
    q, rr = np.linalg.qr(z)
    u = q * (np.diag(rr) / np.abs(np.diag(rr)))
    rows = np.zeros(40, int)
    rows[:12] = 1
    cols = np.bincount(rng.integers(0, 40, 13), minlength=40)
    out.append(case("bs_m40_n12", u, rows, cols))
    # degenerate inputs: the reference returns [1]
    out.append(case("zero_rows", np.ones((2, 2)), [0, 0], [0, 1]))
    out.append(case("zero_cols", np.ones((2, 2)), [1, 0], [0, 0]))
    path = os.path.join(HERE, "laplace_vectors.json")
    with open(path, "w") as f:
        json.dump({"source": "src/permanent_laplace.cpp:43-240 built by oracle/Makefile (make ref)",
                   "cases": out}, f)
    print("wrote", path, len(out), "cases")


if __name__ == "__main__":
    main()
