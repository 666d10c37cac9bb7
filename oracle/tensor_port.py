"""ORACLE — test infrastructure, not product code.

Plain-numpy restatement of the reference's tensor permanent,
``_calculate_tensor_permanent`` (piquasso/_simulators/passive/probabilities.py:193-237):
subset row sums and row |.|^2 sums tabulated over all 2^n column subsets
(each from the subset without its lowest column), then the double loop over
non-empty subset pairs (s, t) with sign (-1)^(n-|s|) (-1)^(n-|t|) and the
product over rows of x R_s conj(R_t) + (1 - x) Q_{s&t}, accumulated one pair
at a time.  It is the timed CPU baseline of the GPU kernel (the reference's
own code is Python and is not run here); tests pin it to the double-double
oracle.  Only ``tests/`` and ``bench.py``'s CPU leg use it.
"""

from __future__ import annotations

import numpy as np

__all__ = ["tensor_perm_port"]


def tensor_perm_port(a, overlap: float) -> complex:
    a = np.asarray(a, dtype=np.complex128)
    n = a.shape[1]
    if n == 0:
        return 1.0 + 0j
    subsets = 1 << n
    sums = np.zeros((subsets, n), dtype=np.complex128)
    abs_sums = np.zeros((subsets, n), dtype=np.complex128)
    sq = np.abs(a) ** 2
    for s in range(1, subsets):
        low = s & -s
        c = low.bit_length() - 1
        sums[s] = sums[s ^ low] + a[:, c]
        abs_sums[s] = abs_sums[s ^ low] + sq[:, c]
    total = 0.0 + 0j
    for s in range(1, subsets):
        sign_s = -1 if (n - bin(s).count("1")) % 2 else 1
        for t in range(1, subsets):
            sign_t = -1 if (n - bin(t).count("1")) % 2 else 1
            rows = overlap * sums[s] * np.conj(sums[t]) + (1.0 - overlap) * abs_sums[s & t]
            total += sign_s * sign_t * np.prod(rows)
    return complex(total)
