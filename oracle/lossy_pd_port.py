"""ORACLE — test infrastructure, not product code.

Plain-numpy restatement of the reference's
``get_lossy_partially_distinguishable_detection_probabilities``
(piquasso/_simulators/passive/probabilities.py:244-384): the B_loss / B_m
matrices, the subset row sums tabulated over all 2^n particle subsets (each
from the subset without its lowest bit, probabilities.py:429-444), and per
subset the product of the n linear row factors as a complex128 polynomial
truncated to the box of the largest requested occupation per mode
(``multiply_by_linear_truncated``, piquasso/_math/polynomial.py:18-38),
accumulated with the sign (-1)^(n - |S|).  It is the timed CPU baseline of
the GPU kernel (the reference itself is Python and is not run here); tests
pin it to the double-double oracle.  ``subsets`` limits the loop to the first
subsets, for a bounded timing sample.
"""

from __future__ import annotations

import numpy as np

from .oracle import general_input_norm, lossy_pd_matrices, uniform_input_norm

__all__ = ["lossy_pd_port", "subset_row_sums"]


def subset_row_sums(matrix) -> np.ndarray:
    n = matrix.shape[0]
    out = np.zeros((1 << n, n), dtype=matrix.dtype)
    for s in range(1, 1 << n):
        low = s & -s
        out[s] = out[s ^ low] + matrix[:, low.bit_length() - 1]
    return out


def _times_linear(poly, const, lin, out):
    out[...] = const * poly
    for axis, c in enumerate(lin):
        np.moveaxis(out, axis, 0)[1:] += c * np.moveaxis(poly, axis, 0)[:-1]
    return out


def lossy_pd_port(occupation_numbers, transmission_matrix, input_occupation, particle_overlap,
                  subsets: int | None = None) -> np.ndarray:
    occ = np.asarray(occupation_numbers, dtype=int)
    if occ.ndim == 1:
        occ = occ[None, :]
    inp = np.asarray(input_occupation, dtype=int)
    n = int(inp.sum())
    probs = np.zeros(len(occ))
    detected = occ.sum(axis=1)
    if n == 0:
        probs[detected == 0] = 1.0
        return probs
    valid = detected <= n
    if not valid.any():
        return probs
    b0, bms, g = lossy_pd_matrices(transmission_matrix, inp, particle_overlap)
    if np.isscalar(particle_overlap):
        norm = complex(uniform_input_norm(inp, float(particle_overlap)))
    else:
        norm = general_input_norm(inp, g)
    loss_sums = subset_row_sums(b0)
    mode_sums = np.array([subset_row_sums(b) for b in bms])
    shape = tuple(occ[valid].max(axis=0) + 1)
    coef = np.zeros(shape, dtype=np.complex128)
    poly = np.zeros(shape, dtype=np.complex128)
    buf = np.zeros_like(poly)
    vacuum = (0,) * len(shape)
    last = (1 << n) if subsets is None else min(1 << n, 1 + int(subsets))
    for s in range(1, last):
        sign = -1 if (n - bin(s).count("1")) % 2 else 1
        poly.fill(0.0)
        poly[vacuum] = 1.0
        for row in range(n):
            _times_linear(poly, loss_sums[s][row], mode_sums[:, s, row], buf)
            poly, buf = buf, poly
        coef += sign * poly
    coef /= norm
    probs[valid] = np.real(coef[tuple(occ[valid].T)])
    return probs
