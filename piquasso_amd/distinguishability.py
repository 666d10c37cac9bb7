"""Detection probabilities of uniformly partially distinguishable photons.

piquasso evaluates them with the tensor permanent of the
multiplicity-expanded interferometer (Tichy, arXiv:1410.7687, Sec. III):
``get_partially_distinguishable_detection_probability`` and
``_calculate_tensor_permanent`` in
piquasso/_simulators/passive/probabilities.py:156-237, a pure-Python double
loop over 4^n column-subset pairs.  Here the same value is summed over
4^(n-1) pairs of sign vectors (Glynn's form, far less cancellation) in the
HIP kernels of ``tensor_perm.hip`` (``pqp_tensor_permanent_batch_c128``):
one lane per sign vector, a Gray walk over the other per wave,
double-double accumulation, and only the real part (the two orders of a
pair give conjugate terms, so the pair sum is real).

Deliberate difference: the reference returns the tensor permanent and the
probability in a complex dtype whose imaginary part is zero up to rounding;
these functions return real floats.
"""

from __future__ import annotations

import ctypes
import math
from typing import Sequence

import numpy as np

from . import _capi

__all__ = ["tensor_permanent", "tensor_permanent_batch", "detection_probability",
           "detection_probabilities", "uniform_input_norm"]

MAX_PARTICLES = 20


def _square(matrix) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(matrix, dtype=np.complex128))
    if a.ndim != 2 or a.shape[0] != a.shape[1]:
        raise ValueError(f"tensor permanent needs a square matrix, got shape {a.shape}")
    return a


def tensor_permanent_batch(matrices, particle_overlaps, with_abs_sum: bool = False):
    """Tensor permanents of many square matrices (each with its own uniform
    particle overlap) from one launch per matrix order; float64 array.  With
    ``with_abs_sum`` also the sums of |term| (about 8 (n + 2) eps times that
    bounds each result's error; the alternating sum cancels heavily)."""
    mats = [_square(m) for m in matrices]
    x = np.ascontiguousarray(np.broadcast_to(np.asarray(particle_overlaps, dtype=np.float64),
                                             (len(mats),)))
    count = len(mats)
    out = np.zeros(count)
    asum = np.zeros(count)
    if count == 0:
        return (out, asum) if with_abs_sum else out
    n = np.array([m.shape[0] for m in mats], dtype=np.int64)
    dp = ctypes.POINTER(ctypes.c_double)
    i64p = ctypes.POINTER(ctypes.c_int64)
    ptrs = (dp * count)(*[m.view(np.float64).ctypes.data_as(dp) if m.size else None for m in mats])
    st = _capi.lib().pqp_tensor_permanent_batch_c128(
        count, ptrs, n.ctypes.data_as(i64p), n.ctypes.data_as(i64p), x.ctypes.data_as(dp),
        out.ctypes.data_as(dp), asum.ctypes.data_as(dp))
    _capi.check(st)
    return (out, asum) if with_abs_sum else out


def tensor_permanent(matrix, particle_overlap: float) -> float:
    """``_calculate_tensor_permanent(A, particle_overlap)``
    (piquasso/_simulators/passive/probabilities.py:193-237) on the GPU."""
    return float(tensor_permanent_batch([matrix], [particle_overlap])[0])


def uniform_input_norm(input_occupation, particle_overlap: float) -> float:
    """prod_j sum_k C(n_j, k) x^k (1 - x)^(n_j - k) k!, the normalisation of
    the input state (probabilities.py:387-398)."""
    norm = 1.0
    for nj in np.asarray(input_occupation, dtype=np.int64).reshape(-1):
        nj = int(nj)
        norm *= sum(math.comb(nj, k) * particle_overlap ** k * (1.0 - particle_overlap) ** (nj - k)
                    * math.factorial(k) for k in range(nj + 1))
    return norm


def _expand(interferometer, rows, cols) -> np.ndarray:
    """Row i repeated rows[i] times, column j cols[j] times
    (piquasso/_math/linalg.py:98-117)."""
    ri = np.repeat(np.arange(len(rows)), rows)
    ci = np.repeat(np.arange(len(cols)), cols)
    return np.asarray(interferometer)[np.ix_(ri, ci)]


def detection_probabilities(occupation_numbers, interferometer, input_occupation: Sequence[int],
                            uniform_particle_overlap: float) -> np.ndarray:
    """Probabilities of detecting each output occupation (rows of
    ``occupation_numbers``) for the input occupation with uniform particle
    overlap: tensor permanent / (input norm x prod of output factorials),
    0 where the photon numbers differ (probabilities.py:156-190).  All
    tensor permanents come from one batched call."""
    u = np.asarray(interferometer, dtype=np.complex128)
    inp = np.asarray(input_occupation, dtype=np.int64).reshape(-1)
    outs = np.asarray(occupation_numbers, dtype=np.int64)
    if outs.ndim == 1:
        outs = outs[None, :]
    d = u.shape[0]
    if u.shape != (d, d) or inp.shape[0] != d or outs.shape[1] != d:
        raise ValueError("occupations must have one entry per mode of a square interferometer")
    if (inp < 0).any() or (outs < 0).any():
        raise ValueError("occupation numbers must be non-negative")
    n = int(inp.sum())
    if n > MAX_PARTICLES:
        raise ValueError(f"at most {MAX_PARTICLES} particles are supported, got {n}")
    x = float(uniform_particle_overlap)
    probs = np.zeros(outs.shape[0])
    keep = np.flatnonzero(outs.sum(axis=1) == n)
    if keep.size == 0:
        return probs
    mats = [_expand(u, outs[k], inp) for k in keep]
    t = tensor_permanent_batch(mats, [x] * len(mats))
    in_norm = uniform_input_norm(inp, x)
    for j, k in enumerate(keep):
        out_fact = float(np.prod([math.factorial(int(v)) for v in outs[k]]))
        probs[k] = t[j] / (in_norm * out_fact)
    return probs


def detection_probability(occupation_number, interferometer, input_occupation,
                          uniform_particle_overlap: float) -> float:
    """``get_partially_distinguishable_detection_probability``
    (probabilities.py:156-190) for one output occupation."""
    return float(detection_probabilities([occupation_number], interferometer, input_occupation,
                                         uniform_particle_overlap)[0])
