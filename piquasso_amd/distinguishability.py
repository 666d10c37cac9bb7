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

The general case — any Gram matrix of the particles' internal states and a
lossy (non-unitary) transmission matrix — is
``get_lossy_partially_distinguishable_detection_probabilities``
(probabilities.py:244-384): the coefficients of perm(B_loss + sum_m x_m B_m)
by Ryser's formula over particle subsets, each subset's product of linear
row factors kept as a truncated polynomial.  Here that subset loop runs in
``perm_poly.hip`` (``pqp_perm_poly_coefficients_c128``), one workgroup per
subset, with the polynomial kept on the monomials that divide a requested
output pattern (the reference keeps the whole box up to the largest
occupation per mode; the coefficients read out are the same).

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
           "detection_probabilities", "uniform_input_norm", "general_input_norm",
           "perm_poly_coefficients", "lossy_detection_probabilities",
           "get_lossy_partially_distinguishable_detection_probabilities"]

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


def perm_poly_coefficients(b0, bms, exponents, with_info: bool = False):
    """Coefficients [x^e] perm(B_0 + sum_m x_m B_m) for every row e of
    ``exponents`` (shape (count, len(bms))), complex128 array, from
    ``pqp_perm_poly_coefficients_c128``.  Rows with total degree above n give
    0.  With ``with_info`` also a dict with the monomial count, links,
    workgroups, whether the polynomial was kept in LDS, the algorithmic FP64
    flops per subset and the number of subsets."""
    b0 = np.ascontiguousarray(np.asarray(b0, dtype=np.complex128))
    n = b0.shape[0]
    if b0.ndim != 2 or b0.shape != (n, n) or n == 0:
        raise ValueError(f"B_0 must be a non-empty square matrix, got shape {b0.shape}")
    bm = np.ascontiguousarray(np.asarray(bms, dtype=np.complex128))
    if bm.size == 0:
        bm = np.zeros((0, n, n), dtype=np.complex128)
    if bm.ndim != 3 or bm.shape[1:] != (n, n):
        raise ValueError(f"B_m must be {n} x {n} matrices, got shape {bm.shape}")
    e = np.atleast_2d(np.asarray(exponents))
    if e.dtype != np.int32 or not e.flags.c_contiguous:
        e = np.ascontiguousarray(e, dtype=np.int32)
    if e.ndim != 2 or e.shape[1] != bm.shape[0]:
        raise ValueError("one exponent per variable in every row")
    out = np.zeros(2 * max(len(e), 1))
    info = np.zeros(6, dtype=np.int64)
    dp = ctypes.POINTER(ctypes.c_double)
    st = _capi.lib().pqp_perm_poly_coefficients_c128(
        n, bm.shape[0], b0.view(np.float64).ctypes.data_as(dp),
        bm.view(np.float64).ctypes.data_as(dp) if bm.size else None, len(e),
        e.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), out.ctypes.data_as(dp),
        info.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    _capi.check(st)
    coef = out[: 2 * len(e)].view(np.complex128).copy()
    if not with_info:
        return coef
    keys = ("monomials", "links", "workgroups", "lds", "flops_per_subset", "subsets")
    return coef, dict(zip(keys, (int(v) for v in info)))


def general_input_norm(input_occupation, particle_overlap_matrix) -> complex:
    """prod over input modes holding more than one particle of
    perm(G; mask, mask), mask selecting that mode's particles
    (``_general_input_norm``, probabilities.py:402-426), each permanent from
    the GPU."""
    from ._math.permanent import permanent

    g = np.asarray(particle_overlap_matrix, dtype=np.complex128)
    norm = 1.0 + 0.0j
    start = 0
    for nj in np.asarray(input_occupation, dtype=np.int64).reshape(-1):
        nj = int(nj)
        if nj > 1:
            mask = np.zeros(len(g), dtype=np.int32)
            mask[start:start + nj] = 1
            norm *= complex(permanent(g, mask, mask))
        start += nj
    return norm


def _lossy_matrices(t, input_occupation, particle_overlap, conjugate_loss_kernel=False):
    """B_loss = G o (I - T^dag T) and B_m = G o (t_m t_m^dag) on the
    first-quantized input modes, t_m = T[m, input_modes]
    (probabilities.py:304-338); with ``conjugate_loss_kernel`` the loss
    kernel is conj(I - T^dag T)."""
    modes = np.repeat(np.arange(len(input_occupation)), input_occupation)
    n = len(modes)
    if np.isscalar(particle_overlap):
        x = float(particle_overlap)
        g = x * np.ones((n, n), dtype=np.complex128) + (1.0 - x) * np.eye(n)
    else:
        g = np.asarray(particle_overlap, dtype=np.complex128)
        if g.shape != (n, n):
            raise ValueError(f"the particle overlap matrix must be {n} x {n}, got {g.shape}")
    kernel = (np.eye(t.shape[1]) - t.conj().T @ t)[np.ix_(modes, modes)]
    if conjugate_loss_kernel:
        kernel = np.conj(kernel)
    tm = t[:, modes]
    bms = g[None, :, :] * (tm[:, :, None] * np.conj(tm)[:, None, :])
    return g * kernel, bms, g


def lossy_detection_probabilities(occupation_numbers, transmission_matrix, input_occupation,
                                  particle_overlap, conjugate_loss_kernel: bool = False
                                  ) -> np.ndarray:
    """Detection probabilities of the output occupations (rows of
    ``occupation_numbers``, one entry per output mode) for the input
    occupation through the transmission matrix T (lossy or not), with a
    scalar uniform particle overlap or a Gram matrix of the particles'
    internal states (first-quantized order): the coefficients of
    perm(B_loss + sum_m x_m B_m) divided by the input norm
    (probabilities.py:244-384).  Outputs with more photons than the input
    get 0; with no input photons the vacuum has probability 1.

    By default B_loss = G o (I - T^dag T) as in the reference.  Then
    sum_m B_m = G o (T^dag T)^T, so when T^dag T is not real (losses after a
    complex interferometer) the outputs with lost photons are off and the
    table does not sum to 1; ``conjugate_loss_kernel=True`` uses
    G o conj(I - T^dag T), which reproduces the probabilities of the unitary
    dilation of T.  Both agree whenever T^dag T is real (unitary T, uniform
    losses, losses before the interferometer)."""
    t = np.asarray(transmission_matrix, dtype=np.complex128)
    if t.ndim != 2:
        raise ValueError("the transmission matrix must be 2-D")
    inp = np.asarray(input_occupation, dtype=np.int64).reshape(-1)
    outs = np.asarray(occupation_numbers)
    if outs.ndim == 1:
        outs = outs[None, :]
    if inp.shape[0] != t.shape[1] or outs.shape[1] != t.shape[0]:
        raise ValueError("occupations must have one entry per input / output mode of T")
    if (inp < 0).any():
        raise ValueError("occupation numbers must be non-negative")
    n = int(inp.sum())
    if n == 0:
        return (outs.sum(axis=1) == 0).astype(np.float64)
    b0, bms, g = _lossy_matrices(t, inp, particle_overlap, conjugate_loss_kernel)
    if np.isscalar(particle_overlap):
        norm = complex(uniform_input_norm(inp, float(particle_overlap)))
    else:
        norm = general_input_norm(inp, g)
    # outputs with more photons than the input (and negative entries, which
    # the planner rejects) are handled in the C planner: their coefficient is 0
    return np.real(perm_poly_coefficients(b0, bms, outs) / norm)


def get_lossy_partially_distinguishable_detection_probabilities(
        occupation_numbers, transmission_matrix, input_occupation, particle_overlap,
        connector=None) -> np.ndarray:
    """The reference's name and argument order
    (probabilities.py:244-250); ``connector`` is accepted and ignored."""
    return lossy_detection_probabilities(occupation_numbers, transmission_matrix,
                                         input_occupation, particle_overlap)
