"""Gradients of permanents, evaluated on the GPU.

For perm(A; r, c) (a holomorphic polynomial in the entries of A)

    G_ij = d perm / d A_ij = r_i c_j perm(A; r - e_i, c - e_j)

(0 where r_i or c_j is 0), the quantity ``grad_perm`` computes
(src/permanent.cpp:271-300) and the JAX extension's backward multiplies by
the cotangent (src/jax_perm/jax_perm_core.cpp:141-208,
piquasso/jax_extensions/permanent.py:73-143).  Here every row i with r_i > 0
is one column-minor item: a single Gray sweep yields perm(A; r - e_i, c - e_j)
for all j at once, and all items of all permanents share one launch.  The
reference's routine assumes a square matrix; this one takes any m_r x m_c.

* :func:`permanent_grad` — G for one permanent;
* :func:`permanent_grad_batch` — permanents and G for many;
* :class:`PermanentBatchFunction` / :func:`torch_permanent_batch` — the
  same for a stack of matrices (B, m_r, m_c) with per-item multiplicities,
  all permanents and gradients from one engine call (the batched backward
  of the JAX extension, src/jax_perm/cuda/kernels.cc.cu:431-630);
* :class:`PermanentFunction` / :func:`torch_permanent` — a
  ``torch.autograd.Function`` with PyTorch's convention for complex inputs:
  ``A.grad = grad_output * conj(G)`` (for the loss |perm|^2 this is
  ``2 perm conj(G)``), the conjugate of JAX's holomorphic
  ``ct * G`` (piquasso/jax_extensions/permanent.py:120-141).
"""

from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np

from . import _capi

__all__ = ["permanent_grad", "permanent_grad_batch", "PermanentFunction", "torch_permanent",
           "PermanentBatchFunction", "torch_permanent_batch"]

_dp = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)


def _prep(matrix, rows, cols):
    a = np.ascontiguousarray(np.asarray(matrix, dtype=np.complex128))
    if a.ndim != 2:
        raise ValueError(f"matrix must be 2-D, got shape {a.shape}")
    r = np.ascontiguousarray(np.asarray(rows).astype(np.int32, casting="unsafe").reshape(-1))
    c = np.ascontiguousarray(np.asarray(cols).astype(np.int32, casting="unsafe").reshape(-1))
    if r.shape[0] != a.shape[0] or c.shape[0] != a.shape[1]:
        raise ValueError(
            f"rows/cols lengths {r.shape[0]}/{c.shape[0]} do not match matrix shape {a.shape}"
        )
    return a, r, c


def permanent_grad_batch(matrices: Sequence, rows: Sequence, cols: Sequence):
    """Permanents and gradients of many (matrix, rows, cols) triples.

    Returns ``(perms, grads)``: a complex128 array of the permanents and a
    list of complex128 arrays, ``grads[k]`` of the shape of ``matrices[k]``.
    Raises ``RuntimeError`` when an item's multiplicities do not sum to the
    same number (as ``permanent`` does)."""
    if not (len(matrices) == len(rows) == len(cols)):
        raise ValueError("matrices, rows and cols must have the same length")
    items = [_prep(m, r, c) for m, r, c in zip(matrices, rows, cols)]
    n = len(items)
    grads = [np.zeros(a.shape, dtype=np.complex128) for a, _, _ in items]
    perms = np.zeros(max(n, 1), dtype=np.complex128)
    if n == 0:
        return perms[:0], grads
    mats = (_dp * n)(*[a.view(np.float64).ctypes.data_as(_dp) for a, _, _ in items])
    rws = (_i32p * n)(*[r.ctypes.data_as(_i32p) for _, r, _ in items])
    cls = (_i32p * n)(*[c.ctypes.data_as(_i32p) for _, _, c in items])
    outs = (_dp * n)(*[g.view(np.float64).ctypes.data_as(_dp) for g in grads])
    m_r = np.array([a.shape[0] for a, _, _ in items], dtype=np.int64)
    m_c = np.array([a.shape[1] for a, _, _ in items], dtype=np.int64)
    _capi.check(
        _capi.lib().pqp_grad_batch_c128(
            n, mats, m_r.ctypes.data_as(_i64p), m_c.ctypes.data_as(_i64p),
            m_c.ctypes.data_as(_i64p), rws, cls, perms.view(np.float64).ctypes.data_as(_dp), outs,
        )
    )
    return perms[:n], grads


def permanent_grad(matrix, rows, cols) -> np.ndarray:
    """d perm(A; rows, cols) / d A as a complex128 matrix of A's shape."""
    _, grads = permanent_grad_batch([matrix], [rows], [cols])
    return grads[0]


def _torch():
    import torch

    return torch


class PermanentFunction:
    """``PermanentFunction.apply(A, rows, cols)``: the permanent of a complex
    torch tensor A with integer multiplicities as a differentiable op.

    The forward evaluates the permanent and the gradient matrix G in one
    call to the GPU engine and keeps G for the backward, which returns
    ``grad_output * conj(G)`` (PyTorch's convention for a holomorphic
    function of complex inputs).  A may live on any device; the result is a
    0-d tensor of A's dtype on A's device."""

    _impl = None

    @classmethod
    def _cls(cls):
        if cls._impl is None:
            torch = _torch()

            class _Perm(torch.autograd.Function):
                @staticmethod
                def forward(ctx, A, rows, cols):
                    a = A.detach().to("cpu", torch.complex128).numpy()
                    r = np.asarray(rows.detach().cpu().numpy() if torch.is_tensor(rows) else rows)
                    c = np.asarray(cols.detach().cpu().numpy() if torch.is_tensor(cols) else cols)
                    if A.requires_grad:
                        perms, grads = permanent_grad_batch([a], [r], [c])
                        ctx.save_for_backward(torch.from_numpy(grads[0]).to(A.device, A.dtype))
                        value = perms[0]
                    else:
                        from ._math.permanent import permanent

                        value = complex(permanent(a, r, c))
                        ctx.save_for_backward(None)
                    return torch.tensor(value, dtype=A.dtype, device=A.device)

                @staticmethod
                def backward(ctx, grad_out):
                    (G,) = ctx.saved_tensors
                    if G is None:
                        return None, None, None
                    return grad_out * G.conj(), None, None

            cls._impl = _Perm
        return cls._impl

    @classmethod
    def apply(cls, A, rows, cols):
        return cls._cls().apply(A, rows, cols)


def torch_permanent(A, rows, cols):
    """Differentiable permanent of a complex torch tensor (see
    :class:`PermanentFunction`)."""
    return PermanentFunction.apply(A, rows, cols)


class PermanentBatchFunction:
    """``PermanentBatchFunction.apply(A, rows, cols)``: permanents of a stack
    of complex matrices A (B, m_r, m_c) with multiplicities rows (B, m_r) and
    cols (B, m_c) as one differentiable op; returns a (B,) tensor of A's
    dtype on A's device.  With ``A.requires_grad`` the forward evaluates all
    permanents and gradient matrices in one ``pqp_grad_batch_c128`` call and
    the backward returns ``grad_output[:, None, None] * conj(G)``."""

    _impl = None

    @classmethod
    def _cls(cls):
        if cls._impl is None:
            torch = _torch()

            def _host(x):
                return np.asarray(x.detach().cpu().numpy() if torch.is_tensor(x) else x)

            class _PermBatch(torch.autograd.Function):
                @staticmethod
                def forward(ctx, A, rows, cols):
                    if A.dim() != 3:
                        raise ValueError(f"expected a (B, m_r, m_c) tensor, got shape {tuple(A.shape)}")
                    a = A.detach().to("cpu", torch.complex128).numpy()
                    r = _host(rows)
                    c = _host(cols)
                    B = a.shape[0]
                    if r.shape != (B, a.shape[1]) or c.shape != (B, a.shape[2]):
                        raise ValueError("rows / cols must have shapes (B, m_r) / (B, m_c)")
                    if A.requires_grad:
                        perms, grads = permanent_grad_batch(list(a), list(r), list(c))
                        G = np.stack(grads) if B else np.zeros(a.shape, dtype=np.complex128)
                        ctx.save_for_backward(torch.from_numpy(G).to(A.device, A.dtype))
                    else:
                        from .engine import permanent_batch

                        perms = permanent_batch(list(a), list(r), list(c)) if B else np.zeros(0, complex)
                        ctx.save_for_backward(None)
                    return torch.from_numpy(np.asarray(perms, dtype=np.complex128)).to(A.device, A.dtype)

                @staticmethod
                def backward(ctx, grad_out):
                    (G,) = ctx.saved_tensors
                    if G is None:
                        return None, None, None
                    return grad_out[:, None, None] * G.conj(), None, None

            cls._impl = _PermBatch
        return cls._impl

    @classmethod
    def apply(cls, A, rows, cols):
        return cls._cls().apply(A, rows, cols)


def torch_permanent_batch(A, rows, cols):
    """Differentiable permanents of a stack of complex matrices (see
    :class:`PermanentBatchFunction`)."""
    return PermanentBatchFunction.apply(A, rows, cols)
