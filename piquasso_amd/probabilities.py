"""Batched output probabilities of a linear-optical (boson-sampling) circuit.

For an m-mode interferometer U, an input occupation s and an output
occupation t (both of length m, same photon number), the transition
amplitude is perm(U; rows=t, cols=s) / sqrt(prod t! prod s!) and the
probability its squared modulus — the quantity piquasso's passive simulator
evaluates one permanent at a time (piquasso/_simulators/passive/utils.py:131-138,
piquasso/_simulators/passive/probabilities.py:26-54).  Here all requested
output patterns go to the GPU as one batch (:class:`PermanentBatch`).
"""

from __future__ import annotations

import math
from typing import Sequence

import numpy as np

from .engine import PermanentBatch

__all__ = ["output_amplitudes", "output_probabilities"]


def _norm(occupation) -> float:
    return float(np.prod([math.factorial(int(k)) for k in occupation]))


def output_amplitudes(unitary, input_state: Sequence[int], output_states,
                      precision: str = "fp64") -> np.ndarray:
    """Amplitudes <t| U |s> for every output pattern t (rows of
    `output_states`), as a complex128 array."""
    u = np.ascontiguousarray(np.asarray(unitary, dtype=np.complex128))
    if u.ndim != 2 or u.shape[0] != u.shape[1]:
        raise ValueError(f"unitary must be square, got shape {u.shape}")
    s = np.asarray(input_state, dtype=np.int64).reshape(-1)
    outs = np.asarray(output_states, dtype=np.int64)
    if outs.ndim == 1:
        outs = outs[None, :]
    m = u.shape[0]
    if s.shape[0] != m or outs.shape[1] != m:
        raise ValueError("occupation vectors must have one entry per mode")
    if (s < 0).any() or (outs < 0).any():
        raise ValueError("occupation numbers must be non-negative")
    n = int(s.sum())
    amps = np.zeros(outs.shape[0], dtype=np.complex128)
    ok = outs.sum(axis=1) == n  # other photon numbers have amplitude 0
    idx = np.flatnonzero(ok)
    if idx.size:
        sel = np.ascontiguousarray(outs[idx], dtype=np.int32)
        with PermanentBatch([u] * idx.size, list(sel), [s] * idx.size, precision=precision) as b:
            b.run("eager")
            perms = b.results()
        # prod t! over the output pattern, from a factorial table (n <= 64)
        fact = np.array([math.factorial(k) for k in range(n + 1)], dtype=np.float64)
        out_norm = np.prod(fact[sel], axis=1)
        amps[idx] = perms / np.sqrt(out_norm * _norm(s))
    return amps


def output_probabilities(unitary, input_state: Sequence[int], output_states,
                         precision: str = "fp64") -> np.ndarray:
    """Probabilities |<t| U |s>|^2 for every output pattern t."""
    return np.abs(output_amplitudes(unitary, input_state, output_states, precision)) ** 2
