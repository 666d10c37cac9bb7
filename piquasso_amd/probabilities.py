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

__all__ = ["output_amplitudes", "output_probabilities", "particle_number_probabilities",
           "calculate_inner_product", "get_ideal_particle_number_probability"]


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


def particle_number_probabilities(output_states, interferometer, occupation_numbers, coefficients,
                                  precision: str = "fp64") -> np.ndarray:
    """|sum_j c_j <t| U |s_j>|^2 for every output pattern t (rows of
    ``output_states``) of the superposition sum_j c_j |s_j> of input Fock
    states (rows of ``occupation_numbers``): the reference's
    ``get_ideal_particle_number_probability`` (probabilities.py:26-54) for
    many outputs, every (output, input) pair of equal photon number in one
    batched launch."""
    u = np.ascontiguousarray(np.asarray(interferometer, dtype=np.complex128))
    if u.ndim != 2 or u.shape[0] != u.shape[1]:
        raise ValueError(f"interferometer must be square, got shape {u.shape}")
    m = u.shape[0]
    ins = np.asarray(occupation_numbers, dtype=np.int64)
    if ins.ndim == 1:
        ins = ins[None, :]
    outs = np.asarray(output_states, dtype=np.int64)
    if outs.ndim == 1:
        outs = outs[None, :]
    coefs = np.asarray(coefficients, dtype=np.complex128).reshape(-1)
    if ins.shape[1] != m or outs.shape[1] != m or coefs.shape[0] != ins.shape[0]:
        raise ValueError("one coefficient per input state and one entry per mode in every state")
    if (ins < 0).any() or (outs < 0).any():
        raise ValueError("occupation numbers must be non-negative")
    n_out, n_in = outs.sum(axis=1), ins.sum(axis=1)
    pair_out, pair_in = np.nonzero(n_out[:, None] == n_in[None, :])
    amps = np.zeros(outs.shape[0], dtype=np.complex128)
    if pair_out.size:
        rows = np.ascontiguousarray(outs[pair_out], dtype=np.int32)
        cols = np.ascontiguousarray(ins[pair_in], dtype=np.int32)
        with PermanentBatch([u] * pair_out.size, list(rows), list(cols), precision=precision) as b:
            b.run("eager")
            perms = b.results()
        top = int(max(outs.max(initial=0), ins.max(initial=0)))
        fact = np.array([math.factorial(k) for k in range(top + 1)], dtype=np.float64)
        norm = np.sqrt(np.prod(fact[rows], axis=1) * np.prod(fact[cols], axis=1))
        np.add.at(amps, pair_out, coefs[pair_in] * perms / norm)
    return np.abs(amps) ** 2


def calculate_inner_product(interferometer, input, output, connector=None) -> complex:
    """<output| U |input> = perm(U; rows=output, cols=input) /
    sqrt(prod output! prod input!) (piquasso/_simulators/passive/utils.py:131-138);
    ``connector`` is accepted and ignored."""
    return complex(output_amplitudes(interferometer, input, [output])[0])


def get_ideal_particle_number_probability(occupation_number, interferometer, occupation_numbers,
                                          coefficients, connector=None) -> float:
    """The reference's name and arguments (probabilities.py:26-54) for one
    output pattern; ``connector`` is accepted and ignored."""
    return float(particle_number_probabilities([occupation_number], interferometer,
                                               occupation_numbers, coefficients)[0])
