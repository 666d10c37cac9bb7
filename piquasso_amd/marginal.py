"""Exact marginal output probabilities and marginal sampling of ideal (or
uniformly lossy) boson sampling, from one matrix-polynomial permanent.

piquasso obtains the marginal distribution on k selected output modes T from
the binomial moments b_alpha = E[prod_j C(X_j, alpha_j)] of their photon
numbers (piquasso/_simulators/passive/marginal.py:117-470), which it builds
as products of Laguerre polynomials.  The same numbers are the coefficients
of a permanent polynomial (that module's own description, :26-50):

    sum_alpha b_alpha c^alpha = (1 / r!) Per( (I + V diag(c) V^dag)_{r, r} ),
    V[a, j] = conj(U[T_j, a]),

with rows and columns repeated by the input occupations r.  Expanded over
the n input particles (first-quantized modes a_k), the matrix polynomial is
B_0 + sum_j c_j B_j with B_0[k, l] = [a_k = a_l] and
B_j[k, l] = conj(U[T_j, a_k]) U[T_j, a_l], so all b_alpha with |alpha| <= n
come from ONE ``pqp_perm_poly_coefficients_c128`` call (Ryser over the
particle subsets on the GPU).  The probabilities then follow from the
binomial transform

    Pr(Y = y) = sum_{beta, |y + beta| <= n} b_{y + beta} prod_j C(y_j + beta_j, y_j) (-1)^beta_j,

and the marginal sampler draws mode by mode from conditional probabilities
exactly as the reference's ``generate_marginal_samples``
(piquasso/_simulators/passive/sampling.py:796-857).
"""

from __future__ import annotations

import math
from typing import Dict, Optional, Sequence, Tuple

import numpy as np

from .distinguishability import perm_poly_coefficients
from .sampling import fock_space_basis

__all__ = ["binomial_moments", "marginal_fock_probabilities", "get_marginal_fock_probabilities",
           "generate_marginal_samples"]


class _Moments:
    """b_alpha over the Fock basis of k modes with at most n photons
    (piquasso's order), with a vectorised rank lookup of compositions."""

    def __init__(self, comps: np.ndarray, values: np.ndarray, n: int):
        self.comps = comps
        self.values = values
        self.n = n
        self.k = comps.shape[1]
        self.base = np.int64(n + 1)
        self.weights = self.base ** np.arange(self.k, dtype=np.int64)
        keys = comps.astype(np.int64) @ self.weights
        self.order = np.argsort(keys, kind="stable")
        self.sorted_keys = keys[self.order]
        self.degree = comps.sum(axis=1)
        top = n + 1
        self.binom = np.array([[math.comb(a, b) for b in range(top)] for a in range(top)],
                              dtype=np.float64)

    def index(self, alphas: np.ndarray) -> np.ndarray:
        keys = alphas.astype(np.int64) @ self.weights
        return self.order[np.searchsorted(self.sorted_keys, keys)]

    def probability(self, y: np.ndarray) -> float:
        """Pr(Y_1..Y_t = y) for the first t = len(y) modes (the others summed
        over), by the binomial transform over beta with |y + beta| <= n."""
        t = len(y)
        full = np.zeros(self.k, dtype=np.int64)
        full[:t] = y
        left = self.n - int(full.sum())
        if left < 0:
            return 0.0
        # beta on the first t modes only, zero on the others
        sel = (self.degree <= left) & (self.comps[:, t:].sum(axis=1) == 0)
        beta = self.comps[sel].astype(np.int64)
        alpha = beta + full
        w = np.prod(self.binom[alpha[:, :t], full[None, :t]], axis=1)
        sign = np.where(beta[:, :t].sum(axis=1) % 2 == 1, -1.0, 1.0)
        return float(np.sum(sign * w * self.values[self.index(alpha)]))


def binomial_moments(input_photons, interferometer, all_modes) -> Tuple[np.ndarray, np.ndarray]:
    """(compositions, b): every alpha over the modes ``all_modes`` with
    |alpha| <= n in piquasso's Fock-basis order, and b_alpha
    (``get_binomial_moments``, marginal.py:165-225), from one GPU call."""
    u = np.asarray(interferometer, dtype=np.complex128)
    inp = np.asarray(input_photons, dtype=np.int64).reshape(-1)
    modes = np.asarray(all_modes, dtype=np.int64).reshape(-1)
    if u.ndim != 2 or u.shape[0] != u.shape[1] or inp.shape[0] != u.shape[1]:
        raise ValueError("the interferometer must be square with one input occupation per mode")
    if (inp < 0).any():
        raise ValueError("occupation numbers must be non-negative")
    n = int(inp.sum())
    comps = fock_space_basis(len(modes), n + 1)
    if n == 0:
        return comps, np.ones(len(comps))
    a = np.repeat(np.arange(len(inp)), inp)
    b0 = (a[:, None] == a[None, :]).astype(np.complex128)
    rows = u[np.ix_(modes, a)]  # rows[j, k] = U[T_j, a_k]
    bms = np.conj(rows)[:, :, None] * rows[:, None, :]
    coef = perm_poly_coefficients(b0, bms, comps)
    r_fact = float(np.prod([math.factorial(int(x)) for x in inp]))
    return comps, coef.real / r_fact


def marginal_fock_probabilities(input_photons, interferometer, postselected_modes: Sequence[int],
                                postselected_photons: Sequence[int],
                                marginal_modes: Sequence[int]) -> Dict[Tuple[int, ...], float]:
    """``get_marginal_fock_probabilities`` (marginal.py:117-162): the joint
    probabilities of every outcome on ``marginal_modes`` (at most n minus the
    postselected photons) together with the postselection, as a dict in
    piquasso's Fock-basis order of the outcomes."""
    ps_modes = tuple(int(m) for m in postselected_modes)
    ps_photons = np.asarray(postselected_photons, dtype=np.int64).reshape(-1)
    mg = tuple(int(m) for m in marginal_modes)
    if set(ps_modes) & set(mg):
        raise ValueError("marginal probabilities cannot be taken on postselected modes")
    if len(ps_photons) != len(ps_modes):
        raise ValueError("one postselected photon number per postselected mode")
    comps, b = binomial_moments(input_photons, interferometer, ps_modes + mg)
    n = int(np.sum(input_photons))
    mom = _Moments(comps, b, n)
    outcomes = fock_space_basis(len(mg), n - int(ps_photons.sum()) + 1)
    out = {}
    for y in outcomes:
        full = np.concatenate([ps_photons, y.astype(np.int64)])
        out[tuple(int(v) for v in y)] = mom.probability(full)
    return out


def get_marginal_fock_probabilities(input_photons, interferometer, postselected_modes,
                                    postselected_photons, marginal_modes):
    """The reference's name and argument order (marginal.py:117-124)."""
    return marginal_fock_probabilities(input_photons, interferometer, postselected_modes,
                                       postselected_photons, marginal_modes)


def generate_marginal_samples(input, interferometer, modes, shots: int, config=None,
                              postselect_data: Sequence = ((), ()), *, rng=None,
                              seed_sequence: Optional[int] = None):
    """Samples of piquasso's ``generate_marginal_samples``
    (sampling.py:796-857) on ``modes``: for every shot and selected mode,
    one ``rng.random()`` threshold against the running sum of the
    conditional probabilities of 0, 1, ... photons given the modes before
    (and the postselection).  The generator is ``rng``, else ``config.rng``,
    else ``default_rng(seed_sequence)``.  Prefix probabilities are memoised
    across shots; each is the binomial transform of the GPU's moments."""
    if rng is None:
        rng = getattr(config, "rng", None) if config is not None else None
    if rng is None:
        rng = np.random.default_rng(seed_sequence)
    n = int(np.sum(input))
    mg = tuple(int(m) for m in modes)
    ps_modes = tuple(int(m) for m in postselect_data[0])
    ps_photons = np.asarray(postselect_data[1], dtype=np.int64).reshape(-1)
    comps, b = binomial_moments(input, interferometer, ps_modes + mg)
    mom = _Moments(comps, b, n)
    memo: Dict[Tuple[int, ...], float] = {}

    def prob(prefix: np.ndarray) -> float:
        key = tuple(int(v) for v in prefix)
        if key not in memo:
            memo[key] = mom.probability(np.asarray(prefix, dtype=np.int64))
        return memo[key]

    ps_prob = prob(ps_photons)
    offset_p = int(ps_photons.sum())
    offset_m = len(ps_photons)
    samples = []
    for _ in range(int(shots)):
        previous = ps_prob
        remaining = n - offset_p
        particles = np.concatenate([ps_photons, np.zeros(len(mg), dtype=np.int64)])
        for j in range(len(mg)):
            threshold = rng.random()
            cumulative = 0.0
            pos = offset_m + j
            probability = previous
            for count in range(remaining + 1):
                particles[pos] = count
                probability = prob(particles[:pos + 1])
                cumulative += probability / previous
                if cumulative >= threshold:
                    break
            remaining -= count
            previous = probability
        samples.append(tuple(int(v) for v in particles[offset_m:]))
    return samples
