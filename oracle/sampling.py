"""ORACLE — test infrastructure, not product code.

A sequential CPU restatement of piquasso's Clifford & Clifford boson sampler
(algorithm B of arXiv:1706.01260 with the generalisation of Brod & Oszmaniec
for multiply occupied inputs), as the reference implements it in
piquasso/_simulators/passive/sampling.py.  The shots run one after the other
and each column-minor problem goes to :func:`oracle.oracle.laplace_dd` (the
double-double restatement of ``permanent_laplace_cpp``) unless another
``laplace`` callable is given.

Random numbers are drawn in the reference's order from the reference's
generators: one ``numpy.random.default_rng(seed + shot)`` per shot
(sampling.py:178-193) and the caller's ``reject_condition`` (which in the
reference draws from the simulator's own generator).  With equal seeds the
samples equal the reference's, which pins this restatement against the seeded
sample lists of the reference's tests
(tests/_simulators/passive/test_measurements.py:237-563).

Only ``tests/`` imports this module.
"""

from __future__ import annotations

import math
import random
from typing import Callable, Optional, Sequence

import numpy as np

from .oracle import laplace_dd

__all__ = ["generate_samples", "generate_lossy_samples", "expanded_lossy_matrix", "result_order"]


class SamplingAborted(RuntimeError):
    """More postselection trials than allowed (the reference raises its
    InvalidSimulation there, sampling.py:275-283)."""


def _first_quantized(occupation) -> np.ndarray:
    """Mode index of every particle, modes ascending (piquasso/_math/indices.py:105-114)."""
    return np.repeat(np.arange(len(occupation)), np.asarray(occupation, dtype=np.int64))


def _pmf(current_input, sample, interferometer, laplace):
    """Distribution of the next output mode (sampling.py:711-749): the column
    minors of U restricted to occupied output rows / input columns, weighted
    by the input multiplicity and by U's entry in every candidate row."""
    in_mask = current_input > 0
    out_mask = sample > 0
    sub = interferometer[np.ix_(out_mask, in_mask)]
    minors = laplace(sub, sample[out_mask], current_input[in_mask])
    in_modes = np.flatnonzero(in_mask)
    d = len(sample)
    pmf = np.empty(d)
    for row in range(d):
        amp = 0j
        for j, mode in enumerate(in_modes[: len(minors)]):
            amp += current_input[mode] * minors[j] * interferometer[row, mode]
        pmf[row] = abs(amp) ** 2
    return pmf / pmf.sum()


def _grow(current_input, pool, rng):
    """One particle of the pool joins the input (sampling.py:197-205)."""
    pick = rng.choice(len(pool))
    current_input[pool[pick]] += 1
    return np.delete(pool, pick)


def _one_sample(d, n, pool, interferometer, rng, reject, laplace):
    """sampling.py:208-236."""
    sample = np.zeros(d, dtype=int)
    current_input = np.zeros(d, dtype=int)
    pool = np.array(pool, dtype=np.int64)
    for _ in range(n):
        if reject():
            continue
        pool = _grow(current_input, pool, rng)
        pmf = _pmf(current_input, sample, interferometer, laplace)
        sample[rng.choice(np.arange(d), p=pmf)] += 1
    return sample


def _one_sample_postselected(d, n, pool, interferometer, rng, reject, laplace,
                             ps_modes, ps_photons, max_trials, track_needed=True):
    """sampling.py:239-325: restart the sample whenever the postselected modes
    receive more photons than asked, or can no longer receive enough."""
    ps_modes = list(ps_modes)
    needed0 = int(np.sum(ps_photons))
    trials = 0
    while True:
        trials += 1
        if trials > max_trials:
            raise SamplingAborted("Too many trials during sample generation.")
        current_input = np.zeros(d, dtype=int)
        sample = np.zeros(d, dtype=int)
        needed = needed0
        left = np.array(ps_photons, dtype=int)
        remaining_pool = np.array(pool, dtype=np.int64)
        failed = False
        for k in range(1, n + 1):
            if reject():
                continue
            remaining_pool = _grow(current_input, remaining_pool, rng)
            pmf = _pmf(current_input, sample, interferometer, laplace)
            mode = rng.choice(np.arange(d), p=pmf)
            sample[mode] += 1
            if mode in ps_modes:
                needed -= 1
                left[ps_modes.index(mode)] -= 1
                if np.any(left < 0):
                    failed = True
                    break
            if track_needed and needed > n - k:
                failed = True
                break
        if not failed:
            break
    if not track_needed:
        return sample
    return np.delete(sample, ps_modes)


def _split_particles(occupation, overlap, rng):
    """Indistinguishable / distinguishable split of every input mode's
    particles (sampling.py:371-399): K of n_j particles stay
    indistinguishable with weight C(n_j, K) x^K (1 - x)^(n_j - K) K!."""
    indist = np.zeros(len(occupation), dtype=int)
    dist = np.zeros(len(occupation), dtype=int)
    for j, nj in enumerate(occupation):
        nj = int(nj)
        if nj == 0:
            continue
        w = np.array([math.comb(nj, q) * overlap ** q * (1.0 - overlap) ** (nj - q) * math.factorial(q)
                      for q in range(nj + 1)], dtype=float)
        q = rng.choice(nj + 1, p=w / w.sum())
        indist[j] = q
        dist[j] = nj - q
    return indist, dist


def _distinguishable_outputs(interferometer, particles, rng, reject):
    """Each distinguishable particle leaves independently by |U_{:,j}|^2
    (sampling.py:402-419)."""
    d = len(particles)
    out = np.zeros(d, dtype=int)
    for mode, count in enumerate(particles):
        p = np.abs(interferometer[:, mode]) ** 2
        p = p / p.sum()
        for _ in range(int(count)):
            target = rng.choice(d, p=p)
            if not reject():
                out[target] += 1
    return out


def times_linear(poly, const, lin, out):
    """(const + sum_a lin[a] x_a) * poly, truncated to poly's shape, as
    multiply_by_linear_truncated (piquasso/_math/polynomial.py:18-38) writes
    it: out = const * poly, then out += lin[a] * (poly shifted along axis a)
    axis by axis.  With out = poly (the in-place call of sampling.py:545-548)
    each shift reads the values already updated."""
    out[...] = const * poly
    for axis, c in enumerate(lin):
        np.moveaxis(out, axis, 0)[1:] += c * np.moveaxis(poly, axis, 0)[:-1]
    return out


def _ps_probability(u, dist, modes, remaining, in_place=True):
    """sampling.py:532-551.  in_place: the reference's in-place polynomial
    updates (each factor's shifted terms read values already scaled by the
    constant); False: the product with a fresh copy per factor, the exact
    probability that the distinguishable particles supply ``remaining``."""
    poly = np.zeros(tuple(np.asarray(remaining) + 1))
    poly[(0,) * len(remaining)] = 1.0
    for mode, mult in enumerate(dist):
        p = np.abs(u[modes, mode]) ** 2
        for _ in range(int(mult)):
            poly = times_linear(poly, 1.0 - p.sum(), p, out=poly if in_place else np.empty_like(poly))
    return poly[tuple(remaining)]


def _ps_table(u, dist, modes, bound):
    """sampling.py:554-582."""
    fq = _first_quantized(dist)
    table = [np.zeros(tuple(np.asarray(bound) + 1)) for _ in range(len(fq) + 1)]
    table[-1][(0,) * len(bound)] = 1.0
    for k in range(len(fq) - 1, -1, -1):
        p = np.abs(u[modes, fq[k]]) ** 2
        times_linear(table[k + 1], 1.0 - p.sum(), p, out=table[k])
    return table


def _ps_conditioned(u, dist, modes, remaining, table, rng):
    """sampling.py:585-655: each distinguishable particle to a free mode, to
    loss, or to a postselected mode still short of photons."""
    d = u.shape[0]
    free = [m for m in range(d) if m not in set(modes)]
    out = np.zeros(d, dtype=int)
    left = np.array(remaining, dtype=int)
    for k, mode in enumerate(_first_quantized(dist)):
        nxt = table[k + 1]
        fut = nxt[tuple(left)]
        pf = np.abs(u[free, mode]) ** 2
        pp = np.abs(u[list(modes), mode]) ** 2
        # clamped at 0 like the product sampler: the reference's choice()
        # raises when rounding leaves this below 0, and is unchanged otherwise
        ploss = max(1.0 - pf.sum() - pp.sum(), 0.0)
        wp = np.zeros(len(modes))
        for a in range(len(modes)):
            if left[a] > 0:
                left[a] -= 1
                wp[a] = pp[a] * nxt[tuple(left)]
                left[a] += 1
        w = np.concatenate([pf * fut, np.array([ploss * fut]), wp])
        w /= np.sum(w)
        i = rng.choice(len(w), p=w)
        if i < len(free):
            out[free[i]] += 1
        elif i > len(free):
            out[modes[i - len(free) - 1]] += 1
            left[i - len(free) - 1] -= 1
    return out


def _one_sample_postselected_overlap(d, occ, u, rng, reject, laplace, modes, photons, max_trials,
                                     overlap, in_place=True):
    """sampling.py:422-529: split, one postselection attempt for the
    indistinguishable particles, accept the distinguishable ones with their
    (in-place evaluated) postselection probability, place them conditioned;
    restart on failure."""
    modes = list(modes)
    photons = np.asarray(photons, dtype=int)
    trials = 0
    while True:
        trials += 1
        if trials > max_trials:
            raise SamplingAborted("Too many trials during sample generation.")
        indist, dist = _split_particles(occ, overlap, rng)
        try:
            ind = _one_sample_postselected(d, int(indist.sum()), _first_quantized(indist), u, rng,
                                           reject, laplace, modes, photons, 1, track_needed=False)
        except SamplingAborted:
            continue
        remaining = photons - ind[modes]
        if rng.random() > _ps_probability(u, dist, modes, remaining, in_place):
            continue
        table = _ps_table(u, dist, modes, photons)
        out = _ps_conditioned(u, dist, modes, remaining, table, rng) + ind
        return np.delete(out, modes)


def generate_samples(input_occupation, shots: int, interferometer, seed: int,
                     reject_condition: Optional[Callable[[], bool]] = None,
                     postselect: Sequence = ((), (), 1000),
                     uniform_particle_overlap: Optional[float] = None,
                     laplace=laplace_dd, ps_acceptance_in_place: bool = True):
    """Samples of ``generate_samples`` (sampling.py:34-107) for the ideal,
    uniformly lossy (``reject_condition``), postselected and uniformly
    partially distinguishable cases, and postselection with uniform partial
    distinguishability; one shot after another.  ``ps_acceptance_in_place``
    False evaluates the distinguishable particles' postselection probability
    exactly instead of in place as the reference does (the product sampler
    deliberately does the former)."""
    occ = np.asarray(input_occupation, dtype=int)
    d = len(occ)
    n = int(occ.sum())
    u = np.asarray(interferometer, dtype=np.complex128)
    reject = reject_condition or (lambda: False)
    ps_modes, ps_photons, max_trials = postselect
    pool = _first_quantized(occ)
    samples = []
    for shot in range(shots):
        rng = np.random.default_rng(seed=seed + shot)
        if uniform_particle_overlap is None:
            if len(ps_modes):
                s = _one_sample_postselected(d, n, pool, u, rng, reject, laplace,
                                             ps_modes, ps_photons, max_trials)
            else:
                s = _one_sample(d, n, pool, u, rng, reject, laplace)
        else:
            if len(ps_modes):
                s = _one_sample_postselected_overlap(d, occ, u, rng, reject, laplace, ps_modes,
                                                     ps_photons, max_trials,
                                                     float(uniform_particle_overlap),
                                                     ps_acceptance_in_place)
                samples.append(tuple(int(v) for v in s))
                continue
            # sampling.py:328-368
            indist, dist = _split_particles(occ, uniform_particle_overlap, rng)
            s = _distinguishable_outputs(u, dist, rng, reject) + _one_sample(
                d, int(indist.sum()), _first_quantized(indist), u, rng, reject, laplace)
        samples.append(tuple(int(v) for v in s))
    return samples


def expanded_lossy_matrix(interferometer) -> np.ndarray:
    """The 2d x 2d unitary dilation of a lossy interferometer used by
    ``generate_lossy_samples`` (sampling.py:756-784): with T = V diag(s) W
    (numpy's SVD), [[V, 0], [0, I]] [[diag s, diag t], [diag t, diag s]]
    [[W, 0], [0, I]], t = sqrt(1 - s^2)."""
    v, s, w = np.linalg.svd(np.asarray(interferometer))
    d = len(s)
    z = np.zeros((d, d))
    eye = np.eye(d)
    t = np.sqrt(1.0 - s ** 2)
    big_v = np.block([[v, z], [z, eye]])
    big_w = np.block([[w, z], [z, eye]])
    mid = np.block([[np.diag(s), np.diag(t)], [np.diag(t), np.diag(s)]])
    return big_v @ mid @ big_w


def generate_lossy_samples(input_occupation, shots: int, interferometer, seed: int,
                           postselect: Sequence = ((), (), 1000), laplace=laplace_dd):
    """``generate_lossy_samples`` (sampling.py:110-146): sample the dilation
    with empty auxiliary inputs, keep the first d output modes."""
    occ = np.asarray(input_occupation, dtype=int)
    d = len(occ)
    big = expanded_lossy_matrix(interferometer)
    full = np.concatenate([occ, np.zeros(d, dtype=int)])
    out = generate_samples(full, shots, big, seed, postselect=postselect, laplace=laplace)
    return [s[:d] for s in out]


def result_order(samples, seed):
    """The order in which the reference's ``Result.samples`` lists the
    sampler's output: binned by first occurrence (piquasso/_utils.py:23-30,
    simulation_steps.py:393-396), each bin repeated by its count, then
    shuffled with ``random.Random(seed)`` (piquasso/api/result.py:124-136)."""
    counts = {}
    for s in samples:
        counts[s] = counts.get(s, 0) + 1
    flat = []
    for s, k in counts.items():
        flat.extend([s] * k)
    random.Random(seed).shuffle(flat)
    return flat
