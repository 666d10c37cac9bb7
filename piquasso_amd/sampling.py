"""Clifford & Clifford boson sampling with the shots in lockstep on the GPU.

piquasso samples a linear-optical circuit with algorithm B of Clifford &
Clifford (arXiv:1706.01260), generalised to multiply occupied inputs: photon
by photon, the next output mode is drawn from a distribution built from the
column minors perm(U_sub; sample, input - e_l) of the interferometer
(piquasso/_simulators/passive/sampling.py:197-236,711-753).  The reference
runs the shots one after the other and calls ``permanent_laplace`` once per
shot and photon.  Here every shot advances one photon per round and each
round's minors of all shots come from ONE launch of the column-minor kernels
(``pqp_laplace_batch_c128``); the bookkeeping around them (input growth,
distributions, draws) is vectorised over the shots.

The samples equal the reference's for equal seeds.  Each shot draws from its
own ``numpy.random.default_rng(seed + shot)`` (sampling.py:178-193) in the
reference's order, so the batching cannot change them:

* the reference's ``rng.choice(len(pool))`` is drawn as
  ``rng.integers(0, len(pool), dtype=np.int64)`` and its
  ``rng.choice(np.arange(d), p=pmf)`` as one ``rng.random()`` located in the
  normalised cumulative sum (searchsorted, side='right') — the draws numpy's
  ``Generator.choice`` makes for those calls (checked draw for draw by the
  tests);
* ``reject_condition`` (uniform losses: the reference's closure draws from
  the simulator's generator, simulation_steps.py:352-360) is called for every
  shot before the rounds start, in the reference's shot-major order, as many
  times as the reference calls it (once per photon).  Where that count is not
  fixed in advance (postselection restarts a shot), the shots run one at a
  time instead.

Supported: ideal and uniformly lossy sampling (sampling.py:208-236),
postselection (:239-325), uniform partial distinguishability
(:328-419), postselection together with uniform partial distinguishability
(:422-655; shots one after another, restarts being data-dependent),
non-uniform losses through the 2d-mode unitary dilation (:110-146,756-784)
and, for any Gram matrix of the particles with losses,
``generate_lossy_and_partially_distinguishable_samples`` (:658-708) from the
GPU's probability table.
"""

from __future__ import annotations

import ctypes
import math
import os
import time
from typing import Callable, Optional, Sequence

import numpy as np

from . import _capi

__all__ = ["generate_samples", "generate_lossy_samples", "expanded_lossy_matrix",
           "generate_lossy_and_partially_distinguishable_samples", "fock_space_basis",
           "postselected_fock_basis", "SampleGenerationError", "last_run"]

# Counters of the most recent generate_samples call in this process: rounds,
# column-minor problems, the seconds spent in the batched minors calls
# (planning, upload, kernels, download) and in total, and, when
# ``count_flops`` is set, the minors' algorithmic FLOP (off by default: it
# costs host time every round).
last_run: dict = {}
count_flops = False


class SampleGenerationError(RuntimeError):
    """A postselected shot needed more trials than allowed (the reference
    raises its ``InvalidSimulation`` there, sampling.py:275-283)."""


def _seed_of(config, seed_sequence):
    if seed_sequence is not None:
        return int(seed_sequence)
    if config is not None and getattr(config, "seed_sequence", None) is not None:
        return int(config.seed_sequence)
    # piquasso's Config draws a fresh seed the same way when none is given
    return int.from_bytes(os.urandom(8), "big")


class _MinorsBatch:
    """Column minors of many (rows, cols) multiplicity patterns over ONE
    interferometer through ``pqp_laplace_batch_c128``: problem s is
    perm(U; rows_s, cols_s - e_l) for every column l; zero multiplicities
    are compacted away by the planner, so no sub-matrix is copied."""

    def __init__(self, u: np.ndarray):
        self.u = np.ascontiguousarray(u, dtype=np.complex128)
        self.d = self.u.shape[0]
        self.lib = _capi.lib()
        self._dp = ctypes.POINTER(ctypes.c_double)
        self._i32p = ctypes.POINTER(ctypes.c_int32)
        self._i64p = ctypes.POINTER(ctypes.c_int64)
        self.seconds = 0.0

    def __call__(self, rows: np.ndarray, cols: np.ndarray) -> np.ndarray:
        t0 = time.perf_counter()
        try:
            return self._run(rows, cols)
        finally:
            self.seconds += time.perf_counter() - t0

    # problems per pqp_laplace_batch_c128 call: bounds the host plan and the
    # device arena (~10 KB per 20-photon problem) whatever the shot count
    max_problems = 8192

    def _run(self, rows: np.ndarray, cols: np.ndarray) -> np.ndarray:
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        count, d = rows.shape
        if count > self.max_problems:
            return np.concatenate([self._run(rows[k:k + self.max_problems], cols[k:k + self.max_problems])
                                   for k in range(0, count, self.max_problems)])
        out = np.zeros((count, 2 * d), dtype=np.float64)
        if count == 0:
            return out.view(np.complex128)
        step_i = np.arange(count, dtype=np.uint64) * np.uint64(4 * d)
        mats = np.full(count, self.u.ctypes.data, dtype=np.uint64)
        rp = np.uint64(rows.ctypes.data) + step_i
        cp = np.uint64(cols.ctypes.data) + step_i
        op = np.uint64(out.ctypes.data) + np.arange(count, dtype=np.uint64) * np.uint64(16 * d)
        dims = np.full(count, d, dtype=np.int64)
        pp = ctypes.POINTER(self._dp)
        st = self.lib.pqp_laplace_batch_c128(
            count, mats.ctypes.data_as(pp), dims.ctypes.data_as(self._i64p),
            dims.ctypes.data_as(self._i64p), dims.ctypes.data_as(self._i64p),
            rp.ctypes.data_as(ctypes.POINTER(self._i32p)),
            cp.ctypes.data_as(ctypes.POINTER(self._i32p)), op.ctypes.data_as(pp),
        )
        _capi.check(st)
        return out.view(np.complex128)


_U32 = np.uint64(0xFFFFFFFF)


class _Draws:
    """The per-shot draws the sampler needs, ``Generator.integers(0, L,
    dtype=np.int64)`` and ``Generator.random()``, evaluated for many shots at
    once from blocks of each shot's raw PCG64 outputs (``random_raw``), by
    numpy's own rules: ``random()`` is the top 53 bits of the next 64-bit
    output times 2^-53; ``integers`` takes a 32-bit half through the bit
    generator's cached-half state (the low half of a fresh output, then its
    high half on the next 32-bit draw) and applies Lemire's
    multiply-and-reject method, and draws nothing for L = 1.  The values and
    their order equal the generator's own calls (tested draw for draw).  The
    generators' states run ahead by the unused part of the blocks, so the
    engine must be the generators' last user."""

    def __init__(self, rngs, block: int):
        S = len(rngs)
        self.bgs = [r.bit_generator for r in rngs]
        self.raw = np.empty((S, max(block, 1)), dtype=np.uint64)
        self.has32 = np.zeros(S, dtype=bool)
        self.cache = np.zeros(S, dtype=np.uint64)
        for k, bg in enumerate(self.bgs):
            st = bg.state
            self.has32[k] = bool(st["has_uint32"])
            self.cache[k] = np.uint64(st["uinteger"])
            self.raw[k] = bg.random_raw(self.raw.shape[1])
        self.ptr = np.zeros(S, dtype=np.int64)

    def _next64(self, shots):
        while (self.ptr[shots] >= self.raw.shape[1]).any():
            more = np.stack([bg.random_raw(self.raw.shape[1]) for bg in self.bgs])
            self.raw = np.concatenate([self.raw, more], axis=1)
        v = self.raw[shots, self.ptr[shots]]
        self.ptr[shots] += 1
        return v

    def _next32(self, shots):
        out = np.empty(len(shots), dtype=np.uint64)
        h = self.has32[shots]
        out[h] = self.cache[shots[h]]
        self.has32[shots[h]] = False
        fresh = shots[~h]
        if fresh.size:
            v = self._next64(fresh)
            out[~h] = v & _U32
            self.cache[fresh] = v >> np.uint64(32)
            self.has32[fresh] = True
        return out

    def doubles(self, shots) -> np.ndarray:
        return (self._next64(shots) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def integers(self, shots, bound) -> np.ndarray:
        L = np.asarray(bound, dtype=np.uint64)
        res = np.zeros(len(shots), dtype=np.int64)
        act = L > 1
        if not act.any():
            return res
        sh, Le = shots[act], L[act]
        m = self._next32(sh) * Le
        left = m & _U32
        thr = (np.uint64(1 << 32) - Le) % Le
        rej = (left < Le) & (left < thr)
        while rej.any():
            i = np.flatnonzero(rej)
            m[i] = self._next32(sh[i]) * Le[i]
            left[i] = m[i] & _U32
            rej[i] = left[i] < thr[i]
        res[act] = (m >> np.uint64(32)).astype(np.int64)
        return res


class _Lockstep:
    """S shots of Clifford & Clifford sampling advanced one photon per round.

    Per shot: ``pool`` (the input particles not yet placed, by mode, in the
    reference's order), ``cur`` (the input grown so far), ``smp`` (the
    output so far), ``step`` (photon iterations done of ``n``).  ``reject``
    maps an array of shot indices to one rejection decision each, in call
    order; ``postselect`` = (modes, photons, max_trials) or None.
    """

    def __init__(self, minors: _MinorsBatch, rngs, pools, reject, postselect=None,
                 track_needed: bool = True, fast_draws: bool = False):
        self.minors = minors
        self.u_t = np.ascontiguousarray(minors.u.T)
        d = minors.d
        S = len(pools)
        self.rngs = rngs
        self._ints = [r.integers for r in rngs]
        self._rand = [r.random for r in rngs]
        self.n = np.array([len(p) for p in pools], dtype=np.int64)
        width = max(int(self.n.max()) if S else 0, 1)
        self.pool0 = np.full((S, width), -1, dtype=np.int64)
        for s, p in enumerate(pools):
            self.pool0[s, : len(p)] = p
        self.reject = reject
        self.ps = postselect
        self.track_needed = track_needed
        # vectorised draws when this engine is the generators' last user
        self.draws = _Draws(rngs, int(self.n.max()) * 3 // 2 + 4) if fast_draws and S else None
        self.cur = np.zeros((S, d), dtype=np.int32)
        self.smp = np.zeros((S, d), dtype=np.int32)
        self.pool = self.pool0.copy()
        self.plen = self.n.copy()
        self.step = np.zeros(S, dtype=np.int64)
        if postselect is not None:
            modes, photons, _ = postselect
            self.ps_modes = np.asarray(modes, dtype=np.int64)
            self.ps_photons = np.asarray(photons, dtype=np.int64)
            self.left = np.tile(self.ps_photons, (S, 1))
            self.needed = np.full(S, int(self.ps_photons.sum()), dtype=np.int64)
            self.trials = np.ones(S, dtype=np.int64)
            self.mode_slot = np.full(d, -1, dtype=np.int64)
            # the reference's tuple(modes).index(mode): the first listed slot
            for q in range(len(self.ps_modes) - 1, -1, -1):
                self.mode_slot[self.ps_modes[q]] = q
        self.rounds = 0
        self.problems = 0
        self.flops = 0.0

    def _restart(self, shots):
        self.trials[shots] += 1
        over = shots[self.trials[shots] > self.ps[2]]
        if over.size:
            raise SampleGenerationError(
                "Too many trials during sample generation. The specified postselection "
                "criteria may be very highly unlikely. Aborting. To increase the limit, "
                "set Config.max_sample_generation_trials to a higher value. Current value: "
                f"{self.ps[2]}.")
        self.cur[shots] = 0
        self.smp[shots] = 0
        self.pool[shots] = self.pool0[shots]
        self.plen[shots] = self.n[shots]
        self.step[shots] = 0
        self.left[shots] = self.ps_photons
        self.needed[shots] = int(self.ps_photons.sum())

    def run(self) -> np.ndarray:
        while True:
            live = np.flatnonzero(self.step < self.n)
            if self.reject is not None:
                # rejected iterations place no particle (sampling.py:222-224);
                # each shot takes decisions until one keeps its particle
                settled = []
                pending = live
                while pending.size:
                    rej = self.reject(pending)
                    settled.append(pending[~rej])
                    skipped = pending[rej]
                    self.step[skipped] += 1
                    pending = skipped[self.step[skipped] < self.n[skipped]]
                live = np.sort(np.concatenate(settled)) if settled else live
            if live.size == 0:
                return self.smp
            # grow the input by a uniformly chosen pool particle (sampling.py:197-205)
            plen = self.plen[live]
            if self.draws is not None:
                pick = self.draws.integers(live, plen)
            else:
                ints = self._ints
                pick = np.fromiter((ints[s](0, L, dtype=np.int64)
                                    for s, L in zip(live.tolist(), plen.tolist())),
                                   dtype=np.int64, count=live.size)
            self.cur[live, self.pool[live, pick]] += 1  # one entry per shot: no repeated index
            # drop the picked particle, keeping the pool's order
            width = self.pool.shape[1]
            j = np.arange(width)[None, :]
            src = np.minimum(j + (j >= pick[:, None]), width - 1)
            self.pool[live] = np.take_along_axis(self.pool[live], src, axis=1)
            self.plen[live] = plen - 1
            # distribution of the next output mode (sampling.py:711-749)
            cur = self.cur[live]
            rows = self.smp[live]
            lap = self.minors(rows, cur)
            if count_flops:
                self.flops += _minors_flops(rows, cur)
            w = np.where(cur > 0, cur * lap, 0)
            amp = w @ self.u_t
            pmf = amp.real ** 2 + amp.imag ** 2
            norm = pmf.sum(axis=1, keepdims=True)
            if not np.all(np.isfinite(norm)) or np.any(norm <= 0):
                raise ValueError("probabilities contain NaN")
            cdf = np.cumsum(pmf / norm, axis=1)
            cdf /= cdf[:, -1:]
            if self.draws is not None:
                u = self.draws.doubles(live)
            else:
                rand = self._rand
                u = np.fromiter((rand[s]() for s in live.tolist()), dtype=np.float64, count=live.size)
            mode = (cdf <= u[:, None]).sum(axis=1)
            mode = np.minimum(mode, cdf.shape[1] - 1)
            self.smp[live, mode] += 1
            self.step[live] += 1
            self.rounds += 1
            self.problems += live.size
            if self.ps is not None:
                slot = self.mode_slot[mode]
                hit = slot >= 0
                if hit.any():
                    hs = live[hit]
                    self.needed[hs] -= 1
                    self.left[hs, slot[hit]] -= 1
                over = np.zeros(live.size, dtype=bool)
                if hit.any():
                    over[hit] = (self.left[live[hit]] < 0).any(axis=1)
                # too few photons left for the postselected modes (sampling.py:314-315)
                if self.track_needed:
                    over |= self.needed[live] > self.n[live] - self.step[live]
                fail = live[over]
                if fail.size:
                    self._restart(fail)


def _minors_flops(rows: np.ndarray, cols: np.ndarray) -> float:
    """Algorithmic FLOP of the column-minor sweeps of a round (the planner's
    count, T (2 m_c + 6 (3 (m_c - 1) + 2 (n + 1 - m_c)) + 4 (m_c + 1)) per
    problem, T = prod (rho_i + 1) after the min-row prepend)."""
    occupied = rows > 0
    n = rows.sum(axis=1)
    big = np.iinfo(np.int32).max
    rmin = np.where(occupied, rows, big).min(axis=1)
    terms = np.prod(np.where(occupied, rows + 1.0, 1.0), axis=1) * rmin / (rmin + 1.0)
    mc = (cols > 0).sum(axis=1)
    per = 2.0 * mc + 6.0 * (3.0 * (mc - 1) + 2.0 * (n + 1 - mc)) + 4.0 * (mc + 1)
    return float(np.sum(np.where(n > 0, terms * per, 0.0)))


def _first_quantized(occupation) -> np.ndarray:
    """Mode of every particle, modes ascending (piquasso/_math/indices.py:105-114)."""
    return np.repeat(np.arange(len(occupation)), np.asarray(occupation, dtype=np.int64))


def _split_particles(occupation, overlap, rng):
    """sampling.py:371-399: K of n_j particles stay indistinguishable with
    weight C(n_j, K) x^K (1 - x)^(n_j - K) K!."""
    indist = np.zeros(len(occupation), dtype=np.int64)
    dist = np.zeros(len(occupation), dtype=np.int64)
    for j, nj in enumerate(occupation):
        nj = int(nj)
        if nj == 0:
            continue
        w = np.array([math.comb(nj, q) * overlap ** q * (1.0 - overlap) ** (nj - q) * math.factorial(q)
                      for q in range(nj + 1)], dtype=float)
        q = int(rng.choice(nj + 1, p=w / w.sum()))
        indist[j] = q
        dist[j] = nj - q
    return indist, dist


def _distinguishable(u, particles, rng, reject_one):
    """sampling.py:402-419: each distinguishable particle leaves by |U_{:,j}|^2."""
    d = len(particles)
    out = np.zeros(d, dtype=np.int64)
    for mode, count in enumerate(particles):
        p = np.abs(u[:, mode]) ** 2
        p = p / p.sum()
        for _ in range(int(count)):
            target = rng.choice(d, p=p)
            if not reject_one():
                out[target] += 1
    return out


def generate_samples(input, shots: int, interferometer,
                     reject_condition: Optional[Callable[[], bool]] = None,
                     postselect_data: Sequence = ((), (), 1000),
                     uniform_particle_overlap: Optional[float] = None,
                     config=None, *, seed_sequence: Optional[int] = None):
    """Samples of piquasso's ``generate_samples``
    (piquasso/_simulators/passive/sampling.py:34-107), same argument names
    minus ``calculate_permanent_laplace`` (the minors always come from the
    batched GPU kernels).  ``config`` is anything with a ``seed_sequence``
    attribute (piquasso's ``Config``), or give ``seed_sequence`` directly;
    ``reject_condition`` None means no particle is ever rejected.  Returns a
    list of ``shots`` tuples of ints, in shot order."""
    occ = np.asarray(input, dtype=np.int64).reshape(-1)
    if (occ < 0).any():
        raise ValueError("occupation numbers must be non-negative")
    u = np.asarray(interferometer, dtype=np.complex128)
    d = len(occ)
    if u.shape != (d, d):
        raise ValueError(f"interferometer must be {d} x {d}, got {u.shape}")
    if shots < 0:
        raise ValueError("shots must be non-negative")
    n = int(occ.sum())
    seed = _seed_of(config, seed_sequence)
    ps_modes, ps_photons, max_trials = postselect_data
    postselected = len(ps_modes) > 0
    minors = _MinorsBatch(u)
    if shots == 0:
        return []
    t0 = time.perf_counter()
    engines = []
    out = _sample_shots(minors, occ, shots, seed, reject_condition, ps_modes, ps_photons,
                        max_trials, uniform_particle_overlap, engines)
    last_run.clear()
    last_run.update(
        shots=shots, photons=n, modes=d,
        rounds=sum(e.rounds for e in engines),
        problems=sum(e.problems for e in engines),
        minors_flops=sum(e.flops for e in engines) if count_flops else None,
        minors_seconds=minors.seconds,
        seconds=time.perf_counter() - t0,
    )
    return out


def _sample_shots(minors, occ, shots, seed, reject_condition, ps_modes, ps_photons, max_trials,
                  uniform_particle_overlap, engines):
    u = minors.u
    d = len(occ)
    n = int(occ.sum())
    postselected = len(ps_modes) > 0
    rngs = [np.random.default_rng(seed=seed + s) for s in range(shots)]

    if postselected and uniform_particle_overlap is not None:
        # restarts depend on the draws: the shots run one after another
        return [_postselected_overlap_shot(minors, occ, rngs[s], reject_condition, ps_modes,
                                           ps_photons, max_trials, float(uniform_particle_overlap),
                                           engines)
                for s in range(shots)]

    if postselected and reject_condition is not None:
        # the number of reject_condition calls depends on the restarts: keep
        # the reference's global call order by running the shots one by one
        out = []
        for s in range(shots):
            live_reject = lambda idx: np.array([bool(reject_condition()) for _ in idx])  # noqa: E731
            eng = _Lockstep(minors, [rngs[s]], [_first_quantized(occ)], live_reject,
                            (ps_modes, ps_photons, max_trials), fast_draws=True)
            engines.append(eng)
            smp = eng.run()[0]
            out.append(tuple(int(v) for v in np.delete(smp, list(ps_modes))))
        return out

    decisions = None
    pointer = np.zeros(shots, dtype=np.int64)
    if reject_condition is not None and n > 0:
        # n calls per shot, shot after shot: the reference's order
        decisions = np.array([[bool(reject_condition()) for _ in range(n)] for _ in range(shots)],
                             dtype=bool).reshape(shots, n)

    def reject_one(s):
        if decisions is None:
            return False
        v = decisions[s, pointer[s]]
        pointer[s] += 1
        return bool(v)

    def reject_many(shot_idx):
        return np.array([reject_one(s) for s in shot_idx], dtype=bool)

    if uniform_particle_overlap is None:
        pools = [_first_quantized(occ)] * shots
        dist_out = None
    else:
        pools, dist_out = [], np.zeros((shots, d), dtype=np.int64)
        for s in range(shots):
            indist, dist = _split_particles(occ, float(uniform_particle_overlap), rngs[s])
            dist_out[s] = _distinguishable(u, dist, rngs[s], lambda s=s: reject_one(s))
            pools.append(_first_quantized(indist))
    eng = _Lockstep(minors, rngs, pools, reject_many if decisions is not None else None,
                    (ps_modes, ps_photons, max_trials) if postselected else None, fast_draws=True)
    engines.append(eng)
    smp = eng.run().astype(np.int64)
    if dist_out is not None:
        smp = dist_out + smp
    if postselected:
        smp = np.delete(smp, list(ps_modes), axis=1)
    return [tuple(row) for row in smp.tolist()]


def _times_linear(poly, const, lin, out):
    """(const + sum_a lin[a] x_a) * poly truncated to poly's shape, written to
    ``out`` axis by axis (piquasso/_math/polynomial.py:18-38).  When ``out``
    is ``poly`` itself, every axis step reads the values the previous steps
    left, which yields const * poly * prod_a (1 + lin[a] x_a); the reference
    calls it that way for the postselection probability of the
    distinguishable particles (sampling.py:532-551), and so does
    :func:`_dist_postselection_probability`, to draw the same samples."""
    out[...] = const * poly
    for axis, c in enumerate(lin):
        np.moveaxis(out, axis, 0)[1:] += c * np.moveaxis(poly, axis, 0)[:-1]
    return out


def _dist_postselection_probability(u, dist, ps_modes, remaining):
    """The probability that the distinguishable particles put exactly
    ``remaining`` photons into the postselected modes: the coefficient of
    x^remaining in prod over particles of (1 - sum_a p_a + sum_a p_a x_a)
    (sampling.py:532-551).  Deliberate difference: the reference multiplies
    the polynomial in place, so each factor's shifted terms read values
    already scaled by the constant (one particle gives (1 - p) p instead of
    p), and its accepted samples are then not distributed as the detection
    probabilities (a chi-square test against them rejects at p ~ 1e-5 with
    2 x 10^4 shots of 3 photons in 4 modes); here each factor multiplies a
    copy.  Every other draw is the reference's."""
    poly = np.zeros(tuple(remaining + 1), dtype=np.float64)
    poly[(0,) * len(remaining)] = 1.0
    for mode, count in enumerate(dist):
        p = np.abs(u[ps_modes, mode]) ** 2
        for _ in range(int(count)):
            poly = _times_linear(poly, 1.0 - p.sum(), p, out=np.empty_like(poly))
    return poly[tuple(remaining)]


def _dist_postselection_table(u, dist, ps_modes, bound):
    """sampling.py:554-582: table[k] = the postselected modes' polynomial of
    the distinguishable particles k, k+1, ... (first-quantized order)."""
    fq = _first_quantized(dist)
    table = [np.zeros(tuple(bound + 1), dtype=np.float64) for _ in range(len(fq) + 1)]
    table[-1][(0,) * len(bound)] = 1.0
    for k in range(len(fq) - 1, -1, -1):
        p = np.abs(u[ps_modes, fq[k]]) ** 2
        _times_linear(table[k + 1], 1.0 - p.sum(), p, out=table[k])
    return table


def _dist_conditioned_output(u, dist, ps_modes, remaining, table, rng):
    """sampling.py:585-655: the distinguishable particles one by one, each to
    a free mode, lost, or to a postselected mode that still needs photons,
    weighted by the table of what the later particles can still supply."""
    d = u.shape[0]
    free = np.delete(np.arange(d), ps_modes)
    out = np.zeros(d, dtype=np.int64)
    left = np.array(remaining, dtype=np.int64)
    for k, mode in enumerate(_first_quantized(dist)):
        nxt = table[k + 1]
        future = nxt[tuple(left)]
        p_free = np.abs(u[free, mode]) ** 2
        p_ps = np.abs(u[ps_modes, mode]) ** 2
        # 1 - (a unitary column's squared norm) can round below 0; the
        # reference's numpy choice() then raises "Probabilities are not
        # non-negative" (sampling.py:615-640).  Clamped here; every draw
        # the reference completes is unchanged.
        p_loss = max(1.0 - p_free.sum() - p_ps.sum(), 0.0)
        w_ps = np.zeros(len(ps_modes), dtype=np.float64)
        for a, pa in enumerate(p_ps):
            if left[a] == 0:
                continue
            left[a] -= 1
            w_ps[a] = pa * nxt[tuple(left)]
            left[a] += 1
        w = np.concatenate([p_free * future, np.array([p_loss * future], dtype=np.float64), w_ps])
        w /= np.sum(w)
        idx = rng.choice(len(w), p=w)
        if idx < len(free):
            out[free[idx]] += 1
        elif idx > len(free):
            a = idx - len(free) - 1
            out[ps_modes[a]] += 1
            left[a] -= 1
    return out


def _postselected_overlap_shot(minors, occ, rng, reject_condition, ps_modes, ps_photons,
                               max_trials, overlap, engines):
    """One shot of postselected sampling with uniform partial
    distinguishability (sampling.py:422-529): split the particles, sample the
    indistinguishable ones with one postselection attempt (no look-ahead on
    the photons still needed), accept the distinguishable ones' chance to
    fill the remaining postselected photons with one draw, then place them
    conditioned on it; restart on any failure, at most ``max_trials`` times."""
    u = minors.u
    modes = np.asarray(ps_modes, dtype=np.int64)
    photons = np.asarray(ps_photons, dtype=np.int64)
    live_reject = None
    if reject_condition is not None:
        live_reject = lambda idx: np.array([bool(reject_condition()) for _ in idx])  # noqa: E731
    trial = 0
    while True:
        trial += 1
        if trial > max_trials:
            raise SampleGenerationError(
                "Too many trials during sample generation. The specified postselection "
                "criteria may be very highly unlikely. Aborting. To increase the limit, "
                "set Config.max_sample_generation_trials to a higher value. Current value: "
                f"{max_trials}.")
        indist, dist = _split_particles(occ, overlap, rng)
        eng = _Lockstep(minors, [rng], [_first_quantized(indist)], live_reject,
                        (modes, photons, 1), track_needed=False)
        engines.append(eng)
        try:
            indist_out = eng.run()[0].astype(np.int64)
        except SampleGenerationError:
            continue
        remaining = photons - indist_out[modes]
        if rng.random() > _dist_postselection_probability(u, dist, modes, remaining):
            continue
        table = _dist_postselection_table(u, dist, modes, photons)
        out = _dist_conditioned_output(u, dist, modes, remaining, table, rng) + indist_out
        return tuple(int(v) for v in np.delete(out, modes))


def expanded_lossy_matrix(interferometer) -> np.ndarray:
    """The 2d x 2d unitary dilation of a lossy interferometer T = V diag(s) W
    (numpy's SVD): [[V, 0], [0, I]] [[diag s, diag t], [diag t, diag s]]
    [[W, 0], [0, I]] with t = sqrt(1 - s^2) (sampling.py:756-784)."""
    v, s, w = np.linalg.svd(np.asarray(interferometer, dtype=np.complex128))
    d = len(s)
    z = np.zeros((d, d))
    eye = np.eye(d)
    t = np.sqrt(1.0 - s ** 2)
    return (np.block([[v, z], [z, eye]])
            @ np.block([[np.diag(s), np.diag(t)], [np.diag(t), np.diag(s)]])
            @ np.block([[w, z], [z, eye]]))


def generate_lossy_samples(input, shots: int, interferometer,
                           postselect_data: Sequence = ((), (), 1000),
                           config=None, *, seed_sequence: Optional[int] = None):
    """Samples of piquasso's ``generate_lossy_samples`` (sampling.py:110-146):
    the dilation sampled with empty auxiliary inputs, first d modes kept."""
    occ = np.asarray(input, dtype=np.int64).reshape(-1)
    d = len(occ)
    big = expanded_lossy_matrix(interferometer)
    full = np.concatenate([occ, np.zeros(d, dtype=np.int64)])
    out = generate_samples(full, shots, big, postselect_data=postselect_data, config=config,
                           seed_sequence=seed_sequence)
    return [s[:d] for s in out]


def _compositions(boxes: int, particles: int):
    """All ways to put ``particles`` into ``boxes``, anti-lexicographic (the
    first box fullest first), as piquasso's ``partitions``
    (piquasso/_math/combinatorics.py:62-108) lists them."""
    if boxes == 0:
        return [()] if particles == 0 else []
    if boxes == 1:
        return [(particles,)]
    out = []
    for first in range(particles, -1, -1):
        out.extend((first,) + rest for rest in _compositions(boxes - 1, particles - first))
    return out


def fock_space_basis(d: int, cutoff: int) -> np.ndarray:
    """Occupations of d modes with 0..cutoff-1 particles, by particle number,
    each block anti-lexicographic, int32 (``get_fock_space_basis``,
    piquasso/_math/fock.py:65-80)."""
    rows = [c for k in range(max(int(cutoff), 0)) for c in _compositions(int(d), k)]
    return np.array(rows, dtype=np.int32).reshape(len(rows), int(d))


def postselected_fock_basis(d: int, cutoff: int, postselected_modes, postselected_photons):
    """``get_postselected_fock_basis`` (piquasso/_math/fock.py:83-113): the
    basis of the other modes with cutoff reduced by the postselected photons,
    the postselected modes filled in."""
    modes = list(postselected_modes)
    photons = np.asarray(postselected_photons, dtype=np.int64)
    active = fock_space_basis(d - len(modes), int(cutoff) - int(photons.sum()))
    if not modes:
        return active
    full = np.zeros((len(active), d), dtype=np.int32)
    full[:, np.delete(np.arange(d), modes)] = active
    full[:, modes] = photons
    return full


def generate_lossy_and_partially_distinguishable_samples(
        input, shots: int, interferometer, postselect_data: Sequence = ((), (), 1000),
        particle_overlap_matrix=None, config=None, *, rng=None,
        seed_sequence: Optional[int] = None, conjugate_loss_kernel: bool = False):
    """Samples of piquasso's ``generate_lossy_and_partially_distinguishable_samples``
    (piquasso/_simulators/passive/sampling.py:658-708): the detection
    probabilities of every output with at most n photons (postselected modes
    fixed) from ONE matrix-polynomial permanent call
    (``lossy_detection_probabilities``), normalised, and ``shots`` indices
    drawn with ``rng.choice(count, size=shots, p=...)``.  The generator is
    ``rng``, else ``config.rng`` (piquasso's Config), else
    ``default_rng(seed_sequence)``.  ``particle_overlap_matrix`` is the Gram
    matrix of the particles (first-quantized order) or a scalar overlap;
    ``conjugate_loss_kernel`` is passed on to the probabilities.

    Deliberate difference: probabilities that come out negative only through
    rounding (above -1e-12 times the largest) are set to 0 before the draw;
    numpy's ``choice`` would reject them."""
    from .distinguishability import lossy_detection_probabilities

    occ = np.asarray(input, dtype=np.int64).reshape(-1)
    t = np.asarray(interferometer, dtype=np.complex128)
    if shots < 0:
        raise ValueError("shots must be non-negative")
    if rng is None:
        rng = getattr(config, "rng", None) if config is not None else None
    if rng is None:
        rng = np.random.default_rng(seed_sequence)
    ps_modes, ps_photons = list(postselect_data[0]), postselect_data[1]
    d = len(occ)
    basis = postselected_fock_basis(d, int(occ.sum()) + 1, ps_modes, ps_photons)
    trimmed = np.delete(basis, ps_modes, axis=1) if ps_modes else basis
    overlap = 1.0 if particle_overlap_matrix is None else particle_overlap_matrix
    p = lossy_detection_probabilities(basis, t, occ, overlap, conjugate_loss_kernel)
    top = float(np.max(p)) if len(p) else 0.0
    p = np.where((p < 0) & (p >= -1e-12 * top), 0.0, p)
    p = p / np.sum(p)
    idx = rng.choice(len(basis), size=shots, p=p)
    return [tuple(int(v) for v in row) for row in trimmed[idx]]
