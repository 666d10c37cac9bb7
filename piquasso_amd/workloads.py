"""Seeded synthetic inputs for the permanent workloads the engine is measured on.

Each generator returns plain numpy data (complex128 matrices, int32
multiplicities) and is deterministic for a given seed (numpy PCG64).  Haar
unitaries are drawn as the QR factor of a complex Gaussian matrix with the
phase of R's diagonal moved into Q (Mezzadri, "How to generate random
matrices from the classical compact groups", Notices AMS 54, 2007).

Workloads (sizes from the engine's benchmark definition):
  plain_single      n = 12: top-left 12 x 12 block of a 24-mode Haar unitary
  boson_sampling    40-mode Haar unitary, 20 photons in modes 0..19, an output
                    pattern of 20 photons over 40 modes (multiplicity <= 4)
  large_single      n x n i.i.d. (N + iN) / sqrt(2n), multiplicities 1
  batched_uneven    4096 submatrices of a 48-mode Haar unitary, n ~ U{8..24},
                    row / column multiplicities in {1, 2, 3} summing to n
  gradient_batch    256 Haar submatrices, n ~ U{12..20}, multiplicities 1
  output_patterns   every output occupation of `photons` photons in `modes`
                    modes of a Haar unitary, input one photon in each of the
                    first modes (probability tables)
  tensor            n x n i.i.d. (N + iN) / sqrt(2n) matrix for the tensor
                    permanent of partially distinguishable photons
  sampling          Clifford & Clifford sampling: 40-mode Haar unitary, one
                    photon in each of modes 0..19 (minors_flops counts the
                    column-minor work of one shot's samples)
"""

from __future__ import annotations

import numpy as np

__all__ = [
    "haar_unitary",
    "plain_single",
    "boson_sampling",
    "large_single",
    "batched_uneven",
    "gradient_batch",
    "sampling",
    "minors_flops",
    "tensor",
    "tensor_flops",
    "output_patterns",
    "glynn_terms",
    "glynn_flops",
]


def haar_unitary(m: int, rng: np.random.Generator) -> np.ndarray:
    z = (rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))) / np.sqrt(2.0)
    q, r = np.linalg.qr(z)
    d = np.diag(r)
    return q * (d / np.abs(d))[None, :]


def plain_single(seed: int = 1):
    rng = np.random.default_rng(seed)
    u = haar_unitary(24, rng)
    a = np.ascontiguousarray(u[:12, :12])
    ones = np.ones(12, dtype=np.int32)
    return a, ones, ones.copy()


def boson_sampling(seed: int = 2, modes: int = 40, photons: int = 20, max_mult: int = 4):
    """Returns (U, rows=output pattern, cols=input pattern)."""
    rng = np.random.default_rng(seed)
    u = haar_unitary(modes, rng)
    cols = np.zeros(modes, dtype=np.int32)
    cols[:photons] = 1
    while True:
        draws = rng.integers(0, modes, size=photons)
        rows = np.bincount(draws, minlength=modes).astype(np.int32)
        if rows.max() <= max_mult:
            break
    return u, rows, cols


def large_single(n: int = 32, seed: int = 3):
    rng = np.random.default_rng(seed)
    a = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(2.0 * n)
    ones = np.ones(n, dtype=np.int32)
    return np.ascontiguousarray(a), ones, ones.copy()


def _multiplicity_pattern(n: int, modes: int, rng: np.random.Generator, max_mult: int = 3):
    """Distinct modes with multiplicities in {1..max_mult} summing to n."""
    mults = []
    left = n
    while left > 0:
        m = int(rng.integers(1, min(max_mult, left) + 1))
        mults.append(m)
        left -= m
    sel = rng.choice(modes, size=len(mults), replace=False)
    return np.sort(sel), np.asarray(mults, dtype=np.int32)[np.argsort(sel)]


def batched_uneven(count: int = 4096, seed: int = 4, modes: int = 48, n_lo: int = 8,
                   n_hi: int = 24, max_mult: int = 3, unitary=None):
    """Returns lists (matrices, rows, cols).  `unitary` replaces the drawn
    Haar unitary (same shape) after the draw, so the selections stay the same;
    a stored unitary makes the inputs bit-identical across machines whose
    LAPACK rounds the QR differently."""
    rng = np.random.default_rng(seed)
    u = haar_unitary(modes, rng)
    if unitary is not None:
        u = np.asarray(unitary, dtype=np.complex128)
        if u.shape != (modes, modes):
            raise ValueError("unitary has the wrong shape")
    mats, rows, cols = [], [], []
    for _ in range(count):
        n = int(rng.integers(n_lo, n_hi + 1))
        rsel, rm = _multiplicity_pattern(n, modes, rng, max_mult)
        csel, cm = _multiplicity_pattern(n, modes, rng, max_mult)
        mats.append(np.ascontiguousarray(u[np.ix_(rsel, csel)]))
        rows.append(rm)
        cols.append(cm)
    return mats, rows, cols


def gradient_batch(count: int = 256, seed: int = 5, modes: int = 48, n_lo: int = 12,
                   n_hi: int = 20):
    rng = np.random.default_rng(seed)
    u = haar_unitary(modes, rng)
    mats, rows, cols = [], [], []
    for _ in range(count):
        n = int(rng.integers(n_lo, n_hi + 1))
        rsel = np.sort(rng.choice(modes, size=n, replace=False))
        csel = np.sort(rng.choice(modes, size=n, replace=False))
        mats.append(np.ascontiguousarray(u[np.ix_(rsel, csel)]))
        rows.append(np.ones(n, dtype=np.int32))
        cols.append(np.ones(n, dtype=np.int32))
    return mats, rows, cols


def _prepended_radices(rows):
    """Gray radices rho_i + 1 after the reference's min-row prepend
    (src/permanent.cpp:54-95), zero multiplicities dropped."""
    r = [int(x) for x in np.asarray(rows).reshape(-1)]
    nz = [x for x in r if x > 0]
    if not nz:
        return []
    m = min(nz)
    i0 = r.index(m)
    out = []
    for i, x in enumerate(r):
        rho = x - (1 if i == i0 else 0)
        if rho > 0:
            out.append(rho + 1)
    return out


def glynn_terms(rows) -> int:
    """Number of Gray-code terms T = prod (rho_i + 1)."""
    t = 1
    for rad in _prepended_radices(rows):
        t *= rad
    return t


def glynn_flops(rows, cols) -> float:
    """Algorithmic FLOP of one permanent: T * (2 m_c + 6 (n - 1) + 4), m_c the
    number of columns with non-zero multiplicity."""
    c = np.asarray(cols).reshape(-1)
    n = int(c.sum())
    if n == 0:
        return 0.0
    mc = int(np.count_nonzero(c))
    return float(glynn_terms(rows)) * (2.0 * mc + 6.0 * (n - 1) + 4.0)


def sampling(seed: int = 6, modes: int = 40, photons: int = 20):
    """Returns (U, input occupation) for boson sampling of `photons` single
    photons in the first modes of a `modes`-mode Haar interferometer."""
    rng = np.random.default_rng(seed)
    u = haar_unitary(modes, rng)
    occ = np.zeros(modes, dtype=np.int64)
    occ[:photons] = 1
    return u, occ


def minors_flops(rows, cols) -> float:
    """Algorithmic FLOP of one column-minor sweep perm(A; r, c - e_l) for all
    l, the planner's count: T terms of (row update 2 m_c, prefix / suffix /
    combine products 3 (m_c - 1) complex multiplies plus 2 (n + 1 - m_c) for
    repeated columns, m_c + 1 weighted accumulations of 4), T from the
    prepended row radices; 0 for an empty row set."""
    rows = np.asarray(rows)
    c = np.asarray(cols)
    if rows.sum() == 0:
        return 0.0
    mc = int(np.count_nonzero(c))
    n = int(rows.sum())
    return float(glynn_terms(rows)) * (2.0 * mc + 6.0 * (3.0 * (mc - 1) + 2.0 * (n + 1 - mc))
                                       + 4.0 * (mc + 1))


def tensor(n: int = 18, seed: int = 7) -> np.ndarray:
    """An n x n complex Gaussian matrix (N + iN) / sqrt(2n)."""
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(2.0 * n)


def tensor_flops(n: int) -> float:
    """Algorithmic FLOP of one tensor permanent in the sign-vector form:
    4^(n-1) / 2 pairs (the conjugate half is not visited) of n factors
    x R conj(R') + Q (complex multiply 6, add 1), n - 1 complex products (6),
    the Gray updates of R' (2 per row) and Q (1 per row) and the weighted
    accumulation (2): 16 n - 4 per pair."""
    if n <= 0:
        return 0.0
    return 4.0 ** (n - 1) / 2.0 * (16.0 * n - 4.0)


def output_patterns(modes: int = 14, photons: int = 7, seed: int = 8):
    """(U, input occupation, all output occupations) for a probability table:
    a Haar unitary, one photon in each of the first `photons` modes, and every
    way of placing `photons` photons in `modes` modes (lexicographic)."""
    import itertools

    rng = np.random.default_rng(seed)
    u = haar_unitary(modes, rng)
    inp = np.zeros(modes, dtype=np.int64)
    inp[:photons] = 1
    outs = np.array([np.bincount(c, minlength=modes)
                     for c in itertools.combinations_with_replacement(range(modes), photons)],
                    dtype=np.int64)
    return u, inp, outs


def lossy_pd(modes: int, photons: int, seed: int = 9):
    """(T, input occupation, Gram matrix, outputs) for a lossy, partially
    distinguishable probability table: T = U diag(sqrt(eta)) (a Haar unitary
    after mode losses eta ~ U[0.6, 1]), one photon in each of the first
    `photons` modes, a random Gram matrix of unit vectors, and every output
    with at most `photons` photons (piquasso's Fock basis order)."""
    from .sampling import fock_space_basis

    rng = np.random.default_rng(seed)
    u = haar_unitary(modes, rng)
    eta = rng.uniform(0.6, 1.0, size=modes)
    t = u @ np.diag(np.sqrt(eta))
    v = rng.standard_normal((photons, photons)) + 1j * rng.standard_normal((photons, photons))
    v /= np.linalg.norm(v, axis=1)[:, None]
    inp = np.zeros(modes, dtype=np.int64)
    inp[:photons] = 1
    return t, inp, v @ v.conj().T, fock_space_basis(modes, photons + 1)
