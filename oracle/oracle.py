"""ORACLE — test infrastructure, not product code.

Python face of the CPU oracle.  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` import this module, and only as
the checker (or the timed CPU baseline); the product package
``piquasso_amd`` never imports it.

Contents
--------
* ctypes bindings of ``liboracle.so`` (built from ``glynn_port.c`` and
  ``glynn_dd.c`` by ``oracle/Makefile``):

  - :func:`perm_port` / :func:`perm_port_c64` — plain-precision restatement of
    ``permanent_cpp<T>`` (src/permanent.cpp:49-264), same Gray order and
    chunking;
  - :func:`perm_dd` — the same sum in double-double: the extended-precision
    arbiter;
  - :func:`grad_dd` — ``grad_perm`` (src/permanent.cpp:271-300) as explicit
    minors in double-double;
  - :func:`laplace_dd` — ``permanent_laplace_cpp`` (src/permanent_laplace.cpp:43-240)
    as separate double-double sums per column;
  - :func:`gray_states` — the reflected n-ary Gray counter
    (src/n_aryGrayCodeCounter.hpp:170-254);
  - :func:`perm_poly_dd` / :func:`lossy_pd_probabilities_dd` — the Ryser
    subset loop of the lossy partially distinguishable detection
    probabilities (piquasso/_simulators/passive/probabilities.py:244-426) in
    double-double, pinned by :func:`perm_poly_bruteforce` (the permutation
    sum with polynomial entries).

* independent pure-Python / numpy formulas used to pin the above:

  - :func:`assym_reduce` — multiplicity expansion
    (piquasso/_math/linalg.py:98-117);
  - :func:`perm_bruteforce` — sum over S_n, exact with ``fractions`` for
    rational inputs;
  - :func:`perm_ryser` — Ryser's inclusion-exclusion formula (Ryser 1963),
    a different algorithm from Glynn's.

* loaders for the reference's own code compiled by ``make ref`` into
  ``oracle/_ref`` (absent unless that build ran): :func:`ref_module` (the
  pybind ``permanent`` module, src: piquasso/_math/permanent.cpp:71-85) and
  :func:`ref_lib` (ctypes shim around ``permanent_cpp`` / ``grad_perm``).
"""

from __future__ import annotations

import ctypes
import importlib.util
import itertools
import math
import os
import sysconfig
from fractions import Fraction
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_lib = None


class OracleUnavailable(RuntimeError):
    pass


def lib():
    """Load ``liboracle.so`` (``make -C oracle`` builds it)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OracleUnavailable(
                f"{LIB_PATH} is missing; run `make -C oracle` first"
            )
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        fp = ctypes.POINTER(ctypes.c_float)
        ip = ctypes.POINTER(ctypes.c_int)
        ci = ctypes.c_int
        for name, argt in (
            ("pqo_perm_port_c128", [dp, ci, ci, ip, ip, ci, dp]),
            ("pqo_perm_port_c64", [fp, ci, ci, ip, ip, ci, fp]),
            ("pqo_perm_dd_c128", [dp, ci, ci, ip, ip, ci, dp]),
            ("pqo_grad_dd_c128", [dp, ci, ci, ip, ip, ci, dp]),
            ("pqo_laplace_dd_c128", [dp, ci, ci, ip, ip, ci, dp]),
            ("pqo_gray_states", [ip, ci, ctypes.c_int64, ctypes.c_int64, ip]),
            ("pqo_tensor_perm_dd", [dp, ci, ctypes.c_double, ci, dp, dp]),
            ("pqo_perm_poly_dd", [dp, dp, ci, ci, ctypes.POINTER(ctypes.c_int32), ci, dp, dp]),
            ("pqo_perm_int_exact", [ctypes.POINTER(ctypes.c_int64), ci, ci, ip, ip, ci,
                                    ctypes.POINTER(ctypes.c_uint64)]),
        ):
            fn = getattr(L, name)
            fn.argtypes = argt
            fn.restype = ci
        L.pqo_num_procs.restype = ci
        L.pqo_num_procs.argtypes = []
        _lib = L
    return _lib


def _prep(matrix, rows, cols, dtype=np.complex128):
    a = np.ascontiguousarray(np.asarray(matrix, dtype=dtype))
    if a.ndim != 2:
        raise ValueError("matrix must be 2-D")
    r = np.ascontiguousarray(np.asarray(rows, dtype=np.int32).reshape(-1))
    c = np.ascontiguousarray(np.asarray(cols, dtype=np.int32).reshape(-1))
    if r.shape[0] != a.shape[0] or c.shape[0] != a.shape[1]:
        raise ValueError("rows/cols lengths must match the matrix shape")
    return a, r, c


def _ptr(x, ct):
    return x.ctypes.data_as(ctypes.POINTER(ct))


def _check(status):
    if status == -1:
        raise RuntimeError("Number of input and output states should be equal")
    if status != 0:
        raise ValueError(f"oracle rejected the arguments (status {status})")


def perm_port(matrix, rows, cols, n_jobs: int = 0) -> complex:
    """Plain FP64 restatement of ``permanent_cpp<double>``."""
    a, r, c = _prep(matrix, rows, cols)
    out = np.zeros(2)
    _check(
        lib().pqo_perm_port_c128(
            _ptr(a.view(np.float64), ctypes.c_double), a.shape[0], a.shape[1],
            _ptr(r, ctypes.c_int), _ptr(c, ctypes.c_int), n_jobs,
            _ptr(out, ctypes.c_double),
        )
    )
    return complex(out[0], out[1])


def perm_port_c64(matrix, rows, cols, n_jobs: int = 0) -> np.complex64:
    """Plain FP32 restatement of ``permanent_cpp<float>``."""
    a, r, c = _prep(matrix, rows, cols, dtype=np.complex64)
    out = np.zeros(2, dtype=np.float32)
    _check(
        lib().pqo_perm_port_c64(
            _ptr(a.view(np.float32), ctypes.c_float), a.shape[0], a.shape[1],
            _ptr(r, ctypes.c_int), _ptr(c, ctypes.c_int), n_jobs,
            _ptr(out, ctypes.c_float),
        )
    )
    return np.complex64(complex(out[0], out[1]))


def perm_dd_parts(matrix, rows, cols, n_jobs: int = 0) -> np.ndarray:
    """Double-double permanent as ``[re_hi, re_lo, im_hi, im_lo]``."""
    a, r, c = _prep(matrix, rows, cols)
    out = np.zeros(4)
    _check(
        lib().pqo_perm_dd_c128(
            _ptr(a.view(np.float64), ctypes.c_double), a.shape[0], a.shape[1],
            _ptr(r, ctypes.c_int), _ptr(c, ctypes.c_int), n_jobs,
            _ptr(out, ctypes.c_double),
        )
    )
    return out


def perm_dd(matrix, rows, cols, n_jobs: int = 0) -> complex:
    """Extended-precision permanent rounded to complex128."""
    p = perm_dd_parts(matrix, rows, cols, n_jobs)
    return complex(p[0] + p[1], p[2] + p[3])


def perm_int_exact(matrix, rows, cols, n_jobs: int = 0) -> int:
    """Exact permanent of an integer matrix as a Python int (multi-word
    integer Glynn sum, oracle/perm_int.c)."""
    a = np.ascontiguousarray(np.asarray(matrix), dtype=np.int64)
    if a.ndim != 2:
        raise ValueError("matrix must be 2-D")
    r = np.ascontiguousarray(np.asarray(rows).reshape(-1), dtype=np.int32)
    c = np.ascontiguousarray(np.asarray(cols).reshape(-1), dtype=np.int32)
    limbs = np.zeros(8, dtype=np.uint64)
    _check(
        lib().pqo_perm_int_exact(
            _ptr(a, ctypes.c_int64), a.shape[0], a.shape[1], _ptr(r, ctypes.c_int),
            _ptr(c, ctypes.c_int), n_jobs, _ptr(limbs, ctypes.c_uint64),
        )
    )
    v = 0
    for k in range(7, -1, -1):
        v = (v << 64) | int(limbs[k])
    return v - (1 << 512) if v >> 511 else v


def grad_dd(matrix, rows, cols, n_jobs: int = 0) -> np.ndarray:
    """d perm / d A_ij = r_i c_j perm(A; r - e_i, c - e_j), double-double."""
    a, r, c = _prep(matrix, rows, cols)
    g = np.zeros(a.shape, dtype=np.complex128)
    _check(
        lib().pqo_grad_dd_c128(
            _ptr(a.view(np.float64), ctypes.c_double), a.shape[0], a.shape[1],
            _ptr(r, ctypes.c_int), _ptr(c, ctypes.c_int), n_jobs,
            _ptr(g.view(np.float64), ctypes.c_double),
        )
    )
    return g


def laplace_dd(matrix, rows, cols, n_jobs: int = 0) -> np.ndarray:
    """Column minors of ``permanent_laplace_cpp`` (src/permanent_laplace.cpp:43-240)
    in double-double: perm(A; r, c - e_l) for c_l > 0, the uncut Glynn sum for
    c_l = 0.  Degenerate inputs (empty matrix, a zero multiplicity sum) give
    the reference's ``[1]`` (src/permanent_laplace.cpp:52-57)."""
    a, r, c = _prep(matrix, rows, cols)
    if a.shape[0] == 0 or a.shape[1] == 0 or r.sum() == 0 or c.sum() == 0:
        return np.ones(1, dtype=np.complex128)
    out = np.zeros(a.shape[1], dtype=np.complex128)
    _check(
        lib().pqo_laplace_dd_c128(
            _ptr(a.view(np.float64), ctypes.c_double), a.shape[0], a.shape[1],
            _ptr(r, ctypes.c_int), _ptr(c, ctypes.c_int), n_jobs,
            _ptr(out.view(np.float64), ctypes.c_double),
        )
    )
    return out


def ref_laplace(matrix, rows, cols) -> np.ndarray:
    """``permanent_laplace_cpp<double>`` of the reference, called directly."""
    a, r, c = _prep(matrix, rows, cols)
    out = np.zeros(max(a.shape[1], 1), dtype=np.complex128)
    size = ref_lib().pqref_laplace_c128(
        _ptr(a.view(np.float64), ctypes.c_double), a.shape[0], a.shape[1],
        _ptr(r, ctypes.c_int), _ptr(c, ctypes.c_int),
        _ptr(out.view(np.float64), ctypes.c_double),
    )
    _check(min(size, 0))  # the shim returns the result length
    return out[:size]


def tensor_perm_dd(matrix, overlap: float, n_jobs: int = 0):
    """Tensor permanent of uniformly partially distinguishable bosons
    (piquasso/_simulators/passive/probabilities.py:193-237) in double-double:
    returns (value as complex, sum over subset pairs of |term|)."""
    a = np.ascontiguousarray(np.asarray(matrix, dtype=np.complex128))
    n = a.shape[0]
    if a.ndim != 2 or a.shape[1] != n:
        raise ValueError("tensor permanent needs a square matrix")
    out = np.zeros(4)
    asum = np.zeros(1)
    _check(lib().pqo_tensor_perm_dd(_ptr(a.view(np.float64), ctypes.c_double), n, float(overlap),
                                    n_jobs, _ptr(out, ctypes.c_double), _ptr(asum, ctypes.c_double)))
    return complex(out[0] + out[1], out[2] + out[3]), float(asum[0])


def tensor_perm_bruteforce(matrix, overlap, exact: bool = False):
    """The tensor permanent by its permutation form: the sum over pairs of
    permutations (sigma, tau) of prod_i A[i, sigma(i)] conj(A[i, tau(i)]),
    each weighted by overlap^(number of rows where sigma(i) != tau(i)) — the
    expansion of the subset formula (a row whose two column choices agree
    gets x + (1 - x) = 1).  With ``exact=True`` the entries and the overlap
    become Fractions and the exact value (re, im) is returned."""
    a = np.asarray(matrix)
    n = a.shape[0]
    if n == 0:
        return (Fraction(1), Fraction(0)) if exact else 1.0 + 0j
    if exact:
        fr = [[(Fraction(float(np.real(a[i, j]))), Fraction(float(np.imag(a[i, j]))))
               for j in range(n)] for i in range(n)]
        x = Fraction(overlap)
    perms = list(itertools.permutations(range(n)))
    tr, ti = (Fraction(0), Fraction(0)) if exact else (0.0, 0.0)
    total = 0j
    for p in perms:
        for q in perms:
            k = sum(1 for i in range(n) if p[i] != q[i])
            if exact:
                pr, pi = x ** k, Fraction(0)
                for i in range(n):
                    ar, ai = fr[i][p[i]]
                    br, bi = fr[i][q[i]]
                    # a * conj(b)
                    cr, ci_ = ar * br + ai * bi, ai * br - ar * bi
                    pr, pi = pr * cr - pi * ci_, pr * ci_ + pi * cr
                tr += pr
                ti += pi
            else:
                prod = overlap ** k + 0j
                for i in range(n):
                    prod *= a[i, p[i]] * np.conj(a[i, q[i]])
                total += prod
    return (tr, ti) if exact else total


def uniform_input_norm(input_occupation, overlap: float) -> float:
    """prod_j sum_k C(n_j, k) x^k (1 - x)^(n_j - k) k!
    (piquasso/_simulators/passive/probabilities.py:387-398)."""
    norm = 1.0
    for nj in input_occupation:
        nj = int(nj)
        norm *= sum(math.comb(nj, k) * overlap ** k * (1.0 - overlap) ** (nj - k) * math.factorial(k)
                    for k in range(nj + 1))
    return norm


def pd_detection_probability(occupation, interferometer, input_occupation, overlap: float) -> float:
    """Detection probability of an output occupation for uniformly partially
    distinguishable input photons (probabilities.py:156-190): the tensor
    permanent of the multiplicity-expanded interferometer over the input
    norm times prod of output factorials; 0 when the photon numbers differ."""
    occ = np.asarray(occupation, dtype=int)
    inp = np.asarray(input_occupation, dtype=int)
    if occ.sum() != inp.sum():
        return 0.0
    a = assym_reduce(np.asarray(interferometer), occ, inp)
    t, _ = tensor_perm_dd(a, overlap)
    den = uniform_input_norm(inp, overlap) * float(np.prod([math.factorial(int(k)) for k in occ]))
    return t.real / den


def perm_poly_dd(b0, bms, exponents):
    """Coefficients [x^e] perm(B_0 + sum_m x_m B_m) for the rows e of
    ``exponents`` (shape (count, len(bms))), by the reference's Ryser subset
    loop with box-truncated polynomials
    (piquasso/_simulators/passive/probabilities.py:341-382) in double-double.
    Returns (coefficients as complex array, sum over subsets of the
    absolute-value polynomial's coefficient)."""
    b0 = np.ascontiguousarray(np.asarray(b0, dtype=np.complex128))
    n = b0.shape[0]
    if b0.shape != (n, n):
        raise ValueError("B_0 must be square")
    bm = np.ascontiguousarray(np.asarray(bms, dtype=np.complex128).reshape(-1, n, n))
    m = bm.shape[0]
    e = np.ascontiguousarray(np.atleast_2d(np.asarray(exponents, dtype=np.int32)))
    if e.shape[1] != m:
        raise ValueError("one exponent per variable")
    out = np.zeros(2 * max(len(e), 1))
    asum = np.zeros(max(len(e), 1))
    _check(lib().pqo_perm_poly_dd(
        _ptr(b0.view(np.float64), ctypes.c_double),
        _ptr(bm.view(np.float64), ctypes.c_double) if m else None, n, m,
        _ptr(e, ctypes.c_int32), len(e), _ptr(out, ctypes.c_double), _ptr(asum, ctypes.c_double)))
    return out[: 2 * len(e)].view(np.complex128).copy(), asum[: len(e)].copy()


def perm_poly_bruteforce(b0, bms, exponent) -> complex:
    """[x^e] perm(B_0 + sum_m x_m B_m) from the definition: the sum over
    permutations sigma of prod_i (B_0 + sum_m x_m B_m)[i, sigma(i)], each
    product expanded as a polynomial (dict from exponent tuples) and the
    coefficient of x^e read off.  Independent of Ryser's formula; n <= 6."""
    b0 = np.asarray(b0, dtype=np.complex128)
    bm = np.asarray(bms, dtype=np.complex128).reshape(-1, *b0.shape)
    n, m = b0.shape[0], bm.shape[0]
    target = tuple(int(v) for v in exponent)
    total = 0j
    for sigma in itertools.permutations(range(n)):
        poly = {(0,) * m: 1.0 + 0j}
        for i in range(n):
            j = sigma[i]
            nxt: dict = {}
            for key, val in poly.items():
                nxt[key] = nxt.get(key, 0j) + val * b0[i, j]
                for v in range(m):
                    k2 = key[:v] + (key[v] + 1,) + key[v + 1:]
                    if k2[v] <= target[v]:
                        nxt[k2] = nxt.get(k2, 0j) + val * bm[v, i, j]
            poly = nxt
        total += poly.get(target, 0j)
    return total


def lossy_pd_matrices(transmission_matrix, input_occupation, particle_overlap,
                      conjugate_loss: bool = False):
    """B_loss and the B_m of the lossy partially distinguishable detection
    probabilities (probabilities.py:304-338): G = x J + (1 - x) I for a
    scalar overlap x or the Gram matrix itself, B_loss = G o (I - T^dag T)
    on the first-quantized input modes, B_m = G o (t_m t_m^dag) with
    t_m = T[m, input_modes].  ``conjugate_loss`` takes the complex conjugate
    of I - T^dag T, the kernel that matches sum_m t_m t_m^dag = (T^dag T)^T."""
    t = np.asarray(transmission_matrix, dtype=np.complex128)
    modes = np.repeat(np.arange(len(input_occupation)), np.asarray(input_occupation, dtype=int))
    n = len(modes)
    if np.isscalar(particle_overlap):
        x = float(particle_overlap)
        g = x * np.ones((n, n), dtype=np.complex128) + (1.0 - x) * np.eye(n)
    else:
        g = np.asarray(particle_overlap, dtype=np.complex128)
    kernel = (np.eye(t.shape[1]) - t.conj().T @ t)[np.ix_(modes, modes)]
    if conjugate_loss:
        kernel = np.conj(kernel)
    b0 = g * kernel
    bms = np.array([g * np.outer(t[m, modes], np.conj(t[m, modes])) for m in range(t.shape[0])])
    return b0, bms, g


def general_input_norm(input_occupation, gram) -> complex:
    """prod over input modes with n_j > 1 of perm(G[block_j, block_j]), block_j
    the particles of mode j (probabilities.py:402-426), in double-double."""
    g = np.asarray(gram, dtype=np.complex128)
    norm = 1.0 + 0j
    start = 0
    for nj in input_occupation:
        nj = int(nj)
        if nj > 1:
            mask = np.zeros(len(g), dtype=int)
            mask[start:start + nj] = 1
            norm *= perm_dd(g, mask, mask)
        start += nj
    return norm


def lossy_pd_probabilities_dd(occupation_numbers, transmission_matrix, input_occupation,
                              particle_overlap, conjugate_loss: bool = False):
    """get_lossy_partially_distinguishable_detection_probabilities
    (probabilities.py:244-384) with the subset sum in double-double: returns
    (probabilities, error scale) per row of ``occupation_numbers``; the
    scale is the absolute-value sum divided by |input norm|."""
    occ = np.asarray(occupation_numbers, dtype=int)
    if occ.ndim == 1:
        occ = occ[None, :]
    inp = np.asarray(input_occupation, dtype=int)
    n = int(inp.sum())
    probs = np.zeros(len(occ))
    scale = np.zeros(len(occ))
    detected = occ.sum(axis=1)
    if n == 0:
        probs[detected == 0] = 1.0
        return probs, scale
    valid = detected <= n
    if not valid.any():
        return probs, scale
    b0, bms, g = lossy_pd_matrices(transmission_matrix, inp, particle_overlap, conjugate_loss)
    if np.isscalar(particle_overlap):
        norm = complex(uniform_input_norm(inp, float(particle_overlap)))
    else:
        norm = general_input_norm(inp, g)
    coef, asum = perm_poly_dd(b0, bms, occ[valid])
    probs[valid] = np.real(coef / norm)
    scale[valid] = asum / abs(norm)
    return probs, scale


def gray_states(limits, offset: int = 0, count: Optional[int] = None) -> np.ndarray:
    """Gray states ``offset .. offset+count-1`` of the reflected mixed-radix
    counter with the given radices (digit 0 least significant)."""
    lim = np.ascontiguousarray(np.asarray(limits, dtype=np.int32))
    total = int(np.prod(lim, dtype=np.int64)) if lim.size else 1
    if count is None:
        count = total - offset
    out = np.zeros((count, lim.size), dtype=np.int32)
    st = lib().pqo_gray_states(
        _ptr(lim, ctypes.c_int), lim.size, offset, count, _ptr(out, ctypes.c_int)
    )
    if st != 0:
        raise ValueError("invalid Gray-state request")
    return out


def num_procs() -> int:
    return int(lib().pqo_num_procs())


# --------------------------------------------------------------------------
# Independent formulas (pure Python / numpy)
# --------------------------------------------------------------------------


def assym_reduce(array, row_reduce_on, col_reduce_on):
    """Repeat row i ``row_reduce_on[i]`` times and column j
    ``col_reduce_on[j]`` times (piquasso/_math/linalg.py:98-117)."""
    ri = np.repeat(np.arange(len(row_reduce_on)), np.asarray(row_reduce_on, int))
    ci = np.repeat(np.arange(len(col_reduce_on)), np.asarray(col_reduce_on, int))
    return np.asarray(array)[np.ix_(ri, ci)]


def perm_bruteforce(matrix, rows=None, cols=None, exact: bool = False):
    """Permanent by the definition, sum over all permutations of S_n.

    With ``exact=True`` every entry is converted to a pair of Fractions (the
    exact value of the double), so the result is the exact permanent of the
    given floating-point matrix, returned as (Fraction re, Fraction im)."""
    a = np.asarray(matrix)
    if rows is not None:
        a = assym_reduce(a, rows, cols)
    n = a.shape[0]
    if a.shape[1] != n:
        raise ValueError("expanded matrix must be square")
    if n == 0:
        return (Fraction(1), Fraction(0)) if exact else 1.0 + 0j
    if exact:
        fr = [[(Fraction(float(np.real(a[i, j]))), Fraction(float(np.imag(a[i, j]))))
               for j in range(n)] for i in range(n)]
        tr, ti = Fraction(0), Fraction(0)
        for p in itertools.permutations(range(n)):
            pr, pi = Fraction(1), Fraction(0)
            for i in range(n):
                xr, xi = fr[i][p[i]]
                pr, pi = pr * xr - pi * xi, pr * xi + pi * xr
            tr += pr
            ti += pi
        return tr, ti
    total = 0j
    for p in itertools.permutations(range(n)):
        prod = 1 + 0j
        for i in range(n):
            prod *= complex(a[i, p[i]])
        total += prod
    return total


def perm_ryser(matrix, rows=None, cols=None) -> complex:
    """Ryser's formula perm(A) = (-1)^n sum_S (-1)^|S| prod_i sum_{j in S} a_ij
    over column subsets S (vectorised in numpy; n <= ~18)."""
    a = np.asarray(matrix, dtype=np.complex128)
    if rows is not None:
        a = assym_reduce(a, rows, cols)
    n = a.shape[0]
    if n == 0:
        return 1.0 + 0j
    masks = np.arange(1, 1 << n, dtype=np.int64)
    bits = ((masks[:, None] >> np.arange(n)[None, :]) & 1).astype(np.float64)
    sizes = bits.sum(axis=1)
    total = 0j
    step = 1 << 14
    for s in range(0, masks.size, step):
        sub = bits[s : s + step]
        rs = sub @ a.T  # (k, n): row sums over chosen columns, per row i
        prod = np.prod(rs, axis=1)
        sign = np.where((n - sizes[s : s + step]) % 2 == 0, 1.0, -1.0)
        total += np.sum(sign * prod)
    return complex(total)


# --------------------------------------------------------------------------
# The reference's own build (oracle/_ref, produced by `make -C oracle ref`)
# --------------------------------------------------------------------------

_ref_mod = None
_ref_lib = None


def ref_available() -> bool:
    return os.path.exists(os.path.join(REF_DIR, "libpqref.so"))


def ref_module():
    """The reference pybind module ``permanent`` built from its sources."""
    global _ref_mod
    if _ref_mod is None:
        path = os.path.join(REF_DIR, "permanent" + sysconfig.get_config_var("EXT_SUFFIX"))
        if not os.path.exists(path):
            raise OracleUnavailable("oracle/_ref has no reference build")
        spec = importlib.util.spec_from_file_location("permanent", path)
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        _ref_mod = mod
    return _ref_mod


def ref_lib():
    """ctypes handle of the shim around the reference C++ functions."""
    global _ref_lib
    if _ref_lib is None:
        path = os.path.join(REF_DIR, "libpqref.so")
        if not os.path.exists(path):
            raise OracleUnavailable("oracle/_ref has no reference build")
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        fp = ctypes.POINTER(ctypes.c_float)
        ip = ctypes.POINTER(ctypes.c_int)
        ci = ctypes.c_int
        L.pqref_perm_c128.argtypes = [dp, ci, ci, ip, ip, dp]
        L.pqref_perm_c64.argtypes = [fp, ci, ci, ip, ip, fp]
        L.pqref_grad_c128.argtypes = [dp, ci, ip, ip, dp]
        L.pqref_laplace_c128.argtypes = [dp, ci, ci, ip, ip, dp]
        for f in (L.pqref_perm_c128, L.pqref_perm_c64, L.pqref_grad_c128,
                  L.pqref_laplace_c128):
            f.restype = ci
        _ref_lib = L
    return _ref_lib


def ref_perm(matrix, rows, cols) -> complex:
    """``permanent_cpp<double>`` of the reference, called directly."""
    a, r, c = _prep(matrix, rows, cols)
    out = np.zeros(2)
    _check(
        ref_lib().pqref_perm_c128(
            _ptr(a.view(np.float64), ctypes.c_double), a.shape[0], a.shape[1],
            _ptr(r, ctypes.c_int), _ptr(c, ctypes.c_int), _ptr(out, ctypes.c_double),
        )
    )
    return complex(out[0], out[1])


def ref_perm_c64(matrix, rows, cols) -> np.complex64:
    """``permanent_cpp<float>`` of the reference, called directly."""
    a, r, c = _prep(matrix, rows, cols, dtype=np.complex64)
    out = np.zeros(2, dtype=np.float32)
    _check(
        ref_lib().pqref_perm_c64(
            _ptr(a.view(np.float32), ctypes.c_float), a.shape[0], a.shape[1],
            _ptr(r, ctypes.c_int), _ptr(c, ctypes.c_int), _ptr(out, ctypes.c_float),
        )
    )
    return np.complex64(complex(out[0], out[1]))


def ref_grad(matrix, rows, cols) -> np.ndarray:
    """``grad_perm`` of the reference (square matrices only)."""
    a, r, c = _prep(matrix, rows, cols)
    if a.shape[0] != a.shape[1]:
        raise ValueError("grad_perm assumes a square matrix")
    g = np.zeros(a.shape, dtype=np.complex128)
    _check(
        ref_lib().pqref_grad_c128(
            _ptr(a.view(np.float64), ctypes.c_double), a.shape[0],
            _ptr(r, ctypes.c_int), _ptr(c, ctypes.c_int),
            _ptr(g.view(np.float64), ctypes.c_double),
        )
    )
    return g
