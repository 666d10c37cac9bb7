"""Exact permanents of integer matrices on the GPU (pqp_permanent_int64 and
the routing of pqp_permanent_c128 / _c64 to it).

The multi-modular kernel's integers are compared with the exact integer
oracle (oracle/perm_int.c, pinned by brute force, Ryser and closed forms in
test_oracle.py) and with closed forms beyond the oracle's reach: perm(J_n) =
n!, perm(J_n - I) = D_n, and the all-ones matrix with multiplicities (n!
again).  The reference's integer case (tests/_math/test_permanent.py:252-282)
is 0.  Results through ``permanent`` are the exact permanent rounded once.
"""

import ctypes
import math

import numpy as np
import pytest

import piquasso_amd as pa
from conftest import glynn_abs_sum
from piquasso_amd import _capi

pytestmark = pytest.mark.gpu


def _derangements(n):
    d = [1, 0]
    for k in range(2, n + 1):
        d.append((k - 1) * (d[k - 1] + d[k - 2]))
    return d[n]


def test_all_ones_and_derangements(gpu):
    for n in list(range(1, 25)) + [32]:
        one = np.ones(n, dtype=int)
        J = np.ones((n, n), dtype=np.int64)
        assert pa.permanent_exact(J, one, one) == math.factorial(n), n
        assert pa.permanent_exact(J - np.eye(n, dtype=np.int64), one, one) == _derangements(n), n
    # all-ones with multiplicities: the expanded matrix is J_n
    assert pa.permanent_exact(np.ones((16, 16), np.int64), [2] * 16, [2] * 16) == math.factorial(32)
    assert pa.permanent_exact(np.ones((3, 5), np.int64), [7, 6, 7], [4, 4, 4, 4, 4]) == \
        math.factorial(20)


def test_random_zero_one_matrices_match_the_oracle(gpu, oracle):
    rng = np.random.default_rng(14)
    for n in range(14, 25):
        a = rng.integers(0, 2, size=(n, n), dtype=np.int64)
        one = np.ones(n, dtype=int)
        expect = oracle.perm_int_exact(a, one, one)
        assert pa.permanent_exact(a, one, one) == expect, n
        # the drop-in entry: the exact value rounded once to FP64
        assert complex(pa.permanent(a.astype(np.complex128), one, one)) == float(expect), n


def test_signed_entries_and_multiplicities_match_the_oracle(gpu, oracle):
    rng = np.random.default_rng(15)
    for t in range(12):
        m_r, m_c = (int(x) for x in rng.integers(3, 9, size=2))
        hi = [1, 3, 100, 2**20][t % 4]
        # entries of 2^20 at n <= 14 keep |perm| below the eight moduli
        n = int(rng.integers(10, 15 if hi == 2**20 else 19))
        a = rng.integers(-hi, hi + 1, size=(m_r, m_c), dtype=np.int64)
        r = np.bincount(rng.integers(0, m_r, n), minlength=m_r)
        c = np.bincount(rng.integers(0, m_c, n), minlength=m_c)
        expect = oracle.perm_int_exact(a, r, c)
        assert pa.permanent_exact(a, r, c) == expect, (t, hi)


def test_the_references_integer_case(gpu):
    """tests/_math/test_permanent.py:252-282: the direct and the expanded
    evaluation are both 0, with dtype complex128."""
    from oracle.oracle import assym_reduce

    a = np.array([[0, 0, -1, -1, -1, 0, -1, 0, 0, -1], [-1, -1, -1, -1, 0, 0, -1, -1, -1, -1],
                  [-1, 0, -1, -1, -1, 0, 0, -1, 0, -1], [-1, 0, -1, -1, -1, -1, 0, -1, -1, -1],
                  [-1, 0, 0, -1, -1, -1, -1, -1, 0, -1], [0, -1, 0, 0, -1, -1, -1, -1, 0, -1],
                  [-1, 0, 0, 0, -1, -1, -1, -1, 0, -1], [0, 0, -1, -1, -1, -1, -1, -1, 0, 0],
                  [0, 0, 0, -1, -1, -1, 0, 0, -1, -1], [0, 0, 0, 0, -1, 0, -1, -1, -1, 0]],
                 dtype=np.complex128)
    rows = np.array([0, 0, 0, 0, 0, 3, 4, 0, 3, 0])
    cols = np.array([1, 0, 0, 0, 0, 0, 1, 7, 0, 1])
    ones = np.ones(10, dtype=int)
    direct = pa.permanent(a, rows=rows, cols=cols)
    indirect = pa.permanent(assym_reduce(a, rows, cols), ones, ones)
    assert direct.dtype == np.complex128
    assert complex(direct) == 0 and complex(indirect) == 0
    assert pa.permanent_exact(a, rows, cols) == 0


def test_rounding_of_large_results(gpu):
    """2 x 2 permanents ad + bc next to 2^94: the exact integer comes back in
    the limbs, and the double is rounded once to nearest, ties to even."""
    one = [1, 1]
    big = 2**47
    # 2^94 + 2^41 is a tie (down to even), + 1 more rounds up, 2^94 + 3 2^41
    # is a tie up to even, 2^94 + 2^41 - 1 rounds down
    for bc, expect in ((2**41, 2.0**94), (2**41 + 1, 2.0**94 + 2.0**42),
                       (3 * 2**41, 2.0**94 + 2.0**43), (2**41 - 1, 2.0**94)):
        assert float(2**94 + bc) == expect  # Python's correctly rounded int -> float
        b = bc // 2**20 if bc % 2**20 == 0 else bc
        c = bc // b
        for sign in (1, -1):
            a = np.array([[sign * big, sign * b], [c, big]], dtype=np.int64)
            exact = sign * (big * big + b * c)
            assert pa.permanent_exact(a, one, one) == exact, (bc, sign)
            got = complex(pa.permanent(a.astype(np.complex128), one, one))
            assert got == sign * expect, (bc, sign, got)


def test_complex64_integer_input_is_exact(gpu):
    """The complex64 path rounds the exact permanent once to FP32 (the
    FP32 sweep itself would round every product beyond 2^24)."""
    for n in (10, 14, 18):
        one = np.ones(n, dtype=int)
        got = pa.permanent(np.ones((n, n), dtype=np.complex64), one, one)
        assert got.dtype == np.complex64
        assert got == np.complex64(np.float32(math.factorial(n))), n


def test_out_dd_carries_the_exact_value(gpu):
    """out_dd's two leading doubles add up to the exact permanent (it needs
    at most 106 bits here)."""
    n = 26
    one = np.ones(n, dtype=np.int32)
    a = np.ones((n, n), dtype=np.complex128)
    out = np.zeros(2)
    dd = np.zeros(4)
    dp = ctypes.POINTER(ctypes.c_double)
    i32p = ctypes.POINTER(ctypes.c_int32)
    st = _capi.lib().pqp_permanent_c128(a.view(np.float64).ctypes.data_as(dp), n, n, n,
                                        one.ctypes.data_as(i32p), one.ctypes.data_as(i32p), 0,
                                        out.ctypes.data_as(dp), dd.ctypes.data_as(dp))
    _capi.check(st)
    from fractions import Fraction
    assert Fraction(dd[0]) + Fraction(dd[1]) == math.factorial(n)
    assert out[0] == dd[0] == float(math.factorial(n)) and out[1] == 0


def test_beyond_the_exact_range_falls_back(gpu):
    """Row multiplicity 56 (C(56, 28) > 2^52) is refused by the exact entry
    and evaluated by the floating-point sweep through ``permanent``, within
    the sweep's rounding bound (n + 2) eps sum|terms|: the Glynn terms of
    one row of multiplicity 56 cancel by about 2^76, so that bound is loose
    compared with the result (the reference's FP64 sum rounds the same way)."""
    a = np.ones((1, 1), dtype=np.int64)
    with pytest.raises(ValueError, match="binomial"):
        pa.permanent_exact(a, [56], [56])
    got = complex(pa.permanent(a.astype(np.complex128), [56], [56]))
    bound = 58 * 2.0**-53 * glynn_abs_sum(a, [56], [56])
    assert abs(got - math.factorial(56)) <= bound
