"""Per-result error bounds on the GPU.

The compensated forward sweeps sum the magnitudes of their terms beside the
terms; every result then carries the bound 4 (n + 2) u sum|term| / 2^(n-1)
plus its final rounding (pqp_last_error_bound, PermanentBatch.error_bounds).
Checked here: the kernels' magnitude sums against the same sum computed on
the CPU (conftest.glynn_abs_sum), and every result within its bound of the
double-double oracle - random inputs with multiplicities, every kernel
width, the large single permanents, the 4096-item batched workload and the
FP32-product variant.  The plain baseline and persistent launches keep no
magnitudes and report NaN; results of the exact integer path carry their
final rounding alone.
"""

import ctypes
import math
import os

import numpy as np
import pytest

import piquasso_amd as pa
from conftest import glynn_abs_sum, load_json
from piquasso_amd import _capi
from piquasso_amd import workloads as W

pytestmark = pytest.mark.gpu

EPS = 2.0**-53
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _expected_bound(a, r, c, u=EPS):
    """The bound with the CPU's sum|term| (the kernels sum |coef| (|Re| +
    |Im|) per term, between 1 and sqrt(2) times |term|)."""
    n = int(np.sum(r))
    return 4 * (n + 2) * u * glynn_abs_sum(a, r, c)


def _matches_a_sweep(bound, a, r, c):
    """True when `bound` (less its final rounding) is the magnitude bound of
    the sweep as given or of the transposed sweep perm(A^T; c, r), whichever
    the planner ran."""
    for ar, rr, cc in ((np.asarray(a), r, c), (np.asarray(a).T, c, r)):
        if W.glynn_terms(rr) > 1 << 22:
            continue  # far too many terms for the planner to choose, or for the CPU sum
        lo = _expected_bound(ar, rr, cc)
        if lo * (1 - 1e-9) <= bound <= math.sqrt(2) * lo * (1 + 1e-9):
            return True
    return False


def test_single_bounds_hold_and_match_the_term_magnitudes(gpu, oracle):
    rng = np.random.default_rng(31)
    for trial in range(40):
        m_r, m_c = (int(x) for x in rng.integers(1, 9, size=2))
        n = int(rng.integers(1, 14))
        a = rng.standard_normal((m_r, m_c)) + 1j * rng.standard_normal((m_r, m_c))
        r = np.bincount(rng.integers(0, m_r, n), minlength=m_r)
        c = np.bincount(rng.integers(0, m_c, n), minlength=m_c)
        got, _, bound = pa.permanent_ex(a, r, c, with_bound=True)
        ref = oracle.perm_dd(a, r, c)
        assert abs(got - ref) <= bound, (trial, abs(got - ref), bound)
        assert _matches_a_sweep(bound - EPS * abs(got), a, r, c), (trial, bound)


def test_every_kernel_width_within_its_bound(gpu, oracle):
    """Widths 1..64, two rows of multiplicity ~m_c/2 (sums cancelling by up
    to 1e9): single calls and one batch, each result within its bound."""
    rng = np.random.default_rng(7)
    mats, rows, cols, refs = [], [], [], []
    for mc in range(1, 65):
        a = (rng.standard_normal((2, mc)) + 1j * rng.standard_normal((2, mc))) / np.sqrt(mc)
        mats.append(a)
        rows.append(np.array([mc // 2, mc - mc // 2]))
        cols.append(np.ones(mc, int))
        refs.append(oracle.perm_dd(a, rows[-1], cols[-1]))
    with pa.PermanentBatch(mats, rows, cols) as b:
        b.run("eager")
        vals, bounds = b.results(), b.error_bounds()
    for k in range(64):
        assert abs(vals[k] - refs[k]) <= bounds[k], (k + 1, abs(vals[k] - refs[k]), bounds[k])
        assert _matches_a_sweep(bounds[k] - EPS * abs(vals[k]), mats[k], rows[k], cols[k]), k
    for k in (0, 15, 31, 47, 63):
        # a single call splits the sweep into other units: same terms, sums
        # in another order
        got, _, bound = pa.permanent_ex(mats[k], rows[k], cols[k], with_bound=True)
        assert abs(got - refs[k]) <= bound, k
        assert abs(bound - bounds[k]) <= 1e-12 * bound, (k, bound, bounds[k])


def test_large_permanents_within_their_bounds(gpu):
    """n = 12, 20, 24, 28, 32 and the 40-mode / 20-photon amplitude against
    the committed extended-precision values."""
    for case in load_json("extended_vectors.json")["cases"]:
        m = case["matrix"]
        a = (np.asarray(m["re"]) + 1j * np.asarray(m["im"])).reshape(m["shape"])
        dd = case["dd"]
        ref = complex(dd[0] + dd[1], dd[2] + dd[3])
        got, _, bound = pa.permanent_ex(a, case["rows"], case["cols"], with_bound=True)
        print(f"{case['id']}: |err| / |perm| {abs(got - ref) / abs(ref):.2e}, "
              f"bound / |perm| {bound / abs(ref):.2e}", flush=True)
        assert abs(got - ref) <= bound, case["id"]


@pytest.mark.parametrize("persistent", [False, True])
def test_batched_workload_within_its_bounds(gpu, persistent):
    fx = np.load(os.path.join(ROOT, "tests", "golden", "batched_uneven_seed4.npz"))
    mats, rows, cols = W.batched_uneven(seed=4, unitary=fx["unitary"])
    dd = fx["dd"]
    ref = (dd[:, 0] + dd[:, 1]) + 1j * (dd[:, 2] + dd[:, 3])
    with pa.PermanentBatch(mats, rows, cols, persistent=persistent) as b:
        b.run("eager")
        vals, bounds = b.results(), b.error_bounds()
        if persistent:
            # persistent launches keep no term magnitudes
            assert np.all(np.isnan(bounds))
            return
        err = np.abs(vals - ref)
        rel = bounds / np.abs(ref)
        print(f"bound / |perm|: median {np.median(rel):.2e}, max {rel.max():.2e}; "
              f"smallest bound / error {np.min(bounds / np.maximum(err, 1e-300)):.1f}", flush=True)
        assert np.all(err <= bounds), int(np.argmax(err - bounds))
        b.run("graph")
        assert np.array_equal(b.error_bounds(), bounds)
        b.run("plain")
        assert np.all(np.isnan(b.error_bounds()))


def test_fp32_products_within_their_bounds(gpu, oracle):
    """FP32 row sums and products (u = 2^-24): the n = 20 amplitude and a
    batch of random items."""
    case = {c["id"]: c for c in load_json("extended_vectors.json")["cases"]}["large_single_n20"]
    m = case["matrix"]
    a = (np.asarray(m["re"]) + 1j * np.asarray(m["im"])).reshape(m["shape"])
    dd = case["dd"]
    ref = complex(dd[0] + dd[1], dd[2] + dd[3])
    got, _, bound = pa.permanent_ex(a, case["rows"], case["cols"], mode="fp32", with_bound=True)
    print(f"n=20 FP32 products: |err| / |perm| {abs(got - ref) / abs(ref):.2e}, "
          f"bound / |perm| {bound / abs(ref):.2e}", flush=True)
    assert abs(got - ref) <= bound
    rng = np.random.default_rng(32)
    mats, rows, cols = [], [], []
    for _ in range(24):
        m_r, m_c = (int(x) for x in rng.integers(2, 9, size=2))
        n = int(rng.integers(4, 14))
        mats.append(rng.standard_normal((m_r, m_c)) + 1j * rng.standard_normal((m_r, m_c)))
        rows.append(np.bincount(rng.integers(0, m_r, n), minlength=m_r))
        cols.append(np.bincount(rng.integers(0, m_c, n), minlength=m_c))
    with pa.PermanentBatch(mats, rows, cols, precision="fp32") as b:
        b.run("eager")
        vals, bounds = b.results(), b.error_bounds()
    for k in range(24):
        ref = oracle.perm_dd(mats[k], rows[k], cols[k])
        assert abs(vals[k] - ref) <= bounds[k], k


def test_plain_exact_and_allocation_free_bounds(gpu):
    a = np.exp(2j * np.pi * np.random.default_rng(33).random((10, 10)))
    one = np.ones(10, int)
    _, _, b_plain = pa.permanent_ex(a, one, one, mode="plain", with_bound=True)
    assert math.isnan(b_plain)
    _, _, b_dd = pa.permanent_ex(a, one, one, with_bound=True)
    assert b_dd > 0
    # the allocation-free entry reports the same bound for the same plan
    L = _capi.lib()
    dp = ctypes.POINTER(ctypes.c_double)
    i32p = ctypes.POINTER(ctypes.c_int32)
    a = np.ascontiguousarray(a)
    r = np.ones(10, dtype=np.int32)
    out = np.zeros(2)
    _capi.check(L.pqp_permanent_c128_noalloc(a.view(np.float64).ctypes.data_as(dp), 10, 10, 10,
                                             r.ctypes.data_as(i32p), r.ctypes.data_as(i32p), 0,
                                             out.ctypes.data_as(dp), None))
    assert L.pqp_last_error_bound() == b_dd
    # the exact integer path: the final rounding alone
    n = 20
    one = np.ones(n, int)
    got, _, bound = pa.permanent_ex(np.ones((n, n)), one, one, with_bound=True)
    assert got == float(math.factorial(n))
    assert bound == abs(got) * EPS
    assert abs(got - math.factorial(n)) <= bound
