"""Pin the CPU oracle (oracle/) before trusting it as the checker.

Every oracle function is checked against the reference's own known-answer
vectors (tests/golden/reference_vectors.json, transcribed from the
reference test-suite), against the reference's C++ compiled from its
sources (oracle/_ref, when that build is present), and against what the
mathematics fixes: exact brute force over S_n in rational arithmetic, Ryser's
formula, closed forms and invariants.  The last test plants plausible
mistakes in a CPU restatement and shows that these checks catch each one.
"""

import itertools
import math
from fractions import Fraction

import numpy as np
import pytest

from conftest import complex_matrix, rel_err


def _case_matrix(case):
    if "shape" in case:
        return complex_matrix(case["matrix"]).reshape(case["shape"])
    return complex_matrix(case["matrix"])


def _expected(case):
    e = case["expected"]
    return complex_matrix(e) if isinstance(e["re"], list) else complex(e["re"], e["im"])


# ---------------------------------------------------------------- golden --


def test_reference_permanent_vectors(oracle, reference_vectors):
    for case in reference_vectors["permanent"]:
        a = _case_matrix(case)
        r, c = case["rows"], case["cols"]
        port = oracle.perm_port(a, r, c)
        dd = oracle.perm_dd(a, r, c)
        check = case.get("check")
        if check == "equals_expanded":
            expanded = oracle.perm_bruteforce(a, r, c)
            assert rel_err(dd, expanded) < 1e-13, case["id"]
            assert rel_err(port, expanded) < 1e-13, case["id"]
            continue
        if check == "conjugate_symmetry":
            assert rel_err(oracle.perm_dd(np.conj(a), r, c), np.conj(dd)) < 1e-15
            continue
        exp = _expected(case)
        if case.get("exact"):
            assert dd == exp, case["id"]
            assert port == exp, case["id"]
        else:
            tol = max(case["rtol"], 1e-12)
            scale = max(abs(exp), 1e-8)
            assert abs(dd - exp) <= tol * scale, (case["id"], dd, exp)
            assert abs(port - exp) <= tol * scale, (case["id"], port, exp)


def test_reference_vectors_are_tighter_than_their_tolerance(oracle, reference_vectors):
    """The transcribed expected values carry ~16 digits; the dd oracle meets
    them far below the reference's own rtol=1e-5."""
    for case in reference_vectors["permanent"]:
        if case.get("check") or case.get("exact") or case["id"] == "trivial_real":
            continue
        a = _case_matrix(case)
        dd = oracle.perm_dd(a, case["rows"], case["cols"])
        assert rel_err(dd, _expected(case)) < 1e-8, case["id"]


def test_reference_gradient_vectors(oracle, reference_vectors):
    for case in reference_vectors["gradient"]:
        a = _case_matrix(case)
        g = oracle.grad_dd(a, case["rows"], case["cols"])
        exp = _expected(case)
        if case.get("exact"):
            assert np.array_equal(g, exp), case["id"]
        else:
            assert np.allclose(g, exp, rtol=case["rtol"], atol=1e-8), case["id"]


def test_reference_probabilities(oracle, reference_vectors):
    for case in reference_vectors["probability"]:
        u = _case_matrix(case)
        inp, out = case["input"], case["output"]
        amp = oracle.perm_dd(u, out, inp)
        norm = np.prod([math.factorial(k) for k in out]) * np.prod([math.factorial(k) for k in inp])
        p = abs(amp) ** 2 / norm
        assert abs(p - case["expected"]) < 1e-8, (case["id"], p, case["expected"])


# ------------------------------------------------- the reference's own code --


def _random_case(rng, m_r, m_c, n):
    a = rng.standard_normal((m_r, m_c)) + 1j * rng.standard_normal((m_r, m_c))
    rows = np.bincount(rng.integers(0, m_r, n), minlength=m_r)
    cols = np.bincount(rng.integers(0, m_c, n), minlength=m_c)
    return a, rows, cols


def test_port_is_bit_exact_with_reference_build(oracle):
    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built (make -C oracle ref needs the reference tree)")
    rng = np.random.default_rng(7)
    jobs = 4 * oracle.num_procs()  # the reference uses 4 * hardware_concurrency chunks
    for trial in range(60):
        m_r, m_c = rng.integers(1, 8, size=2)
        n = int(rng.integers(0, 11))
        a, r, c = _random_case(rng, m_r, m_c, n)
        ref = oracle.ref_perm(a, r, c)
        port = oracle.perm_port(a, r, c, n_jobs=jobs)
        assert port == ref, (trial, port, ref)


def test_port_c64_is_bit_exact_with_reference_build(oracle):
    """The FP32 restatement (permanent_cpp<float>, src/permanent.cpp:266-267)
    reproduces the reference's complex64 results bit for bit, directly and
    through the pybind module's float overload."""
    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built")
    rng = np.random.default_rng(17)
    jobs = 4 * oracle.num_procs()
    m = oracle.ref_module()
    for trial in range(40):
        m_r, m_c = rng.integers(1, 7, size=2)
        n = int(rng.integers(0, 10))
        a, r, c = _random_case(rng, m_r, m_c, n)
        a64 = a.astype(np.complex64)
        port = oracle.perm_port_c64(a64, r, c, n_jobs=jobs)
        assert port == oracle.ref_perm_c64(a64, r, c), trial
        assert port == complex(m.permanent(a64, r, c)), trial


def test_reference_build_sum_mismatch_and_pybind_entry(oracle):
    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built")
    with pytest.raises(RuntimeError):
        oracle.ref_perm(np.eye(2), [1, 1], [1, 0])
    with pytest.raises(RuntimeError):
        oracle.perm_port(np.eye(2), [1, 1], [1, 0])
    with pytest.raises(RuntimeError):
        oracle.perm_dd(np.eye(2), [1, 1], [1, 0])
    m = oracle.ref_module()
    r = m.permanent(np.eye(3, dtype=complex), np.ones(3, int), np.ones(3, int))
    assert isinstance(r, np.ndarray) and r.shape == () and r.dtype == np.complex128


def test_grad_dd_matches_reference_grad_perm(oracle):
    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built")
    rng = np.random.default_rng(11)
    for n in (2, 3, 4, 5):
        a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        r = rng.integers(0, 3, n)
        c = rng.permutation(r)
        g_ref = oracle.ref_grad(a, r, c)
        g = oracle.grad_dd(a, r, c)
        assert np.allclose(g, g_ref, rtol=1e-12, atol=1e-12)


# ---------------------------------------------------------- mathematics --


def test_dd_matches_exact_rational_bruteforce(oracle):
    rng = np.random.default_rng(3)
    for trial in range(25):
        m_r, m_c = rng.integers(1, 5, size=2)
        n = int(rng.integers(1, 7))
        a, r, c = _random_case(rng, m_r, m_c, n)
        er, ei = oracle.perm_bruteforce(a, r, c, exact=True)
        parts = oracle.perm_dd_parts(a, r, c)
        dr = Fraction(parts[0]) + Fraction(parts[1])
        di = Fraction(parts[2]) + Fraction(parts[3])
        mag = math.sqrt(float(er) ** 2 + float(ei) ** 2)
        err = math.sqrt(float(dr - er) ** 2 + float(di - ei) ** 2)
        assert err <= 1e-26 * max(mag, 1.0), (trial, err, mag)


def test_dd_matches_ryser(oracle):
    rng = np.random.default_rng(5)
    for n in range(1, 13):
        a = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
        ones = np.ones(n, int)
        assert rel_err(oracle.perm_dd(a, ones, ones), oracle.perm_ryser(a)) < 1e-11, n


def test_closed_forms(oracle):
    for n in range(1, 11):
        ones = np.ones(n, int)
        J = np.ones((n, n), dtype=complex)
        assert oracle.perm_dd(J, ones, ones) == math.factorial(n)
        assert oracle.perm_port(J, ones, ones) == math.factorial(n)
        # derangements: perm(J - I) = !n
        der = round(math.factorial(n) * sum((-1) ** k / math.factorial(k) for k in range(n + 1)))
        assert oracle.perm_dd(J - np.eye(n), ones, ones) == der
        # scaled: perm(c J) = n! c^n
        cst = 0.3 - 0.7j
        assert rel_err(oracle.perm_dd(cst * J, ones, ones), math.factorial(n) * cst ** n) < 1e-14
    # multiplicity form of J: one row repeated n times, one column repeated n times
    for n in range(1, 9):
        assert oracle.perm_dd(np.ones((1, 1)), [n], [n]) == math.factorial(n)
    rng = np.random.default_rng(9)
    d = rng.standard_normal(7) + 1j * rng.standard_normal(7)
    ones = np.ones(7, int)
    assert rel_err(oracle.perm_dd(np.diag(d), ones, ones), np.prod(d)) < 1e-15
    u = rng.standard_normal(6) + 1j * rng.standard_normal(6)
    v = rng.standard_normal(6) + 1j * rng.standard_normal(6)
    ones = np.ones(6, int)
    assert rel_err(oracle.perm_dd(np.outer(u, v), ones, ones),
                   math.factorial(6) * np.prod(u) * np.prod(v)) < 1e-13
    # block diagonal
    b1 = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
    b2 = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
    blk = np.zeros((7, 7), dtype=complex)
    blk[:3, :3] = b1
    blk[3:, 3:] = b2
    one7 = np.ones(7, int)
    assert rel_err(oracle.perm_dd(blk, one7, one7),
                   oracle.perm_dd(b1, [1] * 3, [1] * 3) * oracle.perm_dd(b2, [1] * 4, [1] * 4)) < 1e-14


def test_invariants(oracle):
    rng = np.random.default_rng(12)
    for trial in range(10):
        m_r, m_c = rng.integers(2, 6, size=2)
        n = int(rng.integers(2, 8))
        a, r, c = _random_case(rng, m_r, m_c, n)
        p = oracle.perm_dd(a, r, c)
        # transpose swaps the roles of rows and columns
        assert rel_err(oracle.perm_dd(a.T, c, r), p) < 1e-13
        # conjugation
        assert rel_err(oracle.perm_dd(np.conj(a), r, c), np.conj(p)) < 1e-13
        # row / column permutation (multiplicities follow their rows / columns)
        pr, pc = rng.permutation(m_r), rng.permutation(m_c)
        assert rel_err(oracle.perm_dd(a[pr][:, pc], r[pr], c[pc]), p) < 1e-13
        # diagonal scaling: perm(D1 A D2) = prod d1^r prod d2^c perm(A)
        d1 = rng.standard_normal(m_r) + 1j * rng.standard_normal(m_r)
        d2 = rng.standard_normal(m_c) + 1j * rng.standard_normal(m_c)
        scaled = oracle.perm_dd(d1[:, None] * a * d2[None, :], r, c)
        assert rel_err(scaled, np.prod(d1 ** r) * np.prod(d2 ** c) * p) < 1e-12
        # multiplicity form equals the expanded matrix with unit multiplicities
        ex = oracle.assym_reduce(a, r, c)
        ones = np.ones(ex.shape[0], int)
        assert rel_err(oracle.perm_dd(ex, ones, ones), p) < 1e-13


def test_unitary_output_distribution_sums_to_one(oracle):
    """sum over output patterns of |perm(U; out, in)|^2 / (prod out! prod in!) = 1."""
    rng = np.random.default_rng(4)
    m, n = 4, 3
    z = rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))
    q, rr = np.linalg.qr(z)
    u = q * (np.diag(rr) / np.abs(np.diag(rr)))[None, :]
    inp = np.array([1, 1, 1, 0])
    total = 0.0
    for out in itertools.product(range(n + 1), repeat=m):
        if sum(out) != n:
            continue
        amp = oracle.perm_dd(u, out, inp)
        total += abs(amp) ** 2 / np.prod([math.factorial(k) for k in out])
    assert abs(total - 1.0) < 1e-13


# ------------------------------------------------------------ Gray code --


def _reflected_gray_recursive(radices):
    """Independent construction: the reflected mixed-radix Gray code of the
    digits (digit 0 least significant) lists, for each value of the top
    digit, the code of the lower digits forward or reversed alternately."""
    if not radices:
        return [()]
    lower = _reflected_gray_recursive(radices[:-1])
    out = []
    for v in range(radices[-1]):
        seq = lower if v % 2 == 0 else lower[::-1]
        out.extend(s + (v,) for s in seq)
    return out


def test_gray_counter_matches_recursive_definition(oracle):
    for radices in ([2, 2, 2], [3, 2], [2, 3, 4], [4, 1, 3], [5], [3, 3, 2, 2]):
        seq = oracle.gray_states(radices)
        ref = np.array(_reflected_gray_recursive(radices))
        assert np.array_equal(seq, ref), radices
        # every state once, one digit moves by +-1 per step
        assert len({tuple(s) for s in seq}) == int(np.prod(radices))
        steps = np.abs(np.diff(seq, axis=0))
        assert np.all(steps.sum(axis=1) == 1)
        # initialising at any offset gives the same state as stepping there
        for off in range(0, len(seq), max(1, len(seq) // 5)):
            assert np.array_equal(oracle.gray_states(radices, off, 1)[0], seq[off])


# ------------------------------------------------- the checks have teeth --


def _glynn_python(a, rows, cols, mistake=None):
    """Small pure-Python Glynn with multiplicities (the formula of
    src/permanent.cpp:41-50), with an optional planted mistake."""
    a = np.asarray(a, dtype=complex)
    rows = list(rows)
    cols = list(cols)
    n = sum(rows)
    if n == 0:
        return 1.0 + 0j
    nz = [x for x in rows if x > 0]
    i0 = rows.index(min(nz))
    if mistake == "no_prepend":
        b0 = np.zeros(a.shape[1], dtype=complex)
        rho = rows[:]
    else:
        b0 = a[i0].copy()
        rho = rows[:]
        rho[i0] -= 1
    total = 0j
    for ks in itertools.product(*[range(x + 1) for x in rho]):
        w = 1
        for x, k in zip(rho, ks):
            w *= math.comb(x, k)
        sign = -1 if sum(ks) % 2 else 1
        if mistake == "sign":
            sign = 1
        cs = b0.copy()
        for i, (x, k) in enumerate(zip(rho, ks)):
            coef = x - 2 * k if mistake != "coef" else x - k
            cs = cs + coef * a[i]
        prod = 1 + 0j
        for j, c in enumerate(cols):
            prod *= cs[j] ** (c if mistake != "col_power" else min(c, 1))
        total += sign * w * prod
    scale = 2.0 ** (n - 1) if mistake != "scale" else 2.0 ** n
    return total / scale


def test_planted_mistakes_are_caught(oracle):
    rng = np.random.default_rng(21)
    a, r, c = _random_case(rng, 4, 4, 6)
    r = np.array([2, 1, 0, 3])
    c = np.array([1, 2, 2, 1])
    truth = oracle.perm_bruteforce(a, r, c)
    assert rel_err(_glynn_python(a, r, c), truth) < 1e-12
    assert rel_err(_glynn_python(a.T, c, r), truth) < 1e-12
    for mistake in ("no_prepend", "sign", "coef", "col_power", "scale"):
        assert rel_err(_glynn_python(a, r, c, mistake), truth) > 1e-3, mistake
    # a transposed operand is caught by the non-square known answers
    asym = np.array([[0.909394 + 0.264435j, 0.00450261 + 0.0188079j, 0.316704 + 0.0490014j],
                     [-0.109435 - 0.117915j, 0.35198 + 0.812257j, 0.312276 + 0.304881j]])
    good = oracle.perm_dd(asym, [2, 1], [1, 1, 1])
    assert rel_err(good, oracle.perm_bruteforce(asym, [2, 1], [1, 1, 1])) < 1e-13
    with pytest.raises(ValueError):
        oracle.perm_dd(asym.T, [2, 1], [1, 1, 1])


# ------------------------------------------------------- column minors --


def _laplace_case(case):
    a = complex_matrix(case["matrix"]).reshape(case["shape"])
    return a, case["rows"], case["cols"], complex_matrix(case["reference"]), complex_matrix(case["dd"])


def test_laplace_dd_matches_reference_fixtures(oracle):
    """laplace_dd against outputs of the reference's permanent_laplace_cpp
    (compiled from its sources) stored in tests/golden/laplace_vectors.json,
    including zero-multiplicity columns and the degenerate [1] cases."""
    from conftest import load_json

    for case in load_json("laplace_vectors.json")["cases"]:
        a, r, c, ref, dd = _laplace_case(case)
        got = oracle.laplace_dd(a, r, c)
        assert got.shape == ref.shape, case["id"]
        scale = max(np.abs(ref).max(), 1e-300)
        assert np.abs(got - ref).max() <= 1e-10 * scale, case["id"]
        assert np.abs(got - dd).max() <= 1e-15 * scale, case["id"]


def test_laplace_dd_matches_reference_build(oracle):
    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built (make -C oracle ref)")
    rng = np.random.default_rng(8)
    for t in range(60):
        m_r, m_c = (int(x) for x in rng.integers(1, 6, size=2))
        n = int(rng.integers(1, 7))
        a = rng.standard_normal((m_r, m_c)) + 1j * rng.standard_normal((m_r, m_c))
        r = np.bincount(rng.integers(0, m_r, n), minlength=m_r)
        c = np.bincount(rng.integers(0, m_c, n + 1), minlength=m_c)
        x, y = oracle.laplace_dd(a, r, c), oracle.ref_laplace(a, r, c)
        assert np.abs(x - y).max() <= 1e-11 * max(1.0, np.abs(x).max()), t


def test_laplace_dd_exact_minors_and_expansion(oracle):
    """Each column minor equals the exact rational brute-force permanent of
    A with that column's multiplicity lowered by one, and the minors expand
    back to the permanent along any row x: perm(A; r, c) =
    sum_l c_l a_xl perm(A; r - e_x, c - e_l)."""
    rng = np.random.default_rng(9)
    for t in range(12):
        m_r, m_c = (int(x) for x in rng.integers(2, 5, size=2))
        n = int(rng.integers(2, 6))
        a = rng.integers(-3, 4, size=(m_r, m_c)) + 1j * rng.integers(-3, 4, size=(m_r, m_c))
        r = np.bincount(rng.integers(0, m_r, n), minlength=m_r)
        c = np.bincount(rng.integers(0, m_c, n), minlength=m_c)
        full = complex(*(float(v) for v in oracle.perm_bruteforce(a, r, c, exact=True)))
        for x in np.flatnonzero(r):
            rx = r.copy()
            rx[x] -= 1
            if rx.sum() == 0:
                continue
            lap = oracle.laplace_dd(a, rx, c)
            for l in np.flatnonzero(c):
                cl = c.copy()
                cl[l] -= 1
                er, ei = oracle.perm_bruteforce(a, rx, cl, exact=True)
                # integer entries: every term and partial sum is exact
                assert lap[l] == complex(float(er), float(ei)), (t, x, l)
            assert np.sum(c * a[x] * lap) == full, (t, x)


# ----------------------------------------------------- exact integers --


def _perm_int_bruteforce(a, rows, cols):
    """Exact permanent of the expanded integer matrix by the definition."""
    ri = np.repeat(np.arange(len(rows)), np.asarray(rows, int))
    ci = np.repeat(np.arange(len(cols)), np.asarray(cols, int))
    b = [[int(a[i][j]) for j in ci] for i in ri]
    n = len(b)
    total = 0
    for p in itertools.permutations(range(n)):
        prod = 1
        for i in range(n):
            prod *= b[i][p[i]]
            if not prod:
                break
        total += prod
    return total


def _perm_int_ryser(b):
    """Ryser's formula in Python integers (an independent method):
    perm(B) = (-1)^n sum_S (-1)^|S| prod_i sum_{j in S} b_ij."""
    n = len(b)
    total = 0
    for mask in range(1, 1 << n):
        cols = [j for j in range(n) if mask >> j & 1]
        prod = 1
        for i in range(n):
            prod *= sum(b[i][j] for j in cols)
            if not prod:
                break
        total += (-1) ** len(cols) * prod
    return (-1) ** n * total


def _derangements(n):
    d = [1, 0]
    for k in range(2, n + 1):
        d.append((k - 1) * (d[k - 1] + d[k - 2]))
    return d[n]


def test_perm_int_exact_matches_bruteforce(oracle):
    """The integer oracle equals the definition on random integer matrices
    with multiplicities, including entries of 2^40 whose terms need several
    64-bit words."""
    rng = np.random.default_rng(21)
    for t in range(40):
        m_r, m_c = (int(x) for x in rng.integers(1, 5, size=2))
        n = int(rng.integers(1, 7))
        hi = 2**40 if t % 4 == 0 else 4
        a = rng.integers(-hi, hi + 1, size=(m_r, m_c), dtype=np.int64)
        r = np.bincount(rng.integers(0, m_r, n), minlength=m_r)
        c = np.bincount(rng.integers(0, m_c, n), minlength=m_c)
        assert oracle.perm_int_exact(a, r, c) == _perm_int_bruteforce(a, r, c), t


def test_perm_int_exact_matches_ryser(oracle):
    rng = np.random.default_rng(22)
    for n in (8, 10, 12):
        b = rng.integers(0, 2, size=(n, n), dtype=np.int64)
        b[rng.random((n, n)) < 0.2] = -1
        one = np.ones(n, dtype=int)
        assert oracle.perm_int_exact(b, one, one) == _perm_int_ryser(b.tolist()), n


def test_perm_int_exact_closed_forms(oracle):
    """perm(J_n) = n!, perm(J_n - I) = D_n (derangements), the all-ones
    matrix with multiplicities is n! whatever the multiplicities, and the
    reference's integer case (tests/_math/test_permanent.py:252-282) is 0."""
    for n in (1, 2, 5, 14, 18, 20):
        one = np.ones(n, dtype=int)
        assert oracle.perm_int_exact(np.ones((n, n), np.int64), one, one) == math.factorial(n)
        assert oracle.perm_int_exact(np.ones((n, n), np.int64) - np.eye(n, dtype=np.int64), one,
                                     one) == _derangements(n)
    assert oracle.perm_int_exact(np.ones((3, 4), np.int64), [5, 0, 7], [2, 3, 4, 3]) == \
        math.factorial(12)
    a = np.array([[0, 0, -1, -1, -1, 0, -1, 0, 0, -1], [-1, -1, -1, -1, 0, 0, -1, -1, -1, -1],
                  [-1, 0, -1, -1, -1, 0, 0, -1, 0, -1], [-1, 0, -1, -1, -1, -1, 0, -1, -1, -1],
                  [-1, 0, 0, -1, -1, -1, -1, -1, 0, -1], [0, -1, 0, 0, -1, -1, -1, -1, 0, -1],
                  [-1, 0, 0, 0, -1, -1, -1, -1, 0, -1], [0, 0, -1, -1, -1, -1, -1, -1, 0, 0],
                  [0, 0, 0, -1, -1, -1, 0, 0, -1, -1], [0, 0, 0, 0, -1, 0, -1, -1, -1, 0]],
                 dtype=np.int64)
    rows = [0, 0, 0, 0, 0, 3, 4, 0, 3, 0]
    cols = [1, 0, 0, 0, 0, 0, 1, 7, 0, 1]
    assert oracle.perm_int_exact(a, rows, cols) == 0
    # n = 20 all-ones: beyond FP64's exact integers, where the double-double
    # oracle is only close
    one = np.ones(20, dtype=int)
    dd = oracle.perm_dd(np.ones((20, 20)), one, one)
    assert abs(dd.real - math.factorial(20)) <= 1e-15 * math.factorial(20)


def test_perm_int_exact_errors(oracle):
    with pytest.raises(Exception):
        oracle.perm_int_exact(np.ones((2, 2), np.int64), [1, 1], [2, 1])
    # a term beyond 511 bits is refused, not wrapped
    with pytest.raises(Exception):
        oracle.perm_int_exact(np.full((1, 1), 2**40, np.int64), [13], [13])


def test_perm_int_planted_mistakes_are_caught(oracle):
    """Plausible slips in an integer Glynn sum (a dropped sign, a missing
    final shift by n - 1) fail the closed form perm(J_n) = n!."""
    def glynn(b, mistake=None):
        n = len(b)
        total = 0
        for mask in range(1 << (n - 1)):
            d = [1] + [(-1) ** (mask >> k & 1) for k in range(n - 1)]
            prod = 1
            for j in range(n):
                prod *= sum(d[i] * b[i][j] for i in range(n))
            sign = 1 if mistake == "sign" else (-1) ** bin(mask).count("1")
            total += sign * prod
        return total if mistake == "shift" else total >> (n - 1)

    n = 6
    j = np.ones((n, n), dtype=int).tolist()
    assert glynn(j) == math.factorial(n)
    assert oracle.perm_int_exact(np.array(j, np.int64), [1] * n, [1] * n) == glynn(j)
    for m in ("sign", "shift"):
        assert glynn(j, m) != math.factorial(n), m
