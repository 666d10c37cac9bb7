"""Parity of the column-minor kernels (permanent_laplace, gradients) with the
oracle (GPU).

* ``permanent_laplace`` against outputs of the reference's
  ``permanent_laplace_cpp`` compiled from its sources (tests/golden/laplace_vectors.json)
  and against the double-double oracle ``laplace_dd``;
* gradients against the reference's known-answer gradient vectors
  (tests/jax_extensions/test_permanent_grad.py, transcribed in
  reference_vectors.json), the double-double explicit minors ``grad_dd``,
  central finite differences, and the Laplace expansion identity;
* the batched gradient mode and the torch autograd wrapper.

Tolerance: 1e-12 relative to the largest entry (or to the absolute-value
bound of the sum where cancellation is strong), far inside the 1e-10 bar.
"""

import numpy as np
import pytest

import piquasso_amd as pa
from conftest import complex_matrix, glynn_abs_sum, load_json
from piquasso_amd import workloads as W

pytestmark = pytest.mark.gpu

TOL = 1e-12


EPS = 2.0 ** -53


def _tol(a, r, c, ref):
    """max(TOL |ref|, 4 (n + 2) eps sum|Glynn terms|): the FP64 products
    round each term, so strong cancellation widens the absolute error."""
    n = int(np.sum(r))
    return max(TOL * abs(ref), 4 * (n + 2) * EPS * glynn_abs_sum(a, r, c))


def _laplace_tols(a, r, c, ref_vec):
    c = np.asarray(c)
    out = np.zeros(len(c))
    for l in range(len(c)):
        cl = c.copy()
        if cl[l] > 0:
            cl[l] -= 1
        out[l] = _tol(a, r, cl, ref_vec[l])
    return out


def test_laplace_reference_fixtures(gpu, oracle):
    for case in load_json("laplace_vectors.json")["cases"]:
        a = complex_matrix(case["matrix"]).reshape(case["shape"])
        r, c = case["rows"], case["cols"]
        ref = complex_matrix(case["reference"])
        dd = complex_matrix(case["dd"])
        got = pa.permanent_laplace(a, r, c)
        assert got.dtype == np.complex128 and got.shape == ref.shape, case["id"]
        if got.shape == (1,) and (np.sum(r) == 0 or np.sum(c) == 0):
            assert got[0] == 1 and ref[0] == 1
            continue
        tols = _laplace_tols(a, r, c, dd)
        assert (np.abs(got - dd) <= tols).all(), (case["id"], got, dd, tols)
        # the reference's own FP64 sum carries the same kind of rounding
        assert (np.abs(got - ref) <= 2 * tols + 1e-10 * np.abs(dd)).all(), case["id"]


def test_laplace_widths_and_multiplicities(gpu, oracle):
    """Column counts over all four minor kernel buckets (1-8, 9-16, 17-24,
    25-32), with multiplicities and zero columns."""
    rng = np.random.default_rng(12)
    for m_c, n in ((3, 4), (8, 9), (11, 10), (16, 14), (20, 15), (27, 14), (32, 13)):
        m_r = int(rng.integers(max(2, n // 3), n + 1))
        a = (rng.standard_normal((m_r, m_c)) + 1j * rng.standard_normal((m_r, m_c))) / np.sqrt(2 * m_c)
        r = np.bincount(rng.integers(0, m_r, n), minlength=m_r)
        c = np.bincount(rng.choice(m_c, size=n + 1, replace=n + 1 > m_c), minlength=m_c)
        got = pa.permanent_laplace(a, r, c)
        lap = oracle.laplace_dd(a, r, c)
        for j in range(m_c):
            cj = c.copy()
            if c[j] > 0:
                cj[j] -= 1
                assert lap[j] == oracle.perm_dd(a, r, cj)
        tols = _laplace_tols(a, r, c, lap)
        assert (np.abs(got - lap) <= tols).all(), (m_c, n, got, lap, tols)


def test_laplace_expansion_identity_n18(gpu):
    """sum_l c_l a_xl perm(A; r - e_x, c - e_l) = perm(A; r, c) along every
    row x, with the forward kernel on the right-hand side."""
    rng = np.random.default_rng(5)
    n = 18
    u = W.haar_unitary(n, rng)
    r = np.ones(n, int)
    c = np.ones(n, int)
    full = complex(pa.permanent(u, r, c))
    for x in (0, 7, n - 1):
        rx = r.copy()
        rx[x] = 0
        lap = pa.permanent_laplace(u, rx, c)
        assert abs(np.sum(c * u[x] * lap) - full) <= 1e-11 * max(abs(full), 1e-3)


def test_laplace_edge_cases(gpu):
    one = pa.permanent_laplace(np.ones((2, 2)), [0, 0], [0, 1])
    assert one.shape == (1,) and one[0] == 1
    assert pa.permanent_laplace(np.zeros((0, 0)), [], []).tolist() == [1]
    # complex64 dispatch: the result is complex64
    a = np.array([[1 + 1j, 2], [3, 4 - 1j]], dtype=np.complex64)
    got = pa.permanent_laplace(a, [1, 0], [1, 1])
    assert got.dtype == np.complex64
    assert np.allclose(got, [a[0, 1], a[0, 0]], rtol=1e-6)
    with pytest.raises(RuntimeError):
        pa.permanent_laplace(np.eye(2), [1, 1], [1, 1])


def _wide_problem(rng, m_c, zero_column):
    """Two rows of multiplicity ~n/2 over m_c columns (few Gray terms), all
    columns of multiplicity 1, or one of 0 and one of 2."""
    n = m_c - 1
    a = (rng.standard_normal((2, m_c)) + 1j * rng.standard_normal((2, m_c))) / np.sqrt(m_c)
    r = np.array([n // 2, n - n // 2])
    c = np.ones(m_c, int)
    if zero_column:
        c[3] = 0
        c[m_c - 2] = 2
    return a, r, c


def test_laplace_beyond_32_columns(gpu, oracle):
    """33..64 columns of non-zero multiplicity: more than the column-minor
    kernels hold, so the minors run as forward items (one per column, the
    uncut product for a zero column), in single calls and in a batch."""
    rng = np.random.default_rng(13)
    probs = [_wide_problem(rng, m_c, z) for m_c, z in ((33, False), (40, True), (48, False), (64, True))]
    batch = pa.permanent_laplace_batch(*zip(*probs))
    for (a, r, c), got_b in zip(probs, batch):
        print(f"laplace width {len(c)}", flush=True)
        got = pa.permanent_laplace(a, r, c)
        lap = oracle.laplace_dd(a, r, c)
        tols = _laplace_tols(a, r, c, lap)
        assert (np.abs(got - lap) <= tols).all(), (len(c), np.abs(got - lap).max())
        assert (np.abs(got_b - lap) <= tols).all(), len(c)
    # a narrow problem in the same batch goes the same way
    a2 = (rng.standard_normal((3, 5)) + 1j * rng.standard_normal((3, 5))) / 3
    r2, c2 = np.array([1, 2, 1]), np.array([1, 1, 1, 1, 1])
    mixed = pa.permanent_laplace_batch([probs[1][0], a2], [probs[1][1], r2], [probs[1][2], c2])
    lap2 = oracle.laplace_dd(a2, r2, c2)
    assert (np.abs(mixed[1] - lap2) <= _laplace_tols(a2, r2, c2, lap2)).all()


def test_gradient_every_width(gpu, oracle):
    """n = m_c = 1..64 over two rows of multiplicity ~n/2 (few Gray terms):
    every single-sweep gradient bucket (1-8, 9-16, 17-19, 20, 21-32) and,
    beyond 32 particles, the forward-item form, against explicit minors."""
    rng = np.random.default_rng(15)
    mats, rows, cols = [], [], []
    for n in list(range(1, 33)) + [33, 40, 48, 64]:
        m_r = 1 if n == 1 else 2
        a = (rng.standard_normal((m_r, n)) + 1j * rng.standard_normal((m_r, n))) / np.sqrt(n)
        mats.append(a)
        rows.append(np.array([n]) if m_r == 1 else np.array([n // 2, n - n // 2]))
        cols.append(np.ones(n, int))
    # one batch per form: a batch holding any permanent above 32 particles
    # runs every item in the column-minor / forward-item form
    grads = pa.permanent_grad_batch(mats[:32], rows[:32], cols[:32])[1]
    grads += pa.permanent_grad_batch(mats[32:], rows[32:], cols[32:])[1]
    for a, r, c, g in zip(mats, rows, cols, grads):
        n = int(np.sum(r))
        ref = oracle.grad_dd(a, r, c)
        for i in np.flatnonzero(r):
            ri = r.copy()
            ri[i] -= 1
            for j in range(n):
                cj = c.copy()
                cj[j] -= 1
                tol = r[i] * c[j] * _tol(a, ri, cj, ref[i, j] / (r[i] * c[j]))
                assert abs(g[i, j] - ref[i, j]) <= tol, (n, i, j, g[i, j], ref[i, j])


def test_gradient_beyond_32_particles_and_columns(gpu, oracle):
    """n = 40 over 40 distinct columns: neither the single-sweep gradient
    kernels (n <= 32) nor the column-minor kernels (<= 32 columns) hold it,
    so G_ij = r_i c_j perm(A; r - e_i, c - e_j) runs as forward items; the
    direct entry and a gradient batch agree with the explicit minors."""
    rng = np.random.default_rng(14)
    a, r, c = _wide_problem(rng, 41, False)
    c = c[:40]
    a = a[:, :40]
    r = np.array([20, 20])
    g = pa.permanent_grad(a, r, c)
    ref = oracle.grad_dd(a, r, c)
    with pa.PermanentBatch([a], [r], [c], gradient=True) as b:
        b.run("eager")
        gb = b.gradients()[0]
        pb = b.results()[0]
    p = complex(pa.permanent(a, r, c))
    assert abs(pb - p) <= _tol(a, r, c, p)
    for i in range(2):
        ri = r.copy()
        ri[i] -= 1
        for j in range(40):
            cj = c.copy()
            cj[j] -= 1
            tol = r[i] * c[j] * _tol(a, ri, cj, ref[i, j] / (r[i] * c[j]))
            assert abs(g[i, j] - ref[i, j]) <= tol, (i, j, g[i, j], ref[i, j])
            assert abs(gb[i, j] - ref[i, j]) <= tol, (i, j)


def test_gradient_reference_vectors(gpu, reference_vectors):
    for case in reference_vectors["gradient"]:
        a = complex_matrix(case["matrix"])
        n = len(case["rows"])
        a = a.reshape(n, -1)
        g = pa.permanent_grad(a, case["rows"], case["cols"])
        exp = complex_matrix(case["expected"]).reshape(a.shape)
        if case.get("exact"):
            assert np.array_equal(g, exp), case["id"]
        else:
            assert np.allclose(g, exp, rtol=case["rtol"], atol=1e-8), case["id"]


def test_gradient_matches_explicit_minors(gpu, oracle):
    """Square and rectangular matrices, multiplicities, zero rows/columns."""
    rng = np.random.default_rng(21)
    for m_r, m_c, n in ((4, 4, 4), (6, 6, 6), (5, 7, 8), (9, 3, 9), (12, 12, 12), (8, 14, 14)):
        a = (rng.standard_normal((m_r, m_c)) + 1j * rng.standard_normal((m_r, m_c))) / np.sqrt(2 * n)
        r = np.bincount(rng.integers(0, m_r, n), minlength=m_r)
        c = np.bincount(rng.integers(0, m_c, n), minlength=m_c)
        g = pa.permanent_grad(a, r, c)
        ref = oracle.grad_dd(a, r, c)
        for i in np.flatnonzero(r):
            ri = r.copy()
            ri[i] -= 1
            for j in np.flatnonzero(c):
                cj = c.copy()
                cj[j] -= 1
                tol = r[i] * c[j] * _tol(a, ri, cj, ref[i, j] / (r[i] * c[j]))
                assert abs(g[i, j] - ref[i, j]) <= tol, (m_r, m_c, n, i, j, g[i, j], ref[i, j])
        assert (g[r == 0] == 0).all() and (g[:, c == 0] == 0).all()


def test_gradient_batch_of_uniform_items(gpu, oracle):
    """A gradient batch (n = 12..20, Haar submatrices): every item against
    explicit minors for the small ones, all against Euler's identity
    sum_ij A_ij G_ij = n perm(A) (the permanent is homogeneous of degree n)."""
    mats, rows, cols = W.gradient_batch(count=48, seed=5)
    perms, grads = pa.permanent_grad_batch(mats, rows, cols)
    for k, (a, r, c) in enumerate(zip(mats, rows, cols)):
        n = int(np.sum(r))
        p = complex(pa.permanent(a, r, c))
        assert abs(perms[k] - p) <= 1e-10 * abs(p)
        euler = np.sum(a * grads[k])
        assert abs(euler - n * p) <= 1e-11 * np.sum(np.abs(a * grads[k])) + 1e-10 * n * abs(p), (k, n)
        if n <= 13:
            ref = oracle.grad_dd(a, r, c)
            assert np.abs(grads[k] - ref).max() <= 1e-10 * np.abs(ref).max(), k


def test_gradient_batch_mode_graph_replays(gpu):
    """PermanentBatch(gradient=True): forward + minors in one hipGraph; every
    replay returns the eager results of permanent_grad_batch."""
    mats, rows, cols = W.gradient_batch(count=32, seed=6)
    perms, grads = pa.permanent_grad_batch(mats, rows, cols)
    with pa.PermanentBatch(mats, rows, cols, gradient=True) as b:
        for mode in ("eager", "graph", "graph", "graph"):
            b.run(mode)
            b.sync()
            vals = b.results()
            g = b.gradients()
            for k in range(len(mats)):
                assert abs(vals[k] - perms[k]) <= 1e-12 * abs(perms[k])
                scale = np.abs(grads[k]).max()
                assert np.abs(g[k] - grads[k]).max() <= 1e-12 * scale, (mode, k)
        ms = b.time("graph", reps=3)
        assert (ms > 0).all()


def test_torch_autograd_matches_finite_differences(gpu):
    """L = |perm(A)|^2: A.grad = 2 perm conj(G) (PyTorch convention), checked
    against central differences dL/dRe A_ij + i dL/dIm A_ij, h = 1e-6."""
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(3)
    n = 6
    a = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(2 * n)
    r = np.array([2, 1, 0, 1, 1, 1])
    c = np.array([1, 1, 2, 0, 1, 1])
    A = torch.tensor(a, dtype=torch.complex128, requires_grad=True)
    p = pa.torch_permanent(A, r, c)
    loss = (p * p.conj()).real
    loss.backward()
    got = A.grad.numpy()

    def L(m):
        return abs(complex(pa.permanent(m, r, c))) ** 2

    h = 1e-6
    fd = np.zeros_like(a)
    for i in range(n):
        for j in range(n):
            e = np.zeros_like(a)
            e[i, j] = h
            d_re = (L(a + e) - L(a - e)) / (2 * h)
            d_im = (L(a + 1j * e) - L(a - 1j * e)) / (2 * h)
            fd[i, j] = d_re + 1j * d_im
    assert np.allclose(got, fd, rtol=1e-6, atol=1e-8)
    G = pa.permanent_grad(a, r, c)
    pv = complex(pa.permanent(a, r, c))
    assert np.allclose(got, 2 * pv * np.conj(G), rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("n_want", [12, 16, 20])
def test_gradient_workload_items_against_finite_differences(gpu, n_want):
    """Items of the 256-gradient workload (workloads.gradient_batch(seed=5),
    n ~ U{12..20}, Haar submatrices) at n = 12, 16 and 20: the batched torch
    op's A.grad for L = |perm(A)|^2 against central differences of L in the
    real and imaginary part of every entry (h = 1e-6, forward permanents on
    the GPU), and against 2 perm conj(G) from the batch."""
    torch = pytest.importorskip("torch")
    mats, rows, cols = W.gradient_batch(count=256, seed=5)
    k = next(i for i, r in enumerate(rows) if int(np.sum(r)) == n_want)
    a, r, c = mats[k], rows[k], cols[k]
    print(f"gradient workload item {k}: n = {n_want}", flush=True)
    A = torch.tensor(a[None], dtype=torch.complex128, requires_grad=True)
    p = pa.torch_permanent_batch(A, r[None], c[None])
    loss = (p * p.conj()).real.sum()
    loss.backward()
    got = A.grad[0].numpy()

    def L(m):
        return abs(complex(pa.permanent(m, r, c))) ** 2

    h = 1e-6
    m_r, m_c = a.shape
    fd = np.zeros_like(a)
    for i in range(m_r):
        for j in range(m_c):
            e = np.zeros_like(a)
            e[i, j] = h
            fd[i, j] = (L(a + e) - L(a - e)) / (2 * h) + 1j * (L(a + 1j * e) - L(a - 1j * e)) / (2 * h)
    scale = np.abs(fd).max()
    assert np.abs(got - fd).max() <= 1e-6 * scale, (k, np.abs(got - fd).max() / scale)
    _, grads = pa.permanent_grad_batch([a], [r], [c])
    pv = complex(pa.permanent(a, r, c))
    assert np.allclose(got, 2 * pv * np.conj(grads[0]), rtol=1e-12, atol=1e-14 * scale)


def test_torch_batched_autograd(gpu):
    """L = sum_b w_b |perm(A_b)|^2 over a stack: the batched op's gradient
    equals the single-matrix op's, item by item, and its values the
    permanents; the forward without requires_grad gives the same values."""
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(8)
    B, n = 5, 7
    a = (rng.standard_normal((B, n, n)) + 1j * rng.standard_normal((B, n, n))) / np.sqrt(2 * n)
    r = rng.integers(0, 3, size=(B, n))
    c = np.zeros((B, n), dtype=np.int64)
    for b in range(B):  # columns with the same photon number as the rows
        for _ in range(int(r[b].sum())):
            c[b, rng.integers(0, n)] += 1
    w = torch.tensor(rng.uniform(0.5, 2.0, size=B))
    A = torch.tensor(a, dtype=torch.complex128, requires_grad=True)
    p = pa.torch_permanent_batch(A, r, c)
    assert p.shape == (B,)
    loss = (w * (p * p.conj()).real).sum()
    loss.backward()
    for b in range(B):
        Ab = torch.tensor(a[b], dtype=torch.complex128, requires_grad=True)
        pb = pa.torch_permanent(Ab, r[b], c[b])
        (w[b] * (pb * pb.conj()).real).backward()
        assert abs(complex(pb.detach()) - complex(p[b].detach())) <= 1e-13 * abs(complex(pb.detach()))
        assert np.allclose(A.grad[b].numpy(), Ab.grad.numpy(), rtol=1e-12, atol=1e-15)
    with torch.no_grad():
        q = pa.torch_permanent_batch(torch.tensor(a), r, c)
    assert np.allclose(q.numpy(), p.detach().numpy(), rtol=1e-13, atol=0)


def test_gradient_errors(gpu):
    with pytest.raises(RuntimeError):
        pa.permanent_grad(np.eye(2), [1, 1], [2, 1])
    # an empty selection has a zero gradient
    assert not pa.permanent_grad(np.ones((2, 2)), [0, 0], [0, 0]).any()
