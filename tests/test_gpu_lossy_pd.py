"""GPU parity of the matrix-polynomial permanent (perm_poly.hip) and of the
lossy, partially distinguishable detection probabilities and sampler built
on it, against the double-double oracle (both the LDS and the global-memory
polynomial paths), the reference tests' closed forms, and invariants at a
size beyond the oracle (normalisation, the indistinguishable and uniform
limits against the GPU's Glynn and tensor-permanent kernels)."""

import math

import numpy as np
import pytest

from test_lossy_pd import (CASES, all_patterns, haar, lossy, random_gram)

pytestmark = pytest.mark.gpu

EPS = np.finfo(float).eps


def _tol(n, scale):
    # FP64 row factors and products, FP64 subset partials summed in dd:
    # a few (n + 2) eps of the absolute-value sum
    return 16 * (n + 2) * EPS * scale + 1e-300


@pytest.mark.parametrize("seed", range(10))
def test_poly_coefficients_match_oracle(gpu, oracle, seed):
    import piquasso_amd as pa

    rng = np.random.default_rng(100 + seed)
    n = 1 + seed % 7
    m = seed % 5
    b0 = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    bm = rng.standard_normal((m, n, n)) + 1j * rng.standard_normal((m, n, n))
    exps = rng.integers(0, 3, size=(12, m)).astype(np.int32)
    if m:
        exps[0] = 0
        exps[1, 0] = n + 1  # total degree above n: exactly 0
    print(f"perm poly n={n} vars={m}", flush=True)
    got, info = pa.perm_poly_coefficients(b0, bm, exps, with_info=True)
    ref, scale = oracle.perm_poly_dd(b0, bm, exps)
    assert info["lds"] == 1
    for e, g, r, s in zip(exps, got, ref, scale):
        if e.sum() > n:
            assert g == 0
        assert abs(g - r) <= _tol(n, s), (e, g, r)


def test_poly_coefficients_global_memory_path(gpu, oracle):
    """6476 monomials (all 0/1 exponents of weight <= 6 over 14 variables)
    exceed the LDS budget: the polynomial lives in global memory."""
    import itertools

    import piquasso_amd as pa

    rng = np.random.default_rng(7)
    n, m = 6, 14
    b0 = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    bm = (rng.standard_normal((m, n, n)) + 1j * rng.standard_normal((m, n, n))) / 2
    exps = np.array([e for e in itertools.product((0, 1), repeat=m) if sum(e) <= n], dtype=np.int32)
    print(f"perm poly global path: {len(exps)} coefficients", flush=True)
    got, info = pa.perm_poly_coefficients(b0, bm, exps, with_info=True)
    assert info["lds"] == 0 and info["monomials"] == len(exps)
    ref, scale = oracle.perm_poly_dd(b0, bm, exps)
    assert np.all(np.abs(got - ref) <= _tol(n, scale))


def test_poly_no_variables_and_repeats(gpu, oracle):
    import piquasso_amd as pa

    rng = np.random.default_rng(3)
    a = rng.standard_normal((5, 5)) + 1j * rng.standard_normal((5, 5))
    got = pa.perm_poly_coefficients(a, np.zeros((0, 5, 5)), np.zeros((3, 0), dtype=np.int32))
    p = complex(pa.permanent(a, np.ones(5, int), np.ones(5, int)))
    assert np.all(np.abs(got - p) <= 1e-13 * abs(p))
    # the same call twice (buffers reused) gives the same bits
    bm = rng.standard_normal((2, 5, 5)) + 0j
    e = [[1, 2], [0, 0], [3, 1]]
    assert np.array_equal(pa.perm_poly_coefficients(a, bm, e), pa.perm_poly_coefficients(a, bm, e))


def test_plan_reuse_across_calls(gpu, oracle):
    """Calls with the same exponent rows reuse the plan and the device
    tables (only the matrices are uploaded); each must still match the
    oracle for its own matrices, also after a call with other rows."""
    import piquasso_amd as pa

    rng = np.random.default_rng(21)
    n, m = 5, 3
    e1 = rng.integers(0, 3, size=(20, m)).astype(np.int32)
    e2 = rng.integers(0, 2, size=(7, m)).astype(np.int32)
    for k, e in enumerate([e1, e1, e2, e1, e1]):
        b0 = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        bm = rng.standard_normal((m, n, n)) + 1j * rng.standard_normal((m, n, n))
        got = pa.perm_poly_coefficients(b0, bm, e)
        ref, scale = oracle.perm_poly_dd(b0, bm, e)
        assert np.all(np.abs(got - ref) <= _tol(n, scale)), k


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_lossy_probabilities_reference_closed_forms(gpu, case):
    import piquasso_amd as pa

    _, t, inp, ov, expected = case
    outs = list(expected)
    got = pa.lossy_detection_probabilities(outs, t, inp, ov)
    np.testing.assert_allclose(got, [expected[o] for o in outs], atol=1e-14)


@pytest.mark.parametrize("overlap", [0.35, "gram"])
@pytest.mark.parametrize("conj", [False, True])
def test_lossy_probabilities_match_oracle(gpu, oracle, overlap, conj):
    import piquasso_amd as pa

    t = lossy(5, 51)
    inp = [1, 0, 2, 1, 0]
    n = sum(inp)
    ov = random_gram(n, 52) if overlap == "gram" else overlap
    outs = all_patterns(5, n)
    print(f"lossy probabilities: {len(outs)} outputs, conj={conj}", flush=True)
    got = pa.lossy_detection_probabilities(outs, t, inp, ov, conjugate_loss_kernel=conj)
    ref, scale = oracle.lossy_pd_probabilities_dd(outs, t, inp, ov, conjugate_loss=conj)
    assert np.all(np.abs(got - ref) <= _tol(n, scale) + 1e-16)
    if conj:
        assert abs(got.sum() - 1.0) < 1e-12


def test_lossy_probabilities_large_table(gpu):
    """8 photons in 8 modes, every output with at most 8 photons (12870
    outputs, global-memory polynomials): normalisation for losses before the
    interferometer; the indistinguishable (G = J) and uniform-overlap limits
    against the Glynn and tensor-permanent kernels on sampled outputs."""
    import piquasso_amd as pa

    d = n = 8
    u = haar(d, 61)
    eta = np.random.default_rng(62).uniform(0.5, 1.0, size=d)
    t = u @ np.diag(np.sqrt(eta))
    inp = [1] * d
    outs = all_patterns(d, n)
    print(f"lossy probabilities large: {len(outs)} outputs", flush=True)
    p = pa.lossy_detection_probabilities(outs, t, inp, random_gram(n, 63))
    assert abs(p.sum() - 1.0) < 1e-11 and p.min() > -1e-14
    full = outs[outs.sum(axis=1) == n][::97]
    pj = pa.lossy_detection_probabilities(full, u, inp, np.ones((n, n)))
    for o, g in zip(full, pj):
        ref = abs(complex(pa.permanent(u, o, inp))) ** 2 / math.prod(math.factorial(int(k)) for k in o)
        assert abs(g - ref) <= 1e-12 * max(ref, 1e-3), o
    px = pa.lossy_detection_probabilities(full, u, inp, 0.5)
    np.testing.assert_allclose(px, pa.detection_probabilities(full, u, inp, 0.5), atol=1e-13)


def test_sampler_matches_oracle_sampler(gpu, oracle):
    """Same generator, same draws: the GPU sampler returns the samples the
    oracle's probabilities give through rng.choice over the postselected
    basis (sampling.py:658-708)."""
    import piquasso_amd as pa

    t = lossy(4, 71)
    inp = [1, 1, 0, 1]
    g = random_gram(3, 72)
    ps = ((1,), (1,), 1000)
    got = pa.generate_lossy_and_partially_distinguishable_samples(
        inp, 400, t, ps, g, rng=np.random.default_rng(5))
    basis = pa.postselected_fock_basis(4, 4, [1], [1])
    p, _ = oracle.lossy_pd_probabilities_dd(basis, t, inp, g)
    idx = np.random.default_rng(5).choice(len(basis), size=400, p=p / p.sum())
    assert got == [tuple(r) for r in np.delete(basis, [1], axis=1)[idx].tolist()]


def test_sampler_reference_support_checks(gpu):
    """The reference's sampling tests for this path
    (tests/_simulators/passive/test_measurements.py:1310-1400)."""
    import piquasso_amd as pa

    g3 = np.array([[1.0, 0.2 + 0.1j, 0.3], [0.2 - 0.1j, 1.0, 0.4 - 0.2j], [0.3, 0.4 + 0.2j, 1.0]])
    # full loss: the vacuum
    s = pa.generate_lossy_and_partially_distinguishable_samples(
        [1, 1, 1], 20, np.zeros((3, 3), dtype=complex), ((), (), 1000), g3, seed_sequence=42)
    assert all(x == (0, 0, 0) for x in s)
    # a permutation with one fully lossy output mode: deterministic
    perm = np.array([[0, 0, 1, 0], [1, 0, 0, 0], [0, 0, 0, 1], [0, 1, 0, 0]], dtype=complex)
    perm[3] = 0.0
    s = pa.generate_lossy_and_partially_distinguishable_samples(
        [1, 1, 1, 0], 20, perm, ((), (), 1000), 0.5, seed_sequence=42)
    assert all(x == (1, 1, 0, 0) for x in s)
    # postselection with a deterministic loss returns the active modes
    t = np.eye(3, dtype=complex)
    t[1, 1] = 0.0
    s = pa.generate_lossy_and_partially_distinguishable_samples(
        [1, 1, 0], 20, t, ((0,), (1,), 1000), 0.5, seed_sequence=42)
    assert all(x == (0, 0) for x in s)
    # beamsplitter, uniform loss, a 2 x 2 Gram matrix: only the six outputs
    bs = np.array([[1, 1], [1, -1]]) / np.sqrt(2) * np.sqrt(0.5)
    s = pa.generate_lossy_and_partially_distinguishable_samples(
        [1, 1], 50, bs, ((), (), 1000), np.array([[1.0, 0.5], [0.5, 1.0]]), seed_sequence=1)
    assert set(s) <= {(0, 0), (1, 0), (0, 1), (2, 0), (1, 1), (0, 2)}


def test_errors(gpu):
    import piquasso_amd as pa

    with pytest.raises(ValueError):
        pa.perm_poly_coefficients(np.eye(25), np.zeros((1, 25, 25)), [[1]])
    with pytest.raises(ValueError):
        pa.perm_poly_coefficients(np.eye(2), np.zeros((1, 2, 2)), [[-1]])
    with pytest.raises(ValueError):
        pa.perm_poly_coefficients(np.eye(2), np.zeros((2, 2, 2)), [[1]])
    with pytest.raises(ValueError):
        pa.lossy_detection_probabilities([[1, 0]], np.eye(2), [1, 1], np.eye(3))
    assert pa.perm_poly_coefficients(np.eye(2), np.zeros((1, 2, 2)), np.zeros((0, 1))).shape == (0,)
