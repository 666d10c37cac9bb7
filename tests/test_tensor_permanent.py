"""Tensor permanent of uniformly partially distinguishable photons: the
double-double oracle (oracle/tensor_dd.c, the subset-pair formula of
piquasso/_simulators/passive/probabilities.py:193-237) against independent
forms — the exact permutation-pair expansion, the x = 1 and x = 0 limits
(|perm A|^2 and perm(|A|^2)), the two-photon HOM dip — and the detection
probabilities built on it against the normalisation of a unitary.  GPU
parity is in test_gpu_tensor.py."""

import itertools
import math

import numpy as np
import pytest


def _rand(n, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(2 * n)


@pytest.mark.parametrize("n", [1, 2, 3, 4])
@pytest.mark.parametrize("x", [0.0, 0.3, 0.75, 1.0])
def test_oracle_matches_exact_permutation_expansion(oracle, n, x):
    a = _rand(n, 10 + n)
    got, _ = oracle.tensor_perm_dd(a, x)
    re, im = oracle.tensor_perm_bruteforce(a, x, exact=True)
    exact = float(re) + 1j * float(im)
    assert abs(got - exact) <= 1e-28 * max(1.0, abs(exact)) + 1e-30
    assert abs(exact.imag) < 1e-30  # the pair sum is real


@pytest.mark.parametrize("n", [2, 5, 7])
def test_oracle_limits(oracle, n):
    a = _rand(n, 30 + n)
    ones = np.ones(n, dtype=int)
    t1, _ = oracle.tensor_perm_dd(a, 1.0)
    t0, _ = oracle.tensor_perm_dd(a, 0.0)
    p = oracle.perm_dd(a, ones, ones)
    pa2 = oracle.perm_dd(np.abs(a) ** 2, ones, ones)
    assert abs(t1 - abs(p) ** 2) <= 1e-14 * abs(p) ** 2
    assert abs(t0 - pa2) <= 1e-14 * abs(pa2)


def test_oracle_empty_and_hom_dip(oracle):
    assert oracle.tensor_perm_dd(np.zeros((0, 0)), 0.5)[0] == 1.0
    # Hong-Ou-Mandel: two photons of overlap x on a 50:50 beam splitter
    # coincide with probability (1 - x^2) / 2 (the swap pair differs in both
    # rows, weight x^2)
    bs = np.array([[1, 1], [1, -1]]) / np.sqrt(2)
    for x in (0.0, 0.25, 0.5, 1.0):
        p = oracle.pd_detection_probability([1, 1], bs, [1, 1], x)
        assert abs(p - (1 - x * x) / 2) < 1e-15


def _patterns(d, n):
    for c in itertools.combinations_with_replacement(range(d), n):
        yield np.bincount(c, minlength=d)


@pytest.mark.parametrize("x", [0.0, 0.4, 1.0])
def test_oracle_probabilities_sum_to_one(oracle, x):
    rng = np.random.default_rng(3)
    z = (rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))) / np.sqrt(2)
    u, r = np.linalg.qr(z)
    inp = [2, 1, 0, 1]
    total = sum(oracle.pd_detection_probability(o, u, inp, x) for o in _patterns(4, 4))
    assert abs(total - 1.0) < 1e-13


def test_uniform_input_norm_matches_oracle():
    from piquasso_amd.distinguishability import uniform_input_norm
    from oracle import oracle as O

    for inp in ([1, 1], [2, 0, 3], [0, 0], [4]):
        for x in (0.0, 0.5, 1.0):
            assert uniform_input_norm(inp, x) == pytest.approx(O.uniform_input_norm(inp, x), rel=1e-15)
    # x = 1: prod n_j!, x = 0: 1
    assert uniform_input_norm([2, 3], 1.0) == math.factorial(2) * math.factorial(3)
    assert uniform_input_norm([2, 3], 0.0) == 1.0


@pytest.mark.parametrize("n", [1, 3, 5])
def test_port_matches_oracle(oracle, n):
    """The plain-numpy restatement of the reference's loop (the CPU
    baseline) computes the same value as the double-double oracle."""
    from oracle.tensor_port import tensor_perm_port

    a = _rand(n, 50 + n)
    for x in (0.0, 0.6):
        ref, asum = oracle.tensor_perm_dd(a, x)
        got = tensor_perm_port(a, x)
        assert abs(got - ref) <= 16 * (n + 2) * np.finfo(float).eps * asum
