"""GPU parity of the tensor-permanent kernels (tensor_perm.hip) against the
double-double oracle, the x = 1 / x = 0 limits at sizes the oracle cannot
reach (against the GPU's own Glynn permanent, an independent kernel and
formula), detection probabilities (HOM dip, normalisation), batches and
errors."""

import itertools

import numpy as np
import pytest

from test_tensor_permanent import _patterns, _rand

pytestmark = pytest.mark.gpu

EPS = np.finfo(float).eps


def _tol(n, abs_sum, ref):
    # FP64 products of n factors, summed in double-double: forward error
    # ~ (3n) eps per term relative to sum |term|
    return max(1e-12 * abs(ref), 8 * (n + 2) * EPS * abs_sum)


@pytest.mark.parametrize("n", list(range(0, 12)))
def test_tensor_permanent_matches_oracle(gpu, oracle, n):
    import piquasso_amd as pa

    a = _rand(n, 100 + n) if n else np.zeros((0, 0), dtype=complex)
    for x in (0.0, 0.37, 1.0):
        print(f"tensor permanent n={n} x={x}", flush=True)
        (got,), (gsum,) = pa.tensor_permanent_batch([a], [x], with_abs_sum=True)
        ref, asum = oracle.tensor_perm_dd(a, x)
        assert abs(got - ref.real) <= _tol(n, asum, ref.real), (n, x, got, ref)
        # the kernel's |term| sum weights the (s, t) / (t, s) pairs by their real parts
        assert gsum <= asum * (1 + 1e-12) + 1e-300


@pytest.mark.parametrize("n", [13, 14, 16])
def test_tensor_permanent_limits_large(gpu, n):
    """x = 1 gives |perm A|^2 and x = 0 gives perm(|A|^2): checked against
    the Glynn kernels at sizes beyond the oracle's reach."""
    import piquasso_amd as pa

    a = _rand(n, 200 + n)
    print(f"tensor permanent limits n={n}", flush=True)
    (t1, t0), (s1, s0) = pa.tensor_permanent_batch([a, a], [1.0, 0.0], with_abs_sum=True)
    p = complex(pa.permanent(a, np.ones(n, int), np.ones(n, int)))
    p2 = complex(pa.permanent(np.abs(a) ** 2, np.ones(n, int), np.ones(n, int)))
    # the subset-pair sum cancels: its error is bounded by its own sum of |terms|
    print(f"  cancellation x=1: {s1 / abs(t1):.2e}, x=0: {s0 / abs(t0):.2e}", flush=True)
    assert abs(t1 - abs(p) ** 2) <= 8 * (n + 2) * EPS * s1 + 1e-13 * abs(p) ** 2
    assert abs(t0 - p2.real) <= 8 * (n + 2) * EPS * s0 + 1e-13 * abs(p2)


def test_tensor_batch_mixed_orders(gpu, oracle):
    import piquasso_amd as pa

    mats = [_rand(n, 300 + k) for k, n in enumerate([5, 9, 2, 9, 7, 0, 10, 1])]
    mats[5] = np.zeros((0, 0), dtype=complex)
    xs = [0.1, 0.5, 0.9, 0.0, 1.0, 0.3, 0.66, 0.2]
    got = pa.tensor_permanent_batch(mats, xs)
    for k, (a, x) in enumerate(zip(mats, xs)):
        ref, asum = oracle.tensor_perm_dd(a, x)
        assert abs(got[k] - ref.real) <= _tol(a.shape[0], asum, ref.real), k
        assert got[k] == pa.tensor_permanent(a, x) or abs(got[k] - pa.tensor_permanent(a, x)) <= 1e-15 * abs(got[k])


def test_detection_probabilities(gpu, oracle):
    import piquasso_amd as pa

    bs = np.array([[1, 1], [1, -1]]) / np.sqrt(2)
    for x in (0.0, 0.3, 1.0):
        assert abs(pa.detection_probability([1, 1], bs, [1, 1], x) - (1 - x * x) / 2) < 1e-14
    rng = np.random.default_rng(4)
    z = (rng.standard_normal((5, 5)) + 1j * rng.standard_normal((5, 5))) / np.sqrt(2)
    u, _ = np.linalg.qr(z)
    inp = [2, 1, 0, 1, 1]
    outs = list(_patterns(5, 5))
    print(f"detection probabilities: {len(outs)} patterns", flush=True)
    for x in (0.0, 0.6, 1.0):
        p = pa.detection_probabilities(outs, u, inp, x)
        assert abs(p.sum() - 1.0) < 1e-12
        for o, pk in list(zip(outs, p))[::17]:
            assert abs(pk - oracle.pd_detection_probability(o, u, inp, x)) < 1e-13
    # x = 1: the indistinguishable probabilities
    ideal = pa.output_probabilities(u, inp, outs)
    np.testing.assert_allclose(pa.detection_probabilities(outs, u, inp, 1.0), ideal, atol=1e-14)
    # a pattern with another photon number has probability 0
    assert pa.detection_probability([1, 0, 0, 0, 0], u, inp, 0.5) == 0.0


def test_tensor_errors(gpu):
    import piquasso_amd as pa

    with pytest.raises(ValueError):
        pa.tensor_permanent(np.zeros((2, 3)), 0.5)
    with pytest.raises(ValueError):
        pa.tensor_permanent(np.zeros((21, 21)), 0.5)
    with pytest.raises(ValueError):
        pa.tensor_permanent(np.eye(2), float("nan"))
    with pytest.raises(ValueError):
        pa.detection_probabilities([[1, 0]], np.eye(2), [1, -1], 0.5)
    assert pa.tensor_permanent_batch([], []).shape == (0,)


def test_tensor_permanent_n18_limits(gpu):
    """n = 18 (4^17 / 2 pairs, ~0.06 s): the x = 1 and x = 0 limits against
    the Glynn kernels."""
    import piquasso_amd as pa
    from piquasso_amd import workloads as W

    a = W.tensor(18)
    print("tensor permanent limits n=18", flush=True)
    (t1, t0), (s1, s0) = pa.tensor_permanent_batch([a, a], [1.0, 0.0], with_abs_sum=True)
    ones = np.ones(18, int)
    p = complex(pa.permanent(a, ones, ones))
    p2 = complex(pa.permanent(np.abs(a) ** 2, ones, ones))
    assert abs(t1 - abs(p) ** 2) <= 8 * 20 * EPS * s1 + 1e-13 * abs(p) ** 2
    assert abs(t0 - p2.real) <= 8 * 20 * EPS * s0 + 1e-13 * abs(p2)
