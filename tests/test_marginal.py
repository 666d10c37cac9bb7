"""Marginal probabilities and marginal sampling (piquasso_amd/marginal.py)
with the double-double oracle's matrix-polynomial permanent standing in for
the GPU call: the binomial moments as permanent-polynomial coefficients
against the brute-force marginal of the full output distribution, the
reference tests' closed forms (tests/_simulators/passive/test_state.py:1066-1123),
postselection, and the marginal sampler's frequencies against the
probabilities.  GPU parity is in test_gpu_marginal.py."""

import itertools
import math

import numpy as np
import pytest

from test_lossy_pd import bs, haar


@pytest.fixture
def marginal_on_oracle(monkeypatch, oracle):
    import piquasso_amd.marginal as M

    def coeffs(b0, bms, exps):
        return oracle.perm_poly_dd(b0, bms, np.asarray(exps, dtype=np.int32))[0]

    monkeypatch.setattr(M, "perm_poly_coefficients", coeffs)
    return M


def brute_marginal(oracle, u, inp, modes, ps_modes=(), ps_photons=()):
    """Marginal of the full output distribution (|perm|^2 / (out! in!) over
    every output with n photons), joint with the postselection."""
    d, n = u.shape[0], sum(inp)
    rf = math.prod(math.factorial(x) for x in inp)
    out = {}
    for o in itertools.product(range(n + 1), repeat=d):
        if sum(o) != n or any(o[m] != p for m, p in zip(ps_modes, ps_photons)):
            continue
        p = abs(oracle.perm_dd(u, o, inp)) ** 2 / (rf * math.prod(math.factorial(x) for x in o))
        key = tuple(o[m] for m in modes)
        out[key] = out.get(key, 0.0) + p
    return out


@pytest.mark.parametrize("inp,modes,ps", [
    ([1, 2, 0, 1], (2, 0), ((), ())),
    ([1, 1, 1, 0], (3,), ((), ())),
    ([2, 0, 1, 1], (0, 1, 3), ((), ())),
    ([1, 1, 1, 1], (2, 3), ((0,), (1,))),
])
def test_marginal_matches_brute_force(marginal_on_oracle, oracle, inp, modes, ps):
    M = marginal_on_oracle
    u = haar(len(inp), 7)
    got = M.marginal_fock_probabilities(inp, u, ps[0], ps[1], modes)
    ref = brute_marginal(oracle, u, inp, modes, *ps)
    for y, p in got.items():
        assert abs(p - ref.get(y, 0.0)) < 1e-14, y
    assert set(ref) <= set(got)


def test_marginal_reference_closed_forms(marginal_on_oracle):
    M = marginal_on_oracle
    got = M.marginal_fock_probabilities([1, 1], np.eye(2), (), (), (0,))
    assert got.keys() == {(0,), (1,), (2,)}
    assert np.allclose([got[(0,)], got[(1,)], got[(2,)]], [0.0, 1.0, 0.0], atol=1e-15)
    th = np.pi / 5
    got = M.marginal_fock_probabilities([1, 1], bs(th), (), (), (0,))
    want = {(0,): np.sin(2 * th) ** 2 / 2, (1,): np.cos(2 * th) ** 2, (2,): np.sin(2 * th) ** 2 / 2}
    for y, p in want.items():
        assert abs(got[y] - p) < 1e-14


def test_binomial_moments_are_moments(marginal_on_oracle, oracle):
    """b_alpha = E[prod_j C(X_j, alpha_j)] over the brute-force marginal."""
    M = marginal_on_oracle
    u = haar(4, 3)
    inp = [1, 0, 2, 1]
    modes = (1, 3)
    comps, b = M.binomial_moments(inp, u, modes)
    dist = brute_marginal(oracle, u, inp, modes)
    for a, ba in zip(comps, b):
        e = sum(p * math.prod(math.comb(x, int(ai)) for x, ai in zip(y, a)) for y, p in dist.items())
        assert abs(ba - e) < 1e-14


@pytest.mark.parametrize("ps", [((), ()), ((0,), (1,))])
def test_marginal_sampler_frequencies(marginal_on_oracle, ps):
    """4000 seeded samples: each outcome's frequency within 5 standard
    deviations of its probability (conditional on the postselection)."""
    M = marginal_on_oracle
    u = haar(4, 11)
    inp = [1, 1, 1, 0]
    modes = (1, 2)
    probs = M.marginal_fock_probabilities(inp, u, ps[0], ps[1], modes)
    total = sum(probs.values())
    samples = M.generate_marginal_samples(inp, u, modes, 4000, postselect_data=ps,
                                          rng=np.random.default_rng(2))
    assert len(samples) == 4000
    counts = {}
    for s in samples:
        counts[s] = counts.get(s, 0) + 1
    assert set(counts) <= {y for y, p in probs.items() if p > 1e-15}
    for y, p in probs.items():
        q = p / total
        sd = math.sqrt(q * (1 - q) / 4000) + 1e-12
        assert abs(counts.get(y, 0) / 4000 - q) <= 5 * sd + 1e-9, y


def test_marginal_sampler_is_seeded(marginal_on_oracle):
    M = marginal_on_oracle
    u = haar(3, 5)
    a = M.generate_marginal_samples([1, 1, 0], u, (0, 2), 50, seed_sequence=9)
    b = M.generate_marginal_samples([1, 1, 0], u, (0, 2), 50, rng=np.random.default_rng(9))
    assert a == b


def test_marginal_errors(marginal_on_oracle):
    M = marginal_on_oracle
    with pytest.raises(ValueError):
        M.marginal_fock_probabilities([1, 1], np.eye(2), (0,), (1,), (0,))
    with pytest.raises(ValueError):
        M.binomial_moments([1, -1], np.eye(2), (0,))
