"""GPU: marginal probabilities and marginal sampling from the binomial
moments of one matrix-polynomial permanent (piquasso_amd/marginal.py),
against the marginal of the full output distribution from the GPU's
batched permanents, the double-double oracle's moments, and a uniformly
lossy interferometer against the lossy-probability path."""

import itertools
import math

import numpy as np
import pytest

from test_lossy_pd import haar

pytestmark = pytest.mark.gpu


def _full_marginal(pa, u, inp, modes, ps_modes=(), ps_photons=()):
    d, n = u.shape[0], int(sum(inp))
    outs = np.array([o for o in itertools.product(range(n + 1), repeat=d) if sum(o) == n])
    keep = np.all(outs[:, list(ps_modes)] == np.asarray(ps_photons, dtype=int), axis=1) if ps_modes \
        else np.ones(len(outs), dtype=bool)
    p = pa.output_probabilities(u, inp, outs[keep])
    out = {}
    for o, pk in zip(outs[keep], p):
        key = tuple(int(o[m]) for m in modes)
        out[key] = out.get(key, 0.0) + pk
    return out


@pytest.mark.parametrize("inp,modes,ps", [
    ([1, 1, 1, 1, 1, 0, 0], (0, 3, 6), ((), ())),
    ([2, 0, 1, 1, 0, 1, 0], (1, 5), ((2,), (1,))),
    ([1, 1, 1, 1, 1, 0, 0], (0, 1, 2, 3, 4, 5), ((), ())),
])
def test_marginal_matches_full_distribution(gpu, inp, modes, ps):
    import piquasso_amd as pa
    from piquasso_amd.marginal import marginal_fock_probabilities

    u = haar(len(inp), 21)
    print(f"marginal: input {inp}, modes {modes}, postselection {ps}", flush=True)
    got = marginal_fock_probabilities(inp, u, ps[0], ps[1], modes)
    ref = _full_marginal(pa, u, inp, modes, *ps)
    for y, p in got.items():
        assert abs(p - ref.get(y, 0.0)) < 1e-13, y
    if not ps[0]:
        assert abs(sum(got.values()) - 1.0) < 1e-12


def test_moments_match_oracle(gpu, oracle):
    import piquasso_amd.marginal as M

    u = haar(5, 4)
    inp = [1, 2, 0, 1, 1]
    comps, b = M.binomial_moments(inp, u, (4, 1, 2))
    a = np.repeat(np.arange(5), inp)
    b0 = (a[:, None] == a[None, :]).astype(complex)
    rows = u[np.ix_([4, 1, 2], a)]
    ref, scale = oracle.perm_poly_dd(b0, np.conj(rows)[:, :, None] * rows[:, None, :], comps)
    rf = math.prod(math.factorial(x) for x in inp)
    np.testing.assert_allclose(b, ref.real / rf, atol=1e-13 * max(1.0, scale.max() / rf))


def test_uniformly_lossy_marginal(gpu):
    """T = sqrt(eta) U: the marginal equals the lossy probabilities of
    indistinguishable photons (overlap 1) summed over the other modes."""
    import piquasso_amd as pa
    from piquasso_amd.marginal import marginal_fock_probabilities

    u = haar(4, 8)
    t = np.sqrt(0.7) * u
    inp = [1, 1, 1, 0]
    got = marginal_fock_probabilities(inp, t, (), (), (1, 3))
    outs = pa.fock_space_basis(4, 4)
    p = pa.lossy_detection_probabilities(outs, t, inp, 1.0)
    ref = {}
    for o, pk in zip(outs, p):
        key = (int(o[1]), int(o[3]))
        ref[key] = ref.get(key, 0.0) + pk
    for y, q in got.items():
        assert abs(q - ref.get(y, 0.0)) < 1e-13, y


def test_marginal_sampler_gpu_equals_oracle_moments(gpu, oracle, monkeypatch):
    """The same generator gives the same samples whether the moments come
    from the GPU or from the double-double oracle."""
    import piquasso_amd.marginal as M

    u = haar(6, 30)
    inp = [1, 1, 0, 1, 1, 0]
    got = M.generate_marginal_samples(inp, u, (0, 2, 5), 300, postselect_data=((1,), (1,)),
                                      seed_sequence=17)

    def coeffs(b0, bms, exps):
        return oracle.perm_poly_dd(b0, bms, np.asarray(exps, dtype=np.int32))[0]

    monkeypatch.setattr(M, "perm_poly_coefficients", coeffs)
    want = M.generate_marginal_samples(inp, u, (0, 2, 5), 300, postselect_data=((1,), (1,)),
                                       seed_sequence=17)
    assert got == want
