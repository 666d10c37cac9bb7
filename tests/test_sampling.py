"""Clifford & Clifford sampling: the oracle's sequential sampler against the
reference's seeded sample lists, the draw equivalences the lockstep sampler
relies on, and the lockstep host logic (with the column minors taken from the
oracle, on the CPU).  The GPU runs are in test_gpu_sampling.py."""

import numpy as np
import pytest

from conftest import load_json


def _cases():
    return load_json("sampling_vectors.json")["cases"]


def _matrix(case):
    u = np.array([[complex(*z) for z in row] for row in case["U"]])
    if "singular_values" in case:
        # both lossy cases build U diag(s) U U^T (test_measurements.py:446,529)
        u = u @ np.diag(case["singular_values"]) @ u @ u.T
    return u


def run_case(case, generate, generate_lossy):
    """One seeded case the way the reference's simulator reaches the sampler
    (piquasso/_simulators/passive/simulation_steps.py:342-367)."""
    u = _matrix(case)
    if case["kind"] == "ideal":
        return generate(case["input"], case["shots"], u, case["seed"], None)
    if case["kind"] == "lossy_nonuniform":
        return generate_lossy(case["input"], case["shots"], u, case["seed"])
    # uniform losses: each particle is lost with 1 - s_0^2, decided by the
    # simulator's generator seeded with the same seed
    g = np.random.default_rng(case["seed"])
    p = np.linalg.svd(u, compute_uv=False)[0] ** 2
    return generate(case["input"], case["shots"], u, case["seed"], lambda: g.uniform() > p)


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c["kind"])
def test_oracle_sampler_matches_reference_samples(case):
    from oracle import sampling as S

    got = run_case(
        case,
        lambda occ, shots, u, seed, rej: S.generate_samples(occ, shots, u, seed, reject_condition=rej),
        lambda occ, shots, u, seed: S.generate_lossy_samples(occ, shots, u, seed),
    )
    assert S.result_order(got, case["seed"]) == [tuple(e) for e in case["expected"]]


def test_oracle_result_order_is_a_permutation():
    from oracle import sampling as S

    samples = [(1, 0), (0, 1), (1, 0), (2, 0), (0, 1)]
    out = S.result_order(samples, 7)
    assert sorted(out) == sorted(samples)


def test_generator_draw_equivalences():
    """The lockstep sampler replaces numpy's Generator.choice by the draws it
    makes: choice(n) == integers(0, n, dtype=int64) and choice(arange(d), p)
    == searchsorted(cumsum(p) / last, random(), 'right')."""
    for seed in range(400):
        r1 = np.random.default_rng(seed)
        r2 = np.random.default_rng(seed)
        for t in range(12):
            n = t % 9 + 1
            assert r1.choice(n) == r2.integers(0, n, dtype=np.int64)
            d = t % 6 + 2
            p = np.abs(np.sin(np.arange(d) + 0.37 * seed + t)) + 1e-3
            p /= p.sum()
            a = r1.choice(np.arange(d), p=p)
            cdf = p.cumsum()
            cdf /= cdf[-1]
            assert a == int(cdf.searchsorted(r2.random(), side="right"))


class _OracleMinors:
    """Stand-in for the GPU minors batch in CPU tests of the lockstep host
    logic: each problem through the oracle's laplace, scattered to all d
    columns like the C ABI does."""

    def __init__(self, u):
        from oracle import oracle as O

        self.O = O
        self.u = np.ascontiguousarray(u, dtype=np.complex128)
        self.d = self.u.shape[0]
        self.seconds = 0.0

    def __call__(self, rows, cols):
        out = np.zeros(rows.shape, dtype=np.complex128)
        for s in range(rows.shape[0]):
            r, c = rows[s], cols[s]
            rm, cm = r > 0, c > 0
            lap = self.O.laplace_dd(self.u[np.ix_(rm, cm)], r[rm], c[cm])
            out[s, np.flatnonzero(cm)[: len(lap)]] = lap
        return out


@pytest.fixture
def lockstep_on_oracle(monkeypatch):
    import piquasso_amd.sampling as P

    monkeypatch.setattr(P, "_MinorsBatch", _OracleMinors)
    return P


def _haar(d, seed):
    rng = np.random.default_rng(seed)
    z = (rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


@pytest.mark.parametrize("variant", ["ideal", "uniform_loss", "postselect", "overlap",
                                     "overlap_loss", "lossy", "postselect_loss",
                                     "postselect_overlap", "postselect_overlap_loss"])
def test_lockstep_host_logic_matches_oracle(lockstep_on_oracle, variant):
    from oracle import sampling as S

    P = lockstep_on_oracle
    d, seed, shots = 6, 11, 24
    occ = [2, 0, 1, 1, 0, 1]
    u = _haar(d, 3)
    kw_p, kw_o = {}, {}
    if variant in ("uniform_loss", "overlap_loss", "postselect_loss"):
        g1 = np.random.default_rng(99)
        g2 = np.random.default_rng(99)
        kw_p["reject_condition"] = lambda: g1.uniform() > 0.7
        kw_o["reject_condition"] = lambda: g2.uniform() > 0.7
    if "overlap" in variant:
        kw_p["uniform_particle_overlap"] = kw_o["uniform_particle_overlap"] = 0.6
    if variant.startswith("postselect"):
        kw_p["postselect_data"] = ((1, 4), (1, 0), 500)
        kw_o["postselect"] = ((1, 4), (1, 0), 500)
    if "postselect_overlap" in variant:
        # the product evaluates the distinguishable particles' postselection
        # probability exactly, not in place as the reference does
        kw_o["ps_acceptance_in_place"] = False
        # the distinguishable particles' loss weight is 1 - sum |T|^2, which
        # rounding can make slightly negative for a unitary (numpy's choice
        # then raises, in the reference as here): a lossy T
        u = 0.9 * u
    if variant == "lossy":
        t = u @ np.diag([0.95, 0.9, 0.8, 0.7, 0.6, 0.5]) @ _haar(d, 4)
        got = P.generate_lossy_samples(occ, shots, t, seed_sequence=seed)
        want = S.generate_lossy_samples(occ, shots, t, seed)
    else:
        got = P.generate_samples(occ, shots, u, seed_sequence=seed, **kw_p)
        want = S.generate_samples(occ, shots, u, seed, **kw_o)
    assert got == want
    assert len(got) == shots


def test_lockstep_particle_count(lockstep_on_oracle):
    P = lockstep_on_oracle
    occ = [1, 1, 1, 0]
    out = P.generate_samples(occ, 10, _haar(4, 1), seed_sequence=5)
    assert all(sum(s) == 3 and len(s) == 4 for s in out)


def test_sampler_argument_errors(lockstep_on_oracle):
    P = lockstep_on_oracle
    u = _haar(3, 0)
    with pytest.raises(ValueError):
        P.generate_samples([1, -1, 0], 2, u, seed_sequence=0)
    with pytest.raises(ValueError):
        P.generate_samples([1, 0], 2, u, seed_sequence=0)
    assert P.generate_samples([1, 1, 0], 0, u, seed_sequence=0) == []


def test_postselection_trial_limit(lockstep_on_oracle):
    P = lockstep_on_oracle
    # identity interferometer: mode 0 can never receive a photon from mode 1
    with pytest.raises(P.SampleGenerationError):
        P.generate_samples([0, 1], 3, np.eye(2), postselect_data=((0,), (1,), 5), seed_sequence=1)


def test_in_place_linear_factor_semantics():
    """The reference multiplies the distinguishable particles' postselection
    polynomial in place (sampling.py:545-548 with polynomial.py:18-38): each
    axis step then reads values already scaled by the constant.  One particle
    and one postselected mode: the coefficient of x is c * p (c = 1 - p), not
    p; with two modes the cross term c * p0 * p1 appears."""
    from oracle.sampling import times_linear

    poly = np.array([1.0, 0.0, 0.0])
    times_linear(poly, 0.7, [0.3], out=poly)
    assert poly.tolist() == [0.7, 0.7 * 0.3, 0.0]
    fresh = times_linear(np.array([1.0, 0.0, 0.0]), 0.7, [0.3], out=np.zeros(3))
    assert fresh.tolist() == [0.7, 0.3, 0.0]
    p2 = np.zeros((2, 2))
    p2[0, 0] = 1.0
    times_linear(p2, 0.5, [0.2, 0.3], out=p2)
    assert np.allclose(p2, [[0.5, 0.5 * 0.3], [0.5 * 0.2, 0.5 * 0.2 * 0.3]])
    # so the reference's acceptance probability for one distinguishable
    # particle reaching a postselected mode with probability 0.3 is 0.21; the
    # product sampler (and the oracle with in_place=False) uses 0.3, and for
    # two particles the exact 2 p (1 - p)
    from oracle.sampling import _ps_probability
    import piquasso_amd.sampling as P

    u = np.array([[np.sqrt(0.7), np.sqrt(0.3)], [np.sqrt(0.3), -np.sqrt(0.7)]], dtype=complex)
    assert np.isclose(_ps_probability(u, [1, 0], [1], np.array([1])), 0.7 * 0.3)
    assert np.isclose(_ps_probability(u, [1, 0], [1], np.array([1]), in_place=False), 0.3)
    assert np.isclose(P._dist_postselection_probability(u, [1, 0], [1], np.array([1])), 0.3)
    assert np.isclose(P._dist_postselection_probability(u, [2, 0], [1], np.array([1])), 2 * 0.3 * 0.7)


def test_postselected_overlap_reference_support(lockstep_on_oracle):
    """The reference's checks for postselection with distinguishable photons
    (tests/_simulators/passive/test_measurements.py:1242-1268): a balanced
    beamsplitter, overlap 0.5, one photon postselected in mode 0 leaves
    exactly one photon in mode 1; with uniform losses 0 or 1."""
    P = lockstep_on_oracle
    bs = np.array([[1, 1], [1, -1]], dtype=complex) / np.sqrt(2)
    out = P.generate_samples([1, 1], 10, bs, postselect_data=((0,), (1,), 1000),
                             uniform_particle_overlap=0.5, seed_sequence=3)
    assert out == [(1,)] * 10
    g = np.random.default_rng(4)
    out = P.generate_samples([1, 1], 10, np.sqrt(0.8) * bs, postselect_data=((0,), (1,), 1000),
                             uniform_particle_overlap=0.5, seed_sequence=3,
                             reject_condition=lambda: g.uniform() > 0.8)
    assert set(out) <= {(0,), (1,)}


def test_vectorised_draws_equal_generator_calls():
    """_Draws gives, shot by shot, the values Generator.integers(0, L,
    dtype=np.int64) and Generator.random() give, in any interleaving: L = 1
    (no bits used), small L, and L near 2^31 and 3 * 2^30 (Lemire's
    rejection taken often), from generators that already hold a cached
    32-bit half."""
    from piquasso_amd.sampling import _Draws

    rs = np.random.default_rng(0)
    for trial in range(60):
        seeds = rs.integers(0, 2 ** 32, 6)
        real = [np.random.default_rng(int(x)) for x in seeds]
        emu = [np.random.default_rng(int(x)) for x in seeds]
        for g1, g2 in zip(real, emu):
            for _ in range(trial % 3):
                g1.integers(0, 5, dtype=np.int64)
                g2.integers(0, 5, dtype=np.int64)
        draws = _Draws(emu, block=4)
        shots = np.arange(6)
        for _ in range(25):
            if rs.random() < 0.5:
                bounds = rs.choice([1, 2, 3, 17, 40, 2 ** 31 + 3, 3 * 2 ** 30], size=6)
                want = [g.integers(0, int(b), dtype=np.int64) for g, b in zip(real, bounds)]
                got = draws.integers(shots, bounds)
            else:
                want = [g.random() for g in real]
                got = draws.doubles(shots)
            assert list(got) == want

