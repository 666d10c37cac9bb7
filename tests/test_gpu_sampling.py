"""GPU: batched column minors (pqp_laplace_batch_c128) and the lockstep
Clifford & Clifford sampler built on them.  The sampler must reproduce the
reference's seeded sample lists exactly and agree draw for draw with the
oracle's sequential sampler (same seeds, so any difference in a minor that
moves a draw shows up as a different sample)."""

import numpy as np
import pytest

from conftest import load_json
from test_sampling import _haar, run_case

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _rand_problem(rng, m_r, m_c, n):
    """A column-minor problem: rows sum to n, cols to n + 1, some zeros."""
    a = (rng.standard_normal((m_r, m_c)) + 1j * rng.standard_normal((m_r, m_c))) / np.sqrt(2 * m_c)
    r = np.zeros(m_r, dtype=np.int32)
    c = np.zeros(m_c, dtype=np.int32)
    for _ in range(n):
        r[rng.integers(0, m_r)] += 1
    for _ in range(n + 1):
        c[rng.integers(0, m_c)] += 1
    return a, r, c


def test_laplace_batch_matches_single_and_oracle(gpu, oracle):
    import piquasso_amd as pa

    rng = np.random.default_rng(5)
    mats, rows, cols = [], [], []
    for k in range(40):
        n = int(rng.integers(0, 13))
        a, r, c = _rand_problem(rng, int(rng.integers(1, 10)), int(rng.integers(1, 12)), n)
        mats.append(a)
        rows.append(r)
        cols.append(c)
    print("laplace batch: 40 problems, n in [0, 12]", flush=True)
    got = pa.permanent_laplace_batch(mats, rows, cols)
    for k, (a, r, c) in enumerate(zip(mats, rows, cols)):
        single = pa.permanent_laplace(a, r, c)
        ref = oracle.laplace_dd(a, r, c)
        assert got[k].shape == ref.shape, k
        # the same sweep, split into work units differently (summation order)
        np.testing.assert_allclose(got[k], single, rtol=0, atol=TOL * max(1.0, np.abs(single).max()))
        scale = max(np.abs(ref).max(), 1e-300)
        assert np.abs(got[k] - ref).max() <= TOL * scale + 1e-15, k


def test_laplace_batch_c_abi_zero_rows(gpu):
    """A problem with no occupied row (the first photon of every shot) has
    the minor 1 in its occupied column, straight from the C ABI."""
    import ctypes

    from piquasso_amd import _capi

    u = _haar(4, 2)
    rows = np.zeros(4, dtype=np.int32)
    cols = np.array([0, 0, 1, 0], dtype=np.int32)
    out = np.zeros(8)
    dp = ctypes.POINTER(ctypes.c_double)
    i32p = ctypes.POINTER(ctypes.c_int32)
    m = (dp * 1)(u.view(np.float64).ctypes.data_as(dp))
    d = np.array([4], dtype=np.int64)
    i64p = ctypes.POINTER(ctypes.c_int64)
    st = _capi.lib().pqp_laplace_batch_c128(
        1, m, d.ctypes.data_as(i64p), d.ctypes.data_as(i64p), d.ctypes.data_as(i64p),
        (i32p * 1)(rows.ctypes.data_as(i32p)), (i32p * 1)(cols.ctypes.data_as(i32p)),
        (dp * 1)(out.ctypes.data_as(dp)))
    _capi.check(st)
    assert out[4] == 1.0 and out[5] == 0.0


def test_laplace_batch_errors(gpu):
    import piquasso_amd as pa

    a = np.eye(2, dtype=np.complex128)
    with pytest.raises(RuntimeError):
        pa.permanent_laplace_batch([a], [[1, 1]], [[1, 1]])  # sum(cols) != sum(rows) + 1
    with pytest.raises(ValueError):
        pa.permanent_laplace_batch([a], [[1, 1]], [[1, 1, 1]])
    assert pa.permanent_laplace_batch([], [], []) == []


@pytest.mark.parametrize("case", load_json("sampling_vectors.json")["cases"], ids=lambda c: c["kind"])
def test_sampler_reproduces_reference_samples(gpu, case):
    import piquasso_amd as pa
    from oracle import sampling as S

    print(f"seeded case {case['kind']}", flush=True)
    got = run_case(
        case,
        lambda occ, shots, u, seed, rej: pa.generate_samples(occ, shots, u, reject_condition=rej,
                                                             seed_sequence=seed),
        lambda occ, shots, u, seed: pa.generate_lossy_samples(occ, shots, u, seed_sequence=seed),
    )
    assert S.result_order(got, case["seed"]) == [tuple(e) for e in case["expected"]]


@pytest.mark.parametrize("variant", ["ideal", "uniform_loss", "postselect", "overlap", "lossy",
                                     "postselect_overlap"])
def test_sampler_matches_oracle_sampler(gpu, variant):
    import piquasso_amd as pa
    from oracle import sampling as S

    d, seed, shots = 10, 2024, 48
    occ = [2, 1, 0, 1, 1, 0, 2, 0, 1, 0]
    u = _haar(d, 17)
    kw_p, kw_o = {}, {}
    if variant == "uniform_loss":
        g1, g2 = np.random.default_rng(8), np.random.default_rng(8)
        kw_p["reject_condition"] = lambda: g1.uniform() > 0.8
        kw_o["reject_condition"] = lambda: g2.uniform() > 0.8
    if variant == "overlap":
        kw_p["uniform_particle_overlap"] = kw_o["uniform_particle_overlap"] = 0.7
    if variant.startswith("postselect"):
        kw_p["postselect_data"] = ((0, 5), (1, 1), 1000)
        kw_o["postselect"] = ((0, 5), (1, 1), 1000)
    if variant == "postselect_overlap":
        kw_p["uniform_particle_overlap"] = kw_o["uniform_particle_overlap"] = 0.7
        kw_o["ps_acceptance_in_place"] = False  # the product's exact acceptance
        u = 0.9 * u  # lossy: the distinguishable particles' loss weight stays positive
    print(f"sampler vs oracle: {variant}, d={d} n={sum(occ)} shots={shots}", flush=True)
    if variant == "lossy":
        t = u @ np.diag(np.linspace(0.95, 0.5, d)) @ _haar(d, 18)
        got = pa.generate_lossy_samples(occ, shots, t, seed_sequence=seed)
        want = S.generate_lossy_samples(occ, shots, t, seed)
    else:
        got = pa.generate_samples(occ, shots, u, seed_sequence=seed, **kw_p)
        want = S.generate_samples(occ, shots, u, seed, **kw_o)
    assert got == want


def test_sampler_larger_case_matches_oracle(gpu):
    """24 modes, 12 photons (multiplicities up to 2), 64 shots: every shot's
    12 draws agree with the sequential oracle sampler."""
    import piquasso_amd as pa
    from oracle import sampling as S

    d = 24
    occ = [2, 1, 1, 0, 2, 1, 0, 1, 1, 0, 1, 2] + [0] * 12
    u = _haar(d, 31)
    print("sampler vs oracle: d=24 n=12 shots=64", flush=True)
    got = pa.generate_samples(occ, 64, u, seed_sequence=77)
    want = S.generate_samples(occ, 64, u, 77)
    assert got == want


def test_sampler_chunked_minors_calls(gpu, monkeypatch):
    """A round's minor problems split over several batched calls give the
    same samples (the problems are independent)."""
    import piquasso_amd as pa
    import piquasso_amd.sampling as P

    u = _haar(8, 41)
    occ = [1, 2, 0, 1, 1, 0, 0, 1]
    want = pa.generate_samples(occ, 40, u, seed_sequence=5)
    monkeypatch.setattr(P._MinorsBatch, "max_problems", 7)
    got = pa.generate_samples(occ, 40, u, seed_sequence=5)
    assert got == want


def test_sampler_edge_cases(gpu):
    """No photons, one shot, every photon in one mode, a single mode, and the
    identity interferometer (photons stay where they are)."""
    import piquasso_amd as pa
    from oracle import sampling as S

    u = _haar(5, 3)
    assert pa.generate_samples([0, 0, 0, 0, 0], 4, u, seed_sequence=1) == [(0, 0, 0, 0, 0)] * 4
    for occ in ([4, 0, 0, 0, 0], [0, 1, 0, 0, 0], [1, 1, 1, 1, 1]):
        got = pa.generate_samples(occ, 1, u, seed_sequence=9)
        assert got == S.generate_samples(occ, 1, u, 9)
        assert sum(got[0]) == sum(occ)
    assert pa.generate_samples([3], 3, np.ones((1, 1)), seed_sequence=2) == [(3,)] * 3
    occ = [2, 0, 1, 3]
    assert pa.generate_samples(occ, 5, np.eye(4), seed_sequence=4) == [tuple(occ)] * 5
