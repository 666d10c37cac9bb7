"""Parity of the HIP forward kernels with the oracle (GPU).

Every value is compared with the double-double oracle (oracle/glynn_dd.c) at
1e-12 relative error for sizes the oracle finishes quickly, bit-exactly for
integer matrices whose terms stay below 2^53, and with the reference's own
known-answer vectors.  The 1e-10 bar of the engine's specification applies
to the large cases in test_gpu_large.py.
"""

import itertools
import math

import numpy as np
import pytest

import piquasso_amd as pa
from conftest import complex_matrix, rel_err
from piquasso_amd import workloads as W

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _case_matrix(case):
    if "shape" in case:
        return complex_matrix(case["matrix"]).reshape(case["shape"])
    return complex_matrix(case["matrix"])


def test_reference_vectors(gpu, oracle, reference_vectors):
    for case in reference_vectors["permanent"]:
        a = _case_matrix(case)
        r, c = case["rows"], case["cols"]
        got = pa.permanent(a, r, c)
        assert got.dtype == np.complex128 and got.shape == ()
        check = case.get("check")
        if check == "conjugate_symmetry":
            assert rel_err(pa.permanent(np.conj(a), r, c), np.conj(got)) < TOL
            continue
        ref = oracle.perm_dd(a, r, c)
        if case.get("exact"):
            assert complex(got) == ref, case["id"]
        else:
            assert abs(complex(got) - ref) <= TOL * max(abs(ref), 1e-300), (case["id"], got, ref)


def test_reference_probabilities(gpu, reference_vectors):
    for case in reference_vectors["probability"]:
        u = _case_matrix(case)
        inp, out = case["input"], case["output"]
        amp = complex(pa.permanent(u, cols=inp, rows=out))
        norm = np.prod([math.factorial(k) for k in out]) * np.prod([math.factorial(k) for k in inp])
        assert abs(abs(amp) ** 2 / norm - case["expected"]) < 1e-8, case["id"]


def test_real_input_dispatch_matches_reference(gpu):
    got = pa.permanent(np.array([[4.2]]), np.ones(1, int), np.ones(1, int))
    assert got.dtype == np.complex128 and complex(got) == 4.2
    got = pa.permanent([[1.5]], [1], [1])
    assert got.dtype == np.complex64 and complex(got) == 1.5


def test_random_with_multiplicities(gpu, oracle):
    rng = np.random.default_rng(100)
    for trial in range(40):
        m_r, m_c = (int(x) for x in rng.integers(1, 10, size=2))
        n = int(rng.integers(0, 15))
        a = (rng.standard_normal((m_r, m_c)) + 1j * rng.standard_normal((m_r, m_c)))
        r = np.bincount(rng.integers(0, m_r, n), minlength=m_r)
        c = np.bincount(rng.integers(0, m_c, n), minlength=m_c)
        got = complex(pa.permanent(a, r, c))
        ref = oracle.perm_dd(a, r, c)
        assert rel_err(got, ref) < TOL, (trial, m_r, m_c, r, c, got, ref)


def _glynn_abs_sum(a, rows, cols):
    """sum_k |term_k| of the Glynn sum (same prepend rule and scaling), for
    the forward error bound |computed - exact| <~ (n + 2) eps sum |term|."""
    rows = [int(x) for x in rows]
    cols = [int(x) for x in cols]
    n = sum(rows)
    nz = [x for x in rows if x > 0]
    i0 = rows.index(min(nz))
    rho = rows[:]
    rho[i0] -= 1
    total = 0.0
    for ks in itertools.product(*[range(x + 1) for x in rho]):
        w = 1
        for x, k in zip(rho, ks):
            w *= math.comb(x, k)
        cs = a[i0].astype(complex)
        for i, (x, k) in enumerate(zip(rho, ks)):
            cs = cs + (x - 2 * k) * a[i]
        total += w * abs(np.prod(cs ** np.asarray(cols)))
    return total / 2.0 ** (n - 1)


def test_every_kernel_width(gpu, oracle):
    """m_c = 1..64 covers every exact-width kernel and the padded ones (>32).

    Two rows of multiplicity ~m_c/2 keep the term count small but make the
    alternating sum cancel by up to 1e9 at m_c = 64, so each case is held to
    its forward error bound 4 (n + 2) eps sum|term| / |perm| (and 1e-12 at
    least), which is what FP64 row sums and products can promise."""
    rng = np.random.default_rng(7)
    mats, rows, cols, refs, tols = [], [], [], [], []
    for mc in range(1, 65):
        a = (rng.standard_normal((2, mc)) + 1j * rng.standard_normal((2, mc))) / np.sqrt(mc)
        r = np.array([mc // 2, mc - mc // 2])
        c = np.ones(mc, int)
        ref = oracle.perm_dd(a, r, c)
        cond = _glynn_abs_sum(a, r, c) / abs(ref)
        mats.append(a)
        rows.append(r)
        cols.append(c)
        refs.append(ref)
        tols.append(max(TOL, 4 * (mc + 2) * 2.0 ** -53 * cond))
    for k in (0, 7, 8, 15, 23, 31, 32, 39, 47, 55, 63):
        got = complex(pa.permanent(mats[k], rows[k], cols[k]))
        assert rel_err(got, refs[k]) < tols[k], (k + 1, rel_err(got, refs[k]), tols[k])
    vals = pa.permanent_batch(mats, rows, cols)
    for k in range(64):
        assert rel_err(vals[k], refs[k]) < tols[k], (k + 1, rel_err(vals[k], refs[k]), tols[k])


def test_integer_matrices_exact(gpu):
    for n in range(1, 14):
        ones = np.ones(n, int)
        J = np.ones((n, n))
        assert complex(pa.permanent(J.astype(complex), ones, ones)) == math.factorial(n)
        der = round(math.factorial(n) * sum((-1) ** k / math.factorial(k) for k in range(n + 1)))
        assert complex(pa.permanent((J - np.eye(n)).astype(complex), ones, ones)) == der
    rng = np.random.default_rng(3)
    for trial in range(10):
        n = int(rng.integers(2, 12))
        a = rng.integers(0, 2, size=(n, n))
        ones = np.ones(n, int)
        exact = _perm_int_ryser(a)
        assert complex(pa.permanent(a.astype(complex), ones, ones)) == exact


def _perm_int_ryser(a):
    n = a.shape[0]
    total = 0
    for mask in range(1, 1 << n):
        cols = [j for j in range(n) if mask >> j & 1]
        prod = 1
        for i in range(n):
            prod *= int(sum(int(a[i, j]) for j in cols))
        total += (-1) ** (n - len(cols)) * prod
    return total


def test_modes_agree(gpu, oracle):
    rng = np.random.default_rng(5)
    for n in (6, 11, 17):
        a = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(2 * n)
        ones = np.ones(n, int)
        ref = oracle.perm_dd(a, ones, ones)
        v_dd, dd = pa.permanent_ex(a, ones, ones, mode="dd")
        v_plain, _ = pa.permanent_ex(a, ones, ones, mode="plain")
        v_32, _ = pa.permanent_ex(a, ones, ones, mode="fp32")
        assert rel_err(v_dd, ref) < TOL
        assert rel_err(complex(dd[0] + dd[1], dd[2] + dd[3]), ref) < TOL
        assert rel_err(v_plain, ref) < 1e-10
        assert rel_err(v_32, ref) < 5e-3


def test_complex64_path(gpu, oracle):
    rng = np.random.default_rng(8)
    a = ((rng.standard_normal((9, 9)) + 1j * rng.standard_normal((9, 9))) / 3).astype(np.complex64)
    r = np.array([1, 1, 2, 0, 1, 1, 1, 1, 1])
    c = np.array([1, 1, 1, 1, 1, 2, 0, 1, 1])
    got = pa.permanent(a, r, c)
    assert got.dtype == np.complex64 and got.shape == ()
    ref = oracle.perm_dd(a.astype(np.complex128), r, c)
    assert rel_err(complex(got), ref) < 1e-4


def test_errors(gpu):
    with pytest.raises(RuntimeError):
        pa.permanent(np.eye(3, dtype=complex), [1, 1, 1], [1, 1, 0])
    with pytest.raises(ValueError):
        pa.permanent(np.eye(3, dtype=complex), [1, 1], [1, 1, 1])
    with pytest.raises(ValueError):
        pa.permanent(np.eye(2, dtype=complex), [2, -1], [1, 0])
    assert complex(pa.permanent(np.zeros((0, 0), dtype=complex), [], [])) == 1.0
    assert complex(pa.permanent(np.eye(3, dtype=complex), [0, 0, 0], [0, 0, 0])) == 1.0


def test_batch_matches_single_and_oracle(gpu, oracle):
    mats, rows, cols = W.batched_uneven(count=96, seed=44, n_lo=4, n_hi=14)
    refs = [oracle.perm_dd(m, r, c) for m, r, c in zip(mats, rows, cols)]
    with pa.PermanentBatch(mats, rows, cols) as b:
        eager = b.run("eager").results()
        graph = [b.run("graph").results() for _ in range(3)]
        plain = b.run("plain").results()
        eager2 = b.run("eager").results()
        graph2 = b.run("graph").results()
    for k in range(len(mats)):
        assert rel_err(eager[k], refs[k]) < TOL, k
        assert rel_err(plain[k], refs[k]) < 1e-10, k
    # deterministic: replays and eager runs give identical bits
    for g in graph + [graph2, eager2]:
        assert np.array_equal(g, eager)
    single = [complex(pa.permanent(m, r, c)) for m, r, c in zip(mats[:8], rows[:8], cols[:8])]
    for k in range(8):
        assert rel_err(single[k], refs[k]) < TOL


def test_static_and_persistent_agree(gpu):
    mats, rows, cols = W.batched_uneven(count=64, seed=45, n_lo=8, n_hi=16)
    with pa.PermanentBatch(mats, rows, cols, persistent=True) as b1:
        v1 = b1.run("eager").results()
    with pa.PermanentBatch(mats, rows, cols, persistent=False) as b2:
        v2 = b2.run("eager").results()
    assert np.array_equal(v1, v2)


def test_inner_size_does_not_change_the_value(gpu, oracle):
    a, r, c = W.large_single(n=16, seed=9)
    ref = oracle.perm_dd(a, r, c)
    for target in (1, 2, 8, 64, 1024, 1 << 15):
        with pa.PermanentBatch([a], [r], [c], target_inner=target) as b:
            got = b.run("eager").results()[0]
        assert rel_err(got, ref) < TOL, target


def test_boson_sampling_probability(gpu, oracle):
    u, rows, cols = W.boson_sampling(seed=2)
    got = complex(pa.permanent(u, rows, cols))
    ref = oracle.perm_dd(u, rows, cols)
    assert rel_err(got, ref) < 1e-11


def test_batched_output_probabilities(gpu, reference_vectors):
    """output_probabilities (one batch for all output patterns) reproduces the
    reference's probability known answers and sums to 1 over every output
    pattern of a unitary (3 photons in 5 modes: 35 patterns)."""
    for case in reference_vectors["probability"]:
        u = _case_matrix(case)
        p = pa.output_probabilities(u, case["input"], [case["output"]])
        assert abs(p[0] - case["expected"]) < 1e-8, case["id"]
    rng = np.random.default_rng(9)
    u = W.haar_unitary(5, rng)
    s = [1, 0, 1, 1, 0]
    outs = [t for t in itertools.product(range(4), repeat=5) if sum(t) == 3]
    p = pa.output_probabilities(u, s, outs)
    assert len(outs) == 35 and abs(p.sum() - 1.0) < 1e-12
    # a pattern with the wrong photon number has probability 0
    assert pa.output_probabilities(u, s, [[1, 1, 1, 1, 0]])[0] == 0.0


@pytest.mark.parametrize("cols", [
    [3, 3, 2, 1],            # multiplicity drops 0, 1, 1, 0: grouped powers
    [4, 1, 1, 2],            # a drop of 2 after sorting (4, 2, 1, 1)
    [4, 1, 1],               # a drop of 3: third snapshot mask
    [5, 1],                  # a drop of 4: count form
    [2] * 6 + [1] * 3,       # one drop, nine columns (bucket 9-16)
    [1, 3, 1, 1, 2, 1, 1, 1, 1, 1, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1],  # 26 columns: count form
])
def test_column_power_forms(gpu, oracle, cols):
    """Column multiplicities through every product form of the sweep kernel:
    prefix snapshots with one, two or three masks, the count form for larger
    drops and for kernels wider than the snapshot form's limit."""
    from conftest import glynn_abs_sum

    rng = np.random.default_rng(len(cols) * 7 + sum(cols))
    c = np.array(cols)
    n = int(c.sum())
    m_r = min(n, 6)
    r = np.bincount(np.concatenate([np.arange(m_r), rng.integers(0, m_r, n - m_r)]), minlength=m_r)
    a = (rng.standard_normal((m_r, len(c))) + 1j * rng.standard_normal((m_r, len(c)))) / np.sqrt(2 * len(c))
    ref = oracle.perm_dd(a, r, c)
    got = complex(pa.permanent(a, r, c))
    tol = max(TOL * abs(ref), 4 * (n + 2) * 2.0 ** -53 * glynn_abs_sum(a, r, c))
    assert abs(got - ref) <= tol, (cols, got, ref)
    # the same item inside a batch (another launch path) gives the same value
    vals = pa.permanent_batch([a, a], [r, r], [c, c])
    assert abs(vals[0] - ref) <= tol and vals[0] == vals[1]


def test_drop_in_argument_forms_agree(gpu):
    """The lean complex128 path and the general path give the same value for
    every argument form (strided / typed multiplicities, lists, views)."""
    import piquasso_amd as pa

    rng = np.random.default_rng(12)
    a = (rng.standard_normal((7, 7)) + 1j * rng.standard_normal((7, 7))) / np.sqrt(14)
    r = np.array([1, 0, 2, 1, 0, 1, 1])
    c = np.array([2, 1, 0, 1, 1, 0, 1])
    ref = complex(pa.permanent(a, r.tolist(), c.tolist()))
    forms = [
        (a, r, c),
        (a, r.astype(np.int32), c.astype(np.uint8)),
        (a, np.repeat(r, 2)[::2], np.repeat(c, 2)[::2]),  # strided multiplicities
        (np.asfortranarray(a), r, c),                        # general path
        (a, list(r), tuple(c)),
    ]
    for m, rr, cc in forms:
        v = pa.permanent(m, rr, cc)
        assert v.shape == () and v.dtype == np.complex128
        assert abs(complex(v) - ref) <= 1e-13 * abs(ref)


def test_batch_shared_matrix_marshalling(gpu):
    """A batch that reuses one matrix object with stacked multiplicities
    (the probability-table form) gives the same permanents as the same items
    passed one by one."""
    rng = np.random.default_rng(21)
    u = W.haar_unitary(9, rng)
    outs = np.array([np.bincount(c, minlength=9)
                     for c in itertools.combinations_with_replacement(range(9), 4)])[::7]
    inp = np.array([1, 1, 0, 1, 0, 0, 1, 0, 0])
    with pa.PermanentBatch([u] * len(outs), outs, [inp] * len(outs)) as b:
        b.run("eager")
        shared = b.results()
    copies = [u.copy() for _ in outs]  # distinct objects: the item-by-item path
    with pa.PermanentBatch(copies, list(outs), [inp] * len(outs)) as b:
        b.run("eager")
        single = b.results()
    np.testing.assert_array_equal(shared, single)
    p = pa.output_probabilities(u, inp, outs)
    assert p.shape == (len(outs),) and np.all(p >= 0)


def test_superposition_particle_number_probabilities(gpu, oracle):
    """particle_number_probabilities: |sum_j c_j <t|U|s_j>|^2 for a
    superposition of input Fock states (the reference's
    get_ideal_particle_number_probability, probabilities.py:26-54), against
    the double-double oracle, one input reducing to output_probabilities,
    the identity interferometer (|c_j|^2 on each input pattern) and the
    normalisation of a unitary."""
    from piquasso_amd.probabilities import (calculate_inner_product,
                                            get_ideal_particle_number_probability)

    rng = np.random.default_rng(12)
    u = W.haar_unitary(5, rng)
    ins = np.array([[1, 0, 1, 1, 0], [0, 2, 0, 0, 1], [3, 0, 0, 0, 0], [1, 1, 0, 0, 0]])
    c = rng.standard_normal(4) + 1j * rng.standard_normal(4)
    c /= np.linalg.norm(c[:3])  # the three 3-photon inputs are normalised; the 2-photon one is extra
    outs = np.array([t for t in itertools.product(range(4), repeat=5) if sum(t) in (2, 3)])
    p = pa.particle_number_probabilities(outs, u, ins, c)
    fact = [math.factorial(k) for k in range(4)]
    for t, pt in list(zip(outs, p))[::9]:
        amp = 0j
        for s, cj in zip(ins, c):
            if s.sum() == t.sum():
                amp += cj * oracle.perm_dd(u, t, s) / math.sqrt(
                    np.prod([fact[k] for k in t]) * np.prod([fact[k] for k in s]))
        assert abs(pt - abs(amp) ** 2) < 1e-13, t
    three = outs.sum(axis=1) == 3
    assert abs(p[three].sum() - 1.0) < 1e-12
    assert abs(p[~three].sum() - abs(c[3]) ** 2) < 1e-12
    one = pa.particle_number_probabilities(outs, u, ins[:1], [1.0])
    np.testing.assert_allclose(one, pa.output_probabilities(u, ins[0], outs), atol=1e-15)
    pid = pa.particle_number_probabilities(ins, np.eye(5), ins, c)
    np.testing.assert_allclose(pid, np.abs(c) ** 2, atol=1e-15)
    assert abs(get_ideal_particle_number_probability(outs[7], u, ins, c) - p[7]) < 1e-15
    amp = calculate_inner_product(u, ins[0], outs[40])
    assert abs(amp - complex(pa.output_amplitudes(u, ins[0], [outs[40]])[0])) == 0


def test_single_calls_around_the_one_block_limit(gpu, oracle):
    """Single permanents just below and above 16 384 Gray terms: one block of
    the device (per-unit start rows where the top digits are binary) or one
    static launch over the device whose last block finalizes the permanent
    (its block counter must be back at zero for the next call).  Calls of
    each kind follow one another, every result against the double-double
    oracle."""
    rng = np.random.default_rng(12)
    cases = [
        (np.ones(12, int), np.ones(12, int)),            # 2^11 terms, binary top digits
        (np.ones(13, int), np.ones(13, int)),            # 2^12
        (np.ones(16, int), np.ones(16, int)),            # 2^15: spread, width 16
        (np.array([1] * 11 + [2] * 3), np.array([2] + [1] * 15)),  # 27 648 terms: spread, mixed radix
        (np.ones(15, int), np.ones(15, int)),            # 2^14: one block at the limit
    ]
    for rep in range(3):
        for r, c in cases:
            a = (rng.standard_normal((len(r), len(c))) + 1j * rng.standard_normal((len(r), len(c)))) / 3
            got = complex(pa.permanent(a, r, c))
            ref = oracle.perm_dd(a, r, c)
            assert rel_err(got, ref) < TOL, (rep, r.tolist(), c.tolist())
    # complex64 input takes the FP32-product variant of the same launches
    a = (rng.standard_normal((16, 16)) + 1j * rng.standard_normal((16, 16))).astype(np.complex64) / 3
    ones = np.ones(16, int)
    ref = oracle.perm_dd(a.astype(np.complex128), ones, ones)
    got = complex(pa.permanent(a, ones, ones))
    assert rel_err(got, ref) < 1e-3
