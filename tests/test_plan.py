"""Host logic (CPU only): the planner's device arena, checked by emulating the
kernels' arithmetic on it in numpy and comparing with the oracle.

pqp_plan_image returns the exact bytes the HIP kernels read.  The emulation
below follows perm_fwd.hip step by step (start row sums per lane from the
B rows and the lane's top Gray digits, inner sweep adding the tabulated
+-2B rows, coefficient table, lane weight and sign, 2^-(n-1)), so a wrong
table, offset, sign, weight or digit split fails here without a GPU."""

import ctypes
import math

import numpy as np
import pytest

from piquasso_amd import _capi

ITEM = np.dtype([
    ("b_off", "<i8"), ("b2_off", "<i8"), ("step_off", "<i8"), ("coef_off", "<i8"),
    ("alpha_off", "<i8"), ("top_off", "<i8"), ("cpow_off", "<i8"), ("cmul_off", "<i8"),
    ("S_in", "<i8"), ("U", "<i8"), ("partial_base", "<i8"), ("units", "<i4"), ("mcp", "<i4"),
    ("mc", "<i4"), ("n", "<i4"), ("L", "<i4"), ("Dt", "<i4"), ("flags", "<i4"), ("cmax", "<i4"),
    ("out_index", "<i4"), ("pad0", "<i4"),
])
assert ITEM.itemsize == 128


def plan_image(a, rows, cols, target_inner=0, min_units=0, flags=0):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.complex128))
    r = np.ascontiguousarray(np.asarray(rows, dtype=np.int32))
    c = np.ascontiguousarray(np.asarray(cols, dtype=np.int32))
    L = _capi.lib()
    need = ctypes.c_int64(0)
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int32)
    args = (a.view(np.float64).ctypes.data_as(dp), a.shape[0], a.shape[1], a.shape[1],
            r.ctypes.data_as(ip), c.ctypes.data_as(ip), flags, target_inner, min_units)
    _capi.check(L.pqp_plan_image(*args, None, 0, ctypes.byref(need)))
    buf = np.zeros(need.value, dtype=np.uint8)
    _capi.check(L.pqp_plan_image(*args, buf.ctypes.data, need.value, ctypes.byref(need)))
    return buf


INNER = np.dtype([("coef", "<f8"), ("sign2", "<f4"), ("row", "<u4")])


def inner_table(buf, d, S):
    """The interleaved inner table: per inner state v its coefficient and the
    update after it, sign2 (+-2) x inner row of B (S + 1 entries, padding
    past the last step)."""
    off = int(d["step_off"])
    t = np.frombuffer(buf[off: off + 16 * (S + 1)].tobytes(), dtype=INNER)
    assert t["sign2"][S - 1] == 0 and t["sign2"][S] == 0 and t["coef"][S] == 0.0
    assert set(np.abs(t["sign2"][: S - 1]).tolist()) <= {2.0}
    return t["coef"][:S].copy(), t[: max(S - 1, 0)].copy()


def emulate(buf):
    """The kernels' arithmetic on the planned arena, in numpy (complex128)."""
    d = np.frombuffer(buf[:128].tobytes(), dtype=ITEM)[0]
    if d["n"] == 0:
        return 1.0 + 0j
    mcp, L, Dt, S, U = int(d["mcp"]), int(d["L"]), int(d["Dt"]), int(d["S_in"]), int(d["U"])

    def f64(off, count):
        return np.frombuffer(buf[off: off + 8 * count].tobytes(), dtype=np.float64)

    B = f64(int(d["b_off"]), (1 + L + Dt) * mcp * 2).view(np.complex128).reshape(1 + L + Dt, mcp)
    assert int(d["b2_off"]) == int(d["b_off"]) + 16 * mcp  # inner rows of B
    alpha = f64(int(d["alpha_off"]), max(L, 1))
    coef, steps = inner_table(buf, d, S)
    top = f64(int(d["top_off"]), 4 * max(Dt, 1)).reshape(-1, 4)
    cmax = int(d["cmax"])
    po = int(d["cpow_off"])
    masks = [int(x) for x in np.frombuffer(buf[po: po + 24].tobytes(), dtype=np.uint64)]
    cpow = np.frombuffer(buf[po + 24: po + 24 + 4 * max(cmax - 1, 1)].tobytes(), dtype=np.int32)
    grouped = bool(d["flags"] & 8)
    mc = int(d["mc"])
    total = 0j
    for u in range(U):
        cs = B[0].copy()
        for k in range(L):
            cs = cs + alpha[k] * B[1 + k]
        w = 1.0
        parity = 0
        q_hi = 0.0
        for t in range(Dt - 1, -1, -1):
            inv_place, radix, rho = top[t, 0], top[t, 1], int(top[t, 2])
            q = math.floor((u + 0.5) * inv_place)
            cdig = int(q - radix * q_hi)
            q_hi = q
            g = int(radix) - 1 - cdig if parity else cdig
            parity ^= g & 1
            w *= math.comb(rho, g)
            cs = cs + (rho - 2 * g) * B[1 + L + t]
        lane_w = -w if (u & 1) else w
        acc = 0j
        for v in range(S):
            if grouped:
                # prefix products, multiplied in again where the multiplicity drops
                pre = np.cumprod(cs)
                p = pre[-1]
                for j in range(mc):
                    for k in range(3):
                        if (masks[k] >> j) & 1:
                            p *= pre[j]
            else:
                p = np.prod(cs)
                for rr in range(2, cmax + 1):
                    p *= np.prod(cs[: cpow[rr - 2]])
            acc += coef[v] * p
            if v + 1 < S:
                cs = cs + float(steps[v]["sign2"]) * B[1 + int(steps[v]["row"])]
        total += lane_w * acc
    return total / 2.0 ** (int(d["n"]) - 1)


def test_plan_emulation_matches_oracle(oracle):
    rng = np.random.default_rng(31)
    for trial in range(40):
        m_r, m_c = (int(x) for x in rng.integers(1, 7, size=2))
        n = int(rng.integers(0, 9))
        a = rng.standard_normal((m_r, m_c)) + 1j * rng.standard_normal((m_r, m_c))
        r = np.bincount(rng.integers(0, m_r, n), minlength=m_r)
        c = np.bincount(rng.integers(0, m_c, n), minlength=m_c)
        ref = oracle.perm_dd(a, r, c)
        for target in (0, 1, 2, 3, 8, 1 << 20):
            for flags in (0, _capi.PQP_FLAG_EXPAND_COLS, _capi.PQP_FLAG_KEEP_COLS):
                got = emulate(plan_image(a, r, c, target_inner=target, flags=flags))
                assert abs(got - ref) <= 1e-11 * max(abs(ref), 1e-300), (trial, target, flags)


def test_plan_splits_large_items(oracle):
    # 2^12 terms: min_units forces many lanes with short inner sweeps
    rng = np.random.default_rng(2)
    n = 13
    a = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    ones = np.ones(n, int)
    buf = plan_image(a, ones, ones, min_units=4)
    d = np.frombuffer(buf[:128].tobytes(), dtype=ITEM)[0]
    assert d["units"] >= 4 and d["S_in"] * d["U"] == 2 ** (n - 1)
    assert d["flags"] & 2  # all top digits binary
    ref = oracle.perm_dd(a, ones, ones)
    assert abs(emulate(buf) - ref) <= 1e-11 * abs(ref)


def test_orientation_choice(oracle):
    """perm(A; r, c) = perm(A^T; c, r): six unit rows against two columns of
    multiplicity 3 sweep 2^5 Gray terms as given and 4 x 4 / 2 transposed, so
    the planner takes the transposed form (the six rows of A become the
    product's columns) unless PQP_FLAG_AS_GIVEN keeps the rows as digits."""
    rng = np.random.default_rng(8)
    a = rng.standard_normal((6, 2)) + 1j * rng.standard_normal((6, 2))
    r, c = np.ones(6, int), np.array([3, 3])
    ref = oracle.perm_dd(a, r, c)
    for flags, mc in ((0, 6), (_capi.PQP_FLAG_AS_GIVEN, 2)):
        buf = plan_image(a, r, c, flags=flags)
        d = np.frombuffer(buf[:128].tobytes(), dtype=ITEM)[0]
        assert d["mc"] == mc, flags
        assert abs(emulate(buf) - ref) <= 1e-12 * abs(ref), flags
    # a square matrix with unit multiplicities keeps its orientation
    d = np.frombuffer(plan_image(a[:2], [1, 1], [1, 1])[:128].tobytes(), dtype=ITEM)[0]
    assert d["mc"] == 2


def test_column_expansion_layout():
    a = np.arange(6, dtype=float).reshape(2, 3) + 0j
    r, c = [2, 3], [3, 0, 2]
    kept = np.frombuffer(plan_image(a, r, c, flags=_capi.PQP_FLAG_KEEP_COLS)[:128].tobytes(), ITEM)[0]
    exp = np.frombuffer(plan_image(a, r, c, flags=_capi.PQP_FLAG_EXPAND_COLS)[:128].tobytes(), ITEM)[0]
    assert (kept["mc"], kept["cmax"], kept["flags"] & 1) == (2, 3, 0)
    assert (exp["mc"], exp["cmax"], exp["flags"] & 1) == (5, 1, 1)


def test_padding_columns_are_neutral(oracle):
    rng = np.random.default_rng(4)
    mc = 37  # padded to a 40-wide kernel
    a = (rng.standard_normal((2, mc)) + 1j * rng.standard_normal((2, mc))) / np.sqrt(mc)
    r = np.array([18, 19])
    c = np.ones(mc, int)
    buf = plan_image(a, r, c)
    d = np.frombuffer(buf[:128].tobytes(), dtype=ITEM)[0]
    assert d["mcp"] == 40 and d["mc"] == 37
    ref = oracle.perm_dd(a, r, c)
    assert abs(emulate(buf) - ref) <= 1e-9 * abs(ref)


def test_plan_errors():
    with pytest.raises(RuntimeError):
        plan_image(np.eye(2), [1, 1], [2, 1])
    with pytest.raises(ValueError):
        plan_image(np.eye(2), [1, -1], [0, 0])
    with pytest.raises(ValueError):
        plan_image(np.ones((1, 65)), [65], np.ones(65))


def emulate_minors(buf, c):
    """Column-minor plans: the products prod_k cs_k^(c_k - [k = l]) of every
    kernel column l (and the uncut product) summed as the minors kernel does,
    returned in source-column order (kernel columns are the non-zero columns
    stably sorted by decreasing multiplicity) plus the uncut sum."""
    d = np.frombuffer(buf[:128].tobytes(), dtype=ITEM)[0]
    assert d["flags"] & 4
    mc, mcp, L, Dt, S, U = (int(d[k]) for k in ("mc", "mcp", "L", "Dt", "S_in", "U"))

    def f64(off, count):
        return np.frombuffer(buf[off: off + 8 * count].tobytes(), dtype=np.float64)

    B = f64(int(d["b_off"]), (1 + L + Dt) * mcp * 2).view(np.complex128).reshape(1 + L + Dt, mcp)
    assert int(d["b2_off"]) == int(d["b_off"]) + 16 * mcp  # inner rows of B
    alpha = f64(int(d["alpha_off"]), max(L, 1))
    coef, steps = inner_table(buf, d, S)
    top = f64(int(d["top_off"]), 4 * max(Dt, 1)).reshape(-1, 4)
    cm = np.frombuffer(buf[int(d["cmul_off"]): int(d["cmul_off"]) + 4 * mc].tobytes(), dtype=np.int32)
    acc = np.zeros(mc + 1, dtype=np.complex128)
    for u in range(U):
        cs = B[0].copy()
        for k in range(L):
            cs = cs + alpha[k] * B[1 + k]
        w, parity, q_hi = 1.0, 0, 0.0
        for t in range(Dt - 1, -1, -1):
            inv_place, radix, rho = top[t, 0], top[t, 1], int(top[t, 2])
            q = math.floor((u + 0.5) * inv_place)
            cdig = int(q - radix * q_hi)
            q_hi = q
            g = int(radix) - 1 - cdig if parity else cdig
            parity ^= g & 1
            w *= math.comb(rho, g)
            cs = cs + (rho - 2 * g) * B[1 + L + t]
        lane_w = -w if (u & 1) else w
        for v in range(S):
            x = cs[:mc]
            for l in range(mc + 1):
                e = cm - (np.arange(mc) == l)
                acc[l] += lane_w * coef[v] * np.prod(x ** e)
            if v + 1 < S:
                cs = cs + float(steps[v]["sign2"]) * B[1 + int(steps[v]["row"])]
    acc /= 2.0 ** (int(d["n"]) - 1)
    order = sorted(np.flatnonzero(np.asarray(c) > 0), key=lambda j: -c[j])
    out = np.full(len(c), acc[mc])
    for l, j in enumerate(order):
        out[j] = acc[l]
    return out, acc[mc]


def test_minors_plan_emulation_matches_oracle(oracle):
    """The column-minor arena (same tables plus the per-column multiplicity
    table) reproduces perm(A; r, c - e_j) for every column j."""
    rng = np.random.default_rng(77)
    for trial in range(25):
        m_r, m_c = (int(x) for x in rng.integers(1, 6, size=2))
        n = int(rng.integers(1, 7))
        a = rng.standard_normal((m_r, m_c)) + 1j * rng.standard_normal((m_r, m_c))
        r = np.bincount(rng.integers(0, m_r, n), minlength=m_r)
        c = np.bincount(rng.integers(0, m_c, n + 1), minlength=m_c)
        for target in (0, 1, 3):
            got, _ = emulate_minors(
                plan_image(a, r, c, target_inner=target, flags=_capi.PQP_FLAG_GRADIENT), c)
            for j in range(m_c):
                if c[j] == 0:
                    continue
                cj = c.copy()
                cj[j] -= 1
                ref = oracle.perm_dd(a, r, cj)
                scale = max(abs(ref), oracle.perm_dd(np.abs(a), r, cj).real)
                assert abs(got[j] - ref) <= 1e-11 * scale, (trial, target, j)


def test_minors_plan_errors():
    with pytest.raises(RuntimeError):  # sum(c) must be sum(r) + 1
        plan_image(np.eye(2), [1, 1], [1, 1], flags=_capi.PQP_FLAG_GRADIENT)
    with pytest.raises(ValueError):  # column minors are planned up to 32 columns
        plan_image(np.ones((1, 40)), [39], np.ones(40), flags=_capi.PQP_FLAG_GRADIENT)


def test_grouped_power_masks():
    """Snapshot masks: bit j of mask k set when c_j - c_{j+1} > k (c_mc = 1),
    columns sorted by decreasing multiplicity; drops above 3 fall back to
    the count form."""
    a = np.ones((3, 4), dtype=complex)
    buf = plan_image(a, [4, 3, 2], [3, 3, 2, 1])
    d = np.frombuffer(buf[:128].tobytes(), dtype=ITEM)[0]
    po = int(d["cpow_off"])
    masks = [int(x) for x in np.frombuffer(buf[po: po + 24].tobytes(), dtype=np.uint64)]
    assert d["flags"] & 8
    # sorted multiplicities 3 3 2 1: drops d = 0 1 1 0
    assert masks == [0b0110, 0, 0]
    # kept as given: planned transposed, the rows 4 3 2 would be the product side
    buf = plan_image(a, [4, 3, 2], [6, 1, 1, 1], flags=_capi.PQP_FLAG_AS_GIVEN)
    d = np.frombuffer(buf[:128].tobytes(), dtype=ITEM)[0]
    assert not d["flags"] & 8  # drop of 5 > 3: count form


def emulate_gradient(buf, a, r, c):
    """Gradient plans (pqp_grad_batch_c128): the arithmetic of perm_grad.hip on
    the planned arena, unit by unit.  Per term the exclusion products
    Y_j = s c_j cs_j^(c_j - 1) prod_(k != j) cs_k^(c_k); the prepended row's
    gradient is the sum of Y over lanes and terms; an inner row's comes from
    the running wave sum V and its accumulator K (K -= (coefficient change) V
    after each term in which the row moves, then K + final coefficient x V);
    a top row's from its lane-constant coefficient times the lane's sum of Y.
    The rows of B are mapped back to rows of A by value (A's rows are
    distinct here), the prepended row and row i0's digit row both to i0."""
    d = np.frombuffer(buf[:128].tobytes(), dtype=ITEM)[0]
    assert d["flags"] & 32
    mc, mcp, L, Dt, S, U = (int(d[k]) for k in ("mc", "mcp", "L", "Dt", "S_in", "U"))

    def f64(off, count):
        return np.frombuffer(buf[off: off + 8 * count].tobytes(), dtype=np.float64)

    B = f64(int(d["b_off"]), (1 + L + Dt) * mcp * 2).view(np.complex128).reshape(1 + L + Dt, mcp)
    alpha = f64(int(d["alpha_off"]), 2 * max(L, 1))
    a_start, a_final = alpha[:L], alpha[L: 2 * L]
    coef, steps = inner_table(buf, d, S)
    top = f64(int(d["top_off"]), 4 * max(Dt, 1)).reshape(-1, 4)
    cm = np.frombuffer(buf[int(d["cmul_off"]): int(d["cmul_off"]) + 4 * mc].tobytes(), dtype=np.int32)
    rows_B = 1 + L + Dt
    GB = np.zeros((rows_B, mc), dtype=np.complex128)
    perm = 0j
    for unit in range(int(d["units"])):
        lanes = [u for u in range(64 * unit, 64 * unit + 64) if u < U]
        cs, lw, tal = {}, {}, {}
        for u in lanes:
            x = B[0].copy()
            for k in range(L):
                x = x + a_start[k] * B[1 + k]
            w, parity, q_hi = 1.0, 0, 0.0
            al = np.zeros(Dt)
            for t in range(Dt - 1, -1, -1):
                inv_place, radix, rho = top[t, 0], top[t, 1], int(top[t, 2])
                q = math.floor((u + 0.5) * inv_place)
                cdig = int(q - radix * q_hi)
                q_hi = q
                g = int(radix) - 1 - cdig if parity else cdig
                parity ^= g & 1
                w *= math.comb(rho, g)
                al[t] = rho - 2 * g
                x = x + al[t] * B[1 + L + t]
            cs[u], lw[u], tal[u] = x, (-w if (u & 1) else w), al
        W = {u: np.zeros(mc, dtype=np.complex128) for u in lanes}
        V = np.zeros(mc, dtype=np.complex128)
        K = np.zeros((max(L, 1), mc), dtype=np.complex128)
        for v in range(S):
            Yw = np.zeros(mc, dtype=np.complex128)
            for u in lanes:
                x = cs[u][:mc]
                cf = coef[v] * lw[u]
                y = np.array([cf * cm[j] * x[j] ** (cm[j] - 1) *
                              np.prod([x[k] ** cm[k] for k in range(mc) if k != j]) for j in range(mc)])
                perm += cf * np.prod(x ** cm)
                W[u] += y
                Yw += y
            V += Yw
            if v + 1 < S:
                s2, row = float(steps[v]["sign2"]), int(steps[v]["row"])
                K[row] -= s2 * V
                for u in lanes:
                    cs[u] = cs[u] + s2 * B[1 + row]
        GB[0] += V
        for i in range(L):
            GB[1 + i] += K[i] + a_final[i] * V
        for t in range(Dt):
            GB[1 + L + t] += sum(tal[u][t] * W[u] for u in lanes)
    scale = 2.0 ** -(int(d["n"]) - 1)
    perm *= scale
    GB *= scale
    # back to A's rows and columns: kernel columns are the non-zero columns
    # stably sorted by decreasing multiplicity, a column of multiplicity c
    # expanded into c copies whose entries add up
    order = sorted(np.flatnonzero(np.asarray(c) > 0), key=lambda j: -c[j])
    kcols = [j for j in order for _ in range(int(c[j]))]
    assert len(kcols) == mc and (cm == 1).all()
    G = np.zeros(a.shape, dtype=np.complex128)
    for b in range(rows_B):
        i = int(np.argmin(np.abs(a[:, kcols] - B[b, :mc]).sum(axis=1)))
        assert np.allclose(a[i, kcols], B[b, :mc])
        for jj, j in enumerate(kcols):
            G[i, j] += GB[b, jj]
    return perm, G


def test_gradient_plan_emulation_matches_oracle(oracle):
    """The single-sweep gradient arena (the forward tables plus the inner rows'
    final coefficients) reproduces the explicit minors r_i c_j perm(A; r - e_i,
    c - e_j), with inner and top digits of higher radix (row multiplicities)
    and columns of multiplicity c planned as c unit columns."""
    rng = np.random.default_rng(91)
    for trial in range(30):
        m_r, m_c = (int(x) for x in rng.integers(1, 6, size=2))
        n = int(rng.integers(1, 8))
        a = rng.standard_normal((m_r, m_c)) + 1j * rng.standard_normal((m_r, m_c))
        r = np.bincount(rng.integers(0, m_r, n), minlength=m_r)
        c = np.bincount(rng.integers(0, m_c, n), minlength=m_c)
        ref = oracle.grad_dd(a, r, c)
        p_ref = oracle.perm_dd(a, r, c)
        for target in (0, 1, 2, 4):
            buf = plan_image(a, r, c, target_inner=target, flags=_capi.PQP_PLAN_GRADIENT_SWEEP)
            p, g = emulate_gradient(buf, a, r, c)
            scale = max(np.abs(ref).max(), 1e-300)
            assert abs(p - p_ref) <= 1e-11 * max(abs(p_ref), 1e-300), (trial, target)
            assert np.abs(g - ref).max() <= 1e-11 * scale, (trial, target, g, ref)


def test_gradient_plan_final_coefficients():
    """The last inner state of the reflected Gray code: with all counter digits
    at radix - 1, a digit is at its top value when the higher Gray digits
    have even sum, else at 0 (rows with multiplicities 3, 1, 2 as inner)."""
    a = np.arange(1, 13, dtype=float).reshape(4, 3) + 0j
    r = [1, 3, 1, 2]
    buf = plan_image(a, r, [3, 2, 2], target_inner=1 << 10, flags=_capi.PQP_PLAN_GRADIENT_SWEEP)
    d = np.frombuffer(buf[:128].tobytes(), dtype=ITEM)[0]
    L = int(d["L"])
    alpha = np.frombuffer(buf[int(d["alpha_off"]): int(d["alpha_off"]) + 16 * L].tobytes(), np.float64)
    rho = alpha[:L].astype(int)
    # walk the Gray sequence: the final state's coefficients
    coef, steps = inner_table(buf, d, int(d["S_in"]))
    g = np.zeros(L, int)
    for s in steps:
        g[int(s["row"])] += 1 if s["sign2"] < 0 else -1
    assert (alpha[L:] == rho - 2 * g).all()
