import json
import math
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running CPU oracle case")


def complex_matrix(d, shape=None):
    re = np.asarray(d["re"], dtype=np.float64)
    im = np.asarray(d["im"], dtype=np.float64)
    a = re + 1j * im
    if shape is not None:
        a = a.reshape(shape)
    return a


def load_json(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def reference_vectors():
    return load_json("reference_vectors.json")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    try:
        O.lib()
    except O.OracleUnavailable as e:  # pragma: no cover - build step missing
        pytest.fail(str(e))
    return O


@pytest.fixture(scope="session")
def gpu():
    """The HIP library, on a machine with a usable device (else the test fails
    loudly: GPU tests must not pass on a silent fallback)."""
    from piquasso_amd import _capi

    n, cus, name = _capi.device_info()
    assert n >= 1
    return {"n_devices": n, "cus": cus, "arch": name}


def rel_err(x, ref):
    x = complex(x)
    ref = complex(ref)
    den = abs(ref)
    return abs(x - ref) / den if den > 0 else abs(x - ref)


def glynn_abs_sum(a, rows, cols):
    """sum_k |term_k| / 2^(n-1) of the Glynn sum with the reference's prepend
    rule (first minimal non-zero row fixed at +1, src/permanent.cpp:54-95).
    Rounding bounds scale with it: |computed - exact| <~ (n + 2) eps times
    this sum, however much the terms cancel.  The column multiplicities need
    not sum to the row multiplicities (column-minor products)."""
    import itertools

    a = np.asarray(a, dtype=np.complex128)
    rows = [int(x) for x in np.asarray(rows).reshape(-1)]
    cols = np.asarray(cols, dtype=np.int64).reshape(-1)
    n = sum(rows)
    if n == 0:
        return 1.0
    nz = [x for x in rows if x > 0]
    i0 = rows.index(min(nz))
    rho = list(rows)
    rho[i0] -= 1
    live = [i for i, x in enumerate(rho) if x > 0]
    if live:
        K = np.array(list(itertools.product(*[range(rho[i] + 1) for i in live])), dtype=np.int64)
    else:
        K = np.zeros((1, 0), dtype=np.int64)
    rv = np.array([rho[i] for i in live], dtype=np.int64)
    w = np.ones(K.shape[0])
    for t, x in enumerate(rv):
        w *= np.array([math.comb(int(x), int(k)) for k in range(x + 1)])[K[:, t]]
    cs = a[i0][None, :] + (rv[None, :] - 2 * K).astype(np.float64) @ a[live] if live else a[i0][None, :]
    mag = np.prod(np.abs(cs) ** cols[None, :], axis=1)
    return float(np.sum(w * mag)) / 2.0 ** (n - 1)
