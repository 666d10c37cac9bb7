import json
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
