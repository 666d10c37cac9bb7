"""GPU: host-forced failures leave the engine usable.  The allocation hook
(pqp_debug_fail_allocation) fails each device or pinned-host allocation of
a batch's creation in turn; every failure must surface as MemoryError (the
C library's PQP_ERR_NOMEM), release what the batch held, and leave later
calls exact.  PQP_FORCE_NO_DEVICE must turn a machine with a device into a
clean PQP_ERR_DEVICE without initialising HIP.  No kernel runs on a failed
path: the failures are decided on the host before any launch."""

import os
import subprocess
import sys

import numpy as np
import pytest

from test_abi import _NO_DEVICE_CHILD

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _items(seed):
    rng = np.random.default_rng(seed)
    out = []
    for m, n in ((5, 6), (7, 9), (4, 12), (8, 8)):
        a = rng.normal(size=(m, m)) + 1j * rng.normal(size=(m, m))
        r = np.zeros(m, dtype=np.int64)
        c = np.zeros(m, dtype=np.int64)
        np.add.at(r, rng.integers(0, m, n), 1)
        np.add.at(c, rng.integers(0, m, n), 1)
        out.append((a, r, c))
    return out


@pytest.mark.parametrize("gradient", [False, True])
def test_each_allocation_failure_is_clean(gpu, oracle, gradient):
    from piquasso_amd import _capi
    from piquasso_amd.engine import PermanentBatch

    items = _items(3 + gradient)
    mats, rows, cols = zip(*items)
    want = np.array([oracle.perm_dd(a, r, c) for a, r, c in items])
    L = _capi.lib()
    failures = 0
    for k in range(64):
        print(f"allocation failure {k}, gradient={gradient}", flush=True)
        L.pqp_debug_fail_allocation(k)
        try:
            b = PermanentBatch(list(mats), list(rows), list(cols), gradient=gradient)
        except MemoryError as e:
            assert "forced failure" in str(e)
            failures += 1
            continue
        finally:
            L.pqp_debug_fail_allocation(-1)
        got = b.run("eager").results()
        assert np.max(np.abs(got - want) / np.abs(want)) < 1e-12
        if gradient:
            for g, (a, r, c) in zip(b.gradients(), items):
                ref = oracle.grad_dd(a, r, c)
                assert np.max(np.abs(g - ref)) <= 1e-11 * np.max(np.abs(ref))
        b.close()
        break
    else:
        pytest.fail("batch creation never succeeded")
    # arena, partials, results, queue heads, pinned stage (a gradient batch
    # is one plan too: its sweep gives the permanents and the gradients)
    assert failures >= 5
    # the hook is off again and the library still computes
    b = PermanentBatch(list(mats), list(rows), list(cols))
    assert np.max(np.abs(b.run("graph").results() - want) / np.abs(want)) < 1e-12
    b.close()


def test_single_call_after_a_forced_failure(gpu, oracle):
    """A single-permanent call that finds its cached buffers too small fails
    cleanly under the hook; either way the next call is exact."""
    import piquasso_amd as pa
    from piquasso_amd import _capi

    a, r, c = _items(9)[1]
    L = _capi.lib()
    L.pqp_debug_fail_allocation(0)
    try:
        pa.permanent(a, r, c)
    except MemoryError as e:
        assert "forced failure" in str(e)
    finally:
        L.pqp_debug_fail_allocation(-1)
    got = pa.permanent(a, r, c)
    assert abs(got - oracle.perm_dd(a, r, c)) / abs(oracle.perm_dd(a, r, c)) < 1e-12


def test_forced_no_device_on_a_machine_with_one(gpu):
    env = dict(os.environ, PQP_FORCE_NO_DEVICE="1")
    print("forced no-device child", flush=True)
    res = subprocess.run([sys.executable, "-c", _NO_DEVICE_CHILD], cwd=ROOT, env=env,
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    lines = res.stdout.splitlines()
    assert lines[0].startswith("status -4 0")
    assert "PQP_FORCE_NO_DEVICE" in lines[1]
    assert lines[2].startswith("raised")
