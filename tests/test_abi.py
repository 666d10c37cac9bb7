"""The C-ABI boundary without a GPU: the library loads, exports every entry
point that include/pqp_permanent.h declares, the ctypes table binds each of
them, and the host-side argument handling (dtype dispatch, validation,
error mapping) behaves like the reference's pybind module.  No compute call
reaches a device here."""

import ctypes
import os
import re

import numpy as np
import pytest

from piquasso_amd import _capi
from piquasso_amd._math.permanent import resolve_dtype

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pqp_permanent.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pqp_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for must in ("pqp_permanent_c128", "pqp_permanent_c64", "pqp_laplace_c128",
                 "pqp_grad_batch_c128", "pqp_batch_create", "pqp_batch_run",
                 "pqp_batch_gradients", "pqp_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_capi.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_ctypes_table_covers_the_header():
    assert set(declared_functions()) == set(_capi.SIGNATURES)


def test_constants_match_the_header():
    """Every PQP_* constant of the header (status codes, flags, run modes) has
    the same value in the Python binding."""
    text = open(HEADER).read()
    defs = dict(re.findall(r"#define\s+(PQP_[A-Z0-9_]+)\s+\(?(-?(?:0x)?[0-9a-fA-F]+)u?\)?", text))
    assert len(defs) >= 15
    for name, value in defs.items():
        if hasattr(_capi, name):
            assert getattr(_capi, name) == int(value, 0), name
    for name in ("PQP_FLAG_AS_GIVEN", "PQP_FLAG_GRADIENT", "PQP_PLAN_GRADIENT_SWEEP", "PQP_RUN_KERNEL",
                 "PQP_ERR_NOKERNEL"):
        assert name in defs and hasattr(_capi, name), name


def test_version_and_device_query_without_gpu():
    L = _capi.lib()
    assert L.pqp_version().decode().startswith("pqp-permanent")
    n = ctypes.c_int(-1)
    st = L.pqp_device_info(ctypes.byref(n), None, None, 0)
    if st != 0:  # no device in this container: a clean error, not a crash
        assert st == _capi.PQP_ERR_DEVICE
        assert L.pqp_last_error().decode()


@pytest.mark.parametrize("dtype,expect", [
    (np.complex64, np.complex64), (np.float32, np.complex64), (np.int8, np.complex64),
    (np.bool_, np.complex64), (np.complex128, np.complex128), (np.float64, np.complex128),
    (np.int64, np.complex128), (np.int32, np.complex128),
])
def test_dtype_dispatch_mirrors_pybind(dtype, expect):
    """The complex dtype the drop-in computes in, for each input dtype, is the
    one the reference's pybind overloads pick (piquasso/_math/permanent.cpp:73-78):
    compared with the reference module built from its sources when that
    build exists (oracle/_ref), else with the expected table."""
    a = np.eye(2, dtype=dtype)
    assert resolve_dtype(a) == expect
    from oracle import oracle as O

    try:
        ref = O.ref_module().permanent
    except O.OracleUnavailable:
        return
    out = ref(a, rows=np.ones(2, dtype=np.int64), cols=np.ones(2, dtype=np.int64))
    assert out.dtype == expect


def test_dtype_dispatch_lists_and_errors():
    assert resolve_dtype([[1.0, 2.0], [3.0, 4.0]]) == np.complex64
    assert resolve_dtype(np.zeros((2, 4), dtype=np.complex128)[:, ::2]) == np.complex128
    with pytest.raises(TypeError):
        resolve_dtype(np.zeros((2, 2), dtype=np.longdouble))
    with pytest.raises(TypeError):
        resolve_dtype(np.zeros((2, 2), dtype=object))


def test_argument_errors_precede_the_device():
    """Sum mismatch -> RuntimeError (the reference's message), bad shapes ->
    ValueError; both raised by the planner before any device is touched."""
    import piquasso_amd as pa

    with pytest.raises(RuntimeError, match="Number of input and output states"):
        pa.permanent(np.eye(2, dtype=np.complex128), [1, 1], [2, 1])
    with pytest.raises(ValueError):
        pa.permanent(np.eye(2, dtype=np.complex128), [1, 1, 0], [1, 1])
    with pytest.raises(ValueError):
        pa.permanent(np.ones(3, dtype=np.complex128), [1], [1])
    with pytest.raises(ValueError):
        pa.permanent(np.eye(2, dtype=np.complex128), [1, -1], [0, 0])
    with pytest.raises(RuntimeError):
        pa.permanent_laplace(np.eye(2), [1, 1], [1, 1])


def test_batch_entry_errors_precede_the_device():
    """The batched column-minor and tensor-permanent entries validate their
    arguments before touching a device."""
    import piquasso_amd as pa

    L = _capi.lib()
    dp = ctypes.POINTER(ctypes.c_double)
    i64p = ctypes.POINTER(ctypes.c_int64)
    assert L.pqp_laplace_batch_c128(-1, None, None, None, None, None, None, None) == _capi.PQP_ERR_INVALID
    assert L.pqp_laplace_batch_c128(1, None, None, None, None, None, None, None) == _capi.PQP_ERR_INVALID
    assert L.pqp_laplace_batch_c128(0, None, None, None, None, None, None, None) == _capi.PQP_OK
    with pytest.raises(RuntimeError):  # sum(cols) != sum(rows) + 1
        pa.permanent_laplace_batch([np.eye(2)], [[1, 1]], [[1, 1]])
    assert L.pqp_tensor_permanent_batch_c128(-1, None, None, None, None, None, None) == _capi.PQP_ERR_INVALID
    a = np.zeros((2, 2), dtype=np.complex128)
    mats = (dp * 1)(a.view(np.float64).ctypes.data_as(dp))
    out = np.zeros(1)

    def tensor(n, ld, x):
        nn = np.array([n], dtype=np.int64)
        ll = np.array([ld], dtype=np.int64)
        xx = np.array([x])
        return L.pqp_tensor_permanent_batch_c128(1, mats, nn.ctypes.data_as(i64p), ll.ctypes.data_as(i64p),
                                                 xx.ctypes.data_as(dp), out.ctypes.data_as(dp), None)

    assert tensor(21, 21, 0.5) == _capi.PQP_ERR_RANGE
    assert tensor(-1, 2, 0.5) == _capi.PQP_ERR_INVALID
    assert tensor(2, 1, 0.5) == _capi.PQP_ERR_INVALID
    assert tensor(2, 2, float("nan")) == _capi.PQP_ERR_INVALID
    with pytest.raises(ValueError):
        pa.tensor_permanent(np.zeros((2, 3)), 0.5)


def test_integer_entry_errors_precede_the_device():
    """pqp_permanent_int64 checks its arguments and the exact path's range
    before it needs a device; an input it accepts reaches the device check."""
    import piquasso_amd as pa

    L = _capi.lib()
    i32p = ctypes.POINTER(ctypes.c_int32)
    i64p = ctypes.POINTER(ctypes.c_int64)
    dp = ctypes.POINTER(ctypes.c_double)
    u64p = ctypes.POINTER(ctypes.c_uint64)

    def call(a, rows, cols, out=True, n_limbs=8):
        a = np.ascontiguousarray(a, dtype=np.int64)
        r = np.ascontiguousarray(rows, dtype=np.int32)
        c = np.ascontiguousarray(cols, dtype=np.int32)
        o = np.zeros(1)
        limbs = np.zeros(8, dtype=np.uint64)
        return L.pqp_permanent_int64(a.ctypes.data_as(i64p), a.shape[0], a.shape[1], a.shape[1],
                                     r.ctypes.data_as(i32p), c.ctypes.data_as(i32p),
                                     o.ctypes.data_as(dp) if out else None,
                                     limbs.ctypes.data_as(u64p), n_limbs)

    one2 = [1, 1]
    assert call(np.eye(2), one2, one2, out=False) == _capi.PQP_ERR_INVALID
    assert call(np.eye(2), one2, one2, n_limbs=4) == _capi.PQP_ERR_INVALID
    assert call(np.eye(2), [1, -1], [0, 0]) == _capi.PQP_ERR_INVALID
    assert call(np.eye(2), one2, [2, 1]) == _capi.PQP_ERR_SUM_MISMATCH
    big = np.ones((2, 2), dtype=np.int64)
    big[0, 0] = 2**48
    assert call(big, one2, one2) == _capi.PQP_ERR_RANGE          # a selected entry > 2^47
    assert "2^47" in L.pqp_last_error().decode()
    assert call(np.full((3, 1), 2**47), [1, 1, 1], [3]) == _capi.PQP_ERR_RANGE  # column bound > 2^48
    assert call(np.ones((1, 1)), [60], [60]) == _capi.PQP_ERR_RANGE  # C(60, 30) > 2^52
    assert "binomial" in L.pqp_last_error().decode()
    # |perm| bound (20 * 2^20)^20 = 2^486
    assert call(np.full((20, 20), 2**20), np.ones(20), np.ones(20)) == _capi.PQP_ERR_RANGE
    assert "2^390" in L.pqp_last_error().decode()
    n_dev = ctypes.c_int(0)
    if L.pqp_device_info(ctypes.byref(n_dev), None, None, 0) != _capi.PQP_OK:
        # unselected large entries are never read; the rest is the device's
        assert call(big, [0, 1], [0, 1]) == _capi.PQP_ERR_DEVICE
    with pytest.raises(ValueError, match="not integers"):
        pa.permanent_exact(np.array([[0.5]]), [1], [1])
    with pytest.raises(ValueError, match="imaginary"):
        pa.permanent_exact(np.array([[1j]]), [1], [1])
    with pytest.raises(RuntimeError):
        pa.permanent_exact(np.eye(2, dtype=int), [1, 1], [2, 1])


def test_fast_binding_of_the_drop_in_call():
    """The CPython binding of pqp_permanent_c128 (csrc/fastcall.c) is built
    beside the library and loaded by permanent(); it reports C-ABI failures
    as their status codes and refuses other argument layouts with TypeError
    (permanent() then takes its general path)."""
    from piquasso_amd._math import permanent as PM

    fast = PM._fast_module()
    assert fast, "lib/_pqp_fast*.so is not built"
    a = np.eye(2, dtype=np.complex128)
    one = np.ones(2, dtype=np.int64)
    assert fast.permanent_c128(a, one, np.array([2, 1], dtype=np.int32)) == _capi.PQP_ERR_SUM_MISMATCH
    for bad in ((a.astype(np.complex64), one, one), (a, one.astype(np.float64), one),
                (a, np.ones(3, dtype=np.int64), one), (np.ones(3, dtype=np.complex128), one, one)):
        with pytest.raises(TypeError):
            fast.permanent_c128(*bad)
    with pytest.raises(ValueError):  # numpy's buffer export refuses a non-contiguous view
        fast.permanent_c128(np.eye(3, dtype=np.complex128)[:, :2], np.ones(3, np.int64), one)


def test_empty_selection_needs_no_device():
    """perm of an empty selection is 1 (src/permanent.cpp:106-108); the
    Laplace entry returns the reference's [1] for degenerate inputs."""
    import piquasso_amd as pa

    assert pa.permanent_laplace(np.ones((2, 2)), [0, 0], [0, 1]).tolist() == [1]
    assert pa.permanent_laplace(np.zeros((0, 3)), [], [1, 0, 0]).tolist() == [1]
    assert pa.permanent_exact(np.zeros((0, 0), dtype=int), [], []) == 1
    assert pa.permanent_exact(np.ones((2, 2), dtype=int), [0, 0], [0, 0]) == 1


def test_no_kernel_spills_to_scratch():
    """Every sweep kernel keeps its state in registers: the build records the
    compiler's resource usage per kernel object (piquasso_amd/csrc/hipcc_kernel.sh);
    a scratch spill in the hot loop costs several times the kernel's speed."""
    import glob

    res = sorted(glob.glob(os.path.join(ROOT, "piquasso_amd", "lib", "obj", "*.res")))
    if not res:
        pytest.skip("no resource records (library not built by piquasso_amd/csrc/Makefile)")
    spills = []
    for path in res:
        name = None
        for line in open(path):
            if line.startswith("Function Name:"):
                name = line.split(":", 1)[1].strip()
            elif line.startswith("ScratchSize") and int(line.split(":")[-1]) != 0:
                spills.append((os.path.basename(path), name))
    assert not spills, spills


def test_poly_entry_argument_errors_without_device():
    """pqp_perm_poly_coefficients_c128 validates before it needs a device:
    order beyond 24, negative exponents, null pointers; rows whose total
    degree exceeds n are answered (0) without any device work."""
    import ctypes

    from piquasso_amd import _capi

    L = _capi.lib()
    dp, i32p, i64p = (ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32),
                      ctypes.POINTER(ctypes.c_int64))
    b = np.zeros((2, 2), dtype=np.complex128)
    bm = np.zeros((1, 2, 2), dtype=np.complex128)
    out = np.full(4, 7.0)
    info = np.zeros(6, dtype=np.int64)

    def call(n, exps, b0=b, bms=bm, o=out):
        e = np.ascontiguousarray(exps, dtype=np.int32)
        return L.pqp_perm_poly_coefficients_c128(
            n, e.shape[1], b0.view(np.float64).ctypes.data_as(dp) if b0 is not None else None,
            bms.view(np.float64).ctypes.data_as(dp), e.shape[0], e.ctypes.data_as(i32p),
            o.ctypes.data_as(dp) if o is not None else None, info.ctypes.data_as(i64p))

    assert call(25, [[1]]) == _capi.PQP_ERR_RANGE
    assert call(2, [[-1]]) == _capi.PQP_ERR_INVALID
    assert call(2, [[1]], b0=None) == _capi.PQP_ERR_INVALID
    assert call(2, [[1]], o=None) == _capi.PQP_ERR_INVALID
    assert call(2, [[3], [4]]) == _capi.PQP_OK  # degree > n: zero, no device needed
    assert list(out) == [0.0, 0.0, 0.0, 0.0]
    assert info[5] == 3 and info[0] == 0
    # the same sizes again (the plan of the previous call is cached) with a
    # null exponent pointer: rejected before the cache is consulted
    st = L.pqp_perm_poly_coefficients_c128(2, 1, b.view(np.float64).ctypes.data_as(dp),
                                           bm.view(np.float64).ctypes.data_as(dp), 2, None,
                                           out.ctypes.data_as(dp), info.ctypes.data_as(i64p))
    assert st == _capi.PQP_ERR_INVALID


def test_noalloc_entry_errors_without_device():
    """pqp_permanent_c128_noalloc checks its arguments and its scratch limit
    before it needs a device: a null output, a sum mismatch, and a permanent
    of 2^40 Gray terms (n = 41), whose plan exceeds the 8 MiB device array."""
    L = _capi.lib()
    i32p = ctypes.POINTER(ctypes.c_int32)
    dp = ctypes.POINTER(ctypes.c_double)

    def call(a, rows, cols, out=True):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        r = np.ascontiguousarray(rows, dtype=np.int32)
        c = np.ascontiguousarray(cols, dtype=np.int32)
        o = np.zeros(2)
        return L.pqp_permanent_c128_noalloc(a.view(np.float64).ctypes.data_as(dp), a.shape[0],
                                            a.shape[1], a.shape[1], r.ctypes.data_as(i32p),
                                            c.ctypes.data_as(i32p), 0,
                                            o.ctypes.data_as(dp) if out else None, None)

    assert call(np.eye(2), [1, 1], [1, 1], out=False) == _capi.PQP_ERR_INVALID
    assert call(np.eye(2), [1, 1], [2, 1]) == _capi.PQP_ERR_SUM_MISMATCH
    assert call(np.zeros((41, 41)), np.ones(41), np.ones(41)) == _capi.PQP_ERR_RANGE
    msg = L.pqp_last_error().decode()
    assert "bytes of device scratch" in msg and "8388608" in msg


_NO_DEVICE_CHILD = r"""
import ctypes, numpy as np
from piquasso_amd import _capi
import piquasso_amd as pa
L = _capi.lib()
n = ctypes.c_int(-1)
print("status", L.pqp_device_info(ctypes.byref(n), None, None, 0), n.value)
print("message", L.pqp_last_error().decode())
try:
    pa.permanent(np.eye(2, dtype=complex), np.array([1, 1]), np.array([1, 1]))
    print("computed")
except RuntimeError as e:
    print("raised", e)
"""


def test_forced_no_device_switch():
    """PQP_FORCE_NO_DEVICE makes every call fail with PQP_ERR_DEVICE, named
    as forced, before the HIP runtime is touched (a fresh process: the
    switch is read once)."""
    import subprocess
    import sys

    env = dict(os.environ, PQP_FORCE_NO_DEVICE="1")
    res = subprocess.run([sys.executable, "-c", _NO_DEVICE_CHILD], cwd=ROOT, env=env,
                         capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr
    lines = res.stdout.splitlines()
    assert lines[0] == f"status {_capi.PQP_ERR_DEVICE} 0"
    assert "PQP_FORCE_NO_DEVICE" in lines[1]
    assert lines[2].startswith("raised") and "PQP_FORCE_NO_DEVICE" in lines[2]


def test_allocation_hook_arms_and_disarms():
    L = _capi.lib()
    assert L.pqp_debug_fail_allocation(3) == 0
    assert L.pqp_debug_fail_allocation(-1) == 0
