"""ctypes binding of ``libpqp_permanent.so`` (the C ABI in include/pqp_permanent.h).

The shared library is built in-tree (``piquasso_amd/lib``) by
``__graft_entry__.build()`` / ``make -C piquasso_amd/csrc``.  There is no
fallback: if the library is missing, :func:`lib` raises, and if no HIP
device is usable every compute entry point returns ``PQP_ERR_DEVICE``, which
the Python layer turns into ``RuntimeError``.
"""

from __future__ import annotations

import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libpqp_permanent.so")

PQP_OK = 0
PQP_ERR_SUM_MISMATCH = -1
PQP_ERR_INVALID = -2
PQP_ERR_RANGE = -3
PQP_ERR_DEVICE = -4
PQP_ERR_NOMEM = -5
PQP_ERR_NOKERNEL = -6

PQP_FLAG_PLAIN = 0x1
PQP_FLAG_FP32 = 0x2
PQP_FLAG_STATIC = 0x4
PQP_FLAG_EXPAND_COLS = 0x8
PQP_FLAG_KEEP_COLS = 0x10
PQP_FLAG_GRADIENT = 0x20
PQP_FLAG_AS_GIVEN = 0x80
PQP_PLAN_GRADIENT_SWEEP = 0x40

PQP_RUN_EAGER = 0
PQP_RUN_GRAPH = 1
PQP_RUN_PLAIN = 2
PQP_RUN_KERNEL = 3

# Every symbol declared in include/pqp_permanent.h, with its ctypes signature.
_c = ctypes
_dp = _c.POINTER(_c.c_double)
_fp = _c.POINTER(_c.c_float)
_i32p = _c.POINTER(_c.c_int32)
_i64p = _c.POINTER(_c.c_int64)
_vp = _c.c_void_p

SIGNATURES = {
    "pqp_version": (_c.c_char_p, []),
    "pqp_last_error": (_c.c_char_p, []),
    "pqp_last_error_bound": (_c.c_double, []),
    "pqp_device_info": (_c.c_int, [_c.POINTER(_c.c_int), _c.POINTER(_c.c_int), _c.c_char_p, _c.c_int]),
    "pqp_debug_fail_allocation": (_c.c_int, [_c.c_int64]),
    "pqp_permanent_c128": (
        _c.c_int,
        [_dp, _c.c_int64, _c.c_int64, _c.c_int64, _i32p, _i32p, _c.c_uint32, _dp, _dp],
    ),
    "pqp_permanent_c128_noalloc": (
        _c.c_int,
        [_dp, _c.c_int64, _c.c_int64, _c.c_int64, _i32p, _i32p, _c.c_uint32, _dp, _dp],
    ),
    "pqp_permanent_c64": (
        _c.c_int,
        [_fp, _c.c_int64, _c.c_int64, _c.c_int64, _i32p, _i32p, _c.c_uint32, _fp, _dp],
    ),
    "pqp_permanent_int64": (
        _c.c_int,
        [_i64p, _c.c_int64, _c.c_int64, _c.c_int64, _i32p, _i32p, _dp,
         _c.POINTER(_c.c_uint64), _c.c_int32],
    ),
    "pqp_batch_create": (
        _c.c_int,
        [
            _c.c_int64, _c.POINTER(_vp), _c.c_int, _i64p, _i64p, _i64p,
            _c.POINTER(_i32p), _c.POINTER(_i32p), _c.c_uint32, _c.c_int64, _c.c_int64,
            _c.POINTER(_vp),
        ],
    ),
    "pqp_batch_run": (_c.c_int, [_vp, _c.c_int]),
    "pqp_batch_sync": (_c.c_int, [_vp]),
    "pqp_batch_results": (_c.c_int, [_vp, _dp, _dp]),
    "pqp_batch_error_bounds": (_c.c_int, [_vp, _dp]),
    "pqp_batch_time": (_c.c_int, [_vp, _c.c_int, _c.c_int, _fp]),
    "pqp_batch_time_span": (_c.c_int, [_vp, _c.c_int, _c.c_int, _fp]),
    "pqp_batch_stats": (_c.c_int, [_vp, _dp, _i64p, _i64p, _i64p, _i64p]),
    "pqp_batch_destroy": (None, [_vp]),
    "pqp_batch_gradients": (_c.c_int, [_vp, _c.POINTER(_dp)]),
    "pqp_laplace_c128": (
        _c.c_int, [_dp, _c.c_int64, _c.c_int64, _c.c_int64, _i32p, _i32p, _dp],
    ),
    "pqp_laplace_batch_c128": (
        _c.c_int,
        [_c.c_int64, _c.POINTER(_dp), _i64p, _i64p, _i64p, _c.POINTER(_i32p),
         _c.POINTER(_i32p), _c.POINTER(_dp)],
    ),
    "pqp_tensor_permanent_c128": (
        _c.c_int, [_dp, _c.c_int64, _c.c_int64, _c.c_double, _dp, _dp],
    ),
    "pqp_tensor_permanent_batch_c128": (
        _c.c_int,
        [_c.c_int64, _c.POINTER(_dp), _i64p, _i64p, _dp, _dp, _dp],
    ),
    "pqp_perm_poly_plan_info": (
        _c.c_int, [_c.c_int64, _c.c_int64, _c.c_int64, _i32p, _i64p],
    ),
    "pqp_perm_poly_coefficients_c128": (
        _c.c_int,
        [_c.c_int64, _c.c_int64, _dp, _dp, _c.c_int64, _i32p, _dp, _i64p],
    ),
    "pqp_grad_batch_c128": (
        _c.c_int,
        [_c.c_int64, _c.POINTER(_dp), _i64p, _i64p, _i64p, _c.POINTER(_i32p),
         _c.POINTER(_i32p), _dp, _c.POINTER(_dp)],
    ),
    "pqp_plan_image": (
        _c.c_int,
        [_dp, _c.c_int64, _c.c_int64, _c.c_int64, _i32p, _i32p, _c.c_uint32,
         _c.c_int64, _c.c_int64, _vp, _c.c_int64, _i64p],
    ),
}

_lib = None
_lock = threading.Lock()


class PermanentLibraryMissing(RuntimeError):
    """The HIP permanent library has not been built."""


def lib():
    """Load the library once; raise loudly when it is not built."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise PermanentLibraryMissing(
                        f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; "
                        "g.build()'` or `make -C piquasso_amd/csrc`"
                    )
                L = ctypes.CDLL(LIB_PATH)
                for name, (res, args) in SIGNATURES.items():
                    fn = getattr(L, name)
                    fn.restype = res
                    fn.argtypes = args
                _lib = L
    return _lib


def last_error() -> str:
    msg = lib().pqp_last_error()
    return msg.decode() if msg else ""


def check(status: int) -> None:
    """Map a C-ABI status to the exception the reference raises (or the
    closest Python one where the reference has undefined behaviour)."""
    if status == PQP_OK:
        return
    msg = last_error()
    if status == PQP_ERR_SUM_MISMATCH:
        # the reference throws a std::string here, surfaced as RuntimeError
        raise RuntimeError(msg or "Number of input and output states should be equal")
    if status == PQP_ERR_INVALID:
        raise ValueError(msg)
    if status == PQP_ERR_RANGE:
        raise ValueError(msg)
    if status == PQP_ERR_NOMEM:
        raise MemoryError(msg)
    raise RuntimeError(f"permanent engine error {status}: {msg}")


def device_info():
    """(n_devices, cu_count, arch name); raises RuntimeError without a device."""
    n = ctypes.c_int(0)
    cu = ctypes.c_int(0)
    name = ctypes.create_string_buffer(64)
    check(lib().pqp_device_info(ctypes.byref(n), ctypes.byref(cu), name, 64))
    return n.value, cu.value, name.value.decode()
