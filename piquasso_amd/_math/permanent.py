"""Drop-in replacement for piquasso's native module ``piquasso._math.permanent``.

``permanent(matrix, rows, cols)`` mirrors the pybind entry of the reference
(piquasso/_math/permanent.cpp:26-41,71-78): the same keyword names, the
same dtype dispatch, a 0-d ``numpy.ndarray`` result (src/numpy_utils.hpp:36-49),
1 for empty / all-zero multiplicities (src/permanent.cpp:106-108) and
``RuntimeError`` when the row and column multiplicities do not sum to the same
number (src/permanent.cpp:100-104).  The value is computed by the HIP kernels
of ``libpqp_permanent.so``; there is no CPU path.

Dtype dispatch, as the reference module built with pybind11 3.0 / numpy 2
resolves its two overloads (``complex<float>`` registered before
``complex<double>``, both ``c_style``, no forcecast):

* a numpy array whose dtype casts *safely* to complex64 (complex64, float32,
  float16, int8/16, uint8, bool) takes the complex64 path and returns
  complex64;
* otherwise a numpy array that casts safely to complex128 (complex128,
  float64, int32/64, ...) takes the complex128 path, also when it is not
  C-contiguous (it is copied);
* any other array dtype is a ``TypeError``;
* non-array inputs (lists, tuples) are converted to complex64.

Multiplicities are cast to int32 like ``py::array_t<int, forcecast>``.

Deliberate differences, where the reference has undefined behaviour: a
``rows``/``cols`` length that does not match the matrix shape, a non-2-D
matrix, or a negative multiplicity raise ``ValueError`` instead of reading
out of bounds.

Integer matrices: a matrix whose selected entries are integers with zero
imaginary parts and whose Glynn terms may leave the floating-point sweep's
exact integers is evaluated by the exact multi-modular kernel instead, so
the result is the exact permanent rounded once (perfect-matching counts at
any size the engine takes).  ``permanent_exact`` returns the integer itself.

``permanent_laplace(matrix, rows, cols)`` mirrors the second pybind entry
(piquasso/_math/permanent.cpp:43-58,79-84; src/permanent_laplace.cpp:43-240):
element l of the result is perm(A; rows, cols - e_l), for a column with
cols[l] = 0 the uncut product sum (what the reference's loop produces there),
and ``[1]`` of length one when the matrix is empty or either multiplicity sum
is 0 (src/permanent_laplace.cpp:52-57).  Same dtype dispatch as
``permanent``; the complex64 path is evaluated with FP64 products and
rounded to complex64 at the end (the reference runs it in FP32).  The
reference does not check ``sum(cols) == sum(rows) + 1`` and then returns a
number that is no permanent; here that raises ``RuntimeError``.
"""

from __future__ import annotations

import ctypes
import importlib.util
import os
import sysconfig
import threading

import numpy as np

from .. import _capi

__all__ = ["permanent", "permanent_exact", "permanent_laplace", "permanent_laplace_batch",
           "resolve_dtype"]


def resolve_dtype(matrix):
    """The complex dtype the reference's pybind overload resolution picks."""
    if isinstance(matrix, np.ndarray):
        if np.can_cast(matrix.dtype, np.complex64, "safe"):
            return np.dtype(np.complex64)
        if np.can_cast(matrix.dtype, np.complex128, "safe"):
            return np.dtype(np.complex128)
        raise TypeError(
            "permanent(): incompatible function arguments; matrix dtype "
            f"{matrix.dtype} converts to neither complex64 nor complex128"
        )
    return np.dtype(np.complex64)


def _multiplicities(values, name, length):
    arr = np.ascontiguousarray(np.asarray(values).astype(np.int32, casting="unsafe").reshape(-1))
    if arr.shape[0] != length:
        raise ValueError(
            f"permanent(): {name} has {arr.shape[0]} entries but the matrix has {length}"
        )
    return arr


_c128_call = None
_fast = None
_tls = threading.local()


def _fast_module():
    """The CPython binding of pqp_permanent_c128 (lib/_pqp_fast*.so, built
    beside the library from csrc/fastcall.c), or False when it is absent;
    the ctypes binding then serves every call."""
    global _fast
    if _fast is None:
        _capi.lib()  # the library itself first: its absence is an error
        path = os.path.join(os.path.dirname(_capi.LIB_PATH),
                            "_pqp_fast" + sysconfig.get_config_var("EXT_SUFFIX"))
        mod = False
        if os.path.exists(path):
            spec = importlib.util.spec_from_file_location("_pqp_fast", path)
            mod = importlib.util.module_from_spec(spec)
            spec.loader.exec_module(mod)
        _fast = mod
    return _fast


def _fast_c128():
    """pqp_permanent_c128 bound with plain address arguments (cheaper to
    call than the pointer-typed binding)."""
    global _c128_call
    if _c128_call is None:
        c = ctypes
        proto = c.CFUNCTYPE(c.c_int, c.c_void_p, c.c_int64, c.c_int64, c.c_int64, c.c_void_p,
                            c.c_void_p, c.c_uint32, c.c_void_p, c.c_void_p)
        _c128_call = proto(("pqp_permanent_c128", _capi.lib()))
    return _c128_call


def _out_buffer():
    buf = getattr(_tls, "out", None)
    if buf is None:
        buf = np.zeros(2)
        _tls.out = buf
        _tls.out_addr = buf.ctypes.data
    return buf, _tls.out_addr


def permanent(matrix, rows, cols):
    """Permanent of ``matrix`` with row multiplicities ``rows`` and column
    multiplicities ``cols`` (Glynn's formula over the reflected n-ary Gray
    code, src/permanent.cpp:49-264), evaluated on the GPU."""
    if type(matrix) is np.ndarray:
        # the common call, a C-contiguous 2-D complex128 matrix with int32 or
        # int64 multiplicity arrays, through the CPython binding (which
        # refuses every other layout with TypeError / ValueError)
        fast = _fast if _fast is not None else _fast_module()
        if fast:
            try:
                v = fast.permanent_c128(matrix, rows, cols)
            except (TypeError, BufferError, ValueError):
                v = None  # another layout: the paths below
            if v is not None:
                if type(v) is int:
                    _capi.check(v)
                return np.array(v, dtype=np.complex128)
    if (type(matrix) is np.ndarray and matrix.dtype == np.complex128 and matrix.ndim == 2
            and matrix.flags.c_contiguous and type(rows) is np.ndarray and type(cols) is np.ndarray
            and rows.ndim == 1 and cols.ndim == 1 and rows.dtype.kind in "iu" and cols.dtype.kind in "iu"):
        # the same call without the binding: addresses through ctypes
        m_r, m_c = matrix.shape
        if rows.shape[0] != m_r or cols.shape[0] != m_c:
            raise ValueError(f"permanent(): rows / cols have {rows.shape[0]} / {cols.shape[0]} "
                             f"entries for a {m_r} x {m_c} matrix")
        r = rows if (rows.dtype == np.int32 and rows.flags.c_contiguous) else rows.astype(np.int32)
        c = cols if (cols.dtype == np.int32 and cols.flags.c_contiguous) else cols.astype(np.int32)
        out, out_addr = _out_buffer()
        st = _fast_c128()(matrix.ctypes.data, m_r, m_c, m_c, r.ctypes.data, c.ctypes.data, 0,
                          out_addr, None)
        if st:
            _capi.check(st)
        return np.array(complex(out[0], out[1]), dtype=np.complex128)
    dtype = resolve_dtype(matrix)
    a = np.ascontiguousarray(np.asarray(matrix, dtype=dtype))
    if a.ndim != 2:
        raise ValueError(f"permanent(): matrix must be 2-D, got shape {a.shape}")
    m_r, m_c = a.shape
    r = _multiplicities(rows, "rows", m_r)
    c = _multiplicities(cols, "cols", m_c)
    L = _capi.lib()
    i32p = ctypes.POINTER(ctypes.c_int32)
    if dtype == np.complex128:
        out = np.zeros(2, dtype=np.float64)
        st = L.pqp_permanent_c128(
            a.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            m_r, m_c, m_c, r.ctypes.data_as(i32p), c.ctypes.data_as(i32p), 0,
            out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), None,
        )
        _capi.check(st)
        return np.array(complex(out[0], out[1]), dtype=np.complex128)
    out = np.zeros(2, dtype=np.float32)
    st = L.pqp_permanent_c64(
        a.view(np.float32).ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
        m_r, m_c, m_c, r.ctypes.data_as(i32p), c.ctypes.data_as(i32p), 0,
        out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), None,
    )
    _capi.check(st)
    return np.array(complex(out[0], out[1]), dtype=np.complex64)


def permanent_exact(matrix, rows, cols) -> int:
    """The exact permanent of an integer matrix with multiplicities, as a
    Python ``int`` (``pqp_permanent_int64``: the Glynn sum of
    src/permanent.cpp:49-264 modulo up to eight primes on the GPU, the
    integer recovered by the Chinese remainder theorem).  ``matrix`` may be
    of any integer or bool dtype, or float / complex with integral values and
    zero imaginary parts.  ``ValueError`` when an entry is not such an
    integer, or outside the exact path's range (selected entries beyond
    2^47, column bounds sum_i r_i |a_ij| beyond 2^48, binomial weights
    prod_i C(r_i, r_i/2) beyond 2^52, a permanent that may exceed 2^390)."""
    a = np.asarray(matrix)
    if a.ndim != 2:
        raise ValueError(f"permanent_exact(): matrix must be 2-D, got shape {a.shape}")
    if a.dtype.kind in "biu":
        if a.dtype.kind == "u" and a.size and int(a.max()) > 2**63 - 1:
            raise ValueError("permanent_exact(): entries beyond 2^47 in magnitude")
        ai = np.ascontiguousarray(a.astype(np.int64))
    elif a.dtype.kind in "fc":
        re = np.real(a)
        if a.dtype.kind == "c" and np.any(np.imag(a) != 0):
            raise ValueError("permanent_exact(): entries with non-zero imaginary parts")
        if not np.all(np.isfinite(re)) or np.any(re != np.rint(re)):
            raise ValueError("permanent_exact(): entries that are not integers")
        if re.size and float(np.max(np.abs(re))) > 2.0**62:
            raise ValueError("permanent_exact(): entries beyond 2^47 in magnitude")
        ai = np.ascontiguousarray(re.astype(np.int64))
    else:
        raise TypeError(f"permanent_exact(): unsupported matrix dtype {a.dtype}")
    m_r, m_c = ai.shape
    r = _multiplicities(rows, "rows", m_r)
    c = _multiplicities(cols, "cols", m_c)
    out = np.zeros(1, dtype=np.float64)
    limbs = np.zeros(8, dtype=np.uint64)
    st = _capi.lib().pqp_permanent_int64(
        ai.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), m_r, m_c, m_c,
        r.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
        c.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
        out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
        limbs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), 8,
    )
    _capi.check(st)
    v = 0
    for k in range(7, -1, -1):
        v = (v << 64) | int(limbs[k])
    return v - (1 << 512) if v >> 511 else v


def permanent_laplace(matrix, rows, cols):
    """All column-minor permanents perm(A; rows, cols - e_l), l = 0..m_c-1,
    from one Gray sweep per work unit on the GPU (prefix / suffix products,
    Lemma 1 of arXiv:2005.04214).  Returns a 1-D ``numpy.ndarray``."""
    dtype = resolve_dtype(matrix)
    a = np.ascontiguousarray(np.asarray(matrix, dtype=np.complex128))
    if a.ndim != 2:
        raise ValueError(f"permanent_laplace(): matrix must be 2-D, got shape {a.shape}")
    m_r, m_c = a.shape
    r = _multiplicities(rows, "rows", m_r)
    c = _multiplicities(cols, "cols", m_c)
    if m_r == 0 or m_c == 0 or int(r.sum()) == 0 or int(c.sum()) == 0:
        return np.ones(1, dtype=dtype)
    out = np.zeros(2 * m_c, dtype=np.float64)
    i32p = ctypes.POINTER(ctypes.c_int32)
    dp = ctypes.POINTER(ctypes.c_double)
    st = _capi.lib().pqp_laplace_c128(
        a.view(np.float64).ctypes.data_as(dp), m_r, m_c, m_c,
        r.ctypes.data_as(i32p), c.ctypes.data_as(i32p), out.ctypes.data_as(dp),
    )
    _capi.check(st)
    return (out[0::2] + 1j * out[1::2]).astype(dtype)


def permanent_laplace_batch(matrices, rows, cols):
    """``[permanent_laplace(m, r, c) for m, r, c in zip(matrices, rows, cols)]``
    evaluated in one launch (``pqp_laplace_batch_c128``): every problem's
    column minors come from the same batched sweep.  Each entry of the
    returned list is a complex128 array with the value and length
    ``permanent_laplace`` gives (``[1]`` for a degenerate problem).  Inputs
    are computed in FP64 whatever their dtype."""
    mats, rs, cs = list(matrices), list(rows), list(cols)
    if not (len(mats) == len(rs) == len(cs)):
        raise ValueError("permanent_laplace_batch(): matrices, rows and cols differ in length")
    out = [None] * len(mats)
    keep_a, keep_r, keep_c, keep_o, idx = [], [], [], [], []
    for k, (m, r, c) in enumerate(zip(mats, rs, cs)):
        a = np.ascontiguousarray(np.asarray(m, dtype=np.complex128))
        if a.ndim != 2:
            raise ValueError(f"permanent_laplace_batch(): item {k} is not 2-D (shape {a.shape})")
        m_r, m_c = a.shape
        rr = _multiplicities(r, "rows", m_r)
        cc = _multiplicities(c, "cols", m_c)
        if m_r == 0 or m_c == 0 or int(rr.sum()) == 0 or int(cc.sum()) == 0:
            out[k] = np.ones(1, dtype=np.complex128)
            continue
        keep_a.append(a)
        keep_r.append(rr)
        keep_c.append(cc)
        keep_o.append(np.zeros(2 * m_c, dtype=np.float64))
        idx.append(k)
    n = len(idx)
    if n:
        dp = ctypes.POINTER(ctypes.c_double)
        i32p = ctypes.POINTER(ctypes.c_int32)
        mats_p = (dp * n)(*[a.view(np.float64).ctypes.data_as(dp) for a in keep_a])
        rows_p = (i32p * n)(*[r.ctypes.data_as(i32p) for r in keep_r])
        cols_p = (i32p * n)(*[c.ctypes.data_as(i32p) for c in keep_c])
        outs_p = (dp * n)(*[o.ctypes.data_as(dp) for o in keep_o])
        m_r = np.array([a.shape[0] for a in keep_a], dtype=np.int64)
        m_c = np.array([a.shape[1] for a in keep_a], dtype=np.int64)
        i64p = ctypes.POINTER(ctypes.c_int64)
        st = _capi.lib().pqp_laplace_batch_c128(
            n, mats_p, m_r.ctypes.data_as(i64p), m_c.ctypes.data_as(i64p),
            m_c.ctypes.data_as(i64p), rows_p, cols_p, outs_p,
        )
        _capi.check(st)
        for k, o in zip(idx, keep_o):
            out[k] = o[0::2] + 1j * o[1::2]
    return out
