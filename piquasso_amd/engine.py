"""Python face of the permanent engine beyond the drop-in ``permanent()``.

* :func:`permanent_ex` — one permanent with an explicit precision mode and
  the double-double parts of the result;
* :class:`PermanentBatch` / :func:`permanent_batch` — many permanents of
  uneven size in one launch per column-width bucket, the buckets running
  concurrently (optionally replayed from a hipGraph).  The reference has no
  batched entry; every item's value equals ``permanent(matrix, rows, cols)``
  on that item.

All arithmetic runs in ``libpqp_permanent.so`` on the GPU.
"""

from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np

from . import _capi

__all__ = ["permanent_ex", "PermanentBatch", "permanent_batch", "MODES"]

# precision modes of a single evaluation
MODES = {
    "dd": 0,                      # FP64 row sums/products, double-double accumulation
    "plain": _capi.PQP_FLAG_PLAIN,  # FP64 everywhere, one plain launch (baseline)
    "fp32": _capi.PQP_FLAG_FP32,    # FP32 row sums/products, dd accumulation
}

_dp = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)


def _prep(matrix, rows, cols, dtype=np.complex128):
    a = np.ascontiguousarray(np.asarray(matrix, dtype=dtype))
    if a.ndim != 2:
        raise ValueError(f"matrix must be 2-D, got shape {a.shape}")
    r = np.ascontiguousarray(np.asarray(rows).astype(np.int32, casting="unsafe").reshape(-1))
    c = np.ascontiguousarray(np.asarray(cols).astype(np.int32, casting="unsafe").reshape(-1))
    if r.shape[0] != a.shape[0] or c.shape[0] != a.shape[1]:
        raise ValueError(
            f"rows/cols lengths {r.shape[0]}/{c.shape[0]} do not match matrix shape {a.shape}"
        )
    return a, r, c


def permanent_ex(matrix, rows, cols, mode: str = "dd", with_bound: bool = False):
    """Return ``(value, dd)`` for one permanent of a complex128 matrix:
    ``value`` the complex result, ``dd`` = [re_hi, re_lo, im_hi, im_lo];
    with ``with_bound`` also a bound on |value - permanent|
    (``pqp_last_error_bound``: the sweep's 4 (n + 2) u sum|terms| / 2^(n-1)
    plus the final rounding; NaN in mode "plain")."""
    if mode not in MODES:
        raise ValueError(f"mode must be one of {sorted(MODES)}")
    a, r, c = _prep(matrix, rows, cols)
    out = np.zeros(2)
    dd = np.zeros(4)
    st = _capi.lib().pqp_permanent_c128(
        a.view(np.float64).ctypes.data_as(_dp), a.shape[0], a.shape[1], a.shape[1],
        r.ctypes.data_as(_i32p), c.ctypes.data_as(_i32p), MODES[mode],
        out.ctypes.data_as(_dp), dd.ctypes.data_as(_dp),
    )
    _capi.check(st)
    if with_bound:
        return complex(out[0], out[1]), dd, float(_capi.lib().pqp_last_error_bound())
    return complex(out[0], out[1]), dd


def _shared_marshal(matrices, rows, cols):
    """Arguments of pqp_batch_create for a batch whose items all use the same
    matrix object and whose multiplicities stack into (count, m_r) and
    (count, m_c) integer arrays; None when the batch is not of that form."""
    n = len(matrices)
    if n < 2 or any(m is not matrices[0] for m in matrices):
        return None
    try:
        a = _prep(matrices[0], np.zeros(np.shape(matrices[0])[0], np.int32),
                  np.zeros(np.shape(matrices[0])[1], np.int32))[0]
        R = np.asarray(rows)
        C = np.asarray(cols)
    except (ValueError, TypeError, IndexError):
        return None
    m_r, m_c = a.shape
    if R.shape != (n, m_r) or C.shape != (n, m_c) or R.dtype.kind not in "iu" or C.dtype.kind not in "iu":
        return None
    R32 = np.ascontiguousarray(R, dtype=np.int32)
    C32 = np.ascontiguousarray(C, dtype=np.int32)
    idx = np.arange(n, dtype=np.uint64)
    mats = np.full(n, a.ctypes.data, dtype=np.uint64)
    rp = np.uint64(R32.ctypes.data) + idx * np.uint64(4 * m_r)
    cp = np.uint64(C32.ctypes.data) + idx * np.uint64(4 * m_c)
    keep = [a, R32, C32, mats, rp, cp]
    mat_arr = mats.ctypes.data_as(ctypes.POINTER(ctypes.c_void_p))
    row_arr = rp.ctypes.data_as(ctypes.POINTER(_i32p))
    col_arr = cp.ctypes.data_as(ctypes.POINTER(_i32p))
    return (keep, mat_arr, row_arr, col_arr, np.full(n, m_r, dtype=np.int64),
            np.full(n, m_c, dtype=np.int64))


class PermanentBatch:
    """A planned batch of permanents resident on the GPU.

    Parameters
    ----------
    matrices, rows, cols:
        sequences of equal length: matrices (complex128, or complex64 which
        selects FP32 products) and their row / column multiplicities.
    precision:
        "fp64" (FP64 row sums and products, double-double accumulation) or
        "fp32" (FP32 row sums and products, double-double accumulation).
    persistent:
        True: one persistent work-stealing launch per column-width bucket;
        False: one block per work unit, units in decreasing cost order;
        None (default): the autotuned choice for this device
        (:func:`piquasso_amd.tune.best_config`), else False.
    target_inner:
        inner Gray states per lane (at most; the largest Gray sub-space of
        the item that fits).  0: the autotuned value for this device when
        total_units is not given either (:mod:`piquasso_amd.tune`), else
        the planner's cost model chooses per item.
    expand_cols:
        None (planner's choice, currently keep), True (columns of multiplicity
        c > 1 become c identical columns) or False (keep one column whose row
        sum is raised to the power c).
    total_units:
        wave work units spread over the batch in proportion to each
        permanent's cost (0: the autotuned count for this device, else 32
        per compute unit).
    gradient:
        also evaluate d perm / d A of every item on each run (column-minor
        sweeps after the permanents, on the same stream and in the same
        graph); read them with :meth:`gradients`.  complex128 / "fp64" only.
    as_given:
        sweep every item with its rows as the Gray digits, as the reference
        does.  By default the planner sweeps perm(A^T; c, r) instead where
        that is clearly cheaper (the same number, rounded differently).
    """

    def __init__(
        self,
        matrices: Sequence,
        rows: Sequence,
        cols: Sequence,
        precision: str = "fp64",
        persistent=None,
        target_inner: int = 0,
        expand_cols=None,
        total_units: int = 0,
        gradient: bool = False,
        as_given: bool = False,
    ):
        if not (len(matrices) == len(rows) == len(cols)):
            raise ValueError("matrices, rows and cols must have the same length")
        if precision not in ("fp64", "fp32"):
            raise ValueError("precision must be 'fp64' or 'fp32'")
        self.count = len(matrices)
        self.precision = precision
        if gradient:
            # gradient sweeps: one block per work unit, the library's unit count
            if persistent is None:
                persistent = False
        elif persistent is None or not total_units:
            from .tune import best_config, mean_log2_terms

            try:
                _, cus, arch = _capi.device_info()
                tuned = best_config(len(matrices), cus, arch, log2_terms=mean_log2_terms(rows))
            except RuntimeError:  # no device: creation fails below with the C library's error
                tuned = {}
            if persistent is None:
                persistent = bool(tuned.get("persistent", False))
            if not total_units:
                total_units = int(tuned.get("total_units", 0))
                # a tuned inner-sweep length replaces the split total_units
                # asks for, so it applies only where the caller gave neither
                if not target_inner:
                    target_inner = int(tuned.get("target_inner", 0))
        self.gradient = bool(gradient)
        n = self.count
        shared = _shared_marshal(matrices, rows, cols)
        if shared is not None:
            # one matrix for every item and multiplicities that stack into
            # two int32 arrays (probability tables): addresses by arithmetic
            self._keep, mat_arr, row_arr, col_arr, m_r, m_c = shared
        else:
            self._keep = []
            mats, rws, cls = [], [], []
            for m, r, c in zip(matrices, rows, cols):
                a, rr, cc = _prep(m, r, c)
                self._keep.extend((a, rr, cc))
                mats.append(a.ctypes.data)
                rws.append(rr.ctypes.data_as(_i32p))
                cls.append(cc.ctypes.data_as(_i32p))
            mat_arr = (ctypes.c_void_p * max(n, 1))(*mats)
            row_arr = (_i32p * max(n, 1))(*rws)
            col_arr = (_i32p * max(n, 1))(*cls)
            m_r = np.array([self._keep[3 * k].shape[0] for k in range(n)] or [0], dtype=np.int64)
            m_c = np.array([self._keep[3 * k].shape[1] for k in range(n)] or [0], dtype=np.int64)
        flags = 0
        if precision == "fp32":
            flags |= _capi.PQP_FLAG_FP32
        if not persistent:
            flags |= _capi.PQP_FLAG_STATIC
        if expand_cols is True:
            flags |= _capi.PQP_FLAG_EXPAND_COLS
        elif expand_cols is False:
            flags |= _capi.PQP_FLAG_KEEP_COLS
        if gradient:
            flags |= _capi.PQP_FLAG_GRADIENT
        if as_given:
            flags |= _capi.PQP_FLAG_AS_GIVEN
        handle = ctypes.c_void_p()
        L = _capi.lib()
        st = L.pqp_batch_create(
            n, mat_arr, 0, m_r.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            m_c.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            m_c.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            row_arr, col_arr, flags, int(target_inner), int(total_units), ctypes.byref(handle),
        )
        _capi.check(st)
        self._h = handle
        self._shapes = [(int(m_r[k]), int(m_c[k])) for k in range(n)]
        self._keep = None  # inputs were copied into the plan

    # -- running ---------------------------------------------------------
    def run(self, mode: str = "graph") -> "PermanentBatch":
        """Enqueue one evaluation: mode "graph" (hipGraph replay), "eager"
        (direct launches) or "plain" (plain FP64 kernel, one launch per
        permanent)."""
        code = {"eager": _capi.PQP_RUN_EAGER, "graph": _capi.PQP_RUN_GRAPH,
                "plain": _capi.PQP_RUN_PLAIN}[mode]
        _capi.check(_capi.lib().pqp_batch_run(self._h, code))
        return self

    def sync(self) -> None:
        _capi.check(_capi.lib().pqp_batch_sync(self._h))

    def results(self, with_dd: bool = False):
        """Results of the last run as complex128 (and the dd parts)."""
        out = np.zeros((max(self.count, 1), 2))
        dd = np.zeros((max(self.count, 1), 4))
        _capi.check(
            _capi.lib().pqp_batch_results(
                self._h, out.ctypes.data_as(_dp), dd.ctypes.data_as(_dp) if with_dd else None
            )
        )
        vals = (out[:, 0] + 1j * out[:, 1])[: self.count]
        return (vals, dd[: self.count]) if with_dd else vals

    def error_bounds(self) -> np.ndarray:
        """Bounds on |result - permanent| of every item from the last run:
        4 (n + 2) u sum|terms| / 2^(n-1) plus the final rounding (u = 2^-53,
        or 2^-24 with FP32 products), from the term magnitudes the sweep
        kernels sum beside the terms.  NaN where the last run kept none (the
        plain baseline, gradient batches)."""
        b = np.zeros(max(self.count, 1))
        _capi.check(_capi.lib().pqp_batch_error_bounds(self._h, b.ctypes.data_as(_dp)))
        return b[: self.count]

    def gradients(self):
        """d perm / d A of every item from the last run (gradient batches):
        a list of complex128 arrays of the items' shapes."""
        grads = [np.zeros(sh, dtype=np.complex128) for sh in self._shapes]
        ptrs = (_dp * max(self.count, 1))(*[g.view(np.float64).ctypes.data_as(_dp) for g in grads])
        _capi.check(_capi.lib().pqp_batch_gradients(self._h, ptrs))
        return grads

    def time(self, mode: str = "graph", reps: int = 5) -> np.ndarray:
        """Milliseconds per evaluation for `reps` evaluations, timed with HIP
        events on the batch stream.  mode "kernel" times the sweep kernels
        alone (the finalize kernel runs outside the events)."""
        code = {"eager": _capi.PQP_RUN_EAGER, "graph": _capi.PQP_RUN_GRAPH,
                "plain": _capi.PQP_RUN_PLAIN, "kernel": _capi.PQP_RUN_KERNEL}[mode]
        ms = np.zeros(max(reps, 1), dtype=np.float32)
        _capi.check(
            _capi.lib().pqp_batch_time(self._h, code, reps, ms.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        )
        return ms[:reps].astype(np.float64)

    def time_span(self, mode: str = "graph", reps: int = 10) -> float:
        """Milliseconds of `reps` evaluations issued back to back between
        one pair of HIP events on the batch stream (after it has drained),
        every kernel of each evaluation included; divide by reps for the
        period of one evaluation.  mode "graph" or "eager"."""
        if mode not in ("graph", "eager"):
            raise ValueError("time_span takes mode 'graph' or 'eager'")
        code = {"eager": _capi.PQP_RUN_EAGER, "graph": _capi.PQP_RUN_GRAPH}[mode]
        ms = ctypes.c_float(0.0)
        _capi.check(_capi.lib().pqp_batch_time_span(self._h, code, int(reps), ctypes.byref(ms)))
        return float(ms.value)

    def stats(self) -> dict:
        flops = ctypes.c_double(0)
        terms = ctypes.c_int64(0)
        units = ctypes.c_int64(0)
        arena = ctypes.c_int64(0)
        lane_terms = ctypes.c_int64(0)
        _capi.check(
            _capi.lib().pqp_batch_stats(
                self._h, ctypes.byref(flops), ctypes.byref(terms), ctypes.byref(units),
                ctypes.byref(arena), ctypes.byref(lane_terms),
            )
        )
        return {"flops": flops.value, "terms": terms.value, "units": units.value,
                "arena_bytes": arena.value, "lane_terms": lane_terms.value,
                "lane_efficiency": terms.value / lane_terms.value if lane_terms.value else 1.0}

    def close(self) -> None:
        if getattr(self, "_h", None):
            _capi.lib().pqp_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def permanent_batch(matrices, rows, cols, precision: str = "fp64",
                    target_inner: int = 0) -> np.ndarray:
    """Permanents of many (matrix, rows, cols) triples in one launch."""
    with PermanentBatch(matrices, rows, cols, precision=precision,
                        target_inner=target_inner) as b:
        b.run("eager")
        return b.results()
