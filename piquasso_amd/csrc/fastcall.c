/*
 * _pqp_fast: the drop-in permanent(matrix, rows, cols) call of
 * piquasso_amd._math.permanent for the common argument types, bound with
 * the CPython C API instead of ctypes.  A single-permanent call is
 * latency-bound (about 10 us on the device side at n <= 12), and ctypes
 * spends about 1 us per array just to find its address; the buffer protocol
 * does that in tens of nanoseconds.
 *
 *   permanent_c128(matrix, rows, cols) -> complex | int
 *
 * matrix: C-contiguous 2-D complex128 buffer; rows, cols: 1-D int32 or
 * int64 buffers of lengths m_r and m_c.  Returns the permanent as a Python
 * complex from pqp_permanent_c128, or, when the C ABI reports a failure,
 * its negative status code as an int (the caller raises the matching
 * exception with pqp_last_error()).  Arguments of any other form raise
 * TypeError, and the caller takes its general path.
 */
#define PY_SSIZE_T_CLEAN
#include <Python.h>

#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "pqp_permanent.h"

static int int_format(const Py_buffer *b) {
  /* 4 for int32, 8 for int64, 0 otherwise */
  const char *f = b->format ? b->format : "B";
  if (*f == '<' || *f == '=' || *f == '@') ++f;
  if (b->itemsize == 4 && (strcmp(f, "i") == 0 || strcmp(f, "l") == 0)) return 4;
  if (b->itemsize == 8 && (strcmp(f, "l") == 0 || strcmp(f, "q") == 0)) return 8;
  return 0;
}

/* the multiplicities as int32: the buffer itself, or a converted copy in
   `stack` (or on the heap when longer); NULL when a value does not fit */
static const int32_t *as_int32(const Py_buffer *b, int width, int32_t *stack, size_t stack_len,
                               int32_t **heap) {
  const Py_ssize_t n = b->shape ? b->shape[0] : 0;
  if (width == 4) return (const int32_t *)b->buf;
  int32_t *dst = stack;
  if ((size_t)n > stack_len) {
    *heap = (int32_t *)malloc((size_t)n * sizeof(int32_t));
    if (!*heap) return NULL;
    dst = *heap;
  }
  const int64_t *src = (const int64_t *)b->buf;
  for (Py_ssize_t i = 0; i < n; ++i) {
    if (src[i] < INT32_MIN || src[i] > INT32_MAX) return NULL;
    dst[i] = (int32_t)src[i];
  }
  return dst;
}

static PyObject *permanent_c128(PyObject *self, PyObject *args) {
  (void)self;
  PyObject *m, *r, *c;
  if (!PyArg_ParseTuple(args, "OOO", &m, &r, &c)) return NULL;
  Py_buffer bm, br, bc;
  const int flags = PyBUF_C_CONTIGUOUS | PyBUF_FORMAT;
  if (PyObject_GetBuffer(m, &bm, flags) < 0) return NULL;
  if (PyObject_GetBuffer(r, &br, flags) < 0) {
    PyBuffer_Release(&bm);
    return NULL;
  }
  if (PyObject_GetBuffer(c, &bc, flags) < 0) {
    PyBuffer_Release(&bm);
    PyBuffer_Release(&br);
    return NULL;
  }
  PyObject *result = NULL;
  int32_t *heap_r = NULL, *heap_c = NULL;
  int32_t stack_r[256], stack_c[256];
  const int wr = int_format(&br), wc = int_format(&bc);
  const char *fm = bm.format ? bm.format : "B";
  if (*fm == '<' || *fm == '=' || *fm == '@') ++fm;
  if (bm.ndim != 2 || bm.itemsize != 16 || strcmp(fm, "Zd") != 0 || br.ndim != 1 ||
      bc.ndim != 1 || !wr || !wc) {
    PyErr_SetString(PyExc_TypeError, "permanent_c128: unsupported argument types");
    goto done;
  }
  const int64_t m_r = bm.shape[0], m_c = bm.shape[1];
  if (br.shape[0] != m_r || bc.shape[0] != m_c) {
    PyErr_SetString(PyExc_TypeError, "permanent_c128: multiplicity lengths do not match the matrix");
    goto done;
  }
  const int32_t *rows = as_int32(&br, wr, stack_r, 256, &heap_r);
  const int32_t *cols = as_int32(&bc, wc, stack_c, 256, &heap_c);
  if (!rows || !cols) {
    PyErr_SetString(PyExc_TypeError, "permanent_c128: multiplicities out of int32 range");
    goto done;
  }
  double out[2];
  int st;
  Py_BEGIN_ALLOW_THREADS
  st = pqp_permanent_c128((const double *)bm.buf, m_r, m_c, m_c, rows, cols, 0, out, NULL);
  Py_END_ALLOW_THREADS
  result = st == PQP_OK ? PyComplex_FromDoubles(out[0], out[1]) : PyLong_FromLong(st);
done:
  free(heap_r);
  free(heap_c);
  PyBuffer_Release(&bm);
  PyBuffer_Release(&br);
  PyBuffer_Release(&bc);
  return result;
}

static PyMethodDef methods[] = {
    {"permanent_c128", permanent_c128, METH_VARARGS,
     "permanent_c128(matrix, rows, cols) -> complex, or the negative C-ABI status"},
    {NULL, NULL, 0, NULL}};

static struct PyModuleDef module = {PyModuleDef_HEAD_INIT, "_pqp_fast",
                                    "CPython binding of the drop-in permanent call", -1, methods,
                                    NULL, NULL, NULL, NULL};

PyMODINIT_FUNC PyInit__pqp_fast(void) { return PyModule_Create(&module); }
