// Test-infrastructure only: a C-ABI driver around the reference's own C++
// permanent code, so that the tests and the CPU baseline can call the real
// reference functions through ctypes.  It is compiled together with the
// reference sources (src/permanent.cpp, src/permanent_laplace.cpp) by
// oracle/Makefile into oracle/_ref/libpqref.so.  Nothing in the product
// (piquasso_amd/) links, loads or calls this file.
//
// Entry points wrap:
//   permanent_cpp<double> / permanent_cpp<float>   src/permanent.cpp:49-269
//   grad_perm                                       src/permanent.cpp:271-300
//   permanent_laplace_cpp<double>                   src/permanent_laplace.cpp:43-240
//
// Every function copies its inputs into buffers it owns (the reference
// wraps raw pointers in non-const Matrix/Vector views), catches the
// reference's `throw std::string` and returns a status: 0 = ok, -1 = the
// reference threw (row/column sum mismatch), -2 = anything else.

#include <complex>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "matrix.hpp"
#include "permanent.hpp"
#include "permanent_laplace.hpp"

namespace {

template <typename T>
int run_perm(const T *a, int m_r, int m_c, const int *rows, const int *cols,
             T *out) {
  try {
    std::vector<std::complex<T>> abuf(static_cast<size_t>(m_r) * m_c);
    for (size_t k = 0; k < abuf.size(); ++k)
      abuf[k] = std::complex<T>(a[2 * k], a[2 * k + 1]);
    std::vector<int> rbuf(rows, rows + m_r), cbuf(cols, cols + m_c);
    Matrix<std::complex<T>> mat(m_r, m_c, abuf.data());
    Vector<int> rv(m_r, rbuf.data()), cv(m_c, cbuf.data());
    std::complex<T> r = permanent_cpp<T>(mat, rv, cv);
    out[0] = r.real();
    out[1] = r.imag();
    return 0;
  } catch (const std::string &) {
    return -1;
  } catch (...) {
    return -2;
  }
}

}  // namespace

extern "C" {

int pqref_perm_c128(const double *a, int m_r, int m_c, const int *rows,
                    const int *cols, double *out) {
  return run_perm<double>(a, m_r, m_c, rows, cols, out);
}

int pqref_perm_c64(const float *a, int m_r, int m_c, const int *rows,
                   const int *cols, float *out) {
  return run_perm<float>(a, m_r, m_c, rows, cols, out);
}

// grad_perm assumes a square n x n matrix (src/permanent.cpp:274-276).
int pqref_grad_c128(const double *a, int n, const int *rows, const int *cols,
                    double *g) {
  try {
    std::vector<std::complex<double>> abuf(static_cast<size_t>(n) * n);
    for (size_t k = 0; k < abuf.size(); ++k)
      abuf[k] = std::complex<double>(a[2 * k], a[2 * k + 1]);
    std::vector<int> rbuf(rows, rows + n), cbuf(cols, cols + n);
    Matrix<std::complex<double>> mat(n, n, abuf.data());
    Vector<int> rv(n, rbuf.data()), cv(n, cbuf.data());
    Matrix<std::complex<double>> gm = grad_perm(mat, rv, cv);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        // grad_perm leaves skipped entries uninitialised; report them as 0.
        std::complex<double> v =
            (rows[i] == 0 || cols[j] == 0) ? std::complex<double>(0.0, 0.0)
                                           : gm(i, j);
        g[2 * (i * n + j)] = v.real();
        g[2 * (i * n + j) + 1] = v.imag();
      }
    return 0;
  } catch (const std::string &) {
    return -1;
  } catch (...) {
    return -2;
  }
}

int pqref_laplace_c128(const double *a, int m_r, int m_c, const int *rows,
                       const int *cols, double *out) {
  try {
    std::vector<std::complex<double>> abuf(static_cast<size_t>(m_r) * m_c);
    for (size_t k = 0; k < abuf.size(); ++k)
      abuf[k] = std::complex<double>(a[2 * k], a[2 * k + 1]);
    std::vector<int> rbuf(rows, rows + m_r), cbuf(cols, cols + m_c);
    Matrix<std::complex<double>> mat(m_r, m_c, abuf.data());
    Vector<int> rv(m_r, rbuf.data()), cv(m_c, cbuf.data());
    Vector<std::complex<double>> res = permanent_laplace_cpp<double>(mat, rv, cv);
    for (size_t k = 0; k < res.size(); ++k) {
      out[2 * k] = res[k].real();
      out[2 * k + 1] = res[k].imag();
    }
    return static_cast<int>(res.size());
  } catch (const std::string &) {
    return -1;
  } catch (...) {
    return -2;
  }
}

}  // extern "C"
