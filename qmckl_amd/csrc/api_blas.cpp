// The reference's public BLAS-like entry points (org/qmckl_blas.org: qmckl_dgemm :940-1118,
// qmckl_adjugate :1686-1795, qmckl_dgemv :3180-3300, qmckl_dger :3310-3420), with its argument
// checks and error codes.  Host pointers in and out, as the reference takes them; the
// arithmetic runs on the context's GPU (kernels_blas.hip, and the Gauss-Jordan kernel of
// kernels_det.hip for adjugates above order 5).  Column-major (Fortran-order) operands.
#include <cstring>

#include "internal.hpp"

namespace qmckl_amd {

namespace {

// elements spanned by a column-major rows x cols block with leading dimension ld
size_t span(int64_t rows, int64_t cols, int64_t ld) { return (size_t)(ld * (cols - 1) + rows); }
// elements spanned by a vector of n entries with stride inc
size_t vspan(int64_t n, int64_t inc) { return (size_t)(inc * (n - 1) + 1); }

struct Staged {
  Context* c;
  qmckl_context h;
  const char* fn;
  hipError_t e = hipSuccess;
  double* put(double* d, const double* src, size_t n) {
    if (e == hipSuccess) e = hipMemcpyAsync(d, src, sizeof(double) * n, hipMemcpyHostToDevice, c->stream);
    return d + n;
  }
  void get(double* dst, const double* d, size_t n) {
    if (e == hipSuccess) e = hipMemcpyAsync(dst, d, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream);
  }
  qmckl_exit_code finish() {
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? QMCKL_SUCCESS : hip_fail(h, e, fn);
  }
};

}  // namespace

}  // namespace qmckl_amd

using namespace qmckl_amd;

extern "C" {

qmckl_exit_code qmckl_dgemm(const qmckl_context context, const char TransA, const char TransB, const int64_t m,
                            const int64_t n, const int64_t k, const double alpha, const double* A, const int64_t lda,
                            const double* B, const int64_t ldb, const double beta, double* const C,
                            const int64_t ldc) {
  const char* fn = "qmckl_dgemm";
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  const bool ta = TransA == 'T' || TransA == 't', tb = TransB == 'T' || TransB == 't';
  if (!ta && TransA != 'N' && TransA != 'n') return fail(context, QMCKL_INVALID_ARG_2, fn, "TransA is not N or T");
  if (!tb && TransB != 'N' && TransB != 'n') return fail(context, QMCKL_INVALID_ARG_3, fn, "TransB is not N or T");
  if (m <= 0) return fail(context, QMCKL_INVALID_ARG_4, fn, "m <= 0");
  if (n <= 0) return fail(context, QMCKL_INVALID_ARG_5, fn, "n <= 0");
  if (k <= 0) return fail(context, QMCKL_INVALID_ARG_6, fn, "k <= 0");
  if (A == nullptr) return fail(context, QMCKL_INVALID_ARG_8, fn, "NULL pointer");
  if (lda <= 0 || lda < (ta ? k : m)) return fail(context, QMCKL_INVALID_ARG_9, fn, "lda below the rows of A");
  if (B == nullptr) return fail(context, QMCKL_INVALID_ARG_10, fn, "NULL pointer");
  if (ldb <= 0 || ldb < (tb ? n : k)) return fail(context, QMCKL_INVALID_ARG_11, fn, "ldb below the rows of B");
  if (C == nullptr) return fail(context, QMCKL_INVALID_ARG_13, fn, "NULL pointer");
  if (ldc <= 0 || ldc < m) return fail(context, QMCKL_INVALID_ARG_14, fn, "ldc < m");
  qmckl_exit_code rc = device_ready(c, context, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  const size_t na = ta ? span(k, m, lda) : span(m, k, lda), nb = tb ? span(n, k, ldb) : span(k, n, ldb),
               nc = span(m, n, ldc);
  Staged st{c, context, fn};
  st.e = c->d_blas.reserve(sizeof(double) * (na + nb + nc));
  double* dA = c->d_blas.as<double>();
  double* dB = st.put(dA, A, na);
  double* dC = st.put(dB, B, nb);
  st.put(dC, C, nc);  // the whole span: the rows past m of each column are written back unchanged
  if (st.e == hipSuccess) st.e = launch_dgemm(ta, tb, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc, c->stream);
  st.get(C, dC, nc);
  return st.finish();
}

qmckl_exit_code qmckl_dgemv(const qmckl_context context, const char Trans, const int64_t m, const int64_t n,
                            const double alpha, const double* A, const int64_t lda, const double* X,
                            const int64_t incx, const double beta, double* const Y, const int64_t incy) {
  const char* fn = "qmckl_dgemv";
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  const bool t = Trans == 'T' || Trans == 't';
  if (!t && Trans != 'N' && Trans != 'n') return fail(context, QMCKL_INVALID_ARG_2, fn, "Trans is not N or T");
  if (m <= 0) return fail(context, QMCKL_INVALID_ARG_3, fn, "m <= 0");
  if (n <= 0) return fail(context, QMCKL_INVALID_ARG_4, fn, "n <= 0");
  if (A == nullptr) return fail(context, QMCKL_INVALID_ARG_6, fn, "NULL pointer");
  if (lda <= 0 || lda < m) return fail(context, QMCKL_INVALID_ARG_7, fn, "lda < m");
  if (X == nullptr) return fail(context, QMCKL_INVALID_ARG_8, fn, "NULL pointer");
  if (incx <= 0) return fail(context, QMCKL_INVALID_ARG_9, fn, "incx <= 0");
  if (Y == nullptr) return fail(context, QMCKL_INVALID_ARG_11, fn, "NULL pointer");
  if (incy <= 0) return fail(context, QMCKL_INVALID_ARG_12, fn, "incy <= 0");
  qmckl_exit_code rc = device_ready(c, context, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  const int64_t lx = t ? m : n, ly = t ? n : m;  // lengths of x and y
  const size_t na = span(m, n, lda), nx = vspan(lx, incx), ny = vspan(ly, incy);
  Staged st{c, context, fn};
  st.e = c->d_blas.reserve(sizeof(double) * (na + nx + ny));
  double* dA = c->d_blas.as<double>();
  double* dX = st.put(dA, A, na);
  double* dY = st.put(dX, X, nx);
  st.put(dY, Y, ny);
  if (st.e == hipSuccess) st.e = launch_dgemv(t, m, n, alpha, dA, lda, dX, incx, beta, dY, incy, c->stream);
  st.get(Y, dY, ny);
  return st.finish();
}

qmckl_exit_code qmckl_dger(const qmckl_context context, const int64_t m, const int64_t n, const double alpha,
                           const double* X, const int64_t incx, const double* Y, const int64_t incy, double* const A,
                           const int64_t lda) {
  const char* fn = "qmckl_dger";
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (m <= 0) return fail(context, QMCKL_INVALID_ARG_2, fn, "m <= 0");
  if (n <= 0) return fail(context, QMCKL_INVALID_ARG_3, fn, "n <= 0");
  if (X == nullptr) return fail(context, QMCKL_INVALID_ARG_5, fn, "NULL pointer");
  if (incx <= 0) return fail(context, QMCKL_INVALID_ARG_6, fn, "incx <= 0");
  if (Y == nullptr) return fail(context, QMCKL_INVALID_ARG_7, fn, "NULL pointer");
  if (incy <= 0) return fail(context, QMCKL_INVALID_ARG_8, fn, "incy <= 0");
  if (A == nullptr) return fail(context, QMCKL_INVALID_ARG_9, fn, "NULL pointer");
  if (lda <= 0 || lda < m) return fail(context, QMCKL_INVALID_ARG_10, fn, "lda < m");
  qmckl_exit_code rc = device_ready(c, context, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  const size_t nx = vspan(m, incx), ny = vspan(n, incy), na = span(m, n, lda);
  Staged st{c, context, fn};
  st.e = c->d_blas.reserve(sizeof(double) * (na + nx + ny));
  double* dA = c->d_blas.as<double>();
  double* dX = st.put(dA, A, na);
  double* dY = st.put(dX, X, nx);
  st.put(dY, Y, ny);
  if (st.e == hipSuccess) st.e = launch_dger(m, n, alpha, dX, incx, dY, incy, dA, lda, c->stream);
  st.get(A, dA, na);
  return st.finish();
}

qmckl_exit_code qmckl_adjugate(const qmckl_context context, const int64_t n, const double* A, const int64_t lda,
                               double* const B, const int64_t ldb, double* det_l) {
  const char* fn = "qmckl_adjugate";
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (n <= 0) return fail(context, QMCKL_INVALID_ARG_2, fn, "n <= 0");
  if (A == nullptr) return fail(context, QMCKL_INVALID_ARG_3, fn, "NULL pointer");
  if (lda < n) return fail(context, QMCKL_INVALID_ARG_4, fn, "lda < n");
  if (B == nullptr) return fail(context, QMCKL_INVALID_ARG_5, fn, "NULL pointer");
  if (ldb < n) return fail(context, QMCKL_INVALID_ARG_6, fn, "ldb < n");
  if (det_l == nullptr) return fail(context, QMCKL_INVALID_ARG_7, fn, "NULL pointer");
  if (n > INT32_MAX / 2) return fail(context, QMCKL_INVALID_ARG_2, fn, "n too large");
  qmckl_exit_code rc = device_ready(c, context, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  const size_t na = span(n, n, lda), nb = span(n, n, ldb), nw = n > 128 ? (size_t)n * n : 0;
  Staged st{c, context, fn};
  st.e = c->d_blas.reserve(sizeof(double) * (na + nb + nw + 1));
  double* dA = c->d_blas.as<double>();
  double* dB = st.put(dA, A, na);
  double* dW = st.put(dB, B, nb);
  double* dD = dW + nw;
  if (st.e == hipSuccess)
    st.e = n <= 5 ? launch_adjugate_small((int)n, dA, lda, dB, ldb, dD, c->stream)
                  : launch_adjugate_gj((int)n, dA, lda, dB, ldb, dD, dW, c->stream);
  double det = 0.0;
  st.get(&det, dD, 1);
  st.get(B, dB, nb);
  rc = st.finish();
  if (rc != QMCKL_SUCCESS) return rc;
  *det_l = det;
  if (n > 5 && det == 0.0) return fail(context, QMCKL_FAILURE, fn, "singular matrix (a zero pivot)");
  return QMCKL_SUCCESS;
}

// Size-safe proxies (org/qmckl_blas.org:1276-1440, 2595-2700): the sizes of the arrays are
// checked as well (at least the elements the operation spans), with the reference's codes.
qmckl_exit_code qmckl_dgemm_safe(const qmckl_context context, const char TransA, const char TransB, const int64_t m,
                                 const int64_t n, const int64_t k, const double alpha, const double* A,
                                 const int64_t size_max_A, const int64_t lda, const double* B, const int64_t size_max_B,
                                 const int64_t ldb, const double beta, double* const C, const int64_t size_max_C,
                                 const int64_t ldc) {
  const char* fn = "qmckl_dgemm_safe";
  if (ctx_of(context) == nullptr) return QMCKL_INVALID_CONTEXT;
  if (m <= 0) return fail(context, QMCKL_INVALID_ARG_4, fn, "m <= 0");
  if (n <= 0) return fail(context, QMCKL_INVALID_ARG_5, fn, "n <= 0");
  if (k <= 0) return fail(context, QMCKL_INVALID_ARG_6, fn, "k <= 0");
  if (lda <= 0) return fail(context, QMCKL_INVALID_ARG_10, fn, "lda <= 0");
  if (ldb <= 0) return fail(context, QMCKL_INVALID_ARG_13, fn, "ldb <= 0");
  if (ldc <= 0) return fail(context, QMCKL_INVALID_ARG_17, fn, "ldc <= 0");
  const bool ta = TransA == 'T' || TransA == 't', tb = TransB == 'T' || TransB == 't';
  if (size_max_A <= 0 || (size_t)size_max_A < (ta ? span(k, m, lda) : span(m, k, lda)))
    return fail(context, QMCKL_INVALID_ARG_9, fn, "size_max_A below the elements of A");
  if (size_max_B <= 0 || (size_t)size_max_B < (tb ? span(n, k, ldb) : span(k, n, ldb)))
    return fail(context, QMCKL_INVALID_ARG_12, fn, "size_max_B below the elements of B");
  if (size_max_C <= 0 || (size_t)size_max_C < span(m, n, ldc))
    return fail(context, QMCKL_INVALID_ARG_16, fn, "size_max_C below the elements of C");
  qmckl_exit_code rc = qmckl_dgemm(context, TransA, TransB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
  // the arguments after A each moved by one (size_max_A), after B by two, after C by three
  if (rc >= QMCKL_INVALID_ARG_9 && rc <= QMCKL_INVALID_ARG_10) rc = (qmckl_exit_code)(rc + 1);
  else if (rc >= QMCKL_INVALID_ARG_11 && rc <= QMCKL_INVALID_ARG_13) rc = (qmckl_exit_code)(rc + 2);
  else if (rc == QMCKL_INVALID_ARG_14) rc = QMCKL_INVALID_ARG_17;
  return rc;
}

qmckl_exit_code qmckl_adjugate_safe(const qmckl_context context, const int64_t n, const double* A,
                                    const int64_t size_max_A, const int64_t lda, double* const B,
                                    const int64_t size_max_B, const int64_t ldb, double* det_l) {
  const char* fn = "qmckl_adjugate_safe";
  if (ctx_of(context) == nullptr) return QMCKL_INVALID_CONTEXT;
  if (n > 0 && (size_max_A < n * n || (lda >= n && (size_t)size_max_A < span(n, n, lda))))
    return fail(context, QMCKL_INVALID_ARG_4, fn, "size_max_A below the elements of A");
  if (n > 0 && (size_max_B < n * n || (ldb >= n && (size_t)size_max_B < span(n, n, ldb))))
    return fail(context, QMCKL_INVALID_ARG_7, fn, "size_max_B below the elements of B");
  qmckl_exit_code rc = qmckl_adjugate(context, n, A, lda, B, ldb, det_l);
  // the arguments after A moved by one (size_max_A), after B by two
  if (rc == QMCKL_INVALID_ARG_4) rc = QMCKL_INVALID_ARG_5;
  else if (rc == QMCKL_INVALID_ARG_5) rc = QMCKL_INVALID_ARG_6;
  else if (rc == QMCKL_INVALID_ARG_6) rc = QMCKL_INVALID_ARG_8;
  else if (rc == QMCKL_INVALID_ARG_7) rc = QMCKL_INVALID_ARG_9;
  return rc;
}

}  // extern "C"
