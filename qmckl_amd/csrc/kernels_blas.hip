// Dense linear algebra behind the reference's public BLAS-like entry points, on column-major
// (Fortran-order) operands as those functions take them: qmckl_dgemm (org/qmckl_blas.org:
// 940-1118), qmckl_dgemv (:3180-3300), qmckl_dger (:3310-3420) and the small-matrix cases of
// qmckl_adjugate (:1686-1790).
#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace qmckl_amd {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int GM = 64, GN = 64, GK = 16;

// A 64 x 64 tile of C = alpha op(A) op(B) + beta C on v_mfma_f64_16x16x4f64: wave w holds rows
// 16 w .. 16 w + 15 of the tile in four 16-column accumulators; op(A) and op(B) are staged in
// LDS GK columns / rows at a time, each read along its memory order, zero past the matrix.  As
// BLAS does, beta = 0 does not read C and alpha = 0 does not use the product.
__global__ void __launch_bounds__(256) k_dgemm(int ta, int tb, int64_t m, int64_t n, int64_t k, double alpha,
                                               const double* __restrict__ A, int64_t lda, const double* __restrict__ B,
                                               int64_t ldb, double beta, double* __restrict__ C, int64_t ldc) {
  __shared__ double As[GM][GK + 1];
  __shared__ double Bs[GK][GN + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, li = lane & 15, lg = lane >> 4;
  const int64_t i0 = (int64_t)blockIdx.x * GM, j0 = (int64_t)blockIdx.y * GN;
  d4 acc[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) acc[c] = d4{0.0, 0.0, 0.0, 0.0};
  for (int64_t p0 = 0; p0 < k; p0 += GK) {
    for (int t = tid; t < GM * GK; t += 256) {
      const int r = ta ? t / GK : t % GM, q = ta ? t % GK : t / GM;  // op(A)(i0 + r, p0 + q)
      const int64_t i = i0 + r, p = p0 + q;
      As[r][q] = (i < m && p < k) ? (ta ? A[p + i * lda] : A[i + p * lda]) : 0.0;
    }
    for (int t = tid; t < GK * GN; t += 256) {
      const int q = tb ? t / GN : t % GK, c = tb ? t % GN : t / GK;  // op(B)(p0 + q, j0 + c)
      const int64_t p = p0 + q, j = j0 + c;
      Bs[q][c] = (p < k && j < n) ? (tb ? B[j + p * ldb] : B[p + j * ldb]) : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < GK / 4; ++s) {
      const double a = As[16 * wave + li][4 * s + lg];
#pragma unroll
      for (int c = 0; c < 4; ++c)
        acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, Bs[4 * s + lg][16 * c + li], acc[c], 0, 0, 0);
    }
    __syncthreads();
  }
  // register q of lane (lg, li) holds row lg + 4 q, column li of its 16 x 16 tile
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t i = i0 + 16 * wave + lg + 4 * q, j = j0 + 16 * c + li;
      if (i < m && j < n) {
        double v = alpha == 0.0 ? 0.0 : alpha * acc[c][q];
        if (beta != 0.0) v += beta * C[i + j * ldc];
        C[i + j * ldc] = v;
      }
    }
}

// y = alpha A x + beta y, A m x n: one thread per row (consecutive rows are consecutive in
// memory)
__global__ void __launch_bounds__(256) k_dgemv_n(int64_t m, int64_t n, double alpha, const double* __restrict__ A,
                                                 int64_t lda, const double* __restrict__ x, int64_t incx, double beta,
                                                 double* __restrict__ y, int64_t incy) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= m) return;
  double s = 0.0;
  for (int64_t j = 0; j < n; ++j) s += A[i + j * lda] * x[j * incx];
  double v = alpha * s;
  if (beta != 0.0) v += beta * y[i * incy];
  y[i * incy] = v;
}

// y = alpha A^T x + beta y, A m x n: one wave per column, lanes along the column, then a
// fixed-order butterfly sum over the wave
__global__ void __launch_bounds__(256) k_dgemv_t(int64_t m, int64_t n, double alpha, const double* __restrict__ A,
                                                 int64_t lda, const double* __restrict__ x, int64_t incx, double beta,
                                                 double* __restrict__ y, int64_t incy) {
  const int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (j >= n) return;  // the whole wave
  double s = 0.0;
  for (int64_t i = lane; i < m; i += 64) s += A[i + j * lda] * x[i * incx];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (lane == 0) {
    double v = alpha * s;
    if (beta != 0.0) v += beta * y[j * incy];
    y[j * incy] = v;
  }
}

// A += alpha x y^T, A m x n
__global__ void __launch_bounds__(256) k_dger(int64_t m, int64_t n, double alpha, const double* __restrict__ x,
                                              int64_t incx, const double* __restrict__ y, int64_t incy,
                                              double* __restrict__ A, int64_t lda) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= m) return;
  for (int64_t j = blockIdx.y; j < n; j += gridDim.y) A[i + j * lda] += x[i * incx] * (alpha * y[j * incy]);
}

// Determinant of the r x r matrix M (row-major, destroyed) by Gaussian elimination with
// partial pivoting; 0 for a zero pivot column.
__device__ double small_det(double* M, int r) {
  double det = 1.0;
  for (int c = 0; c < r; ++c) {
    int p = c;
    for (int q = c + 1; q < r; ++q)
      if (fabs(M[q * r + c]) > fabs(M[p * r + c])) p = q;
    if (M[p * r + c] == 0.0) return 0.0;
    if (p != c) {
      for (int q = 0; q < r; ++q) {
        const double x = M[c * r + q];
        M[c * r + q] = M[p * r + q];
        M[p * r + q] = x;
      }
      det = -det;
    }
    det *= M[c * r + c];
    for (int q = c + 1; q < r; ++q) {
      const double f = M[q * r + c] / M[c * r + c];
      for (int u = c + 1; u < r; ++u) M[q * r + u] -= f * M[c * r + u];
    }
  }
  return det;
}

// The adjugate of an n x n matrix, n <= 5, by cofactors (as the reference's explicit formulas
// for these orders, defined for singular matrices too): thread (i, j) writes B(i, j) =
// (-1)^(i+j) det(A without row j and column i); thread n^2 writes det(A).
__global__ void k_adjugate_small(int n, const double* __restrict__ A, int64_t lda, double* __restrict__ B,
                                 int64_t ldb, double* __restrict__ det) {
  const int t = threadIdx.x;
  double M[25];
  if (t < n * n) {
    const int i = t % n, j = t / n, r = n - 1;
    int rr = 0;
    for (int row = 0; row < n; ++row) {
      if (row == j) continue;
      int cc = 0;
      for (int c = 0; c < n; ++c) {
        if (c == i) continue;
        M[rr * r + cc] = A[row + c * lda];
        ++cc;
      }
      ++rr;
    }
    const double d = small_det(M, r);
    B[i + j * ldb] = ((i + j) & 1) ? -d : d;
  } else if (t == n * n) {
    for (int row = 0; row < n; ++row)
      for (int c = 0; c < n; ++c) M[row * n + c] = A[row + c * lda];
    *det = small_det(M, n);
  }
}

}  // namespace

hipError_t launch_dgemm(bool ta, bool tb, int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t lda,
                        const double* B, int64_t ldb, double beta, double* C, int64_t ldc, hipStream_t s) {
  const int64_t gy = (n + GN - 1) / GN;
  if (gy > 65535) return hipErrorInvalidValue;
  k_dgemm<<<dim3((unsigned)((m + GM - 1) / GM), (unsigned)gy), dim3(256), 0, s>>>(ta, tb, m, n, k, alpha, A, lda, B, ldb,
                                                                                    beta, C, ldc);
  return hipGetLastError();
}

hipError_t launch_dgemv(bool t, int64_t m, int64_t n, double alpha, const double* A, int64_t lda, const double* x,
                        int64_t incx, double beta, double* y, int64_t incy, hipStream_t s) {
  if (t)
    k_dgemv_t<<<dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s>>>(m, n, alpha, A, lda, x, incx, beta, y, incy);
  else
    k_dgemv_n<<<dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s>>>(m, n, alpha, A, lda, x, incx, beta, y, incy);
  return hipGetLastError();
}

hipError_t launch_dger(int64_t m, int64_t n, double alpha, const double* x, int64_t incx, const double* y,
                       int64_t incy, double* A, int64_t lda, hipStream_t s) {
  const unsigned gy = (unsigned)(n < 65535 ? n : 65535);
  k_dger<<<dim3((unsigned)((m + 255) / 256), gy), dim3(256), 0, s>>>(m, n, alpha, x, incx, y, incy, A, lda);
  return hipGetLastError();
}

hipError_t launch_adjugate_small(int n, const double* A, int64_t lda, double* B, int64_t ldb, double* det,
                                 hipStream_t s) {
  k_adjugate_small<<<dim3(1), dim3(64), 0, s>>>(n, A, lda, B, ldb, det);
  return hipGetLastError();
}

}  // namespace qmckl_amd
