// Slater determinants, their inverses and the determinant part of the local energy
// (org/under_development/qmckl_determinant.org, qmckl_local_energy.org), the electrostatic
// potential per walker (org/qmckl_electron.org:1022-1192, 1477-1658), and the
// Sherman-Morrison-Woodbury updates of inverse Slater matrices
// (org/under_development/qmckl_sherman_morrison_woodbury.org), batched over matrices.
//
// One workgroup per matrix.  A Slater matrix of up to kDetLdsMax electrons is inverted in LDS
// (Gauss-Jordan with partial pivoting, the pivot row and column staged in LDS, n steps of
// three barriers); a larger one in its own slice of a global work buffer with the same code.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "internal.hpp"

namespace qmckl_amd {
namespace {

constexpr int kDetBlock = 256;
constexpr int kDetWaves = kDetBlock / 64;
constexpr int kDetLdsMax = 128;  // matrices up to this order are worked on in LDS

// Index of the largest |value| over the workgroup (smallest index on ties); every thread
// gets the result.  v < 0 marks an empty slot.
__device__ int block_argmax(double v, int i, double* sv, int* si) {
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_xor(v, off);
    const int oi = __shfl_xor(i, off);
    if (ov > v || (ov == v && oi < i)) {
      v = ov;
      i = oi;
    }
  }
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sv[wave] = v;
    si[wave] = i;
  }
  __syncthreads();
  double bv = sv[0];
  int bi = si[0];
  for (int q = 1; q < kDetWaves; ++q)
    if (sv[q] > bv || (sv[q] == bv && si[q] < bi)) {
      bv = sv[q];
      bi = si[q];
    }
  return bi;
}

__device__ double block_sum(double v, double* sv) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  const int wave = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sv[wave] = v;
  __syncthreads();
  double s = 0.0;
  for (int q = 0; q < kDetWaves; ++q) s += sv[q];
  return s;
}

// In-place inverse of the n x n matrix A (row stride ld) by Gauss-Jordan elimination with
// partial pivoting; returns det(A) (0 when a pivot column is all zero, A then unspecified).
// rowbuf, col: n doubles; perm: n ints; sv/si: kDetWaves slots.
__device__ double gauss_jordan(double* A, int n, int ld, double* rowbuf, double* col, int* perm, double* sv,
                               int* si) {
  const int t = threadIdx.x;
  const int tx = t & 63, ty = t >> 6;
  double det = 1.0;
  bool singular = false;
  for (int k = 0; k < n; ++k) {
    double best = -1.0;
    int bi = n;
    for (int i = k + t; i < n; i += kDetBlock) {
      const double v = fabs(A[i * ld + k]);
      if (v > best) {
        best = v;
        bi = i;
      }
    }
    const int p = block_argmax(best, bi, sv, si);
    const double pv = A[p * ld + k];
    if (t == 0) perm[k] = pv == 0.0 ? k : p;
    if (pv == 0.0) {  // uniform: every thread read the same pivot
      singular = true;
      __syncthreads();
      continue;
    }
    const double inv_piv = 1.0 / pv;
    det *= (p != k) ? -pv : pv;
    // swap rows k and p, stage the scaled pivot row and the pivot column
    for (int j = t; j < n; j += kDetBlock) {
      const double akj = A[k * ld + j], apj = A[p * ld + j];
      if (p != k) A[p * ld + j] = akj;
      rowbuf[j] = (j == k) ? inv_piv : apj * inv_piv;
    }
    for (int i = t; i < n; i += kDetBlock) col[i] = (i == k) ? 0.0 : (i == p ? A[k * ld + k] : A[i * ld + k]);
    __syncthreads();
    for (int i = ty; i < n; i += kDetWaves) {
      const double ci = col[i];
      if (i == k) {
        for (int j = tx; j < n; j += 64) A[i * ld + j] = rowbuf[j];
      } else {
        for (int j = tx; j < n; j += 64) {
          const double aij = (j == k) ? 0.0 : A[i * ld + j];
          A[i * ld + j] = aij - ci * rowbuf[j];
        }
      }
    }
    __syncthreads();
  }
  // the row interchanges become column interchanges of the inverse, last first
  for (int i = t; i < n; i += kDetBlock)
    for (int k = n - 1; k >= 0; --k) {
      const int p = perm[k];
      if (p != k) {
        const double x = A[i * ld + k];
        A[i * ld + k] = A[i * ld + p];
        A[i * ld + p] = x;
      }
    }
  __syncthreads();
  return singular ? 0.0 : det;
}

template <bool IN_LDS>
__global__ void __launch_bounds__(kDetBlock) k_slater(SlaterArgs a) {
  extern __shared__ double smem[];
  const int n = a.n;
  const int64_t m = blockIdx.x;  // d * walk_num + w
  const int64_t w = m % a.walk_num;
  const int64_t d = m / a.walk_num;
  const int ld = IN_LDS ? (n | 1) : n;
  double* A = IN_LDS ? smem : a.work + m * (int64_t)n * n;
  double* rowbuf = IN_LDS ? smem + n * ld : smem;
  double* col = rowbuf + n;
  double* sv = col + n;
  int* si = reinterpret_cast<int*>(sv + kDetWaves);
  int* perm = si + kDetWaves;
  int* idx = perm + n;
  const int32_t* gidx = a.idx + m * n;
  for (int j = threadIdx.x; j < n; j += kDetBlock) idx[j] = gidx[j];
  __syncthreads();
  const int64_t p0 = w * a.elec_num + a.e0;  // first point of the matrix's electrons
  const int64_t mo = a.mo_num;
  for (int i = threadIdx.x >> 6; i < n; i += kDetWaves) {
    const double* row = a.mo_vgl + (p0 + i) * 5 * mo;
    for (int j = threadIdx.x & 63; j < n; j += 64) A[i * ld + j] = row[idx[j]];
  }
  __syncthreads();
  const double det = gauss_jordan(A, n, ld, rowbuf, col, perm, sv, si);
  const double nan = __builtin_nan("");
  if (threadIdx.x == 0) a.value[m] = det;
  if (a.inv)
    for (int i = threadIdx.x >> 6; i < n; i += kDetWaves)
      for (int j = threadIdx.x & 63; j < n; j += 64)
        a.inv[(m * n + i) * n + j] = det == 0.0 ? nan : A[i * ld + j];
  if (a.kin == nullptr || d != 0) return;
  // grad_e D / D = sum_m grad phi_m(r_e) A^-1[m][e], and the Laplacian likewise
  double lap = 0.0;
  for (int e = threadIdx.x; e < n; e += kDetBlock) {
    const double* row = a.mo_vgl + (p0 + e) * 5 * mo;
    double g0 = 0.0, g1 = 0.0, g2 = 0.0, l = 0.0;
    for (int j = 0; j < n; ++j) {
      const double inv = A[j * ld + e];
      const int q = idx[j];
      g0 += row[mo + q] * inv;
      g1 += row[2 * mo + q] * inv;
      g2 += row[3 * mo + q] * inv;
      l += row[4 * mo + q] * inv;
    }
    double* dr = a.drift + (p0 + e) * 3;
    dr[0] = det == 0.0 ? nan : 2.0 * g0;
    dr[1] = det == 0.0 ? nan : 2.0 * g1;
    dr[2] = det == 0.0 ? nan : 2.0 * g2;
    lap += l;
  }
  const double s = block_sum(lap, sv);
  if (threadIdx.x == 0) a.kin[w] = det == 0.0 ? nan : -0.5 * s;
}

// component c of MO idx[m] at electron e of each matrix: out [d][W][5][n][n]
__global__ void k_slater_vgl(SlaterArgs a, double* out) {
  const int n = a.n;
  const int64_t m = blockIdx.x;
  const int64_t w = m % a.walk_num;
  const int32_t* idx = a.idx + m * n;
  const int64_t p0 = w * a.elec_num + a.e0;
  const int64_t nn = (int64_t)n * n;
  for (int64_t q = threadIdx.x; q < 5 * nn; q += blockDim.x) {
    const int c = (int)(q / nn), e = (int)((q % nn) / n), j = (int)(q % n);
    out[m * 5 * nn + q] = a.mo_vgl[((p0 + e) * 5 + c) * a.mo_num + idx[j]];
  }
}

// Per walker: V_ee (pairs closer than 1e-5 left out) and V_en (distances below 1e-6 left out),
// the cut-offs of org/qmckl_electron.org:1172, 1638; e_pot = V_ee + V_en + vnn; with kin, the
// local energy.
__global__ void __launch_bounds__(kDetBlock) k_potential(PotArgs a) {
  __shared__ double sv[kDetWaves];
  const int64_t w = blockIdx.x;
  const int64_t Ne = a.elec_num, P = a.coord_ld;
  const double* x = a.coord + w * Ne;
  double ee = 0.0, en = 0.0;
  for (int64_t q = threadIdx.x; q < Ne * Ne; q += kDetBlock) {
    const int64_t i = q / Ne, j = q % Ne;
    if (i < j) {
      const double dx = x[i] - x[j], dy = x[P + i] - x[P + j], dz = x[2 * P + i] - x[2 * P + j];
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      if (r > 1e-5) ee += 1.0 / r;
    }
  }
  for (int64_t q = threadIdx.x; q < Ne * a.nucl_num; q += kDetBlock) {
    const int64_t i = q / a.nucl_num, A = q % a.nucl_num;
    const double dx = x[i] - a.nucl_coord[A], dy = x[P + i] - a.nucl_coord[a.nucl_num + A],
                 dz = x[2 * P + i] - a.nucl_coord[2 * a.nucl_num + A];
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    if (r > 1e-6) en -= a.charge[A] / r;
  }
  const double vee = block_sum(ee, sv);
  const double ven = block_sum(en, sv);
  if (threadIdx.x == 0) {
    if (a.vee) a.vee[w] = vee;
    if (a.ven) a.ven[w] = ven;
    const double pot = vee + ven + a.vnn;
    if (a.e_pot) a.e_pot[w] = pot;
    if (a.kin_a) {
      const double kin = a.kin_a[w] + (a.kin_b ? a.kin_b[w] : 0.0);
      if (a.e_kin) a.e_kin[w] = kin;
      if (a.e_local) a.e_local[w] = kin + pot;
    }
  }
}

__global__ void k_adjugate(const double* inv, const double* value, int64_t nn, int64_t count, double* out) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < count; q += (int64_t)gridDim.x * blockDim.x)
    out[q] = value[q / nn] * inv[q];
}

// ---- Sherman-Morrison-Woodbury ----------------------------------------------------------
// S: the working inverse [dim][sld]; the rows of the update go to column index - 1.

// C[i] = sum_j S[i][j] u[j]
__device__ void mat_vec(const double* S, int dim, int sld, const double* u, double* C) {
  for (int i = threadIdx.x; i < dim; i += kDetBlock) {
    double s = 0.0;
    for (int j = 0; j < dim; ++j) s += S[i * sld + j] * u[j];
    C[i] = s;
  }
}

// S -= (C x D) / den with D = row `row` of S (staged in Dbuf)
__device__ void rank1(double* S, int dim, int sld, const double* C, double* Dbuf, int row, double iden) {
  for (int j = threadIdx.x; j < dim; j += kDetBlock) Dbuf[j] = S[row * sld + j] * iden;
  __syncthreads();
  for (int i = threadIdx.x >> 6; i < dim; i += kDetWaves) {
    const double ci = C[i];
    for (int j = threadIdx.x & 63; j < dim; j += 64) S[i * sld + j] -= ci * Dbuf[j];
  }
  __syncthreads();
}

// Woodbury update of K columns at once: C = S U (dim x K), B = 1 + V C (K x K, V selecting
// the K updated rows), S -= C B^-1 V S.  Returns false (S untouched) when |det B| < breakdown.
template <int K>
__device__ bool woodbury(const SmwArgs& a, double* S, int sld, const double* upd, const uint64_t* uidx, double* C,
                         double* u, double* Dbuf, double& factor) {
  const int dim = (int)a.dim;
  const int64_t lds = a.lds;
#pragma unroll
  for (int c = 0; c < K; ++c)
    for (int j = threadIdx.x; j < dim; j += kDetBlock) u[c * dim + j] = upd[c * lds + j];
  __syncthreads();
#pragma unroll
  for (int c = 0; c < K; ++c) mat_vec(S, dim, sld, u + c * dim, C + c * dim);
  __syncthreads();
  int row[K];
  double B[K][K], Bi[K][K];
#pragma unroll
  for (int r = 0; r < K; ++r) row[r] = (int)uidx[r] - 1;
#pragma unroll
  for (int r = 0; r < K; ++r)
#pragma unroll
    for (int c = 0; c < K; ++c) B[r][c] = C[c * dim + row[r]] + (r == c ? 1.0 : 0.0);
  double detB;
  if constexpr (K == 2) {
    detB = B[0][0] * B[1][1] - B[0][1] * B[1][0];
    Bi[0][0] = B[1][1];
    Bi[0][1] = -B[0][1];
    Bi[1][0] = -B[1][0];
    Bi[1][1] = B[0][0];
  } else {
    // transpose of the cofactor matrix
    Bi[0][0] = B[1][1] * B[2][2] - B[1][2] * B[2][1];
    Bi[0][1] = B[0][2] * B[2][1] - B[0][1] * B[2][2];
    Bi[0][2] = B[0][1] * B[1][2] - B[0][2] * B[1][1];
    Bi[1][0] = B[1][2] * B[2][0] - B[1][0] * B[2][2];
    Bi[1][1] = B[0][0] * B[2][2] - B[0][2] * B[2][0];
    Bi[1][2] = B[0][2] * B[1][0] - B[0][0] * B[1][2];
    Bi[2][0] = B[1][0] * B[2][1] - B[1][1] * B[2][0];
    Bi[2][1] = B[0][1] * B[2][0] - B[0][0] * B[2][1];
    Bi[2][2] = B[0][0] * B[1][1] - B[0][1] * B[1][0];
    detB = B[0][0] * Bi[0][0] + B[0][1] * Bi[1][0] + B[0][2] * Bi[2][0];
  }
  if (fabs(detB) < a.breakdown) return false;  // uniform: every thread computed the same B
  const double id = 1.0 / detB;
  factor = detB;
  // F = B^-1 (the K rows of S), staged in Dbuf [K][dim]
  for (int j = threadIdx.x; j < dim; j += kDetBlock) {
    double srow[K];
#pragma unroll
    for (int q = 0; q < K; ++q) srow[q] = S[row[q] * sld + j];
#pragma unroll
    for (int r = 0; r < K; ++r) {
      double f = 0.0;
#pragma unroll
      for (int q = 0; q < K; ++q) f += Bi[r][q] * srow[q];
      Dbuf[r * dim + j] = f * id;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x >> 6; i < dim; i += kDetWaves) {
    double ci[K];
#pragma unroll
    for (int r = 0; r < K; ++r) ci[r] = C[r * dim + i];
    for (int j = threadIdx.x & 63; j < dim; j += 64) {
      double x = S[i * sld + j];
#pragma unroll
      for (int r = 0; r < K; ++r) x -= ci[r] * Dbuf[r * dim + j];
      S[i * sld + j] = x;
    }
  }
  __syncthreads();
  return true;
}

template <bool IN_LDS>
__global__ void __launch_bounds__(kDetBlock) k_smw(SmwArgs a) {
  extern __shared__ double smem[];
  const int64_t m = blockIdx.x;
  const int dim = (int)a.dim;
  const int64_t lds = a.lds;
  const int sld = IN_LDS ? (dim | 1) : dim;
  double* S = IN_LDS ? smem : a.work + m * (int64_t)dim * dim;
  double* C = IN_LDS ? smem + dim * sld : smem;  // [3][dim]
  double* u = C + 3 * dim;                         // [3][dim]
  double* Dbuf = u + 3 * dim;                      // [3][dim]
  const double* upd = a.upd + m * a.n_upd * lds;
  const uint64_t* uidx = a.upd_idx + m * a.n_upd;
  double* sinv = a.s_inv + m * dim * lds;
  // every index must name a row of the matrix, else nothing is touched
  for (int64_t l = 0; l < a.n_upd; ++l)
    if (uidx[l] < 1 || uidx[l] > (uint64_t)dim) {
      if (threadIdx.x == 0) a.status[m] = 2;
      return;
    }
  for (int i = threadIdx.x >> 6; i < dim; i += kDetWaves)
    for (int j = threadIdx.x & 63; j < dim; j += 64) S[i * sld + j] = sinv[i * lds + j];
  __syncthreads();
  double factor = 1.0;
  bool ok = true;
  if (a.kind == 3 || a.kind == 4) {
    ok = a.kind == 3 ? woodbury<2>(a, S, sld, upd, uidx, C, u, Dbuf, factor)
                     : woodbury<3>(a, S, sld, upd, uidx, C, u, Dbuf, factor);
  } else {
    // rank-1 updates in turn; kinds 1 and 2 halve an update whose denominator breaks down
    // and queue the other half (kind 2 returns the queue, kind 1 applies it in further passes)
    const int64_t cap = a.n_upd;
    double *q_upd0 = nullptr, *q_upd1 = nullptr;
    uint64_t *q_idx0 = nullptr, *q_idx1 = nullptr;
    if (a.kind == 1) {
      q_upd0 = a.later_upd + m * 2 * cap * lds;
      q_upd1 = q_upd0 + cap * lds;
      q_idx0 = a.later_idx + m * 2 * cap;
      q_idx1 = q_idx0 + cap;
    } else if (a.kind == 2) {
      q_upd0 = a.later_upd + m * a.later_cap * lds;
      q_idx0 = a.later_idx + m * a.later_cap;
    }
    const double* cur_upd = upd;
    const uint64_t* cur_idx = uidx;
    int64_t cur_n = a.n_upd;
    int64_t base = (a.kind == 2) ? (int64_t)a.later_num[m] : 0;  // kind 2 appends after base
    for (int pass = 0; ok && cur_n > 0; ++pass) {
      if (pass == 64) {
        ok = false;
        break;
      }
      const int qb = pass & 1;
      int64_t later = 0;
      for (int64_t l = 0; l < cur_n; ++l) {
        for (int j = threadIdx.x; j < dim; j += kDetBlock) u[j] = cur_upd[l * lds + j];
        __syncthreads();
        mat_vec(S, dim, sld, u, C);
        __syncthreads();
        const int row = (int)cur_idx[l] - 1;
        double den = 1.0 + C[row];
        if (fabs(den) < a.breakdown) {
          if (a.kind == 0) {
            ok = false;
            break;
          }
          double* qu = qb ? q_upd1 : q_upd0;
          uint64_t* qi = qb ? q_idx1 : q_idx0;
          const int64_t slot = (a.kind == 2) ? base + later : later;
          double* dst = qu + slot * lds;
          // the whole row, padding included (the padding of the update, halved)
          for (int j = threadIdx.x; j < lds; j += kDetBlock) dst[j] = 0.5 * (j < dim ? u[j] : cur_upd[l * lds + j]);
          if (threadIdx.x == 0) {
            qi[slot] = cur_idx[l];
          }
          ++later;
          __syncthreads();
          for (int i = threadIdx.x; i < dim; i += kDetBlock) C[i] *= 0.5;
          __syncthreads();
          den = 1.0 + C[row];
        }
        factor *= den;
        rank1(S, dim, sld, C, Dbuf, row, 1.0 / den);
      }
      if (a.kind == 2) {
        if (threadIdx.x == 0) a.later_num[m] = base + later;
        break;
      }
      // the other halves become the next pass's updates (only kind 1 gets here with later > 0)
      __threadfence_block();
      __syncthreads();
      cur_upd = qb ? q_upd1 : q_upd0;
      cur_idx = qb ? q_idx1 : q_idx0;
      cur_n = later;
    }
  }
  if (threadIdx.x == 0) a.status[m] = ok ? 0 : 1;
  if (!ok) return;
  for (int i = threadIdx.x >> 6; i < dim; i += kDetWaves)
    for (int j = threadIdx.x & 63; j < dim; j += 64) sinv[i * lds + j] = S[i * sld + j];
  if (threadIdx.x == 0 && a.det) a.det[m] *= factor;
}

// The adjugate of one n x n matrix through its inverse, B = det(A) A^-1 (the general case
// of qmckl_adjugate, org/qmckl_blas.org:1785-1790 with :2200-2260), column-major A (lda) and
// B (ldb); det(A) to *det_out.  A singular A (a zero pivot) writes det 0 and leaves B.
template <bool IN_LDS>
__global__ void __launch_bounds__(kDetBlock) k_adjugate_gj(const double* A0, int n, int64_t lda, double* B,
                                                           int64_t ldb, double* det_out, double* work) {
  extern __shared__ double smem[];
  const int ld = IN_LDS ? (n | 1) : n;
  double* A = IN_LDS ? smem : work;
  double* rowbuf = IN_LDS ? smem + n * ld : smem;
  double* col = rowbuf + n;
  double* sv = col + n;
  int* si = reinterpret_cast<int*>(sv + kDetWaves);
  int* perm = si + kDetWaves;
  for (int j = threadIdx.x >> 6; j < n; j += kDetWaves)
    for (int i = threadIdx.x & 63; i < n; i += 64) A[i * ld + j] = A0[i + j * lda];
  __syncthreads();
  const double det = gauss_jordan(A, n, ld, rowbuf, col, perm, sv, si);
  if (threadIdx.x == 0) *det_out = det;
  if (det == 0.0) return;
  for (int j = threadIdx.x >> 6; j < n; j += kDetWaves)
    for (int i = threadIdx.x & 63; i < n; i += 64) B[i + j * ldb] = det * A[i * ld + j];
}

}  // namespace

static size_t slater_lds_bytes(int n, bool in_lds) {
  const size_t ld = in_lds ? (size_t)(n | 1) : 0;
  return sizeof(double) * (ld * n + 2 * n + kDetWaves) + sizeof(int) * (kDetWaves + 2 * n);
}

hipError_t launch_slater(const SlaterArgs& a, int64_t n_mat, hipStream_t s) {
  if (n_mat <= 0 || a.n < 0) return hipSuccess;
  const bool in_lds = a.n <= kDetLdsMax;
  const size_t lds = slater_lds_bytes(a.n, in_lds);
  if (in_lds) {
    hipError_t e = hipFuncSetAttribute((const void*)k_slater<true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    k_slater<true><<<dim3((unsigned)n_mat), dim3(kDetBlock), lds, s>>>(a);
  } else {
    if (a.work == nullptr) return hipErrorInvalidValue;
    k_slater<false><<<dim3((unsigned)n_mat), dim3(kDetBlock), lds, s>>>(a);
  }
  return hipGetLastError();
}

hipError_t launch_slater_vgl(const SlaterArgs& a, int64_t n_mat, double* out, hipStream_t s) {
  if (n_mat <= 0 || a.n <= 0) return hipSuccess;
  k_slater_vgl<<<dim3((unsigned)n_mat), dim3(kDetBlock), 0, s>>>(a, out);
  return hipGetLastError();
}

hipError_t launch_adjugate(const double* inv, const double* value, int64_t nn, int64_t count, double* out,
                           hipStream_t s) {
  if (count <= 0) return hipSuccess;
  const int64_t blocks = std::min<int64_t>((count + 255) / 256, 65536);
  k_adjugate<<<dim3((unsigned)blocks), dim3(256), 0, s>>>(inv, value, nn, count, out);
  return hipGetLastError();
}

hipError_t launch_potential(const PotArgs& a, int64_t walk_num, hipStream_t s) {
  if (walk_num <= 0) return hipSuccess;
  k_potential<<<dim3((unsigned)walk_num), dim3(kDetBlock), 0, s>>>(a);
  return hipGetLastError();
}

static size_t smw_lds_bytes(int64_t dim, bool in_lds) {
  const size_t sld = in_lds ? (size_t)(dim | 1) : 0;
  return sizeof(double) * (sld * dim + 9 * dim);
}

hipError_t launch_smw(const SmwArgs& a, hipStream_t s) {
  if (a.n_mat <= 0) return hipSuccess;
  const bool in_lds = a.dim <= kDetLdsMax;
  const size_t lds = smw_lds_bytes(a.dim, in_lds);
  if (in_lds) {
    hipError_t e = hipFuncSetAttribute((const void*)k_smw<true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    k_smw<true><<<dim3((unsigned)a.n_mat), dim3(kDetBlock), lds, s>>>(a);
  } else {
    if (a.work == nullptr) return hipErrorInvalidValue;
    k_smw<false><<<dim3((unsigned)a.n_mat), dim3(kDetBlock), lds, s>>>(a);
  }
  return hipGetLastError();
}

hipError_t launch_adjugate_gj(int n, const double* A, int64_t lda, double* B, int64_t ldb, double* det, double* work,
                              hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const bool in_lds = n <= kDetLdsMax;
  const size_t lds = slater_lds_bytes(n, in_lds);
  if (in_lds) {
    hipError_t e = hipFuncSetAttribute((const void*)k_adjugate_gj<true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    k_adjugate_gj<true><<<dim3(1), dim3(kDetBlock), lds, s>>>(A, n, lda, B, ldb, det, work);
  } else {
    k_adjugate_gj<false><<<dim3(1), dim3(kDetBlock), lds, s>>>(A, n, lda, B, ldb, det, work);
  }
  return hipGetLastError();
}

}  // namespace qmckl_amd
