// AO -> MO contraction on gfx950 matrix cores: C[m][n] = sum_k A[m][k] B[k][n] with
// A = AO VGL viewed as [5*points][ao], B = MO coefficients transposed [ao][mo_pad] and
// C = MO VGL [5*points][mo] (org/qmckl_mo.org:1905-1940, 2015-2160).
//
// One 256-lane workgroup computes a 128 x (16*NT) tile of C; each of its four waves owns
// 32 rows (two 16-row MFMA tiles) across all NT 16-column tiles.  K is walked in chunks of
// 32 staged through LDS.
//   FP64   : v_mfma_f64_16x16x4_f64, FP64 accumulation (the reference's arithmetic).
//   FP32   : v_mfma_f32_16x16x4_f32; the FP32 accumulators are added into FP64 ones after
//            every K chunk of 32.
//   BF16X3 : both operands are held as three bfloat16 terms (hi, mid, lo); per K chunk the
//            six products lo*hi, hi*lo, mid*mid, mid*hi, hi*mid, hi*hi (smallest first) go
//            through v_mfma_f32_16x16x32_bf16 into an FP32 accumulator that is then added
//            into an FP64 one.
//
// MFMA operand layouts (16x16 blocks, 64 lanes, i = lane % 16, g = lane / 16), measured
// on gfx950:
//   f32 16x16x4  : A[i][g], B[g][i]; result register q holds row 4g + q, column i.
//   f64 16x16x4  : A[i][g], B[g][i]; result register q holds row g + 4q, column i.
//   bf16 16x16x32: A[i][8g..8g+7], B[8g..8g+7][i]; register q holds row 4g + q, column i.
#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace qmckl_amd {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef short s8 __attribute__((ext_vector_type(8)));

constexpr int BM = 128, BK = 32;

template <int NT>
__global__ void __launch_bounds__(256) k_gemm_f64(const double* __restrict__ A, int64_t lda,
                                                  const double* __restrict__ B, int64_t ldb, double* __restrict__ C,
                                                  int64_t ldc, int64_t M, int64_t N, int64_t K) {
  constexpr int BN = 16 * NT;
  __shared__ double As[BM][BK + 1];
  __shared__ double Bs[BK][BN];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lg = lane >> 4;
  const int64_t m0 = (int64_t)blockIdx.x * BM, n0 = (int64_t)blockIdx.y * BN;
  d4 acc[2][NT];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < NT; ++c) acc[r][c] = d4{0.0, 0.0, 0.0, 0.0};

  for (int64_t k0 = 0; k0 < K; k0 += BK) {
    {
      const int r = tid >> 1, c0 = (tid & 1) * 16;
      const int64_t gr = m0 + r;
      const double* src = A + gr * lda + k0 + c0;
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int64_t gk = k0 + c0 + j;
        As[r][c0 + j] = (gr < M && gk < K) ? src[j] : 0.0;
      }
    }
    for (int t = tid; t < BK * BN; t += 256) {
      const int kk = t / BN, cc = t - kk * BN;
      const int64_t gk = k0 + kk, gc = n0 + cc;
      Bs[kk][cc] = (gk < K && gc < ldb) ? B[gk * ldb + gc] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < BK / 4; ++ks) {
      const double a0 = As[wave * 32 + li][ks * 4 + lg];
      const double a1 = As[wave * 32 + 16 + li][ks * 4 + lg];
#pragma unroll
      for (int c = 0; c < NT; ++c) {
        const double b = Bs[ks * 4 + lg][c * 16 + li];
        acc[0][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b, acc[0][c], 0, 0, 0);
        acc[1][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b, acc[1][c], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // FP64 16x16x4 result layout: register q of lane (g, i) holds row g + 4q, column i.
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < NT; ++c) {
      const int64_t col = n0 + c * 16 + li;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int64_t row = m0 + wave * 32 + r * 16 + lg + 4 * q;
        if (row < M && col < N) C[row * ldc + col] = acc[r][c][q];
      }
    }
}

template <int NT>
__global__ void __launch_bounds__(256) k_gemm_f32(const float* __restrict__ A, int64_t lda,
                                                  const float* __restrict__ B, int64_t ldb, double* __restrict__ C,
                                                  int64_t ldc, int64_t M, int64_t N, int64_t K) {
  constexpr int BN = 16 * NT;
  __shared__ float As[BM][BK + 1];
  __shared__ float Bs[BK][BN];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lg = lane >> 4;
  const int64_t m0 = (int64_t)blockIdx.x * BM, n0 = (int64_t)blockIdx.y * BN;
  d4 acc[2][NT];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < NT; ++c) acc[r][c] = d4{0.0, 0.0, 0.0, 0.0};

  for (int64_t k0 = 0; k0 < K; k0 += BK) {
    {
      const int r = tid >> 1, c0 = (tid & 1) * 16;
      const int64_t gr = m0 + r;
      const float* src = A + gr * lda + k0 + c0;
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int64_t gk = k0 + c0 + j;
        As[r][c0 + j] = (gr < M && gk < K) ? src[j] : 0.0f;
      }
    }
    for (int t = tid; t < BK * BN; t += 256) {
      const int kk = t / BN, cc = t - kk * BN;
      const int64_t gk = k0 + kk, gc = n0 + cc;
      Bs[kk][cc] = (gk < K && gc < ldb) ? B[gk * ldb + gc] : 0.0f;
    }
    __syncthreads();
    f4 part[2][NT];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int c = 0; c < NT; ++c) part[r][c] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < BK / 4; ++ks) {
      const float a0 = As[wave * 32 + li][ks * 4 + lg];
      const float a1 = As[wave * 32 + 16 + li][ks * 4 + lg];
#pragma unroll
      for (int c = 0; c < NT; ++c) {
        const float b = Bs[ks * 4 + lg][c * 16 + li];
        part[0][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b, part[0][c], 0, 0, 0);
        part[1][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b, part[1][c], 0, 0, 0);
      }
    }
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int c = 0; c < NT; ++c)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[r][c][q] += (double)part[r][c][q];
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < NT; ++c) {
      const int64_t col = n0 + c * 16 + li;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int64_t row = m0 + wave * 32 + r * 16 + 4 * lg + q;
        if (row < M && col < N) C[row * ldc + col] = acc[r][c][q];
      }
    }
}

constexpr int KP = BK + 8;  // padded row length (bf16 elements) of the LDS tiles: 80 bytes

template <int NT>
__global__ void __launch_bounds__(256) k_gemm_bf16x3(const uint16_t* __restrict__ A, int64_t lda, int64_t a_plane,
                                                     const uint16_t* __restrict__ B, int64_t ldb, int64_t b_plane,
                                                     double* __restrict__ C, int64_t ldc, int64_t M, int64_t N,
                                                     int64_t K) {
  constexpr int BN = 16 * NT;
  __shared__ __attribute__((aligned(16))) uint16_t As[3][BM][KP];
  __shared__ __attribute__((aligned(16))) uint16_t Bt[3][BN][KP];  // transposed: [col][k]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lg = lane >> 4;
  const int64_t m0 = (int64_t)blockIdx.x * BM, n0 = (int64_t)blockIdx.y * BN;
  d4 acc[2][NT];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < NT; ++c) acc[r][c] = d4{0.0, 0.0, 0.0, 0.0};

  for (int64_t k0 = 0; k0 < K; k0 += BK) {
    {
      const int r = tid >> 1, c0 = (tid & 1) * 16;
      const int64_t gr = m0 + r;
#pragma unroll
      for (int pl = 0; pl < 3; ++pl) {
        const uint16_t* src = A + pl * a_plane + gr * lda + k0 + c0;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          const int64_t gk = k0 + c0 + j;
          As[pl][r][c0 + j] = (gr < M && gk < K) ? src[j] : (uint16_t)0;
        }
      }
    }
    for (int t = tid; t < BK * BN; t += 256) {
      const int kk = t / BN, cc = t - kk * BN;
      const int64_t gk = k0 + kk, gc = n0 + cc;
      const bool ok = gk < K && gc < ldb;
#pragma unroll
      for (int pl = 0; pl < 3; ++pl) Bt[pl][cc][kk] = ok ? B[pl * b_plane + gk * ldb + gc] : (uint16_t)0;
    }
    __syncthreads();
    bf8 a[3][2];
#pragma unroll
    for (int pl = 0; pl < 3; ++pl)
#pragma unroll
      for (int r = 0; r < 2; ++r)
        a[pl][r] = __builtin_bit_cast(bf8, *reinterpret_cast<const s8*>(&As[pl][wave * 32 + r * 16 + li][8 * lg]));
#pragma unroll
    for (int c = 0; c < NT; ++c) {
      bf8 b[3];
#pragma unroll
      for (int pl = 0; pl < 3; ++pl)
        b[pl] = __builtin_bit_cast(bf8, *reinterpret_cast<const s8*>(&Bt[pl][c * 16 + li][8 * lg]));
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        f4 p = f4{0.f, 0.f, 0.f, 0.f};
        p = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2][r], b[0], p, 0, 0, 0);
        p = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][r], b[2], p, 0, 0, 0);
        p = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][r], b[1], p, 0, 0, 0);
        p = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][r], b[0], p, 0, 0, 0);
        p = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][r], b[1], p, 0, 0, 0);
        p = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][r], b[0], p, 0, 0, 0);
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[r][c][q] += (double)p[q];
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < NT; ++c) {
      const int64_t col = n0 + c * 16 + li;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int64_t row = m0 + wave * 32 + r * 16 + 4 * lg + q;
        if (row < M && col < N) C[row * ldc + col] = acc[r][c][q];
      }
    }
}

// Number of 16-column tiles per workgroup: the whole (padded) N when it fits in 8 tiles.
int tiles_for(int64_t N) {
  const int64_t t = (N + 15) / 16;
  return (int)(t > 8 ? 8 : (t < 1 ? 1 : t));
}

#define DISPATCH_NT(NT_VALUE, KERNEL, ...)                                                       \
  switch (NT_VALUE) {                                                                            \
    case 1: hipLaunchKernelGGL(KERNEL<1>, __VA_ARGS__); break;                                   \
    case 2: hipLaunchKernelGGL(KERNEL<2>, __VA_ARGS__); break;                                   \
    case 3: hipLaunchKernelGGL(KERNEL<3>, __VA_ARGS__); break;                                   \
    case 4: hipLaunchKernelGGL(KERNEL<4>, __VA_ARGS__); break;                                   \
    case 5: hipLaunchKernelGGL(KERNEL<5>, __VA_ARGS__); break;                                   \
    case 6: hipLaunchKernelGGL(KERNEL<6>, __VA_ARGS__); break;                                   \
    case 7: hipLaunchKernelGGL(KERNEL<7>, __VA_ARGS__); break;                                   \
    default: hipLaunchKernelGGL(KERNEL<8>, __VA_ARGS__); break;                                  \
  }

}  // namespace

hipError_t launch_gemm_f64(const double* A, int64_t lda, const double* B, int64_t ldb, double* C, int64_t ldc,
                           int64_t M, int64_t N, int64_t K, hipStream_t s) {
  if (M <= 0 || N <= 0) return hipSuccess;
  const int nt = tiles_for(N);
  const dim3 grid((unsigned)((M + BM - 1) / BM), (unsigned)((N + 16 * nt - 1) / (16 * nt)));
  DISPATCH_NT(nt, k_gemm_f64, grid, dim3(256), 0, s, A, lda, B, ldb, C, ldc, M, N, K);
  return hipGetLastError();
}

hipError_t launch_gemm_f32(const float* A, int64_t lda, const float* B, int64_t ldb, double* C, int64_t ldc,
                           int64_t M, int64_t N, int64_t K, hipStream_t s) {
  if (M <= 0 || N <= 0) return hipSuccess;
  const int nt = tiles_for(N);
  const dim3 grid((unsigned)((M + BM - 1) / BM), (unsigned)((N + 16 * nt - 1) / (16 * nt)));
  DISPATCH_NT(nt, k_gemm_f32, grid, dim3(256), 0, s, A, lda, B, ldb, C, ldc, M, N, K);
  return hipGetLastError();
}

hipError_t launch_gemm_bf16x3(const uint16_t* A, int64_t lda, int64_t a_plane, const uint16_t* B, int64_t ldb,
                              int64_t b_plane, double* C, int64_t ldc, int64_t M, int64_t N, int64_t K,
                              hipStream_t s) {
  if (M <= 0 || N <= 0) return hipSuccess;
  const int nt = tiles_for(N);
  const dim3 grid((unsigned)((M + BM - 1) / BM), (unsigned)((N + 16 * nt - 1) / (16 * nt)));
  DISPATCH_NT(nt, k_gemm_bf16x3, grid, dim3(256), 0, s, A, lda, a_plane, B, ldb, b_plane, C, ldc, M, N, K);
  return hipGetLastError();
}

}  // namespace qmckl_amd
