// Fused AO VGL -> MO VGL on gfx950 (org/qmckl_ao.org:8438-8572 followed by
// org/qmckl_mo.org:1905-1940): the AO matrix never leaves the chip.
//
// A workgroup owns TP = 16 points, i.e. the 80 rows (point, component) of the product
// MO[p][c][:] = sum_k AO[p][c][k] C_t[k][:], and one 64-lane wave per 16-column tile of
// MOs.  The AOs are walked in chunks of KC = 32:
//   1. the radial VGL of every shell touching the chunk is evaluated for the 16 points
//      (FP64, screened exactly as the reference) into LDS;
//   2. the 16 x 32 AO VGL values of the chunk (5 components each) are assembled in FP64
//      and stored into an LDS tile in the operand precision of the product;
//   3. each wave multiplies the 80 x 32 tile by its 32 x 16 slice of C_t on matrix cores.
// Operand precisions:
//   FP64   v_mfma_f64_16x16x4_f64, FP64 accumulation.
//   FP32   v_mfma_f32_16x16x4_f32; FP32 partial sums of each chunk added into FP64.
//   BF16X3 both operands as three bfloat16 terms (the AO tile by an exact truncation split
//          of float(x), C_t by a rounded split of the FP64 value); per chunk the six
//          leading cross products (smallest first) through v_mfma_f32_16x16x32_bf16 into
//          FP32, then added into FP64.
// A workgroup has 256 lanes; wave w handles the column tiles w, w+4 of its block.
// C_t is read from global memory (it stays in L2): [Kpad][Npad] for FP64/FP32 and three
// transposed bfloat16 planes [3][Npad][Kpad] for BF16X3, zero-padded, so no guards.
#include <hip/hip_runtime.h>

#include "ao_device.hpp"
#include "internal.hpp"

namespace qmckl_amd {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef short s8 __attribute__((ext_vector_type(8)));

constexpr int TP = 16;        // points per workgroup
constexpr int KC = 32;        // AOs per chunk
constexpr int ROWS = 5 * TP;  // product rows per workgroup
constexpr int RT = ROWS / 16; // 16-row MFMA tiles
constexpr int SH = 8;         // doubles per staged (point, shell): 5 radial VGL + dx, dy, dz
constexpr int CT = 2;         // column tiles per wave (8 per workgroup)
constexpr int KP16 = KC + 8;  // padded bf16 row (80 bytes, 16-byte aligned)
// Row strides of the FP64 / FP32 operand tiles, chosen so that the MFMA operand reads
// A[16 r + li][4 ks + lg] of a wave hit distinct LDS banks: rows 34 doubles apart
// (2 mod 32) and 36 floats apart (4 mod 64).
constexpr int LD64 = KC + 2;
constexpr int LD32 = KC + 4;

template <int MODE>
struct Tile;
template <>
struct Tile<0> {
  static constexpr size_t bytes = sizeof(double) * ROWS * LD64;
};
template <>
struct Tile<1> {
  static constexpr size_t bytes = sizeof(float) * ROWS * LD32;
};
template <>
struct Tile<2> {
  static constexpr size_t bytes = sizeof(uint16_t) * 3 * ROWS * KP16;
};

template <int MODE>
__global__ void __launch_bounds__(256) k_ao_mo(AoDevView v, const double* __restrict__ coord, int64_t coord_ld,
                                               int64_t p0, int64_t np, const void* __restrict__ Bv, int64_t Kpad,
                                               int64_t Npad, double* __restrict__ C, int64_t ldc, int64_t N,
                                               int ns_max) {
  extern __shared__ double lds[];
  double* xt = lds;                 // [TP][3]
  double* sh = xt + 3 * TP;         // [TP][ns_max][SH]
  void* at = sh + TP * ns_max * SH;  // operand tile
  const int Nao = (int)v.ao_num;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nthreads = blockDim.x;
  const int li = lane & 15, lg = lane >> 4;
  const int64_t tp0 = p0 + (int64_t)blockIdx.x * TP;  // first point of the tile
  const int ntp = (int)min((int64_t)TP, p0 + np - tp0);
  int ctile[CT];
  bool has[CT];
#pragma unroll
  for (int j = 0; j < CT; ++j) {
    ctile[j] = blockIdx.y * 8 + wave + 4 * j;  // MO column tiles of this wave
    has[j] = ctile[j] * 16 < Npad;
  }
  const bool has_cols = has[0];

  for (int t = tid; t < TP * 3; t += nthreads) {
    const int p = t / 3, d = t - 3 * p;
    xt[t] = p < ntp ? coord[tp0 + p + d * coord_ld] : 0.0;
  }
  __syncthreads();

  d4 acc[CT][RT];
#pragma unroll
  for (int j = 0; j < CT; ++j)
#pragma unroll
    for (int r = 0; r < RT; ++r) acc[j][r] = d4{0.0, 0.0, 0.0, 0.0};

  for (int k0 = 0; k0 < Kpad; k0 += KC) {
    // 1. radial parts of the shells of AOs k0 .. k0+KC-1
    const int kend = min(k0 + KC, Nao);
    const int s_lo = v.ao_shell[k0];
    const int ns = v.ao_shell[kend - 1] - s_lo + 1;
    for (int t = tid; t < TP * ns; t += nthreads) {
      const int p = t / ns, sl = t - p * ns;
      const int s = s_lo + sl;
      const int a = v.shell_nucl[s];
      const double x = xt[3 * p] - v.nucl_coord[a];
      const double y = xt[3 * p + 1] - v.nucl_coord[a + v.nucl_num];
      const double z = xt[3 * p + 2] - v.nucl_coord[a + 2 * v.nucl_num];
      double r[5];
      shell_radial(v, s, a, x, y, z, x * x + y * y + z * z, r);
      double* o = sh + (p * ns_max + sl) * SH;
#pragma unroll
      for (int c = 0; c < 5; ++c) o[c] = r[c];
      o[5] = x;
      o[6] = y;
      o[7] = z;
    }
    __syncthreads();
    // 2. AO VGL of the chunk into the operand tile, row = 5 p + c
    for (int t = tid; t < TP * KC; t += nthreads) {
      const int p = t / KC, kk = t - p * KC;
      const int k = k0 + kk;
      double r[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      if (p < ntp && k < Nao) {
        const double* o = sh + (p * ns_max + (v.ao_shell[k] - s_lo)) * SH;
        ao_from_shell(v, k, o[5], o[6], o[7], o, r);
      }
      if (MODE == 0) {
        double* A = static_cast<double*>(at);
#pragma unroll
        for (int c = 0; c < 5; ++c) A[(5 * p + c) * LD64 + kk] = r[c];
      } else if (MODE == 1) {
        float* A = static_cast<float*>(at);
#pragma unroll
        for (int c = 0; c < 5; ++c) A[(5 * p + c) * LD32 + kk] = (float)r[c];
      } else {
        uint16_t* A = static_cast<uint16_t*>(at);
#pragma unroll
        for (int c = 0; c < 5; ++c) {
          uint16_t h, m, l;
          split3_trunc(r[c], h, m, l);
          const int idx = (5 * p + c) * KP16 + kk;
          A[idx] = h;
          A[ROWS * KP16 + idx] = m;
          A[2 * ROWS * KP16 + idx] = l;
        }
      }
    }
    __syncthreads();
    // 3. product of the chunk on matrix cores
#pragma unroll
    for (int j = 0; j < CT; ++j) {
      if (!has[j]) continue;
      const int n = ctile[j] * 16 + li;
      if (MODE == 0) {
        const double* B = static_cast<const double*>(Bv);
        const double* A = static_cast<const double*>(at);
#pragma unroll
        for (int ks = 0; ks < KC / 4; ++ks) {
          const double b = B[(int64_t)(k0 + 4 * ks + lg) * Npad + n];
#pragma unroll
          for (int r = 0; r < RT; ++r) {
            const double a = A[(16 * r + li) * LD64 + 4 * ks + lg];
            acc[j][r] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[j][r], 0, 0, 0);
          }
        }
      } else if (MODE == 1) {
        const float* B = static_cast<const float*>(Bv);
        const float* A = static_cast<const float*>(at);
        f4 part[RT];
#pragma unroll
        for (int r = 0; r < RT; ++r) part[r] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < KC / 4; ++ks) {
          const float b = B[(int64_t)(k0 + 4 * ks + lg) * Npad + n];
#pragma unroll
          for (int r = 0; r < RT; ++r) {
            const float a = A[(16 * r + li) * LD32 + 4 * ks + lg];
            part[r] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, part[r], 0, 0, 0);
          }
        }
#pragma unroll
        for (int r = 0; r < RT; ++r)
#pragma unroll
          for (int q = 0; q < 4; ++q) acc[j][r][q] += (double)part[r][q];
      } else {
        const uint16_t* B = static_cast<const uint16_t*>(Bv);
        const uint16_t* A = static_cast<const uint16_t*>(at);
        const int64_t plane = Npad * Kpad;
        bf8 b[3];
#pragma unroll
        for (int pl = 0; pl < 3; ++pl)
          b[pl] = __builtin_bit_cast(bf8, *reinterpret_cast<const s8*>(B + pl * plane + (int64_t)n * Kpad + k0 + 8 * lg));
#pragma unroll
        for (int r = 0; r < RT; ++r) {
          bf8 a[3];
#pragma unroll
          for (int pl = 0; pl < 3; ++pl)
            a[pl] = __builtin_bit_cast(bf8, *reinterpret_cast<const s8*>(A + pl * ROWS * KP16 + (16 * r + li) * KP16 + 8 * lg));
          f4 p = f4{0.f, 0.f, 0.f, 0.f};
          p = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], b[0], p, 0, 0, 0);
          p = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[2], p, 0, 0, 0);
          p = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[1], p, 0, 0, 0);
          p = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[0], p, 0, 0, 0);
          p = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[1], p, 0, 0, 0);
          p = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[0], p, 0, 0, 0);
#pragma unroll
          for (int q = 0; q < 4; ++q) acc[j][r][q] += (double)p[q];
        }
      }
    }
  }
  if (!has_cols) return;
#pragma unroll
  for (int j = 0; j < CT; ++j) {
    const int64_t col = (int64_t)ctile[j] * 16 + li;
    if (!has[j] || col >= N) continue;
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        // result rows: f64 MFMA register q holds row g + 4q, f32 (and bf16) row 4g + q
        const int row = 16 * r + (MODE == 0 ? lg + 4 * q : 4 * lg + q);
        if (row < 5 * ntp) C[(tp0 * 5 + row) * ldc + col] = acc[j][r][q];
      }
  }
}

}  // namespace

size_t ao_mo_lds_bytes(int mode, int ns_max) {
  const size_t base = sizeof(double) * (3 * TP + TP * ns_max * SH);
  return base + (mode == 0 ? Tile<0>::bytes : mode == 1 ? Tile<1>::bytes : Tile<2>::bytes);
}

hipError_t launch_ao_mo(const AoDevView& v, const double* coord, int64_t coord_ld, int64_t p0, int64_t np, int mode,
                        const void* B, int64_t Kpad, int64_t Npad, double* C, int64_t ldc, int64_t N, int ns_max,
                        hipStream_t s) {
  if (np <= 0) return hipSuccess;
  if (ns_max < 1 || ns_max > KC) return hipErrorInvalidValue;
  const int nct = (int)(Npad / 16);
  const dim3 grid((unsigned)((np + TP - 1) / TP), (unsigned)((nct + 7) / 8));
  const dim3 block(256);
  const size_t lds = ao_mo_lds_bytes(mode, ns_max);
  if (mode == 0)
    hipLaunchKernelGGL(k_ao_mo<0>, grid, block, lds, s, v, coord, coord_ld, p0, np, B, Kpad, Npad, C, ldc, N, ns_max);
  else if (mode == 1)
    hipLaunchKernelGGL(k_ao_mo<1>, grid, block, lds, s, v, coord, coord_ld, p0, np, B, Kpad, Npad, C, ldc, N, ns_max);
  else
    hipLaunchKernelGGL(k_ao_mo<2>, grid, block, lds, s, v, coord, coord_ld, p0, np, B, Kpad, Npad, C, ldc, N, ns_max);
  return hipGetLastError();
}

}  // namespace qmckl_amd
