// Fused AO VGL -> MO VGL on gfx950 (org/qmckl_ao.org:8438-8572 followed by
// org/qmckl_mo.org:1905-1940): the AO matrix never leaves the chip.
//
// A workgroup of 4 waves owns TP = 16 points, i.e. the 80 rows (point, component) of the
// product MO[p][c][:] = sum_k AO[p][c][k] C_t[k][:] (row 5 p + c), and G <= 8 column tiles of
// 16 MOs (grid.y walks the groups of G tiles).  The 16 points are the same electron of 16
// consecutive walkers (AoMoArgs::tw, te), which the workloads draw around the same nucleus,
// so that the screening is coherent across the tile.  The AOs are walked in chunks of
// KC = 32, grouped by the host so that the shells of a chunk vanish at the same places:
//   1. the radial VGL of every shell of the chunk at the 16 points (FP64, screened exactly as
//      the reference) into LDS; when every primitive of every shell is screened at every
//      point, the chunk's AO values are exact zeros and steps 2 and 3 are skipped;
//   2. the 16 x 32 AO VGL values of the chunk (5 components each), assembled in FP64, into
//      an LDS operand tile in the precision of the product;
//   3. each wave multiplies the tile by its column tiles of C_t on matrix cores.
// The per-chunk basis data (nuclei, shells, primitives and AOs of the chunk: the chunk
// descriptor) is one contiguous record copied to LDS one chunk ahead with the gfx950
// global-to-LDS loads, so the loop issues no dependent global loads.  C_t is stored chunk by
// chunk in the operand layout of the precision and read from L2.
// Operand precisions:
//   FP64   v_mfma_f64_16x16x4_f64, FP64 accumulation.
//   FP32   v_mfma_f32_16x16x4_f32; FP32 sums of each chunk added into FP64.
//   BF16X3 both operands as three bfloat16 terms, each rounded to nearest from the FP64
//          remainder of the terms before it (about 26 significant bits per operand); per
//          chunk the six leading cross products through v_mfma_f32_16x16x32_bf16, the
//          leading one (hi x hi) into an FP32 sum of its own and the five smaller ones
//          (smallest first) into a second, so that the FP32 rounding of the large sum is not
//          compounded by five more; both sums are then added into FP64.
// Why one shared tile, not a per-wave layout: the AO evaluation dominates the vector work,
// and evaluating each AO value once per workgroup at 2-3 waves per SIMD beats layouts that
// keep 25 FP64 accumulator tiles per wave in registers (one wave per SIMD) by 15-30 % at
// every precision on the MI355X.
#include <hip/hip_runtime.h>

#include "ao_device.hpp"
#include "internal.hpp"

namespace qmckl_amd {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef short s8 __attribute__((ext_vector_type(8)));
typedef __attribute__((address_space(3))) void lds_void;

constexpr int KC = AO_MO_KC;  // AOs per chunk
constexpr int NWAVE = 4;      // waves per workgroup

// Copies n bytes (a multiple of 16, both addresses 16-byte aligned) from global memory to
// LDS, 16 bytes per lane, the waves of the block taking 1 KiB slices in turn.  Asynchronous:
// the caller waits (s_waitcnt) and synchronises before reading dst.
__device__ __forceinline__ void lds_copy(void* dst, const void* src, int nbytes, int wave, int lane) {
  const char* s = static_cast<const char*>(src);
  char* d = static_cast<char*>(dst);
  for (int base = wave * 1024; base < nbytes; base += NWAVE * 1024) {
    const int t = base + 16 * lane;
    if (t < nbytes) __builtin_amdgcn_global_load_lds(s + t, (lds_void*)(d + base), 16, 0, 0);
  }
}

struct Chunk {  // views of one chunk descriptor in LDS (layout: AoMoArgs)
  const double *nuc, *shl, *prm, *ao;
  int ns, na;
};
__device__ __forceinline__ Chunk chunk_view(const double* d) {
  Chunk c;
  c.ns = (int)d[0];
  c.na = (int)d[1];
  const int np = (int)d[2];
  c.nuc = d + 4;
  c.shl = c.nuc + 4 * c.na;
  c.prm = c.shl + 2 * c.ns;
  c.ao = c.prm + 2 * np;
  return c;
}

// Radial VGL of shell s of the chunk at displacement (x, y, z) from its nucleus, screened as
// shell_radial (org/qmckl_ao.org:4769-4868); e^x by the degree-DEG series of exp_neg_deg.
// Returns whether any primitive passed the screening (else r is exactly zero).
template <int DEG>
__device__ __forceinline__ bool chunk_radial(const Chunk& c, int s, double x, double y, double z, double cutoff,
                                             double r[5]) {
  const double code = c.shl[2 * s];
  const int nl = (int)code & 255, np = (int)code >> 8, p0 = (int)c.shl[2 * s + 1];
  const double r2 = x * x + y * y + z * z;
  // per primitive: e = c e^(-alpha r^2) and t = -2 alpha e; the gradient is (sum t) (x, y, z)
  // and the Laplacian 3 sum t - 2 sum t alpha r^2, applied once per shell
  double v0 = 0, st = 0, st2 = 0;
  bool live = false;
  if (!(r2 > cutoff * c.nuc[4 * nl + 3])) {
    for (int i = 0; i < np; ++i) {
      const double al = c.prm[2 * (p0 + i)];
      const double ar2 = al * r2;
      if (ar2 > cutoff) continue;
      live = true;
      const double e = c.prm[2 * (p0 + i) + 1] * exp_neg_deg<DEG>(-ar2);
      const double two_a = -2.0 * al * e;
      v0 += e;
      st += two_a;
      st2 += two_a * ar2;
    }
  }
  r[0] = v0;
  r[1] = st * x;
  r[2] = st * y;
  r[3] = st * z;
  r[4] = 3.0 * st - 2.0 * st2;
  return live;
}

// AO VGL of an AO with powers (pa, pb, pc) at displacement (x, y, z) from its nucleus, its
// shell's radial VGL at rs; org/qmckl_ao.org:8530-8565, without the normalisation factor
// ao_factor[k], which the chunk operands of C_t carry (row k scaled by it).
__device__ __forceinline__ void chunk_ao_vgl(double x, double y, double z, int pa, int pb, int pc, const double* rs,
                                             double r[5]) {
  double pv[5];
  monomial_vgl(x + 1.e-20, y + 1.e-20, z + 1.e-20, pa, pb, pc, pv);
  const double s0 = rs[0], s1 = rs[1], s2 = rs[2], s3 = rs[3], s4 = rs[4];
  r[0] = pv[0] * s0;
  r[1] = pv[1] * s0 + pv[0] * s1;
  r[2] = pv[2] * s0 + pv[0] * s2;
  r[3] = pv[3] * s0 + pv[0] * s3;
  r[4] = pv[4] * s0 + pv[0] * s4 + 2.0 * (pv[1] * s1 + pv[2] * s2 + pv[3] * s3);
}
// The kernel.  TP = 16 points per workgroup (5 row tiles, point-major rows 5 p + c), KC = 32
// AOs per chunk.  The 5 x G products of a chunk (row tile r, column tile j) go to the waves
// as: wave w takes
// column tiles w and w + 4 (all row tiles) when they exist, except that with G = 5 the tiles
// (r, 4) are spread over the waves by r mod 4 (at most 7 accumulator tiles per wave).
constexpr int TP = 16;
constexpr int ROWS = 5 * TP;   // product rows per workgroup (point-major, row = 5 p + c)
constexpr int RT = ROWS / 16;  // 16-row MFMA tiles
constexpr int SH = 6;          // doubles per staged (point, shell): 5 radial VGL + padding
constexpr int QS = 64 * NWAVE / KC;  // points between the two of a thread in the AO step
static_assert((64 * NWAVE) % KC == 0 && TP % QS == 0, "AO step: whole points per pass");
constexpr int KP16 = AO_MO_BF_LD;
// Row strides of the FP64 / FP32 operand tiles, so that the operand reads A[16 r + li][4 s
// + lg] of a wave cover the 64 four-byte LDS banks evenly.  Doubles: lane (li, lg) starts at
// bank 2 (34 li + lg) = 4 li + 2 lg mod 64, all 32 even banks once in each half-wave (lg =
// 0, 1 and lg = 2, 3); a stride of 36 (8 li + 2 lg) put two lanes of a half-wave on the same
// banks.  Floats: 36 li mod 64 are 16 distinct multiples of 4, + lg fills all 64 banks.
constexpr int LDA64 = KC + 2;
constexpr int LDA32 = KC + 4;

template <int MODE>
struct TileA;
template <>
struct TileA<0> {
  static constexpr size_t bytes = sizeof(double) * ROWS * LDA64;
};
template <>
struct TileA<1> {
  static constexpr size_t bytes = sizeof(float) * ROWS * LDA32;
};
template <>
struct TileA<2> {
  static constexpr size_t bytes = sizeof(uint16_t) * 3 * ROWS * KP16;
};

// six leading cross products of a (3 planes) x b (3 planes) in FP32: the leading one
// (hi x hi) into `hi`, the five smaller ones, smallest first, into `lo`
__device__ __forceinline__ void bf16x3_product(const bf8 (&a)[3], const bf8 (&b)[3], f4& hi, f4& lo) {
  const f4 z = f4{0.f, 0.f, 0.f, 0.f};
  lo = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], b[0], z, 0, 0, 0);
  hi = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[0], z, 0, 0, 0);
  lo = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[2], lo, 0, 0, 0);
  lo = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[1], lo, 0, 0, 0);
  lo = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[0], lo, 0, 0, 0);
  lo = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[1], lo, 0, 0, 0);
}

// Product of the operand tile's row tile r with column tile j of the chunk's C_t, added
// into acc (FP64 for every mode).
template <int MODE, int G>
__device__ __forceinline__ void tile_product(d4& acc, const void* at, const void* bchunk, int r, int j, int li,
                                             int lg) {
  constexpr int NG16 = 16 * G;
  if constexpr (MODE == 0) {
    const double* A = static_cast<const double*>(at);
    const double* B = static_cast<const double*>(bchunk);
#pragma unroll
    for (int ks = 0; ks < KC / 4; ++ks)
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(A[(16 * r + li) * LDA64 + 4 * ks + lg],
                                                 B[(4 * ks + lg) * NG16 + 16 * j + li], acc, 0, 0, 0);
  } else if constexpr (MODE == 1) {
    const float* A = static_cast<const float*>(at);
    const float* B = static_cast<const float*>(bchunk);
    f4 q = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < KC / 4; ++ks)
      q = __builtin_amdgcn_mfma_f32_16x16x4f32(A[(16 * r + li) * LDA32 + 4 * ks + lg],
                                               B[(4 * ks + lg) * NG16 + 16 * j + li], q, 0, 0, 0);
#pragma unroll
    for (int k = 0; k < 4; ++k) acc[k] += (double)q[k];
  } else {
    const uint16_t* A = static_cast<const uint16_t*>(at);
    const uint16_t* B = static_cast<const uint16_t*>(bchunk);
    bf8 a[3], b[3];
#pragma unroll
    for (int pl = 0; pl < 3; ++pl) {
      a[pl] = __builtin_bit_cast(bf8, *reinterpret_cast<const s8*>(A + pl * ROWS * KP16 + (16 * r + li) * KP16 + 8 * lg));
      b[pl] = __builtin_bit_cast(bf8, *reinterpret_cast<const s8*>(B + ((pl * NG16) + 16 * j + li) * KP16 + 8 * lg));
    }
    f4 hi, lo;
    bf16x3_product(a, b, hi, lo);
#pragma unroll
    for (int k = 0; k < 4; ++k) acc[k] += (double)hi[k] + (double)lo[k];
  }
}

template <int MODE, int G>
__global__ void __launch_bounds__(64 * NWAVE, G > 5 ? 2 : (MODE == 0 ? 3 : 4)) k_ao_mo_tile(AoMoArgs A) {
  extern __shared__ __attribute__((aligned(16))) double lds[];
  constexpr int NMAIN = G < NWAVE ? G : NWAVE;  // waves with a first column tile
  const int tid = threadIdx.x, lane = tid & 63, li = lane & 15, lg = lane >> 4;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  // LDS: descriptor buffers [2][desc_max] (addresses computed from `lds`, so that every
  // access compiles to an LDS instruction), then the points, radial staging and tile
  double* xt = lds + 2 * A.desc_max;                 // [TP][3]
  int64_t* pidx = reinterpret_cast<int64_t*>(xt + 3 * TP);  // [TP] point of each tile row
  double* sh = xt + 4 * TP;                          // [TP][ns_max][SH]
  void* at = sh + TP * A.ns_max * SH;                // operand tile (TileA<MODE>)
  const int group = blockIdx.y;
  const int64_t tp0 = A.p0 + (int64_t)blockIdx.x * TP;  // first position p' of the tile
  const int ntp = (int)min((int64_t)TP, A.p0 + A.np - tp0);
  const char* bsrc = static_cast<const char*>(A.B) + (int64_t)group * A.nchunk * A.b_chunk_bytes;
  const bool has1 = wave < NMAIN, has2 = G > 5 && wave + 4 < G;
  if (tid < TP) {
    // position p' = i tw + w holds point p = w te + i
    const int64_t pp = tp0 + tid, i = pp / A.tw;
    pidx[tid] = tid < ntp ? (pp - i * A.tw) * A.te + i : 0;
  }
  __syncthreads();
  for (int t = tid; t < TP * 3; t += 64 * NWAVE) {
    const int q = t / 3, d = t - 3 * q;
    xt[t] = q < ntp ? A.coord[pidx[q] + d * A.coord_ld] : 0.0;
  }
  // column tile `wave` (all row tiles); then column tile wave + 4 (G > 5) or the tiles
  // (wave, 4), (wave + 4, 4) (G = 5)
  d4 acc1[RT], acc2[G > 5 ? RT : 2];
#pragma unroll
  for (int r = 0; r < RT; ++r) acc1[r] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int r = 0; r < (G > 5 ? RT : 2); ++r) acc2[r] = d4{0.0, 0.0, 0.0, 0.0};
  lds_copy(lds, A.desc + A.desc_off[0], 8 * (A.desc_off[1] - A.desc_off[0]), wave, lane);
  for (int ch = 0; ch < A.nchunk; ++ch) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();  // descriptor ch in LDS; every wave done with chunk ch - 1
    if (ch + 1 < A.nchunk)
      lds_copy(lds + ((ch + 1) & 1) * A.desc_max, A.desc + A.desc_off[ch + 1],
               8 * (A.desc_off[ch + 2] - A.desc_off[ch + 1]), wave, lane);
    const Chunk c = chunk_view(lds + (ch & 1) * A.desc_max);
    // 1. radial VGL of the chunk's shells at the 16 points
    int live = 0;
    for (int t = tid; t < TP * c.ns; t += 64 * NWAVE) {
      const int q = t / c.ns, s = t - q * c.ns;
      const int nl = (int)c.shl[2 * s] & 255;
      const double x = xt[3 * q] - c.nuc[4 * nl], y = xt[3 * q + 1] - c.nuc[4 * nl + 1],
                   z = xt[3 * q + 2] - c.nuc[4 * nl + 2];
      double r[5];
      const bool lv = chunk_radial<MODE == 0 ? 13 : 9>(c, s, x, y, z, A.cutoff, r);
      if (q < ntp) live |= (int)lv;
      double* o = sh + (q * A.ns_max + s) * SH;
#pragma unroll
      for (int k = 0; k < 5; ++k) o[k] = r[k];
    }
    // every shell of the chunk screened at every point of the tile: all its AO values are
    // exact zeros and add nothing to the products (rows past the tile's points are zero too)
    if (!__syncthreads_or(live)) continue;
    // 2. AO VGL of the chunk into the operand tile, row = 5 p + c: a thread keeps AO
    //    kk = tid mod KC (its descriptor decoded once per chunk) for the points
    //    q = tid / KC + QS m
    const int kk = tid & (KC - 1);
    const uint32_t au = (uint32_t)c.ao[2 * kk];
    const int ash = au & 255, apw = au >> 8, anl = (int)c.shl[2 * ash] & 255;
    const int pa = apw & 0xff, pb = (apw >> 8) & 0xff, pc = (apw >> 16) & 0xff;
    const double nxa = c.nuc[4 * anl], nya = c.nuc[4 * anl + 1], nza = c.nuc[4 * anl + 2];
    for (int q = tid / KC; q < TP; q += QS) {
      double r[5];
      chunk_ao_vgl(xt[3 * q] - nxa, xt[3 * q + 1] - nya, xt[3 * q + 2] - nza, pa, pb, pc,
                   sh + (q * A.ns_max + ash) * SH, r);
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        if constexpr (MODE == 0) {
          static_cast<double*>(at)[(5 * q + k) * LDA64 + kk] = r[k];
        } else if constexpr (MODE == 1) {
          static_cast<float*>(at)[(5 * q + k) * LDA32 + kk] = (float)r[k];
        } else {
          uint16_t h, m, l;
          split3(r[k], h, m, l);
          uint16_t* T = static_cast<uint16_t*>(at);
          const int idx = (5 * q + k) * KP16 + kk;
          T[idx] = h;
          T[ROWS * KP16 + idx] = m;
          T[2 * ROWS * KP16 + idx] = l;
        }
      }
    }
    __syncthreads();
    // 3. products of the chunk on matrix cores (C_t read from global memory, where the
    //    chunk's rows stay in L2 for every workgroup)
    const void* bch = bsrc + (int64_t)ch * A.b_chunk_bytes;
    if (has1)
#pragma unroll
      for (int r = 0; r < RT; ++r) tile_product<MODE, G>(acc1[r], at, bch, r, wave, li, lg);
    if constexpr (G > 5) {
      if (has2)
#pragma unroll
        for (int r = 0; r < RT; ++r) tile_product<MODE, G>(acc2[r], at, bch, r, wave + 4, li, lg);
    } else if constexpr (G == 5) {
#pragma unroll
      for (int x = 0; x < 2; ++x)
        if (wave + 4 * x < RT) tile_product<MODE, G>(acc2[x], at, bch, wave + 4 * x, 4, li, lg);
    }
  }
  // result rows: f64 MFMA register k holds row lg + 4 k of the tile, f32 (and bf16) 4 lg + k
  auto store = [&](const d4& v, int r, int j) {
    const int64_t col = (int64_t)(group * G + j) * 16 + li;
    if (col >= A.N) return;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int row = 16 * r + (MODE == 0 ? lg + 4 * k : 4 * lg + k);
      const int q = row / 5;
      if (row < 5 * ntp) A.C[(pidx[q] * 5 + (row - 5 * q)) * A.ldc + col] = v[k];
    }
  };
  if (has1)
#pragma unroll
    for (int r = 0; r < RT; ++r) store(acc1[r], r, wave);
  if constexpr (G > 5) {
    if (has2)
#pragma unroll
      for (int r = 0; r < RT; ++r) store(acc2[r], r, wave + 4);
  } else if constexpr (G == 5) {
#pragma unroll
    for (int x = 0; x < 2; ++x)
      if (wave + 4 * x < RT) store(acc2[x], wave + 4 * x, 4);
  }
}

template <class K>
hipError_t launch_kernel(K kernel, const AoMoArgs& a, dim3 grid, dim3 block, size_t lds, hipStream_t s) {
  hipLaunchKernelGGL(kernel, grid, block, lds, s, a);
  return hipGetLastError();
}

template <int MODE, int G>
hipError_t launch_g(const AoMoArgs& a, size_t lds, hipStream_t s) {
  // dynamic LDS up to the CU's 160 KiB less the kernel's static LDS (the block-wide OR of
  // the chunk screening uses a little)
  static size_t max_dyn = 0;
  if (max_dyn == 0) {
    const void* k = reinterpret_cast<const void*>(&k_ao_mo_tile<MODE, G>);
    hipFuncAttributes fa;
    hipError_t e = hipFuncGetAttributes(&fa, k);
    if (e != hipSuccess) return e;
    const size_t dyn = 160 * 1024 - fa.sharedSizeBytes;
    e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
    if (e != hipSuccess) return e;
    max_dyn = dyn;
  }
  if (lds > max_dyn) return hipErrorInvalidValue;
  const dim3 grid((unsigned)((a.np + TP - 1) / TP), (unsigned)a.ngroup);
  return launch_kernel(k_ao_mo_tile<MODE, G>, a, grid, dim3(64 * NWAVE), lds, s);
}

template <int MODE>
hipError_t launch_mode(const AoMoArgs& a, size_t lds, hipStream_t s) {
  switch (a.G) {
    case 1: return launch_g<MODE, 1>(a, lds, s);
    case 2: return launch_g<MODE, 2>(a, lds, s);
    case 3: return launch_g<MODE, 3>(a, lds, s);
    case 4: return launch_g<MODE, 4>(a, lds, s);
    case 5: return launch_g<MODE, 5>(a, lds, s);
    case 6: return launch_g<MODE, 6>(a, lds, s);
    case 7: return launch_g<MODE, 7>(a, lds, s);
    case 8: return launch_g<MODE, 8>(a, lds, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

size_t ao_mo_lds_bytes(const AoMoArgs& a, int mode) {
  return sizeof(double) * (2 * (size_t)a.desc_max + 4 * TP + (size_t)TP * a.ns_max * SH) +
         (mode == 0 ? TileA<0>::bytes : mode == 1 ? TileA<1>::bytes : TileA<2>::bytes);
}

hipError_t launch_ao_mo(const AoMoArgs& a, int mode, hipStream_t s) {
  if (a.np <= 0) return hipSuccess;
  if (a.G < 1 || a.G > AO_MO_MAX_G || a.nchunk < 1 || a.ns_max < 1 || a.ns_max > KC || a.desc_max % 2 ||
      a.b_chunk_bytes % 16 || a.tw < 1 || a.te < 1 || a.p0 < 0 || a.p0 + a.np > a.tw * a.te)
    return hipErrorInvalidValue;
  const size_t lds = ao_mo_lds_bytes(a, mode);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  if (mode == 0) return launch_mode<0>(a, lds, s);
  if (mode == 1) return launch_mode<1>(a, lds, s);
  return launch_mode<2>(a, lds, s);
}

}  // namespace qmckl_amd
