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

#include <algorithm>
#include <map>
#include <tuple>
#include <vector>

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

// Copies n bytes (a multiple of 16, both addresses 16-byte aligned) from global memory to
// LDS, 16 bytes per lane, the waves of the block taking 1 KiB slices in turn.  Asynchronous:
// the caller waits (s_waitcnt) and synchronises before reading dst.
template <int NW>
__device__ __forceinline__ void lds_copy(void* dst, const void* src, int nbytes, int wave, int lane) {
  const char* s = static_cast<const char*>(src);
  char* d = static_cast<char*>(dst);
  for (int base = wave * 1024; base < nbytes; base += NW * 1024) {
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
template <int DEG, bool SL>
__device__ __forceinline__ bool chunk_radial(const Chunk& c, int s, double x, double y, double z, double cutoff,
                                             double r[5]) {
  const double code = c.shl[2 * s];
  const int pcode = (int)c.shl[2 * s + 1];
  const int nl = (int)code & 255, np = (int)code >> 8, p0 = pcode & 0xffff;
  const double r2 = x * x + y * y + z * z;
  if constexpr (SL) {
    // Slater shells, org/qmckl_ao.org:4930-5040, with the nucleus screening of the AO
    // assembly (:8515, r^2 > cutoff * range) as well as the shell's own (r > cutoff * range)
    const int n = (pcode >> 16) - 128;
    const double range = c.nuc[4 * nl + 3], rr = sqrt(r2);
    double v0 = 0.0, g = 0.0, lap = 0.0, rx = 0.0;
    bool live = false;
    if (!(r2 > cutoff * range) && !(rr > cutoff * range)) {
      if (rr < 1.e-15) {
        if (n == 0)
          for (int i = 0; i < np; ++i) v0 += c.prm[2 * (p0 + i) + 1];
        live = n == 0 && np > 0;
      } else {
        rx = 1.0 / rr;
        for (int i = 0; i < np; ++i) {
          const double al = c.prm[2 * (p0 + i)];
          live = live || !(al * rr > cutoff);
          slater_prim<DEG>(al, c.prm[2 * (p0 + i) + 1], n, rr, rx, cutoff, v0, g, lap);
        }
      }
    }
    r[0] = v0;
    r[1] = g * x * rx;
    r[2] = g * y * rx;
    r[3] = g * z * rx;
    r[4] = lap;
    return live;
  }
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
// The value alone of chunk_ao_vgl (the same products in the same order).
__device__ __forceinline__ double chunk_ao_value(double x, double y, double z, int pa, int pb, int pc,
                                                 const double* rs) {
  double pv[5];
  monomial_vgl(x + 1.e-20, y + 1.e-20, z + 1.e-20, pa, pb, pc, pv);
  return pv[0] * rs[0];
}
// The kernel.  TP = 16 points per workgroup (5 row tiles, point-major rows 5 p + c), KC = 32
// AOs per chunk.  The 5 x G products of a chunk (row tile r, column tile j) go to the waves
// as: wave w takes
// column tiles w and w + 4 (all row tiles) when they exist, except that with G = 5 the tiles
// (r, 4) are spread over the waves by r mod 4 (at most 7 accumulator tiles per wave).
constexpr int TP = 16;
constexpr int ROWS = 5 * TP;   // product rows per workgroup (point-major, row = 5 p + c)
constexpr int SH = 6;          // doubles per staged (point, shell): 5 radial VGL + padding
// points between those of one thread in the AO step, with NW waves per workgroup
template <int NW>
constexpr int qstep() {
  static_assert((64 * NW) % KC == 0 && TP % (64 * NW / KC) == 0, "AO step: whole points per pass");
  return 64 * NW / KC;
}
constexpr int KP16 = AO_MO_BF_LD;
// Row strides of the FP64 / FP32 operand tiles, so that the operand reads A[16 r + li][4 s
// + lg] of a wave cover the 64 four-byte LDS banks evenly.  Doubles: lane (li, lg) starts at
// bank 2 (34 li + lg) = 4 li + 2 lg mod 64, all 32 even banks once in each half-wave (lg =
// 0, 1 and lg = 2, 3); a stride of 36 (8 li + 2 lg) put two lanes of a half-wave on the same
// banks.  Floats: 36 li mod 64 are 16 distinct multiples of 4, + lg fills all 64 banks.
constexpr int LDA64 = KC + 2;
constexpr int LDA32 = KC + 4;

// Bytes of the operand tile of NR rows (ROWS, or TP for the values alone)
template <int MODE, int NR>
constexpr size_t tile_bytes() {
  return MODE == 0 ? sizeof(double) * NR * LDA64 : MODE == 1 ? sizeof(float) * NR * LDA32 : sizeof(uint16_t) * 3 * NR * KP16;
}

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

// Product of row tile r of the operand tile (NR rows) with column tile j of the chunk's C_t,
// added into acc (FP64 for every mode).
template <int MODE, int G, int NR>
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
      a[pl] = __builtin_bit_cast(bf8, *reinterpret_cast<const s8*>(A + pl * NR * KP16 + (16 * r + li) * KP16 + 8 * lg));
      b[pl] = __builtin_bit_cast(bf8, *reinterpret_cast<const s8*>(B + ((pl * NG16) + 16 * j + li) * KP16 + 8 * lg));
    }
    f4 hi, lo;
    bf16x3_product(a, b, hi, lo);
#pragma unroll
    for (int k = 0; k < 4; ++k) acc[k] += (double)hi[k] + (double)lo[k];
  }
}

// Column tiles of a wave: NW = 4 waves take column tiles w and w + 4 when they exist (G <= 8),
// except that with G = 5 the tiles (r, 4) are spread over the waves by r mod 4 (at most 7
// accumulator tiles per wave); NW = 8 waves take column tiles w and w + 8 (G <= 16), so that
// up to 256 MOs need one pass over the AOs.
template <int G, int NW>
constexpr bool second_tile() {
  static_assert((NW == 4 && G >= 1 && G <= 8) || (NW == 8 && G >= 1 && G <= 16), "column tiles per pass");
  return NW == 4 ? G > 5 : G > 8;
}

// Evaluates point tile `tile` x column group `group` of the product described by A (SL:
// Slater shells, a compile-time branch so that the Gaussian kernels keep their registers).
// VO: the MO values alone (org/qmckl_mo.org:988-1100): the radial and AO values without
// their derivatives, a tile of 16 rows (one per point) and one row tile per column tile,
// results [point][mo]; each value is the same sum of the same products as the value row of
// the full evaluation, so both agree bit for bit.
template <int MODE, int G, int NW, bool SL, bool VO>
__device__ __forceinline__ void ao_mo_tile_body(const AoMoArgs& A, const int tile, const int group, double* lds) {
  constexpr int NC = VO ? 1 : 5;          // components per point
  constexpr int NR = NC * TP;             // operand tile rows (row = NC p + c)
  constexpr int RT = NR / 16;             // 16-row MFMA tiles
  constexpr int NMAIN = G < NW ? G : NW;  // waves with a first column tile
  constexpr bool TWO = second_tile<G, NW>();
  constexpr int QS = qstep<NW>();
  const int tid = threadIdx.x, lane = tid & 63, li = lane & 15, lg = lane >> 4;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  // LDS: descriptor buffers [2][desc_max] (addresses computed from `lds`, so that every
  // access compiles to an LDS instruction), then the points, radial staging and tile
  double* xt = lds + 2 * A.desc_max;                 // [TP][3]
  int64_t* pidx = reinterpret_cast<int64_t*>(xt + 3 * TP);  // [TP] point of each tile row
  double* sh = xt + 4 * TP;                          // [TP][ns_max][SH]
  void* at = sh + TP * A.ns_max * SH;                // operand tile (tile_bytes<MODE, NR>)
  const int64_t tp0 = A.p0 + (int64_t)tile * TP;  // first position p' of the tile
  const int ntp = (int)min((int64_t)TP, A.p0 + A.np - tp0);
  const char* bsrc = static_cast<const char*>(A.B) + (int64_t)group * A.nchunk * A.b_chunk_bytes;
  const bool has1 = wave < NMAIN, has2 = TWO && wave + NW < G;
  if (tid < TP) {
    // position p' = i tw + w holds point p = w te + i
    const int64_t pp = tp0 + tid, i = pp / A.tw;
    pidx[tid] = tid < ntp ? (pp - i * A.tw) * A.te + i : 0;
  }
  __syncthreads();
  for (int t = tid; t < TP * 3; t += 64 * NW) {
    const int q = t / 3, d = t - 3 * q;
    xt[t] = q < ntp ? A.coord[pidx[q] + d * A.coord_ld] : 0.0;
  }
  // column tile `wave` (all row tiles); then column tile wave + NW (TWO) or the tiles
  // (wave, 4), (wave + 4, 4) (NW = 4, G = 5)
  d4 acc1[RT], acc2[TWO ? RT : 2];
#pragma unroll
  for (int r = 0; r < RT; ++r) acc1[r] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int r = 0; r < (TWO ? RT : 2); ++r) acc2[r] = d4{0.0, 0.0, 0.0, 0.0};
  lds_copy<NW>(lds, A.desc + A.desc_off[0], 8 * (A.desc_off[1] - A.desc_off[0]), wave, lane);
  for (int ch = 0; ch < A.nchunk; ++ch) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();  // descriptor ch in LDS; every wave done with chunk ch - 1
    if (ch + 1 < A.nchunk)
      lds_copy<NW>(lds + ((ch + 1) & 1) * A.desc_max, A.desc + A.desc_off[ch + 1],
               8 * (A.desc_off[ch + 2] - A.desc_off[ch + 1]), wave, lane);
    const Chunk c = chunk_view(lds + (ch & 1) * A.desc_max);
    // 1. radial VGL of the chunk's shells at the 16 points
    int live = 0;
    for (int t = tid; t < TP * c.ns; t += 64 * NW) {
      const int q = t / c.ns, s = t - q * c.ns;
      const int nl = (int)c.shl[2 * s] & 255;
      const double x = xt[3 * q] - c.nuc[4 * nl], y = xt[3 * q + 1] - c.nuc[4 * nl + 1],
                   z = xt[3 * q + 2] - c.nuc[4 * nl + 2];
      double r[5];
      const bool lv = chunk_radial<MODE == 0 ? 13 : 9, SL>(c, s, x, y, z, A.cutoff, r);
      if (q < ntp) live |= (int)lv;
      double* o = sh + (q * A.ns_max + s) * SH;
#pragma unroll
      for (int k = 0; k < NC; ++k) o[k] = r[k];
    }
    // every shell of the chunk screened at every point of the tile: all its AO values are
    // exact zeros and add nothing to the products (rows past the tile's points are zero too)
    if (!__syncthreads_or(live)) continue;
    // 2. AO VGL of the chunk into the operand tile, row = NC p + c: a thread keeps AO
    //    kk = tid mod KC (its descriptor decoded once per chunk) for the points
    //    q = tid / KC + QS m
    const int kk = tid & (KC - 1);
    const uint32_t au = (uint32_t)c.ao[2 * kk];
    const int ash = au & 255, apw = au >> 8, anl = (int)c.shl[2 * ash] & 255;
    const int pa = apw & 0xff, pb = (apw >> 8) & 0xff, pc = (apw >> 16) & 0xff;
    const double nxa = c.nuc[4 * anl], nya = c.nuc[4 * anl + 1], nza = c.nuc[4 * anl + 2];
    for (int q = tid / KC; q < TP; q += QS) {
      double r[5];
      if constexpr (VO)
        r[0] = chunk_ao_value(xt[3 * q] - nxa, xt[3 * q + 1] - nya, xt[3 * q + 2] - nza, pa, pb, pc,
                              sh + (q * A.ns_max + ash) * SH);
      else
        chunk_ao_vgl(xt[3 * q] - nxa, xt[3 * q + 1] - nya, xt[3 * q + 2] - nza, pa, pb, pc,
                     sh + (q * A.ns_max + ash) * SH, r);
#pragma unroll
      for (int k = 0; k < NC; ++k) {
        if constexpr (MODE == 0) {
          static_cast<double*>(at)[(NC * q + k) * LDA64 + kk] = r[k];
        } else if constexpr (MODE == 1) {
          static_cast<float*>(at)[(NC * q + k) * LDA32 + kk] = (float)r[k];
        } else {
          uint16_t h, m, l;
          split3(r[k], h, m, l);
          uint16_t* T = static_cast<uint16_t*>(at);
          const int idx = (NC * q + k) * KP16 + kk;
          T[idx] = h;
          T[NR * KP16 + idx] = m;
          T[2 * NR * KP16 + idx] = l;
        }
      }
    }
    __syncthreads();
    // 3. products of the chunk on matrix cores (C_t read from global memory, where the
    //    chunk's rows stay in L2 for every workgroup)
    const void* bch = bsrc + (int64_t)ch * A.b_chunk_bytes;
    if (has1)
#pragma unroll
      for (int r = 0; r < RT; ++r) tile_product<MODE, G, NR>(acc1[r], at, bch, r, wave, li, lg);
    if constexpr (TWO) {
      if (has2)
#pragma unroll
        for (int r = 0; r < RT; ++r) tile_product<MODE, G, NR>(acc2[r], at, bch, r, wave + NW, li, lg);
    } else if constexpr (NW == 4 && G == 5) {
#pragma unroll
      for (int x = 0; x < 2; ++x)
        if (wave + 4 * x < RT) tile_product<MODE, G, NR>(acc2[x], at, bch, wave + 4 * x, 4, li, lg);
    }
  }
  // result rows: f64 MFMA register k holds row lg + 4 k of the tile, f32 (and bf16) 4 lg + k
  auto store = [&](const d4& v, int r, int j) {
    const int64_t col = (int64_t)(group * G + j) * 16 + li;
    if (col >= A.N) return;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int row = 16 * r + (MODE == 0 ? lg + 4 * k : 4 * lg + k);
      const int q = row / NC;
      if (row < NC * ntp) A.C[(pidx[q] * NC + (row - NC * q)) * A.ldc + col] = v[k];
    }
  };
  if (has1)
#pragma unroll
    for (int r = 0; r < RT; ++r) store(acc1[r], r, wave);
  if constexpr (TWO) {
    if (has2)
#pragma unroll
      for (int r = 0; r < RT; ++r) store(acc2[r], r, wave + NW);
  } else if constexpr (NW == 4 && G == 5) {
#pragma unroll
    for (int x = 0; x < 2; ++x)
      if (wave + 4 * x < RT) store(acc2[x], wave + 4 * x, 4);
  }
}

// Occupancy bound of a variant: two waves per SIMD when a wave keeps two column tiles (or
// has eight waves), else three (FP64) or four.
template <int MODE, int G, int NW>
constexpr int ao_mo_min_blocks() {
  return NW == 8 ? 1 : (G > 5 ? 2 : (MODE == 0 ? 3 : 4));
}

template <int MODE, int G, int NW, bool SL, bool VO>
__global__ void __launch_bounds__(64 * NW, (ao_mo_min_blocks<MODE, G, NW>())) k_ao_mo_tile(AoMoArgs A) {
  extern __shared__ __attribute__((aligned(16))) double lds[];
  ao_mo_tile_body<MODE, G, NW, SL, VO>(A, blockIdx.x, blockIdx.y, lds);
}

// One launch over the products of several contexts (a batch of molecules): workgroup b
// evaluates item b of `items`, (context, point tile, column group) packed as context |
// tile << 16 | group << 48 (context and group < 2^16, tile < 2^32).  The host orders the
// items by decreasing cost, so that the hardware dispatcher, which hands out workgroups in
// order as CUs free up, leaves the cheapest items for the end.
template <int MODE, int G, int NW, bool SL>
__global__ void __launch_bounds__(64 * NW, (ao_mo_min_blocks<MODE, G, NW>()))
    k_ao_mo_batch(const AoMoArgs* __restrict__ args, const uint64_t* __restrict__ items, int nitems) {
  extern __shared__ __attribute__((aligned(16))) double lds[];
  const uint64_t it = items[blockIdx.x];
  const int m = __builtin_amdgcn_readfirstlane((int)(it & 0xffffu));
  const int tile = __builtin_amdgcn_readfirstlane((int)((it >> 16) & 0xffffffffu));
  const int group = __builtin_amdgcn_readfirstlane((int)(it >> 48));
  ao_mo_tile_body<MODE, G, NW, SL, false>(args[m], tile, group, lds);
}

// Dynamic LDS limit of a kernel: the CU's 160 KiB less its static LDS (the block-wide OR of
// the chunk screening uses a little), set once per kernel.
template <class K>
hipError_t prepare_lds(K kernel, size_t lds, size_t& max_dyn) {
  if (max_dyn == 0) {
    const void* k = reinterpret_cast<const void*>(kernel);
    hipFuncAttributes fa;
    hipError_t e = hipFuncGetAttributes(&fa, k);
    if (e != hipSuccess) return e;
    const size_t dyn = 160 * 1024 - fa.sharedSizeBytes;
    e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
    if (e != hipSuccess) return e;
    max_dyn = dyn;
  }
  return lds > max_dyn ? hipErrorInvalidValue : hipSuccess;
}

template <int MODE, int G, int NW, bool SL, bool VO>
hipError_t launch_g(const AoMoArgs& a, size_t lds, hipStream_t s) {
  static size_t max_dyn = 0;
  hipError_t e = prepare_lds(&k_ao_mo_tile<MODE, G, NW, SL, VO>, lds, max_dyn);
  if (e != hipSuccess) return e;
  const dim3 grid((unsigned)((a.np + TP - 1) / TP), (unsigned)a.ngroup);
  hipLaunchKernelGGL((k_ao_mo_tile<MODE, G, NW, SL, VO>), grid, dim3(64 * NW), lds, s, a);
  return hipGetLastError();
}

template <int MODE, int G, int NW, bool SL>
hipError_t launch_batch_g(const AoMoArgs* d_args, const uint64_t* d_items, int nitems, size_t lds, hipStream_t s) {
  static size_t max_dyn = 0;
  hipError_t e = prepare_lds(&k_ao_mo_batch<MODE, G, NW, SL>, lds, max_dyn);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_ao_mo_batch<MODE, G, NW, SL>), dim3((unsigned)nitems), dim3(64 * NW), lds, s, d_args,
                     d_items, nitems);
  return hipGetLastError();
}

// Calls F::template run<MODE, G, NW, SL>() for a runtime (mode, G, NW, slater) among the
// built variants: Gaussian shells NW = 4 for G = 1..8 and NW = 8 for G = 5..16; Slater shells
// NW = 4 for G = 1..8.
template <class F>
hipError_t dispatch_variant(int mode, int G, int NW, int slater, F&& f) {
  if (slater) {
#define AO_MO_S(M, GG) \
  if (mode == M && G == GG && NW == 4) return f.template run<M, GG, 4, true>();
#define AO_MO_SMODE(M) AO_MO_S(M, 1) AO_MO_S(M, 2) AO_MO_S(M, 3) AO_MO_S(M, 4) AO_MO_S(M, 5) AO_MO_S(M, 6) \
  AO_MO_S(M, 7) AO_MO_S(M, 8)
    AO_MO_SMODE(0)
    AO_MO_SMODE(1)
    AO_MO_SMODE(2)
#undef AO_MO_SMODE
#undef AO_MO_S
    return hipErrorInvalidValue;
  }
#define AO_MO_V(M, GG, W) \
  if (mode == M && G == GG && NW == W) return f.template run<M, GG, W, false>();
#define AO_MO_MODE(M)                                                                                    \
  AO_MO_V(M, 1, 4) AO_MO_V(M, 2, 4) AO_MO_V(M, 3, 4) AO_MO_V(M, 4, 4) AO_MO_V(M, 5, 4) AO_MO_V(M, 6, 4)  \
  AO_MO_V(M, 7, 4) AO_MO_V(M, 8, 4) AO_MO_V(M, 5, 8) AO_MO_V(M, 6, 8) AO_MO_V(M, 7, 8) AO_MO_V(M, 8, 8)  \
  AO_MO_V(M, 9, 8) AO_MO_V(M, 10, 8) AO_MO_V(M, 11, 8) AO_MO_V(M, 12, 8) AO_MO_V(M, 13, 8)             \
  AO_MO_V(M, 14, 8) AO_MO_V(M, 15, 8) AO_MO_V(M, 16, 8)
  AO_MO_MODE(0)
  AO_MO_MODE(1)
  AO_MO_MODE(2)
#undef AO_MO_MODE
#undef AO_MO_V
  return hipErrorInvalidValue;
}

struct RunOne {
  const AoMoArgs& a;
  size_t lds;
  hipStream_t s;
  template <int M, int G, int NW, bool SL>
  hipError_t run() { return a.vo ? launch_g<M, G, NW, SL, true>(a, lds, s) : launch_g<M, G, NW, SL, false>(a, lds, s); }
};

struct RunBatch {
  const AoMoArgs* args;
  const uint64_t* items;
  int n;
  size_t lds;
  hipStream_t s;
  template <int M, int G, int NW, bool SL>
  hipError_t run() { return launch_batch_g<M, G, NW, SL>(args, items, n, lds, s); }
};

}  // namespace

size_t ao_mo_lds_bytes(const AoMoArgs& a, int mode) {
  const size_t tile = a.vo ? (mode == 0 ? tile_bytes<0, TP>() : mode == 1 ? tile_bytes<1, TP>() : tile_bytes<2, TP>())
                           : (mode == 0 ? tile_bytes<0, ROWS>() : mode == 1 ? tile_bytes<1, ROWS>() : tile_bytes<2, ROWS>());
  return sizeof(double) * (2 * (size_t)a.desc_max + 4 * TP + (size_t)TP * a.ns_max * SH) + tile;
}

bool ao_mo_variant_ok(int mode, int G, int NW, int slater) {
  if (mode < 0 || mode > 2) return false;
  return (NW == 4 && G >= 1 && G <= 8) || (!slater && NW == 8 && G >= 5 && G <= 16);
}

hipError_t launch_ao_mo(const AoMoArgs& a, int mode, hipStream_t s) {
  if (a.np <= 0) return hipSuccess;
  if (!ao_mo_variant_ok(mode, a.G, a.nw, a.slater) || a.nchunk < 1 || a.ns_max < 1 || a.ns_max > KC || a.desc_max % 2 ||
      a.b_chunk_bytes % 16 || a.tw < 1 || a.te < 1 || a.p0 < 0 || a.p0 + a.np > a.tw * a.te)
    return hipErrorInvalidValue;
  const size_t lds = ao_mo_lds_bytes(a, mode);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  RunOne r{a, lds, s};
  return dispatch_variant(mode, a.G, a.nw, a.slater, r);
}

hipError_t launch_ao_mo_batch(int mode, int G, int NW, int slater, const AoMoArgs* d_args, const uint64_t* d_items,
                              int nitems, size_t lds, hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  if (!ao_mo_variant_ok(mode, G, NW, slater) || lds > 160 * 1024) return hipErrorInvalidValue;
  RunBatch r{d_args, d_items, nitems, lds, s};
  return dispatch_variant(mode, G, NW, slater, r);
}

int ao_mo_waves(int G) { return G > 8 ? 8 : 4; }

hipError_t mo_batch_build(const std::vector<std::pair<AoMoArgs, int>>& in, MoBatchPlan& plan) {
  mo_batch_free(plan);
  // contexts grouped by shape; within a shape, the costliest contexts' items first (an item
  // costs about its context's number of 32-AO chunks)
  std::map<std::tuple<int, int, int, int>, std::vector<int>> by_shape;
  for (int i = 0; i < (int)in.size(); ++i) {
    const AoMoArgs& a = in[i].first;
    const int mode = in[i].second;
    if (a.np <= 0) continue;
    if (!ao_mo_variant_ok(mode, a.G, a.nw, a.slater) || a.vo || a.nchunk < 1 || a.ns_max < 1 || a.ns_max > KC ||
        a.desc_max % 2 || a.b_chunk_bytes % 16 || a.tw < 1 || a.te < 1 || a.p0 < 0 || a.p0 + a.np > a.tw * a.te ||
        ao_mo_lds_bytes(a, mode) > 160 * 1024)
      return hipErrorInvalidValue;
    by_shape[std::make_tuple(mode, a.G, a.nw, a.slater)].push_back(i);
  }
  hipError_t e = hipSuccess;
  for (auto& kv : by_shape) {
    std::vector<int>& ids = kv.second;
    if (ids.size() > 65535) return hipErrorInvalidValue;
    std::stable_sort(ids.begin(), ids.end(), [&](int x, int y) { return in[x].first.nchunk > in[y].first.nchunk; });
    std::vector<AoMoArgs> args;
    std::vector<uint64_t> items;
    MoBatchLaunch l;
    l.mode = std::get<0>(kv.first);
    l.G = std::get<1>(kv.first);
    l.nw = std::get<2>(kv.first);
    l.slater = std::get<3>(kv.first);
    for (size_t m = 0; m < ids.size(); ++m) {
      const AoMoArgs& a = in[ids[m]].first;
      args.push_back(a);
      l.lds = std::max(l.lds, ao_mo_lds_bytes(a, l.mode));
      const int64_t ntile = (a.np + TP - 1) / TP;
      if (ntile > 0xffffffffLL || a.ngroup > 65535) return hipErrorInvalidValue;
      for (int64_t t = 0; t < ntile; ++t)
        for (int g = 0; g < a.ngroup; ++g) items.push_back((uint64_t)m | ((uint64_t)t << 16) | ((uint64_t)g << 48));
      l.cost += (double)ntile * a.ngroup * a.nchunk * a.G;
    }
    if (items.size() > (size_t)INT32_MAX) return hipErrorInvalidValue;
    l.nitems = (int)items.size();
    e = hipMalloc(reinterpret_cast<void**>(&l.d_args), sizeof(AoMoArgs) * args.size());
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&l.d_items), sizeof(uint64_t) * items.size());
    if (e == hipSuccess)
      e = hipMemcpy(l.d_args, args.data(), sizeof(AoMoArgs) * args.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(l.d_items, items.data(), sizeof(uint64_t) * items.size(), hipMemcpyHostToDevice);
    plan.launches.push_back(l);
    if (e != hipSuccess) break;
  }
  if (e != hipSuccess) mo_batch_free(plan);
  return e;
}

hipError_t mo_batch_launch(const MoBatchPlan& plan, const hipStream_t* s, int n) {
  std::vector<int> order(plan.launches.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int x, int y) { return plan.launches[x].cost > plan.launches[y].cost; });
  std::vector<double> load(n, 0.0);
  for (int i : order) {
    const MoBatchLaunch& l = plan.launches[i];
    const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    load[k] += l.cost;
    const hipError_t e = launch_ao_mo_batch(l.mode, l.G, l.nw, l.slater, l.d_args, l.d_items, l.nitems, l.lds, s[k]);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

void mo_batch_free(MoBatchPlan& plan) {
  for (MoBatchLaunch& l : plan.launches) {
    if (l.d_args) (void)hipFree(l.d_args);
    if (l.d_items) (void)hipFree(l.d_items);
  }
  plan.launches.clear();
}

}  // namespace qmckl_amd
