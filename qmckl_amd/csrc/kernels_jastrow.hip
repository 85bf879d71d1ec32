// CHAMP Jastrow factor on gfx950, FP64 throughout (org/qmckl_jastrow_champ.org).
//
// Three kernels per evaluation, or two when the LDS-resident e-e-n kernel applies (it then
// also computes what k_jas_pair provides, from the pair data it builds anyway):
//   k_jas_pair   one workgroup per walker, one lane per electron i.  e-e term (:3582-3650,
//                :4188-4295), e-n term (:7021-7095, :7500-7590), and the a-independent
//                sums of the e-e-n tensor R[k][c] = sum_{j!=i} g_e(r_ij)^k w_c(ij),
//                w = (1, u_x, u_y, u_z, 1/r_ij), u = (r_i - r_j)/r_ij, k = 1..cord-1.
//   k_jas_een    one workgroup per (walker, tile of TI electrons), one lane per (i, nucleus a).
//                Accumulates over j, for k = 1..cord-1 and q = 1..cord-k,
//                  V[k][q]  = sum_j g_e^k g_a(j)^q       (this is tmp_c,  :12204-12260)
//                  T_c[k][q]= sum_j g_e^k u_c g_a(j)^q
//                  L[k][q]  = sum_j g_e^k g_a(j)^q / r_ij
//                and S[q] = sum_{j!=i} g_a(j)^q (k = 0).  The gradient and Laplacian of
//                g_e^k = exp(-k kappa r) give dtmp_c (:12484-12540):
//                  dP_c = -k kappa T_c,  dP_lap = k kappa (k kappa V - 2 L).
//                The epilogue applies the c-coefficients exactly as factor_een (:14121-14175)
//                and factor_een_gl (:14875-14945) do, with electron i in the role of j there,
//                and reduces over the nuclei of the tile in LDS.  Nothing of size
//                O(Ne^2) or O(Ne Nn cord^2) is written to memory.
//   k_jas_total  value = exp(J) (0 if J < -100) and gl = e^J (grad J, lap J + |grad J|^2)
//                (:16578-16620, :17072-17112).
// Sums are reduced in a fixed order (no atomics), so results are run-to-run reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <map>
#include <vector>
#include <type_traits>

#include "exp_neg.hpp"
#include "internal.hpp"

namespace qmckl_amd {

namespace {

constexpr int kBlock = 256;
constexpr int kPairBlock = 1024;  // k_jas_pair: up to 16 waves per walker, one electron i per thread
// Threads per k_jas_pair<C> block at most: 1024 limits a thread to 128 VGPRs, enough for the
// q = 0 sums of orders up to 6; orders 7 and 8 hold more and take 256 threads (no spill).
constexpr int pair_block(int C) { return C <= 6 ? kPairBlock : 256; }

// ---- (l, k, m) of the c-coefficients, org/qmckl_jastrow_champ.org:12000-12040 ----------
constexpr int dimc(int C) {
  int n = 0;
  for (int p = 2; p <= C; ++p)
    for (int k = p - 1; k >= 0; --k) {
      const int lmax = (k != 0) ? p - k : p - k - 2;
      for (int l = lmax; l >= 0; --l)
        if (((p - k - l) & 1) == 0) ++n;
    }
  return n;
}

template <int C>
struct Lkm {
  static constexpr int dim = dimc(C);
  int l[dim > 0 ? dim : 1], k[dim > 0 ? dim : 1], m[dim > 0 ? dim : 1];
  constexpr Lkm() : l(), k(), m() {
    int n = 0;
    for (int p = 2; p <= C; ++p)
      for (int kk = p - 1; kk >= 0; --kk) {
        const int lmax = (kk != 0) ? p - kk : p - kk - 2;
        for (int ll = lmax; ll >= 0; --ll) {
          if (((p - kk - ll) & 1) == 1) continue;
          l[n] = ll;
          k[n] = kk;
          m[n] = (p - kk - ll) / 2;
          ++n;
        }
      }
  }
};

// Cross-lane exchanges of doubles that stay off the LDS path (a shuffle through
// ds_bpermute costs an LDS round trip): DPP moves within a row of 16 lanes, and gfx950's
// v_permlane16_swap / v_permlane32_swap across rows.  Lane layouts, as measured on the
// MI355X: row_mirror sends lane i of a row to lane 15 - i (xor 15), row_half_mirror lane i
// of 8 to 7 - i (xor 7), quad_perm(1,0,3,2) is xor 1 and quad_perm(2,3,0,1) xor 2;
// permlane16_swap(a, b) returns a with its odd rows (lanes 16-31, 48-63) replaced by the
// even rows of b, and b with its even rows replaced by the odd rows of a; permlane32_swap
// the same with the halves of the wave.
constexpr int DPP_XOR1 = 0xB1, DPP_XOR2 = 0x4E, DPP_XOR7 = 0x141, DPP_XOR15 = 0x140;
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
  const uint64_t b = __builtin_bit_cast(uint64_t, x);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, 0xF, 0xF, false);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
// a + b after the 16-lane (ROWS16) or 32-lane swap: in the even rows (halves) a's own value
// plus b's from the odd ones, in the odd rows (halves) a's value from the even ones plus b's own
template <bool ROWS16>
__device__ __forceinline__ double swap_sum_f64(double a, double b) {
  const uint64_t ua = __builtin_bit_cast(uint64_t, a), ub = __builtin_bit_cast(uint64_t, b);
  uint32_t alo = (uint32_t)ua, ahi = (uint32_t)(ua >> 32), blo = (uint32_t)ub, bhi = (uint32_t)(ub >> 32);
  if constexpr (ROWS16) {
    const auto l = __builtin_amdgcn_permlane16_swap(alo, blo, false, false);
    const auto h = __builtin_amdgcn_permlane16_swap(ahi, bhi, false, false);
    alo = l[0], blo = l[1], ahi = h[0], bhi = h[1];
  } else {
    const auto l = __builtin_amdgcn_permlane32_swap(alo, blo, false, false);
    const auto h = __builtin_amdgcn_permlane32_swap(ahi, bhi, false, false);
    alo = l[0], blo = l[1], ahi = h[0], bhi = h[1];
  }
  return __builtin_bit_cast(double, ((uint64_t)ahi << 32) | alo) + __builtin_bit_cast(double, ((uint64_t)bhi << 32) | blo);
}
// Sum over the 16 lanes of a row, in every lane of it, by the tree of xor 1, 2, 4, 8: after
// the first two steps a quad holds one value, so xor 7 and xor 15 meet the values of xor 4
// and xor 8.
__device__ __forceinline__ double row_sum16(double x) {
  x += dpp_f64<DPP_XOR1>(x);
  x += dpp_f64<DPP_XOR2>(x);
  x += dpp_f64<DPP_XOR7>(x);
  x += dpp_f64<DPP_XOR15>(x);
  return x;
}
// Sum over the 64 lanes of the wave, in every lane, by the tree of xor 1, 2, ..., 32.
__device__ __forceinline__ double wave_sum64(double x) {
  x = row_sum16(x);
  x = swap_sum_f64<true>(x, x);
  return swap_sum_f64<false>(x, x);
}

// A wave-uniform double in scalar registers (its two halves through readfirstlane).
__device__ __forceinline__ double uniform_f64(double x) {
  const uint64_t u = __builtin_bit_cast(uint64_t, x);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u), hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Deterministic sum over the lanes of a block of any multiple of 64 threads: a fixed
// shuffle tree per wave, then the wave partials in wave order.  red holds one double per
// wave.
__device__ __forceinline__ double block_sum_waves(double x, double* red) {
  x = wave_sum64(x);
  const int nw = (int)blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
  __syncthreads();
  double s = 0.0;
  for (int wv = 0; wv < nw; ++wv) s += red[wv];
  __syncthreads();
  return s;
}


// g_a(r_j) = exp(-kappa_a |r_j - R_a|) scratch, laid out as the e-e-n kernels stage it in
// LDS: [W][nucleus tile of 16][j, Ne rounded up to 4][16], zero in the padding.
__host__ __device__ __forceinline__ int64_t gn_index(int64_t w, int a, int j, int NA, int JP) {
  return ((w * NA + (a >> 4)) * JP + j) * 16 + (a & 15);
}
__host__ __device__ __forceinline__ int gn_na(int64_t Nn) { return (int)((Nn + 15) >> 4); }
__host__ __device__ __forceinline__ int gn_jp(int64_t Ne) { return (int)((Ne + 3) & ~3); }

// ------------------------------------------------------------------------------------
// One e-e pair (i, j) with displacement (dx, dy, dz) = r_i - r_j: the value term of the e-e
// factor when j < i (org/qmckl_jastrow_champ.org:3582-3650, the doc loop over j > i' of the
// reference), its gradient and Laplacian with respect to electron i (:4188-4295 with
// org/qmckl_distance.org:1440-1452), and the q = 0 e-e-n sums R0[k-1][c] = sum_j g_e^k w_c
// (k = 1..K1, w = (1, u, 1/r)).  No branch on j == i: the distance is replaced by 1 and
// every gradient, Laplacian and e-e-n term is multiplied by an exact zero (dx = dy = dz = 0,
// ed = em = 0), which leaves the sums unchanged.  Returns g_e(r_ij) = e^(-kappa r_ij) and
// 1/r_ij, or 0 and 0 for j == i.
// |r| and 1/|r| from r2 = |r|^2 with one reciprocal square root: v_rsq_f64 refined by two
// Newton steps y <- y (3/2 - r2 y^2 / 2), within a few ulp of the correctly rounded values,
// instead of a square root and a division; outside [1e-200, 1e200] (coincident points
// included) the library square root and a division.
__device__ __forceinline__ void dist_and_inverse(double r2, double& d, double& rinv) {
  if (r2 >= 1.e-200 && r2 <= 1.e200) {
    double y = __builtin_amdgcn_rsq(r2);
    const double h = 0.5 * r2;
    y = y * __builtin_fma(-h, y * y, 1.5);
    y = y * __builtin_fma(-h, y * y, 1.5);
    rinv = y;
    d = r2 * y;
  } else {
    d = sqrt(r2);
    rinv = 1.0 / d;
  }
}

struct EeAcc {
  double fee = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0, g3 = 0.0;
};

// P(x) = sum_{k=2..n} c_k x^k with its first derivative and half its second derivative, by
// Horner's rule carried for the derivatives (the k = 1 and k = 0 coefficients are 0).
__device__ __forceinline__ void poly2_horner(const double* c, int n, double x, double& p, double& dp, double& hp) {
  p = 0.0;
  dp = 0.0;
  hp = 0.0;
  if (n < 2) return;
  p = c[n];
  for (int k = n - 1; k >= 2; --k) {
    hp = hp * x + dp;
    dp = dp * x + p;
    p = p * x + c[k];
  }
#pragma unroll
  for (int k = 0; k < 2; ++k) {  // the zero coefficients of x^1 and x^0
    hp = hp * x + dp;
    dp = dp * x + p;
    p = p * x;
  }
}
// b: the b_vector (in global memory or its LDS copy)
template <int K1>
__device__ __forceinline__ void ee_pair(const JasDevView& v, const double* b, int i, int j, double dx, double dy,
                                        double dz, double kee, double kinv, EeAcc& s,
                                        double (&r0)[K1 > 0 ? K1 : 1][5], double& ge, double& ri) {
  const int bord = (int)v.bord, up = (int)v.up_num;
  const bool use = j != i;
  double d = 1.0, rinv = 1.0;
  if (use) dist_and_inverse(dx * dx + dy * dy + dz * dz, d, rinv);
  const double e = exp_neg_deg<13, true>(-kee * d);
  const double x = (1.0 - e) * kinv;  // rescaled distance, org/qmckl_distance.org:975-985
  double pq, dq, hq;  // sum_{k>=2} b_k x^k, its derivative, half its second derivative
  poly2_horner(b, bord, x, pq, dq, hq);
  // one reciprocal of the Pade denominator for the value and the derivatives
  const double invd = 1.0 / (1.0 + b[1] * x), invd2 = invd * invd;
  if (j < i) {
    const bool same = (v.spin_independent != 1) && (i < up || j >= up);
    if (v.spin_independent == 1)
      s.fee += b[0] * x * invd - v.asymp_jasb1;
    else if (same)
      s.fee += 0.5 * b[0] * x * invd - v.asymp_jasb0;
    else
      s.fee += b[0] * x * invd - v.asymp_jasb1;
    s.fee += pq;
  }
  const double dd = d > 1.e-20 ? d : 1.e-20;
  const double dinv = d > 1.e-20 ? rinv : 1.e20;  // 1 / dd
  const double ed = use ? ((d > 1.e-20) ? e : exp_neg_deg<13, true>(-kee * dd)) : 0.0;
  const double gx = dx * dinv * ed, gy = dy * dinv * ed, gz = dz * dinv * ed;
  const double gl4 = (2.0 * dinv - kee) * ed;
  const double gc2 = gx * gx + gy * gy + gz * gz;
  double f;
  if (v.spin_independent == 1)
    f = b[0] * invd2;
  else if ((i < up && j < up) || (i >= up && j >= up))
    f = 0.5 * b[0] * invd2;
  else
    f = b[0] * invd2;
  s.g0 += f * gx;
  s.g1 += f * gy;
  s.g2 += f * gz;
  s.g3 += f * (gl4 - 2.0 * b[1] * gc2 * invd);
  // sum_{k>=2} b_k (k x^(k-1) grad x, k x^(k-1) lap x + k (k-1) x^(k-2) |grad x|^2)
  s.g0 += dq * gx;
  s.g1 += dq * gy;
  s.g2 += dq * gz;
  s.g3 += dq * gl4 + 2.0 * hq * gc2;
  const double ux = dx * rinv, uy = dy * rinv, uz = dz * rinv;
  const double em = use ? e : 0.0;
  double gk = 1.0;
#pragma unroll
  for (int k = 1; k <= K1; ++k) {
    gk *= em;
    r0[k - 1][0] += gk;
    r0[k - 1][1] += gk * ux;
    r0[k - 1][2] += gk * uy;
    r0[k - 1][3] += gk * uz;
    r0[k - 1][4] += gk * rinv;
  }
  ge = em;
  ri = use ? rinv : 0.0;
}

// One e-n pair (electron i, nucleus a of type t) with displacement (dx, dy, dz) = r_i - R_a,
// distance d and e = g_a(r_i) = e^(-kappa_a d): the value term of the e-n factor
// (org/qmckl_jastrow_champ.org:7021-7095) and its gradient and Laplacian with respect to
// electron i (:7500-7590).
struct EnAcc {
  double fen = 0.0, h0 = 0.0, h1 = 0.0, h2 = 0.0, h3 = 0.0;
};
// Parameters of the type: av = a_vector row, ka = kappa_en, kinv = 1 / ka, asymp = asymp_jasa.
__device__ __forceinline__ void en_pair_p(const double* av, int aord, double ka, double kinv, double asymp, double dx,
                                          double dy, double dz, double d, double rinv, double e, EnAcc& s) {
  const double x = (1.0 - e) * kinv;
  double pq, dq, hq;  // sum_{p>=2} a_p x^p, its derivative, half its second derivative
  poly2_horner(av, aord, x, pq, dq, hq);
  // one reciprocal of the Pade denominator for the value and the derivatives
  const double invd = 1.0 / (1.0 + av[1] * x), invd2 = invd * invd;
  s.fen += av[0] * x * invd - asymp;
  s.fen += pq;
  const double dd = d > 1.e-20 ? d : 1.e-20;
  const double dinv = d > 1.e-20 ? rinv : 1.e20;  // 1 / dd
  const double ed = (d > 1.e-20) ? e : exp_neg_deg<13, true>(-ka * dd);
  const double gx = dx * dinv * ed, gy = dy * dinv * ed, gz = dz * dinv * ed;
  const double gl4 = (2.0 * dinv - ka) * ed;
  double f = av[0] * invd2;
  const double gc2 = gx * gx + gy * gy + gz * gz;
  s.h0 += f * gx;
  s.h1 += f * gy;
  s.h2 += f * gz;
  s.h3 += f * (gl4 - 2.0 * av[1] * gc2 * invd);
  s.h0 += dq * gx;
  s.h1 += dq * gy;
  s.h2 += dq * gz;
  s.h3 += dq * gl4 + 2.0 * hq * gc2;
}
__device__ __forceinline__ void en_pair(const JasDevView& v, int t, double dx, double dy, double dz, double d,
                                        double rinv, double e, EnAcc& s) {
  const int aord = (int)v.aord;
  const double ka = v.kappa_en[t];
  en_pair_p(v.a_vector + t * (aord + 1), aord, ka, 1.0 / ka, v.asymp_jasa[t], dx, dy, dz, d, rinv, e, s);
}

// ------------------------------------------------------------------------------------
// e-e and e-n terms, and the q = 0 sums of the e-e-n tensor.
// R0 scratch layout: [W][Ne][cord-1][5]; gN: see gn_index.
// ------------------------------------------------------------------------------------
// C (= cord, 0..8) is a template parameter so that only the C - 1 orders of the q = 0 sums
// are held in registers.  The block has 64 ceil(Ne / 64) threads (at most pair_block(C)), one
// electron i per thread.
template <int C>
__device__ __forceinline__ void jas_pair_body(const JasDevView& v, const JasOut& o, double* __restrict__ R0,
                                              double* __restrict__ gN, double* __restrict__ Sg, const int64_t w) {
  constexpr int K1 = C >= 2 ? C - 1 : 0;
  extern __shared__ double lds[];
  const int Ne = (int)v.elec_num, Nn = (int)v.nucl_num;
  const int NAg = gn_na(Nn), JPg = gn_jp(Ne);
  double* xe = lds;            // [Ne][3]
  double* red = lds + 3 * Ne;  // [waves]
  const int nth = (int)blockDim.x;
  for (int t = threadIdx.x; t < Ne; t += nth)
    for (int d = 0; d < 3; ++d) xe[3 * t + d] = v.coord[w * Ne + t + d * v.coord_ld];
  __syncthreads();

  const double kee = v.rescale_ee, kinv = 1.0 / v.rescale_ee;
  double fee = 0.0, fen = 0.0;
  for (int i = threadIdx.x; i < Ne; i += nth) {
    const double xi = xe[3 * i], yi = xe[3 * i + 1], zi = xe[3 * i + 2];
    EeAcc se;
    double r0[K1 > 0 ? K1 : 1][5];
#pragma unroll
    for (int k = 0; k < (K1 > 0 ? K1 : 1); ++k)
#pragma unroll
      for (int c = 0; c < 5; ++c) r0[k][c] = 0.0;
    for (int j = 0; j < Ne; ++j) {
      double ge, ri;
      ee_pair<K1>(v, v.b_vector, i, j, xi - xe[3 * j], yi - xe[3 * j + 1], zi - xe[3 * j + 2], kee, kinv, se, r0, ge, ri);
    }
    fee += se.fee;
    double* og = o.fee_gl + w * 4 * Ne;
    og[i] = se.g0;
    og[i + Ne] = se.g1;
    og[i + 2 * Ne] = se.g2;
    og[i + 3 * Ne] = se.g3;
    if constexpr (K1 > 0) {
      double* rr = R0 + ((w * Ne + i) * K1) * 5;
#pragma unroll
      for (int k = 0; k < K1; ++k)
#pragma unroll
        for (int c = 0; c < 5; ++c) rr[k * 5 + c] = r0[k][c];
    }
    EnAcc sn;
    for (int a = 0; a < Nn; ++a) {
      const int t = (int)v.type_nucl[a];
      const double dx = xi - v.nucl_coord[a], dy = yi - v.nucl_coord[a + Nn], dz = zi - v.nucl_coord[a + 2 * Nn];
      double d, rinv;
      dist_and_inverse(dx * dx + dy * dy + dz * dz, d, rinv);
      const double e = exp_neg_deg<13, true>(-v.kappa_en[t] * d);
      gN[gn_index(w, a, i, NAg, JPg)] = e;  // g_a(r_i), reused by the e-e-n kernels
      en_pair(v, t, dx, dy, dz, d, rinv, e, sn);
    }
    fen += sn.fen;
    double* oh = o.fen_gl + w * 4 * Ne;
    oh[i] = sn.h0;
    oh[i + Ne] = sn.h1;
    oh[i + 2 * Ne] = sn.h2;
    oh[i + 3 * Ne] = sn.h3;
  }
  const double see = block_sum_waves(fee, red);
  const double sen = block_sum_waves(fen, red);
  if (threadIdx.x == 0) {
    o.fee[w] = see;
    o.fen[w] = sen;
  }
  // Sg[w][a][q-1] = sum_i g_a(r_i)^q, q = 1..cord-1: the k = 0 sums of the e-e-n term over
  // all electrons (the barriers of block_sum_waves make this block's gN writes visible).
  if constexpr (K1 > 0) {
    for (int a = threadIdx.x; a < Nn; a += nth) {
      double acc[K1];
#pragma unroll
      for (int q = 0; q < K1; ++q) acc[q] = 0.0;
      const double* ga = gN + gn_index(w, a, 0, NAg, JPg);
#pragma unroll 4
      for (int i = 0; i < Ne; ++i) {
        const double g = ga[16 * i];
        double p = g;
#pragma unroll
        for (int q = 0; q < K1; ++q) {
          acc[q] += p;
          p *= g;
        }
      }
#pragma unroll
      for (int q = 0; q < K1; ++q) Sg[(w * Nn + a) * K1 + q] = acc[q];
    }
  }
  // zero padding of gN: nuclei Nn .. 16 NA - 1 for every j, electrons Ne .. JP - 1 for a < Nn
  const int npad_a = 16 * NAg - Nn;
  for (int t = threadIdx.x; t < npad_a * JPg; t += nth) {
    const int j = t / npad_a, a = Nn + t - j * npad_a;
    gN[gn_index(w, a, j, NAg, JPg)] = 0.0;
  }
  for (int t = threadIdx.x; t < (JPg - Ne) * Nn; t += nth) {
    const int j = Ne + t / Nn, a = t % Nn;
    gN[gn_index(w, a, j, NAg, JPg)] = 0.0;
  }
}

template <int C>
__global__ void __launch_bounds__(pair_block(C)) k_jas_pair(JasDevView v, JasOut o, double* __restrict__ R0,
                                                     double* __restrict__ gN, double* __restrict__ Sg) {
  jas_pair_body<C>(v, o, R0, gN, Sg, blockIdx.x);
}

// ------------------------------------------------------------------------------------
// e-e-n term.  Tile: TI electrons x all nuclei; j walked in chunks of JC.
// Jt scratch: [W][ntiles] partial J_eeN.
// ------------------------------------------------------------------------------------
template <int C>
__device__ __forceinline__ void jas_een_body(const JasDevView& v, const JasOut& o, const double* __restrict__ R0,
                                             const double* __restrict__ gN, double* __restrict__ Jt, int TI, int JC,
                                             int ntiles, const int64_t b) {
  constexpr int NQ = C;  // q = 1..C
  constexpr Lkm<C> T{};
  extern __shared__ double lds[];
  const int Ne = (int)v.elec_num, Nn = (int)v.nucl_num;
  const int64_t w = b / ntiles;
  const int tile = (int)(b - w * ntiles);
  const int i0 = tile * TI;
  const int ti = min(TI, Ne - i0);
  // LDS layout
  double* xt = lds;                // [TI][3] tile electrons
  double* row = xt + 3 * TI;       // [TI][JC][6]: G, ux, uy, uz, 1/r, (j != i)
  double* gn = row + 6 * TI * JC;  // [Nn][JC]: g_a(r_j) = exp(-kappa_a R_ja)
  double* red = gn + Nn * JC;      // [max(ti*Nn,kBlock)][5]
  const double kee = v.rescale_ee;
  const double* coord = v.coord;
  const int64_t cld = v.coord_ld;
  for (int t = threadIdx.x; t < ti; t += kBlock)
    for (int d = 0; d < 3; ++d) xt[3 * t + d] = coord[w * Ne + i0 + t + d * cld];
  __syncthreads();

  const int npairs = ti * Nn;
  for (int base = 0; base < npairs; base += kBlock) {
    const int pair = base + (int)threadIdx.x;
    const bool active = pair < npairs;
    const int il = active ? pair / Nn : 0;
    const int a = active ? pair - il * Nn : 0;
    double V[C][NQ + 1], Tx[C][NQ + 1], Ty[C][NQ + 1], Tz[C][NQ + 1], L[C][NQ + 1];
    double S[NQ + 1];
#pragma unroll
    for (int k = 0; k < C; ++k)
#pragma unroll
      for (int q = 0; q <= NQ; ++q) V[k][q] = Tx[k][q] = Ty[k][q] = Tz[k][q] = L[k][q] = 0.0;
#pragma unroll
    for (int q = 0; q <= NQ; ++q) S[q] = 0.0;

    for (int j0 = 0; j0 < Ne; j0 += JC) {
      const int jc = min(JC, Ne - j0);
      for (int t = threadIdx.x; t < ti * jc; t += kBlock) {
        const int q = t / jc, jj = t - q * jc;
        const int j = j0 + jj;
        const double dx = xt[3 * q] - coord[w * Ne + j];
        const double dy = xt[3 * q + 1] - coord[w * Ne + j + cld];
        const double dz = xt[3 * q + 2] - coord[w * Ne + j + 2 * cld];
        double* r = row + 6 * (q * JC + jj);
        if (i0 + q == j) {
          r[0] = r[1] = r[2] = r[3] = r[4] = r[5] = 0.0;
        } else {
          const double d = sqrt(dx * dx + dy * dy + dz * dz);
          const double rinv = 1.0 / d;
          r[0] = exp(-kee * d);
          r[1] = dx * rinv;
          r[2] = dy * rinv;
          r[3] = dz * rinv;
          r[4] = rinv;
          r[5] = 1.0;
        }
      }
      for (int t = threadIdx.x; t < Nn * jc; t += kBlock) {
        const int aa = t / jc, jj = t - aa * jc;
        gn[aa * JC + jj] = gN[gn_index(w, aa, j0 + jj, gn_na(Nn), gn_jp(Ne))];
      }
      __syncthreads();
      if (active) {
        const double* rp = row + 6 * il * JC;
        const double* gp = gn + a * JC;
        for (int jj = 0; jj < jc; ++jj) {
          const double G = rp[6 * jj], ux = rp[6 * jj + 1], uy = rp[6 * jj + 2], uz = rp[6 * jj + 3],
                       ri = rp[6 * jj + 4], msk = rp[6 * jj + 5];
          const double g = gp[jj];
          double gq[NQ + 1];
          gq[0] = 1.0;
#pragma unroll
          for (int q = 1; q <= NQ; ++q) gq[q] = gq[q - 1] * g;
#pragma unroll
          for (int q = 1; q <= NQ; ++q) S[q] += msk * gq[q];
          double Gk = 1.0;
#pragma unroll
          for (int k = 1; k < C; ++k) {
            Gk *= G;
            const double wx = Gk * ux, wy = Gk * uy, wz = Gk * uz, wl = Gk * ri;
#pragma unroll
            for (int q = 1; q <= C - k; ++q) {
              V[k][q] += Gk * gq[q];
              Tx[k][q] += wx * gq[q];
              Ty[k][q] += wy * gq[q];
              Tz[k][q] += wz * gq[q];
              L[k][q] += wl * gq[q];
            }
          }
        }
      }
      __syncthreads();
    }

    double Ji = 0.0, gli[4] = {0.0, 0.0, 0.0, 0.0};
    if (active) {
      const int i = i0 + il;
      // q = 0 sums (independent of a) and P[0][0] = Ne - 1
      const double* rr = R0 + ((w * Ne + i) * (C - 1)) * 5;
#pragma unroll
      for (int k = 1; k < C; ++k) {
        V[k][0] = rr[(k - 1) * 5 + 0];
        Tx[k][0] = rr[(k - 1) * 5 + 1];
        Ty[k][0] = rr[(k - 1) * 5 + 2];
        Tz[k][0] = rr[(k - 1) * 5 + 3];
        L[k][0] = rr[(k - 1) * 5 + 4];
      }
      V[0][0] = (double)(Ne - 1);
#pragma unroll
      for (int q = 1; q <= NQ; ++q) V[0][q] = S[q];
      // P and dP (4 components) for all (k, q) with k + q <= C
      double P[C][NQ + 1], D[C][NQ + 1][4];
#pragma unroll
      for (int k = 0; k < C; ++k)
#pragma unroll
        for (int q = 0; q <= NQ; ++q) {
          if (k + q > C) continue;
          P[k][q] = V[k][q];
          if (k == 0) {
            D[k][q][0] = D[k][q][1] = D[k][q][2] = D[k][q][3] = 0.0;
          } else {
            const double kk = (double)k * kee;
            D[k][q][0] = -kk * Tx[k][q];
            D[k][q][1] = -kk * Ty[k][q];
            D[k][q][2] = -kk * Tz[k][q];
            D[k][q][3] = kk * (kk * V[k][q] - 2.0 * L[k][q]);
          }
        }
      // g_a(r_i)^p and its gradient/Laplacian, org/qmckl_jastrow_champ.org:11150-11180
      const double ka = v.kappa_en[v.type_nucl[a]];
      const double dx = xt[3 * il] - v.nucl_coord[a], dy = xt[3 * il + 1] - v.nucl_coord[a + Nn],
                   dz = xt[3 * il + 2] - v.nucl_coord[a + 2 * Nn];
      const double R = sqrt(dx * dx + dy * dy + dz * dz);
      const double rinv = 1.0 / R;
      const double g1 = exp(-ka * R);
      double gp[C + 1], dg[C + 1][4];
      gp[0] = 1.0;
#pragma unroll
      for (int p = 1; p <= C; ++p) gp[p] = gp[p - 1] * g1;
#pragma unroll
      for (int p = 0; p <= C; ++p) {
        const double kl = -(double)p * ka;
        const double d0 = kl * (dx * rinv), d1 = kl * (dy * rinv), d2 = kl * (dz * rinv);
        const double d3 = kl * 2.0 * rinv + d0 * d0 + d1 * d1 + d2 * d2;
        dg[p][0] = d0 * gp[p];
        dg[p][1] = d1 * gp[p];
        dg[p][2] = d2 * gp[p];
        dg[p][3] = d3 * gp[p];
      }
      const double* cf = v.cfull + a;
#pragma unroll
      for (int n = 0; n < T.dim; ++n) {
        const int l = T.l[n], k = T.k[n], m = T.m[n];
        double cn = cf[(int64_t)n * Nn];
        if (cn == 0.0) continue;
        Ji += cn * gp[m] * P[k][m + l];
#pragma unroll
        for (int c = 0; c < 4; ++c)
          gli[c] += (P[k][m] * dg[m + l][c] + P[k][m + l] * dg[m][c] + D[k][m][c] * gp[m + l] +
                     D[k][m + l][c] * gp[m]) *
                    cn;
        cn = cn + cn;
        gli[3] += (D[k][m][0] * dg[m + l][0] + D[k][m][1] * dg[m + l][1] + D[k][m][2] * dg[m + l][2] +
                   D[k][m + l][0] * dg[m][0] + D[k][m + l][1] * dg[m][1] + D[k][m + l][2] * dg[m][2]) *
                  cn;
      }
    }
    if (active) {
      double* rd = red + 5 * pair;
      rd[0] = Ji;
      rd[1] = gli[0];
      rd[2] = gli[1];
      rd[3] = gli[2];
      rd[4] = gli[3];
    }
  }
  __syncthreads();
  // Reduce over the nuclei of each electron of the tile (fixed order).
  for (int t = threadIdx.x; t < ti * 4; t += kBlock) {
    const int il = t >> 2, c = t & 3;
    double s = 0.0;
    for (int a = 0; a < Nn; ++a) s += red[5 * (il * Nn + a) + 1 + c];
    o.feen_gl[w * 4 * Ne + (int64_t)c * Ne + i0 + il] = s;
  }
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int p = 0; p < npairs; ++p) s += red[5 * p];
    Jt[w * ntiles + tile] = s;
  }
}

template <int C>
__global__ void __launch_bounds__(kBlock) k_jas_een(JasDevView v, JasOut o, const double* __restrict__ R0,
                                                    const double* __restrict__ gN, double* __restrict__ Jt, int TI,
                                                    int JC, int ntiles) {
  jas_een_body<C>(v, o, R0, gN, Jt, TI, JC, ntiles, blockIdx.x);
}

// ------------------------------------------------------------------------------------
// e-e-n term on FP64 matrix cores (cord 2..6).
// Workgroup = (walker, tile of 16 electrons i).  With w = (1, u_x, u_y, u_z, 1/r_ij) the
// sums over j of k_jas_een are, for k = 1..C-1 and q = 1..C-k,
//   X_c[k][q](i, a) = sum_j A_{k,c}(i, j) B_q(j, a),  A_{k,c} = g_e(r_ij)^k w_c,  B_q = g_a(r_j)^q,
// evaluated per 16-nucleus tile with v_mfma_f64_16x16x4_f64 (rows i, columns a, contraction
// over j four at a time).  X_0 is tmp_c; X_1..3 and X_4 give dtmp_c.  One wave per (k, q):
// five accumulator tiles (the five components) and five products per step for every wave.
// The epilogue of factor_een and factor_een_gl (:14121-14175, :14875-14945) is linear in P
// and dP: every term n = (l, k, m) enters through X[k][m] g_a(r_i)^(m+l) and
// X[k][m+l] g_a(r_i)^m, so the wave owning (k, q) applies the terms in which q is m or m + l
// to all five components at once.  The per-(i, a) sums are reduced over the nuclei across
// lanes and over the waves in LDS, in a fixed order.  The q = 0 sums come from k_jas_pair
// (R0) and are applied by the waves with q = 1; the k = 0 sums S[q] = sum_{j != i} g_a(r_j)^q
// (sums over all j from k_jas_pair less the j = i term) by the waves with k = 1.
// ------------------------------------------------------------------------------------
constexpr int EM = 16;  // electrons i per workgroup

constexpr int NXCD = 8;  // accelerator complex dies of the MI355X, each with its own L2

// Workgroup -> (walker, electron tile).  Workgroups are dealt to the XCDs round-robin, so
// the tiles of a walker are given to workgroups b, b + 8, b + 16, ... of one XCD, which then
// shares the walker's g_a(r_j) table in its L2; walkers past the last full group of 8 keep
// the plain order.
__device__ __forceinline__ void een_block_work(int64_t b, int64_t W, int ntiles, int64_t& w, int& tile) {
  const int64_t full = (W / NXCD) * NXCD * ntiles;
  if (b < full) {
    const int64_t x = b % NXCD, s = b / NXCD;
    w = NXCD * (s / ntiles) + x;
    tile = (int)(s % ntiles);
  } else {
    w = b / ntiles;
    tile = (int)(b - w * ntiles);
  }
}

constexpr int EMFMA_MAX_C = 6;

typedef double d4 __attribute__((ext_vector_type(4)));

// (k, q) of wave w, k-major: (1,1), (1,2), ..., (1,C-1), (2,1), ...
constexpr int een_wave_k(int C, int w) {
  for (int k = 1; k < C; ++k) {
    if (w < C - k) return k;
    w -= C - k;
  }
  return 0;
}
constexpr int een_wave_q(int C, int w) {
  for (int k = 1; k < C; ++k) {
    if (w < C - k) return w + 1;
    w -= C - k;
  }
  return 0;
}
template <int C>
struct EenWaves {
  static constexpr int n = C * (C - 1) / 2;
};

// One pass of wave (K, Q) over nsteps x 4 electrons j held in LDS.  Lane (li, lg) feeds
// A(i = li, j = 4 s + lg) and B(j = 4 s + lg, a = li) of v_mfma_f64_16x16x4_f64.
template <int C, int K, int Q>
__device__ __forceinline__ void een_mfma_pass(d4 (&acc)[5], const double* pdG, const double* pdR, const double* xj,
                                              const double* gch, const double* xi, int JSP, int nsteps, int li,
                                              int lg) {
  for (int st = 0; st < nsteps; ++st) {
    const int jj = 4 * st + lg;
    const double Ge = pdG[jj * EM + li], ri = pdR[jj * EM + li], g = gch[jj * 16 + li];
    const double dx = xi[0] - xj[jj], dy = xi[1] - xj[JSP + jj], dz = xi[2] - xj[2 * JSP + jj];
    double Gk = Ge, gq = g;
#pragma unroll
    for (int k = 2; k <= K; ++k) Gk *= Ge;
#pragma unroll
    for (int q = 2; q <= Q; ++q) gq *= g;
    // g_e^k w with w = (1, u_x, u_y, u_z, 1/r), u = (r_i - r_j) / r, as g_e^k / r times
    // (r_i - r_j); u and 1/r are 0 for j == i (1/r is stored 0 there)
    const double t = Gk * ri;
    acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(Gk, gq, acc[0], 0, 0, 0);
    acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(t * dx, gq, acc[1], 0, 0, 0);
    acc[2] = __builtin_amdgcn_mfma_f64_16x16x4f64(t * dy, gq, acc[2], 0, 0, 0);
    acc[3] = __builtin_amdgcn_mfma_f64_16x16x4f64(t * dz, gq, acc[3], 0, 0, 0);
    acc[4] = __builtin_amdgcn_mfma_f64_16x16x4f64(t, gq, acc[4], 0, 0, 0);
  }
}

// Partial sums of one (i, a) pair, assembled into (J, grad, lap) at the end of the row:
//   J = sJ;  grad_c = h_c s1 - t0[c];  lap = kappa^2 s2 - (2 kappa / R) s1 + s3 - 2 h.t1 - 2 sL
// with h = -kappa_a (r_i - R_a)/R; d_c g^p = p h_c g^p, lap g^p = (p^2 kappa^2 - 2 p kappa/R) g^p.
struct EenRow {
  double sJ = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0, sL = 0.0;
  double t0[3] = {0.0, 0.0, 0.0}, t1[3] = {0.0, 0.0, 0.0};
};

// One role of term n: the value X[k][px] (5 components) times g^PG; JROLE when px = m + l
// (the role that carries the value J).  kk = k kappa_ee.
template <int PG, bool JROLE, bool KZERO>
__device__ __forceinline__ void een_role(EenRow& R, double cn, const double* gp, const double* X, double kk) {
  const double t = cn * gp[PG];
  const double tv = t * X[0];
  if (JROLE) R.sJ += tv;
  R.s1 += (double)PG * tv;
  R.s2 += (double)(PG * PG) * tv;
  if (!KZERO) {
    R.s3 += kk * kk * tv;
    const double tk = kk * t;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      R.t0[c] += tk * X[1 + c];
      R.t1[c] += (double)PG * tk * X[1 + c];
    }
    R.sL += tk * X[4];
  }
}

// Applies every term of wave (K, Q): X = X[K][Q] (from the products), XR = X[K][0] (Q = 1),
// X0 = X[0][Q] value only (K = 1).  Terms are visited at compile time.
template <int C, int K, int Q, int N = 0>
__device__ __forceinline__ void een_fold(EenRow& R, const double* cf, int cfs, const double* gp, const double* X,
                                         const double* XR, const double* X0, double kee) {
  constexpr Lkm<C> T{};
  if constexpr (N < T.dim) {
    constexpr int l = T.l[N], k = T.k[N], m = T.m[N];
    constexpr bool rm = (k == K && m == Q), rml = (k == K && m + l == Q);
    constexpr bool rm0 = (Q == 1 && k == K && m == 0), rml0 = (Q == 1 && k == K && m + l == 0);
    constexpr bool zm = (K == 1 && k == 0 && m == Q), zml = (K == 1 && k == 0 && m + l == Q);
    if constexpr (rm || rml || rm0 || rml0 || zm || zml) {
      const double cn = cf[cfs * N];  // c-coefficient N of this nucleus
      const double kk = (double)K * kee;
      if constexpr (rm) een_role<m + l, false, false>(R, cn, gp, X, kk);
      if constexpr (rml) een_role<m, true, false>(R, cn, gp, X, kk);
      if constexpr (rm0) een_role<m + l, false, false>(R, cn, gp, XR, kk);
      if constexpr (rml0) een_role<m, true, false>(R, cn, gp, XR, kk);
      if constexpr (zm) een_role<m + l, false, true>(R, cn, gp, X0, 0.0);
      if constexpr (zml) een_role<m, true, true>(R, cn, gp, X0, 0.0);
    }
    een_fold<C, K, Q, N + 1>(R, cf, cfs, gp, X, XR, X0, kee);
  }
}

// Wave-uniform dispatch to the (k, q) specialisations.
template <int C, int W = 0>
__device__ __forceinline__ void een_dispatch_pass(int wave, d4 (&acc)[5], const double* pdG, const double* pdR,
                                                  const double* xj, const double* gch, const double* xi, int JSP,
                                                  int nsteps, int li, int lg) {
  if constexpr (W < EenWaves<C>::n) {
    if (wave == W) {
      een_mfma_pass<C, een_wave_k(C, W), een_wave_q(C, W)>(acc, pdG, pdR, xj, gch, xi, JSP, nsteps, li, lg);
      return;
    }
    een_dispatch_pass<C, W + 1>(wave, acc, pdG, pdR, xj, gch, xi, JSP, nsteps, li, lg);
  }
}

template <int C, int W = 0>
__device__ __forceinline__ void een_dispatch_fold(int wave, EenRow& R, const double* cf, int cfs, const double* gp,
                                                  const double* X, const double* r0row, const double* srow,
                                                  double kee) {
  if constexpr (W < EenWaves<C>::n) {
    if (wave == W) {
      constexpr int K = een_wave_k(C, W), Q = een_wave_q(C, W);
      double XR[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, X0[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      if (Q == 1)
#pragma unroll
        for (int c = 0; c < 5; ++c) XR[c] = r0row[(K - 1) * 5 + c];  // X[K][0]
      // X[0][Q], value only.  The subtrahend is g^Q rounded (fma with 0), kept out of any
      // contraction so that S - g^Q is exactly 0 when electron i is the only one.
      if (K == 1) X0[0] = srow[Q - 1] - __builtin_fma(gp[Q - 1], gp[1], 0.0);
      een_fold<C, K, Q>(R, cf, cfs, gp, X, XR, X0, kee);
      return;
    }
    een_dispatch_fold<C, W + 1>(wave, R, cf, cfs, gp, X, r0row, srow, kee);
  }
}

// Dynamic LDS of k_jas_een_mfma, in doubles, for a j range of JS electrons per pass.
struct EenMfmaLds {
  int JS, JSP, C, dim, NW;
  int xt, xj, pdG, pdR, gch, ep, r0t, st, cfl, red, tot, jrow, total;
  __host__ __device__ EenMfmaLds(int JS_, int C_, int dim_, int NW_) : JS(JS_), JSP(JS_ + 4), C(C_), dim(dim_), NW(NW_) {
    int o = 0;
    xt = o;   o += EM * 3;               // tile electrons
    xj = o;   o += 3 * JSP;              // electrons j of the pass, [3][j]
    pdG = o;  o += JSP * EM;             // g_e(r_ij) = exp(-kappa r_ij), [j][i]; 0 for j == i
    pdR = o;  o += JSP * EM;             // 1 / r_ij, [j][i]; 0 for j == i
    gch = o;  o += JSP * 16;             // g_a(r_j) of the nucleus tile, [j][a]
    ep = o;   o += (C + 5) * EM * 16;    // per (i, a): g^1..g^C, h_x, h_y, h_z, 2 kappa/R, kappa^2
    r0t = o;  o += EM * (C - 1) * 5;     // q = 0 sums of the tile electrons
    st = o;   o += 16 * (C - 1);         // sum over all j of g_a(r_j)^q
    cfl = o;  o += dim * 16;             // c-coefficients of the nucleus tile, [n][a]
    red = o;  o += NW * EM * 5;
    tot = o;  o += EM * 5;               // running (J, gl) per row, summed over nuclei
    jrow = o; o += EM;
    total = o;
  }
};

template <int C>
__device__ __forceinline__ void jas_een_mfma_body(const JasDevView& v, const JasOut& o, const double* __restrict__ R0,
                                                  const double* __restrict__ gN, const double* __restrict__ Sg,
                                                  double* __restrict__ Jt, int ntiles, int JS, const int64_t b) {
  constexpr int K1 = C - 1;
  constexpr int NW = EenWaves<C>::n;
  constexpr int NT = 64 * NW;
  constexpr int DIM = Lkm<C>::dim;
  extern __shared__ double lds[];
  const EenMfmaLds L(JS, C, DIM, NW);
  double* xt = lds + L.xt;
  double* xj = lds + L.xj;
  double* pdG = lds + L.pdG;
  double* pdR = lds + L.pdR;
  double* gch = lds + L.gch;
  double* ep = lds + L.ep;
  double* r0t = lds + L.r0t;
  double* st = lds + L.st;
  double* cfl = lds + L.cfl;
  double* red = lds + L.red;
  double* tot = lds + L.tot;
  double* jrow = lds + L.jrow;
  const int Ne = (int)v.elec_num, Nn = (int)v.nucl_num;
  int64_t w;
  int tile;
  een_block_work(b, v.walk_num, ntiles, w, tile);
  const int i0 = tile * EM;
  const int tid = threadIdx.x, lane = tid & 63, li = lane & 15, lg = lane >> 4;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform (k, q) role
  const double kee = v.rescale_ee;
  const double* coord = v.coord;
  const int64_t cld = v.coord_ld;
  const int npass = (Ne + JS - 1) / JS;
  for (int t = tid; t < EM * 3; t += NT) {
    const int il = t / 3, d = t - 3 * il;
    xt[t] = (i0 + il < Ne) ? coord[w * Ne + i0 + il + d * cld] : 0.0;
  }
  for (int t = tid; t < EM * K1 * 5; t += NT) {
    const int il = t / (K1 * 5);
    r0t[t] = (i0 + il < Ne) ? R0[(w * Ne + i0) * K1 * 5 + t] : 0.0;
  }
  for (int t = tid; t < EM * 5; t += NT) tot[t] = 0.0;
  __syncthreads();
  const double xi[3] = {xt[li * 3], xt[li * 3 + 1], xt[li * 3 + 2]};  // electron i of row li

  for (int a0 = 0; a0 < Nn; a0 += 16) {
    // per-tile constants: c-coefficients, the k = 0 sums, and per (i, a) the powers of
    // g_a(r_i) and the factors of their derivatives
    for (int t = tid; t < DIM * 16; t += NT) {
      const int n = t >> 4, al = t & 15;
      cfl[t] = (a0 + al < Nn) ? v.cfull[(int64_t)n * Nn + a0 + al] : 0.0;
    }
    for (int t = tid; t < 16 * K1; t += NT) {
      const int al = t / K1;
      st[t] = (a0 + al < Nn) ? Sg[(w * Nn + a0) * K1 + t] : 0.0;
    }
    for (int t = tid; t < EM * 16; t += NT) {
      const int il = t >> 4, al = t & 15, i = i0 + il, a = a0 + al;
      double gp[C + 1], h[3] = {0.0, 0.0, 0.0}, kr = 0.0, ka2 = 0.0;
#pragma unroll
      for (int p = 0; p <= C; ++p) gp[p] = 0.0;
      if (i < Ne && a < Nn) {
        const double ka = v.kappa_en[v.type_nucl[a]];
        const double dx = xt[il * 3] - v.nucl_coord[a], dy = xt[il * 3 + 1] - v.nucl_coord[a + Nn],
                     dz = xt[il * 3 + 2] - v.nucl_coord[a + 2 * Nn];
        const double R = sqrt(dx * dx + dy * dy + dz * dz);
        const double rinv = 1.0 / R;
        const double g1 = gN[gn_index(w, a, i, gn_na(Nn), gn_jp(Ne))];  // exp(-kappa_a R)
        gp[1] = g1;
#pragma unroll
        for (int p = 2; p <= C; ++p) gp[p] = gp[p - 1] * g1;
        h[0] = -ka * (dx * rinv);
        h[1] = -ka * (dy * rinv);
        h[2] = -ka * (dz * rinv);
        kr = 2.0 * ka * rinv;
        ka2 = ka * ka;
      }
#pragma unroll
      for (int p = 1; p <= C; ++p) ep[(p - 1) * EM * 16 + t] = gp[p];
      ep[(C + 0) * EM * 16 + t] = h[0];
      ep[(C + 1) * EM * 16 + t] = h[1];
      ep[(C + 2) * EM * 16 + t] = h[2];
      ep[(C + 3) * EM * 16 + t] = kr;
      ep[(C + 4) * EM * 16 + t] = ka2;
    }

    d4 acc[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) acc[c] = d4{0.0, 0.0, 0.0, 0.0};

    for (int pass = 0; pass < npass; ++pass) {
      const int j0 = pass * JS;
      const int jn = min(JS, Ne - j0);
      const int jn4 = (jn + 3) & ~3;
      if (npass > 1 || a0 == 0) {
        if (npass > 1) __syncthreads();  // previous pass done with the pair data
        for (int t = tid; t < 3 * jn4; t += NT) {
          const int d = t / jn4, jj = t - d * jn4;
          xj[d * L.JSP + jj] = (jj < jn) ? coord[w * Ne + j0 + jj + d * cld] : 0.0;
        }
        for (int t = tid; t < EM * jn4; t += NT) {
          const int il = t & (EM - 1), jj = t / EM;
          const int i = i0 + il, j = j0 + jj;
          double Ge = 0.0, ri = 0.0;
          if (i < Ne && jj < jn && i != j) {
            const double dx = xt[il * 3] - coord[w * Ne + j];
            const double dy = xt[il * 3 + 1] - coord[w * Ne + j + cld];
            const double dz = xt[il * 3 + 2] - coord[w * Ne + j + 2 * cld];
            double d;
            dist_and_inverse(dx * dx + dy * dy + dz * dz, d, ri);
            Ge = exp_neg_deg<13, true>(-kee * d);
          }
          pdG[t] = Ge;  // [jj][il]
          pdR[t] = ri;
        }
      }
      for (int t = tid; t < 16 * jn4; t += NT) {  // consecutive threads: consecutive doubles
        const int jj = t >> 4, al = t & 15;
        const int a = a0 + al;
        gch[jj * 16 + al] = (a < Nn && jj < jn) ? gN[gn_index(w, a, j0 + jj, gn_na(Nn), gn_jp(Ne))] : 0.0;
      }
      __syncthreads();
      een_dispatch_pass<C>(wave, acc, pdG, pdR, xj, gch, xi, L.JSP, jn4 >> 2, li, lg);
      if (pass + 1 < npass) __syncthreads();  // before gch of the next pass is written
    }

    // Epilogue: lane (li, lg) holds rows lg + 4 r (register r), column a = a0 + li.
    const bool va = a0 + li < Nn;
#pragma unroll 1
    for (int r = 0; r < 4; ++r) {
      const int il = lg + 4 * r, i = i0 + il;
      double out[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      if (va && i < Ne) {
        const int pidx = il * 16 + li;
        double gp[C + 1];
        gp[0] = 1.0;
#pragma unroll
        for (int p = 1; p <= C; ++p) gp[p] = ep[(p - 1) * EM * 16 + pidx];
        double X[5];
#pragma unroll
        for (int c = 0; c < 5; ++c) {
          const d4 x = acc[c];
          X[c] = r == 0 ? x[0] : r == 1 ? x[1] : r == 2 ? x[2] : x[3];
        }
        EenRow R;
        een_dispatch_fold<C>(wave, R, cfl + li, 16, gp, X, r0t + il * K1 * 5, st + li * K1, kee);
        const double h0 = ep[(C + 0) * EM * 16 + pidx], h1 = ep[(C + 1) * EM * 16 + pidx],
                     h2 = ep[(C + 2) * EM * 16 + pidx];
        const double kr = ep[(C + 3) * EM * 16 + pidx], ka2 = ep[(C + 4) * EM * 16 + pidx];
        out[0] = R.sJ;
        out[1] = h0 * R.s1 - R.t0[0];
        out[2] = h1 * R.s1 - R.t0[1];
        out[3] = h2 * R.s1 - R.t0[2];
        out[4] = ka2 * R.s2 - kr * R.s1 + R.s3 - 2.0 * (h0 * R.t1[0] + h1 * R.t1[1] + h2 * R.t1[2]) - 2.0 * R.sL;
      }
      // sum over the 16 nuclei of the tile: lanes li = 0..15 of this row group
#pragma unroll
      for (int c = 0; c < 5; ++c) {
        out[c] = row_sum16(out[c]);
      }
      if (li == 0)
#pragma unroll
        for (int c = 0; c < 5; ++c) red[(wave * EM + il) * 5 + c] = out[c];
    }
    __syncthreads();
    for (int t = tid; t < EM * 5; t += NT) {
      double x = tot[t];
#pragma unroll
      for (int wv = 0; wv < NW; ++wv) x += red[wv * EM * 5 + t];
      tot[t] = x;
    }
    // the next tile rewrites cfl, st, ep and gch only after this barrier
    __syncthreads();
  }
  for (int t = tid; t < EM * 5; t += NT) {
    const int il = t / 5, c = t - 5 * il, i = i0 + il;
    if (c == 0)
      jrow[il] = (i < Ne) ? tot[t] : 0.0;
    else if (i < Ne)
      o.feen_gl[w * 4 * Ne + (int64_t)(c - 1) * Ne + i] = tot[t];
  }
  __syncthreads();
  if (tid == 0) {
    double sj = 0.0;
    for (int il = 0; il < EM; ++il) sj += jrow[il];
    Jt[w * ntiles + tile] = sj;
  }
}

template <int C>
__global__ void __launch_bounds__(64 * EenWaves<C>::n) k_jas_een_mfma(JasDevView v, JasOut o,
                                                                     const double* __restrict__ R0,
                                                                     const double* __restrict__ gN,
                                                                     const double* __restrict__ Sg,
                                                                     double* __restrict__ Jt, int ntiles, int JS) {
  jas_een_mfma_body<C>(v, o, R0, gN, Sg, Jt, ntiles, JS, blockIdx.x);
}

// ------------------------------------------------------------------------------------
// e-e-n on FP64 matrix cores, resident variant: used when the pair data of a tile of 16
// electrons i against every electron j, and g_a(r_j) for every nucleus, fit in LDS together.
// The products and the epilogue are those of k_jas_een_mfma; the work is cut into units
// (k, q, 16-nucleus tile), each a full contraction over j followed by its fold.  Units of the
// same k and tile go in groups of two (q, q + 1) that share the A operand; the groups are
// dealt to EQ_WAVES waves (two per SIMD) so that every wave gets the same number of units,
// and no barrier separates them.  On gfx950 the FP64 matrix instructions and all other
// vector instructions share one issue budget, so the loop keeps its vector work small: the
// data of electron j is one LDS record (g_e(r_ij) and 1/r_ij of the 16 electrons i, then
// x_j, y_j, z_j) addressed from one register, and a group builds its A operand once for
// ten matrix instructions.  Each wave keeps its (i, component) partial sums in registers
// across its groups; the sums over nuclei (lanes) and waves are done once at the end, in a
// fixed order.  The kernel also does the work of k_jas_pair for its electrons: per walker
// g_a(r_j) and the k = 0 sums, per tile the e-e and e-n terms (value, gradient,
// Laplacian) and the q = 0 sums, two electrons i per wave; so that pass and its trip
// through memory (g_a(r_j), R0, S) are not needed.
// ------------------------------------------------------------------------------------
constexpr int EQ_WAVES = 8;
constexpr int JREC = 36;  // doubles per electron-j record: G[16], R[16], x, y, z, pad

// Dynamic LDS of k_jas_een_res, in doubles, every array on a 16-byte boundary.  JP = Ne
// rounded up to 4, NA16 = 16 x number of nucleus tiles.
// Doubles of the pair parameters the resident kernel keeps in LDS (EenResLds::prm): b_vector
// [bord + 1], a_vector [type][aord + 1], per type kappa_en, 1 / kappa_en and asymp_jasa
// [3][type], and the type of each nucleus of the padded table [NA16].
__host__ __device__ inline int een_prm_doubles(const JasDevView& v, int NA16) {
  const int nt = v.type_num > 0 ? (int)v.type_num : 0;
  return (int)v.bord + 1 + nt * ((int)v.aord + 4) + NA16;
}

struct EenResLds {
  int xt, jr, gch, nuc, enr, r0t, st, cfl, red, pee, own, prm, total;
  __host__ __device__ EenResLds(int JP, int NA16, int C, int dim, int nprm) {
    int o = 0;
    auto take = [&o](int n) { const int at = o; o = (o + n + 1) & ~1; return at; };
    xt = take(EM * 3);                 // tile electrons i
    jr = take(JP * JREC);              // electron-j records; G = R = 0 for j == i and padding
    gch = take(NA16 * JP);             // g_a(r_j), [tile][j][16] (as gn_index)
    nuc = take(4 * NA16);              // nucleus x, y, z, kappa_a, [4][a]
    enr = take(EM * NA16);             // 1 / |r_i - R_a|, [i][a]
    r0t = take(EM * (C - 1) * 5);      // q = 0 sums of the tile electrons
    st = take(NA16 * (C - 1));         // sum over all j of g_a(r_j)^q, [a][q-1]
    cfl = take(dim * NA16);            // c-coefficients, [n][a]
    red = take(EQ_WAVES * EM * 5);     // per-wave row sums
    pee = take(2 * EQ_WAVES);          // per-wave e-e and e-n value sums of the tile
    int ngroup = 0;                    // groups of units per nucleus tile
    for (int k = 1; k < C; ++k) ngroup += (C - k + 1) / 2;
    own = take(ngroup * (NA16 / 16));  // wave of each group, as int: full tiles, then tiles of <= 8 rows
    prm = take(nprm);                  // pair parameters (een_prm_doubles)
    total = o;
  }
};

typedef __attribute__((address_space(3))) void lds_void;

// Copies n doubles from global memory to LDS with the gfx950 global-to-LDS loads, 4 bytes
// per lane, for any 8-byte aligned src and dst; the waves of the block take 64-lane slices in
// turn.  The copy is asynchronous: wait (s_waitcnt) and synchronise before reading dst.
__device__ __forceinline__ void lds_copy4(double* dst, const double* src, int n, int wave, int lane) {
  const float* s = reinterpret_cast<const float*>(src);
  float* d = reinterpret_cast<float*>(dst);
  for (int base = wave * 64; base < 2 * n; base += EQ_WAVES * 64) {
    const int t = base + lane;
    if (t < 2 * n) __builtin_amdgcn_global_load_lds(s + t, (lds_void*)(d + base), 4, 0, 0);
  }
}

// The fold of one unit, regrouped: every role of a term that multiplies the same X (the
// unit's products X[K][Q], or X[K][0] when Q = 1, or X[0][Q] when K = 1) differs only in
// t = c_n g^PG, so the roles are first summed into T = sum t, TJ = sum of t over the roles
// that carry J, T1 = sum PG t, T2 = sum PG^2 t, and X is applied once (een_apply).
struct EenTs {
  double T = 0.0, TJ = 0.0, T1 = 0.0, T2 = 0.0;
};

template <int PG, bool JROLE>
__device__ __forceinline__ void een_tadd(EenTs& s, double cn, const double* gp) {
  const double t = cn * gp[PG];
  s.T += t;
  if (JROLE) s.TJ += t;
  if (PG != 0) {
    s.T1 += (double)PG * t;
    s.T2 += (double)(PG * PG) * t;
  }
}

// Role sums of unit (K, Q): m with X[K][Q], r with X[K][0] (Q = 1), z with X[0][Q] (K = 1).
template <int C, int K, int Q, int N = 0>
__device__ __forceinline__ void een_tsums(EenTs& m, EenTs& r, EenTs& z, const double* cf, int cfs, const double* gp) {
  constexpr Lkm<C> T{};
  if constexpr (N < T.dim) {
    constexpr int l = T.l[N], k = T.k[N], mm = T.m[N];
    constexpr bool rm = (k == K && mm == Q), rml = (k == K && mm + l == Q);
    constexpr bool rm0 = (Q == 1 && k == K && mm == 0), rml0 = (Q == 1 && k == K && mm + l == 0);
    constexpr bool zm = (K == 1 && k == 0 && mm == Q), zml = (K == 1 && k == 0 && mm + l == Q);
    if constexpr (rm || rml || rm0 || rml0 || zm || zml) {
      const double cn = cf[cfs * N];  // c-coefficient N of this nucleus
      if constexpr (rm) een_tadd<mm + l, false>(m, cn, gp);
      if constexpr (rml) een_tadd<mm, true>(m, cn, gp);
      if constexpr (rm0) een_tadd<mm + l, false>(r, cn, gp);
      if constexpr (rml0) een_tadd<mm, true>(r, cn, gp);
      if constexpr (zm) een_tadd<mm + l, false>(z, cn, gp);
      if constexpr (zml) een_tadd<mm, true>(z, cn, gp);
    }
    een_tsums<C, K, Q, N + 1>(m, r, z, cf, cfs, gp);
  }
}

// Per-(i, a) factors of the epilogue: h = -kappa_a (r_i - R_a)/R, kr = 2 kappa_a / R, ka2 = kappa_a^2.
struct EenEp {
  double h0, h1, h2, kr, ka2;
};

// Which role sums of unit (K, Q) can be non-zero, from the c-vector layout alone: bit 0 T
// (any term), bit 1 TJ (a term that carries J), bit 2 T1 and T2 (a term with PG != 0).
// ROLE 0: m, 1: r, 2: z, with the conditions of een_tsums.
template <int C, int K, int Q, int ROLE>
constexpr unsigned een_role_mask() {
  constexpr Lkm<C> T{};
  unsigned f = 0;
  for (int N = 0; N < T.dim; ++N) {
    const int l = T.l[N], k = T.k[N], mm = T.m[N];
    bool plain = false, jrole = false;  // een_tadd<mm + l, false> and een_tadd<mm, true>
    if (ROLE == 0) {
      plain = k == K && mm == Q;
      jrole = k == K && mm + l == Q;
    } else if (ROLE == 1) {
      plain = Q == 1 && k == K && mm == 0;
      jrole = Q == 1 && k == K && mm + l == 0;
    } else {
      plain = K == 1 && k == 0 && mm == Q;
      jrole = K == 1 && k == 0 && mm + l == Q;
    }
    if (plain) f |= (mm + l != 0) ? 5u : 1u;
    if (jrole) f |= (mm != 0) ? 7u : 3u;
  }
  return f;
}

// Adds to the row outputs o = (J, grad, lap) the epilogue of role sums s over X, which is
// linear in the partial sums of k_jas_een_mfma (EenRow):
//   o0 += TJ X0,   o_c += h_c T1 X0 - kk T X_c,
//   o4 += ka2 T2 X0 - kr T1 X0 + kk^2 T X0 - 2 kk T1 (h . X_123) - 2 kk T X4;
// VALUE_ONLY (the X[0][Q] roles, k = 0): the terms without kk.  F (een_role_mask) drops the
// terms whose role sums are zero by construction: they would add exact zeros, which the
// compiler must otherwise compute (kk or X could be infinite).
// JONLY: J alone (the values-only evaluation), the same additions into o[0] in the same order.
template <bool VALUE_ONLY, unsigned F = 7u, bool JONLY = false>
__device__ __forceinline__ void een_apply(double (&o)[5], const EenTs& s, const double* X, double kk, const EenEp& e) {
  if constexpr ((F & 1u) != 0) {
    if constexpr ((F & 2u) != 0) o[0] += s.TJ * X[0];
    if constexpr (JONLY) return;
    if constexpr ((F & 4u) != 0) {
      const double t1x = s.T1 * X[0];
      o[1] += e.h0 * t1x;
      o[2] += e.h1 * t1x;
      o[3] += e.h2 * t1x;
      o[4] += e.ka2 * s.T2 * X[0] - e.kr * t1x;
    }
    if constexpr (!VALUE_ONLY) {
      const double kt = kk * s.T;
      o[1] -= kt * X[1];
      o[2] -= kt * X[2];
      o[3] -= kt * X[3];
      if constexpr ((F & 4u) != 0) {
        const double kt1 = kk * s.T1;
        o[4] += kk * kt * X[0] - 2.0 * kt1 * (e.h0 * X[1] + e.h1 * X[2] + e.h2 * X[3]) - 2.0 * kt * X[4];
      } else {
        o[4] += kk * kt * X[0] - 2.0 * kt * X[4];
      }
    }
  }
}

// Groups of units that share k: (k, q) and (k, q + 1) for q = 1, 3, 5, ... (one unit when
// q + 1 > C - k).  A group shares the A operand G^k w of its units; it is the unit of work
// dealt to the waves.  Groups of two come first in this enumeration order: the "pairs"
// (k-major), then the single units.
constexpr int een_npair(int C) {
  int n = 0;
  for (int k = 1; k < C; ++k) n += (C - k) / 2;
  return n;
}
constexpr int een_nsingle(int C) {
  int n = 0;
  for (int k = 1; k < C; ++k) n += (C - k) & 1;
  return n;
}
// (k, first q) of pair p / of single unit p
constexpr int een_pair_k(int C, int p) {
  for (int k = 1; k < C; ++k) {
    if (p < (C - k) / 2) return k;
    p -= (C - k) / 2;
  }
  return 0;
}
constexpr int een_pair_q(int C, int p) {
  for (int k = 1; k < C; ++k) {
    if (p < (C - k) / 2) return 1 + 2 * p;
    p -= (C - k) / 2;
  }
  return 0;
}
constexpr int een_single_k(int C, int p) {
  for (int k = 1; k < C; ++k) {
    if ((C - k) & 1) {
      if (p == 0) return k;
      --p;
    }
  }
  return 0;
}
constexpr int een_single_q(int C, int p) { return C - een_single_k(C, p); }

// The contraction of group (K, Q .. Q + NQ - 1) over nsteps x 4 electrons j; lane (li, lg)
// feeds A(i = li, j = 4 s + lg) and B(j = 4 s + lg, a = li) of v_mfma_f64_16x16x4_f64.
// MINI: the last nucleus tile holds at most 4 nuclei and the products run on
// v_mfma_f64_4x4x4f64 (4 blocks), whose lane l = 16 m + 4 b + n holds A(block b, row n,
// k = m), B(b, k = m, column n) and D(b, row m, column n): with the blocks as the four
// groups of 4 electrons, a lane's A element is the same A(i = li, j = 4 s + lg), its B element
// is B(j = 4 s + lg, a = li mod 4), and its one result is row lg + 4 (li / 4), column li mod 4.
// ROWS: the tile holds at most 8 electrons (the last tile of a walker); the same 4x4x4
// products with the blocks as the four groups of 4 nuclei, once for rows 0-3 and once for
// rows 4-7 (h = 0, 1): lane l = 16 m + 4 b + n feeds A(i = 4 h + n, j = 4 s + m) and
// B(j = 4 s + m, a = 4 b + n), the B element of the full tile, and holds row 4 h + lg,
// column li of the result, so that half the matrix work of a full tile covers its rows.
// MINI2: the last nucleus tile holds at most 2 nuclei and the group is a pair (q, q + 1):
// the four columns of the 4x4x4 products are (nucleus n mod 2, power q + n / 2), so one
// product per component covers both units; lane n < 2 folds its own column (q) with its
// partner's l ^ 2 (q + 1), and lanes n >= 2 fold against the padding nuclei 2, 3 of the
// tile, whose coefficients are zero.
enum { EEN_FULL = 0, EEN_MINI = 1, EEN_ROWS = 2, EEN_MINI2 = 3 };
template <int MODE>
struct EenAccT {
  using type = d4;
};
template <>
struct EenAccT<EEN_MINI> {
  using type = double;
};
template <>
struct EenAccT<EEN_MINI2> {
  using type = double;
};
template <>
struct EenAccT<EEN_ROWS> {
  using type = double[2];
};
template <int MODE>
using EenAcc = typename EenAccT<MODE>::type;

// NC: components of the products, 5 (value, gradient, Laplacian sums) or 1 (the value alone)
template <int K, int Q, int NQ, int NC>
__device__ __forceinline__ void een_res_pass_rows(double (&acc)[NQ][NC][2], const double* jr, const double* gt,
                                                  const double* xt, int nsteps, int li, int lg) {
  const int n4 = li & 3;
  const double* pj = jr + lg * JREC;
  const double* pg = gt + lg * 16 + li;
  const double xa = xt[3 * n4], ya = xt[3 * n4 + 1], za = xt[3 * n4 + 2];
  const double xb = xt[3 * (4 + n4)], yb = xt[3 * (4 + n4) + 1], zb = xt[3 * (4 + n4) + 2];
  for (int st = 0; st < nsteps; ++st) {
    const double Ga = pj[n4], ra = pj[16 + n4], Gb = pj[4 + n4], rb = pj[20 + n4], g = pg[0];
    const double xj = pj[32], yj = pj[33], zj = pj[34];
    pj += 4 * JREC;
    pg += 64;
    double Gka = Ga, Gkb = Gb, gq = g;
#pragma unroll
    for (int k = 2; k <= K; ++k) {
      Gka *= Ga;
      Gkb *= Gb;
    }
#pragma unroll
    for (int q = 2; q <= Q; ++q) gq *= g;
    const double ta = Gka * ra, tb = Gkb * rb;
    const double Aa[5] = {Gka, ta * (xa - xj), ta * (ya - yj), ta * (za - zj), ta};
    const double Ab[5] = {Gkb, tb * (xb - xj), tb * (yb - yj), tb * (zb - zj), tb};
#pragma unroll
    for (int n = 0; n < NQ; ++n) {
      if (n > 0) gq *= g;
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        acc[n][c][0] = __builtin_amdgcn_mfma_f64_4x4x4f64(Aa[c], gq, acc[n][c][0], 0, 0, 0);
        acc[n][c][1] = __builtin_amdgcn_mfma_f64_4x4x4f64(Ab[c], gq, acc[n][c][1], 0, 0, 0);
      }
    }
  }
}

template <int K, int Q, int NQ, int MODE, int NC>
__device__ __forceinline__ void een_res_pass(EenAcc<MODE> (&acc)[NQ][NC], const double* jr, const double* gt,
                                             const double (&xi)[3], int nsteps, int li, int lg) {
  constexpr bool MINI = MODE == EEN_MINI || MODE == EEN_MINI2;
  constexpr bool MINI2 = MODE == EEN_MINI2;
  static_assert(!MINI2 || NQ == 1, "MINI2 holds both units of a pair in one accumulator");
  const double* pj = jr + lg * JREC;
  const double* pi = pj + li;
  const double* pg = gt + lg * 16 + (MINI2 ? (li & 1) : MINI ? (li & 3) : li);
  const bool qup = MINI2 && (li & 2) != 0;  // MINI2: this column holds power q + 1
  // the next step's operands are read before this step's products are issued
  double nGe = pi[0], nri = pi[16], ng = pg[0], nxj = pj[32], nyj = pj[33], nzj = pj[34];
  for (int st = 0; st < nsteps; ++st) {
    const double Ge = nGe, ri = nri, g = ng;
    const double dx = xi[0] - nxj, dy = xi[1] - nyj, dz = xi[2] - nzj;
    const bool more = st + 1 < nsteps;  // the last step re-reads its own records
    pj += more ? 4 * JREC : 0;
    pi += more ? 4 * JREC : 0;
    pg += more ? 64 : 0;
    nGe = pi[0], nri = pi[16], ng = pg[0], nxj = pj[32], nyj = pj[33], nzj = pj[34];
    double Gk = Ge, gq = g;
#pragma unroll
    for (int k = 2; k <= K; ++k) Gk *= Ge;
#pragma unroll
    for (int q = 2; q <= Q; ++q) gq *= g;
    if constexpr (MINI2) gq = qup ? gq * g : gq;
    // g_e^k w with w = (1, u_x, u_y, u_z, 1/r), u = (r_i - r_j) / r, as g_e^k / r times
    // (r_i - r_j); u and 1/r are 0 for j == i (1/r is stored 0 there)
    const double t = Gk * ri;
    const double A[5] = {Gk, t * dx, t * dy, t * dz, t};
#pragma unroll
    for (int n = 0; n < NQ; ++n) {
      if (n > 0) gq *= g;
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        if constexpr (MINI)
          acc[n][c] = __builtin_amdgcn_mfma_f64_4x4x4f64(A[c], gq, acc[n][c], 0, 0, 0);
        else
          acc[n][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(A[c], gq, acc[n][c], 0, 0, 0);
      }
    }
  }
}

// One group: contraction, then the fold of its four rows into ra.  Lane (li, lg) holds rows
// lg + 4 r (register r of each accumulator), column a = 16 at + li.
struct EenResCtx {
  const double *jr, *gt, *xt, *cfl, *r0t, *st, *enr;
  double nx, ny, nz, ka, kee;
  int nsteps, li, lg, i0, Ne, a, NA16;
};

// The fold of unit (K, Q) of row il: its products X with the m roles, X[K][0] (the q = 0 sums
// of the tile electrons) with the r roles when Q = 1, X[0][Q] with the z roles when K = 1.
template <int C, int K, int Q, bool JONLY>
__device__ __forceinline__ void een_fold_unit(const EenResCtx& x, int il, const double (&X)[5], const EenTs& tm,
                                              const EenTs& tr, const EenTs& tz, const double* gp, double kk,
                                              const EenEp& e, double (&out)[5]) {
  constexpr int K1 = C - 1;
  een_apply<false, een_role_mask<C, K, Q, 0>(), JONLY>(out, tm, X, kk, e);
  if constexpr (Q == 1) {
    const double* XR = x.r0t + il * K1 * 5 + (K - 1) * 5;
    een_apply<false, een_role_mask<C, K, Q, 1>(), JONLY>(out, tr, XR, kk, e);
  }
  if constexpr (K == 1) {
    // X[0][Q], value only; g^Q rounded and kept out of contraction (see een_dispatch_fold)
    const double X0[1] = {x.st[x.a * K1 + Q - 1] - __builtin_fma(gp[Q - 1], gp[1], 0.0)};
    een_apply<true, een_role_mask<C, K, Q, 2>(), JONLY>(out, tz, X0, 0.0, e);
  }
}

// Fold of row r of the products (X, 5 components per unit of the group) of lane column
// x.a into out: the role sums of each unit applied to its products, and with q = 1 or k = 1
// the roles on the q = 0 and k = 0 sums.  gcol is the lane's column in the tile.
template <int C, int K, int Q, int NQ, bool JONLY>
__device__ __forceinline__ void een_res_fold_row(const EenResCtx& x, int il, int gcol, const double (&X)[NQ][5],
                                                 double (&out)[5]) {
  const int i = min(x.i0 + il, x.Ne - 1);
  double gp[C + 1];
  gp[0] = 1.0;
  gp[1] = x.gt[i * 16 + gcol];  // g_a(r_i)
#pragma unroll
  for (int p = 2; p <= C; ++p) gp[p] = gp[p - 1] * gp[1];
  const double rinv = x.enr[il * x.NA16 + x.a];
  EenEp e;
  e.h0 = -x.ka * ((x.xt[il * 3] - x.nx) * rinv);
  e.h1 = -x.ka * ((x.xt[il * 3 + 1] - x.ny) * rinv);
  e.h2 = -x.ka * ((x.xt[il * 3 + 2] - x.nz) * rinv);
  e.kr = 2.0 * x.ka * rinv;
  e.ka2 = x.ka * x.ka;
  const double kk = (double)K * x.kee;
#pragma unroll
  for (int n = 0; n < NQ; ++n) {
    EenTs tm, tr, tz;
    if (n == 0) {
      een_tsums<C, K, Q>(tm, tr, tz, x.cfl + x.a, x.NA16, gp);
      een_fold_unit<C, K, Q, JONLY>(x, il, X[n], tm, tr, tz, gp, kk, e, out);
    } else {
      constexpr int Q1 = NQ > 1 ? Q + 1 : Q;
      een_tsums<C, K, Q1>(tm, tr, tz, x.cfl + x.a, x.NA16, gp);
      een_fold_unit<C, K, Q1, JONLY>(x, il, X[n], tm, tr, tz, gp, kk, e, out);
    }
  }
}

// One group: contraction, then the fold.  Full tile: lane (li, lg) holds rows lg + 4 r
// (register r of each accumulator) of column a = 16 at + li, folded into ra[r].  MINI: one
// row lg + 4 (li / 4) of column a = 16 at + li mod 4 per lane, folded into that row of ra.  No branch on
// the row or the nucleus, so the rows interleave: a padding nucleus has zero c-coefficients,
// kappa, 1/R, products and k = 0 sums and adds exactly 0; a row past the last electron is
// discarded by the caller.
template <int C, int K, int Q, int NQ, int MODE, int NC>
__device__ __forceinline__ void een_res_group(const EenResCtx& x, const double (&xi)[3], double (&ra)[4][5]) {
  constexpr bool MINI = MODE == EEN_MINI;
  constexpr bool JONLY = NC == 1;
  if constexpr (MODE == EEN_ROWS) {
    double acc[NQ][NC][2];
#pragma unroll
    for (int n = 0; n < NQ; ++n)
#pragma unroll
      for (int c = 0; c < NC; ++c) acc[n][c][0] = acc[n][c][1] = 0.0;
    een_res_pass_rows<K, Q, NQ, NC>(acc, x.jr, x.gt, x.xt, x.nsteps, x.li, x.lg);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      double X[NQ][5];
#pragma unroll
      for (int n = 0; n < NQ; ++n)
#pragma unroll
        for (int c = 0; c < 5; ++c) X[n][c] = c < NC ? acc[n][c < NC ? c : 0][h] : 0.0;
      een_res_fold_row<C, K, Q, NQ, JONLY>(x, x.lg + 4 * h, x.li, X, ra[h]);
    }
    return;
  } else if constexpr (MODE == EEN_MINI2) {
    double acc[1][NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[0][c] = 0.0;
    een_res_pass<K, Q, 1, EEN_MINI2, NC>(acc, x.jr, x.gt, xi, x.nsteps, x.li, x.lg);
    double X[2][5];
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      X[0][c] = c < NC ? acc[0][c < NC ? c : 0] : 0.0;
      X[1][c] = c < NC ? dpp_f64<DPP_XOR2>(acc[0][c < NC ? c : 0]) : 0.0;  // column n ^ 2: power q + 1
    }
    double out[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    een_res_fold_row<C, K, Q, 2, JONLY>(x, x.lg + 4 * (x.li >> 2), x.li & 3, X, out);
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int c = 0; c < NC; ++c) ra[r][c] += (r == (x.li >> 2)) ? out[c] : 0.0;
  } else {
  EenAcc<MODE> acc[NQ][NC];
#pragma unroll
  for (int n = 0; n < NQ; ++n)
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      if constexpr (MINI)
        acc[n][c] = 0.0;
      else
        acc[n][c] = d4{0.0, 0.0, 0.0, 0.0};
    }
  een_res_pass<K, Q, NQ, MODE, NC>(acc, x.jr, x.gt, xi, x.nsteps, x.li, x.lg);
  if constexpr (MINI) {
    double X[NQ][5];
#pragma unroll
    for (int n = 0; n < NQ; ++n)
#pragma unroll
      for (int c = 0; c < 5; ++c) X[n][c] = c < NC ? acc[n][c < NC ? c : 0] : 0.0;
    double out[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    een_res_fold_row<C, K, Q, NQ, JONLY>(x, x.lg + 4 * (x.li >> 2), x.li & 3, X, out);
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int c = 0; c < NC; ++c) ra[r][c] += (r == (x.li >> 2)) ? out[c] : 0.0;
  } else {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      double X[NQ][5];
#pragma unroll
      for (int n = 0; n < NQ; ++n)
#pragma unroll
        for (int c = 0; c < 5; ++c) X[n][c] = c < NC ? acc[n][c < NC ? c : 0][r] : 0.0;
      een_res_fold_row<C, K, Q, NQ, JONLY>(x, x.lg + 4 * r, x.li, X, ra[r]);
    }
  }
  }
}

template <int C, int NC, int G = 0>
__device__ __forceinline__ void een_res_dispatch(int g, int mode, const EenResCtx& x, const double (&xi)[3],
                                                 double (&ra)[4][5]) {
  constexpr int NPR = een_npair(C);
  if constexpr (G < NPR + een_nsingle(C)) {
    if (g == G) {
      constexpr int K = G < NPR ? een_pair_k(C, G) : een_single_k(C, G - NPR);
      constexpr int Q = G < NPR ? een_pair_q(C, G) : een_single_q(C, G - NPR);
      constexpr int NQ = G < NPR ? 2 : 1;
      if constexpr (NQ == 2) {
        if (mode == EEN_MINI2) {
          een_res_group<C, K, Q, NQ, EEN_MINI2, NC>(x, xi, ra);
          return;
        }
      }
      if (mode == EEN_MINI || mode == EEN_MINI2)
        een_res_group<C, K, Q, NQ, EEN_MINI, NC>(x, xi, ra);
      else if (mode == EEN_ROWS)
        een_res_group<C, K, Q, NQ, EEN_ROWS, NC>(x, xi, ra);
      else
        een_res_group<C, K, Q, NQ, EEN_FULL, NC>(x, xi, ra);
      return;
    }
    een_res_dispatch<C, NC, G + 1>(g, mode, x, xi, ra);
  }
}

// Butterfly reduce-scatter of vec[0 .. N) over the 32 lanes l32 of a half-wave, step S
// first: at each step a lane keeps the half of its entries selected by bit S of l32, adds
// the partner's copy of that half (a lane that differs in bit S), and hands over the other
// half.  After the
// steps S = 16, 8, 4, 2, 1, vec[m] of lane l32 is the sum over the 32 lanes of entry
// l32 * N / 32 + m, added in a fixed order.
template <int N, int S>
__device__ __forceinline__ void een_butterfly(double (&vec)[N], int l32) {
  if constexpr (S >= 1) {
    constexpr int H = N * S / 32;  // half of the N S / 16 entries held before step S
    if constexpr (S == 16) {
      // the rows of 16 lanes swap halves: the lower row keeps vec[m] and receives the upper
      // row's vec[m], the upper row keeps vec[m + H] and receives the lower row's
#pragma unroll
      for (int m = 0; m < H; ++m) vec[m] = swap_sum_f64<true>(vec[m], vec[m + H]);
    } else {
      // within a row the partner is l32 ^ 15, ^ 7, ^ 2, ^ 1 (S = 8, 4, 2, 1): it differs in
      // bit S and holds the same entries; with the rows these partners reach all 32 lanes
      constexpr int CTRL = S == 8 ? DPP_XOR15 : S == 4 ? DPP_XOR7 : S == 2 ? DPP_XOR2 : DPP_XOR1;
      const bool up = (l32 & S) != 0;
#pragma unroll
      for (int m = 0; m < H; ++m) {
        const double keep = up ? vec[m + H] : vec[m];
        const double send = up ? vec[m] : vec[m + H];
        vec[m] = keep + dpp_f64<CTRL>(send);
      }
    }
    een_butterfly<N, S / 2>(vec, l32);
  }
}

// Persistent: the grid holds one workgroup per CU (the LDS allows no more) and workgroup b
// takes the items (walker, tile) in [b N / grid, (b + 1) N / grid) in walker-major order, so
// the per-walker inputs (g_a(r_j) of every nucleus, x_j, the q sums) are copied to LDS once
// per walker instead of once per tile, the constant ones (nucleus table, c-coefficients)
// once per workgroup, and no workgroup launch separates two tiles.
// VAL: the values alone (e-e, e-n and e-e-n factors): the products of the value component
// only (one matrix instruction of five per step), the fold of J alone, no GL written; the
// values equal those of the full evaluation bit for bit (the same operations on them).
template <int C, bool VAL>
__device__ __forceinline__ void jas_een_res_body(const JasDevView& v, const JasOut& o, double* __restrict__ Jt,
                                                 double* __restrict__ JtP, int ntiles, const int bx, const int gx) {
  constexpr int K1 = C - 1;
  constexpr int NC = VAL ? 1 : 5;
  constexpr int NT = 64 * EQ_WAVES;
  constexpr int DIM = Lkm<C>::dim;
  extern __shared__ double lds[];
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int nitem = (int)v.walk_num * ntiles;  // < 2^31 (checked by the launcher)
  const int it0 = (int)((int64_t)nitem * bx / gx), it1 = (int)((int64_t)nitem * (bx + 1) / gx);
  constexpr int NPR = een_npair(C), NSG = een_nsingle(C);
  {
    // Once per workgroup (ready at the first item's barrier): the constant inputs, and the
    // owner wave of every group of units of a tile.  Waves w and w + 4 of the workgroup
    // share a SIMD (measured on gfx950), on which the FP64 matrix instructions and all other
    // vector instructions issue one at a time; a SIMD is best used while both its waves have
    // work.  So every wave gets the same number of groups (one more for some when they do
    // not divide) and the SIMDs the same cost: the groups, in decreasing cost (the pairs of
    // every full nucleus tile, its single units, then, when the last tile holds at most 4
    // nuclei, its pairs and single units on the 4x4x4 products), go to the least loaded SIMD
    // that has room, and there to the wave with fewer groups.  A group's cost is its steps
    // times the matrix cycles of a step (64 per 16x16x4, 16 per 4x4x4 product) plus about
    // 4.5 cycles per vector instruction, plus its fold.  A second table serves the electron
    // tiles of at most 8 rows, whose full nucleus tiles run on the 4x4x4 products by rows.
    const int Nn = (int)v.nucl_num, NA = gn_na(Nn), NA16 = 16 * NA;
    const EenResLds L(gn_jp((int)v.elec_num), NA16, C, DIM, een_prm_doubles(v, NA16));
    lds_copy4(lds + L.nuc, v.nuc_pad, 4 * NA16, wave, threadIdx.x & 63);
    lds_copy4(lds + L.cfl, v.cfull_pad, DIM * NA16, wave, threadIdx.x & 63);
    {
      // the pair parameters, read by every pair of every tile: from LDS, the reads of the
      // Horner sums issue without waiting on global memory
      double* prm = lds + L.prm;
      const int nb = (int)v.bord + 1, na1 = (int)v.aord + 1, nt = v.type_num > 0 ? (int)v.type_num : 0;
      for (int t = threadIdx.x; t < nb; t += NT) prm[t] = v.b_vector[t];
      for (int t = threadIdx.x; t < nt * na1; t += NT) prm[nb + t] = v.a_vector[t];
      for (int t = threadIdx.x; t < nt; t += NT) {
        const double ka = v.kappa_en[t];
        prm[nb + nt * na1 + t] = ka;
        prm[nb + nt * (na1 + 1) + t] = 1.0 / ka;
        prm[nb + nt * (na1 + 2) + t] = v.asymp_jasa[t];
      }
      for (int a = threadIdx.x; a < NA16; a += NT) prm[nb + nt * (na1 + 3) + a] = a < Nn ? (double)v.type_nucl[a] : 0.0;
    }
    if (threadIdx.x == 0) {
      static_assert(EQ_WAVES == 8, "two waves per SIMD: waves w and w + 4");
      const bool mini = Nn - 16 * (NA - 1) <= 4, mini2 = Nn - 16 * (NA - 1) <= 2;
      const int NAF = mini ? NA - 1 : NA;
      const int n0 = NPR * NAF, n1 = n0 + NSG * NAF, n3 = n1 + (mini ? NPR + NSG : 0);
      const int nsteps = gn_jp((int)v.elec_num) >> 2;
      const int cap = (n3 + EQ_WAVES - 1) / EQ_WAVES;  // groups per wave, at most
      int* own = reinterpret_cast<int*>(lds + L.own);
      for (int rows = 0; rows < 2; ++rows) {
        int simd_load[4] = {0, 0, 0, 0}, wave_load[EQ_WAVES], wave_cnt[EQ_WAVES];
        for (int wv = 0; wv < EQ_WAVES; ++wv) wave_load[wv] = wave_cnt[wv] = 0;
        for (int it = 0; it < n3; ++it) {
          // matrix cycles and vector instructions of one step, and the fold
          int mc, vi, fold;
          if (it < n0) {
            mc = rows ? 320 : 640, vi = rows ? 25 : 14, fold = 1200;
          } else if (it < n1) {
            mc = rows ? 160 : 320, vi = rows ? 22 : 13, fold = 700;
          } else if (it < n1 + NPR) {
            mc = mini2 ? 80 : 160, vi = 14, fold = 1200;
          } else {
            mc = 80, vi = 13, fold = 700;
          }
          const int cost = nsteps * (mc + (9 * vi) / 2) + fold;
          int sb = -1;
          for (int sd = 0; sd < 4; ++sd) {
            const bool room = wave_cnt[sd] < cap || wave_cnt[sd + 4] < cap;
            if (room && (sb < 0 || simd_load[sd] < simd_load[sb])) sb = sd;
          }
          const int wa = sb, wb = sb + 4;
          int owner;
          if (wave_cnt[wa] >= cap)
            owner = wb;
          else if (wave_cnt[wb] >= cap)
            owner = wa;
          else if (wave_cnt[wa] != wave_cnt[wb])
            owner = wave_cnt[wa] < wave_cnt[wb] ? wa : wb;
          else
            owner = wave_load[wa] <= wave_load[wb] ? wa : wb;
          simd_load[sb] += cost;
          wave_load[owner] += cost;
          ++wave_cnt[owner];
          own[rows * n3 + it] = owner;
        }
      }
    }
  }
  int wcur = -1;
  for (int item = it0; item < it1; ++item) {
    // The sizes and the thread index pass through an empty asm statement, so that what
    // derives from them (LDS offsets, bounds, lane offsets) is recomputed per item instead
    // of being held in registers across the loop.
    // (readfirstlane: the view may come from a job table in memory)
    int Ne = __builtin_amdgcn_readfirstlane((int)v.elec_num), Nn = __builtin_amdgcn_readfirstlane((int)v.nucl_num),
        tidl = threadIdx.x;
    asm volatile("" : "+s"(Ne), "+s"(Nn), "+v"(tidl));
    const int tid = tidl, lane = tid & 63, li = lane & 15, lg = lane >> 4;
    const int JP = gn_jp(Ne), NA = gn_na(Nn), NA16 = 16 * NA;
    const EenResLds L(JP, NA16, C, DIM, een_prm_doubles(v, NA16));
    double* xt = lds + L.xt;
    double* jr = lds + L.jr;
    double* gch = lds + L.gch;
    double* nuc = lds + L.nuc;
    double* enr = lds + L.enr;
    double* r0t = lds + L.r0t;
    double* st = lds + L.st;
    double* cfl = lds + L.cfl;
    double* red = lds + L.red;
    const double* coord = v.coord;
    const int64_t cld = v.coord_ld;
    const bool mini = Nn - 16 * (NA - 1) <= 4;   // nuclei in the last tile
    const bool mini2 = Nn - 16 * (NA - 1) <= 2;  // pairs of units in one 4x4x4 product
    const int NAF = mini ? NA - 1 : NA;         // tiles on the 16x16x4 products
    const int n0 = NPR * NAF, n1 = n0 + NSG * NAF, n2 = n1 + (mini ? NPR : 0), n3 = n2 + (mini ? NSG : 0);
    const int w32 = item / ntiles;
    const int tile = item - w32 * ntiles;
    const int64_t w = w32;
    const int i0 = tile * EM;
    // 1. inputs.  Per walker (when it changes): x_j, then g_a(r_j) = e^(-kappa_a |r_j - R_a|)
    // for every nucleus and electron j (zero in the padding), then the k = 0 sums
    // st[a][q-1] = sum_j g_a(r_j)^q.  Per tile: x_i.
    if (w32 != wcur) {
      wcur = w32;
      for (int t = tid; t < 4 * JP; t += NT) {
        const int d = t / JP, jj = t - d * JP;  // d = 3: the pad slot
        jr[jj * JREC + 32 + d] = (d < 3 && jj < Ne) ? coord[w * Ne + jj + d * cld] : 0.0;
      }
      __builtin_amdgcn_s_waitcnt(0);  // the nucleus table and c-coefficients (first item)
      __syncthreads();
      for (int t = tid; t < NA16 * JP; t += NT) {  // gch[(tile * JP + j) * 16 + a mod 16]
        const int al = t & 15, jj = (t >> 4) % JP, a = 16 * ((t >> 4) / JP) + al;
        double g = 0.0;
        if (jj < Ne && a < Nn) {
          const double* pj = jr + jj * JREC + 32;
          const double dx = pj[0] - nuc[a], dy = pj[1] - nuc[NA16 + a], dz = pj[2] - nuc[2 * NA16 + a];
          g = exp_neg_deg<13, true>(-nuc[3 * NA16 + a] * sqrt(dx * dx + dy * dy + dz * dz));
        }
        gch[t] = g;
      }
      __syncthreads();
      for (int a = tid; a < NA16; a += NT) {
        double acc[K1];
#pragma unroll
        for (int q = 0; q < K1; ++q) acc[q] = 0.0;
        const double* ga = gch + (a >> 4) * JP * 16 + (a & 15);
        for (int jj = 0; jj < Ne; ++jj) {
          const double g = ga[16 * jj];
          double p = g;
#pragma unroll
          for (int q = 0; q < K1; ++q) {
            acc[q] += p;
            p *= g;
          }
        }
#pragma unroll
        for (int q = 0; q < K1; ++q) st[a * K1 + q] = acc[q];
      }
    }
    for (int t = tid; t < EM * 3; t += NT) {
      const int il = t / 3, d = t - 3 * il;
      xt[t] = (i0 + il < Ne) ? coord[w * Ne + i0 + il + d * cld] : 0.0;
    }
    __syncthreads();
    // 2. pairs of the tile: wave wv takes electrons i = i0 + 2 wv and i0 + 2 wv + 1, 32 lanes
    // each over the electrons j and the nuclei a.  They give the pair records g_e(r_ij),
    // 1/r_ij of the units, the e-n inverse distances of the fold, and the e-e and e-n
    // factors with their gradients and Laplacians and the q = 0 e-e-n sums of electron i,
    // summed over the 32 lanes by a fixed butterfly (after it, lane l32 holds entry
    // l32 * NVP / 32 + m of the NVP-entry vector).
    {
      // lane index and kappa through an empty asm statement: what derives from them lives
      // in this phase only
      int tp = threadIdx.x;
      double kee = uniform_f64(v.rescale_ee);
      asm volatile("" : "+v"(tp), "+s"(kee));
      const int il = 2 * wave + ((tp & 63) >> 5), l32 = tp & 31, i = i0 + il;
      const double kinv = 1.0 / kee;
      // the pair parameters from their LDS copy
      const double* prm = lds + L.prm;
      const int nb = (int)v.bord + 1, aord = (int)v.aord, nt = v.type_num > 0 ? (int)v.type_num : 0;
      const double xi = xt[il * 3], yi = xt[il * 3 + 1], zi = xt[il * 3 + 2];
      EeAcc se;
      EnAcc sn;
      double r0[K1][5];
#pragma unroll
      for (int k = 0; k < K1; ++k)
#pragma unroll
        for (int c = 0; c < 5; ++c) r0[k][c] = 0.0;
      // two electrons j per iteration, so that their independent chains interleave
      for (int j0 = l32; j0 < JP; j0 += 64) {
        const int j1 = j0 + 32;
        double ge0 = 0.0, ri0 = 0.0, ge1 = 0.0, ri1 = 0.0;
        if (i < Ne && j0 < Ne) {
          const double* pj = jr + j0 * JREC + 32;
          ee_pair<K1>(v, prm, i, j0, xi - pj[0], yi - pj[1], zi - pj[2], kee, kinv, se, r0, ge0, ri0);
        }
        if (i < Ne && j1 < Ne) {
          const double* pj = jr + j1 * JREC + 32;
          ee_pair<K1>(v, prm, i, j1, xi - pj[0], yi - pj[1], zi - pj[2], kee, kinv, se, r0, ge1, ri1);
        }
        jr[j0 * JREC + il] = ge0;
        jr[j0 * JREC + 16 + il] = ri0;
        if (j1 < JP) {
          jr[j1 * JREC + il] = ge1;
          jr[j1 * JREC + 16 + il] = ri1;
        }
      }
      for (int a = l32; a < NA16; a += 32) {
        double rinv = 0.0;
        if (i < Ne && a < Nn) {
          const double dx = xi - nuc[a], dy = yi - nuc[NA16 + a], dz = zi - nuc[2 * NA16 + a];
          double d;
          dist_and_inverse(dx * dx + dy * dy + dz * dz, d, rinv);
          const int t = (int)prm[nb + nt * (aord + 4) + a];
          en_pair_p(prm + nb + t * (aord + 1), aord, prm[nb + nt * (aord + 1) + t], prm[nb + nt * (aord + 2) + t],
                    prm[nb + nt * (aord + 3) + t], dx, dy, dz, d, rinv, gch[((a >> 4) * JP + i) * 16 + (a & 15)], sn);
        }
        enr[il * NA16 + a] = rinv;
      }
      // the tile's e-e and e-n values: sums over the wave, then over the waves at the end
      const double pe = wave_sum64(se.fee), pn = wave_sum64(sn.fen);
      if (lane == 0) {
        lds[L.pee + 2 * wave] = pe;
        lds[L.pee + 2 * wave + 1] = pn;
      }
      // per-electron vector: R0 (5 K1), e-e GL (4), e-n GL (4), zero padding to NVP
      constexpr int NV = 5 * K1 + 8, NVP = NV <= 32 ? 32 : 64;
      double vec[NVP];
#pragma unroll
      for (int k = 0; k < K1; ++k)
#pragma unroll
        for (int c = 0; c < 5; ++c) vec[5 * k + c] = r0[k][c];
      vec[5 * K1] = se.g0;
      vec[5 * K1 + 1] = se.g1;
      vec[5 * K1 + 2] = se.g2;
      vec[5 * K1 + 3] = se.g3;
      vec[5 * K1 + 4] = sn.h0;
      vec[5 * K1 + 5] = sn.h1;
      vec[5 * K1 + 6] = sn.h2;
      vec[5 * K1 + 7] = sn.h3;
#pragma unroll
      for (int n = NV; n < NVP; ++n) vec[n] = 0.0;
      een_butterfly<NVP, 16>(vec, l32);
#pragma unroll
      for (int m = 0; m < NVP / 32; ++m) {
        const int idx = l32 * (NVP / 32) + m;
        if (idx < 5 * K1) {
          r0t[il * K1 * 5 + idx] = vec[m];
        } else if (!VAL && idx < NV && i < Ne) {
          double* gl = (idx < 5 * K1 + 4 ? o.fee_gl : o.fen_gl) + w * 4 * Ne;
          gl[(int64_t)((idx - 5 * K1) & 3) * Ne + i] = vec[m];
        }
      }
    }
    __syncthreads();

    // 3. the groups of units this wave owns (owner table above)
    double ra[4][5];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int c = 0; c < 5; ++c) ra[r][c] = 0.0;
    const bool rows = Ne - i0 <= 8;  // a tile of at most 8 electrons
    const int* own = reinterpret_cast<const int*>(lds + L.own) + (rows ? n3 : 0);
    for (int it = 0; it < n3; ++it) {
      if (__builtin_amdgcn_readfirstlane(own[it]) != wave) continue;
      // the lane indices pass through an empty asm statement once per group, so that the
      // addresses derived from them live for one group, not across the unit loop
      int tg = threadIdx.x;
      asm volatile("" : "+v"(tg));
      const int gli = tg & 15, glg = (tg >> 4) & 3;
      const double xi[3] = {xt[gli * 3], xt[gli * 3 + 1], xt[gli * 3 + 2]};  // electron i of row gli
      EenResCtx x;
      x.jr = jr;
      x.xt = xt;
      x.r0t = r0t;
      x.enr = enr;
      x.cfl = cfl;
      x.st = st;
      x.kee = v.rescale_ee;
      x.nsteps = JP >> 2;
      x.li = gli;
      x.lg = glg;
      x.i0 = i0;
      x.Ne = Ne;
      x.NA16 = NA16;
      int at, g;
      if (it < n0) {
        at = it / (NPR > 0 ? NPR : 1), g = it - at * NPR;  // n0 = 0 when NPR = 0
      } else if (it < n1) {
        const int r = it - n0;
        at = r / NSG, g = NPR + r - at * NSG;
      } else if (it < n2) {
        at = NA - 1, g = it - n1;
      } else {
        at = NA - 1, g = NPR + it - n2;
      }
      const bool mu = it >= n1;  // the last nucleus tile on the 4x4x4 products
      const int mode = mu ? (mini2 ? EEN_MINI2 : EEN_MINI) : rows ? EEN_ROWS : EEN_FULL;
      x.gt = gch + at * JP * 16;
      x.a = 16 * at + (mu ? (gli & 3) : gli);
      x.nx = nuc[x.a];
      x.ny = nuc[NA16 + x.a];
      x.nz = nuc[2 * NA16 + x.a];
      x.ka = nuc[3 * NA16 + x.a];
      een_res_dispatch<C, NC>(g, mode, x, xi, ra);
    }
    // 4. sums over the nuclei (lanes li) and the waves
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const double y = row_sum16(ra[r][c]);
        if (li == 0) red[(wave * EM + lg + 4 * r) * 5 + c] = y;
      }
    __syncthreads();
    double* jrow = xt;  // free after the units
    for (int u = tid; u < EM * NC; u += NT) {
      const int t = VAL ? 5 * u : u;  // entry (il, c) of the reduction rows
      double y = 0.0;
#pragma unroll
      for (int wv = 0; wv < EQ_WAVES; ++wv) y += red[wv * EM * 5 + t];
      const int il = t / 5, c = t - 5 * il, i = i0 + il;
      if (c == 0)
        jrow[il] = (i < Ne) ? y : 0.0;
      else if (i < Ne)
        o.feen_gl[w * 4 * Ne + (int64_t)(c - 1) * Ne + i] = y;
    }
    __syncthreads();
    if (tid == 0) {
      double sj = 0.0, se = 0.0, sn = 0.0;
      for (int il = 0; il < EM; ++il) sj += jrow[il];
      for (int wv = 0; wv < EQ_WAVES; ++wv) {
        se += lds[L.pee + 2 * wv];
        sn += lds[L.pee + 2 * wv + 1];
      }
      Jt[w * ntiles + tile] = sj;
      JtP[2 * (w * ntiles + tile)] = se;
      JtP[2 * (w * ntiles + tile) + 1] = sn;
    }
    __syncthreads();  // xt, jr, enr, r0t, red and pee are rewritten by the next item
  }
}

template <int C, bool VAL>
__global__ void __launch_bounds__(64 * EQ_WAVES) k_jas_een_res(JasDevView v, JasOut o, double* __restrict__ Jt,
                                                              double* __restrict__ JtP, int ntiles) {
  jas_een_res_body<C, VAL>(v, o, Jt, JtP, ntiles, blockIdx.x, gridDim.x);
}

// value_only: the factors and e^J alone (no GL read or written)
__device__ __forceinline__ void jas_total_body(const JasDevView& v, const JasOut& o, const double* __restrict__ Jt,
                                               const double* __restrict__ JtP, int ntiles, int have_een, int fused,
                                               int value_only, const int64_t w) {
  __shared__ double sval;
  const int Ne = (int)v.elec_num;
  if (threadIdx.x == 0) {
    double een = 0.0;
    if (have_een)
      for (int t = 0; t < ntiles; ++t) een += Jt[w * ntiles + t];
    o.feen[w] = een;
    if (fused) {  // the e-e and e-n values of the resident e-e-n kernel, per tile
      double fee = 0.0, fen = 0.0;
      for (int t = 0; t < ntiles; ++t) {
        fee += JtP[2 * (w * ntiles + t)];
        fen += JtP[2 * (w * ntiles + t) + 1];
      }
      o.fee[w] = fee;
      o.fen[w] = fen;
    }
    const double J = o.fee[w] + o.fen[w] + een;
    const double val = (J < -100.0) ? 0.0 : exp(J);
    o.value[w] = val;
    sval = val;
  }
  if (value_only) return;
  __syncthreads();
  const double val = sval;
  const double* a = o.fee_gl + w * 4 * Ne;
  const double* b = o.fen_gl + w * 4 * Ne;
  double* cg = o.feen_gl + w * 4 * Ne;
  double* g = o.gl + w * 4 * Ne;
  for (int i = threadIdx.x; i < Ne; i += kBlock) {
    double s[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      if (!have_een) cg[c * Ne + i] = 0.0;
      s[c] = a[c * Ne + i] + b[c * Ne + i] + cg[c * Ne + i];
    }
    s[3] = s[3] + s[0] * s[0] + s[1] * s[1] + s[2] * s[2];
#pragma unroll
    for (int c = 0; c < 4; ++c) g[c * Ne + i] = s[c] * val;
  }
}

__global__ void __launch_bounds__(kBlock) k_jas_total(JasDevView v, JasOut o, const double* __restrict__ Jt,
                                                      const double* __restrict__ JtP, int ntiles, int have_een,
                                                      int fused, int value_only) {
  jas_total_body(v, o, Jt, JtP, ntiles, have_een, fused, value_only, blockIdx.x);
}

// ------------------------------------------------------------------------------------
// e-e-n on FP64 matrix cores, pass-based and balanced (cord 4 and 5, systems too large for
// the resident kernel): the products of k_jas_een_mfma, taken over two 16-nucleus tiles at
// a time.  The (k, q) units of a tile are grouped as in the resident kernel -- pairs (q, q + 1)
// of the same k, which share the A operand (ten matrix instructions per A), and single
// units -- and every group of the two tiles gets a wave of its own, dealt so that the waves
// of every SIMD (waves w, w + 4, w + 8) carry the same matrix load: at C = 5 12 waves, the 8
// pairs on waves 0-7 and the 4 singles on waves 8-11 (10 + 10 + 5 = 25 products per step on
// every SIMD, three waves each to hide one another's operand reads); at C = 4 8 waves, the 4
// pairs on waves 0-3 and the 4 singles on waves 4-7 (15 per SIMD).  One wave per (k, q), as
// in k_jas_een_mfma, puts 10 waves on 4 SIMDs, 3, 3, 2 and 2 of them.  The accumulators of a
// wave's group live across the passes over electrons j; the fold is that of k_jas_een_mfma.
// ------------------------------------------------------------------------------------
constexpr int bal_waves(int C) { return C == 5 ? 12 : 8; }

// the group of wave w of C: a pair (tile, index) on the first 2 npair waves, else a single
// (tile, index); tile -1: none
template <int C>
constexpr int bal_pair_tile(int w) {
  return w < 2 * een_npair(C) ? w / een_npair(C) : -1;
}
template <int C>
constexpr int bal_pair_idx(int w) {
  return w % een_npair(C);
}
template <int C>
constexpr int bal_single_tile(int w) {
  return w >= 2 * een_npair(C) ? (w - 2 * een_npair(C)) / een_nsingle(C) : -1;
}
template <int C>
constexpr int bal_single_idx(int w) {
  return (w - 2 * een_npair(C)) % een_nsingle(C);
}
static_assert(2 * een_npair(5) + 2 * een_nsingle(5) == bal_waves(5) &&
                  2 * een_npair(4) + 2 * een_nsingle(4) == bal_waves(4),
              "one group per wave");

// Dynamic LDS of k_jas_een_bal, in doubles, for passes of JS electrons j.
struct EenBalLds {
  int JSP, xt, xj, pdG, pdR, gch, ep, r0t, st, cfl, red, tot, jrow, total;
  __host__ __device__ EenBalLds(int JS, int C, int dim) : JSP(JS + 4) {
    int o = 0;
    xt = o;   o += EM * 3;                  // tile electrons
    xj = o;   o += 3 * JSP;                 // electrons j of the pass, [3][j]
    pdG = o;  o += JSP * EM;                // g_e(r_ij), [j][i]; 0 for j == i
    pdR = o;  o += JSP * EM;                // 1 / r_ij, [j][i]; 0 for j == i
    gch = o;  o += 2 * JSP * 16;            // g_a(r_j) of the two nucleus tiles, [tile][j][a]
    ep = o;   o += 2 * (C + 5) * EM * 16;   // per tile and (i, a): g^1..g^C, h, 2 kappa/R, kappa^2
    r0t = o;  o += EM * (C - 1) * 5;        // q = 0 sums of the tile electrons
    st = o;   o += 2 * 16 * (C - 1);        // per tile: sum over all j of g_a(r_j)^q
    cfl = o;  o += 2 * dim * 16;            // per tile: c-coefficients, [n][a]
    red = o;  o += bal_waves(C) * EM * 5;
    tot = o;  o += EM * 5;                  // running (J, gl) per row, summed over nuclei
    jrow = o; o += EM;
    total = o;
  }
};

// y^n for n = 1..4 (wave-uniform n) by selects: y, y^2, y^2 y, y^2 y^2 -- the products of the
// repeated multiplication of een_mfma_pass for n <= 3, (y y)(y y) for n = 4
__device__ __forceinline__ double een_pow4(double y, int n) {
  const double y2 = y * y;
  const double y3 = y2 * y, y4 = y2 * y2;
  return n == 1 ? y : n == 2 ? y2 : n == 3 ? y3 : y4;
}

// One pass of a wave's group over steps [st0, st1) of 4 electrons j (the first NC components:
// 5, or 1 for the values alone): units (K, Q) into a and, for a pair
// (two), (K, Q + 1) into b, on the nucleus tile g_tile.  (K, Q) are wave-uniform run-time
// values: one code path for all waves keeps the registers of one group.
template <int NC>
__device__ __forceinline__ void een_bal_pass(d4 (&a)[5], d4 (&b)[5], int two, int K, int Q, const double* g_tile,
                                             const double* pdG, const double* pdR, const double* xj,
                                             const double* xi, int JSP, int st0, int st1, int li, int lg) {
#pragma unroll 1
  for (int st = st0; st < st1; ++st) {
    const int jj = 4 * st + lg;
    const double Ge = pdG[jj * EM + li], ri = pdR[jj * EM + li], g = g_tile[jj * 16 + li];
    const double dx = xi[0] - xj[jj], dy = xi[1] - xj[JSP + jj], dz = xi[2] - xj[2 * JSP + jj];
    const double Gk = een_pow4(Ge, K), gq = een_pow4(g, Q);
    const double t = Gk * ri;
    const double A[5] = {Gk, t * dx, t * dy, t * dz, t};
#pragma unroll
    for (int c = 0; c < NC; ++c) a[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(A[c], gq, a[c], 0, 0, 0);
    if (two) {
      const double gq1 = gq * g;
#pragma unroll
      for (int c = 0; c < NC; ++c) b[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(A[c], gq1, b[c], 0, 0, 0);
    }
  }
}

// The fold of unit (K, Q) at row register r into R (the roles of een_fold; the q = 0 sums
// when Q = 1, the k = 0 sums when K = 1).
template <int C, int K, int Q>
__device__ __forceinline__ void een_bal_fold(EenRow& R, const d4 (&acc)[5], int r, const double* cf, const double* gp,
                                             const double* r0row, const double* srow, double kee) {
  double X[5], XR[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, X0[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    const d4 x = acc[c];
    X[c] = r == 0 ? x[0] : r == 1 ? x[1] : r == 2 ? x[2] : x[3];
  }
  if (Q == 1)
#pragma unroll
    for (int c = 0; c < 5; ++c) XR[c] = r0row[(K - 1) * 5 + c];
  // the subtrahend is g^Q rounded (fma with 0), so that S - g^Q is exactly 0 when electron
  // i is the only one
  if (K == 1) X0[0] = srow[Q - 1] - __builtin_fma(gp[Q - 1], gp[1], 0.0);
  een_fold<C, K, Q>(R, cf, 16, gp, X, XR, X0, kee);
}

// (J, gradient, Laplacian) of a row from its folded sums and the (i, a) factors e (stride es)
__device__ __forceinline__ void een_bal_out(const EenRow& R, const double* e, int es, double out[5]) {
  const double h0 = e[0], h1 = e[es], h2 = e[2 * es], kr = e[3 * es], ka2 = e[4 * es];
  out[0] += R.sJ;
  out[1] += h0 * R.s1 - R.t0[0];
  out[2] += h1 * R.s1 - R.t0[1];
  out[3] += h2 * R.s1 - R.t0[2];
  out[4] += ka2 * R.s2 - kr * R.s1 + R.s3 - 2.0 * (h0 * R.t1[0] + h1 * R.t1[1] + h2 * R.t1[2]) - 2.0 * R.sL;
}

// The groups of wave W at compile time (for its folds).
template <int C, int W>
struct EenBalWave {
  static constexpr int PT = bal_pair_tile<C>(W), ST = bal_single_tile<C>(W);
  static constexpr int PK = PT >= 0 ? een_pair_k(C, bal_pair_idx<C>(W)) : 1;
  static constexpr int PQ = PT >= 0 ? een_pair_q(C, bal_pair_idx<C>(W)) : 1;
  static constexpr int SK = ST >= 0 ? een_single_k(C, bal_single_idx<C>(W)) : 1;
  static constexpr int SQ = ST >= 0 ? een_single_q(C, bal_single_idx<C>(W)) : 1;
};

// The folds of wave W's groups at row register r (row il = lg + 4 r) into out[5].
template <int C, int W = 0>
__device__ __forceinline__ void een_bal_dispatch_fold(int wave, const d4 (&pa)[5], const d4 (&pb)[5], int r, int il,
                                                      int li, int a0, int Nn, int ntile, const double* ep,
                                                      const double* cfl, const double* r0t, const double* st,
                                                      double kee, double out[5]) {
  if constexpr (W < bal_waves(C)) {
    if (wave == W) {
      typedef EenBalWave<C, W> B;
      constexpr int K1 = C - 1, DIM = Lkm<C>::dim, EPS = EM * 16;
      auto group = [&](int tile, auto fold) {
        if (tile >= ntile || a0 + 16 * tile + li >= Nn) return;
        const double* e = ep + tile * (C + 5) * EPS + il * 16 + li;
        double gp[C + 1];
        gp[0] = 1.0;
#pragma unroll
        for (int p = 1; p <= C; ++p) gp[p] = e[(p - 1) * EPS];
        EenRow R;
        fold(R, gp, cfl + tile * DIM * 16 + li, st + tile * 16 * K1 + li * K1);
        een_bal_out(R, e + C * EPS, EPS, out);
      };
      if constexpr (B::PT >= 0)
        group(B::PT, [&](EenRow& R, const double* gp, const double* cf, const double* srow) {
          een_bal_fold<C, B::PK, B::PQ>(R, pa, r, cf, gp, r0t + il * K1 * 5, srow, kee);
          een_bal_fold<C, B::PK, B::PQ + 1>(R, pb, r, cf, gp, r0t + il * K1 * 5, srow, kee);
        });
      if constexpr (B::ST >= 0)
        group(B::ST, [&](EenRow& R, const double* gp, const double* cf, const double* srow) {
          een_bal_fold<C, B::SK, B::SQ>(R, pa, r, cf, gp, r0t + il * K1 * 5, srow, kee);
        });
      return;
    }
    een_bal_dispatch_fold<C, W + 1>(wave, pa, pb, r, il, li, a0, Nn, ntile, ep, cfl, r0t, st, kee, out);
  }
}

// VAL: the values alone (the value component of the products only, J folded alone, no GL
// written); the value products stay zero in the other components, so J is the full
// evaluation's bit for bit.
template <int C, bool VAL>
__device__ __forceinline__ void jas_een_bal_body(const JasDevView& v, const JasOut& o, const double* __restrict__ R0,
                                                 const double* __restrict__ gN, const double* __restrict__ Sg,
                                                 double* __restrict__ Jt, int ntiles, int JS, const int64_t b) {
  static_assert(C == 4 || C == 5, "balanced pass-based kernel: cord 4 and 5");
  constexpr int K1 = C - 1;
  constexpr int NC = VAL ? 1 : 5;
  constexpr int NWV = bal_waves(C);
  constexpr int NT = 64 * NWV;
  constexpr int DIM = Lkm<C>::dim;
  constexpr int EPS = EM * 16;
  extern __shared__ double lds[];
  const EenBalLds L(JS, C, DIM);
  double* xt = lds + L.xt;
  double* xj = lds + L.xj;
  double* pdG = lds + L.pdG;
  double* pdR = lds + L.pdR;
  double* gch = lds + L.gch;
  double* ep = lds + L.ep;
  double* r0t = lds + L.r0t;
  double* st = lds + L.st;
  double* cfl = lds + L.cfl;
  double* red = lds + L.red;
  double* tot = lds + L.tot;
  double* jrow = lds + L.jrow;
  const int Ne = (int)v.elec_num, Nn = (int)v.nucl_num;
  int64_t w;
  int tile;
  een_block_work(b, v.walk_num, ntiles, w, tile);
  const int i0 = tile * EM;
  const int tid = threadIdx.x, lane = tid & 63, li = lane & 15, lg = lane >> 4;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const double kee = v.rescale_ee;
  const double* coord = v.coord;
  const int64_t cld = v.coord_ld;
  const int npass = (Ne + JS - 1) / JS;
  for (int t = tid; t < EM * 3; t += NT) {
    const int il = t / 3, d = t - 3 * il;
    xt[t] = (i0 + il < Ne) ? coord[w * Ne + i0 + il + d * cld] : 0.0;
  }
  for (int t = tid; t < EM * K1 * 5; t += NT) {
    const int il = t / (K1 * 5);
    r0t[t] = (i0 + il < Ne) ? R0[(w * Ne + i0) * K1 * 5 + t] : 0.0;
  }
  for (int t = tid; t < EM * 5; t += NT) tot[t] = 0.0;
  __syncthreads();
  const double xi[3] = {xt[li * 3], xt[li * 3 + 1], xt[li * 3 + 2]};  // electron i of row li
  // this wave's group: nucleus tile, (k, q), pair or single
  int g_t = 0, g_k = 1, g_q = 1, g_two = 0, g_idx = 0;
  for (int wv = 0; wv < NWV; ++wv)
    if (wv == wave) {
      g_two = bal_pair_tile<C>(wv) >= 0;
      if (g_two) {
        g_t = bal_pair_tile<C>(wv);
        g_idx = bal_pair_idx<C>(wv);
        g_k = een_pair_k(C, g_idx);
        g_q = een_pair_q(C, g_idx);
      } else {
        g_t = bal_single_tile<C>(wv);
        g_idx = bal_single_idx<C>(wv);
        g_k = een_single_k(C, g_idx);
        g_q = een_single_q(C, g_idx);
      }
    }
  // With an odd number of nucleus tiles the last iteration holds one tile: the waves of the
  // second tile's groups then take the second half of the electron steps of the same group
  // on the first tile (its partner wave the first half), and hand their partial sums to the
  // partner through LDS (the pair records' space, free by then), when that space holds a
  // pair group's sums for every pair (40 doubles per lane each).
  const bool split_ok = L.JSP >= 40 * een_npair(C);

  for (int a0 = 0; a0 < Nn; a0 += 32) {
    const int ntile = a0 + 16 < Nn ? 2 : 1;
    // per-tile constants of the two tiles: c-coefficients, the k = 0 sums, and per (i, a)
    // the powers of g_a(r_i) and the factors of their derivatives
    for (int t = tid; t < 2 * DIM * 16; t += NT) {
      const int tt = t / (DIM * 16), n = (t >> 4) % DIM, al = t & 15, a = a0 + 16 * tt + al;
      cfl[t] = (a < Nn) ? v.cfull[(int64_t)n * Nn + a] : 0.0;
    }
    for (int t = tid; t < 2 * 16 * K1; t += NT) {
      const int a = a0 + t / K1;
      st[t] = (a < Nn) ? Sg[(w * Nn + a0) * K1 + t] : 0.0;
    }
    for (int t = tid; t < 2 * EPS; t += NT) {
      const int tt = t / EPS, u = t - tt * EPS, il = u >> 4, al = u & 15, i = i0 + il, a = a0 + 16 * tt + al;
      double gp[C + 1], h[3] = {0.0, 0.0, 0.0}, kr = 0.0, ka2 = 0.0;
#pragma unroll
      for (int p = 0; p <= C; ++p) gp[p] = 0.0;
      if (i < Ne && a < Nn) {
        const double ka = v.kappa_en[v.type_nucl[a]];
        const double dx = xt[il * 3] - v.nucl_coord[a], dy = xt[il * 3 + 1] - v.nucl_coord[a + Nn],
                     dz = xt[il * 3 + 2] - v.nucl_coord[a + 2 * Nn];
        const double R = sqrt(dx * dx + dy * dy + dz * dz);
        const double rinv = 1.0 / R;
        const double g1 = gN[gn_index(w, a, i, gn_na(Nn), gn_jp(Ne))];  // exp(-kappa_a R)
        gp[1] = g1;
#pragma unroll
        for (int p = 2; p <= C; ++p) gp[p] = gp[p - 1] * g1;
        h[0] = -ka * (dx * rinv);
        h[1] = -ka * (dy * rinv);
        h[2] = -ka * (dz * rinv);
        kr = 2.0 * ka * rinv;
        ka2 = ka * ka;
      }
      double* e = ep + tt * (C + 5) * EPS + u;
#pragma unroll
      for (int p = 1; p <= C; ++p) e[(p - 1) * EPS] = gp[p];
      e[(C + 0) * EPS] = h[0];
      e[(C + 1) * EPS] = h[1];
      e[(C + 2) * EPS] = h[2];
      e[(C + 3) * EPS] = kr;
      e[(C + 4) * EPS] = ka2;
    }

    d4 pa[5], pb[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) pa[c] = pb[c] = d4{0.0, 0.0, 0.0, 0.0};

    for (int pass = 0; pass < npass; ++pass) {
      const int j0 = pass * JS;
      const int jn = min(JS, Ne - j0);
      const int jn4 = (jn + 3) & ~3;
      if (npass > 1 || a0 == 0) {
        if (npass > 1) __syncthreads();  // previous pass done with the pair data
        for (int t = tid; t < 3 * jn4; t += NT) {
          const int d = t / jn4, jj = t - d * jn4;
          xj[d * L.JSP + jj] = (jj < jn) ? coord[w * Ne + j0 + jj + d * cld] : 0.0;
        }
        for (int t = tid; t < EM * jn4; t += NT) {
          const int il = t & (EM - 1), jj = t / EM;
          const int i = i0 + il, j = j0 + jj;
          double Ge = 0.0, ri = 0.0;
          if (i < Ne && jj < jn && i != j) {
            const double dx = xt[il * 3] - coord[w * Ne + j];
            const double dy = xt[il * 3 + 1] - coord[w * Ne + j + cld];
            const double dz = xt[il * 3 + 2] - coord[w * Ne + j + 2 * cld];
            double d;
            dist_and_inverse(dx * dx + dy * dy + dz * dz, d, ri);
            Ge = exp_neg_deg<13, true>(-kee * d);
          }
          pdG[t] = Ge;  // [jj][il]
          pdR[t] = ri;
        }
      }
      for (int t = tid; t < 2 * 16 * jn4; t += NT) {  // consecutive threads: consecutive doubles
        const int tt = t / (16 * jn4), u = t - tt * 16 * jn4, jj = u >> 4, al = u & 15;
        const int a = a0 + 16 * tt + al;
        gch[tt * L.JSP * 16 + jj * 16 + al] =
            (a < Nn && jj < jn) ? gN[gn_index(w, a, j0 + jj, gn_na(Nn), gn_jp(Ne))] : 0.0;
      }
      __syncthreads();
      const bool split = ntile == 1 && split_ok, helper = split && g_t == 1;
      if (g_t < ntile || helper) {
        const int ns = jn4 >> 2, h = split ? ns / 2 : ns;
        een_bal_pass<NC>(pa, pb, g_two, g_k, g_q, gch + (helper ? 0 : g_t) * L.JSP * 16, pdG, pdR, xj, xi, L.JSP,
                         helper ? h : 0, helper ? ns : h, li, lg);
      }
      if (pass + 1 < npass) __syncthreads();  // before gch of the next pass is written
    }
    if (ntile == 1 && split_ok) {
      // partial sums of the helpers (g_t == 1) into their partners (g_t == 0, same group):
      // the pairs' sums first, then the singles', slot g_idx of 40 (20) doubles per lane
      double* sbuf = lds + L.pdG;
      for (int round = 0; round < 2; ++round) {
        const bool mine = (round == 0) == (g_two != 0);
        const int nc = g_two ? 10 : 5;
        double* slot = sbuf + (size_t)g_idx * nc * 4 * 64 + lane;
        __syncthreads();  // every product done; the previous round's slots read
        if (mine && g_t == 1) {
#pragma unroll
          for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              slot[(c * 4 + r) * 64] = pa[c][r];
              if (g_two) slot[((5 + c) * 4 + r) * 64] = pb[c][r];
            }
        }
        __syncthreads();
        if (mine && g_t == 0) {
#pragma unroll
          for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              pa[c][r] += slot[(c * 4 + r) * 64];
              if (g_two) pb[c][r] += slot[((5 + c) * 4 + r) * 64];
            }
        }
      }
    }

    // Epilogue: lane (li, lg) holds rows lg + 4 r (register r), column a = a0 + 16 tile + li.
#pragma unroll 1
    for (int r = 0; r < 4; ++r) {
      const int il = lg + 4 * r, i = i0 + il;
      double out[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      if (i < Ne) een_bal_dispatch_fold<C>(wave, pa, pb, r, il, li, a0, Nn, ntile, ep, cfl, r0t, st, kee, out);
      // sum over the 16 nuclei of the tile(s): lanes li = 0..15 of this row group
#pragma unroll
      for (int c = 0; c < NC; ++c) out[c] = row_sum16(out[c]);
      if (li == 0)
#pragma unroll
        for (int c = 0; c < NC; ++c) red[(wave * EM + il) * 5 + c] = out[c];
    }
    __syncthreads();
    for (int u = tid; u < EM * NC; u += NT) {
      const int t = VAL ? 5 * u : u;  // entry (il, c) of the row sums
      double x = tot[t];
#pragma unroll
      for (int wv = 0; wv < NWV; ++wv) x += red[wv * EM * 5 + t];
      tot[t] = x;
    }
    // the next tile pair rewrites cfl, st, ep and gch only after this barrier
    __syncthreads();
  }
  for (int u = tid; u < EM * NC; u += NT) {
    const int t = VAL ? 5 * u : u;
    const int il = t / 5, c = t - 5 * il, i = i0 + il;
    if (c == 0)
      jrow[il] = (i < Ne) ? tot[t] : 0.0;
    else if (!VAL && i < Ne)
      o.feen_gl[w * 4 * Ne + (int64_t)(c - 1) * Ne + i] = tot[t];
  }
  __syncthreads();
  if (tid == 0) {
    double sj = 0.0;
    for (int il = 0; il < EM; ++il) sj += jrow[il];
    Jt[w * ntiles + tile] = sj;
  }
}

template <int C, bool VAL>
__global__ void __launch_bounds__(64 * bal_waves(C)) k_jas_een_bal(JasDevView v, JasOut o, const double* __restrict__ R0,
                                                              const double* __restrict__ gN,
                                                              const double* __restrict__ Sg, double* __restrict__ Jt,
                                                              int ntiles, int JS) {
  jas_een_bal_body<C, VAL>(v, o, R0, gN, Sg, Jt, ntiles, JS, blockIdx.x);
}

// ------------------------------------------------------------------------------------
// Grouped launches over the Jastrow evaluations of several contexts (a batch of molecules):
// one launch per kernel (and cord), in which workgroup b serves job m, the one whose block
// range [starts[m], starts[m + 1]) holds b, as that job's workgroup b - starts[m] of
// starts[m + 1] - starts[m].  The job's view and outputs are read from the job table in
// place (wave-uniform loads; a local copy of the view would live in scratch memory).
// ------------------------------------------------------------------------------------
__device__ __forceinline__ int job_of(const int* __restrict__ starts, int njobs, int b) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (starts[mid] <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  return __builtin_amdgcn_readfirstlane(lo);
}

template <int C>
__global__ void __launch_bounds__(pair_block(C)) k_jas_pair_batch(const JasJob* __restrict__ jobs,
                                                           const int* __restrict__ starts, int njobs) {
  const int m = job_of(starts, njobs, blockIdx.x);
  const JasJob& J = jobs[m];
  jas_pair_body<C>(J.v, J.o, J.R0, J.gN, J.Sg, (int64_t)blockIdx.x - starts[m]);
}

template <int C>
__global__ void __launch_bounds__(kBlock) k_jas_een_batch(const JasJob* __restrict__ jobs,
                                                          const int* __restrict__ starts, int njobs) {
  const int m = job_of(starts, njobs, blockIdx.x);
  const JasJob& J = jobs[m];
  jas_een_body<C>(J.v, J.o, J.R0, J.gN, J.Jt, J.TI, J.JC, J.ntiles, (int64_t)blockIdx.x - starts[m]);
}

template <int C>
__global__ void __launch_bounds__(64 * EenWaves<C>::n) k_jas_een_mfma_batch(const JasJob* __restrict__ jobs,
                                                                           const int* __restrict__ starts,
                                                                           int njobs) {
  const int m = job_of(starts, njobs, blockIdx.x);
  const JasJob& J = jobs[m];
  jas_een_mfma_body<C>(J.v, J.o, J.R0, J.gN, J.Sg, J.Jt, J.ntiles, J.JC, (int64_t)blockIdx.x - starts[m]);
}

template <int C>
__global__ void __launch_bounds__(64 * bal_waves(C)) k_jas_een_bal_batch(const JasJob* __restrict__ jobs,
                                                                    const int* __restrict__ starts, int njobs) {
  const int m = job_of(starts, njobs, blockIdx.x);
  const JasJob& J = jobs[m];
  jas_een_bal_body<C, false>(J.v, J.o, J.R0, J.gN, J.Sg, J.Jt, J.ntiles, J.JC, (int64_t)blockIdx.x - starts[m]);
}

// The persistent resident kernel over several contexts: each context gets its own share of
// the workgroups (one per CU), in proportion to its work, and runs its (walker, tile) items
// on them as a launch of its own would.
template <int C>
__global__ void __launch_bounds__(64 * EQ_WAVES) k_jas_een_res_batch(const JasJob* __restrict__ jobs,
                                                                    const int* __restrict__ starts, int njobs) {
  const int m = job_of(starts, njobs, blockIdx.x);
  const JasJob& J = jobs[m];
  jas_een_res_body<C, false>(J.v, J.o, J.Jt, J.JtP, J.ntiles, (int)blockIdx.x - starts[m], starts[m + 1] - starts[m]);
}

__global__ void __launch_bounds__(kBlock) k_jas_total_batch(const JasJob* __restrict__ jobs,
                                                            const int* __restrict__ starts, int njobs) {
  const int m = job_of(starts, njobs, blockIdx.x);
  const JasJob& J = jobs[m];
  jas_total_body(J.v, J.o, J.Jt, J.JtP, J.ntiles, J.have_een, J.fused, 0, (int64_t)blockIdx.x - starts[m]);
}

struct EenShape {
  bool mfma;  // matrix-core kernel (cord <= EMFMA_MAX_C)
  bool res;   // its resident variant (everything of a walker tile in LDS)
  bool bal;   // else, at cord 4 and 5, the balanced pass-based variant
  int TI, JC, ntiles;
  size_t lds;
};

EenShape een_shape(const JasDevView& v) {
  EenShape s;
  const int Ne = (int)v.elec_num, Nn = (int)v.nucl_num;
  s.mfma = v.cord >= 2 && v.cord <= EMFMA_MAX_C;
  if (s.mfma) {
    // electrons j per LDS pass: all of them when they fit (up to about 310 electrons at
    // cord 5), else the fewest passes that fit, of equal length (the pair records of the
    // tile are rebuilt for every pass and nucleus tile)
    const int C = (int)v.cord;
    s.TI = EM;
    const int ne4 = (Ne + 3) & ~3;
    const size_t res_lds0 =
        sizeof(double) * (size_t)EenResLds(gn_jp(Ne), 16 * gn_na(Nn), C, (int)v.dim_c, een_prm_doubles(v, 16 * gn_na(Nn))).total;
    // the balanced variant where k_jas_een_mfma needs more than one pass over j as well (it
    // holds fewer electrons per pass: with one pass, about 230-300 electrons at C = 5, the
    // one-wave-per-unit kernel builds the tile's pair records once and is as fast or faster)
    const bool mfma_one_pass =
        sizeof(double) * (size_t)EenMfmaLds(ne4, C, (int)v.dim_c, C * (C - 1) / 2).total <= 160 * 1024;
    s.bal = res_lds0 > 160 * 1024 && (C == 4 || C == 5) && !mfma_one_pass;
    auto fits = [&](int js) {
      const size_t n = s.bal ? (size_t)EenBalLds(js, C, (int)v.dim_c).total
                             : (size_t)EenMfmaLds(js, C, (int)v.dim_c, C * (C - 1) / 2).total;
      return sizeof(double) * n <= 160 * 1024;
    };
    int npass = 1, js = ne4;
    while (!fits(js) && js > 4) {
      ++npass;
      js = (((ne4 + npass - 1) / npass) + 3) & ~3;
    }
    s.JC = js;
    s.ntiles = (Ne + EM - 1) / EM;
    s.res = res_lds0 <= 160 * 1024;
    s.lds = s.res   ? res_lds0
            : s.bal ? sizeof(double) * (size_t)EenBalLds(s.JC, C, (int)v.dim_c).total
                    : sizeof(double) * (size_t)EenMfmaLds(s.JC, C, (int)v.dim_c, C * (C - 1) / 2).total;
    return s;
  }
  s.res = false;
  s.bal = false;
  s.TI = kBlock / Nn;
  if (s.TI < 1) s.TI = 1;
  if (s.TI > Ne) s.TI = Ne;
  s.JC = 4096 / Nn;
  if (s.JC > 64) s.JC = 64;
  if (s.JC < 4) s.JC = 4;
  s.ntiles = (Ne + s.TI - 1) / s.TI;
  const int npairs = s.TI * Nn;
  const int nred = npairs > kBlock ? npairs : kBlock;
  s.lds = sizeof(double) * (3 * s.TI + 6 * s.TI * s.JC + (size_t)Nn * s.JC + 5 * (size_t)nred);
  return s;
}

// ------------------------------------------------------------------------------------
// Intermediates with the reference layouts (materialised for the getters only).
// ------------------------------------------------------------------------------------
__device__ __forceinline__ void ee_vec(const JasDevView& v, int64_t w, int i, int j, double& dx, double& dy,
                                       double& dz, double& d) {
  const int Ne = (int)v.elec_num;
  const double* c = v.coord;
  const int64_t L = v.coord_ld;
  dx = c[w * Ne + i] - c[w * Ne + j];
  dy = c[w * Ne + i + L] - c[w * Ne + j + L];
  dz = c[w * Ne + i + 2 * L] - c[w * Ne + j + 2 * L];
  d = sqrt(dx * dx + dy * dy + dz * dz);
}

__device__ __forceinline__ void en_vec(const JasDevView& v, int64_t w, int i, int a, double& dx, double& dy,
                                       double& dz, double& d) {
  const int Ne = (int)v.elec_num, Nn = (int)v.nucl_num;
  const double* c = v.coord;
  const int64_t L = v.coord_ld;
  dx = c[w * Ne + i] - v.nucl_coord[a];
  dy = c[w * Ne + i + L] - v.nucl_coord[a + Nn];
  dz = c[w * Ne + i + 2 * L] - v.nucl_coord[a + 2 * Nn];
  d = sqrt(dx * dx + dy * dy + dz * dz);
}

// een_rescaled_e_gl element: d/d r_i (component c) of exp(-kappa r_ij)^k (diagonal: 0)
__device__ __forceinline__ double een_e_gl(const JasDevView& v, int64_t w, int k, int i, int j, int c) {
  if (i == j || k == 0) return 0.0;
  double dx, dy, dz, d;
  ee_vec(v, w, i, j, dx, dy, dz, d);
  const double kl = -(double)k * v.rescale_ee;
  const double rinv = 1.0 / d;
  double g = 1.0;
  const double e = exp(-v.rescale_ee * d);
  for (int p = 0; p < k; ++p) g *= e;
  if (c == 0) return kl * (dx * rinv) * g;
  if (c == 1) return kl * (dy * rinv) * g;
  if (c == 2) return kl * (dz * rinv) * g;
  return kl * (2.0 * rinv + kl) * g;
}

__device__ __forceinline__ double een_e_val(const JasDevView& v, int64_t w, int k, int i, int j) {
  if (i == j) return 0.0;
  double dx, dy, dz, d;
  ee_vec(v, w, i, j, dx, dy, dz, d);
  const double e = exp(-v.rescale_ee * d);
  double g = 1.0;
  for (int p = 0; p < k; ++p) g *= e;
  return g;
}

__device__ __forceinline__ double een_n_val(const JasDevView& v, int64_t w, int q, int a, int i) {
  double dx, dy, dz, d;
  en_vec(v, w, i, a, dx, dy, dz, d);
  const double e = exp(-v.kappa_en[v.type_nucl[a]] * d);
  double g = 1.0;
  for (int p = 0; p < q; ++p) g *= e;
  return g;
}

__global__ void k_jas_inter(JasDevView v, int kind, double* __restrict__ out, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int Ne = (int)v.elec_num, Nn = (int)v.nucl_num, C = (int)v.cord;
  int64_t r = t;
  double res = 0.0;
  switch (kind) {
    case J_EE_RESCALED: {  // [W][Ne(j)][Ne(i)]
      const int i = (int)(r % Ne); r /= Ne;
      const int j = (int)(r % Ne); r /= Ne;
      double dx, dy, dz, d;
      ee_vec(v, r, i, j, dx, dy, dz, d);
      res = (1.0 - exp(-v.rescale_ee * d)) * (1.0 / v.rescale_ee);
      break;
    }
    case J_EE_RESCALED_GL: {  // [W][Ne(j)][Ne(i)][4]
      const int c = (int)(r % 4); r /= 4;
      const int i = (int)(r % Ne); r /= Ne;
      const int j = (int)(r % Ne); r /= Ne;
      double dx, dy, dz, d;
      ee_vec(v, r, i, j, dx, dy, dz, d);
      const double dd = d > 1.e-20 ? d : 1.e-20;
      const double dinv = 1.0 / dd, e = exp(-v.rescale_ee * dd);
      res = c == 0 ? dx * dinv * e : c == 1 ? dy * dinv * e : c == 2 ? dz * dinv * e : (2.0 * dinv - v.rescale_ee) * e;
      break;
    }
    case J_EN_RESCALED: {  // [W][Nn][Ne]
      const int i = (int)(r % Ne); r /= Ne;
      const int a = (int)(r % Nn); r /= Nn;
      double dx, dy, dz, d;
      en_vec(v, r, i, a, dx, dy, dz, d);
      const double ka = v.kappa_en[v.type_nucl[a]];
      res = (1.0 - exp(-ka * d)) * (1.0 / ka);
      break;
    }
    case J_EN_RESCALED_GL: {  // [W][Nn][Ne][4]
      const int c = (int)(r % 4); r /= 4;
      const int i = (int)(r % Ne); r /= Ne;
      const int a = (int)(r % Nn); r /= Nn;
      double dx, dy, dz, d;
      en_vec(v, r, i, a, dx, dy, dz, d);
      const double ka = v.kappa_en[v.type_nucl[a]];
      const double dd = d > 1.e-20 ? d : 1.e-20;
      const double dinv = 1.0 / dd, e = exp(-ka * dd);
      res = c == 0 ? dx * dinv * e : c == 1 ? dy * dinv * e : c == 2 ? dz * dinv * e : (2.0 * dinv - ka) * e;
      break;
    }
    case J_EEN_E: {  // [W][C+1][Ne(j)][Ne(i)]
      const int i = (int)(r % Ne); r /= Ne;
      const int j = (int)(r % Ne); r /= Ne;
      const int k = (int)(r % (C + 1)); r /= (C + 1);
      res = een_e_val(v, r, k, i, j);
      break;
    }
    case J_EEN_E_GL: {  // [W][C+1][Ne(j)][4][Ne(i)]
      const int i = (int)(r % Ne); r /= Ne;
      const int c = (int)(r % 4); r /= 4;
      const int j = (int)(r % Ne); r /= Ne;
      const int k = (int)(r % (C + 1)); r /= (C + 1);
      res = een_e_gl(v, r, k, i, j, c);
      break;
    }
    case J_EEN_N: {  // [W][C+1][Nn][Ne]
      const int i = (int)(r % Ne); r /= Ne;
      const int a = (int)(r % Nn); r /= Nn;
      const int q = (int)(r % (C + 1)); r /= (C + 1);
      res = een_n_val(v, r, q, a, i);
      break;
    }
    case J_EEN_N_GL: {  // [W][C+1][Nn][4][Ne]
      const int i = (int)(r % Ne); r /= Ne;
      const int c = (int)(r % 4); r /= 4;
      const int a = (int)(r % Nn); r /= Nn;
      const int q = (int)(r % (C + 1)); r /= (C + 1);
      double dx, dy, dz, d;
      en_vec(v, r, i, a, dx, dy, dz, d);
      const double kl = -(double)q * v.kappa_en[v.type_nucl[a]];
      const double rinv = 1.0 / d;
      const double d0 = kl * (dx * rinv), d1 = kl * (dy * rinv), d2 = kl * (dz * rinv);
      const double d3 = kl * 2.0 * rinv + d0 * d0 + d1 * d1 + d2 * d2;
      const double g = een_n_val(v, r, q, a, i);
      res = (c == 0 ? d0 : c == 1 ? d1 : c == 2 ? d2 : d3) * g;
      break;
    }
    case J_TMP_C: {  // [W][C][C+1][Nn][Ne]
      const int i = (int)(r % Ne); r /= Ne;
      const int a = (int)(r % Nn); r /= Nn;
      const int q = (int)(r % (C + 1)); r /= (C + 1);
      const int k = (int)(r % C); r /= C;
      if (k + q <= C)
        for (int m = 0; m < Ne; ++m) res += een_e_val(v, r, k, i, m) * een_n_val(v, r, q, a, m);
      break;
    }
    case J_DTMP_C: {  // [W][C][C+1][Nn][4][Ne]
      const int i = (int)(r % Ne); r /= Ne;
      const int c = (int)(r % 4); r /= 4;
      const int a = (int)(r % Nn); r /= Nn;
      const int q = (int)(r % (C + 1)); r /= (C + 1);
      const int k = (int)(r % C); r /= C;
      if (k + q <= C)
        for (int m = 0; m < Ne; ++m) res += een_e_gl(v, r, k, i, m, c) * een_n_val(v, r, q, a, m);
      break;
    }
    default:
      break;
  }
  out[t] = res;
}

__global__ void k_ee_distance(const double* __restrict__ coord, int64_t W, int64_t Ne, double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= W * Ne * Ne) return;
  const int64_t i = t % Ne, j = (t / Ne) % Ne, w = t / (Ne * Ne);
  const int64_t L = W * Ne;
  const double dx = coord[w * Ne + i] - coord[w * Ne + j];
  const double dy = coord[w * Ne + i + L] - coord[w * Ne + j + L];
  const double dz = coord[w * Ne + i + 2 * L] - coord[w * Ne + j + 2 * L];
  out[t] = sqrt(dx * dx + dy * dy + dz * dz);
}

__global__ void k_en_distance(const double* __restrict__ coord, int64_t P, int64_t ld, const double* __restrict__ nc,
                              int64_t Nn, double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P * Nn) return;
  const int64_t a = t % Nn, p = t / Nn;
  const double dx = coord[p] - nc[a], dy = coord[p + ld] - nc[a + Nn], dz = coord[p + 2 * ld] - nc[a + 2 * Nn];
  out[t] = sqrt(dx * dx + dy * dy + dz * dz);
}

}  // namespace

// Scratch of the Jastrow kernels: gN, R0, Jt and Sg, each starting on a 256-byte boundary
// (the e-e-n kernel copies gN into LDS 16 bytes per lane).

// Regions of the Jastrow scratch, in doubles, each 256-byte aligned.  The pair kernel's
// outputs (gN, R0, Sg) exist only when it runs; the resident e-e-n kernel, which does its
// work, writes only the per-tile sums Jt and JtP.  Every kernel writes each scratch element
// before any kernel reads it, so the scratch needs no clearing.
struct ScratchLayout {
  int64_t gN, R0, Jt, Sg, JtP, total;
  explicit ScratchLayout(const JasDevView& v) {
    auto up = [](int64_t n) { return (n + 31) & ~int64_t(31); };
    const int64_t C = v.cord, W = v.walk_num;
    const bool fused = een_fused(v), pair = !fused && C >= 2;
    gN = 0;
    R0 = up(gN + (fused ? 0 : W * 16 * gn_na(v.nucl_num) * gn_jp(v.elec_num)));
    Jt = up(R0 + (pair ? W * v.elec_num * (C - 1) * 5 : 0));
    Sg = up(Jt + W * een_shape(v).ntiles);
    JtP = up(Sg + (pair ? W * v.nucl_num * (C - 1) : 0));
    total = up(JtP + (fused ? 2 * W * een_shape(v).ntiles : 0));
  }
};

int64_t jastrow_scratch_doubles(const JasDevView& v) { return ScratchLayout(v).total; }

struct ScratchView {
  double *R0, *gN, *Jt, *Sg, *JtP;
};

ScratchView scratch_view(const JasDevView& v, const JasOut& o) {
  const ScratchLayout l(v);
  ScratchView sv;
  sv.gN = o.scratch + l.gN;
  sv.R0 = o.scratch + l.R0;
  sv.Jt = o.scratch + l.Jt;
  sv.Sg = o.scratch + l.Sg;
  sv.JtP = o.scratch + l.JtP;
  return sv;
}

// Whether the evaluation runs the persistent LDS-resident e-e-n kernel, which then also
// computes the e-e and e-n terms and the sums k_jas_pair would provide (k_jas_pair is not
// launched; k_jas_total adds the per-tile value sums).
bool een_fused(const JasDevView& v) {
  if (!(v.cord >= 2 && v.dim_c > 0)) return false;
  const EenShape sh = een_shape(v);
  return sh.mfma && sh.res && v.walk_num * sh.ntiles <= INT32_MAX;
}

hipError_t launch_jastrow_pair(const JasDevView& v, const JasOut& o, hipStream_t s) {
  if (een_fused(v)) return hipSuccess;  // done by the resident e-e-n kernel
  const ScratchView sv = scratch_view(v, o);
  const int pair_threads = (int)std::min<int64_t>(pair_block((int)v.cord), 64 * ((v.elec_num + 63) / 64));
  const size_t lds_pair = sizeof(double) * (3 * v.elec_num + kPairBlock / 64);
  if (lds_pair > 160 * 1024) return hipErrorInvalidValue;
  switch (v.cord) {
#define PAIR_CASE(CC)                                                                                         \
  case CC:                                                                                                    \
    hipLaunchKernelGGL(k_jas_pair<CC>, dim3((unsigned)v.walk_num), dim3(pair_threads), lds_pair, s, v, o, sv.R0,   \
                       sv.gN, sv.Sg);                                                                         \
    break;
    PAIR_CASE(0)
    PAIR_CASE(1)
    PAIR_CASE(2)
    PAIR_CASE(3)
    PAIR_CASE(4)
    PAIR_CASE(5)
    PAIR_CASE(6)
    PAIR_CASE(7)
    PAIR_CASE(8)
#undef PAIR_CASE
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Workgroups of the persistent e-e-n kernel: one per CU of the current device (its LDS
// allows no more), at most one per item.
static unsigned een_res_grid(int64_t nitem) {
  static std::atomic<int> ncu[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = -1;
  int n = (dev >= 0 && dev < 64) ? ncu[dev].load() : 0;
  if (n <= 0) {
    if (dev < 0 || hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
      n = 256;
    if (dev >= 0 && dev < 64) ncu[dev].store(n);
  }
  return (unsigned)std::min<int64_t>(nitem, n);
}

hipError_t launch_jastrow_een(const JasDevView& v, const JasOut& o, hipStream_t s) {
  const ScratchView sv = scratch_view(v, o);
  const int C = (int)v.cord;
  if (!(C >= 2 && v.dim_c > 0)) return hipSuccess;
  const EenShape sh = een_shape(v);
  const dim3 grid((unsigned)(v.walk_num * sh.ntiles));
  if (sh.mfma) {
    if (sh.lds > 160 * 1024) return hipErrorInvalidValue;
    if (sh.res && v.walk_num * sh.ntiles > INT32_MAX) return hipErrorInvalidValue;  // 32-bit item index
#define EEN_MFMA_CASE(CC)                                                                                      \
  case CC: {                                                                                                   \
    static bool attr_set = false;                                                                              \
    if (!attr_set) {                                                                                           \
      hipError_t ea = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_jas_een_mfma<CC>),                  \
                                          hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);             \
      if (ea == hipSuccess)                                                                                    \
        ea = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_jas_een_res<CC, false>),                     \
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);                      \
      if (ea == hipSuccess && (CC == 4 || CC == 5))                                                            \
        ea = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_jas_een_bal<CC == 4 ? 4 : 5, false>),         \
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);                      \
      if (ea != hipSuccess) return ea;                                                                         \
      attr_set = true;                                                                                         \
    }                                                                                                          \
    if (sh.res)                                                                                                \
      hipLaunchKernelGGL((k_jas_een_res<CC, false>), dim3(een_res_grid(v.walk_num * sh.ntiles)), dim3(64 * EQ_WAVES), \
                         sh.lds, s, v, o, sv.Jt, sv.JtP, sh.ntiles);                                           \
    else if (sh.bal)                                                                                           \
      hipLaunchKernelGGL((k_jas_een_bal<CC == 4 ? 4 : 5, false>), grid, dim3(64 * bal_waves(CC)), sh.lds, s, v, o, \
                         sv.R0, sv.gN, sv.Sg, sv.Jt, sh.ntiles, sh.JC);                                         \
    else                                                                                                       \
      hipLaunchKernelGGL(k_jas_een_mfma<CC>, grid, dim3(64 * EenWaves<CC>::n), sh.lds, s, v, o, sv.R0, sv.gN,    \
                         sv.Sg, sv.Jt, sh.ntiles, sh.JC);                                                       \
    break;                                                                                                     \
  }
    switch (C) {
      EEN_MFMA_CASE(2)
      EEN_MFMA_CASE(3)
      EEN_MFMA_CASE(4)
      EEN_MFMA_CASE(5)
      EEN_MFMA_CASE(6)
      default:
        return hipErrorInvalidValue;
    }
#undef EEN_MFMA_CASE
    return hipGetLastError();
  }
  if (sh.lds > 160 * 1024) return hipErrorInvalidValue;
#define EEN_CASE(CC)                                                                                     \
  case CC:                                                                                               \
    hipLaunchKernelGGL(k_jas_een<CC>, grid, dim3(kBlock), sh.lds, s, v, o, sv.R0, sv.gN, sv.Jt, sh.TI, sh.JC, \
                       sh.ntiles);                                                                       \
    break;
  switch (C) {
    EEN_CASE(2)
    EEN_CASE(3)
    EEN_CASE(4)
    EEN_CASE(5)
    EEN_CASE(6)
    EEN_CASE(7)
    EEN_CASE(8)
    default:
      return hipErrorInvalidValue;
  }
#undef EEN_CASE
  return hipGetLastError();
}

hipError_t launch_jastrow_total(const JasDevView& v, const JasOut& o, hipStream_t s) {
  const ScratchView sv = scratch_view(v, o);
  const int have_een = (v.cord >= 2 && v.dim_c > 0) ? 1 : 0;
  const EenShape sh = een_shape(v);
  hipLaunchKernelGGL(k_jas_total, dim3((unsigned)v.walk_num), dim3(kBlock), 0, s, v, o, sv.Jt, sv.JtP, sh.ntiles,
                     have_een, een_fused(v) ? 1 : 0, 0);
  return hipGetLastError();
}

bool jastrow_values_only_ok(const JasDevView& v) {
  if (een_fused(v)) return true;
  if (!(v.cord >= 2 && v.dim_c > 0)) return false;
  const EenShape sh = een_shape(v);
  return sh.mfma && sh.bal;
}

hipError_t launch_jastrow_values(const JasDevView& v, const JasOut& o, hipStream_t s) {
  if (!jastrow_values_only_ok(v)) return hipErrorInvalidValue;
  const ScratchView sv = scratch_view(v, o);
  const EenShape sh = een_shape(v);
  if (sh.lds > 160 * 1024) return hipErrorInvalidValue;
  if (!een_fused(v)) {
    // the pair kernel (e-e, e-n terms with their GL, the e-e-n inputs), the balanced e-e-n
    // kernel on the values, the totals without GL
    hipError_t e = launch_jastrow_pair(v, o, s);
    if (e != hipSuccess) return e;
#define EEN_BAL_VAL(CC)                                                                                        \
  {                                                                                                            \
    static bool attr_set = false;                                                                              \
    if (!attr_set) {                                                                                           \
      const hipError_t ea = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_jas_een_bal<CC, true>),       \
                                                hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);       \
      if (ea != hipSuccess) return ea;                                                                         \
      attr_set = true;                                                                                         \
    }                                                                                                          \
    hipLaunchKernelGGL((k_jas_een_bal<CC, true>), dim3((unsigned)(v.walk_num * sh.ntiles)),                    \
                       dim3(64 * bal_waves(CC)), sh.lds, s, v, o, sv.R0, sv.gN, sv.Sg, sv.Jt, sh.ntiles, sh.JC); \
  }
    if (v.cord == 4)
      EEN_BAL_VAL(4)
    else
      EEN_BAL_VAL(5)
#undef EEN_BAL_VAL
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_jas_total, dim3((unsigned)v.walk_num), dim3(kBlock), 0, s, v, o, sv.Jt, sv.JtP, sh.ntiles, 1,
                       0, 1);
    return hipGetLastError();
  }
  switch (v.cord) {
#define EEN_VAL_CASE(CC)                                                                                       \
  case CC: {                                                                                                   \
    static bool attr_set = false;                                                                              \
    if (!attr_set) {                                                                                           \
      const hipError_t ea = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_jas_een_res<CC, true>),       \
                                                hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);       \
      if (ea != hipSuccess) return ea;                                                                         \
      attr_set = true;                                                                                         \
    }                                                                                                          \
    hipLaunchKernelGGL((k_jas_een_res<CC, true>), dim3(een_res_grid(v.walk_num * sh.ntiles)),                  \
                       dim3(64 * EQ_WAVES), sh.lds, s, v, o, sv.Jt, sv.JtP, sh.ntiles);                        \
    break;                                                                                                     \
  }
    EEN_VAL_CASE(2)
    EEN_VAL_CASE(3)
    EEN_VAL_CASE(4)
    EEN_VAL_CASE(5)
    EEN_VAL_CASE(6)
#undef EEN_VAL_CASE
    default:
      return hipErrorInvalidValue;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_jas_total, dim3((unsigned)v.walk_num), dim3(kBlock), 0, s, v, o, sv.Jt, sv.JtP, sh.ntiles, 1,
                     1, 1);
  return hipGetLastError();
}

// ---- grouped launches of a batch of contexts -------------------------------------------
namespace {

enum { JB_PAIR = 0, JB_EEN_VALU = 1, JB_EEN_MFMA = 2, JB_EEN_BAL = 3, JB_EEN_RES = 4, JB_TOTAL = 5 };

template <class K>
hipError_t allow_lds(K kernel) {
  return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                             160 * 1024);
}

template <int C>
hipError_t launch_batch_c(const JasBatchLaunch& l, hipStream_t s) {
  const dim3 g((unsigned)l.nblocks), b((unsigned)l.threads);
  hipError_t e = hipSuccess;
  switch (l.kind) {
    case JB_PAIR:
      if ((e = allow_lds(&k_jas_pair_batch<C>)) != hipSuccess) return e;
      hipLaunchKernelGGL(k_jas_pair_batch<C>, g, b, l.lds, s, l.d_jobs, l.d_starts, l.njobs);
      break;
    case JB_EEN_VALU:
      if constexpr (C >= 2) {
        if ((e = allow_lds(&k_jas_een_batch<C>)) != hipSuccess) return e;
        hipLaunchKernelGGL(k_jas_een_batch<C>, g, b, l.lds, s, l.d_jobs, l.d_starts, l.njobs);
        break;
      }
      return hipErrorInvalidValue;
    case JB_EEN_MFMA:
      if constexpr (C >= 2 && C <= EMFMA_MAX_C) {
        if ((e = allow_lds(&k_jas_een_mfma_batch<C>)) != hipSuccess) return e;
        hipLaunchKernelGGL(k_jas_een_mfma_batch<C>, g, b, l.lds, s, l.d_jobs, l.d_starts, l.njobs);
        break;
      }
      return hipErrorInvalidValue;
    case JB_EEN_BAL:
      if constexpr (C == 4 || C == 5) {
        if ((e = allow_lds(&k_jas_een_bal_batch<C>)) != hipSuccess) return e;
        hipLaunchKernelGGL(k_jas_een_bal_batch<C>, g, b, l.lds, s, l.d_jobs, l.d_starts, l.njobs);
        break;
      }
      return hipErrorInvalidValue;
    case JB_EEN_RES:
      if constexpr (C >= 2 && C <= EMFMA_MAX_C) {
        if ((e = allow_lds(&k_jas_een_res_batch<C>)) != hipSuccess) return e;
        hipLaunchKernelGGL(k_jas_een_res_batch<C>, g, b, l.lds, s, l.d_jobs, l.d_starts, l.njobs);
        break;
      }
      return hipErrorInvalidValue;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Workgroups of the resident kernel per job: about kResBlocksPerCu per CU in all, shared in
// proportion to the jobs' work, each job's count a divisor of its walker count (at least
// one) so that every workgroup takes whole walkers (the walker's tables are built once per
// workgroup).  Many workgroups, not one per CU: the batch's other kernels run beside this
// one, and the dispatcher balances workgroups as CUs free up where a static share would not.
constexpr int kResBlocksPerCu = 4;
std::vector<int> res_shares(const std::vector<double>& cost, const std::vector<int64_t>& walkers, int ncu) {
  const int n = (int)cost.size();
  double sum = 0.0;
  for (double c : cost) sum += c;
  std::vector<int> sh(n, 1);
  for (int m = 0; m < n; ++m) {
    const double x = sum > 0 ? (double)kResBlocksPerCu * ncu * cost[m] / sum : 1.0;
    const int64_t W = walkers[m];
    int64_t best = 1;
    for (int64_t d = 1; d <= W && d <= (int64_t)std::ceil(x); ++d)
      if (W % d == 0) best = d;
    sh[m] = (int)best;
  }
  return sh;
}

}  // namespace

hipError_t jastrow_batch_build(const std::vector<std::pair<JasDevView, JasOut>>& in, JasBatchPlan& plan) {
  jastrow_batch_free(plan);
  // family key (kind, C) -> jobs, block counts, threads, LDS
  struct Fam {
    std::vector<JasJob> jobs;
    std::vector<int64_t> blocks, walkers;
    std::vector<double> cost;
    int threads = 0;
    size_t lds = 0;
  };
  std::map<std::pair<int, int>, Fam> fam;
  for (const auto& vo : in) {
    const JasDevView& v = vo.first;
    const JasOut& o = vo.second;
    const ScratchView sv = scratch_view(v, o);
    const EenShape sh = een_shape(v);
    const int C = (int)v.cord;
    const bool have_een = C >= 2 && v.dim_c > 0, fused = een_fused(v);
    JasJob J;
    J.v = v;
    J.o = o;
    J.R0 = sv.R0;
    J.gN = sv.gN;
    J.Sg = sv.Sg;
    J.Jt = sv.Jt;
    J.JtP = sv.JtP;
    J.ntiles = sh.ntiles;
    J.TI = sh.TI;
    J.JC = sh.JC;
    J.have_een = have_een ? 1 : 0;
    J.fused = fused ? 1 : 0;
    const int64_t W = v.walk_num, Ne = v.elec_num;
    auto add = [&](int kind, int64_t blocks, int threads, size_t lds, double cost) {
      Fam& f = fam[{kind, kind == JB_TOTAL ? 0 : C}];
      f.jobs.push_back(J);
      f.blocks.push_back(blocks);
      f.walkers.push_back(W);
      f.cost.push_back(cost);
      f.threads = std::max(f.threads, threads);
      f.lds = std::max(f.lds, lds);
    };
    if (!fused) {
      const size_t lds_pair = sizeof(double) * (3 * Ne + kPairBlock / 64);
      if (lds_pair > 160 * 1024) return hipErrorInvalidValue;
      add(JB_PAIR, W, (int)std::min<int64_t>(pair_block(C), 64 * ((Ne + 63) / 64)), lds_pair, 0.0);
    }
    if (have_een) {
      if (sh.lds > 160 * 1024) return hipErrorInvalidValue;
      if (fused)
        add(JB_EEN_RES, 0, 64 * EQ_WAVES, sh.lds,
            (double)W * sh.ntiles * gn_jp(Ne) * 16.0 * gn_na(v.nucl_num));
      else if (sh.bal)
        add(JB_EEN_BAL, W * sh.ntiles, 64 * bal_waves(C), sh.lds, 0.0);
      else if (sh.mfma)
        add(JB_EEN_MFMA, W * sh.ntiles, 64 * (C * (C - 1) / 2), sh.lds, 0.0);
      else
        add(JB_EEN_VALU, W * sh.ntiles, kBlock, sh.lds, 0.0);
    }
    add(JB_TOTAL, W, kBlock, 0, 0.0);
  }
  int ncu = 256, dev = 0;
  if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  hipError_t e = hipSuccess;
  for (auto& kv : fam) {  // in (kind, C) order: pair, e-e-n kernels, total
    Fam& f = kv.second;
    const int n = (int)f.jobs.size();
    if (kv.first.first == JB_EEN_RES) {
      const std::vector<int> sh = res_shares(f.cost, f.walkers, ncu);
      for (int m = 0; m < n; ++m) f.blocks[m] = sh[m];
    }
    std::vector<int> starts(n + 1, 0);
    int64_t acc = 0;
    for (int m = 0; m < n; ++m) {
      starts[m] = (int)acc;
      acc += f.blocks[m];
    }
    if (acc > INT32_MAX) return hipErrorInvalidValue;
    starts[n] = (int)acc;
    JasBatchLaunch l;
    l.kind = kv.first.first;
    l.C = kv.first.second;
    l.njobs = n;
    l.nblocks = (int)acc;
    l.threads = f.threads;
    l.lds = f.lds;
    e = hipMalloc(reinterpret_cast<void**>(&l.d_jobs), sizeof(JasJob) * n);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&l.d_starts), sizeof(int) * (n + 1));
    if (e == hipSuccess) e = hipMemcpy(l.d_jobs, f.jobs.data(), sizeof(JasJob) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(l.d_starts, starts.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice);
    plan.launches.push_back(l);
    if (e != hipSuccess) break;
  }
  if (e != hipSuccess) jastrow_batch_free(plan);
  return e;
}

hipError_t jastrow_batch_launch(const JasBatchPlan& plan, hipStream_t s, hipStream_t s_res, hipEvent_t res_done) {
  // the resident e-e-n kernel on s_res (it needs nothing from the other kernels); the pair
  // and other e-e-n kernels on s; the totals on s once s_res is done
  hipError_t e = hipSuccess;
  bool res = false;
  for (const JasBatchLaunch& l : plan.launches) {
    if (l.nblocks <= 0) continue;
    hipStream_t st = l.kind == JB_EEN_RES ? s_res : s;
    if (l.kind == JB_EEN_RES) res = true;
    if (l.kind == JB_TOTAL) {
      if (res) {
        e = hipEventRecord(res_done, s_res);
        if (e == hipSuccess) e = hipStreamWaitEvent(s, res_done, 0);
        if (e != hipSuccess) return e;
        res = false;
      }
      hipLaunchKernelGGL(k_jas_total_batch, dim3((unsigned)l.nblocks), dim3(kBlock), 0, st, l.d_jobs, l.d_starts,
                         l.njobs);
      e = hipGetLastError();
    } else {
      switch (l.C) {
        case 0: e = launch_batch_c<0>(l, st); break;
        case 1: e = launch_batch_c<1>(l, st); break;
        case 2: e = launch_batch_c<2>(l, st); break;
        case 3: e = launch_batch_c<3>(l, st); break;
        case 4: e = launch_batch_c<4>(l, st); break;
        case 5: e = launch_batch_c<5>(l, st); break;
        case 6: e = launch_batch_c<6>(l, st); break;
        case 7: e = launch_batch_c<7>(l, st); break;
        case 8: e = launch_batch_c<8>(l, st); break;
        default: e = hipErrorInvalidValue;
      }
    }
    if (e != hipSuccess) return e;
  }
  if (res) {  // no totals launched (cannot happen: every job has one); join anyway
    e = hipEventRecord(res_done, s_res);
    if (e == hipSuccess) e = hipStreamWaitEvent(s, res_done, 0);
  }
  return e;
}

void jastrow_batch_free(JasBatchPlan& plan) {
  for (JasBatchLaunch& l : plan.launches) {
    if (l.d_jobs) (void)hipFree(l.d_jobs);
    if (l.d_starts) (void)hipFree(l.d_starts);
  }
  plan.launches.clear();
}

int64_t jastrow_inter_size(const JasDevView& v, int kind) {
  const int64_t W = v.walk_num, Ne = v.elec_num, Nn = v.nucl_num, C = v.cord;
  switch (kind) {
    case J_EE_RESCALED: return W * Ne * Ne;
    case J_EE_RESCALED_GL: return W * Ne * Ne * 4;
    case J_EN_RESCALED: return W * Nn * Ne;
    case J_EN_RESCALED_GL: return W * Nn * Ne * 4;
    case J_EEN_E: return W * (C + 1) * Ne * Ne;
    case J_EEN_E_GL: return W * (C + 1) * Ne * 4 * Ne;
    case J_EEN_N: return W * (C + 1) * Nn * Ne;
    case J_EEN_N_GL: return W * (C + 1) * Nn * 4 * Ne;
    case J_TMP_C: return W * C * (C + 1) * Nn * Ne;
    case J_DTMP_C: return W * C * (C + 1) * Nn * 4 * Ne;
    default: return 0;
  }
}

hipError_t launch_jastrow_inter(const JasDevView& v, int kind, double* out, hipStream_t s) {
  const int64_t n = jastrow_inter_size(v, kind);
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_jas_inter, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, v, kind, out, n);
  return hipGetLastError();
}

hipError_t launch_ee_distance(const double* coord, int64_t W, int64_t Ne, double* out, hipStream_t s) {
  const int64_t n = W * Ne * Ne;
  hipLaunchKernelGGL(k_ee_distance, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, coord, W, Ne, out);
  return hipGetLastError();
}

hipError_t launch_en_distance(const double* coord, int64_t P, int64_t ld, const double* nucl_coord, int64_t Nn,
                              double* out, hipStream_t s) {
  const int64_t n = P * Nn;
  hipLaunchKernelGGL(k_en_distance, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, coord, P, ld, nucl_coord, Nn, out);
  return hipGetLastError();
}

}  // namespace qmckl_amd
