// Adjoint of the CHAMP Jastrow for the walker-summed functional
//   L = sum_w Phi_w,  Phi_w = wJ J_w + wL E_w,
//   E_w = -1/2 sum_i (lap_i J + |grad_i J|^2) + V_ee + V_en + V_nn
// (ln Psi_J and a Jastrow-only local-energy surrogate), with respect to the Jastrow
// parameters a [type][aord+1], b [bord+1], c [type][dim_c] and the nuclear coordinates.
// No counterpart in the reference; the forward quantities are those of
// org/qmckl_jastrow_champ.org (factor_ee :3582-3650, factor_en :7021-7095, factor_een
// :14121-14175 and their GLs) and org/qmckl_electron.org:1022-1192, 1477-1658 (potentials).
//
// With G_i = grad_i J fixed from the forward pass, dPhi = wJ dJ - wL/2 sum_i d lap_i J
// - wL sum_i G_i . dG_i.  J is a sum of terms tau (e-e pairs, e-n pairs, e-e-n triples),
// so dL/dtheta is the sum over the terms of dW_tau/dtheta for the per-term functional
//   W_tau = wJ tau + sum_p h_p . grad_p tau - wL/2 sum_p lap_p tau,  h_p = -wL G_p,
// p over the electrons of the term.  Each term's W is evaluated in closed form from the
// distances of that term: nothing of the forward pass is kept but G.  Parameter
// derivatives come from the analytic derivatives of the pair functions (e-e, e-n) or, for
// the linear c, from W of each basis term; nuclear-coordinate derivatives by forward-mode
// differentiation in the three coordinates of the term's nucleus (D3 below): every e-n and
// e-e-n term involves exactly one nucleus.
//
// One workgroup per walker; a thread takes the electrons i = tid + k blockDim.  Sums are
// reduced in a fixed order (shuffles within a wave, then the waves in order), so results
// are run-to-run reproducible.  Outputs per walker (summed over walkers afterwards):
//   [0] Phi_w, then dPhi/da [type][aord+1], dPhi/db [bord+1], dPhi/dc [type][dim_c],
//   dPhi/dR [nucl][3] (V_nn's share is added by the host).
#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace qmckl_amd {
namespace {

constexpr int kAdjBlock = 256;
constexpr int kAdjWaves = kAdjBlock / 64;
constexpr int kAdjMaxOrd = 8;  // aord, bord <= 8 (ord + 1 accumulators)

// A value with its derivatives in the three coordinates of one nucleus.
struct D3 {
  double v, d0, d1, d2;
};
__device__ __forceinline__ D3 mk(double v) { return D3{v, 0.0, 0.0, 0.0}; }
__device__ __forceinline__ D3 operator+(D3 a, D3 b) { return D3{a.v + b.v, a.d0 + b.d0, a.d1 + b.d1, a.d2 + b.d2}; }
__device__ __forceinline__ D3 operator-(D3 a, D3 b) { return D3{a.v - b.v, a.d0 - b.d0, a.d1 - b.d1, a.d2 - b.d2}; }
__device__ __forceinline__ D3 operator*(D3 a, D3 b) {
  return D3{a.v * b.v, a.d0 * b.v + a.v * b.d0, a.d1 * b.v + a.v * b.d1, a.d2 * b.v + a.v * b.d2};
}
__device__ __forceinline__ D3 operator*(double s, D3 a) { return D3{s * a.v, s * a.d0, s * a.d1, s * a.d2}; }
__device__ __forceinline__ D3 operator+(double s, D3 a) { return D3{s + a.v, a.d0, a.d1, a.d2}; }
__device__ __forceinline__ D3 inv(D3 a) {
  const double r = 1.0 / a.v, r2 = -r * r;
  return D3{r, r2 * a.d0, r2 * a.d1, r2 * a.d2};
}
__device__ __forceinline__ D3 dexp(D3 a) {
  const double e = exp(a.v);
  return D3{e, e * a.d0, e * a.d1, e * a.d2};
}
__device__ __forceinline__ D3 dsqrt(D3 a) {
  const double s = sqrt(a.v), h = 0.5 / s;
  return D3{s, h * a.d0, h * a.d1, h * a.d2};
}

// Pair function of the e-e and e-n terms at x (the rescaled distance),
//   u(x) = s p0 x / (1 + p1 x) + sum_{q=2..ord} p_q x^q - u_inf,
// u_inf its limit at x = 1/kappa: values u, u', u'' (D3 in x) and, for q = 0..ord, their
// derivatives in p_q (values only).
template <class Q>
__device__ __forceinline__ void pair_fn(const double* p, int ord, double s, double kinv, D3 x, D3& u, D3& u1, D3& u2,
                                        Q&& dq) {
  const D3 den = inv(1.0 + p[1] * x), den2 = den * den, den3 = den2 * den;
  const double xi = kinv, di = 1.0 / (1.0 + p[1] * xi);
  u = (s * p[0]) * (x * den) - mk(s * p[0] * xi * di);
  u1 = (s * p[0]) * den2;
  u2 = (-2.0 * s * p[0] * p[1]) * den3;
  // parameter derivatives of (u, u', u'')
  dq(0, s * x.v * den.v - s * xi * di, s * den2.v, -2.0 * s * p[1] * den3.v);
  dq(1, -s * p[0] * x.v * x.v * den2.v + s * p[0] * xi * xi * di * di, -2.0 * s * p[0] * x.v * den3.v,
     -2.0 * s * p[0] * den3.v + 6.0 * s * p[0] * p[1] * x.v * den3.v * den.v);
  D3 xq1 = x, xq2 = mk(1.0);  // x^(q-1), x^(q-2)
  double xiq = xi;             // xi^(q-1)
#pragma unroll
  for (int q = 2; q <= kAdjMaxOrd; ++q) {
    if (q > ord) break;
    const D3 xq = xq1 * x;
    xiq *= xi;
    u = u + p[q] * (xq - mk(xiq));
    u1 = u1 + ((double)q * p[q]) * xq1;
    u2 = u2 + ((double)(q * (q - 1)) * p[q]) * xq2;
    dq(q, xq.v - xiq, (double)q * xq1.v, (double)(q * (q - 1)) * xq2.v);
    xq2 = xq1;
    xq1 = xq;
  }
}

// Fixed-order sum over the workgroup of the N values of each thread; the result is valid
// in thread 0.  red: LDS [kAdjWaves][N].
template <int N>
__device__ __forceinline__ void block_sums(double (&vals)[N], double* red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < N; ++q) {
    double x = vals[q];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    if (lane == 0) red[wave * N + q] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < N; ++q) {
      double s = 0.0;
      for (int w = 0; w < kAdjWaves; ++w) s += red[w * N + q];
      vals[q] = s;
    }
  }
  __syncthreads();
}

__global__ void __launch_bounds__(kAdjBlock) k_jas_adjoint(AdjArgs A) {
  constexpr int NACC = (kAdjMaxOrd + 1) + 3;  // a derivatives, force (registers)
  extern __shared__ double lds[];
  const JasDevView& v = A.v;
  const int Ne = (int)v.elec_num, Nn = (int)v.nucl_num, up = (int)v.up_num;
  const int ao1 = (int)v.aord + 1, bo1 = (int)v.bord + 1, T = A.ntype;
  const int DC = v.cord >= 2 ? (int)v.dim_c : 0;
  const int64_t w = blockIdx.x;
  double* xe = lds;                      // [Ne][3]
  double* hh = xe + 3 * Ne;              // [Ne][3]: h_i = -wL G_i
  double* acc = hh + 3 * Ne;             // [nout] this walker's sums
  double* red = acc + A.nout;            // [kAdjWaves][NACC]
  double* cacc = red + kAdjWaves * NACC;  // [kAdjBlock][DC] per-thread d/dc sums of one nucleus
  int* lkm = reinterpret_cast<int*>(cacc + kAdjBlock * DC);  // [3][DC]: l, k, m of term n
  const int off_a = 1, off_b = off_a + T * ao1, off_c = off_b + bo1, off_r = off_c + T * DC;
  const double wJ = A.wJ, wL = A.wL;
  for (int t = threadIdx.x; t < Ne; t += kAdjBlock) {
    for (int d = 0; d < 3; ++d) {
      xe[3 * t + d] = v.coord[w * Ne + t + d * v.coord_ld];
      const int64_t g = (w * 4 + d) * Ne + t;
      hh[3 * t + d] = -wL * (A.fee_gl[g] + A.fen_gl[g] + A.feen_gl[g]);
    }
  }
  for (int t = threadIdx.x; t < A.nout; t += kAdjBlock) acc[t] = 0.0;
  if (threadIdx.x == 0) {
    // (l, k, m) of the c-coefficients in the order of org/qmckl_jastrow_champ.org:12000-12040
    int n = 0;
    for (int p = 2; p <= (int)v.cord; ++p)
      for (int k = p - 1; k >= 0; --k)
        for (int l = (k != 0 ? p - k : p - k - 2); l >= 0; --l)
          if ((p - k - l) % 2 == 0 && n < DC) {
            lkm[n] = l;
            lkm[DC + n] = k;
            lkm[2 * DC + n] = (p - k - l) / 2;
            ++n;
          }
  }
  __syncthreads();

  // ---- value Phi_w and the e-e terms (b derivatives) -----------------------------------
  double vb[2 + kAdjMaxOrd];  // vb[0]: Phi, vb[1 + q]: d/db_q
#pragma unroll
  for (int q = 0; q < 2 + kAdjMaxOrd; ++q) vb[q] = 0.0;
  const double kee = v.rescale_ee, kei = 1.0 / kee;
  for (int i = threadIdx.x; i < Ne; i += kAdjBlock) {
    const int64_t g = w * 4 * Ne + i;
    double G2 = 0.0, Lp = A.fee_gl[g + 3 * Ne] + A.fen_gl[g + 3 * Ne] + A.feen_gl[g + 3 * Ne];
    for (int d = 0; d < 3; ++d) {
      const double gd = A.fee_gl[g + d * Ne] + A.fen_gl[g + d * Ne] + A.feen_gl[g + d * Ne];
      G2 += gd * gd;
    }
    double phi = -0.5 * wL * (Lp + G2);
    for (int a = 0; a < Nn; ++a) {  // V_en, org/qmckl_electron.org:1638-1644
      const double dx = xe[3 * i] - v.nucl_coord[a], dy = xe[3 * i + 1] - v.nucl_coord[a + Nn],
                   dz = xe[3 * i + 2] - v.nucl_coord[a + 2 * Nn];
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      if (r > 1e-6) phi -= wL * A.charge[a] / r;
    }
    for (int j = i + 1; j < Ne; ++j) {
      const double dx = xe[3 * i] - xe[3 * j], dy = xe[3 * i + 1] - xe[3 * j + 1], dz = xe[3 * i + 2] - xe[3 * j + 2];
      const double r = sqrt(dx * dx + dy * dy + dz * dz), ri = 1.0 / r;
      if (r > 1e-5) phi += wL * ri;  // V_ee, org/qmckl_electron.org:1172-1176
      const double E = exp(-kee * r), x = (1.0 - E) * kei;
      const bool same = (j < up) || (i >= up);
      const double s = (v.spin_independent != 1 && same) ? 0.5 : 1.0;
      // h_i . grad_i u + h_j . grad_j u = u' E (h_i - h_j) . r_hat;  lap_i u = lap_j u
      const double hr = ((hh[3 * i] - hh[3 * j]) * dx + (hh[3 * i + 1] - hh[3 * j + 1]) * dy +
                         (hh[3 * i + 2] - hh[3 * j + 2]) * dz) * ri;
      const double c1 = E * hr - wL * (2.0 * ri - kee) * E, c2 = -wL * E * E;
      D3 u, u1, u2;
      pair_fn(v.b_vector, (int)v.bord, s, kei, mk(x), u, u1, u2,
              [&](int q, double du, double du1, double du2) { vb[1 + q] += wJ * du + c1 * du1 + c2 * du2; });
    }
    vb[0] += phi;
  }
  if (threadIdx.x == 0) vb[0] += wJ * (A.fee[w] + A.fen[w] + A.feen[w]);
  block_sums(vb, red);
  if (threadIdx.x == 0) {
    acc[0] += vb[0];
#pragma unroll
    for (int q = 0; q <= kAdjMaxOrd; ++q)
      if (q < bo1) acc[off_b + q] += vb[1 + q];
  }

  // ---- per nucleus: e-n terms (a derivatives) and e-e-n terms (c derivatives), force -----
  double* cmine = cacc + threadIdx.x * DC;
  for (int a = 0; a < Nn; ++a) {
    const int t = (int)v.type_nucl[a];
    const double ka = v.kappa_en[t], kai = 1.0 / ka;
    const double* av = v.a_vector + t * ao1;
    const double Ra[3] = {v.nucl_coord[a], v.nucl_coord[a + Nn], v.nucl_coord[a + 2 * Nn]};
    // s: d/da_q [kAdjMaxOrd + 1], d/dR_a [3]
    constexpr int SF = kAdjMaxOrd + 1;
    double s[NACC];
#pragma unroll
    for (int q = 0; q < NACC; ++q) s[q] = 0.0;
    for (int n = 0; n < DC; ++n) cmine[n] = 0.0;
    for (int i = threadIdx.x; i < Ne; i += kAdjBlock) {
      // electron i to nucleus a, differentiable in R_a: d(r_i - R_a)/dR_a = -1
      const D3 dxi{xe[3 * i] - Ra[0], -1.0, 0.0, 0.0}, dyi{xe[3 * i + 1] - Ra[1], 0.0, -1.0, 0.0},
          dzi{xe[3 * i + 2] - Ra[2], 0.0, 0.0, -1.0};
      const D3 Ri = dsqrt(dxi * dxi + dyi * dyi + dzi * dzi), Rii = inv(Ri);
      const D3 Vi = dexp((-ka) * Ri);
      // e-n term: W = wJ u + u' E (h_i . R_hat) - wL/2 (u'' E^2 + u' (2/R - kappa) E)
      {
        const D3 x = kai * (1.0 + ((-1.0) * Vi));
        const D3 hR = (hh[3 * i] * dxi + hh[3 * i + 1] * dyi + hh[3 * i + 2] * dzi) * Rii;
        const D3 c1 = Vi * hR - (0.5 * wL) * ((2.0 * Rii - mk(ka)) * Vi), c2 = (-0.5 * wL) * (Vi * Vi);
        D3 u, u1, u2;
        pair_fn(av, (int)v.aord, 1.0, kai, x, u, u1, u2,
                [&](int q, double du, double du1, double du2) { s[q] += wJ * du + c1.v * du1 + c2.v * du2; });
        const D3 W = wJ * u + u1 * c1 + u2 * c2;
        s[SF] += W.d0;
        s[SF + 1] += W.d1;
        s[SF + 2] += W.d2;
        // V_en = -Z / |r_i - R_a|: dV/dR_a = -Z (r_i - R_a) / |r_i - R_a|^3
        if (Ri.v > 1e-6) {
          const double z3 = -wL * A.charge[a] * Rii.v * Rii.v * Rii.v;
          s[SF] += z3 * dxi.v;
          s[SF + 1] += z3 * dyi.v;
          s[SF + 2] += z3 * dzi.v;
        }
      }
      if (DC == 0) continue;
      // e-e-n terms tau = V_i^m E_ij^k V_j^(m+l) over the ordered pairs (i, j != i)
      const D3 Yix = (-ka) * (dxi * Rii), Yiy = (-ka) * (dyi * Rii), Yiz = (-ka) * (dzi * Rii);  // grad_i ln V_i
      const D3 hYi = hh[3 * i] * Yix + hh[3 * i + 1] * Yiy + hh[3 * i + 2] * Yiz;
      for (int j = 0; j < Ne; ++j) {
        if (j == i) continue;
        const double ex = xe[3 * i] - xe[3 * j], ey = xe[3 * i + 1] - xe[3 * j + 1], ez = xe[3 * i + 2] - xe[3 * j + 2];
        const double rij = sqrt(ex * ex + ey * ey + ez * ez), rinv = 1.0 / rij;
        // X = grad_i ln E_ij = -kappa_ee r_hat_ij; grad_j ln E_ij = -X
        const double Xx = -kee * ex * rinv, Xy = -kee * ey * rinv, Xz = -kee * ez * rinv;
        const D3 dxj{xe[3 * j] - Ra[0], -1.0, 0.0, 0.0}, dyj{xe[3 * j + 1] - Ra[1], 0.0, -1.0, 0.0},
            dzj{xe[3 * j + 2] - Ra[2], 0.0, 0.0, -1.0};
        const D3 Rj = dsqrt(dxj * dxj + dyj * dyj + dzj * dzj), Rji = inv(Rj);
        const D3 Yjx = (-ka) * (dxj * Rji), Yjy = (-ka) * (dyj * Rji), Yjz = (-ka) * (dzj * Rji);
        const double hX = (hh[3 * i] - hh[3 * j]) * Xx + (hh[3 * i + 1] - hh[3 * j + 1]) * Xy +
                          (hh[3 * i + 2] - hh[3 * j + 2]) * Xz;  // h_i . X + h_j . (-X)
        const D3 hYj = hh[3 * j] * Yjx + hh[3 * j + 1] * Yjy + hh[3 * j + 2] * Yjz;
        const D3 XYi = Xx * Yix + Xy * Yiy + Xz * Yiz;             // X . Y_i
        const D3 XYj = (-Xx) * Yjx + (-Xy) * Yjy + (-Xz) * Yjz;    // (-X) . Y_j
        const double kee2 = kee * kee, ka2 = ka * ka;
#pragma unroll 1
        for (int n = 0; n < DC; ++n) {
          const int l = lkm[n], k = lkm[DC + n], m = lkm[2 * DC + n], ml = m + l;
          // A_i = k X + m Y_i, A_j = -k X + (m + l) Y_j (unit r_hat: |X|^2 = kee^2, |Y|^2 = ka^2)
          // lap_i ln tau = -2 k kee / r_ij - 2 m ka / R_i; lap_j: -2 k kee / r_ij - 2 (m + l) ka / R_j
          const D3 tau = dexp(mk(-(double)k * kee * rij) + (-(double)m * ka) * Ri + (-(double)ml * ka) * Rj);
          const D3 hA = mk((double)k * hX) + (double)m * hYi + (double)ml * hYj;
          const D3 A2 = mk((double)(2 * k * k) * kee2 + (double)(m * m + ml * ml) * ka2) +
                        (2.0 * k * m) * XYi + (2.0 * k * ml) * XYj;
          const D3 lap = mk(-4.0 * k * kee * rinv) + (-2.0 * m * ka) * Rii + (-2.0 * ml * ka) * Rji;
          const D3 W = tau * (mk(wJ) + hA + (-0.5 * wL) * (A2 + lap));
          const double cn = v.cfull[n * Nn + a];
          cmine[n] += W.v;
          s[SF] += cn * W.d0;
          s[SF + 1] += cn * W.d1;
          s[SF + 2] += cn * W.d2;
        }
      }
    }
    block_sums(s, red);  // its barriers also order the cacc writes before the sums below
    if (threadIdx.x == 0) {
#pragma unroll
      for (int p = 0; p <= kAdjMaxOrd; ++p)
        if (p < ao1) acc[off_a + t * ao1 + p] += s[p];
#pragma unroll
      for (int d = 0; d < 3; ++d) acc[off_r + 3 * a + d] += s[SF + d];
    }
    for (int n = threadIdx.x; n < DC; n += kAdjBlock) {  // d/dc_n: the threads' sums in order
      double cs = 0.0;
      for (int q = 0; q < kAdjBlock; ++q) cs += cacc[q * DC + n];
      acc[off_c + t * DC + n] += cs;
    }
    __syncthreads();
  }
  for (int q = threadIdx.x; q < A.nout; q += kAdjBlock) A.out[w * A.nout + q] = acc[q];
}

}  // namespace

int adjoint_outputs(const JasDevView& v, int ntype) {
  const int dc = v.cord >= 2 ? (int)v.dim_c : 0;
  return 1 + ntype * (int)(v.aord + 1) + (int)(v.bord + 1) + ntype * dc + 3 * (int)v.nucl_num;
}

hipError_t launch_jastrow_adjoint(const AdjArgs& a, hipStream_t s) {
  const JasDevView& v = a.v;
  if (v.aord < 1 || v.aord > kAdjMaxOrd || v.bord < 1 || v.bord > kAdjMaxOrd) return hipErrorInvalidValue;
  const int dc = v.cord >= 2 ? (int)v.dim_c : 0;
  const size_t lds = sizeof(double) * (6 * (size_t)v.elec_num + a.nout + kAdjWaves * (kAdjMaxOrd + 4) +
                                       (size_t)kAdjBlock * dc) +
                     sizeof(int) * 3 * dc;
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  static bool attr = false;
  if (!attr) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_jas_adjoint),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (e != hipSuccess) return e;
    attr = true;
  }
  hipLaunchKernelGGL(k_jas_adjoint, dim3((unsigned)v.walk_num), dim3(kAdjBlock), lds, s, a);
  return hipGetLastError();
}

}  // namespace qmckl_amd
