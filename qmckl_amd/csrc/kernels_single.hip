// Single-electron moves of the CHAMP Jastrow (org/qmckl_jastrow_champ_single.org): for
// electron mu of every walker moved from r_mu to r', the changes of the e-e, e-n and e-e-n
// factors and of their gradients and Laplacians, and the intermediates of the new position.
//
// Only the terms that involve electron mu change: the e-e pairs (mu, j), the e-n pairs
// (mu, a) and the e-e-n terms V_i^m E_ij^k V_j^(m+l) with i = mu or j = mu (V = exp(-kappa_a
// R), E = exp(-kappa_ee r), the ordered-pair form of org/qmckl_jastrow_champ.org:14121-14175).
// Each is evaluated at the old and the new position and the difference accumulated, so a
// move costs O(Ne Nn dim_c) per walker instead of the O(Ne^2 Nn) of a full evaluation.
// Gradients and Laplacians of the terms come from their log-derivatives:
//   grad_i tau = tau A_i, lap_i tau = tau (|A_i|^2 + lap_i ln tau).
//
// One workgroup per walker, a thread per electron j (strided); the sums over j for
// electron mu are reduced in a fixed order.  Output layouts are the reference getters' (see
// SingleOut in internal.hpp).
#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace qmckl_amd {
namespace {

constexpr int kSingleBlock = 256;
constexpr int kSingleWaves = kSingleBlock / 64;
constexpr int kSingleMaxC = 8;

// e-e / e-n pair function s p0 x / (1 + p1 x) + sum_{q>=2} p_q x^q (no constant) and its
// first and second derivatives in x.
__device__ __forceinline__ void pade_poly(const double* p, int ord, double s, double x, double& u, double& u1,
                                          double& u2) {
  const double den = 1.0 / (1.0 + p[1] * x);
  u = s * p[0] * x * den;
  u1 = s * p[0] * den * den;
  u2 = -2.0 * s * p[0] * p[1] * den * den * den;
  double xq2 = 1.0, xq1 = x;  // x^(q-2), x^(q-1)
  for (int q = 2; q <= ord; ++q) {
    u2 += (double)(q * (q - 1)) * p[q] * xq2;
    u1 += (double)q * p[q] * xq1;
    xq2 = xq1;
    xq1 *= x;
    u += p[q] * xq1;
  }
}

// Value, gradient (in the coordinates of the first point) and Laplacian of u(x(|d|)),
// x = (1 - e^(-kappa r)) / kappa, for the displacement d of the first point from the second.
__device__ __forceinline__ void pair_vgl(const double* p, int ord, double s, double kappa, const double d[3],
                                         double out[5]) {
  const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]), ri = 1.0 / r;
  const double E = exp(-kappa * r), x = (1.0 - E) / kappa;
  double u, u1, u2;
  pade_poly(p, ord, s, x, u, u1, u2);
  const double g = u1 * E * ri;  // u' grad x = u' E d / r
  out[0] = u;
  out[1] = g * d[0];
  out[2] = g * d[1];
  out[3] = g * d[2];
  out[4] = u2 * E * E + u1 * (2.0 * ri - kappa) * E;
}

// Fixed-order block sum of n <= N values per thread; the result in thread 0.
template <int N>
__device__ __forceinline__ void single_sums(double (&v)[N], double* red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < N; ++q) {
    double x = v[q];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    if (lane == 0) red[wave * N + q] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < N; ++q) {
      double s = 0.0;
      for (int w = 0; w < kSingleWaves; ++w) s += red[w * N + q];
      v[q] = s;
    }
  }
  __syncthreads();
}

__global__ void __launch_bounds__(kSingleBlock) k_jas_single(JasDevView v, SingleArgs S) {
  extern __shared__ double lds[];
  const int Ne = (int)v.elec_num, Nn = (int)v.nucl_num, up = (int)v.up_num, C = (int)v.cord;
  const int DC = C >= 2 ? (int)v.dim_c : 0, mu = (int)S.num;
  const int64_t w = blockIdx.x;
  double* xe = lds;                              // [Ne][3] old positions
  double* red = xe + 3 * Ne;                     // [kSingleWaves][16]
  int* lkm = reinterpret_cast<int*>(red + kSingleWaves * 16);  // [3][DC]
  for (int t = threadIdx.x; t < Ne; t += kSingleBlock)
    for (int d = 0; d < 3; ++d) xe[3 * t + d] = v.coord[w * Ne + t + d * v.coord_ld];
  if (threadIdx.x == 0) {
    // (l, k, m) of the c-coefficients, org/qmckl_jastrow_champ.org:12000-12040
    int n = 0;
    for (int p = 2; p <= C; ++p)
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
  const double xn[3] = {S.coord[3 * w], S.coord[3 * w + 1], S.coord[3 * w + 2]};  // new position
  const double xo[3] = {xe[3 * mu], xe[3 * mu + 1], xe[3 * mu + 2]};             // old position
  const double kee = v.rescale_ee;
  const SingleOut& o = S.out;
  // sums for electron mu: [0..2] dJ_ee, dJ_en, dJ_een; [3..6] d(grad, lap)_ee of mu,
  // [7..10] _en, [11..14] _een
  double acc[15];
#pragma unroll
  for (int q = 0; q < 15; ++q) acc[q] = 0.0;

  // ---- e-e pairs (mu, j) and e-e-n terms with electron j ---------------------------------
  for (int j = threadIdx.x; j < Ne; j += kSingleBlock) {
    const double dn[3] = {xe[3 * j] - xn[0], xe[3 * j + 1] - xn[1], xe[3 * j + 2] - xn[2]};  // r_j - r'
    const double rn = sqrt(dn[0] * dn[0] + dn[1] * dn[1] + dn[2] * dn[2]);
    // new-position intermediates (the row of electron mu of the new matrices)
    o.ee_dist[w * Ne + j] = j == mu ? 0.0 : rn;
    const double En = exp(-kee * rn);
    o.ee_resc[w * Ne + j] = j == mu ? 0.0 : (1.0 - En) / kee;
    {
      // gradient in r_j of f(|r_j - r'|) and its Laplacian, [W][Ne][4]
      const double ri = 1.0 / (rn > 1e-20 ? rn : 1e-20), Ec = exp(-kee * (rn > 1e-20 ? rn : 1e-20));
      double* g = o.ee_resc_gl + (w * Ne + j) * 4;
      g[0] = dn[0] * ri * Ec;
      g[1] = dn[1] * ri * Ec;
      g[2] = dn[2] * ri * Ec;
      g[3] = (2.0 * ri - kee) * Ec;
      if (j == mu) g[0] = g[1] = g[2] = g[3] = 0.0;
    }
    for (int q = 0; q <= C; ++q) {
      // een_rescaled_single_e [W][C+1][Ne]: E'^q (0 for j = mu); _gl [W][C+1][4][Ne]: the
      // gradient in r' of E'^q and its Laplacian
      const double eq = j == mu ? 0.0 : exp(-(double)q * kee * rn);
      o.een_e[(w * (C + 1) + q) * Ne + j] = eq;
      const double ri = j == mu ? 0.0 : 1.0 / rn, kl = -(double)q * kee;
      double* g = o.een_e_gl + ((w * (C + 1) + q) * 4) * Ne + j;
      g[0] = -kl * dn[0] * ri * eq;  // d/dr' of exp(kl |r_j - r'|)
      g[Ne] = -kl * dn[1] * ri * eq;
      g[2 * Ne] = -kl * dn[2] * ri * eq;
      g[3 * Ne] = kl * (2.0 * ri + kl) * eq;
    }
    double dgee[4] = {0.0, 0.0, 0.0, 0.0}, dgeen[4] = {0.0, 0.0, 0.0, 0.0};
    if (j != mu) {
      const bool same = (j < up) == (mu < up);
      const double s = (v.spin_independent != 1 && same) ? 0.5 : 1.0;
      const double dold[3] = {xe[3 * j] - xo[0], xe[3 * j + 1] - xo[1], xe[3 * j + 2] - xo[2]};
      double pn[5], po[5];
      pair_vgl(v.b_vector, (int)v.bord, s, kee, dn, pn);    // in the coordinates of j
      pair_vgl(v.b_vector, (int)v.bord, s, kee, dold, po);
      acc[0] += pn[0] - po[0];
      for (int c = 0; c < 4; ++c) dgee[c] = pn[1 + c] - po[1 + c];
      // electron mu: grad_mu u = -grad_j u, same Laplacian
      for (int c = 0; c < 3; ++c) acc[3 + c] -= dgee[c];
      acc[6] += dgee[3];
      // e-e-n terms with electrons (mu, j), both orders
      if (DC > 0) {
        const double ro = sqrt(dold[0] * dold[0] + dold[1] * dold[1] + dold[2] * dold[2]);
        for (int a = 0; a < Nn; ++a) {
          const int t = (int)v.type_nucl[a];
          const double ka = v.kappa_en[t];
          const double Ra[3] = {v.nucl_coord[a], v.nucl_coord[a + Nn], v.nucl_coord[a + 2 * Nn]};
          const double dj[3] = {xe[3 * j] - Ra[0], xe[3 * j + 1] - Ra[1], xe[3 * j + 2] - Ra[2]};
          const double Rj = sqrt(dj[0] * dj[0] + dj[1] * dj[1] + dj[2] * dj[2]), Rji = 1.0 / Rj;
          const double bj[3] = {-ka * dj[0] * Rji, -ka * dj[1] * Rji, -ka * dj[2] * Rji};  // grad_j ln V_j
          double Sj[4] = {0.0, 0.0, 0.0, 0.0}, Smu[4] = {0.0, 0.0, 0.0, 0.0}, SJ = 0.0;
          for (int side = 0; side < 2; ++side) {  // 0: new position of mu, 1: old (subtracted)
            const double* pm = side == 0 ? xn : xo;
            const double sg = side == 0 ? 1.0 : -1.0;
            const double r = side == 0 ? rn : ro, rinv = 1.0 / r;
            const double* dd = side == 0 ? dn : dold;  // r_j - r_mu
            const double dm[3] = {pm[0] - Ra[0], pm[1] - Ra[1], pm[2] - Ra[2]};
            const double Rm = sqrt(dm[0] * dm[0] + dm[1] * dm[1] + dm[2] * dm[2]), Rmi = 1.0 / Rm;
            const double bm[3] = {-ka * dm[0] * Rmi, -ka * dm[1] * Rmi, -ka * dm[2] * Rmi};  // grad_mu ln V_mu
            // grad_mu ln E = +kee (r_j - r_mu) / r = -grad_j ln E
            const double am[3] = {kee * dd[0] * rinv, kee * dd[1] * rinv, kee * dd[2] * rinv};
            for (int n = 0; n < DC; ++n) {
              const double cn = v.cfull[n * Nn + a];
              if (cn == 0.0) continue;
              const int l = lkm[n], k = lkm[DC + n], m = lkm[2 * DC + n], ml = m + l;
              const double lnE = -(double)k * kee * r;
              // tau_A = V_mu^m E^k V_j^(m+l), tau_B = V_j^m E^k V_mu^(m+l)
              const double tA = exp(lnE - (double)m * ka * Rm - (double)ml * ka * Rj);
              const double tB = exp(lnE - (double)ml * ka * Rm - (double)m * ka * Rj);
              double AmA[3], AjA[3], AmB[3], AjB[3];
              for (int c = 0; c < 3; ++c) {
                AmA[c] = k * am[c] + m * bm[c];
                AjA[c] = -k * am[c] + ml * bj[c];
                AmB[c] = k * am[c] + ml * bm[c];
                AjB[c] = -k * am[c] + m * bj[c];
              }
              const double lk = -2.0 * k * kee * rinv;
              const double LmA = AmA[0] * AmA[0] + AmA[1] * AmA[1] + AmA[2] * AmA[2] + lk - 2.0 * m * ka * Rmi;
              const double LjA = AjA[0] * AjA[0] + AjA[1] * AjA[1] + AjA[2] * AjA[2] + lk - 2.0 * ml * ka * Rji;
              const double LmB = AmB[0] * AmB[0] + AmB[1] * AmB[1] + AmB[2] * AmB[2] + lk - 2.0 * ml * ka * Rmi;
              const double LjB = AjB[0] * AjB[0] + AjB[1] * AjB[1] + AjB[2] * AjB[2] + lk - 2.0 * m * ka * Rji;
              const double wA = sg * cn * tA, wB = sg * cn * tB;
              SJ += wA + wB;
              for (int c = 0; c < 3; ++c) {
                Smu[c] += wA * AmA[c] + wB * AmB[c];
                Sj[c] += wA * AjA[c] + wB * AjB[c];
              }
              Smu[3] += wA * LmA + wB * LmB;
              Sj[3] += wA * LjA + wB * LjB;
            }
          }
          acc[2] += SJ;
          for (int c = 0; c < 4; ++c) {
            acc[11 + c] += Smu[c];
            dgeen[c] += Sj[c];
          }
        }
      }
    }
    // the other electrons' rows; electron mu's row is written after the sums below
    if (j != mu) {
      double* g = o.d_ee_gl + (w * Ne + j) * 4;  // [W][Ne][4]
      for (int c = 0; c < 4; ++c) g[c] = dgee[c];
      o.d_en_gl[(w * Ne + j) * 4 + 0] = o.d_en_gl[(w * Ne + j) * 4 + 1] = 0.0;
      o.d_en_gl[(w * Ne + j) * 4 + 2] = o.d_en_gl[(w * Ne + j) * 4 + 3] = 0.0;
      for (int c = 0; c < 4; ++c) {
        o.d_een_gl[(w * 4 + c) * Ne + j] = dgeen[c];  // [W][4][Ne]
        o.d_een_g[(w * 4 + c) * Ne + j] = 0.0;
      }
    }
  }

  // ---- e-n pairs (mu, a) -----------------------------------------------------------------
  for (int a = threadIdx.x; a < Nn; a += kSingleBlock) {
    const int t = (int)v.type_nucl[a];
    const double ka = v.kappa_en[t];
    const double dnw[3] = {xn[0] - v.nucl_coord[a], xn[1] - v.nucl_coord[a + Nn], xn[2] - v.nucl_coord[a + 2 * Nn]};
    const double dol[3] = {xo[0] - v.nucl_coord[a], xo[1] - v.nucl_coord[a + Nn], xo[2] - v.nucl_coord[a + 2 * Nn]};
    const double Rn = sqrt(dnw[0] * dnw[0] + dnw[1] * dnw[1] + dnw[2] * dnw[2]);
    o.en_dist[w * Nn + a] = Rn;
    const double Vn = exp(-ka * Rn);
    o.en_resc[w * Nn + a] = (1.0 - Vn) / ka;
    {
      const double Rc = Rn > 1e-20 ? Rn : 1e-20, ri = 1.0 / Rc, Ec = exp(-ka * Rc);
      double* g = o.en_resc_gl + (w * Nn + a) * 4;  // [W][Nn][4], gradient in r'
      g[0] = dnw[0] * ri * Ec;
      g[1] = dnw[1] * ri * Ec;
      g[2] = dnw[2] * ri * Ec;
      g[3] = (2.0 * ri - ka) * Ec;
    }
    for (int q = 0; q <= C; ++q) {
      // een_rescaled_single_n [W][C+1][Nn]: V'^q; _gl [W][C+1][Nn][4]
      const double vq = q == 0 ? 1.0 : exp(-(double)q * ka * Rn);
      o.een_n[(w * (C + 1) + q) * Nn + a] = vq;
      const double ri = 1.0 / Rn, kl = -(double)q * ka;
      double* g = o.een_n_gl + ((w * (C + 1) + q) * Nn + a) * 4;
      g[0] = kl * dnw[0] * ri * vq;
      g[1] = kl * dnw[1] * ri * vq;
      g[2] = kl * dnw[2] * ri * vq;
      g[3] = (kl * 2.0 * ri + kl * kl) * vq;
    }
    double pn[5], po[5];
    pair_vgl(v.a_vector + t * (v.aord + 1), (int)v.aord, 1.0, ka, dnw, pn);
    pair_vgl(v.a_vector + t * (v.aord + 1), (int)v.aord, 1.0, ka, dol, po);
    acc[1] += pn[0] - po[0];
    for (int c = 0; c < 4; ++c) acc[7 + c] += pn[1 + c] - po[1 + c];
  }

  single_sums(acc, red);
  if (threadIdx.x == 0) {
    o.d_ee[w] = acc[0];
    o.d_en[w] = acc[1];
    o.d_een[w] = acc[2];
    for (int c = 0; c < 4; ++c) {
      o.d_ee_gl[(w * Ne + mu) * 4 + c] = acc[3 + c];
      o.d_en_gl[(w * Ne + mu) * 4 + c] = acc[7 + c];
      o.d_een_gl[(w * 4 + c) * Ne + mu] = acc[11 + c];
      o.d_een_g[(w * 4 + c) * Ne + mu] = c < 3 ? acc[11 + c] : 0.0;  // the gradient of mu only
    }
  }
}

// Accepts the move: electron mu of every walker takes its new position (device coordinates
// [3][W*Ne]), and the factors and factor GLs of the context take their deltas; the value
// and GL of e^J are formed again from them (org/qmckl_jastrow_champ.org:16578-16625,
// 17072-17110).
__global__ void __launch_bounds__(kSingleBlock) k_jas_single_accept(JasDevView v, SingleArgs S, JasOut J,
                                                                    double* coord) {
  const int Ne = (int)v.elec_num, mu = (int)S.num;
  const int64_t w = blockIdx.x;
  const SingleOut& o = S.out;
  __shared__ double sval;
  if (threadIdx.x == 0) {
    for (int d = 0; d < 3; ++d) coord[w * Ne + mu + d * v.coord_ld] = S.coord[3 * w + d];
    J.fee[w] += o.d_ee[w];
    J.fen[w] += o.d_en[w];
    J.feen[w] += o.d_een[w];
    const double Jt = J.fee[w] + J.fen[w] + J.feen[w];
    const double val = (Jt < -100.0) ? 0.0 : exp(Jt);
    J.value[w] = val;
    sval = val;
  }
  __syncthreads();
  const double val = sval;
  for (int i = threadIdx.x; i < Ne; i += kSingleBlock) {
    double s[4];
    for (int c = 0; c < 4; ++c) {
      const int64_t g = (w * 4 + c) * Ne + i;
      J.fee_gl[g] += o.d_ee_gl[(w * Ne + i) * 4 + c];
      J.fen_gl[g] += o.d_en_gl[(w * Ne + i) * 4 + c];
      J.feen_gl[g] += o.d_een_gl[g];
      s[c] = J.fee_gl[g] + J.fen_gl[g] + J.feen_gl[g];
    }
    s[3] = s[3] + s[0] * s[0] + s[1] * s[1] + s[2] * s[2];
    for (int c = 0; c < 4; ++c) J.gl[(w * 4 + c) * Ne + i] = s[c] * val;
  }
}

size_t single_lds(const JasDevView& v) {
  const int dc = v.cord >= 2 ? (int)v.dim_c : 0;
  return sizeof(double) * (3 * (size_t)v.elec_num + kSingleWaves * 16) + sizeof(int) * 3 * dc;
}

}  // namespace

int64_t single_out_doubles(const JasDevView& v) {
  const int64_t W = v.walk_num, Ne = v.elec_num, Nn = v.nucl_num, C1 = v.cord + 1;
  return W * (Ne + Nn + Ne + 4 * Ne + Nn + 4 * Nn + C1 * Ne + 4 * C1 * Ne + C1 * Nn + 4 * C1 * Nn + 3 + 4 * Ne +
              4 * Ne + 4 * Ne + 4 * Ne);
}

SingleOut single_out_view(const JasDevView& v, double* base) {
  const int64_t W = v.walk_num, Ne = v.elec_num, Nn = v.nucl_num, C1 = v.cord + 1;
  SingleOut o;
  double* p = base;
  auto take = [&](int64_t n) {
    double* r = p;
    p += n;
    return r;
  };
  o.ee_dist = take(W * Ne);
  o.en_dist = take(W * Nn);
  o.ee_resc = take(W * Ne);
  o.ee_resc_gl = take(W * Ne * 4);
  o.en_resc = take(W * Nn);
  o.en_resc_gl = take(W * Nn * 4);
  o.een_e = take(W * C1 * Ne);
  o.een_e_gl = take(W * C1 * 4 * Ne);
  o.een_n = take(W * C1 * Nn);
  o.een_n_gl = take(W * C1 * Nn * 4);
  o.d_ee = take(W);
  o.d_en = take(W);
  o.d_een = take(W);
  o.d_ee_gl = take(W * Ne * 4);
  o.d_en_gl = take(W * Ne * 4);
  o.d_een_gl = take(W * 4 * Ne);
  o.d_een_g = take(W * 4 * Ne);
  return o;
}

hipError_t launch_jastrow_single(const JasDevView& v, const SingleArgs& s, hipStream_t st) {
  if (v.cord > kSingleMaxC || s.num < 0 || s.num >= v.elec_num) return hipErrorInvalidValue;
  const size_t lds = single_lds(v);
  if (lds > 64 * 1024) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_jas_single, dim3((unsigned)v.walk_num), dim3(kSingleBlock), lds, st, v, s);
  return hipGetLastError();
}

hipError_t launch_jastrow_single_accept(const JasDevView& v, const SingleArgs& s, const JasOut& o, double* coord,
                                        hipStream_t st) {
  hipLaunchKernelGGL(k_jas_single_accept, dim3((unsigned)v.walk_num), dim3(kSingleBlock), 0, st, v, s, o, coord);
  return hipGetLastError();
}

}  // namespace qmckl_amd
