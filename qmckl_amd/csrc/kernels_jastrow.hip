// CHAMP Jastrow factor on gfx950, FP64 throughout (org/qmckl_jastrow_champ.org).
//
// Three kernels per evaluation:
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

#include "internal.hpp"

namespace qmckl_amd {

namespace {

constexpr int kBlock = 256;

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

__device__ __forceinline__ double block_sum(double x, double* red) {
  // Deterministic tree reduction over the 256 lanes of the block.
  red[threadIdx.x] = x;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// ------------------------------------------------------------------------------------
// e-e and e-n terms, and the q = 0 sums of the e-e-n tensor.
// R0 scratch layout: [W][Ne][cord-1][5]; gN: [W][Nn][Ne] = exp(-kappa_a |r_i - R_a|).
// ------------------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_jas_pair(JasDevView v, JasOut o, double* __restrict__ R0,
                                                     double* __restrict__ gN) {
  extern __shared__ double lds[];
  const int Ne = (int)v.elec_num, Nn = (int)v.nucl_num;
  const int64_t w = blockIdx.x;
  double* xe = lds;            // [Ne][3]
  double* red = lds + 3 * Ne;  // [kBlock]
  for (int t = threadIdx.x; t < Ne; t += kBlock)
    for (int d = 0; d < 3; ++d) xe[3 * t + d] = v.coord[w * Ne + t + d * v.coord_ld];
  __syncthreads();

  const double kee = v.rescale_ee, kinv = 1.0 / v.rescale_ee;
  const double* b = v.b_vector;
  const int bord = (int)v.bord, aord = (int)v.aord, C = (int)v.cord;
  const int up = (int)v.up_num;
  double fee = 0.0, fen = 0.0;
  for (int i = threadIdx.x; i < Ne; i += kBlock) {
    const double xi = xe[3 * i], yi = xe[3 * i + 1], zi = xe[3 * i + 2];
    double g0 = 0, g1 = 0, g2 = 0, g3 = 0;
    double r0[7][5];
#pragma unroll
    for (int k = 0; k < 7; ++k)
#pragma unroll
      for (int c = 0; c < 5; ++c) r0[k][c] = 0.0;
    for (int j = 0; j < Ne; ++j) {
      if (j == i) continue;
      const double dx = xi - xe[3 * j], dy = yi - xe[3 * j + 1], dz = zi - xe[3 * j + 2];
      const double d = sqrt(dx * dx + dy * dy + dz * dz);
      const double e = exp(-kee * d);
      const double x = (1.0 - e) * kinv;  // rescaled distance, org/qmckl_distance.org:975-985
      // value: the pair (j, i) with j < i, as in the doc loop over j > i' of the reference
      if (j < i) {
        const bool same = (v.spin_independent != 1) && (i < up || j >= up);
        if (v.spin_independent == 1)
          fee += b[0] * x / (1.0 + b[1] * x) - v.asymp_jasb1;
        else if (same)
          fee += 0.5 * b[0] * x / (1.0 + b[1] * x) - v.asymp_jasb0;
        else
          fee += b[0] * x / (1.0 + b[1] * x) - v.asymp_jasb1;
        double xk = x;
        for (int k = 2; k <= bord; ++k) {
          xk *= x;
          fee += b[k] * xk;
        }
      }
      // gradient/Laplacian w.r.t. electron i, org/qmckl_distance.org:1440-1452
      const double dd = d > 1.e-20 ? d : 1.e-20;
      const double dinv = 1.0 / dd;
      const double ed = (d > 1.e-20) ? e : exp(-kee * dd);
      const double gx = dx * dinv * ed, gy = dy * dinv * ed, gz = dz * dinv * ed;
      const double gl4 = (2.0 * dinv - kee) * ed;
      const double denom = 1.0 + b[1] * x;
      const double invd = 1.0 / denom, invd2 = invd * invd;
      const double gc2 = gx * gx + gy * gy + gz * gz;
      double f;
      if (v.spin_independent == 1)
        f = b[0] * invd2;
      else if ((i < up && j < up) || (i >= up && j >= up))
        f = 0.5 * b[0] * invd2;
      else
        f = b[0] * invd2;
      g0 += f * gx;
      g1 += f * gy;
      g2 += f * gz;
      g3 += f * (gl4 - 2.0 * b[1] * gc2 * invd);
      double kf = 2.0, xp = 1.0;
      for (int k = 2; k <= bord; ++k) {
        f = b[k] * kf * xp;
        g0 += f * x * gx;
        g1 += f * x * gy;
        g2 += f * x * gz;
        g3 += f * (x * gl4 + (kf - 1.0) * gc2);
        xp *= x;
        kf += 1.0;
      }
      // a-independent e-e-n sums (k >= 1); g_e = exp(-kappa r), diagonal excluded
      const double rinv = 1.0 / d;
      const double ux = dx * rinv, uy = dy * rinv, uz = dz * rinv;
      double gk = 1.0;
#pragma unroll
      for (int k = 1; k < 8; ++k) {
        if (k >= C) break;
        gk *= e;
        r0[k - 1][0] += gk;
        r0[k - 1][1] += gk * ux;
        r0[k - 1][2] += gk * uy;
        r0[k - 1][3] += gk * uz;
        r0[k - 1][4] += gk * rinv;
      }
    }
    double* og = o.fee_gl + w * 4 * Ne;
    og[i] = g0;
    og[i + Ne] = g1;
    og[i + 2 * Ne] = g2;
    og[i + 3 * Ne] = g3;
    if (C >= 2) {
      double* rr = R0 + ((w * Ne + i) * (C - 1)) * 5;
#pragma unroll
      for (int k = 0; k < 7; ++k) {
        if (k >= C - 1) break;
#pragma unroll
        for (int c = 0; c < 5; ++c) rr[k * 5 + c] = r0[k][c];
      }
    }

    // e-n term, org/qmckl_jastrow_champ.org:7021-7095 and :7500-7590
    double h0 = 0, h1 = 0, h2 = 0, h3 = 0;
    for (int a = 0; a < Nn; ++a) {
      const int t = (int)v.type_nucl[a];
      const double* av = v.a_vector + t * (aord + 1);
      const double ka = v.kappa_en[t];
      const double dx = xi - v.nucl_coord[a], dy = yi - v.nucl_coord[a + Nn], dz = zi - v.nucl_coord[a + 2 * Nn];
      const double d = sqrt(dx * dx + dy * dy + dz * dz);
      const double e = exp(-ka * d);
      gN[(w * Nn + a) * Ne + i] = e;  // g_a(r_i), reused by the e-e-n kernel
      const double x = (1.0 - e) * (1.0 / ka);
      fen += av[0] * x / (1.0 + av[1] * x) - v.asymp_jasa[t];
      double xk = x;
      for (int p = 2; p <= aord; ++p) {
        xk *= x;
        fen += av[p] * xk;
      }
      const double dd = d > 1.e-20 ? d : 1.e-20;
      const double dinv = 1.0 / dd;
      const double ed = (d > 1.e-20) ? e : exp(-ka * dd);
      const double gx = dx * dinv * ed, gy = dy * dinv * ed, gz = dz * dinv * ed;
      const double gl4 = (2.0 * dinv - ka) * ed;
      const double denom = 1.0 + av[1] * x;
      const double invd = 1.0 / denom, invd2 = invd * invd;
      double f = av[0] * invd2;
      const double gc2 = gx * gx + gy * gy + gz * gz;
      h0 += f * gx;
      h1 += f * gy;
      h2 += f * gz;
      h3 += f * (gl4 - 2.0 * av[1] * gc2 * invd);
      double kf = 2.0, xp = 1.0;
      for (int k = 2; k <= aord; ++k) {
        f = av[k] * kf * xp;
        h0 += f * x * gx;
        h1 += f * x * gy;
        h2 += f * x * gz;
        h3 += f * (x * gl4 + (kf - 1.0) * gc2);
        xp *= x;
        kf += 1.0;
      }
    }
    double* oh = o.fen_gl + w * 4 * Ne;
    oh[i] = h0;
    oh[i + Ne] = h1;
    oh[i + 2 * Ne] = h2;
    oh[i + 3 * Ne] = h3;
  }
  const double see = block_sum(fee, red);
  const double sen = block_sum(fen, red);
  if (threadIdx.x == 0) {
    o.fee[w] = see;
    o.fen[w] = sen;
  }
}

// ------------------------------------------------------------------------------------
// e-e-n term.  Tile: TI electrons x all nuclei; j walked in chunks of JC.
// Jt scratch: [W][ntiles] partial J_eeN.
// ------------------------------------------------------------------------------------
template <int C>
__global__ void __launch_bounds__(kBlock) k_jas_een(JasDevView v, JasOut o, const double* __restrict__ R0,
                                                    const double* __restrict__ gN, double* __restrict__ Jt, int TI,
                                                    int JC, int ntiles) {
  constexpr int NQ = C;  // q = 1..C
  constexpr Lkm<C> T{};
  extern __shared__ double lds[];
  const int Ne = (int)v.elec_num, Nn = (int)v.nucl_num;
  const int64_t w = blockIdx.x / ntiles;
  const int tile = blockIdx.x - (int)(w * ntiles);
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
        gn[aa * JC + jj] = gN[(w * Nn + aa) * Ne + j0 + jj];
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

__global__ void __launch_bounds__(kBlock) k_jas_total(JasDevView v, JasOut o, const double* __restrict__ Jt,
                                                      int ntiles, int have_een) {
  __shared__ double sval;
  const int Ne = (int)v.elec_num;
  const int64_t w = blockIdx.x;
  if (threadIdx.x == 0) {
    double een = 0.0;
    if (have_een)
      for (int t = 0; t < ntiles; ++t) een += Jt[w * ntiles + t];
    o.feen[w] = een;
    const double J = o.fee[w] + o.fen[w] + een;
    const double val = (J < -100.0) ? 0.0 : exp(J);
    o.value[w] = val;
    sval = val;
  }
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

struct EenShape {
  int TI, JC, ntiles;
  size_t lds;
};

EenShape een_shape(const JasDevView& v) {
  EenShape s;
  const int Ne = (int)v.elec_num, Nn = (int)v.nucl_num;
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

__global__ void k_en_distance(const double* __restrict__ coord, int64_t P, const double* __restrict__ nc, int64_t Nn,
                              double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P * Nn) return;
  const int64_t a = t % Nn, p = t / Nn;
  const double dx = coord[p] - nc[a], dy = coord[p + P] - nc[a + Nn], dz = coord[p + 2 * P] - nc[a + 2 * Nn];
  out[t] = sqrt(dx * dx + dy * dy + dz * dz);
}

}  // namespace

int64_t jastrow_scratch_doubles(const JasDevView& v) {
  const int64_t C = v.cord;
  const int64_t r0 = C >= 2 ? v.walk_num * v.elec_num * (C - 1) * 5 : 0;
  const EenShape s = een_shape(v);
  return r0 + v.walk_num * v.nucl_num * v.elec_num + v.walk_num * s.ntiles;
}

struct ScratchView {
  double *R0, *gN, *Jt;
};

ScratchView scratch_view(const JasDevView& v, const JasOut& o) {
  const int64_t C = v.cord;
  const int64_t r0n = C >= 2 ? v.walk_num * v.elec_num * (C - 1) * 5 : 0;
  ScratchView sv;
  sv.R0 = o.scratch;
  sv.gN = o.scratch + r0n;
  sv.Jt = sv.gN + v.walk_num * v.nucl_num * v.elec_num;
  return sv;
}

hipError_t launch_jastrow_pair(const JasDevView& v, const JasOut& o, hipStream_t s) {
  const ScratchView sv = scratch_view(v, o);
  const size_t lds_pair = sizeof(double) * (3 * v.elec_num + kBlock);
  if (lds_pair > 160 * 1024) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_jas_pair, dim3((unsigned)v.walk_num), dim3(kBlock), lds_pair, s, v, o, sv.R0, sv.gN);
  return hipGetLastError();
}

hipError_t launch_jastrow_een(const JasDevView& v, const JasOut& o, hipStream_t s) {
  const ScratchView sv = scratch_view(v, o);
  const int C = (int)v.cord;
  if (!(C >= 2 && v.dim_c > 0)) return hipSuccess;
  const EenShape sh = een_shape(v);
  if (sh.lds > 160 * 1024) return hipErrorInvalidValue;
  const dim3 grid((unsigned)(v.walk_num * sh.ntiles));
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
  hipLaunchKernelGGL(k_jas_total, dim3((unsigned)v.walk_num), dim3(kBlock), 0, s, v, o, sv.Jt, sh.ntiles, have_een);
  return hipGetLastError();
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

hipError_t launch_en_distance(const double* coord, int64_t P, const double* nucl_coord, int64_t Nn, double* out,
                              hipStream_t s) {
  const int64_t n = P * Nn;
  hipLaunchKernelGGL(k_en_distance, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, coord, P, nucl_coord, Nn, out);
  return hipGetLastError();
}

}  // namespace qmckl_amd
