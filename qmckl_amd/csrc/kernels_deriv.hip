// Jastrow parameter derivatives (org/qmckl_jastrow_champ.org) and Jastrow nuclear forces
// (org/qmckl_forces.org) on gfx950, FP64.
//
// Inputs are the reference-layout intermediates materialised by k_jas_inter (rescaled
// distances, een_rescaled_e/_n with their GL, tmp_c, dtmp_c) plus the e-n distances.  Each
// kernel computes one output element per thread, with the sums the reference's doc routine
// accumulates into that element; quantities the reference reduces over walkers are written
// per walker first and summed over walkers in a second pass (k_sum_walkers), in walker order,
// so results are run-to-run reproducible.  The arithmetic follows the reference's, including
// three places where it differs from the exact derivative (documented with the oracle):
//   factor_ee_pderiv, B0 of opposite-spin pairs: asymp_jasb_pderiv(2,1) (:5259);
//   factor_ee_gl_pderiv, Laplacian of the B0 derivative: +2 B1 |grad x|^2 f1/(1 + B1 x) (:5584);
//   factor_en_pderiv, A1: a_vector(2) in place of a_vector(1) (:8502).
//
// Layouts (C row-major):
//   ee  [W][Ne(j)][Ne(i)]        eeg  [W][Ne(j)][Ne(i)][4]
//   en  [W][Nn][Ne]              eng  [W][Nn][Ne][4]          edist [W*Ne][Nn]
//   een_e [W][C+1][Ne(j)][Ne(i)] een_e_gl [W][C+1][Ne(j)][4][Ne(i)]
//   een_n [W][C+1][Nn][Ne]       een_n_gl [W][C+1][Nn][4][Ne]
//   tmp_c [W][C][C+1][Nn][Ne]    dtmp_c [W][C][C+1][Nn][4][Ne]
//   ftmp_c  [W][C][C+1][Nn][4][Ne]          (forces_tmp_c)
//   fdtmp_c [W][C][C+1][4(k)][Nn][4(kk)][Ne] (forces_dtmp_c)
//   feen_n  [W][C+1][3(n)][Nn][4(m)][Ne]    (forces_een_rescaled_n_gl)
#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace qmckl_amd {

namespace {

struct Idx {
  int64_t W, Ne, Nn, C;
  __device__ int64_t ee(int64_t i, int64_t j, int64_t w) const { return (w * Ne + j) * Ne + i; }
  __device__ int64_t eeg(int c, int64_t i, int64_t j, int64_t w) const { return ((w * Ne + j) * Ne + i) * 4 + c; }
  __device__ int64_t en(int64_t i, int64_t a, int64_t w) const { return (w * Nn + a) * Ne + i; }
  __device__ int64_t eng(int c, int64_t i, int64_t a, int64_t w) const { return ((w * Nn + a) * Ne + i) * 4 + c; }
  __device__ int64_t edist(int64_t a, int64_t i, int64_t w) const { return (w * Ne + i) * Nn + a; }
  __device__ int64_t eene(int64_t i, int64_t j, int64_t l, int64_t w) const { return ((w * (C + 1) + l) * Ne + j) * Ne + i; }
  __device__ int64_t eeneg(int64_t i, int c, int64_t j, int64_t l, int64_t w) const {
    return (((w * (C + 1) + l) * Ne + j) * 4 + c) * Ne + i;
  }
  __device__ int64_t eenn(int64_t i, int64_t a, int64_t q, int64_t w) const { return ((w * (C + 1) + q) * Nn + a) * Ne + i; }
  __device__ int64_t eenng(int64_t i, int c, int64_t a, int64_t q, int64_t w) const {
    return (((w * (C + 1) + q) * Nn + a) * 4 + c) * Ne + i;
  }
  __device__ int64_t tmpc(int64_t i, int64_t a, int64_t q, int64_t k, int64_t w) const {
    return (((w * C + k) * (C + 1) + q) * Nn + a) * Ne + i;
  }
  __device__ int64_t dtmpc(int64_t i, int c, int64_t a, int64_t q, int64_t k, int64_t w) const {
    return ((((w * C + k) * (C + 1) + q) * Nn + a) * 4 + c) * Ne + i;
  }
  __device__ int64_t fdtmpc(int64_t i, int kk, int64_t a, int k, int64_t q, int64_t l, int64_t w) const {
    return (((((w * C + l) * (C + 1) + q) * 4 + k) * Nn + a) * 4 + kk) * Ne + i;
  }
  __device__ int64_t feenn(int64_t i, int m, int64_t a, int n, int64_t l, int64_t w) const {
    return ((((w * (C + 1) + l) * 3 + n) * Nn + a) * 4 + m) * Ne + i;
  }
};

__device__ __forceinline__ int64_t gtid() { return (int64_t)blockIdx.x * blockDim.x + threadIdx.x; }

// ---- parameter derivatives, per walker -------------------------------------------------
// factor_ee_pderiv, :5194-5273.  part [W][bord+1]; ajb [2][bord+1] = asymp_jasb_pderiv(p+1, s+1).
__global__ void k_ee_pderiv(Idx X, int64_t up, int64_t bord, const double* __restrict__ b, const double* __restrict__ ee,
                            const double* __restrict__ ajb, int32_t spin_indep, double* __restrict__ part) {
  const int64_t n = bord + 1, t = gtid();
  if (t >= X.W * n) return;
  const int64_t p = t % n, w = t / n;
  double acc = 0.0;
  for (int64_t j = 0; j < X.Ne; ++j)
    for (int64_t i = 0; i < j; ++i) {
      const double x = ee[X.ee(i, j, w)];
      const double den = 1.0 + b[1] * x;
      if (p == 0) {
        if (spin_indep == 1)
          acc += x / den - ajb[n];
        else if (j < up || i >= up)
          acc += 0.5 * x / den - ajb[0];
        else
          acc += x / den - ajb[1];
      } else if (p == 1) {
        if (spin_indep == 1)
          acc += -b[0] * x * x / (den * den) - ajb[n + 1];
        else if (j < up || i >= up)
          acc += -0.5 * b[0] * x * x / (den * den) - ajb[1];
        else
          acc += -b[0] * x * x / (den * den) - ajb[n + 1];
      } else {
        double xk = x;
        for (int64_t k = 2; k <= p; ++k) xk *= x;
        acc += xk - ajb[p];
      }
    }
  part[t] = acc;
}

// factor_ee_gl_pderiv, :5494-5605.  part [W][bord+1][Ne][4].
__global__ void k_ee_gl_pderiv(Idx X, int64_t up, int64_t bord, const double* __restrict__ b,
                               const double* __restrict__ ee, const double* __restrict__ eeg, int32_t spin_indep,
                               double* __restrict__ part) {
  const int64_t n = bord + 1, t = gtid();
  if (t >= X.W * n * X.Ne * 4) return;
  int64_t r = t;
  const int c = (int)(r % 4); r /= 4;
  const int64_t i = r % X.Ne; r /= X.Ne;
  const int64_t p = r % n;
  const int64_t w = r / n;
  double acc = 0.0;
  for (int64_t j = 0; j < X.Ne; ++j) {
    if (i == j) continue;
    const double x = ee[X.ee(i, j, w)];
    const double inv = 1.0 / (1.0 + b[1] * x), inv2 = inv * inv, inv3 = inv2 * inv;
    double dx[4];
    for (int cc = 0; cc < 4; ++cc) dx[cc] = eeg[X.eeg(cc, i, j, w)];
    const double gc2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
    double f1, f2;
    if (spin_indep == 1) {
      f1 = inv2;
      f2 = -2.0 * b[0] * x * inv3;
    } else if ((i < up && j < up) || (i >= up && j >= up)) {
      f1 = 0.5 * inv2;
      f2 = -1.0 * b[0] * x * inv3;
    } else {
      f1 = inv2;
      f2 = -2.0 * b[0] * x * inv3;
    }
    if (p == 0) {
      acc += (c < 3) ? f1 * dx[c] : f1 * dx[3] + 2.0 * b[1] * gc2 * f1 * inv;
    } else if (p == 1) {
      acc += (c < 3) ? f2 * dx[c] : 2.0 * f1 * b[0] * (-inv * x * dx[3] - inv2 * gc2 * (1 - 2.0 * x * b[1]));
    } else {
      double xp = 1.0;  // x^(p-2)
      for (int64_t k = 2; k < p; ++k) xp *= x;
      const double kf = (double)p, f = kf * xp;
      acc += (c < 3) ? f * x * dx[c] : f * (x * dx[3] + (kf - 1.0) * gc2);
    }
  }
  part[t] = acc;
}

// factor_en_pderiv, :8436-8520.  part [W][ntype][aord+1]; aja [ntype][aord+1].
__global__ void k_en_pderiv(Idx X, int64_t ntype, int64_t aord, const int64_t* __restrict__ type,
                            const double* __restrict__ a, const double* __restrict__ en, const double* __restrict__ aja,
                            double* __restrict__ part) {
  const int64_t n = aord + 1, t = gtid();
  if (t >= X.W * ntype * n) return;
  const int64_t p = t % n, ty = (t / n) % ntype, w = t / (n * ntype);
  const double* av = a + ty * n;
  double acc = 0.0;
  for (int64_t at = 0; at < X.Nn; ++at) {
    if (type[at] != ty) continue;
    for (int64_t i = 0; i < X.Ne; ++i) {
      const double x = en[X.en(i, at, w)];
      const double den = 1.0 + av[1] * x;
      if (p == 0)
        acc += x / den - aja[ty * n];
      else if (p == 1)
        acc += -av[1] * x * x / (den * den) - aja[ty * n + 1];
      else {
        double xk = x;
        for (int64_t k = 2; k <= p; ++k) xk *= x;
        acc += xk - aja[ty * n + p];
      }
    }
  }
  part[t] = acc;
}

// factor_en_gl_pderiv, :8735-8825.  part [W][ntype][aord+1][Ne][4].
__global__ void k_en_gl_pderiv(Idx X, int64_t ntype, int64_t aord, const int64_t* __restrict__ type,
                               const double* __restrict__ a, const double* __restrict__ en,
                               const double* __restrict__ eng, double* __restrict__ part) {
  const int64_t n = aord + 1, t = gtid();
  if (t >= X.W * ntype * n * X.Ne * 4) return;
  int64_t r = t;
  const int c = (int)(r % 4); r /= 4;
  const int64_t i = r % X.Ne; r /= X.Ne;
  const int64_t p = r % n; r /= n;
  const int64_t ty = r % ntype;
  const int64_t w = r / ntype;
  const double* av = a + ty * n;
  double acc = 0.0;
  for (int64_t at = 0; at < X.Nn; ++at) {
    if (type[at] != ty) continue;
    const double x = en[X.en(i, at, w)];
    const double inv = 1.0 / (1.0 + av[1] * x), inv2 = inv * inv, inv3 = inv2 * inv;
    double dx[4];
    for (int cc = 0; cc < 4; ++cc) dx[cc] = eng[X.eng(cc, i, at, w)];
    const double gc2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
    if (p == 0) {
      acc += (c < 3) ? inv2 * dx[c] : inv2 * (dx[3] - 2.0 * av[1] * gc2 * inv);
    } else if (p == 1) {
      const double f = -2.0 * av[0] * inv3;
      acc += (c < 3) ? f * x * dx[c] : f * (x * dx[3] + inv * (gc2 - 2 * gc2 * x * av[1]));
    } else {
      double xp = 1.0;
      for (int64_t k = 2; k < p; ++k) xp *= x;
      const double kf = (double)p, f = kf * xp;
      acc += (c < 3) ? f * x * dx[c] : f * (x * dx[3] + (kf - 1.0) * gc2);
    }
  }
  part[t] = acc;
}

// factor_een_pderiv, :15931-15975.  part [W][ntype][dim]; lkpm [4][dim].
__global__ void k_een_pderiv(Idx X, int64_t ntype, int64_t dim, const int64_t* __restrict__ type,
                             const int64_t* __restrict__ lkpm, const double* __restrict__ tmp_c,
                             const double* __restrict__ een_n, double* __restrict__ part) {
  const int64_t t = gtid();
  if (t >= X.W * ntype * dim) return;
  const int64_t n = t % dim, ty = (t / dim) % ntype, w = t / (dim * ntype);
  const int64_t l = lkpm[n], k = lkpm[dim + n], m = lkpm[3 * dim + n];
  double acc = 0.0;
  for (int64_t at = 0; at < X.Nn; ++at) {
    if (type[at] != ty) continue;
    double s = 0.0;
    for (int64_t j = 0; j < X.Ne; ++j) s += een_n[X.eenn(j, at, m, w)] * tmp_c[X.tmpc(j, at, m + l, k, w)];
    acc += s;
  }
  part[t] = acc;
}

// factor_een_gl_pderiv, :16255-16305.  part [W][ntype][dim][Ne][4].
__global__ void k_een_gl_pderiv(Idx X, int64_t ntype, int64_t dim, const int64_t* __restrict__ type,
                                const int64_t* __restrict__ lkpm, const double* __restrict__ tmp_c,
                                const double* __restrict__ dtmp_c, const double* __restrict__ een_n,
                                const double* __restrict__ een_n_gl, double* __restrict__ part) {
  const int64_t t = gtid();
  if (t >= X.W * ntype * dim * X.Ne * 4) return;
  int64_t r = t;
  const int c = (int)(r % 4); r /= 4;
  const int64_t j = r % X.Ne; r /= X.Ne;
  const int64_t n = r % dim; r /= dim;
  const int64_t ty = r % ntype;
  const int64_t w = r / ntype;
  const int64_t l = lkpm[n], k = lkpm[dim + n], m = lkpm[3 * dim + n];
  double acc = 0.0;
  for (int64_t at = 0; at < X.Nn; ++at) {
    if (type[at] != ty) continue;
    acc += tmp_c[X.tmpc(j, at, m, k, w)] * een_n_gl[X.eenng(j, c, at, m + l, w)] +
           tmp_c[X.tmpc(j, at, m + l, k, w)] * een_n_gl[X.eenng(j, c, at, m, w)] +
           dtmp_c[X.dtmpc(j, c, at, m, k, w)] * een_n[X.eenn(j, at, m + l, w)] +
           dtmp_c[X.dtmpc(j, c, at, m + l, k, w)] * een_n[X.eenn(j, at, m, w)];
    if (c == 3) {
      double s = 0.0;
      for (int cc = 0; cc < 3; ++cc)
        s += dtmp_c[X.dtmpc(j, cc, at, m, k, w)] * een_n_gl[X.eenng(j, cc, at, m + l, w)] +
             dtmp_c[X.dtmpc(j, cc, at, m + l, k, w)] * een_n_gl[X.eenng(j, cc, at, m, w)];
      acc += s * 2;
    }
  }
  part[t] = acc;
}

// out[e] = scale * sum_w in[w][e], walkers in order.
__global__ void k_sum_walkers(const double* __restrict__ in, int64_t W, int64_t n, double scale,
                              double* __restrict__ out) {
  const int64_t e = gtid();
  if (e >= n) return;
  double s = 0.0;
  for (int64_t w = 0; w < W; ++w) s += in[w * n + e];
  out[e] = s * scale;
}

// ---- forces of the e-n term, org/qmckl_forces.org:732-1620 --------------------------------
// forces_jastrow_en [W][Nn][3].
__global__ void k_forces_en(Idx X, int64_t aord, const int64_t* __restrict__ type, const double* __restrict__ a,
                            const double* __restrict__ en, const double* __restrict__ eng, double* __restrict__ out) {
  const int64_t t = gtid();
  if (t >= X.W * X.Nn * 3) return;
  const int c = (int)(t % 3);
  const int64_t at = (t / 3) % X.Nn, w = t / (3 * X.Nn);
  const double* av = a + type[at] * (aord + 1);
  double acc = 0.0;
  for (int64_t i = 0; i < X.Ne; ++i) {
    const double x1 = en[X.en(i, at, w)];
    const double inv = 1.0 / (1.0 + av[1] * x1);
    const double dx = -eng[X.eng(c, i, at, w)];
    acc += av[0] * inv * inv * dx;
    double kf = 2.0, x = 1.0;
    for (int64_t k = 2; k <= aord; ++k) {
      acc += av[k] * kf * x * x1 * dx;
      x *= x1;
      kf += 1.0;
    }
  }
  out[t] = acc;
}

// forces_jastrow_en_g [W][Nn][3(l)][Ne][3(m)].
__global__ void k_forces_en_g(Idx X, int64_t aord, const int64_t* __restrict__ type, const double* __restrict__ a,
                              const double* __restrict__ kappa, const double* __restrict__ edist,
                              const double* __restrict__ en, const double* __restrict__ eng, double* __restrict__ out) {
  const int64_t t = gtid();
  if (t >= X.W * X.Nn * 3 * X.Ne * 3) return;
  int64_t r = t;
  const int m = (int)(r % 3); r /= 3;
  const int64_t i = r % X.Ne; r /= X.Ne;
  const int l = (int)(r % 3); r /= 3;
  const int64_t at = r % X.Nn;
  const int64_t w = r / X.Nn;
  const int64_t ty = type[at];
  const double* av = a + ty * (aord + 1);
  const double kap = kappa[ty];
  const double dist = edist[X.edist(at, i, w)];
  const double expk = exp(kap * dist), invdist = 1.0 / dist;
  const double x1 = en[X.en(i, at, w)];
  const double inv = 1.0 / (1.0 + av[1] * x1), inv2 = inv * inv;
  const double dm = eng[X.eng(m, i, at, w)], dl = eng[X.eng(l, i, at, w)];
  double f = av[0] * inv2;
  double acc = 0.0;
  if (m == l) acc -= f * invdist / expk;
  acc += f * dm * dl * expk * (invdist + kap);
  acc += 2.0 * f * inv * av[1] * dm * dl;
  double kf = 2.0, x = 1.0;
  for (int64_t k = 2; k <= aord; ++k) {
    f = av[k] * kf * x;
    const double f2 = av[k] * kf * x * (kf - 1.0);
    if (m == l) acc -= f * x1 * invdist / expk;
    acc += -f2 * dm * dl + f * x1 * dm * dl * kap * expk + f * x1 * dm * dl * invdist * expk;
    x *= x1;
    kf += 1.0;
  }
  out[t] = acc;
}

// forces_jastrow_en_l [W][Nn][3].
__global__ void k_forces_en_l(Idx X, int64_t aord, const int64_t* __restrict__ type, const double* __restrict__ a,
                              const double* __restrict__ kappa, const double* __restrict__ edist,
                              const double* __restrict__ en, const double* __restrict__ eng, double* __restrict__ out) {
  const int64_t t = gtid();
  if (t >= X.W * X.Nn * 3) return;
  const int m = (int)(t % 3);
  const int64_t at = (t / 3) % X.Nn, w = t / (3 * X.Nn);
  const int64_t ty = type[at];
  const double* av = a + ty * (aord + 1);
  const double kap = kappa[ty];
  double acc = 0.0;
  for (int64_t i = 0; i < X.Ne; ++i) {
    const double dist = edist[X.edist(at, i, w)];
    const double expk = exp(kap * dist), invdist = 1.0 / dist;
    const double x1 = en[X.en(i, at, w)];
    const double inv = 1.0 / (1.0 + av[1] * x1), inv2 = inv * inv;
    double f = av[0] * inv2;
    double dx[4];
    for (int c = 0; c < 4; ++c) dx[c] = eng[X.eng(c, i, at, w)];
    const double gc2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
    acc += f * dx[m] * (kap * (2 * invdist - kap) + 2 * invdist * invdist);
    acc += 2.0 * f * inv * av[1] * dx[m] * dx[3];
    acc -= 4.0 * f * inv * av[1] * kap * dx[m] / expk;
    acc -= 6.0 * f * inv * inv * av[1] * av[1] * gc2 * dx[m];
    double kf = 2.0, x = 1.0;
    for (int64_t k = 2; k <= aord; ++k) {
      f = av[k] * kf * x;
      acc += -f * dx[m] * dx[3] * (kf - 1.0) - f / x1 * (kf - 1.0) * (kf - 2.0) * dx[m] / expk / expk +
             f * x1 * kap * dx[m] * dx[3] * expk + f * x1 * 2 * dx[m] * invdist * invdist +
             2 * f * (kf - 1.0) * dx[m] * kap / expk;
      x *= x1;
      kf += 1.0;
    }
  }
  out[t] = acc;
}

// ---- forces of the e-e-n term, org/qmckl_forces.org:2352-4345 -----------------------------
// forces_tmp_c [W][C][C+1][Nn][4][Ne]: -een_rescaled_e(:,:,k) . een_rescaled_n_gl(:,:,:,q),
// q = 1..C-k (:2352-2385); 0 elsewhere.
__global__ void k_forces_tmp_c(Idx X, const double* __restrict__ een_e, const double* __restrict__ een_n_gl,
                               double* __restrict__ out) {
  const int64_t t = gtid();
  const int64_t C = X.C;
  if (t >= X.W * C * (C + 1) * X.Nn * 4 * X.Ne) return;
  int64_t r = t;
  const int64_t i = r % X.Ne; r /= X.Ne;
  const int c = (int)(r % 4); r /= 4;
  const int64_t at = r % X.Nn; r /= X.Nn;
  const int64_t q = r % (C + 1); r /= (C + 1);
  const int64_t k = r % C;
  const int64_t w = r / C;
  double s = 0.0;
  if (q >= 1 && q <= C - k)
    for (int64_t m = 0; m < X.Ne; ++m) s += een_e[X.eene(i, m, k, w)] * een_n_gl[X.eenng(m, c, at, q, w)];
  out[t] = -s;
}

// forces_dtmp_c [W][C][C+1][4(k)][Nn][4(kk)][Ne] (:2688-2725).
__global__ void k_forces_dtmp_c(Idx X, const double* __restrict__ een_e_gl, const double* __restrict__ een_n_gl,
                                double* __restrict__ out) {
  const int64_t t = gtid();
  const int64_t C = X.C;
  if (t >= X.W * C * (C + 1) * 4 * X.Nn * 4 * X.Ne) return;
  int64_t r = t;
  const int64_t i = r % X.Ne; r /= X.Ne;
  const int kk = (int)(r % 4); r /= 4;
  const int64_t at = r % X.Nn; r /= X.Nn;
  const int k = (int)(r % 4); r /= 4;
  const int64_t m = r % (C + 1); r /= (C + 1);
  const int64_t l = r % C;
  const int64_t w = r / C;
  double s = 0.0;
  if (m >= 1 && l + m <= C)
    for (int64_t j = 0; j < X.Ne; ++j) s -= een_e_gl[X.eeneg(i, kk, j, l, w)] * een_n_gl[X.eenng(j, k, at, m, w)];
  out[t] = s;
}

// forces_een_rescaled_n_gl [W][C+1][3(n)][Nn][4(m)][Ne] (:3441-3520).
__global__ void k_forces_een_n(Idx X, const int64_t* __restrict__ type, const double* __restrict__ kappa,
                               const double* __restrict__ edist, const double* __restrict__ een_n,
                               const double* __restrict__ een_n_gl, double* __restrict__ out) {
  const int64_t t = gtid();
  const int64_t C = X.C;
  if (t >= X.W * (C + 1) * 3 * X.Nn * 4 * X.Ne) return;
  int64_t r = t;
  const int64_t i = r % X.Ne; r /= X.Ne;
  const int m = (int)(r % 4); r /= 4;
  const int64_t at = r % X.Nn; r /= X.Nn;
  const int n = (int)(r % 3); r /= 3;
  const int64_t l = r % (C + 1);
  const int64_t w = r / (C + 1);
  double v = 0.0;
  if (l >= 1) {
    const double kl = (double)l * kappa[type[at]];
    const double d = edist[X.edist(at, i, w)];
    const double g = een_n[X.eenn(i, at, l, w)];
    const double gm = een_n_gl[X.eenng(i, m, at, l, w)], gn = een_n_gl[X.eenng(i, n, at, l, w)];
    if (m == n) v += kl * g / d;
    if (m < 3)
      v -= gm * gn / (kl * g * d) + gm * gn / g;
    else
      v += 2.0 * gn / (d * d) - gm * gn / g;
  }
  out[t] = v;
}

// forces_jastrow_een [W][Nn][3] (:3168-3210).
__global__ void k_forces_een(Idx X, int64_t dim, const double* __restrict__ cfull, const int64_t* __restrict__ lkpm,
                             const double* __restrict__ een_n, const double* __restrict__ een_n_gl,
                             const double* __restrict__ tmp_c, const double* __restrict__ ftmp_c,
                             double* __restrict__ out) {
  const int64_t t = gtid();
  if (t >= X.W * X.Nn * 3) return;
  const int ii = (int)(t % 3);
  const int64_t at = (t / 3) % X.Nn, w = t / (3 * X.Nn);
  double acc = 0.0;
  for (int64_t n = 0; n < dim; ++n) {
    const int64_t l = lkpm[n], k = lkpm[dim + n], m = lkpm[3 * dim + n];
    const double cn = cfull[n * X.Nn + at];
    if (cn == 0.0) continue;
    for (int64_t j = 0; j < X.Ne; ++j)
      acc += (een_n[X.eenn(j, at, m, w)] * ftmp_c[X.dtmpc(j, ii, at, m + l, k, w)] -
              een_n_gl[X.eenng(j, ii, at, m, w)] * tmp_c[X.tmpc(j, at, m + l, k, w)]) *
             cn;
  }
  out[t] = acc;
}

// forces_jastrow_een_g [W][3(jj)][Nn][3(ii)][Ne] (:3918-3965).
__global__ void k_forces_een_g(Idx X, int64_t dim, const double* __restrict__ cfull, const int64_t* __restrict__ lkpm,
                               const double* __restrict__ tmp_c, const double* __restrict__ dtmp_c,
                               const double* __restrict__ ftmp_c, const double* __restrict__ fdtmp_c,
                               const double* __restrict__ feen_n, const double* __restrict__ een_n,
                               const double* __restrict__ een_n_gl, double* __restrict__ out) {
  const int64_t t = gtid();
  if (t >= X.W * 3 * X.Nn * 3 * X.Ne) return;
  int64_t r = t;
  const int64_t i = r % X.Ne; r /= X.Ne;
  const int ii = (int)(r % 3); r /= 3;
  const int64_t at = r % X.Nn; r /= X.Nn;
  const int jj = (int)(r % 3);
  const int64_t w = r / 3;
  double acc = 0.0;
  for (int64_t n = 0; n < dim; ++n) {
    const int64_t l = lkpm[n], k = lkpm[dim + n], m = lkpm[3 * dim + n];
    const double cn = cfull[n * X.Nn + at];
    if (cn == 0.0) continue;
    acc += (tmp_c[X.tmpc(i, at, m, k, w)] * feen_n[X.feenn(i, ii, at, jj, m + l, w)] +
            ftmp_c[X.dtmpc(i, jj, at, m, k, w)] * een_n_gl[X.eenng(i, ii, at, m + l, w)] +
            tmp_c[X.tmpc(i, at, m + l, k, w)] * feen_n[X.feenn(i, ii, at, jj, m, w)] +
            ftmp_c[X.dtmpc(i, jj, at, m + l, k, w)] * een_n_gl[X.eenng(i, ii, at, m, w)] -
            dtmp_c[X.dtmpc(i, ii, at, m, k, w)] * een_n_gl[X.eenng(i, jj, at, m + l, w)] +
            fdtmp_c[X.fdtmpc(i, ii, at, jj, m, k, w)] * een_n[X.eenn(i, at, m + l, w)] -
            dtmp_c[X.dtmpc(i, ii, at, m + l, k, w)] * een_n_gl[X.eenng(i, jj, at, m, w)] +
            fdtmp_c[X.fdtmpc(i, ii, at, jj, m + l, k, w)] * een_n[X.eenn(i, at, m, w)]) *
           cn;
  }
  out[t] = acc;
}

// forces_jastrow_een_l [W][3][Nn] (:4282-4345).
__global__ void k_forces_een_l(Idx X, int64_t dim, const double* __restrict__ cfull, const int64_t* __restrict__ lkpm,
                               const double* __restrict__ tmp_c, const double* __restrict__ dtmp_c,
                               const double* __restrict__ ftmp_c, const double* __restrict__ fdtmp_c,
                               const double* __restrict__ feen_n, const double* __restrict__ een_n,
                               const double* __restrict__ een_n_gl, double* __restrict__ out) {
  const int64_t t = gtid();
  if (t >= X.W * 3 * X.Nn) return;
  const int64_t at = t % X.Nn;
  const int ii = (int)((t / X.Nn) % 3);
  const int64_t w = t / (3 * X.Nn);
  double acc = 0.0;
  for (int64_t n = 0; n < dim; ++n) {
    const int64_t l = lkpm[n], k = lkpm[dim + n], m = lkpm[3 * dim + n];
    double cn = cfull[n * X.Nn + at];
    if (cn == 0.0) continue;
    for (int64_t i = 0; i < X.Ne; ++i)
      acc += (tmp_c[X.tmpc(i, at, m, k, w)] * feen_n[X.feenn(i, 3, at, ii, m + l, w)] +
              ftmp_c[X.dtmpc(i, ii, at, m, k, w)] * een_n_gl[X.eenng(i, 3, at, m + l, w)] +
              tmp_c[X.tmpc(i, at, m + l, k, w)] * feen_n[X.feenn(i, 3, at, ii, m, w)] +
              ftmp_c[X.dtmpc(i, ii, at, m + l, k, w)] * een_n_gl[X.eenng(i, 3, at, m, w)] -
              dtmp_c[X.dtmpc(i, 3, at, m, k, w)] * een_n_gl[X.eenng(i, ii, at, m + l, w)] +
              fdtmp_c[X.fdtmpc(i, 3, at, ii, m, k, w)] * een_n[X.eenn(i, at, m + l, w)] -
              dtmp_c[X.dtmpc(i, 3, at, m + l, k, w)] * een_n_gl[X.eenng(i, ii, at, m, w)] +
              fdtmp_c[X.fdtmpc(i, 3, at, ii, m + l, k, w)] * een_n[X.eenn(i, at, m, w)]) *
             cn;
    cn = cn + cn;
    for (int64_t i = 0; i < X.Ne; ++i)
      for (int jj = 0; jj < 3; ++jj)
        acc += (dtmp_c[X.dtmpc(i, jj, at, m, k, w)] * feen_n[X.feenn(i, jj, at, ii, m + l, w)] +
                dtmp_c[X.dtmpc(i, jj, at, m + l, k, w)] * feen_n[X.feenn(i, jj, at, ii, m, w)] +
                fdtmp_c[X.fdtmpc(i, jj, at, ii, m, k, w)] * een_n_gl[X.eenng(i, jj, at, m + l, w)] +
                fdtmp_c[X.fdtmpc(i, jj, at, ii, m + l, k, w)] * een_n_gl[X.eenng(i, jj, at, m, w)]) *
               cn;
  }
  out[t] = acc;
}

inline dim3 grid_for(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace

hipError_t launch_deriv(const DerivArgs& d, int what, double* out, hipStream_t s) {
  const Idx X{d.W, d.Ne, d.Nn, d.C};
  const dim3 B(256);
  int64_t n = 0;
  switch (what) {
    case DV_EE_PDERIV:
      n = d.W * (d.bord + 1);
      hipLaunchKernelGGL(k_ee_pderiv, grid_for(n), B, 0, s, X, d.up, d.bord, d.b, d.ee, d.ajb, d.spin_indep, out);
      break;
    case DV_EE_GL_PDERIV:
      n = d.W * (d.bord + 1) * d.Ne * 4;
      hipLaunchKernelGGL(k_ee_gl_pderiv, grid_for(n), B, 0, s, X, d.up, d.bord, d.b, d.ee, d.eeg, d.spin_indep, out);
      break;
    case DV_EN_PDERIV:
      n = d.W * d.ntype * (d.aord + 1);
      hipLaunchKernelGGL(k_en_pderiv, grid_for(n), B, 0, s, X, d.ntype, d.aord, d.type, d.a, d.en, d.aja, out);
      break;
    case DV_EN_GL_PDERIV:
      n = d.W * d.ntype * (d.aord + 1) * d.Ne * 4;
      hipLaunchKernelGGL(k_en_gl_pderiv, grid_for(n), B, 0, s, X, d.ntype, d.aord, d.type, d.a, d.en, d.eng, out);
      break;
    case DV_EEN_PDERIV:
      n = d.W * d.ntype * d.dim;
      hipLaunchKernelGGL(k_een_pderiv, grid_for(n), B, 0, s, X, d.ntype, d.dim, d.type, d.lkpm, d.tmp_c, d.een_n, out);
      break;
    case DV_EEN_GL_PDERIV:
      n = d.W * d.ntype * d.dim * d.Ne * 4;
      hipLaunchKernelGGL(k_een_gl_pderiv, grid_for(n), B, 0, s, X, d.ntype, d.dim, d.type, d.lkpm, d.tmp_c, d.dtmp_c,
                         d.een_n, d.een_n_gl, out);
      break;
    case DV_FORCES_EN:
      n = d.W * d.Nn * 3;
      hipLaunchKernelGGL(k_forces_en, grid_for(n), B, 0, s, X, d.aord, d.type, d.a, d.en, d.eng, out);
      break;
    case DV_FORCES_EN_G:
      n = d.W * d.Nn * 3 * d.Ne * 3;
      hipLaunchKernelGGL(k_forces_en_g, grid_for(n), B, 0, s, X, d.aord, d.type, d.a, d.kappa, d.edist, d.en, d.eng, out);
      break;
    case DV_FORCES_EN_L:
      n = d.W * d.Nn * 3;
      hipLaunchKernelGGL(k_forces_en_l, grid_for(n), B, 0, s, X, d.aord, d.type, d.a, d.kappa, d.edist, d.en, d.eng, out);
      break;
    case DV_FORCES_TMP_C:
      n = d.W * d.C * (d.C + 1) * d.Nn * 4 * d.Ne;
      hipLaunchKernelGGL(k_forces_tmp_c, grid_for(n), B, 0, s, X, d.een_e, d.een_n_gl, out);
      break;
    case DV_FORCES_DTMP_C:
      n = d.W * d.C * (d.C + 1) * 4 * d.Nn * 4 * d.Ne;
      hipLaunchKernelGGL(k_forces_dtmp_c, grid_for(n), B, 0, s, X, d.een_e_gl, d.een_n_gl, out);
      break;
    case DV_FORCES_EEN_N:
      n = d.W * (d.C + 1) * 3 * d.Nn * 4 * d.Ne;
      hipLaunchKernelGGL(k_forces_een_n, grid_for(n), B, 0, s, X, d.type, d.kappa, d.edist, d.een_n, d.een_n_gl, out);
      break;
    case DV_FORCES_EEN:
      n = d.W * d.Nn * 3;
      hipLaunchKernelGGL(k_forces_een, grid_for(n), B, 0, s, X, d.dim, d.cfull, d.lkpm, d.een_n, d.een_n_gl, d.tmp_c,
                         d.ftmp_c, out);
      break;
    case DV_FORCES_EEN_G:
      n = d.W * 3 * d.Nn * 3 * d.Ne;
      hipLaunchKernelGGL(k_forces_een_g, grid_for(n), B, 0, s, X, d.dim, d.cfull, d.lkpm, d.tmp_c, d.dtmp_c, d.ftmp_c,
                         d.fdtmp_c, d.feen_n, d.een_n, d.een_n_gl, out);
      break;
    case DV_FORCES_EEN_L:
      n = d.W * 3 * d.Nn;
      hipLaunchKernelGGL(k_forces_een_l, grid_for(n), B, 0, s, X, d.dim, d.cfull, d.lkpm, d.tmp_c, d.dtmp_c, d.ftmp_c,
                         d.fdtmp_c, d.feen_n, d.een_n, d.een_n_gl, out);
      break;
    default:
      return hipErrorInvalidValue;
  }
  if (n <= 0) return hipSuccess;
  return hipGetLastError();
}

int64_t deriv_size(const DerivArgs& d, int what) {
  switch (what) {
    case DV_EE_PDERIV: return d.W * (d.bord + 1);
    case DV_EE_GL_PDERIV: return d.W * (d.bord + 1) * d.Ne * 4;
    case DV_EN_PDERIV: return d.W * d.ntype * (d.aord + 1);
    case DV_EN_GL_PDERIV: return d.W * d.ntype * (d.aord + 1) * d.Ne * 4;
    case DV_EEN_PDERIV: return d.W * d.ntype * d.dim;
    case DV_EEN_GL_PDERIV: return d.W * d.ntype * d.dim * d.Ne * 4;
    case DV_FORCES_EN: return d.W * d.Nn * 3;
    case DV_FORCES_EN_G: return d.W * d.Nn * 3 * d.Ne * 3;
    case DV_FORCES_EN_L: return d.W * d.Nn * 3;
    case DV_FORCES_TMP_C: return d.W * d.C * (d.C + 1) * d.Nn * 4 * d.Ne;
    case DV_FORCES_DTMP_C: return d.W * d.C * (d.C + 1) * 4 * d.Nn * 4 * d.Ne;
    case DV_FORCES_EEN_N: return d.W * (d.C + 1) * 3 * d.Nn * 4 * d.Ne;
    case DV_FORCES_EEN: return d.W * d.Nn * 3;
    case DV_FORCES_EEN_G: return d.W * 3 * d.Nn * 3 * d.Ne;
    case DV_FORCES_EEN_L: return d.W * 3 * d.Nn;
    default: return 0;
  }
}

hipError_t launch_sum_walkers(const double* in, int64_t W, int64_t n, double scale, double* out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sum_walkers, grid_for(n), dim3(256), 0, s, in, W, n, scale, out);
  return hipGetLastError();
}

}  // namespace qmckl_amd
