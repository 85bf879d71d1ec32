/*
 * qmckl_oracle_deriv.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C FP64 restatement of the reference's Jastrow parameter derivatives
 * (org/qmckl_jastrow_champ.org) and Jastrow nuclear forces (org/qmckl_forces.org), as the
 * checker of the GPU implementation.  The inputs are the reference's intermediates in its C
 * layouts (as oracle_jastrow produces them); every function cites the reference function it
 * restates and follows its arithmetic, including these properties of the reference:
 *   - factor_ee/en (_gl)_pderiv are averaged over walkers, factor_een (_gl)_pderiv summed;
 *   - factor_ee_pderiv subtracts asymp_jasb_pderiv(2,1) (not (1,2)) from the B0 derivative
 *     of opposite-spin pairs (org/qmckl_jastrow_champ.org:5259);
 *   - factor_ee_gl_pderiv adds +2 B1 |grad x|^2 f1/(1 + B1 x) to the Laplacian of the B0
 *     derivative, where the exact derivative has -2 (org/qmckl_jastrow_champ.org:5584);
 *   - factor_en_pderiv uses a_vector(2) where the derivative of a_1 x/(1 + a_2 x) with
 *     respect to a_2 has a_1 (org/qmckl_jastrow_champ.org:8502);
 *   - the `if (abs(x) < 1.d-12) continue` guards are no-ops (Fortran `continue`).
 *
 * Layouts (C row-major; Fortran (i, j, ...) is C [...][j][i]):
 *   ee  [W][Ne(j)][Ne(i)]        eeg  [W][Ne(j)][Ne(i)][4]
 *   en  [W][Nn][Ne]              eng  [W][Nn][Ne][4]          edist [W*Ne][Nn]
 *   een_e [W][C+1][Ne(j)][Ne(i)] een_e_gl [W][C+1][Ne(j)][4][Ne(i)]
 *   een_n [W][C+1][Nn][Ne]       een_n_gl [W][C+1][Nn][4][Ne]
 *   tmp_c [W][C][C+1][Nn][Ne]    dtmp_c [W][C][C+1][Nn][4][Ne]
 *   cfull [dim][Nn]              lkpm [4][dim]
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define EE(i, j, w) ee[((int64_t)(w) * Ne + (j)) * Ne + (i)]
#define EEG(c, i, j, w) eeg[(((int64_t)(w) * Ne + (j)) * Ne + (i)) * 4 + (c)]
#define EN(i, a, w) en[((int64_t)(w) * Nn + (a)) * Ne + (i)]
#define ENG(c, i, a, w) eng[(((int64_t)(w) * Nn + (a)) * Ne + (i)) * 4 + (c)]
#define EDIST(a, i, w) edist[((int64_t)(w) * Ne + (i)) * Nn + (a)]
#define EENE(i, j, l, w) een_e[(((int64_t)(w) * (C + 1) + (l)) * Ne + (j)) * Ne + (i)]
#define EENEG(i, c, j, l, w) een_e_gl[((((int64_t)(w) * (C + 1) + (l)) * Ne + (j)) * 4 + (c)) * Ne + (i)]
#define EENN(i, a, q, w) een_n[(((int64_t)(w) * (C + 1) + (q)) * Nn + (a)) * Ne + (i)]
#define EENNG(i, c, a, q, w) een_n_gl[((((int64_t)(w) * (C + 1) + (q)) * Nn + (a)) * 4 + (c)) * Ne + (i)]
#define TMPC(i, a, q, k, w) tmp_c[((((int64_t)(w) * C + (k)) * (C + 1) + (q)) * Nn + (a)) * Ne + (i)]
#define DTMPC(i, c, a, q, k, w) dtmp_c[(((((int64_t)(w) * C + (k)) * (C + 1) + (q)) * Nn + (a)) * 4 + (c)) * Ne + (i)]
#define FTMPC(i, c, a, q, k, w) ftmp_c[(((((int64_t)(w) * C + (k)) * (C + 1) + (q)) * Nn + (a)) * 4 + (c)) * Ne + (i)]
#define FDTMPC(i, kk, a, k, q, l, w) \
  fdtmp_c[((((((int64_t)(w) * C + (l)) * (C + 1) + (q)) * 4 + (k)) * Nn + (a)) * 4 + (kk)) * Ne + (i)]
#define FEENN(i, m, a, n, l, w) feen_n[(((((int64_t)(w) * (C + 1) + (l)) * 3 + (n)) * Nn + (a)) * 4 + (m)) * Ne + (i)]
#define CF(a, n) cfull[(int64_t)(n) * Nn + (a)]
#define LK(n, col) lkpm[(int64_t)(col) * dim + (n)]

/* org/qmckl_jastrow_champ.org:4948-4994 (asymp_jasb_pderiv_doc): out [2][bord+1]. */
void oracle_asymp_jasb_pderiv(int64_t bord, const double *b, double kappa_ee, int32_t spin_independent,
                              double *out) {
  const double ki = 1.0 / kappa_ee;
  const double one = ki / (1.0 + b[1] * ki);
  const double two = -b[0] * ki * ki / ((1.0 + b[1] * ki) * (1.0 + b[1] * ki));
  const int64_t n = bord + 1;
  out[0] = spin_independent == 1 ? one : 0.5 * one; /* (1,1) */
  out[n] = one;                                     /* (1,2) */
  if (bord >= 1) {
    out[1] = spin_independent == 1 ? two : 0.5 * two; /* (2,1) */
    out[n + 1] = two;                                 /* (2,2) */
  }
  double x = ki;
  for (int64_t p = 2; p <= bord; ++p) {
    x *= ki;
    out[p] = x;
    out[n + p] = x;
  }
}

/* org/qmckl_jastrow_champ.org:8192-8215 (asymp_jasa_pderiv): out [ntype][aord+1]. */
void oracle_asymp_jasa_pderiv(int64_t aord, int64_t ntype, const double *a, const double *kappa_en, double *out) {
  for (int64_t t = 0; t < ntype; ++t) {
    const double *av = a + t * (aord + 1);
    double *o = out + t * (aord + 1);
    const double ki = 1.0 / kappa_en[t];
    o[0] = ki / (1.0 + av[1] * ki);
    o[1] = -av[0] * ki * ki / ((1.0 + av[1] * ki) * (1.0 + av[1] * ki));
    double x = ki;
    for (int64_t p = 2; p <= aord; ++p) {
      x *= ki;
      o[p] = x;
    }
  }
}

/* org/qmckl_jastrow_champ.org:5194-5273 (factor_ee_pderiv_doc): out [bord+1], walker mean. */
void oracle_factor_ee_pderiv(int64_t W, int64_t Ne, int64_t up, int64_t bord, const double *b, const double *ee,
                             const double *ajb, int32_t spin_independent, double *out) {
  const int64_t n = bord + 1;
#define AJB(p, s) ajb[(int64_t)(s) * n + (p)]
  for (int64_t p = 0; p < n; ++p) out[p] = 0.0;
  for (int64_t w = 0; w < W; ++w)
    for (int64_t j = 0; j < Ne; ++j)
      for (int64_t i = 0; i < j; ++i) {
        const double x = EE(i, j, w);
        const double den = 1.0 + b[1] * x;
        if (spin_independent == 1) {
          out[0] += x / den - AJB(0, 1);
          out[1] += -b[0] * x * x / (den * den) - AJB(1, 1);
        } else if (j < up || i >= up) {
          out[0] += 0.5 * x / den - AJB(0, 0);
          out[1] += -0.5 * b[0] * x * x / (den * den) - AJB(1, 0);
        } else {
          out[0] += x / den - AJB(1, 0);
          out[1] += -b[0] * x * x / (den * den) - AJB(1, 1);
        }
        double xk = x;
        for (int64_t k = 2; k <= bord; ++k) {
          xk *= x;
          out[k] += xk - AJB(k, 0);
        }
      }
  for (int64_t p = 0; p < n; ++p) out[p] /= (double)W;
#undef AJB
}

/* org/qmckl_jastrow_champ.org:5494-5605 (factor_ee_gl_pderiv_doc): out [bord+1][Ne][4], walker mean. */
void oracle_factor_ee_gl_pderiv(int64_t W, int64_t Ne, int64_t up, int64_t bord, const double *b, const double *ee,
                                const double *eeg, int32_t spin_independent, double *out) {
  const int64_t n = bord + 1;
#define O(c, i, p) out[((int64_t)(p) * Ne + (i)) * 4 + (c)]
  memset(out, 0, sizeof(double) * n * Ne * 4);
  for (int64_t w = 0; w < W; ++w)
    for (int64_t j = 0; j < Ne; ++j)
      for (int64_t i = 0; i < Ne; ++i) {
        if (i == j) continue;
        double x = EE(i, j, w);
        const double inv = 1.0 / (1.0 + b[1] * x), inv2 = inv * inv, inv3 = inv2 * inv;
        const double dx[4] = {EEG(0, i, j, w), EEG(1, i, j, w), EEG(2, i, j, w), EEG(3, i, j, w)};
        const double gc2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
        double f1, f2;
        if (spin_independent == 1) {
          f1 = inv2;
          f2 = -2.0 * b[0] * x * inv3;
        } else if ((i < up && j < up) || (i >= up && j >= up)) {
          f1 = 0.5 * inv2;
          f2 = -1.0 * b[0] * x * inv3;
        } else {
          f1 = inv2;
          f2 = -2.0 * b[0] * x * inv3;
        }
        for (int c = 0; c < 3; ++c) O(c, i, 0) += f1 * dx[c];
        O(3, i, 0) += f1 * dx[3] + 2.0 * b[1] * gc2 * f1 * inv;
        for (int c = 0; c < 3; ++c) O(c, i, 1) += f2 * dx[c];
        O(3, i, 1) += 2.0 * f1 * b[0] * (-inv * x * dx[3] - inv2 * gc2 * (1 - 2.0 * x * b[1]));
        double kf = 2.0;
        const double x1 = x;
        x = 1.0;
        for (int64_t k = 2; k < n; ++k) {
          const double f = kf * x;
          for (int c = 0; c < 3; ++c) O(c, i, k) += f * x1 * dx[c];
          O(3, i, k) += f * (x1 * dx[3] + (kf - 1.0) * gc2);
          x *= x1;
          kf += 1.0;
        }
      }
  for (int64_t t = 0; t < n * Ne * 4; ++t) out[t] /= (double)W;
#undef O
}

/* org/qmckl_jastrow_champ.org:8436-8520 (factor_en_pderiv_doc): out [ntype][aord+1], walker mean. */
void oracle_factor_en_pderiv(int64_t W, int64_t Ne, int64_t Nn, int64_t ntype, const int64_t *type, int64_t aord,
                             const double *a, const double *en, const double *aja, double *out) {
  const int64_t n = aord + 1;
  memset(out, 0, sizeof(double) * ntype * n);
  for (int64_t w = 0; w < W; ++w)
    for (int64_t at = 0; at < Nn; ++at)
      for (int64_t i = 0; i < Ne; ++i) {
        const int64_t t = type[at];
        const double *av = a + t * n, *aj = aja + t * n;
        double *o = out + t * n;
        double x = EN(i, at, w);
        const double den = 1.0 + av[1] * x;
        o[0] += x / den - aj[0];
        o[1] += -av[1] * x * x / (den * den) - aj[1];
        for (int64_t p = 2; p <= aord; ++p) {
          x *= EN(i, at, w);
          o[p] += x - aj[p];
        }
      }
  for (int64_t t = 0; t < ntype * n; ++t) out[t] /= (double)W;
}

/* org/qmckl_jastrow_champ.org:8735-8825 (factor_en_gl_pderiv_doc): out [ntype][aord+1][Ne][4], walker mean. */
void oracle_factor_en_gl_pderiv(int64_t W, int64_t Ne, int64_t Nn, int64_t ntype, const int64_t *type, int64_t aord,
                                const double *a, const double *en, const double *eng, double *out) {
  const int64_t n = aord + 1;
#define O(c, i, k, t) out[(((int64_t)(t) * n + (k)) * Ne + (i)) * 4 + (c)]
  memset(out, 0, sizeof(double) * ntype * n * Ne * 4);
  for (int64_t w = 0; w < W; ++w)
    for (int64_t at = 0; at < Nn; ++at)
      for (int64_t i = 0; i < Ne; ++i) {
        const int64_t t = type[at];
        const double *av = a + t * n;
        double x = EN(i, at, w);
        const double inv = 1.0 / (1.0 + av[1] * x), inv2 = inv * inv, inv3 = inv2 * inv;
        const double dx[4] = {ENG(0, i, at, w), ENG(1, i, at, w), ENG(2, i, at, w), ENG(3, i, at, w)};
        const double gc2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
        double f = inv2;
        for (int c = 0; c < 3; ++c) O(c, i, 0, t) += f * dx[c];
        O(3, i, 0, t) += f * (dx[3] - 2.0 * av[1] * gc2 * inv);
        f = -2.0 * av[0] * inv3;
        for (int c = 0; c < 3; ++c) O(c, i, 1, t) += f * x * dx[c];
        O(3, i, 1, t) += f * (x * dx[3] + inv * (gc2 - 2 * gc2 * x * av[1]));
        double kf = 2.0;
        const double x1 = x;
        x = 1.0;
        for (int64_t k = 2; k < n; ++k) {
          f = kf * x;
          for (int c = 0; c < 3; ++c) O(c, i, k, t) += f * x1 * dx[c];
          O(3, i, k, t) += f * (x1 * dx[3] + (kf - 1.0) * gc2);
          x *= x1;
          kf += 1.0;
        }
      }
  for (int64_t t = 0; t < ntype * n * Ne * 4; ++t) out[t] /= (double)W;
#undef O
}

/* org/qmckl_jastrow_champ.org:15931-15975 (factor_een_pderiv_doc): out [ntype][dim], summed over walkers. */
void oracle_factor_een_pderiv(int64_t W, int64_t Ne, int64_t Nn, int64_t ntype, const int64_t *type, int64_t C,
                              int64_t dim, const int64_t *lkpm, const double *tmp_c, const double *een_n, double *out) {
  memset(out, 0, sizeof(double) * ntype * dim);
  if (C == 0) return;
  for (int64_t w = 0; w < W; ++w)
    for (int64_t n = 0; n < dim; ++n) {
      const int64_t l = LK(n, 0), k = LK(n, 1), m = LK(n, 3);
      for (int64_t a = 0; a < Nn; ++a) {
        double acc = 0.0;
        for (int64_t j = 0; j < Ne; ++j) acc += EENN(j, a, m, w) * TMPC(j, a, m + l, k, w);
        out[type[a] * dim + n] += acc;
      }
    }
}

/* org/qmckl_jastrow_champ.org:16255-16305 (factor_een_gl_pderiv_doc): out [ntype][dim][Ne][4], summed over walkers. */
void oracle_factor_een_gl_pderiv(int64_t W, int64_t Ne, int64_t Nn, int64_t ntype, const int64_t *type, int64_t C,
                                 int64_t dim, const int64_t *lkpm, const double *tmp_c, const double *dtmp_c,
                                 const double *een_n, const double *een_n_gl, double *out) {
#define O(c, j, n, t) out[(((int64_t)(t) * dim + (n)) * Ne + (j)) * 4 + (c)]
  memset(out, 0, sizeof(double) * ntype * dim * Ne * 4);
  if (C == 0) return;
  for (int64_t w = 0; w < W; ++w)
    for (int64_t n = 0; n < dim; ++n) {
      const int64_t l = LK(n, 0), k = LK(n, 1), m = LK(n, 3);
      for (int64_t a = 0; a < Nn; ++a) {
        const int64_t t = type[a];
        for (int c = 0; c < 4; ++c)
          for (int64_t j = 0; j < Ne; ++j)
            O(c, j, n, t) += TMPC(j, a, m, k, w) * EENNG(j, c, a, m + l, w) + TMPC(j, a, m + l, k, w) * EENNG(j, c, a, m, w) +
                             DTMPC(j, c, a, m, k, w) * EENN(j, a, m + l, w) + DTMPC(j, c, a, m + l, k, w) * EENN(j, a, m, w);
        for (int64_t j = 0; j < Ne; ++j)
          O(3, j, n, t) += (DTMPC(j, 0, a, m, k, w) * EENNG(j, 0, a, m + l, w) + DTMPC(j, 1, a, m, k, w) * EENNG(j, 1, a, m + l, w) +
                            DTMPC(j, 2, a, m, k, w) * EENNG(j, 2, a, m + l, w) + DTMPC(j, 0, a, m + l, k, w) * EENNG(j, 0, a, m, w) +
                            DTMPC(j, 1, a, m + l, k, w) * EENNG(j, 1, a, m, w) + DTMPC(j, 2, a, m + l, k, w) * EENNG(j, 2, a, m, w)) *
                           2;
      }
    }
#undef O
}

/* org/qmckl_forces.org:732-790 (forces_jastrow_en_doc): out [W][Nn][3]. */
void oracle_forces_jastrow_en(int64_t W, int64_t Ne, int64_t Nn, const int64_t *type, int64_t aord, const double *a,
                              const double *en, const double *eng, double *out) {
  memset(out, 0, sizeof(double) * W * Nn * 3);
  for (int64_t w = 0; w < W; ++w)
    for (int64_t at = 0; at < Nn; ++at)
      for (int64_t i = 0; i < Ne; ++i) {
        const double *av = a + type[at] * (aord + 1);
        double *o = out + (w * Nn + at) * 3;
        double x = EN(i, at, w);
        const double inv = 1.0 / (1.0 + av[1] * x), inv2 = inv * inv;
        const double dx[3] = {-ENG(0, i, at, w), -ENG(1, i, at, w), -ENG(2, i, at, w)};
        double f = av[0] * inv2;
        for (int c = 0; c < 3; ++c) o[c] += f * dx[c];
        double kf = 2.0;
        const double x1 = x;
        x = 1.0;
        for (int64_t k = 2; k <= aord; ++k) {
          f = av[k] * kf * x;
          for (int c = 0; c < 3; ++c) o[c] += f * x1 * dx[c];
          x *= x1;
          kf += 1.0;
        }
      }
}

/* org/qmckl_forces.org:1123-1200 (forces_jastrow_en_g_doc): out [W][Nn][3(l)][Ne][3(m)]. */
void oracle_forces_jastrow_en_g(int64_t W, int64_t Ne, int64_t Nn, const int64_t *type, int64_t aord, const double *a,
                                const double *kappa_en, const double *edist, const double *en, const double *eng,
                                double *out) {
#define O(m, i, l, a_, w) out[((((int64_t)(w) * Nn + (a_)) * 3 + (l)) * Ne + (i)) * 3 + (m)]
  memset(out, 0, sizeof(double) * W * Nn * 3 * Ne * 3);
  for (int64_t w = 0; w < W; ++w)
    for (int64_t at = 0; at < Nn; ++at)
      for (int64_t i = 0; i < Ne; ++i) {
        const int64_t t = type[at];
        const double *av = a + t * (aord + 1);
        const double kap = kappa_en[t];
        const double expk = exp(kap * EDIST(at, i, w));
        const double invdist = 1.0 / EDIST(at, i, w);
        double x = EN(i, at, w);
        const double inv = 1.0 / (1.0 + av[1] * x), inv2 = inv * inv;
        double f = av[0] * inv2;
        const double dx[3] = {ENG(0, i, at, w), ENG(1, i, at, w), ENG(2, i, at, w)};
        for (int m = 0; m < 3; ++m)
          for (int l = 0; l < 3; ++l) {
            if (m == l) O(m, i, l, at, w) -= f * invdist / expk;
            O(m, i, l, at, w) += f * dx[m] * dx[l] * expk * (invdist + kap);
            O(m, i, l, at, w) += 2.0 * f * inv * av[1] * dx[m] * dx[l];
          }
        double kf = 2.0;
        const double x1 = x;
        x = 1.0;
        for (int64_t k = 2; k <= aord; ++k) {
          f = av[k] * kf * x;
          const double f2 = av[k] * kf * x * (kf - 1.0);
          for (int m = 0; m < 3; ++m)
            for (int l = 0; l < 3; ++l) {
              if (m == l) O(m, i, l, at, w) -= f * x1 * invdist / expk;
              O(m, i, l, at, w) += -f2 * dx[m] * dx[l] + f * x1 * dx[m] * dx[l] * kap * expk +
                                   f * x1 * dx[m] * dx[l] * invdist * expk;
            }
          x *= x1;
          kf += 1.0;
        }
      }
#undef O
}

/* org/qmckl_forces.org:1547-1620 (forces_jastrow_en_l_doc): out [W][Nn][3]. */
void oracle_forces_jastrow_en_l(int64_t W, int64_t Ne, int64_t Nn, const int64_t *type, int64_t aord, const double *a,
                                const double *kappa_en, const double *edist, const double *en, const double *eng,
                                double *out) {
  memset(out, 0, sizeof(double) * W * Nn * 3);
  for (int64_t w = 0; w < W; ++w)
    for (int64_t at = 0; at < Nn; ++at)
      for (int64_t i = 0; i < Ne; ++i) {
        const int64_t t = type[at];
        const double *av = a + t * (aord + 1);
        const double kap = kappa_en[t];
        double *o = out + (w * Nn + at) * 3;
        const double expk = exp(kap * EDIST(at, i, w));
        const double invdist = 1.0 / EDIST(at, i, w);
        double x = EN(i, at, w);
        const double inv = 1.0 / (1.0 + av[1] * x), inv2 = inv * inv;
        double f = av[0] * inv2;
        const double dx[4] = {ENG(0, i, at, w), ENG(1, i, at, w), ENG(2, i, at, w), ENG(3, i, at, w)};
        const double gc2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
        for (int m = 0; m < 3; ++m) {
          o[m] += f * dx[m] * (kap * (2 * invdist - kap) + 2 * invdist * invdist);
          o[m] += 2.0 * f * inv * av[1] * dx[m] * dx[3];
          o[m] -= 4.0 * f * inv * av[1] * kap * dx[m] / expk;
          o[m] -= 6.0 * f * inv * inv * av[1] * av[1] * gc2 * dx[m];
        }
        double kf = 2.0;
        const double x1 = x;
        x = 1.0;
        for (int64_t k = 2; k <= aord; ++k) {
          f = av[k] * kf * x;
          for (int m = 0; m < 3; ++m)
            o[m] += -f * dx[m] * dx[3] * (kf - 1.0) - f / x1 * (kf - 1.0) * (kf - 2.0) * dx[m] / expk / expk +
                    f * x1 * kap * dx[m] * dx[3] * expk + f * x1 * 2 * dx[m] * invdist * invdist +
                    2 * f * (kf - 1.0) * dx[m] * kap / expk;
          x *= x1;
          kf += 1.0;
        }
      }
}

/* org/qmckl_forces.org:2352-2385 (forces_tmp_c): out [W][C][C+1][Nn][4][Ne];
 * forces_tmp_c(:,:,:,j,i) = -een_rescaled_e(:,:,i) . een_rescaled_n_gl(:,:,:,j), j = 1..C-i;
 * the other entries stay 0 as the reference's zero-filled allocation leaves them. */
void oracle_forces_tmp_c(int64_t W, int64_t Ne, int64_t Nn, int64_t C, const double *een_e, const double *een_n_gl,
                         double *ftmp_c) {
  memset(ftmp_c, 0, sizeof(double) * W * C * (C + 1) * Nn * 4 * Ne);
  for (int64_t w = 0; w < W; ++w)
    for (int64_t i = 0; i < C; ++i)
      for (int64_t j = 1; j <= C - i; ++j)
        for (int64_t a = 0; a < Nn; ++a)
          for (int c = 0; c < 4; ++c)
            for (int64_t r = 0; r < Ne; ++r) {
              double s = 0.0;
              for (int64_t m = 0; m < Ne; ++m) s += EENE(r, m, i, w) * EENNG(m, c, a, j, w);
              FTMPC(r, c, a, j, i, w) = -s;
            }
}

/* org/qmckl_forces.org:2688-2725 (forces_dtmp_c_doc): out [W][C][C+1][4(k)][Nn][4(kk)][Ne]. */
void oracle_forces_dtmp_c(int64_t W, int64_t Ne, int64_t Nn, int64_t C, const double *een_e_gl, const double *een_n_gl,
                          double *fdtmp_c) {
  memset(fdtmp_c, 0, sizeof(double) * W * C * (C + 1) * 4 * Nn * 4 * Ne);
  for (int64_t w = 0; w < W; ++w)
    for (int64_t l = 0; l < C; ++l)
      for (int64_t m = 1; m <= C; ++m) {
        if (l + m > C) continue;
        for (int k = 0; k < 4; ++k)
          for (int64_t a = 0; a < Nn; ++a)
            for (int kk = 0; kk < 4; ++kk)
              for (int64_t i = 0; i < Ne; ++i) {
                double s = 0.0;
                for (int64_t j = 0; j < Ne; ++j) s -= EENEG(i, kk, j, l, w) * EENNG(j, k, a, m, w);
                FDTMPC(i, kk, a, k, m, l, w) = s;
              }
      }
}

/* org/qmckl_forces.org:3441-3520 (forces_een_rescaled_n_gl): out [W][C+1][3(n)][Nn][4(m)][Ne]. */
void oracle_forces_een_n(int64_t W, int64_t Ne, int64_t Nn, const int64_t *type, int64_t C, const double *kappa_en,
                         const double *edist, const double *een_n, const double *een_n_gl, double *feen_n) {
  memset(feen_n, 0, sizeof(double) * W * (C + 1) * 3 * Nn * 4 * Ne);
  for (int64_t w = 0; w < W; ++w)
    for (int64_t i = 0; i < Ne; ++i)
      for (int64_t a = 0; a < Nn; ++a)
        for (int64_t l = 1; l <= C; ++l) {
          const double kl = (double)l * kappa_en[type[a]];
          const double d = EDIST(a, i, w);
          for (int m = 0; m < 4; ++m)
            for (int n = 0; n < 3; ++n) {
              double v = 0.0;
              if (m == n) v += kl * EENN(i, a, l, w) / d;
              if (m < 3)
                v -= EENNG(i, m, a, l, w) * EENNG(i, n, a, l, w) / (kl * EENN(i, a, l, w) * d) +
                     EENNG(i, m, a, l, w) * EENNG(i, n, a, l, w) / EENN(i, a, l, w);
              else
                v += 2.0 * EENNG(i, n, a, l, w) / (d * d) - EENNG(i, m, a, l, w) * EENNG(i, n, a, l, w) / EENN(i, a, l, w);
              FEENN(i, m, a, n, l, w) = v;
            }
        }
}

/* org/qmckl_forces.org:3168-3210 (forces_jastrow_een): out [W][Nn][3]. */
void oracle_forces_jastrow_een(int64_t W, int64_t Ne, int64_t Nn, int64_t C, int64_t dim, const double *cfull,
                               const int64_t *lkpm, const double *een_n, const double *een_n_gl, const double *tmp_c,
                               const double *ftmp_c, double *out) {
  memset(out, 0, sizeof(double) * W * Nn * 3);
  for (int64_t w = 0; w < W; ++w)
    for (int64_t n = 0; n < dim; ++n) {
      const int64_t l = LK(n, 0), k = LK(n, 1), m = LK(n, 3);
      for (int64_t a = 0; a < Nn; ++a) {
        const double cn = CF(a, n);
        if (cn == 0.0) continue;
        for (int64_t j = 0; j < Ne; ++j)
          for (int ii = 0; ii < 3; ++ii) {
            const double acc = EENN(j, a, m, w) * FTMPC(j, ii, a, m + l, k, w) - EENNG(j, ii, a, m, w) * TMPC(j, a, m + l, k, w);
            out[(w * Nn + a) * 3 + ii] += acc * cn;
          }
      }
    }
}

/* org/qmckl_forces.org:3918-3965 (forces_jastrow_een_g): out [W][3(jj)][Nn][3(ii)][Ne]. */
void oracle_forces_jastrow_een_g(int64_t W, int64_t Ne, int64_t Nn, int64_t C, int64_t dim, const double *cfull,
                                 const int64_t *lkpm, const double *tmp_c, const double *dtmp_c, const double *ftmp_c,
                                 const double *fdtmp_c, const double *feen_n, const double *een_n,
                                 const double *een_n_gl, double *out) {
#define O(i, ii, a, jj, w) out[((((int64_t)(w) * 3 + (jj)) * Nn + (a)) * 3 + (ii)) * Ne + (i)]
  memset(out, 0, sizeof(double) * W * 3 * Nn * 3 * Ne);
  for (int64_t w = 0; w < W; ++w)
    for (int64_t n = 0; n < dim; ++n) {
      const int64_t l = LK(n, 0), k = LK(n, 1), m = LK(n, 3);
      for (int64_t a = 0; a < Nn; ++a) {
        const double cn = CF(a, n);
        if (cn == 0.0) continue;
        for (int ii = 0; ii < 3; ++ii)
          for (int64_t i = 0; i < Ne; ++i)
            for (int jj = 0; jj < 3; ++jj)
              O(i, ii, a, jj, w) += (TMPC(i, a, m, k, w) * FEENN(i, ii, a, jj, m + l, w) + FTMPC(i, jj, a, m, k, w) * EENNG(i, ii, a, m + l, w) +
                                     TMPC(i, a, m + l, k, w) * FEENN(i, ii, a, jj, m, w) + FTMPC(i, jj, a, m + l, k, w) * EENNG(i, ii, a, m, w) -
                                     DTMPC(i, ii, a, m, k, w) * EENNG(i, jj, a, m + l, w) + FDTMPC(i, ii, a, jj, m, k, w) * EENN(i, a, m + l, w) -
                                     DTMPC(i, ii, a, m + l, k, w) * EENNG(i, jj, a, m, w) + FDTMPC(i, ii, a, jj, m + l, k, w) * EENN(i, a, m, w)) *
                                    cn;
      }
    }
#undef O
}

/* org/qmckl_forces.org:4282-4345 (forces_jastrow_een_l): out [W][3][Nn]. */
void oracle_forces_jastrow_een_l(int64_t W, int64_t Ne, int64_t Nn, int64_t C, int64_t dim, const double *cfull,
                                 const int64_t *lkpm, const double *tmp_c, const double *dtmp_c, const double *ftmp_c,
                                 const double *fdtmp_c, const double *feen_n, const double *een_n,
                                 const double *een_n_gl, double *out) {
#define O(a, ii, w) out[((int64_t)(w) * 3 + (ii)) * Nn + (a)]
  memset(out, 0, sizeof(double) * W * 3 * Nn);
  for (int64_t w = 0; w < W; ++w)
    for (int64_t n = 0; n < dim; ++n) {
      const int64_t l = LK(n, 0), k = LK(n, 1), m = LK(n, 3);
      for (int64_t a = 0; a < Nn; ++a) {
        double cn = CF(a, n);
        if (cn == 0.0) continue;
        for (int ii = 0; ii < 3; ++ii)
          for (int64_t i = 0; i < Ne; ++i)
            O(a, ii, w) += (TMPC(i, a, m, k, w) * FEENN(i, 3, a, ii, m + l, w) + FTMPC(i, ii, a, m, k, w) * EENNG(i, 3, a, m + l, w) +
                            TMPC(i, a, m + l, k, w) * FEENN(i, 3, a, ii, m, w) + FTMPC(i, ii, a, m + l, k, w) * EENNG(i, 3, a, m, w) -
                            DTMPC(i, 3, a, m, k, w) * EENNG(i, ii, a, m + l, w) + FDTMPC(i, 3, a, ii, m, k, w) * EENN(i, a, m + l, w) -
                            DTMPC(i, 3, a, m + l, k, w) * EENNG(i, ii, a, m, w) + FDTMPC(i, 3, a, ii, m + l, k, w) * EENN(i, a, m, w)) *
                           cn;
        cn = cn + cn;
        for (int ii = 0; ii < 3; ++ii)
          for (int64_t i = 0; i < Ne; ++i)
            for (int jj = 0; jj < 3; ++jj)
              O(a, ii, w) += (DTMPC(i, jj, a, m, k, w) * FEENN(i, jj, a, ii, m + l, w) + DTMPC(i, jj, a, m + l, k, w) * FEENN(i, jj, a, ii, m, w) +
                              FDTMPC(i, jj, a, ii, m, k, w) * EENNG(i, jj, a, m + l, w) +
                              FDTMPC(i, jj, a, ii, m + l, k, w) * EENNG(i, jj, a, m, w)) *
                             cn;
      }
    }
#undef O
}
