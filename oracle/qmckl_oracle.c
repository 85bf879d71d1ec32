/*
 * qmckl_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, FP64 restatement of the reference library's per-walker wavefunction
 * evaluation (the "doc" code path of TREX-CoE/qmckl).  It is the checker the parity
 * tests and the CPU baseline of bench.py compare the HIP product against; nothing in
 * the product links, loads or calls it.
 *
 * Every function cites the reference file:line it restates.  Array layouts are the
 * reference's C layouts (row-major as seen through its C API):
 *   point coordinates     coord[3][point_num]          (point = walker*elec_num + electron)
 *   nucleus coordinates   nucl_coord[3][nucl_num]
 *   AO VGL                ao_vgl[point][5][ao_num]
 *   MO VGL                mo_vgl[point][5][mo_num]
 *   Jastrow GL            gl[walker][4][elec_num]
 *
 * How each function is pinned (golden vectors of the reference tests, closed forms,
 * finite differences) is listed in the module docstring of oracle/__init__.py and in
 * the tests under tests/.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define SUCCESS 0
#define INVALID_ARG 1

/* Screening constant: cutoff = -log(epsilon), epsilon = 2^(1-precision).
 * org/qmckl_ao.org:118-120 (Fortran `cutoff = -dlog(qmckl_get_numprec_epsilon(context))`),
 * org/qmckl_numprec.org:348-365 (epsilon = 2^(1-p)). */
double oracle_cutoff(int32_t precision) { return -log(ldexp(1.0, 1 - precision)); }

/* ------------------------------------------------------------------------------------ */
/* Basis finalisation: org/qmckl_ao.org:2965-3183 (qmckl_finalize_basis).               */
/* ------------------------------------------------------------------------------------ */
typedef struct {
  double *coef_normalized; /* [prim_num]  coefficient*prim_factor*shell_factor (:3090-3110) */
  double *nucleus_range;   /* [nucl_num]  max over primitives of (r_power+1)/alpha (:3150-3175) */
  int32_t *nucleus_max_ang_mom; /* [nucl_num] (:3113-3130) */
  int64_t *ao_index;       /* [shell_num] first AO of each shell (:8470-8480) */
} basis_extra;

static void finalize_basis(int64_t nucl_num, int64_t shell_num, int64_t prim_num,
                           const int64_t *nucleus_index, const int64_t *nucleus_shell_num,
                           const int32_t *shell_ang_mom, const int64_t *shell_prim_num,
                           const int64_t *shell_prim_index, const double *shell_factor,
                           const double *exponent, const double *coefficient,
                           const double *prim_factor, const int32_t *r_power, basis_extra *b) {
  b->coef_normalized = (double *)calloc(prim_num, sizeof(double));
  b->nucleus_range = (double *)calloc(nucl_num, sizeof(double));
  b->nucleus_max_ang_mom = (int32_t *)calloc(nucl_num, sizeof(int32_t));
  b->ao_index = (int64_t *)calloc(shell_num, sizeof(int64_t));
  for (int64_t s = 0; s < shell_num; ++s)
    for (int64_t p = shell_prim_index[s]; p < shell_prim_index[s] + shell_prim_num[s]; ++p)
      b->coef_normalized[p] = coefficient[p] * prim_factor[p] * shell_factor[s];
  int64_t k = 0;
  for (int64_t a = 0; a < nucl_num; ++a) {
    b->nucleus_max_ang_mom[a] = 0;
    b->nucleus_range[a] = 0.;
    for (int64_t s = nucleus_index[a]; s < nucleus_index[a] + nucleus_shell_num[a]; ++s) {
      const int l = shell_ang_mom[s];
      if (l > b->nucleus_max_ang_mom[a]) b->nucleus_max_ang_mom[a] = l;
      b->ao_index[s] = k;
      k += (l + 1) * (l + 2) / 2;
      for (int64_t p = shell_prim_index[s]; p < shell_prim_index[s] + shell_prim_num[s]; ++p) {
        /* :3195-3205: range = (r_power + 1) / alpha; r_power defaults to 0 (:2982-3008) */
        const double n_val = (double)((r_power ? r_power[s] : 0) + 1);
        const double range = n_val / exponent[p];
        if (range > b->nucleus_range[a]) b->nucleus_range[a] = range;
      }
    }
  }
}

static void free_basis(basis_extra *b) {
  free(b->coef_normalized);
  free(b->nucleus_range);
  free(b->nucleus_max_ang_mom);
  free(b->ao_index);
}

/* Primitive VGL, org/qmckl_ao.org:4463-4520 (qmckl_compute_ao_basis_primitive_gaussian_vgl).
 * prim_vgl[point][5][prim_num].  nucleus_prim_index[nucl_num+1]. */
int oracle_prim_vgl(int64_t nucl_num, int64_t prim_num, const int64_t *nucleus_prim_index,
                    const double *nucl_coord, const double *exponent, int32_t precision,
                    int64_t point_num, const double *coord, double *prim_vgl) {
  const double cutoff = oracle_cutoff(precision);
  memset(prim_vgl, 0, sizeof(double) * point_num * 5 * prim_num);
  for (int64_t a = 0; a < nucl_num; ++a)
    for (int64_t p = nucleus_prim_index[a]; p < nucleus_prim_index[a + 1]; ++p)
      for (int64_t ip = 0; ip < point_num; ++ip) {
        const double x = coord[ip] - nucl_coord[a];
        const double y = coord[ip + point_num] - nucl_coord[a + nucl_num];
        const double z = coord[ip + 2 * point_num] - nucl_coord[a + 2 * nucl_num];
        const double r2 = x * x + y * y + z * z;
        const double ar2 = exponent[p] * r2;
        if (ar2 > cutoff) continue;
        const double v = exp(-ar2);
        const double two_a = -2.0 * exponent[p] * v;
        double *o = prim_vgl + ip * 5 * prim_num + p;
        o[0] = v;
        o[prim_num] = two_a * x;
        o[2 * prim_num] = two_a * y;
        o[3 * prim_num] = two_a * z;
        o[4 * prim_num] = two_a * (3.0 - 2.0 * ar2);
      }
  return SUCCESS;
}

/* Slater shells of nucleus a at displacement (x, y, z), org/qmckl_ao.org:4930-5040
 * (qmckl_compute_ao_basis_shell_slater_vgl): the nucleus is skipped when r > cutoff*range
 * (r, not r^2); at r < 1e-15 a shell of r_power 0 takes the sum of its coefficients as value
 * (gradient and Laplacian 0); else each primitive with alpha r <= cutoff adds
 * v = c r^n e^(-alpha r) (n = r_power, the power by repeated multiplication), the gradient
 * (n/r - alpha) v (x, y, z)/r and the Laplacian (n^2 - 2 n alpha r + alpha r (alpha r - 2) + n)
 * v / r^2. */
static void slater_shells(int64_t a, double x, double y, double z, int64_t shell_num,
                          const int64_t *nucleus_index, const int64_t *nucleus_shell_num,
                          const int64_t *shell_prim_num, const int64_t *shell_prim_index,
                          const double *exponent, const int32_t *r_power, const basis_extra *b,
                          double cutoff, double *shell) {
  const double r = sqrt(x * x + y * y + z * z);
  if (r > cutoff * b->nucleus_range[a]) return;
  for (int64_t s = nucleus_index[a]; s < nucleus_index[a] + nucleus_shell_num[a]; ++s) {
    double v0 = 0., v1 = 0., v2 = 0., v3 = 0., v4 = 0.;
    if (r < 1.e-15) {
      for (int64_t p = shell_prim_index[s]; p < shell_prim_index[s] + shell_prim_num[s]; ++p)
        if (r_power[s] == 0) v0 += b->coef_normalized[p];
    } else {
      const double r_inv = 1.0 / r, n = (double)r_power[s];
      for (int64_t p = shell_prim_index[s]; p < shell_prim_index[s] + shell_prim_num[s]; ++p) {
        const double ar = exponent[p] * r;
        if (ar > cutoff) continue;
        double rn = 1.0;
        for (int32_t e = 0; e < r_power[s]; ++e) rn *= r;
        const double v = b->coef_normalized[p] * rn * exp(-ar);
        v0 += v;
        const double g = (n * r_inv - exponent[p]) * v;
        v1 += g * x * r_inv;
        v2 += g * y * r_inv;
        v3 += g * z * r_inv;
        v4 += r_inv * r_inv * (n * n - 2.0 * n * ar + ar * (ar - 2.0) + n) * v;
      }
    }
    shell[s] = v0;
    shell[s + shell_num] = v1;
    shell[s + 2 * shell_num] = v2;
    shell[s + 3 * shell_num] = v3;
    shell[s + 4 * shell_num] = v4;
  }
}

/* Shell VGL of one point, org/qmckl_ao.org:4769-4868
 * (qmckl_compute_ao_basis_shell_gaussian_vgl): nucleus skipped when r^2 > cutoff*range,
 * primitive skipped when alpha r^2 > cutoff; Slater shells (slater != 0) as slater_shells.
 * shell[5][shell_num]. */
static void shell_vgl_point(double px, double py, double pz, int64_t nucl_num, int64_t shell_num,
                            const double *nucl_coord, const int64_t *nucleus_index,
                            const int64_t *nucleus_shell_num, const int64_t *shell_prim_num,
                            const int64_t *shell_prim_index, const double *exponent,
                            const int32_t *r_power, int32_t slater, const basis_extra *b, double cutoff,
                            double *shell) {
  memset(shell, 0, sizeof(double) * 5 * shell_num);
  for (int64_t a = 0; a < nucl_num; ++a) {
    const double x = px - nucl_coord[a];
    const double y = py - nucl_coord[a + nucl_num];
    const double z = pz - nucl_coord[a + 2 * nucl_num];
    if (slater) {
      slater_shells(a, x, y, z, shell_num, nucleus_index, nucleus_shell_num, shell_prim_num, shell_prim_index,
                    exponent, r_power, b, cutoff, shell);
      continue;
    }
    const double r2 = x * x + y * y + z * z;
    if (r2 > cutoff * b->nucleus_range[a]) continue;
    for (int64_t s = nucleus_index[a]; s < nucleus_index[a] + nucleus_shell_num[a]; ++s) {
      double v0 = 0., v1 = 0., v2 = 0., v3 = 0., v4 = 0.;
      for (int64_t p = shell_prim_index[s]; p < shell_prim_index[s] + shell_prim_num[s]; ++p) {
        const double ar2 = exponent[p] * r2;
        if (ar2 > cutoff) continue;
        const double v = b->coef_normalized[p] * exp(-ar2);
        const double two_a = -2.0 * exponent[p] * v;
        v0 += v;
        v1 += two_a * x;
        v2 += two_a * y;
        v3 += two_a * z;
        v4 += two_a * (3.0 - 2.0 * ar2);
      }
      shell[s] = v0;
      shell[s + shell_num] = v1;
      shell[s + 2 * shell_num] = v2;
      shell[s + 3 * shell_num] = v3;
      shell[s + 4 * shell_num] = v4;
    }
  }
}

/* Cartesian polynomial VGL, org/qmckl_ao.org:6233-6331 (qmckl_ao_polynomial_vgl_doc):
 * Y = (X - R) + 1e-20; order: increasing l, then a = l..0, b = l-a..0, c = l-a-b.
 * vgl[n][5]. Returns the number of polynomials. */
static int64_t polynomial_vgl(const double X[3], const double R[3], int32_t lmax, double *vgl) {
  double Y[3];
  for (int i = 0; i < 3; ++i) Y[i] = (X[i] - R[i]) + 1.e-20;
  double pw[3][34]; /* pw[d][e+2] = Y[d]^e for e = -2..lmax (negative powers set to 1) */
  for (int d = 0; d < 3; ++d) {
    pw[d][0] = pw[d][1] = pw[d][2] = 1.0;
    for (int e = 1; e <= lmax; ++e) pw[d][e + 2] = pw[d][e + 1] * Y[d];
  }
  int64_t n = 0;
  vgl[0] = 1.0; vgl[1] = vgl[2] = vgl[3] = vgl[4] = 0.0;
  n = 1;
  if (lmax >= 1) {
    for (int d = 0; d < 3; ++d) {
      double *v = vgl + 5 * (1 + d);
      v[0] = pw[d][3];
      v[1] = v[2] = v[3] = v[4] = 0.0;
      v[1 + d] = 1.0;
    }
    n = 4;
  }
  for (int l = 2; l <= lmax; ++l) {
    for (int a = l; a >= 0; --a) {
      for (int bb = l - a; bb >= 0; --bb) {
        const int c = l - a - bb;
        const double da = a, db = bb, dc = c;
        double *v = vgl + 5 * n;
        double xy = pw[0][a + 2] * pw[1][bb + 2];
        double yz = pw[1][bb + 2] * pw[2][c + 2];
        double xz = pw[0][a + 2] * pw[2][c + 2];
        v[0] = xy * pw[2][c + 2];
        xy = dc * xy;
        xz = db * xz;
        yz = da * yz;
        v[1] = pw[0][a + 1] * yz;
        v[2] = pw[1][bb + 1] * xz;
        v[3] = pw[2][c + 1] * xy;
        v[4] = (da - 1.0) * pw[0][a] * yz + (db - 1.0) * pw[1][bb] * xz + (dc - 1.0) * pw[2][c] * xy;
        ++n;
      }
    }
  }
  return n;
}

/* Shell VGL for points, org/qmckl_ao.org:4769-4868.  shell_vgl[point][5][shell_num]. */
int oracle_shell_vgl(int64_t nucl_num, int64_t shell_num, int64_t prim_num, const double *nucl_coord,
                     const int64_t *nucleus_index, const int64_t *nucleus_shell_num,
                     const int32_t *shell_ang_mom, const int64_t *shell_prim_num,
                     const int64_t *shell_prim_index, const double *shell_factor,
                     const double *exponent, const double *coefficient, const double *prim_factor,
                     const int32_t *r_power, int32_t precision, int32_t slater, int64_t point_num,
                     const double *coord, double *shell_vgl) {
  basis_extra b;
  finalize_basis(nucl_num, shell_num, prim_num, nucleus_index, nucleus_shell_num, shell_ang_mom,
                 shell_prim_num, shell_prim_index, shell_factor, exponent, coefficient, prim_factor,
                 r_power, &b);
  const double cutoff = oracle_cutoff(precision);
  for (int64_t ip = 0; ip < point_num; ++ip)
    shell_vgl_point(coord[ip], coord[ip + point_num], coord[ip + 2 * point_num], nucl_num, shell_num,
                    nucl_coord, nucleus_index, nucleus_shell_num, shell_prim_num, shell_prim_index,
                    exponent, r_power, slater, &b, cutoff, shell_vgl + ip * 5 * shell_num);
  free_basis(&b);
  return SUCCESS;
}

/* AO VGL, org/qmckl_ao.org:8438-8572 (qmckl_compute_ao_vgl_doc), with the shells from
 * :4769-4868 (Gaussian) or :4930-5040 (Slater) and the polynomials from :6233-6331; the
 * nucleus screening of this routine (r^2 > cutoff*range) applies to both shell types.
 * ao_vgl[point][5][ao_num]. */
int oracle_ao_vgl(int64_t nucl_num, int64_t shell_num, int64_t prim_num, int64_t ao_num,
                  const double *nucl_coord, const int64_t *nucleus_index,
                  const int64_t *nucleus_shell_num, const int32_t *shell_ang_mom,
                  const int64_t *shell_prim_num, const int64_t *shell_prim_index,
                  const double *shell_factor, const double *exponent, const double *coefficient,
                  const double *prim_factor, const int32_t *r_power, const double *ao_factor,
                  int32_t precision, int32_t slater, int64_t point_num, const double *coord, double *ao_vgl) {
  basis_extra b;
  finalize_basis(nucl_num, shell_num, prim_num, nucleus_index, nucleus_shell_num, shell_ang_mom,
                 shell_prim_num, shell_prim_index, shell_factor, exponent, coefficient, prim_factor,
                 r_power, &b);
  const double cutoff = oracle_cutoff(precision);
  int32_t lmax_all = 0;
  for (int64_t a = 0; a < nucl_num; ++a)
    if (b.nucleus_max_ang_mom[a] > lmax_all) lmax_all = b.nucleus_max_ang_mom[a];
  const int64_t npoly_max = (int64_t)(lmax_all + 1) * (lmax_all + 2) * (lmax_all + 3) / 6;

#pragma omp parallel
  {
    double *shell = (double *)malloc(sizeof(double) * 5 * shell_num);
    double *poly = (double *)malloc(sizeof(double) * 5 * npoly_max);
#pragma omp for schedule(static)
    for (int64_t ip = 0; ip < point_num; ++ip) {
      const double X[3] = {coord[ip], coord[ip + point_num], coord[ip + 2 * point_num]};
      shell_vgl_point(X[0], X[1], X[2], nucl_num, shell_num, nucl_coord, nucleus_index,
                      nucleus_shell_num, shell_prim_num, shell_prim_index, exponent, r_power, slater, &b, cutoff,
                      shell);
      double *out = ao_vgl + ip * 5 * ao_num;
      memset(out, 0, sizeof(double) * 5 * ao_num);
      for (int64_t a = 0; a < nucl_num; ++a) {
        const double R[3] = {nucl_coord[a], nucl_coord[a + nucl_num], nucl_coord[a + 2 * nucl_num]};
        const double x = X[0] - R[0], y = X[1] - R[1], z = X[2] - R[2];
        const double r2 = x * x + y * y + z * z;
        if (r2 > cutoff * b.nucleus_range[a]) continue;
        polynomial_vgl(X, R, b.nucleus_max_ang_mom[a], poly);
        for (int64_t s = nucleus_index[a]; s < nucleus_index[a] + nucleus_shell_num[a]; ++s) {
          const int l = shell_ang_mom[s];
          const int64_t il0 = (int64_t)l * (l + 1) * (l + 2) / 6;
          const int64_t nl = (l + 1) * (l + 2) / 2;
          const double s1 = shell[s], s2 = shell[s + shell_num], s3 = shell[s + 2 * shell_num],
                       s4 = shell[s + 3 * shell_num], s5 = shell[s + 4 * shell_num];
          for (int64_t q = 0; q < nl; ++q) {
            const int64_t k = b.ao_index[s] + q;
            const double *p = poly + 5 * (il0 + q);
            const double f = ao_factor[k];
            out[k] = p[0] * s1 * f;
            out[k + ao_num] = (p[1] * s1 + p[0] * s2) * f;
            out[k + 2 * ao_num] = (p[2] * s1 + p[0] * s3) * f;
            out[k + 3 * ao_num] = (p[3] * s1 + p[0] * s4) * f;
            out[k + 4 * ao_num] =
                (p[4] * s1 + p[0] * s5 + 2.0 * (p[1] * s2 + p[2] * s3 + p[3] * s4)) * f;
          }
        }
      }
    }
    free(shell);
    free(poly);
  }
  free_basis(&b);
  return SUCCESS;
}

/* MO VGL, org/qmckl_mo.org:1905-1940 (qmckl_compute_mo_basis_mo_vgl_doc):
 * mo[p][c][j] = sum_{k : ao[p][0][k] != 0} ao[p][c][k] * coefficient[j][k].
 * coefficient[mo_num][ao_num] (C order, :135-160).  mo_vgl[point][5][mo_num]. */
int oracle_mo_vgl(int64_t ao_num, int64_t mo_num, int64_t point_num, const double *coefficient,
                  const double *ao_vgl, double *mo_vgl) {
  double *ct = (double *)malloc(sizeof(double) * ao_num * mo_num); /* coefficient_t [ao][mo], :355-360 */
  for (int64_t i = 0; i < ao_num; ++i)
    for (int64_t j = 0; j < mo_num; ++j) ct[i * mo_num + j] = coefficient[j * ao_num + i];
#pragma omp parallel for schedule(static)
  for (int64_t ip = 0; ip < point_num; ++ip) {
    const double *av = ao_vgl + ip * 5 * ao_num;
    double *mv = mo_vgl + ip * 5 * mo_num;
    memset(mv, 0, sizeof(double) * 5 * mo_num);
    for (int64_t k = 0; k < ao_num; ++k) {
      if (av[k] == 0.0) continue;
      const double c1 = av[k], c2 = av[k + ao_num], c3 = av[k + 2 * ao_num],
                   c4 = av[k + 3 * ao_num], c5 = av[k + 4 * ao_num];
      const double *ck = ct + k * mo_num;
      for (int64_t j = 0; j < mo_num; ++j) {
        mv[j] += ck[j] * c1;
        mv[j + mo_num] += ck[j] * c2;
        mv[j + 2 * mo_num] += ck[j] * c3;
        mv[j + 3 * mo_num] += ck[j] * c4;
        mv[j + 4 * mo_num] += ck[j] * c5;
      }
    }
  }
  free(ct);
  return SUCCESS;
}

/* e-e distances, org/qmckl_electron.org:913-960 (qmckl_compute_ee_distance) via
 * org/qmckl_distance.org:488 (qmckl_distance).  ee[walk][elec][elec]. */
int oracle_ee_distance(int64_t elec_num, int64_t walk_num, const double *coord, double *ee) {
  const int64_t P = elec_num * walk_num;
  for (int64_t w = 0; w < walk_num; ++w)
    for (int64_t j = 0; j < elec_num; ++j)
      for (int64_t i = 0; i < elec_num; ++i) {
        const int64_t pi = w * elec_num + i, pj = w * elec_num + j;
        const double x = coord[pi] - coord[pj], y = coord[pi + P] - coord[pj + P],
                     z = coord[pi + 2 * P] - coord[pj + 2 * P];
        ee[(w * elec_num + j) * elec_num + i] = sqrt(x * x + y * y + z * z);
      }
  return SUCCESS;
}

/* e-n distances, org/qmckl_electron.org:1357-1390 (qmckl_compute_en_distance).
 * en[walk*elec][nucl]. */
int oracle_en_distance(int64_t elec_num, int64_t nucl_num, int64_t walk_num, const double *coord,
                       const double *nucl_coord, double *en) {
  const int64_t P = elec_num * walk_num;
  for (int64_t p = 0; p < P; ++p)
    for (int64_t a = 0; a < nucl_num; ++a) {
      const double x = coord[p] - nucl_coord[a], y = coord[p + P] - nucl_coord[a + nucl_num],
                   z = coord[p + 2 * P] - nucl_coord[a + 2 * nucl_num];
      en[p * nucl_num + a] = sqrt(x * x + y * y + z * z);
    }
  return SUCCESS;
}

/* ------------------------------------------------------------------------------------ */
/* CHAMP Jastrow                                                                         */
/* ------------------------------------------------------------------------------------ */

/* dim_c_vector, org/qmckl_jastrow_champ.org:11365-11405. */
int64_t oracle_dim_c_vector(int64_t cord_num) {
  int64_t n = 0;
  for (int64_t p = 2; p <= cord_num; ++p)
    for (int64_t k = p - 1; k >= 0; --k) {
      const int64_t lmax = (k != 0) ? p - k : p - k - 2;
      for (int64_t l = lmax; l >= 0; --l)
        if (((p - k - l) & 1) == 0) ++n;
    }
  return n;
}

/* lkpm_combined_index[4][dim_c], org/qmckl_jastrow_champ.org:12000-12040. */
int oracle_lkpm(int64_t cord_num, int64_t *lkpm) {
  const int64_t dim = oracle_dim_c_vector(cord_num);
  int64_t kk = 0;
  for (int64_t p = 2; p <= cord_num; ++p)
    for (int64_t k = p - 1; k >= 0; --k) {
      const int64_t lmax = (k != 0) ? p - k : p - k - 2;
      for (int64_t l = lmax; l >= 0; --l) {
        if (((p - k - l) & 1) == 1) continue;
        lkpm[kk] = l;
        lkpm[kk + dim] = k;
        lkpm[kk + 2 * dim] = p;
        lkpm[kk + 3 * dim] = (p - k - l) / 2;
        ++kk;
      }
    }
  return SUCCESS;
}

/* asymp_jasa[type], org/qmckl_jastrow_champ.org:5830-5875. a_vector[type][aord+1]. */
int oracle_asymp_jasa(int64_t aord, int64_t type_num, const double *a_vector,
                      const double *rescale_en, double *asymp_jasa) {
  for (int64_t t = 0; t < type_num; ++t) {
    const double *av = a_vector + t * (aord + 1);
    const double kinv = 1.0 / rescale_en[t];
    asymp_jasa[t] = av[0] * kinv / (1.0 + av[1] * kinv);
    double x = kinv;
    for (int64_t p = 2; p <= aord; ++p) {
      x *= kinv;
      asymp_jasa[t] += av[p] * x;
    }
  }
  return SUCCESS;
}

/* asymp_jasb[2], org/qmckl_jastrow_champ.org:2271-2320 (doc). */
int oracle_asymp_jasb(int64_t bord, const double *b_vector, double rescale_ee,
                      int32_t spin_independent, double *asymp_jasb) {
  const double kinv = 1.0 / rescale_ee;
  const double asym_one = b_vector[0] * kinv / (1.0 + b_vector[1] * kinv);
  asymp_jasb[0] = (spin_independent == 1) ? asym_one : 0.5 * asym_one;
  asymp_jasb[1] = asym_one;
  double x = kinv;
  for (int64_t p = 2; p <= bord; ++p) {
    x *= kinv;
    asymp_jasb[0] += b_vector[p] * x;
    asymp_jasb[1] += b_vector[p] * x;
  }
  return SUCCESS;
}

static double ipow(double x, int64_t n) {
  double r = 1.0;
  for (int64_t i = 0; i < n; ++i) r *= x;
  return r;
}

/* Full CHAMP Jastrow for all walkers.  Every output pointer may be NULL.
 *   coord[3][walk*elec], nucl_coord[3][nucl]
 *   a_vector[type][aord+1], b_vector[bord+1], c_vector[type][dim_c], rescale_en[type]
 * Intermediates (reference layouts):
 *   ee_rescaled [W][Ne][Ne]          org/qmckl_jastrow_champ.org:2685-2710, org/qmckl_distance.org:882-1027
 *   ee_rescaled_gl [W][Ne][Ne][4]    :3077-3120, org/qmckl_distance.org:1345-1509
 *   en_rescaled [W][Nn][Ne]          :6080-6135
 *   en_rescaled_gl [W][Nn][Ne][4]    :6503-6560
 *   een_e [W][c+1][Ne][Ne]           :9077-9130
 *   een_e_gl [W][c+1][Ne][4][Ne]     :9587-9680
 *   een_n [W][c+1][Nn][Ne]           :10701-10770
 *   een_n_gl [W][c+1][Nn][4][Ne]     :11090-11180
 *   tmp_c [W][c][c+1][Nn][Ne]        :12204-12260
 *   dtmp_c [W][c][c+1][Nn][4][Ne]    :12484-12540
 * Results:
 *   factor_ee [W], factor_ee_gl [W][4][Ne]   :3582-3650, :4188-4295
 *   factor_en [W], factor_en_gl [W][4][Ne]   :7021-7095, :7500-7590
 *   factor_een [W], factor_een_gl [W][4][Ne] :14121-14175, :14875-14945
 *   value [W], gl [W][4][Ne]                 :16578-16620, :17072-17112
 */
int oracle_jastrow(int64_t walk_num, int64_t elec_num, int64_t up_num, int64_t nucl_num,
                   const double *coord, const double *nucl_coord, int64_t type_num,
                   const int64_t *type_nucl_vector, int64_t aord, int64_t bord, int64_t cord,
                   const double *a_vector, const double *b_vector, const double *c_vector,
                   double rescale_ee, const double *rescale_en, int32_t spin_independent,
                   double *ee_rescaled, double *ee_rescaled_gl, double *en_rescaled,
                   double *en_rescaled_gl, double *een_e, double *een_e_gl, double *een_n,
                   double *een_n_gl, double *tmp_c, double *dtmp_c, double *factor_ee,
                   double *factor_ee_gl, double *factor_en, double *factor_en_gl,
                   double *factor_een, double *factor_een_gl, double *value, double *gl) {
  const int64_t Ne = elec_num, Nn = nucl_num, W = walk_num, C = cord, P = W * Ne;
  const int64_t dim_c = oracle_dim_c_vector(cord);
  double asymp_jasb[2];
  double *asymp_jasa = (double *)malloc(sizeof(double) * type_num);
  oracle_asymp_jasb(bord, b_vector, rescale_ee, spin_independent, asymp_jasb);
  oracle_asymp_jasa(aord, type_num, a_vector, rescale_en, asymp_jasa);
  int64_t *lkpm = (int64_t *)malloc(sizeof(int64_t) * 4 * (dim_c > 0 ? dim_c : 1));
  oracle_lkpm(cord, lkpm);
  /* c_vector_full[dim_c][Nn], :11865-11895 */
  double *cfull = (double *)malloc(sizeof(double) * (dim_c > 0 ? dim_c : 1) * Nn);
  for (int64_t n = 0; n < dim_c; ++n)
    for (int64_t a = 0; a < Nn; ++a) cfull[n * Nn + a] = c_vector[type_nucl_vector[a] * dim_c + n];

  /* Fewer walkers than threads: one thread walks the walkers and the whole team shares the
     P-tensor sums of each (see below); otherwise one walker per thread. */
  const int inner = W < omp_get_max_threads();
#pragma omp parallel if (!inner)
  {
    double *xe = (double *)malloc(sizeof(double) * 3 * Ne);
    double *rr = (double *)malloc(sizeof(double) * Ne * Ne);       /* ee distance [j][i] */
    double *fr = (double *)malloc(sizeof(double) * Ne * Ne);       /* ee rescaled [j][i] */
    double *fg = (double *)malloc(sizeof(double) * Ne * Ne * 4);   /* ee rescaled gl [j][i][4] */
    double *R = (double *)malloc(sizeof(double) * Ne * Nn);        /* en distance [i][a] */
    double *fn = (double *)malloc(sizeof(double) * Nn * Ne);       /* en rescaled [a][i] */
    double *fng = (double *)malloc(sizeof(double) * Nn * Ne * 4);  /* en rescaled gl [a][i][4] */
    double *ge = (double *)malloc(sizeof(double) * (C + 1) * Ne * Ne);      /* [k][j][i] */
    double *geg = (double *)malloc(sizeof(double) * (C + 1) * Ne * 4 * Ne); /* [k][j][4][i] */
    double *gn = (double *)malloc(sizeof(double) * (C + 1) * Nn * Ne);      /* [q][a][i] */
    double *gng = (double *)malloc(sizeof(double) * (C + 1) * Nn * 4 * Ne); /* [q][a][4][i] */
    const int64_t ntc = (C > 0 ? C : 1) * (C + 1) * Nn * Ne;
    double *tc = (double *)malloc(sizeof(double) * ntc);           /* [k][q][a][i] */
    double *dtc = (double *)malloc(sizeof(double) * ntc * 4);      /* [k][q][a][4][i] */
    double *gee = (double *)malloc(sizeof(double) * 4 * Ne);
    double *gen = (double *)malloc(sizeof(double) * 4 * Ne);
    double *geen = (double *)malloc(sizeof(double) * 4 * Ne);

#pragma omp for schedule(dynamic)
    for (int64_t w = 0; w < W; ++w) {
      for (int64_t i = 0; i < Ne; ++i)
        for (int d = 0; d < 3; ++d) xe[3 * i + d] = coord[w * Ne + i + d * P];

      /* --- e-e distances and rescaled distances ---------------------------------- */
      const double kinv = 1.0 / rescale_ee;
      for (int64_t j = 0; j < Ne; ++j)
        for (int64_t i = 0; i < Ne; ++i) {
          const double x = xe[3 * i] - xe[3 * j], y = xe[3 * i + 1] - xe[3 * j + 1],
                       z = xe[3 * i + 2] - xe[3 * j + 2];
          const double d = sqrt(x * x + y * y + z * z);
          rr[j * Ne + i] = d;
          fr[j * Ne + i] = (1.0 - exp(-rescale_ee * d)) * kinv;
          const double dd = d > 1.e-20 ? d : 1.e-20;
          const double dinv = 1.0 / dd, e = exp(-rescale_ee * dd);
          double *g = fg + (j * Ne + i) * 4;
          g[0] = x * dinv * e;
          g[1] = y * dinv * e;
          g[2] = z * dinv * e;
          g[3] = (2.0 * dinv - rescale_ee) * e;
        }

      /* --- J_ee and its GL ------------------------------------------------------- */
      double fee = 0.0;
      for (int64_t j = 0; j < Ne; ++j)
        for (int64_t i = 0; i < j; ++i) {
          const double x = fr[j * Ne + i];
          if (spin_independent == 1) {
            fee += b_vector[0] * x / (1.0 + b_vector[1] * x) - asymp_jasb[1];
          } else if (j < up_num || i >= up_num) {
            fee += 0.5 * b_vector[0] * x / (1.0 + b_vector[1] * x) - asymp_jasb[0];
          } else {
            fee += b_vector[0] * x / (1.0 + b_vector[1] * x) - asymp_jasb[1];
          }
          double xk = x;
          for (int64_t k = 2; k <= bord; ++k) {
            xk *= x;
            fee += b_vector[k] * xk;
          }
        }
      memset(gee, 0, sizeof(double) * 4 * Ne);
      for (int64_t j = 0; j < Ne; ++j)
        for (int64_t i = 0; i < Ne; ++i) {
          if (i == j) continue;
          double x = fr[j * Ne + i];
          const double denom = 1.0 + b_vector[1] * x;
          const double invd = 1.0 / denom, invd2 = invd * invd;
          const double *dx = fg + (j * Ne + i) * 4;
          const double gc2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
          double f;
          if (spin_independent == 1)
            f = b_vector[0] * invd2;
          else if ((i < up_num && j < up_num) || (i >= up_num && j >= up_num))
            f = 0.5 * b_vector[0] * invd2;
          else
            f = b_vector[0] * invd2;
          gee[i] += f * dx[0];
          gee[i + Ne] += f * dx[1];
          gee[i + 2 * Ne] += f * dx[2];
          gee[i + 3 * Ne] += f * (dx[3] - 2.0 * b_vector[1] * gc2 * invd);
          double kf = 2.0, x1 = x;
          x = 1.0;
          for (int64_t k = 2; k <= bord; ++k) {
            f = b_vector[k] * kf * x;
            gee[i] += f * x1 * dx[0];
            gee[i + Ne] += f * x1 * dx[1];
            gee[i + 2 * Ne] += f * x1 * dx[2];
            gee[i + 3 * Ne] += f * (x1 * dx[3] + (kf - 1.0) * gc2);
            x *= x1;
            kf += 1.0;
          }
        }

      /* --- e-n distances and rescaled distances ---------------------------------- */
      for (int64_t i = 0; i < Ne; ++i)
        for (int64_t a = 0; a < Nn; ++a) {
          const double x = xe[3 * i] - nucl_coord[a], y = xe[3 * i + 1] - nucl_coord[a + Nn],
                       z = xe[3 * i + 2] - nucl_coord[a + 2 * Nn];
          const double d = sqrt(x * x + y * y + z * z);
          const double kap = rescale_en[type_nucl_vector[a]];
          R[i * Nn + a] = d;
          fn[a * Ne + i] = (1.0 - exp(-kap * d)) * (1.0 / kap);
          const double dd = d > 1.e-20 ? d : 1.e-20;
          const double dinv = 1.0 / dd, e = exp(-kap * dd);
          double *g = fng + (a * Ne + i) * 4;
          g[0] = x * dinv * e;
          g[1] = y * dinv * e;
          g[2] = z * dinv * e;
          g[3] = (2.0 * dinv - kap) * e;
        }

      /* --- J_en and its GL ------------------------------------------------------- */
      double fen = 0.0;
      memset(gen, 0, sizeof(double) * 4 * Ne);
      for (int64_t a = 0; a < Nn; ++a) {
        const int64_t t = type_nucl_vector[a];
        const double *av = a_vector + t * (aord + 1);
        for (int64_t i = 0; i < Ne; ++i) {
          double x = fn[a * Ne + i];
          fen += av[0] * x / (1.0 + av[1] * x) - asymp_jasa[t];
          for (int64_t p = 2; p <= aord; ++p) {
            x *= fn[a * Ne + i];
            fen += av[p] * x;
          }
          x = fn[a * Ne + i];
          const double denom = 1.0 + av[1] * x;
          const double invd = 1.0 / denom, invd2 = invd * invd;
          const double *dx = fng + (a * Ne + i) * 4;
          double f = av[0] * invd2;
          const double gc2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
          gen[i] += f * dx[0];
          gen[i + Ne] += f * dx[1];
          gen[i + 2 * Ne] += f * dx[2];
          gen[i + 3 * Ne] += f * (dx[3] - 2.0 * av[1] * gc2 * invd);
          double kf = 2.0, x1 = x;
          x = 1.0;
          for (int64_t k = 2; k <= aord; ++k) {
            f = av[k] * kf * x;
            gen[i] += f * x1 * dx[0];
            gen[i + Ne] += f * x1 * dx[1];
            gen[i + 2 * Ne] += f * x1 * dx[2];
            gen[i + 3 * Ne] += f * (x1 * dx[3] + (kf - 1.0) * gc2);
            x *= x1;
            kf += 1.0;
          }
        }
      }

      /* --- e-e-n: rescaled powers ------------------------------------------------ */
      for (int64_t k = 0; k <= C; ++k)
        for (int64_t j = 0; j < Ne; ++j) {
          for (int64_t i = 0; i < Ne; ++i)
            ge[(k * Ne + j) * Ne + i] = ipow(exp(-rescale_ee * rr[j * Ne + i]), k);
          ge[(k * Ne + j) * Ne + j] = 0.0;
        }
      memset(geg, 0, sizeof(double) * (C + 1) * Ne * 4 * Ne);
      for (int64_t k = 1; k <= C; ++k) {
        const double kl = -(double)k * rescale_ee;
        for (int64_t j = 0; j < Ne; ++j)
          for (int64_t i = 0; i < Ne; ++i) {
            if (i == j) continue;
            const double rinv = 1.0 / rr[j * Ne + i];
            const double g = ge[(k * Ne + j) * Ne + i];
            double *o = geg + ((k * Ne + j) * 4) * Ne + i;
            for (int d = 0; d < 3; ++d)
              o[d * Ne] = kl * ((xe[3 * i + d] - xe[3 * j + d]) * rinv) * g;
            o[3 * Ne] = kl * (2.0 * rinv + kl) * g;
          }
      }
      for (int64_t a = 0; a < Nn; ++a) {
        const double kap = rescale_en[type_nucl_vector[a]];
        for (int64_t i = 0; i < Ne; ++i) {
          gn[(0 * Nn + a) * Ne + i] = 1.0;
          if (C >= 1) gn[(1 * Nn + a) * Ne + i] = exp(-kap * R[i * Nn + a]);
        }
        for (int64_t q = 2; q <= C; ++q)
          for (int64_t i = 0; i < Ne; ++i)
            gn[(q * Nn + a) * Ne + i] = gn[((q - 1) * Nn + a) * Ne + i] * gn[(1 * Nn + a) * Ne + i];
      }
      for (int64_t q = 0; q <= C; ++q)
        for (int64_t a = 0; a < Nn; ++a) {
          const double kl = -(double)q * rescale_en[type_nucl_vector[a]];
          for (int64_t i = 0; i < Ne; ++i) {
            const double rinv = 1.0 / R[i * Nn + a];
            double d[4];
            for (int c = 0; c < 3; ++c) d[c] = kl * ((xe[3 * i + c] - nucl_coord[a + c * Nn]) * rinv);
            d[3] = kl * 2.0 * rinv;
            d[3] = d[3] + d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            const double g = gn[(q * Nn + a) * Ne + i];
            for (int c = 0; c < 4; ++c) gng[((q * Nn + a) * 4 + c) * Ne + i] = d[c] * g;
          }
        }

      /* --- e-e-n: P tensors (tmp_c, dtmp_c) -------------------------------------- */
      /* tmp_c[k][q][a][i] = sum_m een_e[k][m][i] een_n[q][a][m] and dtmp_c likewise with
         een_e_gl, each element summed over m in ascending order (q > c - k stay zero).
         The (k, a) blocks are independent; with fewer walkers than threads the team shares
         them. */
      memset(tc, 0, sizeof(double) * ntc);
      memset(dtc, 0, sizeof(double) * ntc * 4);
#pragma omp parallel for collapse(2) schedule(dynamic) if (inner)
      for (int64_t k = 0; k < C; ++k)
        for (int64_t a = 0; a < Nn; ++a)
          for (int64_t q = 0; q <= C - k; ++q) {
            double *t = tc + ((k * (C + 1) + q) * Nn + a) * Ne;
            double *dt = dtc + (((k * (C + 1) + q) * Nn + a) * 4) * Ne;
            for (int64_t m = 0; m < Ne; ++m) {
              const double r = gn[(q * Nn + a) * Ne + m];
              const double *g = ge + (k * Ne + m) * Ne;
              const double *dg = geg + ((k * Ne + m) * 4) * Ne;
              for (int64_t i = 0; i < Ne; ++i) t[i] += g[i] * r;
              for (int c = 0; c < 4; ++c)
                for (int64_t i = 0; i < Ne; ++i) dt[c * Ne + i] += dg[c * Ne + i] * r;
            }
          }

      /* --- J_eeN and its GL ------------------------------------------------------ */
      double feen = 0.0;
      memset(geen, 0, sizeof(double) * 4 * Ne);
      if (C > 0) {
        for (int64_t n = 0; n < dim_c; ++n) {
          const int64_t l = lkpm[n], k = lkpm[n + dim_c], m = lkpm[n + 3 * dim_c];
          for (int64_t a = 0; a < Nn; ++a) {
            double cn = cfull[n * Nn + a];
            if (cn == 0.0) continue;
            double accu = 0.0;
            for (int64_t j = 0; j < Ne; ++j)
              accu += gn[(m * Nn + a) * Ne + j] * tc[((k * (C + 1) + m + l) * Nn + a) * Ne + j];
            feen += accu * cn;
            for (int c = 0; c < 4; ++c)
              for (int64_t j = 0; j < Ne; ++j)
                geen[c * Ne + j] +=
                    (tc[((k * (C + 1) + m) * Nn + a) * Ne + j] * gng[(((m + l) * Nn + a) * 4 + c) * Ne + j] +
                     tc[((k * (C + 1) + m + l) * Nn + a) * Ne + j] * gng[((m * Nn + a) * 4 + c) * Ne + j] +
                     dtc[(((k * (C + 1) + m) * Nn + a) * 4 + c) * Ne + j] * gn[((m + l) * Nn + a) * Ne + j] +
                     dtc[(((k * (C + 1) + m + l) * Nn + a) * 4 + c) * Ne + j] * gn[(m * Nn + a) * Ne + j]) *
                    cn;
            cn = cn + cn;
            for (int64_t j = 0; j < Ne; ++j) {
              double s = 0.0;
              for (int c = 0; c < 3; ++c)
                s += dtc[(((k * (C + 1) + m) * Nn + a) * 4 + c) * Ne + j] *
                         gng[(((m + l) * Nn + a) * 4 + c) * Ne + j];
              for (int c = 0; c < 3; ++c)
                s += dtc[(((k * (C + 1) + m + l) * Nn + a) * 4 + c) * Ne + j] *
                         gng[((m * Nn + a) * 4 + c) * Ne + j];
              geen[3 * Ne + j] += s * cn;
            }
          }
        }
      }

      /* --- totals ---------------------------------------------------------------- */
      const double jsum = fee + fen + feen;
      const double val = (jsum < -100.0) ? 0.0 : exp(jsum);
      if (factor_ee) factor_ee[w] = fee;
      if (factor_en) factor_en[w] = fen;
      if (factor_een) factor_een[w] = feen;
      if (value) value[w] = val;
      if (factor_ee_gl) memcpy(factor_ee_gl + w * 4 * Ne, gee, sizeof(double) * 4 * Ne);
      if (factor_en_gl) memcpy(factor_en_gl + w * 4 * Ne, gen, sizeof(double) * 4 * Ne);
      if (factor_een_gl) memcpy(factor_een_gl + w * 4 * Ne, geen, sizeof(double) * 4 * Ne);
      if (gl) {
        double *o = gl + w * 4 * Ne;
        for (int c = 0; c < 4; ++c)
          for (int64_t i = 0; i < Ne; ++i) o[c * Ne + i] = gee[c * Ne + i] + gen[c * Ne + i] + geen[c * Ne + i];
        for (int64_t i = 0; i < Ne; ++i)
          o[3 * Ne + i] += o[i] * o[i] + o[Ne + i] * o[Ne + i] + o[2 * Ne + i] * o[2 * Ne + i];
        for (int64_t i = 0; i < 4 * Ne; ++i) o[i] *= val;
      }
      if (ee_rescaled) memcpy(ee_rescaled + w * Ne * Ne, fr, sizeof(double) * Ne * Ne);
      if (ee_rescaled_gl) memcpy(ee_rescaled_gl + w * Ne * Ne * 4, fg, sizeof(double) * Ne * Ne * 4);
      if (en_rescaled) memcpy(en_rescaled + w * Nn * Ne, fn, sizeof(double) * Nn * Ne);
      if (en_rescaled_gl) memcpy(en_rescaled_gl + w * Nn * Ne * 4, fng, sizeof(double) * Nn * Ne * 4);
      if (een_e) memcpy(een_e + w * (C + 1) * Ne * Ne, ge, sizeof(double) * (C + 1) * Ne * Ne);
      if (een_e_gl) memcpy(een_e_gl + w * (C + 1) * Ne * 4 * Ne, geg, sizeof(double) * (C + 1) * Ne * 4 * Ne);
      if (een_n) memcpy(een_n + w * (C + 1) * Nn * Ne, gn, sizeof(double) * (C + 1) * Nn * Ne);
      if (een_n_gl) memcpy(een_n_gl + w * (C + 1) * Nn * 4 * Ne, gng, sizeof(double) * (C + 1) * Nn * 4 * Ne);
      if (tmp_c) memcpy(tmp_c + w * ntc, tc, sizeof(double) * ntc);
      if (dtmp_c) memcpy(dtmp_c + w * ntc * 4, dtc, sizeof(double) * ntc * 4);
    }
    free(xe); free(rr); free(fr); free(fg); free(R); free(fn); free(fng);
    free(ge); free(geg); free(gn); free(gng); free(tc); free(dtc);
    free(gee); free(gen); free(geen);
  }
  free(asymp_jasa);
  free(lkpm);
  free(cfull);
  return SUCCESS;
}

/* ------------------------------------------------------------------------------------ */
/* Hessians, org/qmckl_ao.org:5257-5585 (shells) and :9397-9718 (AOs).                  */
/* ------------------------------------------------------------------------------------ */

/* Derivatives of the radial part S = sum_p c_p exp(-alpha_p r^2) of shell s at r = (x,y,z):
 * h[i][j] = d_i d_j S, t[i] = d_i lap S (org/qmckl_ao.org:5257-5265, 5540-5584), and its
 * value, gradient and Laplacian in vgl[5]; same screening as the shell VGL. */
static void shell_hessian_at(const double r[3], double r2, int64_t s, const int64_t *shell_prim_num,
                             const int64_t *shell_prim_index, const double *exponent, const basis_extra *b,
                             double cutoff, double vgl[5], double h[3][3], double t[3]) {
  for (int i = 0; i < 5; ++i) vgl[i] = 0.0;
  for (int i = 0; i < 3; ++i) {
    t[i] = 0.0;
    for (int j = 0; j < 3; ++j) h[i][j] = 0.0;
  }
  for (int64_t p = shell_prim_index[s]; p < shell_prim_index[s] + shell_prim_num[s]; ++p) {
    const double ar2 = exponent[p] * r2;
    if (ar2 > cutoff) continue;
    const double v = b->coef_normalized[p] * exp(-ar2);
    const double ta = 2.0 * exponent[p];
    vgl[0] += v;
    for (int i = 0; i < 3; ++i) vgl[1 + i] += -ta * r[i] * v;
    vgl[4] += (ta * ta * r2 - 3.0 * ta) * v;
    for (int i = 0; i < 3; ++i) {
      for (int j = 0; j < 3; ++j) h[i][j] += (ta * ta * r[i] * r[j] - (i == j ? ta : 0.0)) * v;
      t[i] += (5.0 * ta * ta * r[i] - ta * ta * ta * r2 * r[i]) * v;
    }
  }
}

static double ipw(double y, int e) {
  if (e < 0) return 0.0;
  double r = 1.0;
  for (int i = 0; i < e; ++i) r *= y;
  return r;
}

/* Derivatives of the monomial M = y0^n0 y1^n1 y2^n2: value, gradient g, Hessian h, Laplacian
 * and the gradient t of its Laplacian (the role of qmckl_ao_polynomial_hessian, :9585-9630). */
static void monomial_derivs(const double Y[3], const int n[3], double *v, double g[3], double h[3][3],
                            double *lap, double t[3]) {
  double p[3][4], d1[3], d2[3], d3[3];
  for (int d = 0; d < 3; ++d) {
    const double e = n[d];
    p[d][0] = ipw(Y[d], n[d]);
    d1[d] = e * ipw(Y[d], n[d] - 1);
    d2[d] = e * (e - 1.0) * ipw(Y[d], n[d] - 2);
    d3[d] = e * (e - 1.0) * (e - 2.0) * ipw(Y[d], n[d] - 3);
  }
  *v = p[0][0] * p[1][0] * p[2][0];
  for (int i = 0; i < 3; ++i) {
    const int j = (i + 1) % 3, k = (i + 2) % 3;
    g[i] = d1[i] * p[j][0] * p[k][0];
    h[i][i] = d2[i] * p[j][0] * p[k][0];
    h[i][j] = h[j][i] = d1[i] * d1[j] * p[k][0];
    t[i] = d3[i] * p[j][0] * p[k][0] + d1[i] * (d2[j] * p[k][0] + p[j][0] * d2[k]);
  }
  *lap = h[0][0] + h[1][1] + h[2][2];
}

/* shell_hessian[point][3][4][shell_num]: [i][j] = d_i d_j S (j < 3), [i][3] = d_i lap S. */
int oracle_shell_hessian(int64_t nucl_num, int64_t shell_num, int64_t prim_num, const double *nucl_coord,
                         const int64_t *nucleus_index, const int64_t *nucleus_shell_num,
                         const int32_t *shell_ang_mom, const int64_t *shell_prim_num,
                         const int64_t *shell_prim_index, const double *shell_factor, const double *exponent,
                         const double *coefficient, const double *prim_factor, const int32_t *r_power,
                         int32_t precision, int64_t point_num, const double *coord, double *out) {
  basis_extra b;
  finalize_basis(nucl_num, shell_num, prim_num, nucleus_index, nucleus_shell_num, shell_ang_mom,
                 shell_prim_num, shell_prim_index, shell_factor, exponent, coefficient, prim_factor, r_power,
                 &b);
  const double cutoff = oracle_cutoff(precision);
  memset(out, 0, sizeof(double) * point_num * 12 * shell_num);
  for (int64_t ip = 0; ip < point_num; ++ip)
    for (int64_t a = 0; a < nucl_num; ++a) {
      const double r[3] = {coord[ip] - nucl_coord[a], coord[ip + point_num] - nucl_coord[a + nucl_num],
                           coord[ip + 2 * point_num] - nucl_coord[a + 2 * nucl_num]};
      const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
      if (r2 > cutoff * b.nucleus_range[a]) continue;
      for (int64_t s = nucleus_index[a]; s < nucleus_index[a] + nucleus_shell_num[a]; ++s) {
        double vgl[5], h[3][3], t[3];
        shell_hessian_at(r, r2, s, shell_prim_num, shell_prim_index, exponent, &b, cutoff, vgl, h, t);
        for (int i = 0; i < 3; ++i) {
          double *o = out + (ip * 3 + i) * 4 * shell_num + s;
          for (int j = 0; j < 3; ++j) o[j * shell_num] = h[i][j];
          o[3 * shell_num] = t[i];
        }
      }
    }
  free_basis(&b);
  return SUCCESS;
}

/* ao_hessian[3][point][4][ao_num]: [i][p][j][k] = d_i d_j AO_k (j < 3), [i][p][3][k] =
 * d_i lap AO_k, AO_k = f_k M_k S_s by the product rule of org/qmckl_ao.org:9686-9709. */
int oracle_ao_hessian(int64_t nucl_num, int64_t shell_num, int64_t prim_num, int64_t ao_num,
                      const double *nucl_coord, const int64_t *nucleus_index, const int64_t *nucleus_shell_num,
                      const int32_t *shell_ang_mom, const int64_t *shell_prim_num, const int64_t *shell_prim_index,
                      const double *shell_factor, const double *exponent, const double *coefficient,
                      const double *prim_factor, const int32_t *r_power, const double *ao_factor,
                      int32_t precision, int64_t point_num, const double *coord, double *out) {
  basis_extra b;
  finalize_basis(nucl_num, shell_num, prim_num, nucleus_index, nucleus_shell_num, shell_ang_mom,
                 shell_prim_num, shell_prim_index, shell_factor, exponent, coefficient, prim_factor, r_power,
                 &b);
  const double cutoff = oracle_cutoff(precision);
  memset(out, 0, sizeof(double) * 3 * point_num * 4 * ao_num);
#pragma omp parallel for schedule(static)
  for (int64_t ip = 0; ip < point_num; ++ip)
    for (int64_t a = 0; a < nucl_num; ++a) {
      const double r[3] = {coord[ip] - nucl_coord[a], coord[ip + point_num] - nucl_coord[a + nucl_num],
                           coord[ip + 2 * point_num] - nucl_coord[a + 2 * nucl_num]};
      const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
      if (r2 > cutoff * b.nucleus_range[a]) continue;
      for (int64_t s = nucleus_index[a]; s < nucleus_index[a] + nucleus_shell_num[a]; ++s) {
        double S[5], Sh[3][3], St[3];
        shell_hessian_at(r, r2, s, shell_prim_num, shell_prim_index, exponent, &b, cutoff, S, Sh, St);
        const int l = shell_ang_mom[s];
        int64_t k = b.ao_index[s];
        for (int na = l; na >= 0; --na)
          for (int nb = l - na; nb >= 0; --nb, ++k) {
            const int n[3] = {na, nb, l - na - nb};
            double M, Mg[3], Mh[3][3], Ml, Mt[3];
            monomial_derivs(r, n, &M, Mg, Mh, &Ml, Mt);
            const double f = ao_factor[k];
            for (int i = 0; i < 3; ++i) {
              double *o = out + ((int64_t)i * point_num + ip) * 4 * ao_num + k;
              for (int j = 0; j < 3; ++j)
                o[j * ao_num] = (Mh[i][j] * S[0] + M * Sh[i][j] + Mg[j] * S[1 + i] + Mg[i] * S[1 + j]) * f;
              double cross = 0.0;
              for (int j = 0; j < 3; ++j) cross += Mh[j][i] * S[1 + j] + Mg[j] * Sh[j][i];
              o[3 * ao_num] = (Mt[i] * S[0] + M * St[i] + Ml * S[1 + i] + Mg[i] * S[4] + 2.0 * cross) * f;
            }
          }
      }
    }
  free_basis(&b);
  return SUCCESS;
}
