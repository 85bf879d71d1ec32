/*
 * A C99 caller of include/qmckl.h, compiled with `gcc -std=c99 -pedantic -Werror` and
 * linked against libqmckl_amd.so by tests/test_cabi.py.  It drives the host side of the
 * API the way the reference's own C tests do (org/qmckl_context.org:733-760,
 * org/qmckl_nucleus.org:1090-1100, org/qmckl_ao.org:2457-2675,
 * org/qmckl_jastrow_champ.org:2420-2480): context life cycle, error record, thread count,
 * numerical precision, nuclei, electrons, an AO basis and the Jastrow parameters, with
 * their getters.  No call reaches a kernel, so it runs without a GPU.  Prints "ok" and
 * exits 0 when every check holds; otherwise prints the failed check and exits 1.
 */
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "qmckl.h"

static int failures = 0;
#define CHECK(cond)                                                   \
  do {                                                                \
    if (!(cond)) {                                                    \
      fprintf(stderr, "line %d: check failed: %s\n", __LINE__, #cond); \
      ++failures;                                                     \
    }                                                                 \
  } while (0)

int main(void) {
  qmckl_context ctx = qmckl_context_create();
  CHECK(ctx != QMCKL_NULL_CONTEXT);
  CHECK(qmckl_context_check(ctx) == ctx);

  /* thread count (org/qmckl_context.org:733-760) */
  int32_t nt = 0;
  CHECK(qmckl_context_get_num_threads(ctx, &nt) == QMCKL_SUCCESS && nt > 0);
  CHECK(qmckl_context_set_num_threads(ctx, 2) == QMCKL_SUCCESS);
  CHECK(qmckl_context_get_num_threads(ctx, &nt) == QMCKL_SUCCESS && nt == 2);
  CHECK(qmckl_context_set_num_threads(ctx, 0) == QMCKL_INVALID_ARG_2);
  CHECK(qmckl_context_unset_num_threads(ctx) == QMCKL_SUCCESS);
  CHECK(qmckl_context_get_num_threads(ctx, &nt) == QMCKL_SUCCESS && nt > 0);
  CHECK(qmckl_context_unset_num_threads(QMCKL_NULL_CONTEXT) == QMCKL_INVALID_CONTEXT);

  /* error record */
  CHECK(strcmp(qmckl_string_of_error(QMCKL_INVALID_ARG_2), "Invalid argument 2") == 0);
  CHECK(qmckl_failwith(ctx, QMCKL_FAILURE, "caller", "message") == QMCKL_FAILURE);
  {
    qmckl_exit_code code = 0;
    char fn[QMCKL_MAX_FUN_LEN], msg[QMCKL_MAX_MSG_LEN];
    CHECK(qmckl_get_error(ctx, &code, fn, msg) == QMCKL_SUCCESS);
    CHECK(code == QMCKL_FAILURE && strcmp(fn, "caller") == 0 && strcmp(msg, "message") == 0);
  }

  /* precision */
  CHECK(qmckl_get_numprec_precision(ctx) == 53);
  CHECK(qmckl_set_numprec_precision(ctx, 41) == QMCKL_SUCCESS);
  CHECK(qmckl_get_numprec_epsilon(ctx) == ldexp(1.0, -40));

  /* nuclei: two Z = 7 at distance 2.07, repulsion 49 / 2.07 */
  const double charge[2] = {7.0, 7.0};
  const double nucl_N[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 2.07}; /* [num][3] */
  double out[64];
  CHECK(qmckl_get_nucleus_charge(ctx, out, 2) == QMCKL_NOT_PROVIDED);
  CHECK(qmckl_set_nucleus_num(ctx, 2) == QMCKL_SUCCESS);
  CHECK(qmckl_set_nucleus_num(ctx, 2) == QMCKL_ALREADY_SET);
  CHECK(qmckl_set_nucleus_charge(ctx, charge, 1) == QMCKL_INVALID_ARG_3);
  CHECK(qmckl_set_nucleus_charge(ctx, charge, 2) == QMCKL_SUCCESS);
  CHECK(qmckl_set_nucleus_coord(ctx, 'N', nucl_N, 6) == QMCKL_SUCCESS);
  CHECK(qmckl_get_nucleus_coord(ctx, 'T', out, 6) == QMCKL_SUCCESS);
  CHECK(out[0] == 0.0 && out[1] == 0.0 && out[4] == 0.0 && out[5] == 2.07);
  {
    double rep = 0.0;
    CHECK(qmckl_get_nucleus_repulsion(ctx, &rep) == QMCKL_SUCCESS);
    CHECK(fabs(rep - 49.0 / 2.07) < 1e-13);
  }

  /* electrons: 2 walkers of 3 up + 2 down */
  CHECK(qmckl_set_electron_num(ctx, 3, 2) == QMCKL_SUCCESS);
  {
    double coord[2 * 5 * 3], back[2 * 5 * 3];
    int64_t i, n = 0;
    for (i = 0; i < 30; ++i) coord[i] = 0.1 * (double)i - 1.0;
    /* the size check is qmckl_set_point's (org/qmckl_electron.org:364, org/qmckl_point.org:306-310) */
    CHECK(qmckl_set_electron_coord(ctx, 'N', 2, coord, 29) == QMCKL_INVALID_ARG_4);
    CHECK(qmckl_set_electron_coord(ctx, 'N', 2, coord, 30) == QMCKL_SUCCESS);
    CHECK(qmckl_get_electron_walk_num(ctx, &n) == QMCKL_SUCCESS && n == 2);
    CHECK(qmckl_get_electron_coord(ctx, 'N', back, 30) == QMCKL_SUCCESS);
    CHECK(memcmp(coord, back, sizeof coord) == 0);
    CHECK(qmckl_get_electron_coord(ctx, 'T', back, 30) == QMCKL_SUCCESS);
    CHECK(back[1] == coord[3] && back[10] == coord[1] && back[29] == coord[29]);
  }

  /* AO basis: one s shell of two primitives per nucleus */
  {
    const int64_t nucleus_index[2] = {0, 1}, nucleus_shell_num[2] = {1, 1};
    const int32_t shell_ang_mom[2] = {0, 0};
    const int64_t shell_prim_num[2] = {2, 2}, shell_prim_index[2] = {0, 2};
    const double shell_factor[2] = {1.0, 1.0}, ao_factor[2] = {1.0, 1.0};
    const double exponent[4] = {5.0, 0.5, 5.0, 0.5}, coefficient[4] = {0.3, 0.7, 0.3, 0.7};
    const double prim_factor[4] = {1.0, 1.0, 1.0, 1.0};
    int64_t n = 0;
    CHECK(qmckl_set_ao_basis_type(ctx, 'X') == QMCKL_INVALID_ARG_2);
    CHECK(qmckl_set_ao_basis_type(ctx, 'G') == QMCKL_SUCCESS);
    CHECK(qmckl_set_ao_basis_shell_num(ctx, 2) == QMCKL_SUCCESS);
    CHECK(qmckl_set_ao_basis_prim_num(ctx, 4) == QMCKL_SUCCESS);
    CHECK(qmckl_set_ao_basis_nucleus_index(ctx, nucleus_index, 2) == QMCKL_SUCCESS);
    CHECK(qmckl_set_ao_basis_nucleus_shell_num(ctx, nucleus_shell_num, 2) == QMCKL_SUCCESS);
    CHECK(qmckl_set_ao_basis_shell_ang_mom(ctx, shell_ang_mom, 2) == QMCKL_SUCCESS);
    CHECK(qmckl_set_ao_basis_shell_prim_num(ctx, shell_prim_num, 2) == QMCKL_SUCCESS);
    CHECK(qmckl_set_ao_basis_shell_prim_index(ctx, shell_prim_index, 2) == QMCKL_SUCCESS);
    CHECK(qmckl_set_ao_basis_shell_factor(ctx, shell_factor, 2) == QMCKL_SUCCESS);
    CHECK(qmckl_set_ao_basis_exponent(ctx, exponent, 4) == QMCKL_SUCCESS);
    CHECK(qmckl_set_ao_basis_coefficient(ctx, coefficient, 4) == QMCKL_SUCCESS);
    CHECK(qmckl_set_ao_basis_prim_factor(ctx, prim_factor, 4) == QMCKL_SUCCESS);
    CHECK(!qmckl_ao_basis_provided(ctx));
    CHECK(qmckl_set_ao_basis_ao_num(ctx, 2) == QMCKL_SUCCESS);
    CHECK(qmckl_set_ao_basis_ao_factor(ctx, ao_factor, 2) == QMCKL_SUCCESS);
    CHECK(qmckl_ao_basis_provided(ctx));
    CHECK(qmckl_get_ao_basis_ao_num(ctx, &n) == QMCKL_SUCCESS && n == 2);
    CHECK(qmckl_get_ao_basis_exponent(ctx, out, 4) == QMCKL_SUCCESS && out[1] == 0.5 && out[2] == 5.0);
    CHECK(qmckl_set_ao_basis_exponent(ctx, exponent, 4) == QMCKL_ALREADY_SET);
  }

  /* CHAMP Jastrow parameters */
  {
    const int64_t type_nucl_vector[2] = {0, 0};
    const double a_vector[6] = {0.0, 0.0, -0.38, -0.03, 0.02, 0.004};
    const double b_vector[6] = {0.5, 0.15, 0.07, -0.04, 0.02, 0.001};
    const double rescale_en[1] = {0.6};
    double c_vector[23];
    int64_t i, dim = 0;
    for (i = 0; i < 23; ++i) c_vector[i] = 0.01 * (double)(i % 7) - 0.03;
    CHECK(qmckl_set_jastrow_champ_spin_independent(ctx, 0) == QMCKL_SUCCESS);
    CHECK(qmckl_set_jastrow_champ_aord_num(ctx, 5) == QMCKL_SUCCESS);
    CHECK(qmckl_set_jastrow_champ_bord_num(ctx, 5) == QMCKL_SUCCESS);
    CHECK(qmckl_set_jastrow_champ_cord_num(ctx, 5) == QMCKL_SUCCESS);
    CHECK(qmckl_get_jastrow_champ_dim_c_vector(ctx, &dim) == QMCKL_SUCCESS && dim == 23);
    CHECK(qmckl_set_jastrow_champ_type_nucl_num(ctx, 1) == QMCKL_SUCCESS);
    CHECK(qmckl_set_jastrow_champ_type_nucl_vector(ctx, type_nucl_vector, 2) == QMCKL_SUCCESS);
    CHECK(qmckl_set_jastrow_champ_a_vector(ctx, a_vector, 6) == QMCKL_SUCCESS);
    CHECK(qmckl_set_jastrow_champ_b_vector(ctx, b_vector, 6) == QMCKL_SUCCESS);
    CHECK(qmckl_set_jastrow_champ_c_vector(ctx, c_vector, 23) == QMCKL_SUCCESS);
    CHECK(qmckl_set_jastrow_champ_rescale_factor_en(ctx, rescale_en, 1) == QMCKL_SUCCESS);
    CHECK(!qmckl_jastrow_champ_provided(ctx));
    CHECK(qmckl_set_jastrow_champ_rescale_factor_ee(ctx, 0.6) == QMCKL_SUCCESS);
    CHECK(qmckl_jastrow_champ_provided(ctx));
    CHECK(qmckl_get_jastrow_champ_b_vector(ctx, out, 6) == QMCKL_SUCCESS && out[1] == 0.15);
    /* asymp_jasb[1] = b0 / k / (1 + b1 / k) + sum_{k>=2} b_k / k^k  (org/qmckl_jastrow_champ.org:2340-2380) */
    CHECK(qmckl_get_jastrow_champ_asymp_jasb(ctx, out, 2) == QMCKL_SUCCESS);
    {
      const double ki = 1.0 / 0.6;
      double ref = 0.5 * ki / (1.0 + 0.15 * ki), x = ki;
      int k;
      for (k = 2; k <= 5; ++k) {
        x *= ki;
        ref += b_vector[k] * x;
      }
      CHECK(fabs(out[1] - ref) < 1e-14 * fabs(ref));
    }
  }

  /* a copy carries the host state */
  {
    qmckl_context c2 = qmckl_context_copy(ctx);
    int64_t n = 0;
    CHECK(c2 != QMCKL_NULL_CONTEXT && c2 != ctx);
    CHECK(qmckl_get_electron_walk_num(c2, &n) == QMCKL_SUCCESS && n == 2);
    CHECK(qmckl_jastrow_champ_provided(c2) && qmckl_ao_basis_provided(c2));
    CHECK(qmckl_context_destroy(c2) == QMCKL_SUCCESS);
  }

  CHECK(qmckl_context_destroy(ctx) == QMCKL_SUCCESS);
  CHECK(qmckl_context_check(ctx) == QMCKL_NULL_CONTEXT);
  if (failures) return 1;
  printf("ok\n");
  return 0;
}
