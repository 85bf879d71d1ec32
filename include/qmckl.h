/*
 * qmckl.h -- C-ABI of the MI355X-native per-walker wavefunction evaluator.
 *
 * This header declares the subset of the TREX-CoE/qmckl public API that the walker
 * evaluation path needs (context, errors, numerical precision, points, nuclei,
 * electrons, contracted-Gaussian AO basis, MO basis, CHAMP Jastrow), with the
 * reference's names, argument order, argument meaning, return codes and host-pointer
 * FP64 semantics, so that C callers, the reference's Fortran `bind(C)` module and its
 * SWIG Python wrapper can link against libqmckl_amd.so instead of libqmckl.so.
 * Each group cites the reference file:line whose declaration it replaces.
 *
 * Entry points whose name starts with `qmckl_amd_` have no counterpart in the reference:
 * they expose device residency (inputs/outputs in GPU memory), the HIP stream and the
 * precision of the AO->MO product.  They are listed at the end.
 *
 * Layouts are C row-major:
 *   point coordinates    'N': [num][3]           'T': [3][num]
 *   AO VGL               [point][5][ao_num]      (value, d/dx, d/dy, d/dz, Laplacian)
 *   MO VGL               [point][5][mo_num]
 *   Jastrow value        [walk_num]
 *   Jastrow GL           [walk_num][4][elec_num] (d/dx, d/dy, d/dz, Laplacian of e^J)
 * with point = walker * elec_num + electron, up-spin electrons first.
 */
#ifndef QMCKL_AMD_H
#define QMCKL_AMD_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Types and constants: org/qmckl_context.org:104-105, org/qmckl_error.org:67,152-181,406-407 */
typedef int64_t qmckl_context;
typedef int32_t qmckl_exit_code;

#define QMCKL_NULL_CONTEXT ((qmckl_context)0)

#define QMCKL_SUCCESS ((qmckl_exit_code)0)
#define QMCKL_INVALID_ARG_1 ((qmckl_exit_code)1)
#define QMCKL_INVALID_ARG_2 ((qmckl_exit_code)2)
#define QMCKL_INVALID_ARG_3 ((qmckl_exit_code)3)
#define QMCKL_INVALID_ARG_4 ((qmckl_exit_code)4)
#define QMCKL_INVALID_ARG_5 ((qmckl_exit_code)5)
#define QMCKL_INVALID_ARG_6 ((qmckl_exit_code)6)
#define QMCKL_INVALID_ARG_7 ((qmckl_exit_code)7)
#define QMCKL_INVALID_ARG_8 ((qmckl_exit_code)8)
#define QMCKL_INVALID_ARG_9 ((qmckl_exit_code)9)
#define QMCKL_INVALID_ARG_10 ((qmckl_exit_code)10)
#define QMCKL_INVALID_ARG_11 ((qmckl_exit_code)11)
#define QMCKL_INVALID_ARG_12 ((qmckl_exit_code)12)
#define QMCKL_INVALID_ARG_13 ((qmckl_exit_code)13)
#define QMCKL_INVALID_ARG_14 ((qmckl_exit_code)14)
#define QMCKL_INVALID_ARG_15 ((qmckl_exit_code)15)
#define QMCKL_INVALID_ARG_16 ((qmckl_exit_code)16)
#define QMCKL_INVALID_ARG_17 ((qmckl_exit_code)17)
#define QMCKL_INVALID_ARG_18 ((qmckl_exit_code)18)
#define QMCKL_INVALID_ARG_19 ((qmckl_exit_code)19)
#define QMCKL_INVALID_ARG_20 ((qmckl_exit_code)20)
#define QMCKL_FAILURE ((qmckl_exit_code)101)
#define QMCKL_ERRNO ((qmckl_exit_code)102)
#define QMCKL_INVALID_CONTEXT ((qmckl_exit_code)103)
#define QMCKL_ALLOCATION_FAILED ((qmckl_exit_code)104)
#define QMCKL_DEALLOCATION_FAILED ((qmckl_exit_code)105)
#define QMCKL_NOT_PROVIDED ((qmckl_exit_code)106)
#define QMCKL_OUT_OF_BOUNDS ((qmckl_exit_code)107)
#define QMCKL_ALREADY_SET ((qmckl_exit_code)108)
#define QMCKL_INVALID_EXIT_CODE ((qmckl_exit_code)109)

#define QMCKL_MAX_FUN_LEN 256
#define QMCKL_MAX_MSG_LEN 1024

/* ---- Context: org/qmckl_context.org:221, 253, 302, 428, 467, 609, 648, 695, 769 */
qmckl_context qmckl_context_check(const qmckl_context context);
qmckl_exit_code qmckl_context_touch(const qmckl_context context);
qmckl_context qmckl_context_create(void);
void qmckl_lock(qmckl_context context);
void qmckl_unlock(qmckl_context context);
qmckl_context qmckl_context_copy(const qmckl_context context);
qmckl_exit_code qmckl_context_get_num_threads(const qmckl_context context, int32_t* const num_threads);
qmckl_exit_code qmckl_context_set_num_threads(const qmckl_context context, const int32_t num_threads);
qmckl_exit_code qmckl_context_unset_num_threads(const qmckl_context context);
qmckl_exit_code qmckl_context_destroy(const qmckl_context context);

/* ---- Errors: org/qmckl_error.org:241, 433, 493, 550, 606, 660 */
const char* qmckl_string_of_error(const qmckl_exit_code error);
qmckl_exit_code qmckl_set_error(qmckl_context context, const qmckl_exit_code exit_code,
                                const char* function_name, const char* message);
qmckl_exit_code qmckl_get_error(qmckl_context context, qmckl_exit_code* exit_code,
                                char* function_name, char* message);
qmckl_exit_code qmckl_failwith(qmckl_context context, const qmckl_exit_code exit_code,
                               const char* function, const char* message);
qmckl_exit_code qmckl_last_error(qmckl_context context, char* buffer);
qmckl_exit_code qmckl_check(qmckl_context context, qmckl_exit_code rc);

/* ---- Numerical precision: org/qmckl_numprec.org:169, 224, 260, 314, 351, 412 */
qmckl_exit_code qmckl_set_numprec_precision(const qmckl_context context, const int precision);
int32_t qmckl_get_numprec_precision(const qmckl_context context);
qmckl_exit_code qmckl_set_numprec_range(const qmckl_context context, const int range);
int32_t qmckl_get_numprec_range(const qmckl_context context);
double qmckl_get_numprec_epsilon(const qmckl_context context);
int32_t qmckl_test_precision_64(double a, double b);
int32_t qmckl_test_precision_32(float a, float b);

/* ---- Points: org/qmckl_point.org:143, 188, 276 */
qmckl_exit_code qmckl_get_point_num(const qmckl_context context, int64_t* const num);
qmckl_exit_code qmckl_get_point(const qmckl_context context, const char transp, double* const coord,
                                const int64_t size_max);
qmckl_exit_code qmckl_set_point(qmckl_context context, const char transp, const int64_t num,
                                const double* coord, const int64_t size_max);

/* ---- Nuclei: org/qmckl_nucleus.org:161, 216, 285, 367, 417, 461, 529, 780, 954 */
qmckl_exit_code qmckl_get_nucleus_num(const qmckl_context context, int64_t* const num);
qmckl_exit_code qmckl_get_nucleus_charge(const qmckl_context context, double* const charge,
                                         const int64_t size_max);
qmckl_exit_code qmckl_get_nucleus_coord(const qmckl_context context, const char transp,
                                        double* const coord, const int64_t size_max);
bool qmckl_nucleus_provided(const qmckl_context context);
qmckl_exit_code qmckl_set_nucleus_num(qmckl_context context, const int64_t num);
qmckl_exit_code qmckl_set_nucleus_charge(qmckl_context context, const double* charge,
                                         const int64_t size_max);
qmckl_exit_code qmckl_set_nucleus_coord(qmckl_context context, const char transp,
                                        const double* coord, const int64_t size_max);
qmckl_exit_code qmckl_get_nucleus_nn_distance(qmckl_context context, double* distance,
                                              const int64_t size_max);
qmckl_exit_code qmckl_get_nucleus_repulsion(qmckl_context context, double* const energy);

/* ---- Electrons: org/qmckl_electron.org:197, 222, 406, 504, 543, 776, 1035, 1203, 1489 */
bool qmckl_electron_provided(const qmckl_context context);
qmckl_exit_code qmckl_set_electron_num(qmckl_context context, const int64_t up_num,
                                       const int64_t down_num);
qmckl_exit_code qmckl_set_electron_coord(qmckl_context context, const char transp,
                                         const int64_t walk_num, const double* coord,
                                         const int64_t size_max);
qmckl_exit_code qmckl_get_electron_num(const qmckl_context context, int64_t* const num);
qmckl_exit_code qmckl_get_electron_up_num(const qmckl_context context, int64_t* const up_num);
qmckl_exit_code qmckl_get_electron_down_num(const qmckl_context context, int64_t* const down_num);
qmckl_exit_code qmckl_get_electron_walk_num(const qmckl_context context, int64_t* const walk_num);
qmckl_exit_code qmckl_get_electron_coord(const qmckl_context context, const char transp,
                                         double* const coord, const int64_t size_max);
qmckl_exit_code qmckl_get_electron_ee_distance(qmckl_context context, double* const distance,
                                               const int64_t size_max);
qmckl_exit_code qmckl_get_electron_ee_potential(qmckl_context context, double* const ee_potential);
qmckl_exit_code qmckl_get_electron_en_distance(qmckl_context context, double* const distance,
                                               const int64_t size_max);
qmckl_exit_code qmckl_get_electron_en_potential(qmckl_context context, double* const en_potential);

/* ---- AO basis setters: org/qmckl_ao.org:422-1264 */
qmckl_exit_code qmckl_set_ao_basis_type(qmckl_context context, const char basis_type);
qmckl_exit_code qmckl_set_ao_basis_shell_num(qmckl_context context, const int64_t shell_num);
qmckl_exit_code qmckl_set_ao_basis_prim_num(qmckl_context context, const int64_t prim_num);
qmckl_exit_code qmckl_set_ao_basis_nucleus_shell_num(qmckl_context context, const int64_t* nucleus_shell_num,
                                                     const int64_t size_max);
qmckl_exit_code qmckl_set_ao_basis_nucleus_index(qmckl_context context, const int64_t* nucleus_index,
                                                 const int64_t size_max);
qmckl_exit_code qmckl_set_ao_basis_shell_ang_mom(qmckl_context context, const int32_t* shell_ang_mom,
                                                 const int64_t size_max);
qmckl_exit_code qmckl_set_ao_basis_shell_prim_num(qmckl_context context, const int64_t* shell_prim_num,
                                                  const int64_t size_max);
qmckl_exit_code qmckl_set_ao_basis_shell_prim_index(qmckl_context context, const int64_t* shell_prim_index,
                                                    const int64_t size_max);
qmckl_exit_code qmckl_set_ao_basis_shell_factor(qmckl_context context, const double* shell_factor,
                                                const int64_t size_max);
qmckl_exit_code qmckl_set_ao_basis_exponent(qmckl_context context, const double* exponent,
                                            const int64_t size_max);
qmckl_exit_code qmckl_set_ao_basis_coefficient(qmckl_context context, const double* coefficient,
                                               const int64_t size_max);
qmckl_exit_code qmckl_set_ao_basis_prim_factor(qmckl_context context, const double* prim_factor,
                                               const int64_t size_max);
qmckl_exit_code qmckl_set_ao_basis_ao_num(qmckl_context context, const int64_t ao_num);
qmckl_exit_code qmckl_set_ao_basis_ao_factor(qmckl_context context, const double* ao_factor,
                                             const int64_t size_max);
qmckl_exit_code qmckl_set_ao_basis_r_power(qmckl_context context, const int32_t* r_power,
                                           const int64_t size_max);
qmckl_exit_code qmckl_set_ao_basis_cartesian(qmckl_context context, const bool cartesian);

/* ---- AO basis getters: org/qmckl_ao.org:1483-2248 */
qmckl_exit_code qmckl_get_ao_basis_type(const qmckl_context context, char* const basis_type);
qmckl_exit_code qmckl_get_ao_basis_shell_num(const qmckl_context context, int64_t* const shell_num);
qmckl_exit_code qmckl_get_ao_basis_prim_num(const qmckl_context context, int64_t* const prim_num);
qmckl_exit_code qmckl_get_ao_basis_nucleus_shell_num(const qmckl_context context, int64_t* const nucleus_shell_num,
                                                     const int64_t size_max);
qmckl_exit_code qmckl_get_ao_basis_nucleus_index(const qmckl_context context, int64_t* const nucleus_index,
                                                 const int64_t size_max);
qmckl_exit_code qmckl_get_ao_basis_r_power(const qmckl_context context, int32_t* const r_power,
                                           const int64_t size_max);
qmckl_exit_code qmckl_get_ao_basis_shell_ang_mom(const qmckl_context context, int32_t* const shell_ang_mom,
                                                 const int64_t size_max);
qmckl_exit_code qmckl_get_ao_basis_shell_prim_num(const qmckl_context context, int64_t* const shell_prim_num,
                                                  const int64_t size_max);
qmckl_exit_code qmckl_get_ao_basis_shell_prim_index(const qmckl_context context, int64_t* const shell_prim_index,
                                                    const int64_t size_max);
qmckl_exit_code qmckl_get_ao_basis_shell_factor(const qmckl_context context, double* const shell_factor,
                                                const int64_t size_max);
qmckl_exit_code qmckl_get_ao_basis_exponent(const qmckl_context context, double* const exponent,
                                            const int64_t size_max);
qmckl_exit_code qmckl_get_ao_basis_coefficient(const qmckl_context context, double* const coefficient,
                                               const int64_t size_max);
qmckl_exit_code qmckl_get_ao_basis_prim_factor(const qmckl_context context, double* const prim_factor,
                                               const int64_t size_max);
qmckl_exit_code qmckl_get_ao_basis_ao_num(const qmckl_context context, int64_t* const ao_num);
qmckl_exit_code qmckl_get_ao_basis_ao_factor(const qmckl_context context, double* const ao_factor,
                                             const int64_t size_max);
bool qmckl_ao_basis_provided(const qmckl_context context);

/* ---- AO values: org/qmckl_ao.org:3417, 3486, 3547, 3609, 3675, 3736 */
qmckl_exit_code qmckl_get_ao_basis_primitive_vgl(qmckl_context context, double* const primitive_vgl,
                                                 const int64_t size_max);
qmckl_exit_code qmckl_get_ao_basis_shell_vgl(qmckl_context context, double* const shell_vgl,
                                             const int64_t size_max);
qmckl_exit_code qmckl_get_ao_basis_ao_vgl(qmckl_context context, double* const ao_vgl, const int64_t size_max);
qmckl_exit_code qmckl_get_ao_basis_ao_vgl_inplace(qmckl_context context, double* const ao_vgl,
                                                  const int64_t size_max);
qmckl_exit_code qmckl_get_ao_basis_ao_value(qmckl_context context, double* const ao_value,
                                            const int64_t size_max);
qmckl_exit_code qmckl_get_ao_basis_ao_value_inplace(qmckl_context context, double* const ao_value,
                                                    const int64_t size_max);

/* ---- MO basis: org/qmckl_mo.org:192, 705, 742, 799, 873, 992, 1048, 1628, 1684, 3479 */
qmckl_exit_code qmckl_set_mo_basis_mo_num(qmckl_context context, const int64_t mo_num);
qmckl_exit_code qmckl_set_mo_basis_coefficient(qmckl_context context, const double* coefficient,
                                               const int64_t size_max);
qmckl_exit_code qmckl_get_mo_basis_mo_num(const qmckl_context context, int64_t* mo_num);
qmckl_exit_code qmckl_get_mo_basis_coefficient(const qmckl_context context, double* const coefficient,
                                               const int64_t size_max);
bool qmckl_mo_basis_provided(const qmckl_context context);
qmckl_exit_code qmckl_mo_basis_select_mo(const qmckl_context context, const int32_t* keep,
                                         const int64_t size_max);
qmckl_exit_code qmckl_get_mo_basis_mo_value(qmckl_context context, double* const mo_value,
                                            const int64_t size_max);
qmckl_exit_code qmckl_get_mo_basis_mo_value_inplace(qmckl_context context, double* const mo_value,
                                                    const int64_t size_max);
qmckl_exit_code qmckl_get_mo_basis_mo_vgl(qmckl_context context, double* const mo_vgl, const int64_t size_max);
qmckl_exit_code qmckl_get_mo_basis_mo_vgl_inplace(qmckl_context context, double* const mo_vgl,
                                                  const int64_t size_max);
qmckl_exit_code qmckl_mo_basis_rescale(qmckl_context context, const double scaling_factor);

/* ---- CHAMP Jastrow parameters: org/qmckl_jastrow_champ.org:531, 1439, 1462 */
qmckl_exit_code qmckl_set_jastrow_champ_rescale_factor_ee(qmckl_context context, const double kappa_ee);
qmckl_exit_code qmckl_set_jastrow_champ_rescale_factor_en(qmckl_context context, const double* kappa_en,
                                                          const int64_t size_max);
qmckl_exit_code qmckl_set_jastrow_champ_aord_num(qmckl_context context, const int64_t aord_num);
qmckl_exit_code qmckl_set_jastrow_champ_bord_num(qmckl_context context, const int64_t bord_num);
qmckl_exit_code qmckl_set_jastrow_champ_cord_num(qmckl_context context, const int64_t cord_num);
qmckl_exit_code qmckl_set_jastrow_champ_type_nucl_num(qmckl_context context, const int64_t type_nucl_num);
qmckl_exit_code qmckl_set_jastrow_champ_type_nucl_vector(qmckl_context context, const int64_t* type_nucl_vector,
                                                         const int64_t size_max);
qmckl_exit_code qmckl_set_jastrow_champ_a_vector(qmckl_context context, const double* a_vector,
                                                 const int64_t size_max);
qmckl_exit_code qmckl_set_jastrow_champ_b_vector(qmckl_context context, const double* b_vector,
                                                 const int64_t size_max);
qmckl_exit_code qmckl_set_jastrow_champ_c_vector(qmckl_context context, const double* c_vector,
                                                 const int64_t size_max);
qmckl_exit_code qmckl_set_jastrow_champ_spin_independent(qmckl_context context, const int32_t spin_independent);
qmckl_exit_code qmckl_get_jastrow_champ_aord_num(qmckl_context context, int64_t* const aord_num);
qmckl_exit_code qmckl_get_jastrow_champ_bord_num(qmckl_context context, int64_t* const bord_num);
qmckl_exit_code qmckl_get_jastrow_champ_cord_num(qmckl_context context, int64_t* const bord_num);
qmckl_exit_code qmckl_get_jastrow_champ_type_nucl_num(qmckl_context context, int64_t* const type_nucl_num);
qmckl_exit_code qmckl_get_jastrow_champ_type_nucl_vector(qmckl_context context, int64_t* const type_nucl_num,
                                                         const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_a_vector(qmckl_context context, double* const a_vector,
                                                 const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_b_vector(qmckl_context context, double* const b_vector,
                                                 const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_c_vector(qmckl_context context, double* const c_vector,
                                                 const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_rescale_factor_ee(const qmckl_context context,
                                                          double* const rescale_factor_ee);
qmckl_exit_code qmckl_get_jastrow_champ_rescale_factor_en(const qmckl_context context,
                                                          double* const rescale_factor_en, const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_dim_c_vector(qmckl_context context, int64_t* const dim_c_vector);
qmckl_exit_code qmckl_get_jastrow_champ_spin_independent(qmckl_context context, int32_t* const spin_independent);
bool qmckl_jastrow_champ_provided(const qmckl_context context);

/* ---- CHAMP Jastrow intermediates and results:
 *      org/qmckl_jastrow_champ.org:2143, 2553, 2944, 3405, 4013, 5677, 5931, 6356, 6839, 7323,
 *      8938, 9438, 10556, 10935, 13836, 14394, 16409, 16767, 16823 */
qmckl_exit_code qmckl_get_jastrow_champ_asymp_jasb(qmckl_context context, double* const asymp_jasb,
                                                   const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_asymp_jasa(qmckl_context context, double* const asymp_jasa,
                                                   const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_ee_distance_rescaled(qmckl_context context, double* const distance_rescaled,
                                                             int64_t const size_max);
qmckl_exit_code qmckl_get_jastrow_champ_ee_distance_rescaled_gl(qmckl_context context,
                                                                double* const distance_rescaled_gl,
                                                                const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_en_distance_rescaled(qmckl_context context, double* const distance_rescaled,
                                                             const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_en_distance_rescaled_gl(qmckl_context context,
                                                                double* const distance_rescaled_gl,
                                                                const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_een_rescaled_e(qmckl_context context, double* const een_rescaled_e,
                                                       const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_een_rescaled_e_gl(qmckl_context context, double* const een_rescaled_e_gl,
                                                          const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_een_rescaled_n(qmckl_context context, double* const een_rescaled_n,
                                                       const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_een_rescaled_n_gl(qmckl_context context, double* const een_rescaled_n_gl,
                                                          const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_tmp_c(qmckl_context context, double* const tmp_c, const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_dtmp_c(qmckl_context context, double* const dtmp_c, const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_factor_ee(qmckl_context context, double* const factor_ee,
                                                  const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_factor_ee_gl(qmckl_context context, double* const factor_ee_gl,
                                                     const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_factor_en(qmckl_context context, double* const factor_en,
                                                  const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_factor_en_gl(qmckl_context context, double* const factor_en_gl,
                                                     const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_factor_een(qmckl_context context, double* const factor_een,
                                                   const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_factor_een_gl(qmckl_context context, double* const factor_een_gl,
                                                      const int64_t size_max);
/* org/qmckl_jastrow_champ.org:14403 */
qmckl_exit_code qmckl_get_jastrow_champ_factor_een_grad(qmckl_context context, double* const factor_een_grad,
                                                        const int64_t size_max);
/* The value and factor getters evaluate the values alone (no GL; one of the five e-e-n
 * products per step, where the LDS-resident e-e-n kernel applies) unless the caller asked for
 * the GL after its previous value request, in which case they run the full evaluation that
 * the GL request will read; the values are the same bit for bit either way. */
qmckl_exit_code qmckl_get_jastrow_champ_value(qmckl_context context, double* const value, const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_gl(qmckl_context context, double* const gl, const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_grad(qmckl_context context, double* const grad, const int64_t size_max);

/* ---- CHAMP Jastrow parameter derivatives: org/qmckl_jastrow_champ.org:4807, 5031, 5332, 8041,
 *      8269, 8573, 15751, 16055.  Summed over walkers and divided by walk_num for the e-e
 *      and e-n terms, summed over walkers for the e-e-n term, as the reference.
 *      asymp_jasb_pderiv [2][bord+1]      asymp_jasa_pderiv [type][aord+1]
 *      factor_ee_pderiv  [bord+1]         factor_ee_gl_pderiv  [bord+1][elec][4]
 *      factor_en_pderiv  [type][aord+1]   factor_en_gl_pderiv  [type][aord+1][elec][4]
 *      factor_een_pderiv [type][dim_c]    factor_een_gl_pderiv [type][dim_c][elec][4] */
qmckl_exit_code qmckl_get_jastrow_champ_asymp_jasb_pderiv(qmckl_context context, double* const asymp_jasb_pderiv,
                                                          const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_asymp_jasa_pderiv(qmckl_context context, double* const asymp_jasa_pderiv,
                                                          const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_factor_ee_pderiv(qmckl_context context, double* const factor_ee_pderiv,
                                                         const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_factor_ee_gl_pderiv(qmckl_context context, double* const factor_ee_gl_pderiv,
                                                            const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_factor_en_pderiv(qmckl_context context, double* const factor_en_pderiv,
                                                         const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_factor_en_gl_pderiv(qmckl_context context, double* const factor_en_gl_pderiv,
                                                            const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_factor_een_pderiv(qmckl_context context, double* const factor_een_pderiv,
                                                          const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_factor_een_gl_pderiv(qmckl_context context,
                                                             double* const factor_een_gl_pderiv,
                                                             const int64_t size_max);

/* ---- Forces of the Jastrow (derivatives with respect to the nuclear coordinates):
 *      org/qmckl_forces.org:566, 947, 1372, 2984, 3715, 4079.
 *      forces_jastrow_en   [walk][nucl][3]              d J_en / d R_a
 *      forces_jastrow_en_g [walk][nucl][3][elec][3]     d (grad_i J_en)_m / d R_a,l  as [w][a][l][i][m]
 *      forces_jastrow_en_l [walk][nucl][3]              d (sum_i lap_i J_en) / d R_a
 *      forces_jastrow_een  [walk][nucl][3]
 *      forces_jastrow_een_g [walk][3][nucl][3][elec]    d (grad_i J_een)_m / d R_a,l as [w][l][a][m][i]
 *      forces_jastrow_een_l [walk][3][nucl] */
qmckl_exit_code qmckl_get_forces_jastrow_en(qmckl_context context, double* const forces_jastrow_en,
                                            const int64_t size_max);
qmckl_exit_code qmckl_get_forces_jastrow_en_g(qmckl_context context, double* const forces_jastrow_en_g,
                                              const int64_t size_max);
qmckl_exit_code qmckl_get_forces_jastrow_en_l(qmckl_context context, double* const forces_jastrow_en_l,
                                              const int64_t size_max);
qmckl_exit_code qmckl_get_forces_jastrow_een(qmckl_context context, double* const forces_jastrow_een,
                                             const int64_t size_max);
qmckl_exit_code qmckl_get_forces_jastrow_een_g(qmckl_context context, double* const forces_jastrow_een_g,
                                               const int64_t size_max);
qmckl_exit_code qmckl_get_forces_jastrow_een_l(qmckl_context context, double* const forces_jastrow_een_l,
                                               const int64_t size_max);

/* ---- Single-electron moves of the CHAMP Jastrow: org/qmckl_jastrow_champ_single.org:452,
 * 589, 629, 904, 1240, 1572, 1868, 2741, 3058, 3682, 4350, 5219, 5513, 5881, 6229, 7088, 7381,
 * 7754, 8081, 8906.  qmckl_set_single_point takes the new position of electron num (0-based)
 * for every walker, [walk_num][3] ('N') or [3][walk_num] ('T'); the getters return, for that
 * move, the distances and rescaled distances of the new position and the changes of the e-e,
 * e-n and e-e-n factors and of their gradients and Laplacians (delta_ee_gl, delta_en_gl
 * [walk_num][elec_num][4]; delta_een_gl [walk_num][4][elec_num]; delta_een_g the moved
 * electron's gradient alone).  qmckl_get_jastrow_champ_single_accept moves the electron. */
qmckl_exit_code qmckl_set_single_point(qmckl_context context, const char transp, const int64_t num,
                                       const double* coord, const int64_t size_max);
qmckl_exit_code qmckl_single_touch(const qmckl_context context);
qmckl_exit_code qmckl_get_single_electron_ee_distance(qmckl_context context, double* const distance,
                                                      const int64_t size_max);
qmckl_exit_code qmckl_get_single_electron_en_distance(qmckl_context context, double* distance,
                                                      const int64_t size_max);
qmckl_exit_code qmckl_get_een_rescaled_single_e(qmckl_context context, double* const distance_rescaled,
                                                const int64_t size_max);
qmckl_exit_code qmckl_get_een_rescaled_single_n(qmckl_context context, double* const distance_rescaled,
                                                const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_single_een(qmckl_context context, double* const delta_een,
                                                   const int64_t size_max);
qmckl_exit_code qmckl_get_een_rescaled_single_n_gl(qmckl_context context, double* const distance_rescaled,
                                                   const int64_t size_max);
qmckl_exit_code qmckl_get_een_rescaled_single_e_gl(qmckl_context context, double* const distance_rescaled,
                                                   const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_single_een_gl(qmckl_context context, double* const delta_een_gl,
                                                      const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_single_een_g(qmckl_context context, double* const delta_een_g,
                                                     const int64_t size_max);
qmckl_exit_code qmckl_get_ee_rescaled_single(qmckl_context context, double* const distance_rescaled,
                                             const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_single_ee(qmckl_context context, double* const delta_ee,
                                                  const int64_t size_max);
qmckl_exit_code qmckl_get_ee_rescaled_single_gl(qmckl_context context, double* const distance_rescaled_gl,
                                                const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_single_ee_gl(qmckl_context context, double* const delta_ee_gl,
                                                     const int64_t size_max);
qmckl_exit_code qmckl_get_en_rescaled_single(qmckl_context context, double* const distance_rescaled,
                                             const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_single_en(qmckl_context context, double* const delta_en,
                                                  const int64_t size_max);
qmckl_exit_code qmckl_get_en_rescaled_single_gl(qmckl_context context, double* distance_rescaled_gl,
                                                const size_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_single_en_gl(qmckl_context context, double* const delta_en_gl,
                                                     const int64_t size_max);
qmckl_exit_code qmckl_get_jastrow_champ_single_accept(qmckl_context context);

/* ---- TREXIO input (org/qmckl_trexio.org:1104-1190) ---------------------------------------
 * Sets the electron numbers, nuclei, AO basis and MO coefficients of `context` from the TREXIO
 * file `file_name` (at most size_max characters).  The TEXT back end (a directory of
 * <group>.txt files) is read without libtrexio; any other path fails with QMCKL_INVALID_ARG_2,
 * as the reference does when no back end opens the file.  Inconsistent index data return
 * QMCKL_FAILURE.  An invalid context returns QMCKL_INVALID_CONTEXT (the reference returns 0). */
qmckl_exit_code qmckl_trexio_read(const qmckl_context context, const char* file_name, const int64_t size_max);

/* ---- Orbital Hessians, MO cusp correction, forces on the orbitals ----------------------
 * shell Hessian [point][3][4][shell] and AO Hessian [3][point][4][ao] ([i][j] = d_i d_j for
 * j < 3, [i][3] = d_i of the Laplacian): org/qmckl_ao.org:5268-5305, 9397-9435.
 * MO Hessian [3][point][3][mo]: org/qmckl_mo.org:2366-2420.
 * qmckl_set_mo_basis_r_cusp (org/qmckl_mo.org:192-196, 479-700): per-nucleus radii inside which
 * the nucleus' s AOs are replaced by a fitted cubic in r; every MO value/VGL/Hessian getter then
 * returns the corrected MOs (fitted again by qmckl_mo_basis_select_mo).
 * Forces (derivatives in the nuclear coordinates): AO values [nucl][point][3][ao], MO values
 * [nucl][point][3][mo], MO gradients [nucl][3][point][3][mo], MO Laplacians
 * [nucl][3][point][mo]: org/qmckl_forces.org:5127-5198, 5473-5535, 5917-5935, 6497-6535. */
qmckl_exit_code qmckl_get_ao_basis_shell_hessian(qmckl_context context, double* const shell_hessian,
                                                 const int64_t size_max);
qmckl_exit_code qmckl_get_ao_basis_ao_hessian(qmckl_context context, double* const ao_hessian, const int64_t size_max);
qmckl_exit_code qmckl_get_mo_basis_mo_hessian(qmckl_context context, double* const mo_hessian, const int64_t size_max);
qmckl_exit_code qmckl_set_mo_basis_r_cusp(qmckl_context context, const double* r_cusp, const int64_t size_max);
qmckl_exit_code qmckl_get_forces_ao_value(qmckl_context context, double* const forces_ao_value, const int64_t size_max);
qmckl_exit_code qmckl_get_forces_mo_value(qmckl_context context, double* const forces_mo_value, const int64_t size_max);
qmckl_exit_code qmckl_get_forces_mo_value_inplace(qmckl_context context, double* const forces_mo_value,
                                                  const int64_t size_max);
qmckl_exit_code qmckl_get_forces_mo_g(qmckl_context context, double* const forces_mo_g, const int64_t size_max);
qmckl_exit_code qmckl_get_forces_mo_g_inplace(qmckl_context context, double* const forces_mo_g, const int64_t size_max);
qmckl_exit_code qmckl_get_forces_mo_l(qmckl_context context, double* const forces_mo_l, const int64_t size_max);

/* ---- Slater determinants (org/under_development/qmckl_determinant.org:344-556, 569-640,
 * 1308-1495).  The reference does not build this module; names and argument lists are its own.
 * mo_index_alpha/beta: [det_num][walk_num][up_num / down_num], 1-based MO indices, set after
 * the walkers (qmckl_set_electron_coord).  Getters without size_max write (like the reference)
 * det_vgl [det][walk][5][n][n] (component c of MO mo_index[d][w][m] - 1 at electron e, [e][m]
 * inner), det_inv_matrix and det_adj_matrix [det][walk][n][n] (the inverse of that value
 * matrix, NaN when it is singular, and det times it), det [det][walk]. */
qmckl_exit_code qmckl_set_determinant_type(const qmckl_context context, const char t);
qmckl_exit_code qmckl_set_determinant_det_num_alpha(const qmckl_context context, const int64_t det_num_alpha);
qmckl_exit_code qmckl_set_determinant_det_num_beta(const qmckl_context context, const int64_t det_num_beta);
qmckl_exit_code qmckl_set_determinant_mo_index_alpha(const qmckl_context context, const int64_t* mo_index_alpha,
                                                     const int64_t size_max);
qmckl_exit_code qmckl_set_determinant_mo_index_beta(const qmckl_context context, const int64_t* mo_index_beta,
                                                    const int64_t size_max);
bool qmckl_determinant_provided(const qmckl_context context);
qmckl_exit_code qmckl_get_det_vgl_alpha(qmckl_context context, double* const det_vgl_alpha);
qmckl_exit_code qmckl_get_det_vgl_beta(qmckl_context context, double* const det_vgl_beta);
qmckl_exit_code qmckl_get_det_inv_matrix_alpha(qmckl_context context, double* const det_inv_matrix_alpha);
qmckl_exit_code qmckl_get_det_inv_matrix_beta(qmckl_context context, double* const det_inv_matrix_beta);
qmckl_exit_code qmckl_get_det_adj_matrix_alpha(qmckl_context context, double* const det_adj_matrix_alpha);
qmckl_exit_code qmckl_get_det_adj_matrix_beta(qmckl_context context, double* const det_adj_matrix_beta);
qmckl_exit_code qmckl_get_det_alpha(qmckl_context context, double* const det_value_alpha);
qmckl_exit_code qmckl_get_det_beta(qmckl_context context, double* const det_value_beta);

/* ---- Local energy of one determinant per spin (org/under_development/qmckl_local_energy.org:
 * 203-1713): kinetic -1/2 sum_e lap_e D / D, potential V_ee + V_en + V_nn, local energy their
 * sum, all [walk_num]; drift vector 2 grad_e D / D [walk_num][elec_num][3].  QMCKL_FAILURE
 * when det_num_alpha or det_num_beta is not 1. */
bool qmckl_local_energy_provided(const qmckl_context context);
qmckl_exit_code qmckl_get_kinetic_energy(qmckl_context context, double* const kinetic_energy);
qmckl_exit_code qmckl_get_potential_energy(qmckl_context context, double* const potential_energy);
qmckl_exit_code qmckl_get_local_energy(qmckl_context context, double* const local_energy, const int64_t size_max);
qmckl_exit_code qmckl_get_drift_vector(qmckl_context context, double* const drift_vector);

/* ---- BLAS-like operations (org/qmckl_blas.org).  Column-major (Fortran-order) operands, as the
 * reference's; host pointers; computed on the context's GPU.
 * qmckl_dgemm (:984-999): C = alpha op(A) op(B) + beta C, op = 'N' or 'T', C m x n; lda, ldb,
 *   ldc at least the rows of A, B, C as stored; beta = 0 does not read C.
 * qmckl_dgemv (:3210-3225): Y = alpha op(A) X + beta Y, A m x n, strides incx, incy > 0.
 * qmckl_dger (:3340-3352): A = alpha X Y^T + A, A m x n.
 * qmckl_adjugate (:1710-1718): B = adj(A) of the n x n matrix A, *det_l = det(A); orders up to
 *   5 by cofactors (singular matrices included), above through the inverse (a singular A
 *   there returns QMCKL_FAILURE with *det_l = 0 and B unchanged). */
qmckl_exit_code qmckl_dgemm(const qmckl_context context, const char TransA, const char TransB, const int64_t m,
                            const int64_t n, const int64_t k, const double alpha, const double* A, const int64_t lda,
                            const double* B, const int64_t ldb, const double beta, double* const C,
                            const int64_t ldc);
qmckl_exit_code qmckl_dgemv(const qmckl_context context, const char Trans, const int64_t m, const int64_t n,
                            const double alpha, const double* A, const int64_t lda, const double* X,
                            const int64_t incx, const double beta, double* const Y, const int64_t incy);
qmckl_exit_code qmckl_dger(const qmckl_context context, const int64_t m, const int64_t n, const double alpha,
                           const double* X, const int64_t incx, const double* Y, const int64_t incy, double* const A,
                           const int64_t lda);
qmckl_exit_code qmckl_adjugate(const qmckl_context context, const int64_t n, const double* A, const int64_t lda,
                               double* const B, const int64_t ldb, double* det_l);
/* The size-safe proxies (:1276-1294, :2595-2605): the array sizes (in elements) are checked
 * against what the operation spans, then as qmckl_dgemm / qmckl_adjugate, the error codes
 * numbered by these functions' own argument positions. */
qmckl_exit_code qmckl_dgemm_safe(const qmckl_context context, const char TransA, const char TransB, const int64_t m,
                                 const int64_t n, const int64_t k, const double alpha, const double* A,
                                 const int64_t size_max_A, const int64_t lda, const double* B, const int64_t size_max_B,
                                 const int64_t ldb, const double beta, double* const C, const int64_t size_max_C,
                                 const int64_t ldc);
qmckl_exit_code qmckl_adjugate_safe(const qmckl_context context, const int64_t n, const double* A,
                                    const int64_t size_max_A, const int64_t lda, double* const B,
                                    const int64_t size_max_B, const int64_t ldb, double* det_l);

/* ---- Sherman-Morrison-Woodbury (org/under_development/qmckl_sherman_morrison_woodbury.org:
 * 299-347, 980-1036, 1623-1668, 2269-2314, 2896-2945).  Slater_inv [Dim][LDS]; Updates
 * [N_updates][LDS]; Updates_index 1-based: update l adds Updates[l] to column
 * Updates_index[l] - 1 of the Slater matrix.  determinant may be NULL.  A breakdown (a
 * denominator or 2x2 / 3x3 determinant below `breakdown`) returns QMCKL_FAILURE with
 * Slater_inv and *determinant unchanged.  The _doc names are the same computation.  Each call
 * runs on the context's GPU; qmckl_amd_sm_update_device batches many inverses. */
qmckl_exit_code qmckl_sm_naive(const qmckl_context context, const uint64_t LDS, const uint64_t Dim,
                               const uint64_t N_updates, const double* Updates, const uint64_t* Updates_index,
                               const double breakdown, double* Slater_inv, double* determinant);
qmckl_exit_code qmckl_sm_naive_doc(const qmckl_context context, const uint64_t LDS, const uint64_t Dim,
                                   const uint64_t N_updates, const double* Updates, const uint64_t* Updates_index,
                                   const double breakdown, double* Slater_inv, double* determinant);
qmckl_exit_code qmckl_sm_splitting(const qmckl_context context, const uint64_t LDS, const uint64_t Dim,
                                   const uint64_t N_updates, const double* Updates, const uint64_t* Updates_index,
                                   const double breakdown, double* Slater_inv, double* determinant);
qmckl_exit_code qmckl_sm_splitting_doc(const qmckl_context context, const uint64_t LDS, const uint64_t Dim,
                                       const uint64_t N_updates, const double* Updates,
                                       const uint64_t* Updates_index, const double breakdown, double* Slater_inv,
                                       double* determinant);
/* The halves of the updates that break down are appended to later_updates [.][LDS] and
 * later_index from position *later on, and *later is advanced. */
qmckl_exit_code qmckl_sm_splitting_core(const qmckl_context context, const uint64_t LDS, const uint64_t Dim,
                                        const uint64_t N_updates, const double* Updates,
                                        const uint64_t* Updates_index, const double breakdown, double* Slater_inv,
                                        double* later_updates, uint64_t* later_index, uint64_t* later,
                                        double* determinant);
qmckl_exit_code qmckl_sm_splitting_core_doc(const qmckl_context context, const uint64_t LDS, const uint64_t Dim,
                                            const uint64_t N_updates, const double* Updates,
                                            const uint64_t* Updates_index, const double breakdown,
                                            double* Slater_inv, double* later_updates, uint64_t* later_index,
                                            uint64_t* later, double* determinant);
qmckl_exit_code qmckl_woodbury_2x2(const qmckl_context context, const uint64_t LDS, const uint64_t Dim,
                                   const double* Updates, const uint64_t* Updates_index, const double breakdown,
                                   double* Slater_inv, double* determinant);
qmckl_exit_code qmckl_woodbury_2x2_doc(const qmckl_context context, const uint64_t LDS, const uint64_t Dim,
                                       const double* Updates, const uint64_t* Updates_index, const double breakdown,
                                       double* Slater_inv, double* determinant);
qmckl_exit_code qmckl_woodbury_3x3(const qmckl_context context, const uint64_t LDS, const uint64_t Dim,
                                   const double* Updates, const uint64_t* Updates_index, const double breakdown,
                                   double* Slater_inv, double* determinant);
qmckl_exit_code qmckl_woodbury_3x3_doc(const qmckl_context context, const uint64_t LDS, const uint64_t Dim,
                                       const double* Updates, const uint64_t* Updates_index, const double breakdown,
                                       double* Slater_inv, double* determinant);

/* ======================================================================================
 * Extensions without a counterpart in the reference (device residency, streams, MO GEMM
 * precision).  Device pointers are HIP device addresses on the context's device.
 * ====================================================================================== */

/* Precision of the AO->MO product.  Until it is set, the numerical precision decides:
 * QMCKL_AMD_MO_FP64 (the reference's arithmetic) above 26 bits, QMCKL_AMD_MO_BF16X3 at 26
 * bits or fewer (where the reference's optimised build computes the product in single
 * precision, org/qmckl_mo.org:2028-2040); qmckl_amd_get_mo_precision returns the mode in use.
 * FP32: the AO VGL and the MO coefficients are rounded to FP32 and multiplied on FP32 matrix
 * cores, FP32 sums per 32 AOs added in FP64.  BF16X3: both operands are split into three
 * bfloat16 terms and the six leading cross products are accumulated in FP32 on bf16 matrix
 * cores, added in FP64 per 32 AOs.  AO values, distances and the whole Jastrow stay FP64 in
 * every mode. */
#define QMCKL_AMD_MO_FP64 ((int32_t)0)
#define QMCKL_AMD_MO_FP32 ((int32_t)1)
#define QMCKL_AMD_MO_BF16X3 ((int32_t)2)
qmckl_exit_code qmckl_amd_set_mo_precision(qmckl_context context, const int32_t mode);
qmckl_exit_code qmckl_amd_get_mo_precision(qmckl_context context, int32_t* const mode);
/* Shape of the fused AO->MO product: col_tiles column tiles of 16 MOs per pass over the AOs
 * (1..16; 0 restores the default) on `waves` waves per workgroup (4 for 1..8 tiles, 8 for
 * 5..16).  The default takes every tile in one pass up to 8 tiles on 4 waves, else passes of
 * at most 16 tiles on 8 waves.  A change of shape rebuilds the product's operand layout (a
 * constant-data change).  qmckl_amd_get_mo_tile returns the shape in use and the number of
 * passes. */
qmckl_exit_code qmckl_amd_set_mo_tile(qmckl_context context, const int32_t col_tiles, const int32_t waves);
qmckl_exit_code qmckl_amd_get_mo_tile(qmckl_context context, int32_t* const col_tiles, int32_t* const waves,
                                      int32_t* const passes);
/* Autotuner of the product shape: groups the contexts into size buckets (device, MO
 * precision, column tiles, 32-AO chunks and 16-point tiles rounded up to powers of two),
 * times every built shape (1 to 4 passes on 4 or 8 waves) on the first context of each bucket
 * with its current points (HIP events on its stream; the context's MO VGL buffer is
 * overwritten), and sets the fastest shape on every context of the bucket.  The choice is
 * kept per process: a bucket is timed once.  shapes (NULL or [n][3]): the column tiles,
 * waves and passes chosen for each context.  Needs the MO basis and the points set. */
qmckl_exit_code qmckl_amd_autotune(const qmckl_context* contexts, const int64_t n, int32_t* const shapes);

/* HIP stream (a hipStream_t, or NULL for the context's own stream) used by every kernel
 * of this context. */
qmckl_exit_code qmckl_amd_set_stream(qmckl_context context, void* stream);
/* Waits for all work of the context. */
qmckl_exit_code qmckl_amd_synchronize(qmckl_context context);

/* qmckl_set_electron_coord with coordinates already in device memory. */
qmckl_exit_code qmckl_amd_set_electron_coord_device(qmckl_context context, const char transp,
                                                    const int64_t walk_num, const double* d_coord,
                                                    const int64_t size_max);
/* Results written straight into caller-owned device buffers (same layouts and size_max
 * rules as the host getters; nothing is copied to the host). */
qmckl_exit_code qmckl_amd_get_ao_basis_ao_vgl_device(qmckl_context context, double* const d_ao_vgl,
                                                     const int64_t size_max);
qmckl_exit_code qmckl_amd_get_mo_basis_mo_vgl_device(qmckl_context context, double* const d_mo_vgl,
                                                     const int64_t size_max);
/* MO values [point][mo] alone (qmckl_get_mo_basis_mo_value into device memory): the value
 * rows of a current MO VGL, else a values-only product (no AO derivatives, a fifth of the
 * matrix work) equal bit for bit to the value rows of the full evaluation. */
qmckl_exit_code qmckl_amd_get_mo_basis_mo_value_device(qmckl_context context, double* const d_mo_value,
                                                       const int64_t size_max);
qmckl_exit_code qmckl_amd_get_jastrow_champ_value_device(qmckl_context context, double* const d_value,
                                                         const int64_t size_max);
qmckl_exit_code qmckl_amd_get_jastrow_champ_gl_device(qmckl_context context, double* const d_gl,
                                                      const int64_t size_max);

/* Batched evaluation of several contexts (e.g. molecules of different sizes) replayed from
 * one HIP graph.  qmckl_amd_batch_record evaluates, for every context i, the step
 *   qmckl_amd_set_electron_coord_device(contexts[i], 'T', walk_num_i, d_coord[i], ...)
 *   qmckl_amd_get_mo_basis_mo_vgl_device(..., d_mo_vgl[i], ...)     (skipped if d_mo_vgl is NULL)
 *   qmckl_amd_get_jastrow_champ_value_device(..., d_value[i], ...)  (skipped if d_value is NULL)
 *   qmckl_amd_get_jastrow_champ_gl_device(..., d_gl[i], ...)        (skipped if d_gl is NULL)
 * once eagerly (walk_num_i is the context's current walker count; its coordinates must have
 * been set), then records the same work for all the contexts as grouped launches (one per
 * kernel and shape, each covering every context) into a graph that forks from and joins
 * back to `stream` (a hipStream_t; NULL: a stream of the batch's own);
 * qmckl_amd_batch_replay launches that graph on it, with the same results as the eager
 * steps, bit for bit.  Each replay reads the current contents of the d_coord[i] buffers.
 * The contexts' streams are set to the batch's stream until qmckl_amd_batch_destroy; the
 * contexts must be distinct and outlive the batch, and kernel profiling must be off.
 * qmckl_amd_batch_replay returns QMCKL_FAILURE without launching anything when a context has
 * changed since the recording: its walker, electron or MO count, its MO precision, any
 * constant data (a setter other than the coordinates, a new product shape), or the address
 * or size of any of its device buffers (a getter that allocated or grew one). */
qmckl_exit_code qmckl_amd_batch_record(const qmckl_context* contexts, const int64_t n, void* stream,
                                       const double* const* d_coord, double* const* d_mo_vgl,
                                       double* const* d_value, double* const* d_gl, int64_t* const batch);
qmckl_exit_code qmckl_amd_batch_replay(const int64_t batch);
/* Waits for the batch's replays (its stream; with a NULL `stream` the batch owns one). */
qmckl_exit_code qmckl_amd_batch_synchronize(const int64_t batch);
qmckl_exit_code qmckl_amd_batch_destroy(const int64_t batch);

/* Differentiation (no counterpart in the reference): L = sum_w [weight_j ln Psi_J(w) +
 * weight_el E_L(w)] over the context's walkers, with the Jastrow-only local-energy surrogate
 * E_L = -1/2 sum_i (lap_i J + |grad_i J|^2) + V_ee + V_en + V_nn (J = ln Psi_J), and its exact
 * gradient by a hand-written adjoint kernel: grad_a [type_nucl_num][aord_num+1], grad_b
 * [bord_num+1], grad_c [type_nucl_num][dim_c_vector] (the layouts of the a, b and c vectors) and
 * grad_nucl [nucl_num][3] (d/dR of the nuclear coordinates).  Needs 1 <= aord_num, bord_num <= 8.
 * Host buffers; synchronous. */
qmckl_exit_code qmckl_amd_get_jastrow_champ_adjoint(qmckl_context context, const double weight_j,
                                                    const double weight_el, double* const value,
                                                    double* const grad_a, const int64_t size_a,
                                                    double* const grad_b, const int64_t size_b,
                                                    double* const grad_c, const int64_t size_c,
                                                    double* const grad_nucl, const int64_t size_nucl);

/* No counterpart in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The Jastrow value
 * e^J [walk_num] and GL [walk_num][4][elec_num] of walk_num walkers at the device coordinates
 * d_coord ('T' layout [3][walk_num*elec_num]), with the context's parameters, into the device
 * buffers d_value and d_gl; the same kernels and results as qmckl_amd_get_jastrow_champ_value_device
 * and _gl_device.  It calls no allocating routine: every intermediate goes to a 64 MiB static
 * device array of the library (one per GPU, held until the process ends), and walk_num larger
 * than that scratch holds answers QMCKL_INVALID_ARG_2.  The context's Jastrow tables must
 * already be on the device and current (any device evaluation of its Jastrow uploads them;
 * otherwise QMCKL_NOT_PROVIDED).  The context's own coordinates and cached results are not
 * touched.  Runs on the context's stream; calls that overlap on one GPU share the scratch
 * and race. */
qmckl_exit_code qmckl_amd_get_jastrow_champ_device_noalloc(qmckl_context context, const int64_t walk_num,
                                                           const double* d_coord, const int64_t coord_size,
                                                           double* const d_value, double* const d_gl,
                                                           const int64_t gl_size);

/* Kernel timing: when profiling is on, every kernel launch of the context is bracketed by
 * HIP events on the context's stream.  qmckl_amd_get_kernel_times waits for the stream,
 * returns the accumulated milliseconds and launch counts per slot and resets them.
 * Slots: 0 AO VGL, 1 AO->MO GEMM, 2 Jastrow e-e/e-n, 3 Jastrow e-e-n, 4 Jastrow totals. */
#define QMCKL_AMD_KERNEL_SLOTS 5
qmckl_exit_code qmckl_amd_set_profiling(qmckl_context context, const int32_t on);
qmckl_exit_code qmckl_amd_get_kernel_times(qmckl_context context, double* const ms, int64_t* const counts,
                                           const int64_t size_max);

/* Batched Sherman-Morrison-Woodbury on n_mat device-resident inverses (no counterpart in the
 * reference: its kernels take one matrix).  d_slater_inv [n_mat][Dim][LDS], d_updates
 * [n_mat][N][LDS], d_updates_index [n_mat][N] (1-based), d_determinant [n_mat] (or NULL),
 * d_status [n_mat]: 0 applied, 1 breakdown, 2 an index outside 1..Dim; a matrix and its
 * determinant are written only when its status is 0.  N is 2 or 3 for the Woodbury kinds.
 * Ordered on the context's stream: call qmckl_amd_synchronize before reading the results. */
#define QMCKL_AMD_SM_NAIVE ((int32_t)0)
#define QMCKL_AMD_SM_SPLITTING ((int32_t)1)
#define QMCKL_AMD_WOODBURY_2X2 ((int32_t)3)
#define QMCKL_AMD_WOODBURY_3X3 ((int32_t)4)
qmckl_exit_code qmckl_amd_sm_update_device(qmckl_context context, const int32_t kind, const int64_t n_mat,
                                           const uint64_t LDS, const uint64_t Dim, const uint64_t N_updates,
                                           const double* d_updates, const uint64_t* d_updates_index,
                                           const double breakdown, double* d_slater_inv, double* d_determinant,
                                           int32_t* d_status);

#ifdef __cplusplus
}
#endif

#endif /* QMCKL_AMD_H */
