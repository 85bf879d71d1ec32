// Internal state of a context and the interface between the host runtime (api_*.cpp)
// and the HIP kernels (kernels_*.hip).  Not part of the public C-ABI.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/qmckl.h"

namespace qmckl_amd {

// Growable device allocation owned by a context.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  // Makes the buffer at least `n` bytes; the contents are not kept.
  hipError_t reserve(size_t n);
  void release();
  template <class T> T* as() const { return static_cast<T*>(ptr); }
};

// Basis tables laid out for the AO kernel (built when the basis is complete).
struct AoTables {
  int64_t nucl_num = 0, shell_num = 0, prim_num = 0, ao_num = 0;
  int32_t lmax = 0;
  // per shell
  std::vector<int32_t> shell_nucl, shell_l, shell_prim_start, shell_prim_count;
  std::vector<int32_t> shell_ao_start;
  // per primitive (shell order)
  std::vector<double> prim_expo, prim_coef;  // coef = coefficient*prim_factor*shell_factor
  // per AO
  std::vector<int32_t> ao_shell;
  std::vector<int32_t> ao_pow;  // a | b<<8 | c<<16
  std::vector<double> ao_factor;
  // per nucleus
  std::vector<double> nucl_range;  // max (r_power + 1)/alpha over its primitives
  std::vector<int32_t> nucl_max_l;
  int32_t chunk_shells_max = 1;  // most shells met by the AOs of one 32-AO chunk
  bool slater = false;            // 'S' basis: Slater shells
  std::vector<int32_t> shell_rpow;  // r_power per shell
};

// Device copies of the constant data (uploaded lazily before the first kernel).
struct DevConst {
  DevBuf shell_nucl, shell_l, shell_prim_start, shell_prim_count, shell_ao_start, shell_rpow;
  DevBuf prim_expo, prim_coef, ao_shell, ao_pow, ao_factor, nucl_range;
  DevBuf nucl_coord;  // [3][nucl]
  DevBuf coef_t;      // MO coefficients transposed [k_pad][mo_pad] (FP64, zero-padded)
  // C_t in the per-chunk operand layouts of the fused AO->MO kernel, [group][chunk][...]:
  // FP64 / FP32 [32][16 G], BF16X3 [3][16 G][AO_MO_BF_LD] (zero-padded), row k scaled by
  // ao_factor[k]
  DevBuf coef_ch64, coef_ch32, coef_chbf;
  int64_t mo_pad = 0, k_pad = 0;
  int ao_mo_g = 1, ao_mo_ngroup = 1;  // column tiles per pass of the fused kernel, passes
  int ao_mo_nw = 4;                   // its waves per workgroup
  // 32-AO chunk descriptors of the fused kernel (AoMoArgs), offsets in doubles
  DevBuf ao_desc, ao_desc_off;
  int ao_desc_max = 0, ao_nchunk = 0, ao_ns_max = 1;
  DevBuf jas_a, jas_b, jas_cfull, jas_type, jas_kappa_en, jas_asymp_a;
  DevBuf jas_cfull_pad, jas_nuc_pad;  // nucleus tables zero-padded to whole 16-nucleus tiles
  DevBuf jas_charge;                  // nuclear charges (adjoint: V_en)
  DevBuf prim_nucl_c, prim_expo_c;  // caller-order primitives (primitive VGL getter)
  uint64_t date = 0;      // context date of the basis upload
  uint64_t jas_date = 0;  // context date of the Jastrow upload
};

struct ProfEvent {
  int slot;
  hipEvent_t start, stop;
};

struct Context {
  uint32_t tag = 0xBEEFFACEu;
  std::recursive_mutex mutex;
  uint64_t date = 1;        // bumped by every setter
  uint64_t const_date = 1;  // bumped by every setter except the coordinate setters
  // host threads of the reference's OpenMP loops; the evaluation runs on the GPU, so this is
  // only recorded (the host copies of large results pick their own thread count)
  static constexpr int32_t kDefaultNumThreads = 1;
  int32_t num_threads = kDefaultNumThreads;

  // error record, org/qmckl_error.org:433-470
  qmckl_exit_code err_code = QMCKL_SUCCESS;
  std::string err_function, err_message;

  // numprec, org/qmckl_numprec.org:101-147
  int32_t precision = 53, range = 11;

  // point, org/qmckl_point.org:105-109: coord is [3][num]
  int64_t point_num = 0;
  std::vector<double> point_coord;
  bool point_on_device = false;  // last coordinates came from device memory
  uint64_t point_date = 0;

  // nucleus, org/qmckl_nucleus.org
  int64_t nucl_num = 0;
  std::vector<double> nucl_charge, nucl_coord;  // coord [3][num]
  uint32_t nucl_uninit = 0x7;
  bool nucl_provided = false;

  // electron, org/qmckl_electron.org:144-165
  int64_t elec_num = 0, up_num = 0, down_num = 0, walk_num = 0;
  uint32_t elec_uninit = 0x1;
  bool elec_provided = false;

  // AO basis, org/qmckl_ao.org:298-381
  char ao_type = 0;
  int64_t shell_num = 0, prim_num = 0, ao_num = 0;
  std::vector<int64_t> nucleus_index, nucleus_shell_num, shell_prim_num, shell_prim_index;
  std::vector<int32_t> shell_ang_mom, r_power;
  std::vector<double> shell_factor, exponent, coefficient, prim_factor, ao_factor;
  bool ao_cartesian = true;
  uint32_t ao_uninit = (1u << 14) - 1;
  bool ao_provided = false;
  AoTables tables;

  // MO basis, org/qmckl_mo.org:135-160
  int64_t mo_num = 0;
  std::vector<double> mo_coef;  // [mo][ao]
  uint32_t mo_uninit = 0x3;
  bool mo_provided = false;
  // MO product precision set by qmckl_amd_set_mo_precision, or -1: by the numerical
  // precision (mo_mode)
  int32_t mo_precision = -1;
  // requested shape of the fused AO->MO product: column tiles per pass and waves per
  // workgroup (0: the default for the MO count)
  int32_t mo_tile_g = 0, mo_tile_nw = 0;

  // CHAMP Jastrow, org/qmckl_jastrow_champ.org:407-488
  int64_t aord = -1, bord = -1, cord = -1, dim_c = -1, type_nucl_num = -1;
  int32_t spin_independent = -1;
  std::vector<int64_t> type_nucl_vector;
  std::vector<double> a_vector, b_vector, c_vector, rescale_en;
  double rescale_ee = 0.0;
  uint32_t jas_uninit = (1u << 11) - 1;
  bool jas_provided = false;
  double asymp_jasb[2] = {0, 0};
  std::vector<double> asymp_jasa;

  // ---- device side ---------------------------------------------------------------
  int device = -1;
  hipStream_t stream = nullptr;      // stream in use
  hipStream_t own_stream = nullptr;  // created by the context
  // pinned staging of large device-to-host copies (copy_to_host), allocated on first use
  void* stage[2] = {nullptr, nullptr};
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
  DevConst dc;
  DevBuf d_coord;  // [3][point_num]
  uint64_t d_coord_date = 0;

  DevBuf d_ao_vgl;  // [P][5][ao] FP64
  uint64_t ao_vgl_date = 0;
  DevBuf d_ao_work;  // chunk scratch for the AO->MO product
  DevBuf d_mo_vgl;   // [P][5][mo]
  uint64_t mo_vgl_date = 0;
  DevBuf d_mo_value;  // [P][mo]
  uint64_t mo_value_date = 0;

  DevBuf d_ee_dist, d_en_dist;
  uint64_t ee_dist_date = 0, en_dist_date = 0;

  // Jastrow results
  DevBuf d_fee, d_fen, d_feen;           // [W]
  DevBuf d_fee_gl, d_fen_gl, d_feen_gl;  // [W][4][Ne]
  DevBuf d_jvalue, d_jgl;                // [W], [W][4][Ne]
  DevBuf d_jscratch;                     // partial sums
  uint64_t jastrow_date = 0;
  uint64_t jvalue_date = 0;  // the factors and e^J alone (values-only evaluation)
  // dates of the last GL request (any consumer of the full evaluation) and of the last
  // values-only request that evaluated: a caller that asked for the GL after its previous
  // value request gets the full evaluation on the next one
  uint64_t jgl_asked_date = 0, jval_asked_date = 0;
  // Jastrow intermediates (materialised only when asked for)
  DevBuf d_jint;
  // derivative getters: intermediates (JasInter kinds, then e-n distances and the force
  // intermediates), per-walker values, outputs and small tables, each valid at deriv_date
  static constexpr int N_DERIV_BUF = 20;
  DevBuf d_deriv[N_DERIV_BUF];
  uint64_t deriv_date[N_DERIV_BUF] = {};
  uint64_t jint_date = 0;
  int jint_kind = -1;

  // single-electron move, org/qmckl_jastrow_champ_single.org:440-586
  int64_t single_num = -1;
  std::vector<double> single_coord;  // [walk_num][3]
  uint64_t single_date = 0;          // bumped by qmckl_set_single_point and qmckl_single_touch
  DevBuf d_single_coord, d_single_out;
  uint64_t single_out_date = 0, single_out_ctx_date = 0;  // single_date and date of d_single_out

  // Slater determinants, org/under_development/qmckl_determinant.org:115-556
  char det_type = 0;
  int64_t det_num_alpha = 0, det_num_beta = 0;
  std::vector<int64_t> mo_index_alpha, mo_index_beta;  // [det][walk][n], 1-based MO indices
  int64_t mo_index_walk[2] = {0, 0};                   // walk_num each index array was set for
  uint32_t det_uninit = 0x1F;
  uint64_t det_set_date = 1;  // bumped by the determinant setters
  DevBuf d_det_idx;           // int32 0-based indices, alpha then beta
  uint64_t det_idx_date = 0, det_idx_const = 0;  // det_set_date and const_date of d_det_idx
  DevBuf d_det_value, d_det_inv, d_det_work, d_det_out;  // values, inverses, work, getter staging
  DevBuf d_det_drift, d_det_kin, d_energy;               // [W][Ne][3], [2][W], [5][W]
  uint64_t det_date = 0, det_seen = 0, energy_date = 0;
  DevBuf d_pot_charge;
  uint64_t pot_charge_date = 0;
  DevBuf d_smw;  // Sherman-Morrison-Woodbury staging and scratch
  DevBuf d_blas;  // operands of the BLAS-like entry points

  // MO cusp correction, org/qmckl_mo.org:479-700: r_cusp [Nn] (empty: none), cusp_param
  // [Nn][4][mo]; device: r_cusp, params, the s-AO table of each nucleus
  std::vector<double> r_cusp, cusp_param;
  DevBuf d_cusp;
  int64_t cusp_s_total = 0;
  bool cusp_uploaded = false;
  // orbital Hessians and forces: AO Hessian [3][P][4][ao] (valid at d_ao_hess_date), outputs,
  // first AO of each nucleus [Nn + 1]
  DevBuf d_ao_hess, d_orb_out, d_nucl_ao;
  uint64_t ao_hess_date = 0, nucl_ao_date = 0;

  // kernel timing (profiling extension)
  bool profiling = false;
  std::vector<ProfEvent> prof_pending;
  double prof_ms[QMCKL_AMD_KERNEL_SLOTS] = {0};
  int64_t prof_count[QMCKL_AMD_KERNEL_SLOTS] = {0};
};

// Brackets the kernel launches of its scope with HIP events when profiling is on.
struct ProfScope {
  Context* c;
  int slot;
  hipEvent_t start = nullptr;
  ProfScope(Context* ctx, int s);
  ~ProfScope();
};
enum { SLOT_AO = 0, SLOT_MO = 1, SLOT_JPAIR = 2, SLOT_JEEN = 3, SLOT_JTOTAL = 4 };

// ---- helpers shared by the api_*.cpp files ----------------------------------------
// The MO product precision in use: the one set with qmckl_amd_set_mo_precision, else, as the
// reference's optimised build switches its MO product to single precision at precision <= 26
// (org/qmckl_mo.org:2028-2040), the 3xBF16 product there (more accurate than a single-precision
// product and faster) and FP64 above.
inline int32_t mo_mode(const Context* c) {
  if (c->mo_precision >= 0) return c->mo_precision;
  return c->precision <= 26 ? QMCKL_AMD_MO_BF16X3 : QMCKL_AMD_MO_FP64;
}
// A change of constant data (basis, MOs, Jastrow parameters, nuclei, precision).
inline void touch_const(Context* c) {
  c->date += 1;
  c->const_date = c->date;
}
Context* ctx_of(qmckl_context context);  // nullptr if not a live context
// Every device allocation a context owns (the constant tables, inputs, results, scratch).
std::vector<DevBuf*> device_buffers(Context* c);
qmckl_exit_code fail(qmckl_context context, qmckl_exit_code code, const char* fn, const char* msg);
// Makes the device, stream and constant tables ready.
qmckl_exit_code device_ready(Context* c, qmckl_context handle, const char* fn);
qmckl_exit_code upload_coords(Context* c, qmckl_context handle, const char* fn);
qmckl_exit_code hip_fail(qmckl_context handle, hipError_t e, const char* fn);
// Device-to-host copy of `bytes` from d_src (on c's device) into the caller's host memory,
// synchronous.  Below 16 MiB one hipMemcpyAsync on c->stream; above, through c's two
// pinned 32 MiB staging buffers: the device-to-host copy of chunk k + 1 overlaps the host
// threads' copy of chunk k into dst.
hipError_t copy_to_host(Context* c, void* dst, const void* d_src, size_t bytes);

// providers
struct AoDevView;
AoDevView ao_view(Context* c);  // the AO tables of c on its device
qmckl_exit_code ao_ready(Context* c, qmckl_context handle, const char* fn);  // tables and coordinates uploaded
qmckl_exit_code provide_ao_vgl(qmckl_context handle, Context* c, const char* fn);
qmckl_exit_code provide_mo_vgl(qmckl_context handle, Context* c, double* d_out, const char* fn);
struct AoMoArgs;
// The fused product's arguments for c's current points into d_out (tables and coordinates
// uploaded first); mode is c's MO precision.
qmckl_exit_code mo_product_args(qmckl_context handle, Context* c, double* d_out, const char* fn, AoMoArgs& a,
                                int& mode);
qmckl_exit_code provide_jastrow(qmckl_context handle, Context* c, const char* fn);
// The factors and e^J alone, without the GL where the values-only kernel applies
qmckl_exit_code provide_jastrow_value(qmckl_context handle, Context* c, const char* fn);
qmckl_exit_code provide_ee_distance(qmckl_context handle, Context* c, const char* fn);
qmckl_exit_code provide_en_distance(qmckl_context handle, Context* c, const char* fn);
// e_kin, e_pot, e_local, V_ee, V_en per walker in c->d_energy ([5][W]); kinetic: with the
// Slater determinants (api_det.cpp)
qmckl_exit_code provide_energy(Context* c, qmckl_context handle, const char* fn, bool kinetic);
// With r_cusp set, the cusp correction of a device MO VGL [P][5][mo] (hessian = 0) or MO
// Hessian [3][P][3][mo] (hessian = 1) of c's points, in place (api_orb.cpp)
qmckl_exit_code apply_mo_cusp(Context* c, qmckl_context handle, double* out, int hessian, const char* fn);

// ---- kernel launchers (kernels_*.hip) ----------------------------------------------
struct AoDevView {
  int64_t nucl_num, shell_num, ao_num;
  const int32_t *shell_nucl, *shell_l, *shell_prim_start, *shell_prim_count, *shell_ao_start;
  const double *prim_expo, *prim_coef;
  const int32_t *ao_shell, *ao_pow;
  const double *ao_factor, *nucl_range, *nucl_coord;
  double cutoff;
  int slater;                // Slater ('S') shells: r^n e^(-alpha r) radial parts
  const int32_t* shell_rpow;  // r_power per shell (the n of Slater shells)
};
// out_kind: 0 = FP64 [P][5][ao] (ld = ao_num), 1 = FP32, 2 = bf16x3 planes
hipError_t launch_ao_vgl(const AoDevView& v, const double* coord, int64_t coord_ld, int64_t p0,
                         int64_t np, void* out, int out_kind, int64_t ld_out, int64_t plane_stride,
                         int values_only, hipStream_t s);
hipError_t launch_shell_vgl(const AoDevView& v, const double* coord, int64_t P, double* out, hipStream_t s);
hipError_t launch_prim_vgl(const AoDevView& v, int64_t prim_num, const int32_t* prim_nucl,
                           const double* prim_expo, const double* coord, int64_t P, double* out,
                           hipStream_t s);

// C[m][n] = sum_k A[m][k] B[k][n], row-major; A has leading dim lda, B is [K][ldb], C ldc.
hipError_t launch_gemm_f64(const double* A, int64_t lda, const double* B, int64_t ldb, double* C,
                           int64_t ldc, int64_t M, int64_t N, int64_t K, hipStream_t s);
hipError_t launch_gemm_f32(const float* A, int64_t lda, const float* B, int64_t ldb, double* C,
                           int64_t ldc, int64_t M, int64_t N, int64_t K, hipStream_t s);
// A and B given as three bf16 planes each (plane stride in elements).
hipError_t launch_gemm_bf16x3(const uint16_t* A, int64_t lda, int64_t a_plane, const uint16_t* B,
                              int64_t ldb, int64_t b_plane, double* C, int64_t ldc, int64_t M,
                              int64_t N, int64_t K, hipStream_t s);
hipError_t launch_split_bf16x3(const double* src, int64_t n, uint16_t* dst, int64_t plane, hipStream_t s);
// Fused AO VGL -> MO VGL for points p0 .. p0+np-1 (mode 0 FP64, 1 FP32, 2 BF16X3).
// The AOs are taken in chunks of AO_MO_KC; chunk c is described by desc[desc_off[c] ..
// desc_off[c+1]) (doubles, an even count):
//   [ns, na, np, 0], na nuclei [x, y, z, range], ns shells [nucleus + 256 nprim, first prim],
//   np primitives [alpha, coef], AO_MO_KC AOs [shell + 256 (a | b << 8 | c << 16), factor]
// (indices local to the chunk; AOs past ao_num have factor 0; the kernel does not read the
// factor).  B holds C_t chunk by chunk in the operand layout of the mode (see DevConst), its
// row k scaled by ao_factor[k] (zero past ao_num), b_chunk_bytes per (group, chunk); the MO
// columns are taken G tiles of 16 at a time, ngroup groups.  C: [points][5][ldc], columns < N.
constexpr int AO_MO_KC = 32;
constexpr int AO_MO_MAX_G = 16;  // column tiles per pass (8 per 4-wave, 16 per 8-wave workgroup)
constexpr int AO_MO_BF_LD = 40;  // bf16 row of a BF16X3 C_t chunk: 32 AOs + 8 padding
struct AoMoArgs {
  const double* coord;
  int64_t coord_ld, p0, np;
  // the points are taken in the order p' = i tw + w of the point p = w te + i (tw te points:
  // tw walkers of te electrons), so that a workgroup's 16 points are the same electron of
  // consecutive walkers; tw = 1 keeps the points' own order.  p0, np select p'.
  int64_t tw, te;
  const double* desc;
  const int64_t* desc_off;
  int desc_max, ns_max, nchunk;
  const void* B;
  int b_chunk_bytes, G, ngroup;
  int nw;      // waves per workgroup: 4 (G <= 8) or 8 (5 <= G <= 16)
  int slater;  // Slater shells (the shell records carry r_power)
  int vo;      // the MO values alone into C [point][mo] (single launches only, not batches)
  double* C;
  int64_t ldc, N;
  double cutoff;
};
size_t ao_mo_lds_bytes(const AoMoArgs& a, int mode);
hipError_t launch_ao_mo(const AoMoArgs& a, int mode, hipStream_t s);
// The default shape of the product for nct column tiles of 16 MOs: every tile in one pass of
// 4 waves up to 8 tiles, else passes of at most 16 tiles on 8 waves.
void ao_mo_default_shape(int64_t nct, int& G, int& nw);
bool ao_mo_variant_ok(int mode, int G, int nw, int slater = 0);  // a built (G, nw) shape
int ao_mo_waves(int G);  // the default waves per workgroup for G column tiles
// One launch over several contexts' products (AoMoArgs device array), one workgroup per item;
// items packed as context | point tile << 16 | column group << 48, all of one (mode, G, nw).
hipError_t launch_ao_mo_batch(int mode, int G, int nw, int slater, const AoMoArgs* d_args, const uint64_t* d_items,
                              int nitems, size_t lds, hipStream_t s);
// The products of a batch of contexts as one grouped launch per shape (mode, G, nw): built
// once (device argument and item tables allocated), launched per step, freed with the batch.
struct MoBatchLaunch {
  int mode = 0, G = 0, nw = 0, slater = 0;
  AoMoArgs* d_args = nullptr;
  uint64_t* d_items = nullptr;
  int nitems = 0;
  size_t lds = 0;
  double cost = 0.0;  // items x chunks x column tiles, for dealing the launches to streams
};
struct MoBatchPlan {
  std::vector<MoBatchLaunch> launches;
};
hipError_t mo_batch_build(const std::vector<std::pair<AoMoArgs, int>>& in, MoBatchPlan& plan);
// the launches dealt to the streams (n of them), the costliest first to the least loaded
hipError_t mo_batch_launch(const MoBatchPlan& plan, const hipStream_t* s, int n);
void mo_batch_free(MoBatchPlan& plan);
hipError_t launch_to_f32(const double* src, int64_t n, float* dst, hipStream_t s);

struct JasDevView {
  int64_t walk_num, elec_num, up_num, nucl_num, aord, bord, cord, dim_c;
  int64_t type_num;  // nucleus types (rows of a_vector)
  int32_t spin_independent;
  double rescale_ee;
  const double* coord;  // [3][W*Ne]
  int64_t coord_ld;     // = W*Ne
  const double* nucl_coord;  // [3][Nn]
  const int64_t* type_nucl;  // [Nn]
  const double* kappa_en;    // [type]
  const double* a_vector;    // [type][aord+1]
  const double* b_vector;    // [bord+1]
  const double* cfull;       // [dim_c][Nn]
  const double* cfull_pad;   // [dim_c][NA16], NA16 = 16 ceil(Nn / 16); zero for a >= Nn
  const double* nuc_pad;     // [4][NA16]: x, y, z, kappa_a of nucleus a; zero for a >= Nn
  double asymp_jasb0, asymp_jasb1;
  const double* asymp_jasa;  // [type]
};
struct JasOut {
  double *fee, *fen, *feen;           // [W]
  double *fee_gl, *fen_gl, *feen_gl;  // [W][4][Ne]
  double *value, *gl;                 // [W], [W][4][Ne]
  double* scratch;                    // partial sums, size jastrow_scratch_doubles()
};
// One context's Jastrow evaluation as the grouped (batch) kernels take it: view, outputs,
// scratch regions and the e-e-n tile shape.
struct JasJob {
  JasDevView v;
  JasOut o;
  double *R0, *gN, *Sg, *Jt, *JtP;
  int ntiles, TI, JC, have_een, fused;
};

int64_t jastrow_scratch_doubles(const JasDevView& v);
// The Jastrow view of c for walk_num walkers at the device coordinates coord
// [3][walk_num * elec_num], on the device tables of c; false (and nothing uploaded or
// allocated) when those tables are not on the device or older than c's constant data.
bool jastrow_view_resident(Context* c, int64_t walk_num, const double* coord, JasDevView& v);
// The Jastrow view of c's walkers after uploading its tables and coordinates.
qmckl_exit_code jastrow_view_ready(Context* c, qmckl_context handle, const char* fn, JasDevView& v);
// c's device factor, GL and value buffers (valid after provide_jastrow).
JasOut jastrow_out(Context* c);
// The three kernels of one evaluation, in this order (same stream).
hipError_t launch_jastrow_pair(const JasDevView& v, const JasOut& o, hipStream_t s);
hipError_t launch_jastrow_een(const JasDevView& v, const JasOut& o, hipStream_t s);
hipError_t launch_jastrow_total(const JasDevView& v, const JasOut& o, hipStream_t s);
// The factors and e^J alone (no GL) in one resident e-e-n launch and the total; only where
// that kernel applies (jastrow_values_only_ok), else the full evaluation gives them.
bool jastrow_values_only_ok(const JasDevView& v);
// Whether the LDS-resident e-e-n kernel applies (it then also does the pair kernel's work).
bool een_fused(const JasDevView& v);
hipError_t launch_jastrow_values(const JasDevView& v, const JasOut& o, hipStream_t s);
// Grouped Jastrow launches over a batch of contexts (one launch per kernel and cord): built
// once (device job tables allocated), launched per step, freed with the batch.
struct JasBatchLaunch {
  int kind = 0, C = 0;
  JasJob* d_jobs = nullptr;
  int* d_starts = nullptr;
  int njobs = 0, nblocks = 0, threads = 0;
  size_t lds = 0;
};
struct JasBatchPlan {
  std::vector<JasBatchLaunch> launches;
};
hipError_t jastrow_batch_build(const std::vector<std::pair<JasDevView, JasOut>>& in, JasBatchPlan& plan);
// the resident e-e-n kernel on s_res, everything else on s; s waits for s_res (res_done)
// before the totals
hipError_t jastrow_batch_launch(const JasBatchPlan& plan, hipStream_t s, hipStream_t s_res, hipEvent_t res_done);
void jastrow_batch_free(JasBatchPlan& plan);
// Gradient of L = sum_w [wJ J_w + wL E_w] (E_w = -1/2 sum_i (lap_i J + |grad_i J|^2) + V_ee
// + V_en; V_nn is added by the caller) in the Jastrow parameters and nuclear coordinates,
// per walker into out [W][nout]: [0] the walker's value, then d/da [type][aord+1], d/db
// [bord+1], d/dc [type][dim_c], d/dR [nucl][3] (kernels_adjoint.hip).  The factors and
// factor GLs of the same coordinates must be current.
struct AdjArgs {
  JasDevView v;
  int ntype;
  const double *fee, *fen, *feen;           // [W]
  const double *fee_gl, *fen_gl, *feen_gl;  // [W][4][Ne]
  const double* charge;                     // [nucl]
  double wJ, wL;
  double* out;
  int nout;
};
int adjoint_outputs(const JasDevView& v, int ntype);
hipError_t launch_jastrow_adjoint(const AdjArgs& a, hipStream_t s);
// Single-electron moves (kernels_single.hip): outputs for the move of electron `num` of
// every walker to coord [W][3] (device), in the layouts of the reference's getters
// (org/qmckl_jastrow_champ_single.org): distances and rescaled distances of the new
// position (ee [W][Ne], en [W][Nn], their GLs [W][Ne][4], [W][Nn][4]), een_rescaled
// single e [W][C+1][Ne] / e_gl [W][C+1][4][Ne] / n [W][C+1][Nn] / n_gl [W][C+1][Nn][4], and
// the changes of the factors [W] and of their GLs (ee, en: [W][Ne][4]; een: [W][4][Ne]; een_g:
// the moved electron's gradient alone, [W][4][Ne]).
struct SingleOut {
  double *ee_dist, *en_dist, *ee_resc, *ee_resc_gl, *en_resc, *en_resc_gl;
  double *een_e, *een_e_gl, *een_n, *een_n_gl;
  double *d_ee, *d_en, *d_een, *d_ee_gl, *d_en_gl, *d_een_gl, *d_een_g;
};
struct SingleArgs {
  int64_t num;
  const double* coord;  // [W][3]
  SingleOut out;
};
int64_t single_out_doubles(const JasDevView& v);
SingleOut single_out_view(const JasDevView& v, double* base);
hipError_t launch_jastrow_single(const JasDevView& v, const SingleArgs& s, hipStream_t st);
// Moves electron num of every walker in coord [3][W*Ne] and adds the changes to the
// factors and factor GLs of o, then forms its value and GL again.
hipError_t launch_jastrow_single_accept(const JasDevView& v, const SingleArgs& s, const JasOut& o, double* coord,
                                        hipStream_t st);
// Intermediates with the reference layouts; kind selects which one.
enum JasInter {
  J_EE_RESCALED = 0, J_EE_RESCALED_GL, J_EN_RESCALED, J_EN_RESCALED_GL,
  J_EEN_E, J_EEN_E_GL, J_EEN_N, J_EEN_N_GL, J_TMP_C, J_DTMP_C
};
int64_t jastrow_inter_size(const JasDevView& v, int kind);
hipError_t launch_jastrow_inter(const JasDevView& v, int kind, double* out, hipStream_t s);

// Jastrow parameter derivatives and nuclear forces (kernels_deriv.hip).  Inputs are the
// reference-layout intermediates of launch_jastrow_inter and the e-n distances [W*Ne][Nn].
enum DerivKind {
  DV_EE_PDERIV = 0, DV_EE_GL_PDERIV, DV_EN_PDERIV, DV_EN_GL_PDERIV, DV_EEN_PDERIV, DV_EEN_GL_PDERIV,
  DV_FORCES_EN, DV_FORCES_EN_G, DV_FORCES_EN_L, DV_FORCES_TMP_C, DV_FORCES_DTMP_C, DV_FORCES_EEN_N,
  DV_FORCES_EEN, DV_FORCES_EEN_G, DV_FORCES_EEN_L
};
struct DerivArgs {
  int64_t W, Ne, Nn, up, C, dim, ntype, aord, bord;
  int32_t spin_indep;
  const int64_t* type;   // [Nn]
  const double *a, *b, *kappa, *cfull;  // [type][aord+1], [bord+1], [type], [dim][Nn]
  const int64_t* lkpm;   // [4][dim]
  const double *ajb, *aja;  // asymp_jasb_pderiv [2][bord+1], asymp_jasa_pderiv [type][aord+1]
  const double *ee, *eeg, *en, *eng, *edist, *een_e, *een_e_gl, *een_n, *een_n_gl, *tmp_c, *dtmp_c;
  const double *ftmp_c, *fdtmp_c, *feen_n;
};
// Per-walker values of `what` (the walker-summed derivatives before the sum over walkers).
hipError_t launch_deriv(const DerivArgs& d, int what, double* out, hipStream_t s);
int64_t deriv_size(const DerivArgs& d, int what);
hipError_t launch_sum_walkers(const double* in, int64_t W, int64_t n, double scale, double* out, hipStream_t s);

// ---- Slater determinants, local energy, Sherman-Morrison-Woodbury (kernels_det.hip) -----
// One Slater matrix per (determinant, walker) of one spin: M[e][m] = mo_vgl value of MO
// idx[d][w][m] at electron e0 + e of walker w.
struct SlaterArgs {
  const double* mo_vgl;  // [W*Ne][5][mo_num]
  int64_t mo_num, elec_num, walk_num;
  int n, e0;             // the spin's electrons e0 .. e0 + n - 1
  const int32_t* idx;    // [det][W][n], 0-based MO indices (checked by the caller)
  double* value;         // [det][W]: det M
  double* inv;           // [det][W][n][n]: M^-1 (NaN when M is singular), or null
  double* work;          // [det * W][n][n] working matrices, needed when n > 128
  double* drift;         // [W][Ne][3]: 2 grad_e D / D of determinant 0, or null
  double* kin;           // [W]: -1/2 sum_e lap_e D / D of determinant 0 (with drift), or null
};
hipError_t launch_slater(const SlaterArgs& a, int64_t n_mat, hipStream_t s);
// adjugate B = det(A) A^-1 of one column-major n x n matrix; det(A) to *det (0: singular, B
// untouched); work: n * n doubles when n > 128
hipError_t launch_adjugate_gj(int n, const double* A, int64_t lda, double* B, int64_t ldb, double* det, double* work,
                              hipStream_t s);
// column-major BLAS-like operations (kernels_blas.hip)
hipError_t launch_dgemm(bool ta, bool tb, int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t lda,
                        const double* B, int64_t ldb, double beta, double* C, int64_t ldc, hipStream_t s);
hipError_t launch_dgemv(bool t, int64_t m, int64_t n, double alpha, const double* A, int64_t lda, const double* x,
                        int64_t incx, double beta, double* y, int64_t incy, hipStream_t s);
hipError_t launch_dger(int64_t m, int64_t n, double alpha, const double* x, int64_t incx, const double* y,
                       int64_t incy, double* A, int64_t lda, hipStream_t s);
hipError_t launch_adjugate_small(int n, const double* A, int64_t lda, double* B, int64_t ldb, double* det,
                                 hipStream_t s);
// out [det][W][5][n][n]: component c of MO idx[m] at electron e, [e][m] inner
hipError_t launch_slater_vgl(const SlaterArgs& a, int64_t n_mat, double* out, hipStream_t s);
// out[m] = value[m / nn] * inv[m] for m < count (adjugates of n x n matrices, nn = n * n)
hipError_t launch_adjugate(const double* inv, const double* value, int64_t nn, int64_t count, double* out,
                           hipStream_t s);
// Per walker, from the coordinates [3][coord_ld]: V_ee, V_en, e_pot = V_ee + V_en + vnn and,
// when kin_a is given, e_kin = kin_a + kin_b and e_local = e_kin + e_pot (null outputs skipped).
struct PotArgs {
  const double* coord;
  int64_t coord_ld, elec_num, nucl_num;
  const double *nucl_coord, *charge;  // [3][Nn], [Nn]
  double vnn;
  const double *kin_a, *kin_b;
  double *vee, *ven, *e_pot, *e_kin, *e_local;
};
hipError_t launch_potential(const PotArgs& a, int64_t walk_num, hipStream_t s);
// Sherman-Morrison-Woodbury on n_mat inverses [dim][lds] with n_upd updates [n_upd][lds] each
// (1-based indices).  kind: 0 naive, 1 splitting, 2 splitting core (halves of the updates that
// break down appended to later_* after later_num[m], stride later_cap per matrix), 3 Woodbury
// 2x2, 4 Woodbury 3x3.  status[m]: 0 applied, 1 breakdown, 2 an index outside 1..dim; the
// matrix and its determinant are written only when 0.
struct SmwArgs {
  int kind;
  int64_t n_mat, lds, dim, n_upd;
  const double* upd;
  const uint64_t* upd_idx;
  double breakdown;
  double* s_inv;
  double* det;          // [n_mat] or null
  int32_t* status;      // [n_mat]
  double* later_upd;    // kind 1: [n_mat][2][n_upd][lds] queue; kind 2: [n_mat][later_cap][lds]
  uint64_t* later_idx;  // kind 1: [n_mat][2][n_upd]; kind 2: [n_mat][later_cap]
  uint64_t* later_num;  // kind 2: [n_mat]
  int64_t later_cap;
  double* work;  // [n_mat][dim][dim] when dim > 128
};
hipError_t launch_smw(const SmwArgs& a, hipStream_t s);

// ---- orbital Hessians, forces on the orbitals, MO cusp correction (kernels_orb.hip) ----
// [3][P][4][ao]: d_i d_j AO (j < 3), d_i lap AO (j = 3)
hipError_t launch_ao_hessian(const AoDevView& v, const double* coord, int64_t P, double* out, hipStream_t s);
// [P][3][4][shell]: the same for the radial parts of the shells
hipError_t launch_shell_hessian(const AoDevView& v, const double* coord, int64_t P, double* out, hipStream_t s);
// out[a][u][v][i] (strides oa, ou, ov) = sign * sum_{k_start[a] <= k < k_start[a+1]}
// src[base + u su + v sv + k] * ct[k][i] (ct row stride ldct), for a < na, u < U, v < V, i < mo
struct ContractArgs {
  const double* src;
  int64_t base, su, sv, U, V;
  const double* ct;
  int64_t ldct, mo;
  const int64_t* k_start;  // [na + 1], device
  int64_t na;
  double* out;
  int64_t oa, ou, ov;
  double sign;
};
hipError_t launch_orb_contract(const ContractArgs& c, hipStream_t s);
// [Nn][P][3][ao]: -grad AO for the AOs on the nucleus, 0 for the others
hipError_t launch_forces_ao_value(const AoDevView& v, const double* ao_vgl, int64_t P, double* out, hipStream_t s);
// Cusp correction in place: at every point closer than r_cusp[A] to nucleus A, the s AOs of A
// (s_ao[s_start[A] .. s_start[A+1]), at most cusp_max_s_ao()) taken out of the MOs and the
// polynomial param[A][0..3][i] r^n added; out is the MO VGL [P][5][mo] (hessian = 0) or the MO
// Hessian [3][P][3][mo] (hessian = 1).
struct CuspArgs {
  const double* coord;  // [3][P]
  int64_t P, mo, nucl_num;
  const double *r_cusp, *param;  // [Nn], [Nn][4][mo]
  const int32_t *s_start, *s_ao;
  const double* ct;  // [ao][ldct] (FP64 C^T)
  int64_t ldct;
  AoDevView v;
  double* out;
  int hessian;
};
int cusp_max_s_ao();
hipError_t launch_mo_cusp(const CuspArgs& c, hipStream_t s);

hipError_t launch_ee_distance(const double* coord, int64_t W, int64_t Ne, double* out, hipStream_t s);
// e-n distances [P][Nn] of the P points coord[p + d ld] (d = 0, 1, 2)
hipError_t launch_en_distance(const double* coord, int64_t P, int64_t ld, const double* nucl_coord, int64_t Nn,
                              double* out, hipStream_t s);

}  // namespace qmckl_amd
