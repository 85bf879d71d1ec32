// Slater determinants, the local energy and the Sherman-Morrison-Woodbury updates: the C-ABI
// of org/under_development/qmckl_determinant.org, qmckl_local_energy.org and
// qmckl_sherman_morrison_woodbury.org over the kernels of kernels_det.hip.
//
// The reference builds none of these modules; the entry points keep its names and argument
// lists.  Where its Fortran departs from the quantities its text defines, these follow the
// text, in C layouts:
//   * det_vgl [det][W][5][n][n] holds component c of MO mo_index[d][w][m] - 1 (1-based
//     indices, as its test sets them) at electron e of the spin, [e][m] inner; the MO values
//     are the walker's own (the Fortran reads mo_vgl without a walker offset);
//   * det_inv_matrix is the inverse M^-1 of that C-layout value matrix, det_adj_matrix
//     det(M) M^-1, det the determinant;
//   * the kinetic energy is -1/2 sum_e lap_e D / D = -1/2 sum_e sum_m lap phi_m(r_e)
//     M^-1[m][e] (the Fortran contracts M^-1 transposed), the drift 2 grad_e D / D as
//     [W][Ne][3]; both for one determinant per spin (the Fortran sums determinants without
//     coefficients), QMCKL_FAILURE otherwise;
//   * a Sherman-Morrison-Woodbury call that breaks down returns QMCKL_FAILURE and leaves the
//     inverse and the determinant as they were (the reference leaves them part-updated).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.hpp"

using namespace qmckl_amd;

namespace {

constexpr uint32_t DET_TYPE = 1u, DET_NUM_A = 2u, DET_NUM_B = 4u, DET_IDX_A = 8u, DET_IDX_B = 16u;

qmckl_exit_code det_set(Context* c, uint32_t mask) {
  c->det_uninit &= ~mask;
  c->det_set_date += 1;
  return QMCKL_SUCCESS;
}

int64_t spin_n(const Context* c, int spin) { return spin == 0 ? c->up_num : c->down_num; }
int64_t spin_dets(const Context* c, int spin) { return spin == 0 ? c->det_num_alpha : c->det_num_beta; }

// Offsets (in doubles) of the spin's values and inverses in d_det_value / d_det_inv.
int64_t value_off(const Context* c, int spin) { return spin == 0 ? 0 : c->det_num_alpha * c->walk_num; }
int64_t inv_off(const Context* c, int spin) {
  return spin == 0 ? 0 : c->det_num_alpha * c->walk_num * c->up_num * c->up_num;
}
int64_t idx_off(const Context* c, int spin) { return spin == 0 ? 0 : c->det_num_alpha * c->walk_num * c->up_num; }

SlaterArgs slater_args(Context* c, int spin) {
  SlaterArgs a;
  a.mo_vgl = c->d_mo_vgl.as<double>();
  a.mo_num = c->mo_num;
  a.elec_num = c->elec_num;
  a.walk_num = c->walk_num;
  a.n = (int)spin_n(c, spin);
  a.e0 = spin == 0 ? 0 : (int)c->up_num;
  a.idx = c->d_det_idx.as<int32_t>() + idx_off(c, spin);
  a.value = c->d_det_value.as<double>() + value_off(c, spin);
  a.inv = c->d_det_inv.as<double>() + inv_off(c, spin);
  a.work = c->d_det_work.as<double>();
  a.drift = nullptr;
  a.kin = nullptr;
  return a;
}

// The determinants of both spins, their inverses and (one determinant per spin) the drift and
// the kinetic terms, on the device and current for the context.
qmckl_exit_code provide_det(Context* c, qmckl_context handle, const char* fn) {
  if (c->det_uninit != 0) return fail(handle, QMCKL_NOT_PROVIDED, fn, "determinant data not set");
  if (!c->mo_provided || !c->elec_provided || c->walk_num <= 0) return fail(handle, QMCKL_NOT_PROVIDED, fn, nullptr);
  const int64_t W = c->walk_num;
  for (int spin = 0; spin < 2; ++spin)
    if (c->mo_index_walk[spin] != W)
      return fail(handle, QMCKL_FAILURE, fn, "the MO indices were set for another number of walkers");
  qmckl_exit_code rc = provide_mo_vgl(handle, c, nullptr, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  hipError_t e = hipSuccess;
  if (c->det_idx_date != c->det_set_date || c->det_idx_const != c->const_date) {
    // checked here: a kernel reads mo_vgl at these indices
    std::vector<int32_t> idx;
    idx.reserve(c->mo_index_alpha.size() + c->mo_index_beta.size());
    for (const std::vector<int64_t>* v : {&c->mo_index_alpha, &c->mo_index_beta})
      for (int64_t m : *v) {
        if (m < 1 || m > c->mo_num) return fail(handle, QMCKL_FAILURE, fn, "MO index outside 1..mo_num");
        idx.push_back((int32_t)(m - 1));
      }
    e = c->d_det_idx.reserve(sizeof(int32_t) * std::max<size_t>(idx.size(), 1));
    if (e == hipSuccess && !idx.empty())
      e = hipMemcpyAsync(c->d_det_idx.ptr, idx.data(), sizeof(int32_t) * idx.size(), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(handle, e, fn);
    c->det_idx_date = c->det_set_date;
    c->det_idx_const = c->const_date;
  }
  if (c->det_date == c->date && c->det_seen == c->det_set_date) return QMCKL_SUCCESS;
  const int64_t na = c->up_num, nb = c->down_num, Ne = c->elec_num;
  const int64_t n_val = (c->det_num_alpha + c->det_num_beta) * W;
  const int64_t n_inv = c->det_num_alpha * W * na * na + c->det_num_beta * W * nb * nb;
  int64_t n_work = 0;
  for (int spin = 0; spin < 2; ++spin)
    if (spin_n(c, spin) > 128) n_work = std::max(n_work, spin_dets(c, spin) * W * spin_n(c, spin) * spin_n(c, spin));
  e = c->d_det_value.reserve(sizeof(double) * std::max<int64_t>(n_val, 1));
  if (e == hipSuccess) e = c->d_det_inv.reserve(sizeof(double) * std::max<int64_t>(n_inv, 1));
  if (e == hipSuccess) e = c->d_det_work.reserve(sizeof(double) * std::max<int64_t>(n_work, 1));
  if (e == hipSuccess) e = c->d_det_drift.reserve(sizeof(double) * W * Ne * 3);
  if (e == hipSuccess) e = c->d_det_kin.reserve(sizeof(double) * 2 * W);
  if (e != hipSuccess) return hip_fail(handle, e, fn);
  const bool single = c->det_num_alpha == 1 && c->det_num_beta == 1;
  for (int spin = 0; spin < 2 && e == hipSuccess; ++spin) {
    SlaterArgs a = slater_args(c, spin);
    if (single) {
      a.drift = c->d_det_drift.as<double>();
      a.kin = c->d_det_kin.as<double>() + spin * W;
    }
    e = launch_slater(a, spin_dets(c, spin) * W, c->stream);
  }
  if (e != hipSuccess) return hip_fail(handle, e, fn);
  c->det_date = c->date;
  c->det_seen = c->det_set_date;
  return QMCKL_SUCCESS;
}

}  // namespace

namespace qmckl_amd {

// e_kin, e_pot, e_local, V_ee, V_en ([5][W] in d_energy) on the device.  kinetic: with the
// determinants; without nuclei only V_ee is meaningful.
qmckl_exit_code provide_energy(Context* c, qmckl_context handle, const char* fn, bool kinetic) {
  if (!c->elec_provided || c->walk_num <= 0 || (kinetic && !c->nucl_provided))
    return fail(handle, QMCKL_NOT_PROVIDED, fn, nullptr);
  qmckl_exit_code rc;
  if (kinetic) {
    if (c->det_uninit == 0 && (c->det_num_alpha != 1 || c->det_num_beta != 1))
      return fail(handle, QMCKL_FAILURE, fn, "the local energy is defined for one determinant per spin");
    rc = provide_det(c, handle, fn);
    if (rc != QMCKL_SUCCESS) return rc;
    if (c->energy_date == c->date && c->det_seen == c->det_set_date) return QMCKL_SUCCESS;
  }
  rc = device_ready(c, handle, fn);
  if (rc == QMCKL_SUCCESS) rc = upload_coords(c, handle, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  const int64_t W = c->walk_num, Nn = c->nucl_provided ? c->nucl_num : 0;
  hipError_t e = c->d_energy.reserve(sizeof(double) * 5 * W);
  if (e == hipSuccess && Nn > 0 && c->pot_charge_date != c->const_date) {
    // nuclear charges and coordinates [Nn] + [3][Nn]
    std::vector<double> buf(c->nucl_charge);
    buf.insert(buf.end(), c->nucl_coord.begin(), c->nucl_coord.end());
    e = c->d_pot_charge.reserve(sizeof(double) * buf.size());
    if (e == hipSuccess)
      e = hipMemcpyAsync(c->d_pot_charge.ptr, buf.data(), sizeof(double) * buf.size(), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) c->pot_charge_date = c->const_date;
  }
  if (e != hipSuccess) return hip_fail(handle, e, fn);
  double vnn = 0.0;
  if (Nn > 0) {
    rc = qmckl_get_nucleus_repulsion(handle, &vnn);
    if (rc != QMCKL_SUCCESS) return rc;
  }
  double* en = c->d_energy.as<double>();
  PotArgs p;
  p.coord = c->d_coord.as<double>();
  p.coord_ld = c->point_num;
  p.elec_num = c->elec_num;
  p.nucl_num = Nn;
  p.charge = Nn > 0 ? c->d_pot_charge.as<double>() : nullptr;
  p.nucl_coord = Nn > 0 ? p.charge + Nn : nullptr;
  p.vnn = vnn;
  p.kin_a = kinetic ? c->d_det_kin.as<double>() : nullptr;
  p.kin_b = kinetic ? c->d_det_kin.as<double>() + W : nullptr;
  p.e_kin = en;
  p.e_pot = en + W;
  p.e_local = en + 2 * W;
  p.vee = en + 3 * W;
  p.ven = en + 4 * W;
  e = launch_potential(p, W, c->stream);
  if (e != hipSuccess) return hip_fail(handle, e, fn);
  if (kinetic) c->energy_date = c->date;
  return QMCKL_SUCCESS;
}

}  // namespace qmckl_amd

namespace {

Context* locked_ctx(qmckl_context context) { return ctx_of(context); }

qmckl_exit_code copy_dev(Context* c, qmckl_context context, void* dst, const void* src, size_t bytes, const char* fn) {
  const hipError_t e = copy_to_host(c, dst, src, bytes);
  if (e != hipSuccess) return hip_fail(context, e, fn);
  return QMCKL_SUCCESS;
}

enum DetOut { OUT_VGL, OUT_INV, OUT_ADJ, OUT_VALUE };

qmckl_exit_code get_det(qmckl_context context, double* out, int spin, DetOut what, const char* fn) {
  Context* c = locked_ctx(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (out == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "NULL pointer");
  qmckl_exit_code rc = provide_det(c, context, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  const int64_t W = c->walk_num, n = spin_n(c, spin), nd = spin_dets(c, spin);
  const int64_t nn = n * n;
  SlaterArgs a = slater_args(c, spin);
  hipError_t e = hipSuccess;
  switch (what) {
    case OUT_VALUE:
      return copy_dev(c, context, out, a.value, sizeof(double) * nd * W, fn);
    case OUT_INV:
      return copy_dev(c, context, out, a.inv, sizeof(double) * nd * W * nn, fn);
    case OUT_ADJ:
      e = c->d_det_out.reserve(sizeof(double) * std::max<int64_t>(nd * W * nn, 1));
      if (e == hipSuccess) e = launch_adjugate(a.inv, a.value, nn, nd * W * nn, c->d_det_out.as<double>(), c->stream);
      if (e != hipSuccess) return hip_fail(context, e, fn);
      return copy_dev(c, context, out, c->d_det_out.ptr, sizeof(double) * nd * W * nn, fn);
    case OUT_VGL:
      e = c->d_det_out.reserve(sizeof(double) * std::max<int64_t>(nd * W * 5 * nn, 1));
      if (e == hipSuccess) e = launch_slater_vgl(a, nd * W, c->d_det_out.as<double>(), c->stream);
      if (e != hipSuccess) return hip_fail(context, e, fn);
      return copy_dev(c, context, out, c->d_det_out.ptr, sizeof(double) * nd * W * 5 * nn, fn);
  }
  return QMCKL_FAILURE;
}

// which: 0 e_kin, 1 e_pot, 2 e_local, 3 V_ee, 4 V_en
qmckl_exit_code get_energy(qmckl_context context, double* out, int which, const char* fn) {
  Context* c = locked_ctx(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (out == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "NULL pointer");
  if (which != 3 && !c->nucl_provided) return fail(context, QMCKL_NOT_PROVIDED, fn, nullptr);
  qmckl_exit_code rc = provide_energy(c, context, fn, which == 0 || which == 2);
  if (rc != QMCKL_SUCCESS) return rc;
  return copy_dev(c, context, out, c->d_energy.as<double>() + which * c->walk_num, sizeof(double) * c->walk_num, fn);
}

qmckl_exit_code set_mo_index(qmckl_context context, int spin, const int64_t* idx, int64_t size_max, const char* fn) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (idx == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "NULL pointer");
  const uint32_t num_mask = spin == 0 ? DET_NUM_A : DET_NUM_B;
  if ((c->det_uninit & num_mask) != 0) return fail(context, QMCKL_NOT_PROVIDED, fn, "set the number of determinants first");
  if (!c->elec_provided || c->walk_num <= 0)
    return fail(context, QMCKL_NOT_PROVIDED, fn, "set the electron coordinates (walkers) first");
  const int64_t need = c->walk_num * spin_dets(c, spin) * spin_n(c, spin);
  if (size_max < need) return fail(context, QMCKL_INVALID_ARG_3, fn, "input array too small");
  std::vector<int64_t>& dst = spin == 0 ? c->mo_index_alpha : c->mo_index_beta;
  dst.assign(idx, idx + need);
  c->mo_index_walk[spin] = c->walk_num;
  return det_set(c, spin == 0 ? DET_IDX_A : DET_IDX_B);
}

// ---- Sherman-Morrison-Woodbury through the device -------------------------------------

struct SmwHost {
  int kind;
  uint64_t LDS, Dim, N;
  const double* upd;
  const uint64_t* idx;
  double breakdown;
  double* s_inv;
  double* det;
  double* later_upd;
  uint64_t* later_idx;
  uint64_t* later;
};

// Argument positions of the reference signatures: naive, splitting, splitting_core
// (ctx, LDS, Dim, N, Updates, Updates_index, breakdown, Slater_inv, ...), Woodbury
// (ctx, LDS, Dim, Updates, Updates_index, breakdown, Slater_inv, determinant).
qmckl_exit_code smw_host(qmckl_context context, const SmwHost& h, const char* fn) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  const bool wb = h.kind >= 3;
  const int a_upd = wb ? 4 : 5, a_idx = wb ? 5 : 6, a_inv = wb ? 7 : 8;
  if (h.Dim < 1) return fail(context, QMCKL_INVALID_ARG_3, fn, "Dim < 1");
  if (h.LDS < h.Dim) return fail(context, QMCKL_INVALID_ARG_2, fn, "LDS < Dim");
  if (h.N < 1) return fail(context, QMCKL_INVALID_ARG_4, fn, "N_updates < 1");
  if (h.upd == nullptr) return fail(context, (qmckl_exit_code)a_upd, fn, "NULL pointer");
  if (h.idx == nullptr) return fail(context, (qmckl_exit_code)a_idx, fn, "NULL pointer");
  if (h.s_inv == nullptr) return fail(context, (qmckl_exit_code)a_inv, fn, "NULL pointer");
  if (h.kind == 2 && (h.later_upd == nullptr || h.later_idx == nullptr || h.later == nullptr))
    return fail(context, QMCKL_INVALID_ARG_9, fn, "NULL pointer");
  for (uint64_t l = 0; l < h.N; ++l)
    if (h.idx[l] < 1 || h.idx[l] > h.Dim) return fail(context, (qmckl_exit_code)a_idx, fn, "index outside 1..Dim");
  qmckl_exit_code rc = device_ready(c, context, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  const uint64_t base = h.kind == 2 ? *h.later : 0;
  const uint64_t cap = h.kind == 2 ? base + h.N : 2 * h.N;
  // staging: [upd N*LDS][s_inv Dim*LDS][later cap*LDS][work Dim*Dim][det][status|later_num][idx N][later_idx cap]
  const size_t n_upd = h.N * h.LDS, n_inv = h.Dim * h.LDS, n_lat = cap * h.LDS;
  const size_t n_work = h.Dim > 128 ? h.Dim * h.Dim : 0;
  const size_t doubles = n_upd + n_inv + n_lat + n_work + 2;
  const size_t bytes = sizeof(double) * doubles + sizeof(uint64_t) * (h.N + cap + 2);
  hipError_t e = c->d_smw.reserve(bytes);
  if (e != hipSuccess) return hip_fail(context, e, fn);
  double* d = c->d_smw.as<double>();
  double* d_upd = d;
  double* d_inv = d_upd + n_upd;
  double* d_lat = d_inv + n_inv;
  double* d_work = d_lat + n_lat;
  double* d_det = d_work + n_work;
  int32_t* d_status = reinterpret_cast<int32_t*>(d_det + 1);
  uint64_t* d_latnum = reinterpret_cast<uint64_t*>(d_det + 2);
  uint64_t* d_idx = d_latnum + 1;
  uint64_t* d_latidx = d_idx + h.N + 1;
  const double det0 = h.det ? *h.det : 1.0;
  e = hipMemcpyAsync(d_upd, h.upd, sizeof(double) * n_upd, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(d_inv, h.s_inv, sizeof(double) * n_inv, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(d_idx, h.idx, sizeof(uint64_t) * h.N, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(d_det, &det0, sizeof(double), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(d_latnum, &base, sizeof(uint64_t), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess && h.kind == 2 && base > 0) {
    e = hipMemcpyAsync(d_lat, h.later_upd, sizeof(double) * base * h.LDS, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_latidx, h.later_idx, sizeof(uint64_t) * base, hipMemcpyHostToDevice, c->stream);
  }
  SmwArgs a;
  a.kind = h.kind;
  a.n_mat = 1;
  a.lds = (int64_t)h.LDS;
  a.dim = (int64_t)h.Dim;
  a.n_upd = (int64_t)h.N;
  a.upd = d_upd;
  a.upd_idx = d_idx;
  a.breakdown = h.breakdown;
  a.s_inv = d_inv;
  a.det = d_det;
  a.status = d_status;
  a.later_upd = d_lat;
  a.later_idx = d_latidx;
  a.later_num = d_latnum;
  a.later_cap = (int64_t)cap;
  a.work = d_work;
  if (e == hipSuccess) e = launch_smw(a, c->stream);
  int32_t status = -1;
  double det1 = det0;
  uint64_t later = base;
  if (e == hipSuccess) e = hipMemcpyAsync(&status, d_status, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_fail(context, e, fn);
  if (status == 1) return fail(context, QMCKL_FAILURE, fn, "breakdown: a denominator fell below the threshold");
  if (status != 0) return fail(context, QMCKL_FAILURE, fn, "update index outside 1..Dim");
  e = hipMemcpyAsync(h.s_inv, d_inv, sizeof(double) * n_inv, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(&det1, d_det, sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess && h.kind == 2) e = hipMemcpyAsync(&later, d_latnum, sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess && h.kind == 2 && later > base) {
    e = hipMemcpyAsync(h.later_upd + base * h.LDS, d_lat + base * h.LDS, sizeof(double) * (later - base) * h.LDS,
                       hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(h.later_idx + base, d_latidx + base, sizeof(uint64_t) * (later - base), hipMemcpyDeviceToHost,
                         c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  }
  if (e != hipSuccess) return hip_fail(context, e, fn);
  if (h.det) *h.det = det1;
  if (h.kind == 2) *h.later = later;
  return QMCKL_SUCCESS;
}

qmckl_exit_code smw(qmckl_context ctx, int kind, uint64_t LDS, uint64_t Dim, uint64_t N, const double* U,
                    const uint64_t* idx, double breakdown, double* S, double* det, const char* fn) {
  SmwHost h{kind, LDS, Dim, N, U, idx, breakdown, S, det, nullptr, nullptr, nullptr};
  return smw_host(ctx, h, fn);
}

}  // namespace

extern "C" {

// ---- determinant setters, qmckl_determinant.org:344-556 -------------------------------
qmckl_exit_code qmckl_set_determinant_type(const qmckl_context context, const char t) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (t != 'G' && t != 'S') return fail(context, QMCKL_INVALID_ARG_2, "qmckl_set_determinant_type", nullptr);
  c->det_type = t;
  return det_set(c, DET_TYPE);
}

qmckl_exit_code qmckl_set_determinant_det_num_alpha(const qmckl_context context, const int64_t det_num_alpha) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (det_num_alpha <= 0)
    return fail(context, QMCKL_INVALID_ARG_2, "qmckl_set_determinant_det_num_alpha", "det_num_alpha <= 0");
  c->det_num_alpha = det_num_alpha;
  c->mo_index_alpha.clear();
  c->mo_index_walk[0] = 0;
  c->det_uninit |= DET_IDX_A;
  return det_set(c, DET_NUM_A);
}

qmckl_exit_code qmckl_set_determinant_det_num_beta(const qmckl_context context, const int64_t det_num_beta) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (det_num_beta <= 0)
    return fail(context, QMCKL_INVALID_ARG_2, "qmckl_set_determinant_det_num_beta", "det_num_beta <= 0");
  c->det_num_beta = det_num_beta;
  c->mo_index_beta.clear();
  c->mo_index_walk[1] = 0;
  c->det_uninit |= DET_IDX_B;
  return det_set(c, DET_NUM_B);
}

qmckl_exit_code qmckl_set_determinant_mo_index_alpha(const qmckl_context context, const int64_t* mo_index_alpha,
                                                     const int64_t size_max) {
  return set_mo_index(context, 0, mo_index_alpha, size_max, "qmckl_set_determinant_mo_index_alpha");
}

qmckl_exit_code qmckl_set_determinant_mo_index_beta(const qmckl_context context, const int64_t* mo_index_beta,
                                                    const int64_t size_max) {
  return set_mo_index(context, 1, mo_index_beta, size_max, "qmckl_set_determinant_mo_index_beta");
}

bool qmckl_determinant_provided(const qmckl_context context) {
  Context* c = ctx_of(context);
  if (!c) return false;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  return c->det_uninit == 0;
}

// ---- determinant getters, qmckl_determinant.org:569-640, 1308-1495 ---------------------
qmckl_exit_code qmckl_get_det_vgl_alpha(qmckl_context context, double* const det_vgl_alpha) {
  return get_det(context, det_vgl_alpha, 0, OUT_VGL, "qmckl_get_det_vgl_alpha");
}
qmckl_exit_code qmckl_get_det_vgl_beta(qmckl_context context, double* const det_vgl_beta) {
  return get_det(context, det_vgl_beta, 1, OUT_VGL, "qmckl_get_det_vgl_beta");
}
qmckl_exit_code qmckl_get_det_inv_matrix_alpha(qmckl_context context, double* const det_inv_matrix_alpha) {
  return get_det(context, det_inv_matrix_alpha, 0, OUT_INV, "qmckl_get_det_inv_matrix_alpha");
}
qmckl_exit_code qmckl_get_det_inv_matrix_beta(qmckl_context context, double* const det_inv_matrix_beta) {
  return get_det(context, det_inv_matrix_beta, 1, OUT_INV, "qmckl_get_det_inv_matrix_beta");
}
qmckl_exit_code qmckl_get_det_adj_matrix_alpha(qmckl_context context, double* const det_adj_matrix_alpha) {
  return get_det(context, det_adj_matrix_alpha, 0, OUT_ADJ, "qmckl_get_det_adj_matrix_alpha");
}
qmckl_exit_code qmckl_get_det_adj_matrix_beta(qmckl_context context, double* const det_adj_matrix_beta) {
  return get_det(context, det_adj_matrix_beta, 1, OUT_ADJ, "qmckl_get_det_adj_matrix_beta");
}
qmckl_exit_code qmckl_get_det_alpha(qmckl_context context, double* const det_value_alpha) {
  return get_det(context, det_value_alpha, 0, OUT_VALUE, "qmckl_get_det_alpha");
}
qmckl_exit_code qmckl_get_det_beta(qmckl_context context, double* const det_value_beta) {
  return get_det(context, det_value_beta, 1, OUT_VALUE, "qmckl_get_det_beta");
}

// ---- local energy, qmckl_local_energy.org:203-1713 -------------------------------------
bool qmckl_local_energy_provided(const qmckl_context context) {
  Context* c = ctx_of(context);
  if (!c) return false;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  return c->elec_provided && c->nucl_provided && c->mo_provided && c->det_uninit == 0;
}

qmckl_exit_code qmckl_get_kinetic_energy(qmckl_context context, double* const kinetic_energy) {
  return get_energy(context, kinetic_energy, 0, "qmckl_get_kinetic_energy");
}

qmckl_exit_code qmckl_get_potential_energy(qmckl_context context, double* const potential_energy) {
  return get_energy(context, potential_energy, 1, "qmckl_get_potential_energy");
}

qmckl_exit_code qmckl_get_local_energy(qmckl_context context, double* const local_energy, const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  {
    std::lock_guard<std::recursive_mutex> g(c->mutex);
    if (local_energy != nullptr && size_max < c->walk_num)
      return fail(context, QMCKL_INVALID_ARG_3, "qmckl_get_local_energy", "Array too small");
  }
  return get_energy(context, local_energy, 2, "qmckl_get_local_energy");
}

qmckl_exit_code qmckl_get_drift_vector(qmckl_context context, double* const drift_vector) {
  const char* fn = "qmckl_get_drift_vector";
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (drift_vector == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "NULL pointer");
  if (c->det_uninit == 0 && (c->det_num_alpha != 1 || c->det_num_beta != 1))
    return fail(context, QMCKL_FAILURE, fn, "the drift vector is defined for one determinant per spin");
  qmckl_exit_code rc = provide_det(c, context, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  return copy_dev(c, context, drift_vector, c->d_det_drift.ptr, sizeof(double) * c->walk_num * c->elec_num * 3, fn);
}

// ---- Sherman-Morrison-Woodbury, qmckl_sherman_morrison_woodbury.org ----------------------
qmckl_exit_code qmckl_sm_naive(const qmckl_context context, const uint64_t LDS, const uint64_t Dim,
                               const uint64_t N_updates, const double* Updates, const uint64_t* Updates_index,
                               const double breakdown, double* Slater_inv, double* determinant) {
  return smw(context, 0, LDS, Dim, N_updates, Updates, Updates_index, breakdown, Slater_inv, determinant,
             "qmckl_sm_naive");
}
qmckl_exit_code qmckl_sm_naive_doc(const qmckl_context context, const uint64_t LDS, const uint64_t Dim,
                                   const uint64_t N_updates, const double* Updates, const uint64_t* Updates_index,
                                   const double breakdown, double* Slater_inv, double* determinant) {
  return smw(context, 0, LDS, Dim, N_updates, Updates, Updates_index, breakdown, Slater_inv, determinant,
             "qmckl_sm_naive_doc");
}
qmckl_exit_code qmckl_sm_splitting(const qmckl_context context, const uint64_t LDS, const uint64_t Dim,
                                   const uint64_t N_updates, const double* Updates, const uint64_t* Updates_index,
                                   const double breakdown, double* Slater_inv, double* determinant) {
  return smw(context, 1, LDS, Dim, N_updates, Updates, Updates_index, breakdown, Slater_inv, determinant,
             "qmckl_sm_splitting");
}
qmckl_exit_code qmckl_sm_splitting_doc(const qmckl_context context, const uint64_t LDS, const uint64_t Dim,
                                       const uint64_t N_updates, const double* Updates,
                                       const uint64_t* Updates_index, const double breakdown, double* Slater_inv,
                                       double* determinant) {
  return smw(context, 1, LDS, Dim, N_updates, Updates, Updates_index, breakdown, Slater_inv, determinant,
             "qmckl_sm_splitting_doc");
}
qmckl_exit_code qmckl_sm_splitting_core(const qmckl_context context, const uint64_t LDS, const uint64_t Dim,
                                        const uint64_t N_updates, const double* Updates,
                                        const uint64_t* Updates_index, const double breakdown, double* Slater_inv,
                                        double* later_updates, uint64_t* later_index, uint64_t* later,
                                        double* determinant) {
  SmwHost h{2, LDS, Dim, N_updates, Updates, Updates_index, breakdown, Slater_inv, determinant, later_updates,
            later_index, later};
  return smw_host(context, h, "qmckl_sm_splitting_core");
}
qmckl_exit_code qmckl_sm_splitting_core_doc(const qmckl_context context, const uint64_t LDS, const uint64_t Dim,
                                            const uint64_t N_updates, const double* Updates,
                                            const uint64_t* Updates_index, const double breakdown,
                                            double* Slater_inv, double* later_updates, uint64_t* later_index,
                                            uint64_t* later, double* determinant) {
  SmwHost h{2, LDS, Dim, N_updates, Updates, Updates_index, breakdown, Slater_inv, determinant, later_updates,
            later_index, later};
  return smw_host(context, h, "qmckl_sm_splitting_core_doc");
}
qmckl_exit_code qmckl_woodbury_2x2(const qmckl_context context, const uint64_t LDS, const uint64_t Dim,
                                   const double* Updates, const uint64_t* Updates_index, const double breakdown,
                                   double* Slater_inv, double* determinant) {
  return smw(context, 3, LDS, Dim, 2, Updates, Updates_index, breakdown, Slater_inv, determinant,
             "qmckl_woodbury_2x2");
}
qmckl_exit_code qmckl_woodbury_2x2_doc(const qmckl_context context, const uint64_t LDS, const uint64_t Dim,
                                       const double* Updates, const uint64_t* Updates_index, const double breakdown,
                                       double* Slater_inv, double* determinant) {
  return smw(context, 3, LDS, Dim, 2, Updates, Updates_index, breakdown, Slater_inv, determinant,
             "qmckl_woodbury_2x2_doc");
}
qmckl_exit_code qmckl_woodbury_3x3(const qmckl_context context, const uint64_t LDS, const uint64_t Dim,
                                   const double* Updates, const uint64_t* Updates_index, const double breakdown,
                                   double* Slater_inv, double* determinant) {
  return smw(context, 4, LDS, Dim, 3, Updates, Updates_index, breakdown, Slater_inv, determinant,
             "qmckl_woodbury_3x3");
}
qmckl_exit_code qmckl_woodbury_3x3_doc(const qmckl_context context, const uint64_t LDS, const uint64_t Dim,
                                       const double* Updates, const uint64_t* Updates_index, const double breakdown,
                                       double* Slater_inv, double* determinant) {
  return smw(context, 4, LDS, Dim, 3, Updates, Updates_index, breakdown, Slater_inv, determinant,
             "qmckl_woodbury_3x3_doc");
}

// Batched updates of n_mat device-resident inverses on the context's stream (extension).
qmckl_exit_code qmckl_amd_sm_update_device(qmckl_context context, const int32_t kind, const int64_t n_mat,
                                           const uint64_t LDS, const uint64_t Dim, const uint64_t N_updates,
                                           const double* d_updates, const uint64_t* d_updates_index,
                                           const double breakdown, double* d_slater_inv, double* d_determinant,
                                           int32_t* d_status) {
  const char* fn = "qmckl_amd_sm_update_device";
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (kind != QMCKL_AMD_SM_NAIVE && kind != QMCKL_AMD_SM_SPLITTING && kind != QMCKL_AMD_WOODBURY_2X2 &&
      kind != QMCKL_AMD_WOODBURY_3X3)
    return fail(context, QMCKL_INVALID_ARG_2, fn, "unknown kind");
  if (n_mat < 0) return fail(context, QMCKL_INVALID_ARG_3, fn, "n_mat < 0");
  if (Dim < 1 || LDS < Dim) return fail(context, QMCKL_INVALID_ARG_4, fn, "need 1 <= Dim <= LDS");
  const uint64_t N = kind == QMCKL_AMD_WOODBURY_2X2 ? 2 : kind == QMCKL_AMD_WOODBURY_3X3 ? 3 : N_updates;
  if (N < 1) return fail(context, QMCKL_INVALID_ARG_6, fn, "N_updates < 1");
  if (n_mat == 0) return QMCKL_SUCCESS;
  if (!d_updates || !d_updates_index || !d_slater_inv || !d_status)
    return fail(context, QMCKL_INVALID_ARG_7, fn, "NULL pointer");
  qmckl_exit_code rc = device_ready(c, context, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  const size_t q = kind == QMCKL_AMD_SM_SPLITTING ? (size_t)n_mat * 2 * N : 0;
  const size_t work = Dim > 128 ? (size_t)n_mat * Dim * Dim : 0;
  hipError_t e = c->d_smw.reserve(sizeof(double) * (q * LDS + work + 1) + sizeof(uint64_t) * (q + 1));
  if (e != hipSuccess) return hip_fail(context, e, fn);
  SmwArgs a;
  a.kind = kind;
  a.n_mat = n_mat;
  a.lds = (int64_t)LDS;
  a.dim = (int64_t)Dim;
  a.n_upd = (int64_t)N;
  a.upd = d_updates;
  a.upd_idx = d_updates_index;
  a.breakdown = breakdown;
  a.s_inv = d_slater_inv;
  a.det = d_determinant;
  a.status = d_status;
  a.later_upd = c->d_smw.as<double>();
  a.work = a.later_upd + q * LDS;
  a.later_idx = reinterpret_cast<uint64_t*>(a.work + work + 1);
  a.later_num = nullptr;
  a.later_cap = 0;
  e = launch_smw(a, c->stream);
  if (e != hipSuccess) return hip_fail(context, e, fn);
  return QMCKL_SUCCESS;
}

}  // extern "C"
