// Single-electron moves of the CHAMP Jastrow: the C-ABI of
// org/qmckl_jastrow_champ_single.org (qmckl_set_single_point, the single-move getters,
// qmckl_get_jastrow_champ_single_accept) over the kernels of kernels_single.hip.
//
// One device pass (launch_jastrow_single) evaluates every quantity of the move for all
// walkers; each getter copies its part.  Results are recomputed when the single point or
// the context changed since.
#include <algorithm>
#include <cstring>
#include <vector>

#include "internal.hpp"

using namespace qmckl_amd;

namespace {

// The move's quantities on the device, current for the single point and the context.
qmckl_exit_code provide_single(Context* c, qmckl_context handle, const char* fn, SingleOut& out) {
  if (!c->jas_provided) return fail(handle, QMCKL_NOT_PROVIDED, fn, nullptr);
  if (c->single_num < 0 || (int64_t)c->single_coord.size() != 3 * c->walk_num)
    return fail(handle, QMCKL_NOT_PROVIDED, fn, "qmckl_set_single_point has not been called for these walkers");
  JasDevView v;
  qmckl_exit_code rc = jastrow_view_ready(c, handle, fn, v);
  if (rc != QMCKL_SUCCESS) return rc;
  const int64_t n = single_out_doubles(v);
  hipError_t e = c->d_single_out.reserve(sizeof(double) * n);
  if (e != hipSuccess) return hip_fail(handle, e, fn);
  out = single_out_view(v, c->d_single_out.as<double>());
  if (c->single_out_date == c->single_date && c->single_out_ctx_date == c->date) return QMCKL_SUCCESS;
  e = c->d_single_coord.reserve(sizeof(double) * c->single_coord.size());
  if (e == hipSuccess)
    e = hipMemcpyAsync(c->d_single_coord.ptr, c->single_coord.data(), sizeof(double) * c->single_coord.size(),
                       hipMemcpyHostToDevice, c->stream);
  SingleArgs s;
  s.num = c->single_num;
  s.coord = c->d_single_coord.as<double>();
  s.out = out;
  if (e == hipSuccess) e = launch_jastrow_single(v, s, c->stream);
  if (e != hipSuccess) return hip_fail(handle, e, fn);
  c->single_out_date = c->single_date;
  c->single_out_ctx_date = c->date;
  return QMCKL_SUCCESS;
}

// Copies `n` doubles of the move's output `which` to the caller's buffer.
qmckl_exit_code get_single(qmckl_context context, double* out, int64_t size_max, const char* fn,
                           double* SingleOut::*which, int64_t n_per_walker) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (out == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "NULL pointer");
  const int64_t n = n_per_walker * c->walk_num;
  if (size_max < n) return fail(context, QMCKL_INVALID_ARG_3, fn, "Array too small");
  SingleOut so;
  qmckl_exit_code rc = provide_single(c, context, fn, so);
  if (rc != QMCKL_SUCCESS) return rc;
  const hipError_t e = copy_to_host(c, out, so.*which, sizeof(double) * n);
  if (e != hipSuccess) return hip_fail(context, e, fn);
  return QMCKL_SUCCESS;
}

}  // namespace

extern "C" {

// org/qmckl_jastrow_champ_single.org:452-557.  coord: [walk_num][3] for 'N', [3][walk_num]
// for 'T'.
qmckl_exit_code qmckl_set_single_point(qmckl_context context, const char transp, const int64_t num,
                                       const double* coord, const int64_t size_max) {
  const char* fn = "qmckl_set_single_point";
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (num < 0) return fail(context, QMCKL_INVALID_ARG_3, fn, "Incorrect point number");
  if (transp != 'N' && transp != 'T') return fail(context, QMCKL_INVALID_ARG_2, fn, "transp should be 'N' or 'T'");
  if (coord == nullptr) return fail(context, QMCKL_INVALID_ARG_3, fn, "coord is a NULL pointer");
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (num >= c->elec_num) return fail(context, QMCKL_INVALID_ARG_3, fn, "Incorrect point number");
  const int64_t W = c->walk_num;
  if (size_max < 3 * W) return fail(context, QMCKL_INVALID_ARG_4, fn, "Array too small");
  c->single_coord.resize(3 * W);
  for (int64_t w = 0; w < W; ++w)
    for (int d = 0; d < 3; ++d) c->single_coord[3 * w + d] = transp == 'N' ? coord[3 * w + d] : coord[d * W + w];
  c->single_num = num;
  c->single_date += 1;
  return QMCKL_SUCCESS;
}

// org/qmckl_jastrow_champ_single.org:588-614
qmckl_exit_code qmckl_single_touch(const qmckl_context context) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  c->date += 1;
  c->single_date += 1;
  return QMCKL_SUCCESS;
}

// org/qmckl_jastrow_champ_single.org:628-690, 903-965: distances of the new position
qmckl_exit_code qmckl_get_single_electron_ee_distance(qmckl_context context, double* const distance,
                                                      const int64_t size_max) {
  Context* c = ctx_of(context);
  return get_single(context, distance, size_max, "qmckl_get_single_electron_ee_distance", &SingleOut::ee_dist,
                    c ? c->elec_num : 0);
}
qmckl_exit_code qmckl_get_single_electron_en_distance(qmckl_context context, double* distance,
                                                      const int64_t size_max) {
  Context* c = ctx_of(context);
  return get_single(context, distance, size_max, "qmckl_get_single_electron_en_distance", &SingleOut::en_dist,
                    c ? c->nucl_num : 0);
}

// :1239-1278, 1571-1610, 2740-2779, 3057-3096: e-e-n rescaled distances of the new position
qmckl_exit_code qmckl_get_een_rescaled_single_e(qmckl_context context, double* const distance_rescaled,
                                                const int64_t size_max) {
  Context* c = ctx_of(context);
  return get_single(context, distance_rescaled, size_max, "qmckl_get_een_rescaled_single_e", &SingleOut::een_e,
                    c ? (c->cord + 1) * c->elec_num : 0);
}
qmckl_exit_code qmckl_get_een_rescaled_single_n(qmckl_context context, double* const distance_rescaled,
                                                const int64_t size_max) {
  Context* c = ctx_of(context);
  return get_single(context, distance_rescaled, size_max, "qmckl_get_een_rescaled_single_n", &SingleOut::een_n,
                    c ? (c->cord + 1) * c->nucl_num : 0);
}
qmckl_exit_code qmckl_get_een_rescaled_single_e_gl(qmckl_context context, double* const distance_rescaled,
                                                   const int64_t size_max) {
  Context* c = ctx_of(context);
  return get_single(context, distance_rescaled, size_max, "qmckl_get_een_rescaled_single_e_gl",
                    &SingleOut::een_e_gl, c ? 4 * (c->cord + 1) * c->elec_num : 0);
}
qmckl_exit_code qmckl_get_een_rescaled_single_n_gl(qmckl_context context, double* const distance_rescaled,
                                                   const int64_t size_max) {
  Context* c = ctx_of(context);
  return get_single(context, distance_rescaled, size_max, "qmckl_get_een_rescaled_single_n_gl",
                    &SingleOut::een_n_gl, c ? 4 * (c->cord + 1) * c->nucl_num : 0);
}

// :1867-1900, 3681-3720, 4349-4390: changes of the e-e-n factor, its GL, its gradient for the
// moved electron
qmckl_exit_code qmckl_get_jastrow_champ_single_een(qmckl_context context, double* const delta_een,
                                                   const int64_t size_max) {
  return get_single(context, delta_een, size_max, "qmckl_get_jastrow_champ_single_een", &SingleOut::d_een, 1);
}
qmckl_exit_code qmckl_get_jastrow_champ_single_een_gl(qmckl_context context, double* const delta_een_gl,
                                                      const int64_t size_max) {
  Context* c = ctx_of(context);
  return get_single(context, delta_een_gl, size_max, "qmckl_get_jastrow_champ_single_een_gl", &SingleOut::d_een_gl,
                    c ? 4 * c->elec_num : 0);
}
qmckl_exit_code qmckl_get_jastrow_champ_single_een_g(qmckl_context context, double* const delta_een_g,
                                                     const int64_t size_max) {
  Context* c = ctx_of(context);
  return get_single(context, delta_een_g, size_max, "qmckl_get_jastrow_champ_single_een_g", &SingleOut::d_een_g,
                    c ? 4 * c->elec_num : 0);
}

// :5218-5250, 5512-5550, 5880-5920, 6228-6260: e-e
qmckl_exit_code qmckl_get_ee_rescaled_single(qmckl_context context, double* const distance_rescaled,
                                             const int64_t size_max) {
  Context* c = ctx_of(context);
  return get_single(context, distance_rescaled, size_max, "qmckl_get_ee_rescaled_single", &SingleOut::ee_resc,
                    c ? c->elec_num : 0);
}
qmckl_exit_code qmckl_get_jastrow_champ_single_ee(qmckl_context context, double* const delta_ee,
                                                  const int64_t size_max) {
  return get_single(context, delta_ee, size_max, "qmckl_get_jastrow_champ_single_ee", &SingleOut::d_ee, 1);
}
qmckl_exit_code qmckl_get_ee_rescaled_single_gl(qmckl_context context, double* const distance_rescaled_gl,
                                                const int64_t size_max) {
  Context* c = ctx_of(context);
  return get_single(context, distance_rescaled_gl, size_max, "qmckl_get_ee_rescaled_single_gl",
                    &SingleOut::ee_resc_gl, c ? 4 * c->elec_num : 0);
}
qmckl_exit_code qmckl_get_jastrow_champ_single_ee_gl(qmckl_context context, double* const delta_ee_gl,
                                                     const int64_t size_max) {
  Context* c = ctx_of(context);
  return get_single(context, delta_ee_gl, size_max, "qmckl_get_jastrow_champ_single_ee_gl", &SingleOut::d_ee_gl,
                    c ? 4 * c->elec_num : 0);
}

// :7087-7120, 7380-7415, 7753-7790, 8080-8115: e-n
qmckl_exit_code qmckl_get_en_rescaled_single(qmckl_context context, double* const distance_rescaled,
                                             const int64_t size_max) {
  Context* c = ctx_of(context);
  return get_single(context, distance_rescaled, size_max, "qmckl_get_en_rescaled_single", &SingleOut::en_resc,
                    c ? c->nucl_num : 0);
}
qmckl_exit_code qmckl_get_jastrow_champ_single_en(qmckl_context context, double* const delta_en,
                                                  const int64_t size_max) {
  return get_single(context, delta_en, size_max, "qmckl_get_jastrow_champ_single_en", &SingleOut::d_en, 1);
}
qmckl_exit_code qmckl_get_en_rescaled_single_gl(qmckl_context context, double* distance_rescaled_gl,
                                                const size_t size_max) {
  Context* c = ctx_of(context);
  return get_single(context, distance_rescaled_gl, (int64_t)size_max, "qmckl_get_en_rescaled_single_gl",
                    &SingleOut::en_resc_gl, c ? 4 * c->nucl_num : 0);
}
qmckl_exit_code qmckl_get_jastrow_champ_single_en_gl(qmckl_context context, double* const delta_en_gl,
                                                     const int64_t size_max) {
  Context* c = ctx_of(context);
  return get_single(context, delta_en_gl, size_max, "qmckl_get_jastrow_champ_single_en_gl", &SingleOut::d_en_gl,
                    c ? 4 * c->elec_num : 0);
}

// org/qmckl_jastrow_champ_single.org:8905-9160: the move is accepted.  Electron num of every
// walker takes its new position; the context's Jastrow factors, their GLs, e^J and its GL
// take their new values from the changes (no new full evaluation); every other cached
// result is recomputed on demand.
qmckl_exit_code qmckl_get_jastrow_champ_single_accept(qmckl_context context) {
  const char* fn = "qmckl_get_jastrow_champ_single_accept";
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  qmckl_exit_code rc = provide_jastrow(context, c, fn);  // the factors of the current positions
  if (rc != QMCKL_SUCCESS) return rc;
  SingleOut so;
  rc = provide_single(c, context, fn, so);
  if (rc != QMCKL_SUCCESS) return rc;
  JasDevView v;
  rc = jastrow_view_ready(c, context, fn, v);
  if (rc != QMCKL_SUCCESS) return rc;
  SingleArgs s;
  s.num = c->single_num;
  s.coord = c->d_single_coord.as<double>();
  s.out = so;
  hipError_t e = launch_jastrow_single_accept(v, s, jastrow_out(c), c->d_coord.as<double>(), c->stream);
  if (e != hipSuccess) return hip_fail(context, e, fn);
  // the host copy of the coordinates (when it is the current one)
  const int64_t W = c->walk_num, Ne = c->elec_num, P = c->point_num;
  if (!c->point_on_device)
    for (int64_t w = 0; w < W; ++w)
      for (int d = 0; d < 3; ++d) c->point_coord[d * P + w * Ne + c->single_num] = c->single_coord[3 * w + d];
  // new date: every cached result but the Jastrow's is stale; the device coordinates are current
  c->date += 1;
  c->point_date = c->date;
  c->d_coord_date = c->date;
  c->jastrow_date = c->date;
  c->single_date += 1;
  return QMCKL_SUCCESS;
}

}  // extern "C"
