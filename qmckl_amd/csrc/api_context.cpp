// Context, error record, numerical precision, points, nuclei and electrons.
// Host side of the C-ABI; device work is delegated to kernels_*.hip.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>
#include <thread>

#include "internal.hpp"

namespace qmckl_amd {

static std::mutex g_registry_mutex;
static std::set<Context*> g_registry;

Context* ctx_of(qmckl_context context) {
  if (context == QMCKL_NULL_CONTEXT) return nullptr;
  Context* c = reinterpret_cast<Context*>(context);
  std::lock_guard<std::mutex> g(g_registry_mutex);
  if (g_registry.count(c) == 0 || c->tag != 0xBEEFFACEu) return nullptr;
  return c;
}

qmckl_exit_code fail(qmckl_context context, qmckl_exit_code code, const char* fn, const char* msg) {
  return qmckl_failwith(context, code, fn, msg);
}

hipError_t DevBuf::reserve(size_t n) {
  if (n <= bytes && ptr != nullptr) return hipSuccess;
  release();
  if (n == 0) n = 8;
  hipError_t e = hipMalloc(&ptr, n);
  if (e != hipSuccess) {
    ptr = nullptr;
    bytes = 0;
    return e;
  }
  bytes = n;
  return hipSuccess;
}

void DevBuf::release() {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
}

qmckl_exit_code hip_fail(qmckl_context handle, hipError_t e, const char* fn) {
  char msg[256];
  snprintf(msg, sizeof msg, "HIP error %d: %s", (int)e, hipGetErrorString(e));
  const qmckl_exit_code code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
                                   ? QMCKL_ALLOCATION_FAILED
                                   : QMCKL_FAILURE;
  return fail(handle, code, fn, msg);
}

qmckl_exit_code device_ready(Context* c, qmckl_context handle, const char* fn) {
  hipError_t e;
  if (c->device < 0) {
    int dev = 0;
    e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(handle, e, fn);
    c->device = dev;
  } else {
    e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(handle, e, fn);
  }
  if (c->stream == nullptr) {
    if (c->own_stream == nullptr) {
      e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
      if (e != hipSuccess) return hip_fail(handle, e, fn);
    }
    c->stream = c->own_stream;
  }
  return QMCKL_SUCCESS;
}

qmckl_exit_code upload_coords(Context* c, qmckl_context handle, const char* fn) {
  if (c->point_num <= 0) return fail(handle, QMCKL_NOT_PROVIDED, fn, "no points set");
  if (c->d_coord_date >= c->point_date) return QMCKL_SUCCESS;
  qmckl_exit_code rc = device_ready(c, handle, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  const size_t bytes = sizeof(double) * 3 * c->point_num;
  hipError_t e = c->d_coord.reserve(bytes);
  if (e != hipSuccess) return hip_fail(handle, e, fn);
  e = hipMemcpyAsync(c->d_coord.ptr, c->point_coord.data(), bytes, hipMemcpyHostToDevice, c->stream);
  if (e != hipSuccess) return hip_fail(handle, e, fn);
  c->d_coord_date = c->point_date;
  return QMCKL_SUCCESS;
}

ProfScope::ProfScope(Context* ctx, int s) : c(ctx), slot(s) {
  if (!c->profiling || c->stream == nullptr) return;
  if (hipEventCreate(&start) != hipSuccess) {
    start = nullptr;
    return;
  }
  (void)hipEventRecord(start, c->stream);
}

ProfScope::~ProfScope() {
  if (start == nullptr) return;
  hipEvent_t stop = nullptr;
  if (hipEventCreate(&stop) != hipSuccess) {
    (void)hipEventDestroy(start);
    return;
  }
  (void)hipEventRecord(stop, c->stream);
  c->prof_pending.push_back({slot, start, stop});
}

static void drain_profile(Context* c) {
  for (ProfEvent& p : c->prof_pending) {
    float ms = 0.f;
    if (hipEventSynchronize(p.stop) == hipSuccess && hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
      c->prof_ms[p.slot] += ms;
      c->prof_count[p.slot] += 1;
    }
    (void)hipEventDestroy(p.start);
    (void)hipEventDestroy(p.stop);
  }
  c->prof_pending.clear();
}

namespace {

constexpr size_t kStageBytes = size_t(32) << 20;  // per pinned staging buffer
constexpr size_t kStagedMin = size_t(16) << 20;   // smaller copies go straight to dst

// memcpy of n bytes split over up to 8 host threads (one for n < 4 MiB); a slice whose
// thread cannot be started is copied by the calling thread, so nothing throws across the
// C-ABI
void host_copy(char* dst, const char* src, size_t n) {
  const unsigned hw = std::thread::hardware_concurrency();
  const int T = n < (size_t(4) << 20) ? 1 : (int)std::min<unsigned>(8u, hw ? hw : 1u);
  if (T == 1) {
    std::memcpy(dst, src, n);
    return;
  }
  const size_t part = ((n + T - 1) / T + 63) & ~size_t(63);
  std::thread th[8];
  for (int t = 1; t < T; ++t) {
    const size_t o = (size_t)t * part;
    if (o >= n) continue;
    const size_t len = std::min(part, n - o);
    try {
      th[t] = std::thread([=] { std::memcpy(dst + o, src + o, len); });
    } catch (...) {
      std::memcpy(dst + o, src + o, len);
    }
  }
  std::memcpy(dst, src, std::min(part, n));
  for (int t = 1; t < T; ++t)
    if (th[t].joinable()) th[t].join();
}

}  // namespace

hipError_t copy_to_host(Context* c, void* dst, const void* d_src, size_t bytes) {
  hipError_t e = hipSuccess;
  if (bytes < kStagedMin) {
    e = hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e;
  }
  for (int b = 0; b < 2 && e == hipSuccess; ++b) {
    if (!c->stage[b]) e = hipHostMalloc(&c->stage[b], kStageBytes, hipHostMallocDefault);
    if (e == hipSuccess && !c->stage_ev[b]) e = hipEventCreateWithFlags(&c->stage_ev[b], hipEventDisableTiming);
  }
  if (e != hipSuccess) return e;
  const char* src = static_cast<const char*>(d_src);
  char* out = static_cast<char*>(dst);
  const size_t nch = (bytes + kStageBytes - 1) / kStageBytes;
  // chunk k goes to stage[k & 1]; that buffer's previous chunk (k - 2) was copied out to
  // dst in the iteration before, so it is free when chunk k is enqueued
  for (size_t k = 0; k <= nch && e == hipSuccess; ++k) {
    if (k < nch) {
      const size_t off = k * kStageBytes, len = std::min(kStageBytes, bytes - off);
      e = hipMemcpyAsync(c->stage[k & 1], src + off, len, hipMemcpyDeviceToHost, c->stream);
      if (e == hipSuccess) e = hipEventRecord(c->stage_ev[k & 1], c->stream);
    }
    if (e == hipSuccess && k > 0) {
      const size_t off = (k - 1) * kStageBytes, len = std::min(kStageBytes, bytes - off);
      e = hipEventSynchronize(c->stage_ev[(k - 1) & 1]);
      if (e == hipSuccess) host_copy(out + off, static_cast<const char*>(c->stage[(k - 1) & 1]), len);
    }
  }
  if (e != hipSuccess) (void)hipStreamSynchronize(c->stream);  // no copy left in flight
  return e;
}

std::vector<DevBuf*> device_buffers(Context* c) {
  DevConst& d = c->dc;
  std::vector<DevBuf*> bufs = {&d.shell_nucl, &d.shell_l, &d.shell_prim_start, &d.shell_prim_count,
                    &d.shell_ao_start, &d.shell_rpow, &d.prim_expo, &d.prim_coef, &d.ao_shell, &d.ao_pow,
                    &d.ao_factor, &d.nucl_range, &d.nucl_coord, &d.coef_t, &d.coef_ch64,
                    &d.coef_ch32, &d.coef_chbf, &d.ao_desc, &d.ao_desc_off, &d.jas_a, &d.jas_b,
                    &d.jas_cfull, &d.jas_cfull_pad, &d.jas_nuc_pad, &d.jas_type, &d.jas_kappa_en,
                    &c->d_coord, &c->d_ao_vgl, &c->d_ao_work, &c->d_mo_vgl, &c->d_mo_value, &c->d_ee_dist,
                    &c->d_en_dist, &c->d_fee, &c->d_fen, &c->d_feen, &c->d_fee_gl, &c->d_fen_gl,
                    &c->d_feen_gl, &c->d_jvalue, &c->d_jgl, &c->d_jscratch, &c->d_jint,
                    &d.jas_asymp_a, &d.prim_nucl_c, &d.prim_expo_c, &d.jas_charge,
                                &c->d_single_coord, &c->d_single_out, &c->d_det_idx, &c->d_det_value,
                                &c->d_det_inv, &c->d_det_work, &c->d_det_out, &c->d_det_drift,
                                &c->d_det_kin, &c->d_energy, &c->d_pot_charge, &c->d_smw, &c->d_cusp,
                                &c->d_ao_hess, &c->d_orb_out, &c->d_nucl_ao};
  for (DevBuf& b : c->d_deriv) bufs.push_back(&b);
  return bufs;
}

static void release_device(Context* c) {
  drain_profile(c);
  for (DevBuf* b : device_buffers(c)) b->release();
  for (int b = 0; b < 2; ++b) {
    if (c->stage[b]) (void)hipHostFree(c->stage[b]);
    if (c->stage_ev[b]) (void)hipEventDestroy(c->stage_ev[b]);
    c->stage[b] = nullptr;
    c->stage_ev[b] = nullptr;
  }
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  c->own_stream = nullptr;
  c->stream = nullptr;
}

}  // namespace qmckl_amd

using namespace qmckl_amd;

extern "C" {

/* ------------------------------------------------------------------ context */
// org/qmckl_context.org:221-240
qmckl_context qmckl_context_check(const qmckl_context context) {
  return ctx_of(context) ? context : QMCKL_NULL_CONTEXT;
}

// org/qmckl_context.org:253-278
qmckl_exit_code qmckl_context_touch(const qmckl_context context) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  touch_const(c);
  return QMCKL_SUCCESS;
}

// org/qmckl_context.org:302-398
qmckl_context qmckl_context_create(void) {
  Context* c = new (std::nothrow) Context();
  if (!c) return QMCKL_NULL_CONTEXT;
  std::lock_guard<std::mutex> g(g_registry_mutex);
  g_registry.insert(c);
  return reinterpret_cast<qmckl_context>(c);
}

// org/qmckl_context.org:428-458
void qmckl_lock(qmckl_context context) {
  Context* c = ctx_of(context);
  if (c) c->mutex.lock();
}
void qmckl_unlock(qmckl_context context) {
  Context* c = ctx_of(context);
  if (c) c->mutex.unlock();
}

// org/qmckl_context.org:467-575: deep copy of the host state; the copy owns no device
// memory until its first computation.
qmckl_context qmckl_context_copy(const qmckl_context context) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_NULL_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  Context* n = new (std::nothrow) Context();
  if (!n) return QMCKL_NULL_CONTEXT;
#define CP(x) n->x = c->x
  CP(date); CP(const_date); CP(num_threads); CP(precision); CP(range);
  CP(point_num); CP(point_coord); CP(point_date);
  CP(nucl_num); CP(nucl_charge); CP(nucl_coord); CP(nucl_uninit); CP(nucl_provided);
  CP(elec_num); CP(up_num); CP(down_num); CP(walk_num); CP(elec_uninit); CP(elec_provided);
  CP(ao_type); CP(shell_num); CP(prim_num); CP(ao_num); CP(nucleus_index); CP(nucleus_shell_num);
  CP(shell_prim_num); CP(shell_prim_index); CP(shell_ang_mom); CP(r_power); CP(shell_factor);
  CP(exponent); CP(coefficient); CP(prim_factor); CP(ao_factor); CP(ao_cartesian); CP(ao_uninit);
  CP(ao_provided); CP(tables);
  CP(mo_num); CP(mo_coef); CP(mo_uninit); CP(mo_provided); CP(mo_precision); CP(mo_tile_g); CP(mo_tile_nw);
  CP(aord); CP(bord); CP(cord); CP(dim_c); CP(type_nucl_num); CP(spin_independent);
  CP(type_nucl_vector); CP(a_vector); CP(b_vector); CP(c_vector); CP(rescale_en); CP(rescale_ee);
  CP(jas_uninit); CP(jas_provided); CP(asymp_jasa);
  CP(single_num); CP(single_coord); CP(single_date);
  CP(det_type); CP(det_num_alpha); CP(det_num_beta); CP(mo_index_alpha); CP(mo_index_beta);
  CP(det_uninit); CP(det_set_date); CP(r_cusp); CP(cusp_param);
#undef CP
  n->mo_index_walk[0] = c->mo_index_walk[0];
  n->mo_index_walk[1] = c->mo_index_walk[1];
  n->asymp_jasb[0] = c->asymp_jasb[0];
  n->asymp_jasb[1] = c->asymp_jasb[1];
  if (c->point_on_device && c->point_num > 0) {
    // Bring device-only coordinates back to the host so the copy is self-contained.
    // The device coordinates were written on c->stream (a non-blocking stream, or a batch
    // replay ordered on it): read them back on that stream, then wait for it.
    n->point_coord.resize(3 * c->point_num);
    hipError_t e = hipMemcpyAsync(n->point_coord.data(), c->d_coord.ptr, sizeof(double) * 3 * c->point_num,
                                  hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
      delete n;
      return QMCKL_NULL_CONTEXT;
    }
  }
  n->point_date = n->date;
  std::lock_guard<std::mutex> g2(g_registry_mutex);
  g_registry.insert(n);
  return reinterpret_cast<qmckl_context>(n);
}

// org/qmckl_context.org:599-760
qmckl_exit_code qmckl_context_get_num_threads(const qmckl_context context, int32_t* const num_threads) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (num_threads == nullptr) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_context_get_num_threads", "null pointer");
  *num_threads = c->num_threads;
  return QMCKL_SUCCESS;
}

qmckl_exit_code qmckl_context_set_num_threads(const qmckl_context context, const int32_t num_threads) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (num_threads <= 0) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_context_set_num_threads", "num_threads <= 0");
  c->num_threads = num_threads;  // host threads are not used by this implementation
  return QMCKL_SUCCESS;
}

// org/qmckl_context.org:690-719: back to the default thread count of a new context
qmckl_exit_code qmckl_context_unset_num_threads(const qmckl_context context) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  c->num_threads = Context::kDefaultNumThreads;
  return QMCKL_SUCCESS;
}

// org/qmckl_context.org:769-815
qmckl_exit_code qmckl_context_destroy(const qmckl_context context) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  {
    std::lock_guard<std::mutex> g(g_registry_mutex);
    g_registry.erase(c);
  }
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  release_device(c);
  c->tag = 0;
  delete c;
  return QMCKL_SUCCESS;
}

/* ------------------------------------------------------------------ errors */
// org/qmckl_error.org:243-380
const char* qmckl_string_of_error(const qmckl_exit_code error) {
  static const char* args[] = {"Invalid argument 1", "Invalid argument 2", "Invalid argument 3",
                               "Invalid argument 4", "Invalid argument 5", "Invalid argument 6",
                               "Invalid argument 7", "Invalid argument 8", "Invalid argument 9",
                               "Invalid argument 10", "Invalid argument 11", "Invalid argument 12",
                               "Invalid argument 13", "Invalid argument 14", "Invalid argument 15",
                               "Invalid argument 16", "Invalid argument 17", "Invalid argument 18",
                               "Invalid argument 19", "Invalid argument 20"};
  if (error == QMCKL_SUCCESS) return "Success";
  if (error >= 1 && error <= 20) return args[error - 1];
  switch (error) {
    case QMCKL_FAILURE: return "Failure";
    case QMCKL_ERRNO: return strerror(errno);
    case QMCKL_INVALID_CONTEXT: return "Invalid context";
    case QMCKL_ALLOCATION_FAILED: return "Allocation failed";
    case QMCKL_DEALLOCATION_FAILED: return "De-allocation failed";
    case QMCKL_NOT_PROVIDED: return "Not provided";
    case QMCKL_OUT_OF_BOUNDS: return "Index out of bounds";
    case QMCKL_ALREADY_SET: return "Already set";
    case QMCKL_INVALID_EXIT_CODE: return "Invalid exit code";
    default: return "Unknown error";
  }
}

// org/qmckl_error.org:433-470
qmckl_exit_code qmckl_set_error(qmckl_context context, const qmckl_exit_code exit_code,
                                const char* function_name, const char* message) {
  if (exit_code == QMCKL_SUCCESS) return QMCKL_INVALID_ARG_2;
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  c->err_code = exit_code;
  c->err_function = function_name ? std::string(function_name).substr(0, QMCKL_MAX_FUN_LEN - 1) : "";
  c->err_message = message ? std::string(message).substr(0, QMCKL_MAX_MSG_LEN - 1) : "";
  return QMCKL_SUCCESS;
}

// org/qmckl_error.org:493-530
qmckl_exit_code qmckl_get_error(qmckl_context context, qmckl_exit_code* exit_code, char* function_name,
                                char* message) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (!exit_code || !function_name || !message) return QMCKL_INVALID_ARG_2;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  *exit_code = c->err_code;
  memcpy(function_name, c->err_function.c_str(), c->err_function.size() + 1);
  memcpy(message, c->err_message.c_str(), c->err_message.size() + 1);
  return QMCKL_SUCCESS;
}

// org/qmckl_error.org:550-590
qmckl_exit_code qmckl_failwith(qmckl_context context, const qmckl_exit_code exit_code, const char* function,
                               const char* message) {
  if (exit_code == QMCKL_SUCCESS) return QMCKL_INVALID_EXIT_CODE;
  if (ctx_of(context) == nullptr) return QMCKL_INVALID_CONTEXT;
  (void)qmckl_set_error(context, exit_code, function, message);
  return exit_code;
}

// org/qmckl_error.org:606-637
qmckl_exit_code qmckl_last_error(qmckl_context context, char* buffer) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (!buffer) return QMCKL_INVALID_ARG_2;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (c->err_code == QMCKL_SUCCESS) {
    buffer[0] = 0;
  } else {
    sprintf(buffer, "Error -- %s -- in %s\n%s", qmckl_string_of_error(c->err_code),
            c->err_function.c_str(), c->err_message.c_str());
  }
  return QMCKL_SUCCESS;
}

// org/qmckl_error.org:660-690
qmckl_exit_code qmckl_check(qmckl_context context, qmckl_exit_code rc) {
  if (rc != QMCKL_SUCCESS) {
    char buffer[QMCKL_MAX_FUN_LEN + QMCKL_MAX_MSG_LEN + 64];
    if (qmckl_last_error(context, buffer) == QMCKL_SUCCESS) fprintf(stderr, "===\n%s\n===\n", buffer);
  }
  return rc;
}

/* ------------------------------------------------------------------ numprec */
// org/qmckl_numprec.org:170-340
qmckl_exit_code qmckl_set_numprec_precision(const qmckl_context context, const int precision) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (precision < 2) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_set_numprec_precision", "precision < 2");
  if (precision > 53) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_set_numprec_precision", "precision > 53");
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  c->precision = precision;
  touch_const(c);  // the AO screening radius depends on it
  return QMCKL_SUCCESS;
}

int32_t qmckl_get_numprec_precision(const qmckl_context context) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  return c->precision;
}

qmckl_exit_code qmckl_set_numprec_range(const qmckl_context context, const int range) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (range < 2) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_set_numprec_range", "range < 2");
  if (range > 11) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_set_numprec_range", "range > 11");
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  c->range = range;
  return QMCKL_SUCCESS;
}

int32_t qmckl_get_numprec_range(const qmckl_context context) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  return c->range;
}

// org/qmckl_numprec.org:348-365: epsilon = 2^(1-precision)
double qmckl_get_numprec_epsilon(const qmckl_context context) {
  const int32_t p = qmckl_get_numprec_precision(context);
  return ldexp(1.0, 1 - p);
}

// org/qmckl_numprec.org:395-440: bits of agreement from the distance in units in the last place
int32_t qmckl_test_precision_64(double a, double b) {
  int64_t ia, ib;
  memcpy(&ia, &a, 8);
  memcpy(&ib, &b, 8);
  int64_t diff;
  if ((ia < 0) != (ib < 0) && a != 0.0 && b != 0.0) {
    const int64_t da = ia < 0 ? INT64_MAX + ia : INT64_MAX - ia;
    const int64_t db = ib < 0 ? INT64_MAX + ib : INT64_MAX - ib;
    diff = da + db;
  } else {
    diff = ia - ib;
    diff = diff > 0 ? diff : -diff;
  }
  if (diff == 0) return 53;
  int32_t result = 53;
  for (int i = 0; i < 53 && diff != 0; ++i) {
    diff >>= 1;
    result--;
  }
  return result;
}

int32_t qmckl_test_precision_32(float a, float b) { return qmckl_test_precision_64((double)a, (double)b); }

/* ------------------------------------------------------------------ points */
// org/qmckl_point.org:143-170
qmckl_exit_code qmckl_get_point_num(const qmckl_context context, int64_t* const num) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (num == nullptr) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_get_point_num", "num is a null pointer");
  *num = c->point_num;
  return QMCKL_SUCCESS;
}

static qmckl_exit_code sync_host_points(Context* c, qmckl_context context, const char* fn) {
  if (!c->point_on_device) return QMCKL_SUCCESS;
  c->point_coord.resize(3 * c->point_num);
  hipError_t e = hipMemcpyAsync(c->point_coord.data(), c->d_coord.ptr, sizeof(double) * 3 * c->point_num,
                                hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_fail(context, e, fn);
  c->point_on_device = false;
  return QMCKL_SUCCESS;
}

// org/qmckl_point.org:188-264
qmckl_exit_code qmckl_get_point(const qmckl_context context, const char transp, double* const coord,
                                const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (transp != 'N' && transp != 'T')
    return fail(context, QMCKL_INVALID_ARG_2, "qmckl_get_point", "transp should be 'N' or 'T'");
  if (coord == nullptr) return fail(context, QMCKL_INVALID_ARG_3, "qmckl_get_point", "coord is a null pointer");
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  const int64_t n = c->point_num;
  if (size_max < 3 * n) return fail(context, QMCKL_INVALID_ARG_4, "qmckl_get_point", "Array too small");
  qmckl_exit_code rc = sync_host_points(c, context, "qmckl_get_point");
  if (rc != QMCKL_SUCCESS) return rc;
  if (transp == 'T') {
    memcpy(coord, c->point_coord.data(), sizeof(double) * 3 * n);
  } else {
    for (int64_t i = 0; i < n; ++i)
      for (int d = 0; d < 3; ++d) coord[3 * i + d] = c->point_coord[d * n + i];
  }
  return QMCKL_SUCCESS;
}

// org/qmckl_point.org:288-375
qmckl_exit_code qmckl_set_point(qmckl_context context, const char transp, const int64_t num, const double* coord,
                                const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (num <= 0) return fail(context, QMCKL_INVALID_ARG_3, "qmckl_set_point", "Number of points should be >0.");
  if (size_max < 3 * num) return fail(context, QMCKL_INVALID_ARG_4, "qmckl_set_point", "Array too small");
  if (transp != 'N' && transp != 'T')
    return fail(context, QMCKL_INVALID_ARG_2, "qmckl_set_point", "transp should be 'N' or 'T'");
  if (coord == nullptr) return fail(context, QMCKL_INVALID_ARG_3, "qmckl_set_point", "coord is a NULL pointer");
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  c->point_num = num;
  c->point_coord.resize(3 * num);
  if (transp == 'T') {
    memcpy(c->point_coord.data(), coord, sizeof(double) * 3 * num);
  } else {
    for (int64_t i = 0; i < num; ++i)
      for (int d = 0; d < 3; ++d) c->point_coord[d * num + i] = coord[3 * i + d];
  }
  c->point_on_device = false;
  c->date += 1;  // coordinates only: the device constant tables stay valid
  c->point_date = c->date;
  return QMCKL_SUCCESS;
}

/* ------------------------------------------------------------------ nucleus */
// org/qmckl_nucleus.org:367-380
bool qmckl_nucleus_provided(const qmckl_context context) {
  Context* c = ctx_of(context);
  return c ? c->nucl_provided : false;
}

static qmckl_exit_code nucl_post(Context* c, uint32_t mask) {
  c->nucl_uninit &= ~mask;
  c->nucl_provided = (c->nucl_uninit == 0);
  touch_const(c);
  return QMCKL_SUCCESS;
}

// org/qmckl_nucleus.org:417-460
qmckl_exit_code qmckl_set_nucleus_num(qmckl_context context, const int64_t num) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (!(c->nucl_uninit & 1u)) return fail(context, QMCKL_ALREADY_SET, "qmckl_set_nucleus_*", nullptr);
  if (num <= 0) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_set_nucleus_num", "num <= 0");
  c->nucl_num = num;
  return nucl_post(c, 1u);
}

// org/qmckl_nucleus.org:461-525
qmckl_exit_code qmckl_set_nucleus_charge(qmckl_context context, const double* charge, const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (!(c->nucl_uninit & 2u)) return fail(context, QMCKL_ALREADY_SET, "qmckl_set_nucleus_*", nullptr);
  if (charge == nullptr)
    return fail(context, QMCKL_INVALID_ARG_2, "qmckl_set_nucleus_charge", "charge is a null pointer");
  if (c->nucl_num <= 0) return fail(context, QMCKL_FAILURE, "qmckl_set_nucleus_charge", "nucl_num is not set");
  if (c->nucl_num > size_max) return fail(context, QMCKL_INVALID_ARG_3, "qmckl_set_nucleus_charge", "Array too small");
  c->nucl_charge.assign(charge, charge + c->nucl_num);
  return nucl_post(c, 2u);
}

// org/qmckl_nucleus.org:529-600 (may be set again: nuclear derivatives move the nuclei)
qmckl_exit_code qmckl_set_nucleus_coord(qmckl_context context, const char transp, const double* coord,
                                        const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  const int64_t n = c->nucl_num;
  if (n <= 0) return fail(context, QMCKL_FAILURE, "qmckl_set_nucleus_coord", "nucl_num is not set");
  if (transp != 'N' && transp != 'T')
    return fail(context, QMCKL_INVALID_ARG_2, "qmckl_set_nucleus_coord", "transp should be 'N' or 'T'");
  if (coord == nullptr) return fail(context, QMCKL_INVALID_ARG_3, "qmckl_set_nucleus_coord", "coord is a null pointer");
  if (size_max < 3 * n) return fail(context, QMCKL_INVALID_ARG_4, "qmckl_set_nucleus_coord", "Array too small");
  c->nucl_coord.resize(3 * n);
  if (transp == 'T') {
    memcpy(c->nucl_coord.data(), coord, sizeof(double) * 3 * n);
  } else {
    for (int64_t i = 0; i < n; ++i)
      for (int d = 0; d < 3; ++d) c->nucl_coord[d * n + i] = coord[3 * i + d];
  }
  return nucl_post(c, 4u);
}

// org/qmckl_nucleus.org:161-200
qmckl_exit_code qmckl_get_nucleus_num(const qmckl_context context, int64_t* const num) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (num == nullptr) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_get_nucleus_num", "num is a null pointer");
  if (c->nucl_uninit & 1u) {
    *num = 0;
    return fail(context, QMCKL_NOT_PROVIDED, "qmckl_get_nucleus_num", "nucleus data is not provided");
  }
  *num = c->nucl_num;
  return QMCKL_SUCCESS;
}

// org/qmckl_nucleus.org:216-270
qmckl_exit_code qmckl_get_nucleus_charge(const qmckl_context context, double* const charge, const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (charge == nullptr) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_get_nucleus_charge", "charge is a null pointer");
  if (c->nucl_uninit & 2u) return fail(context, QMCKL_NOT_PROVIDED, "qmckl_get_nucleus_charge", "nucleus data is not provided");
  if (size_max < c->nucl_num) return fail(context, QMCKL_INVALID_ARG_3, "qmckl_get_nucleus_charge", "Array too small");
  memcpy(charge, c->nucl_charge.data(), sizeof(double) * c->nucl_num);
  return QMCKL_SUCCESS;
}

// org/qmckl_nucleus.org:285-350
qmckl_exit_code qmckl_get_nucleus_coord(const qmckl_context context, const char transp, double* const coord,
                                        const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (transp != 'N' && transp != 'T')
    return fail(context, QMCKL_INVALID_ARG_2, "qmckl_get_nucleus_coord", "transp should be 'N' or 'T'");
  if (coord == nullptr) return fail(context, QMCKL_INVALID_ARG_3, "qmckl_get_nucleus_coord", "coord is a null pointer");
  if (c->nucl_uninit & 4u) return fail(context, QMCKL_NOT_PROVIDED, "qmckl_get_nucleus_coord", "nucleus data is not provided");
  const int64_t n = c->nucl_num;
  if (size_max < 3 * n) return fail(context, QMCKL_INVALID_ARG_4, "qmckl_get_nucleus_coord", "Array too small");
  if (transp == 'T') {
    memcpy(coord, c->nucl_coord.data(), sizeof(double) * 3 * n);
  } else {
    for (int64_t i = 0; i < n; ++i)
      for (int d = 0; d < 3; ++d) coord[3 * i + d] = c->nucl_coord[d * n + i];
  }
  return QMCKL_SUCCESS;
}

// org/qmckl_nucleus.org:776-931: nn_distance[a][b] (a few nuclei: computed on the host)
qmckl_exit_code qmckl_get_nucleus_nn_distance(qmckl_context context, double* distance, const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (!c->nucl_provided) return fail(context, QMCKL_NOT_PROVIDED, "qmckl_get_nucleus_nn_distance", nullptr);
  const int64_t n = c->nucl_num;
  if (distance == nullptr) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_get_nucleus_nn_distance", "null pointer");
  if (size_max < n * n) return fail(context, QMCKL_INVALID_ARG_3, "qmckl_get_nucleus_nn_distance", "Array too small");
  const double* x = c->nucl_coord.data();
  for (int64_t j = 0; j < n; ++j)
    for (int64_t i = 0; i < n; ++i) {
      const double dx = x[i] - x[j], dy = x[n + i] - x[n + j], dz = x[2 * n + i] - x[2 * n + j];
      distance[j * n + i] = sqrt(dx * dx + dy * dy + dz * dz);
    }
  return QMCKL_SUCCESS;
}

// org/qmckl_nucleus.org:946-1090: V_nn = sum_{a<b} Z_a Z_b / R_ab
qmckl_exit_code qmckl_get_nucleus_repulsion(qmckl_context context, double* const energy) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (!c->nucl_provided) return fail(context, QMCKL_NOT_PROVIDED, "qmckl_get_nucleus_repulsion", nullptr);
  if (energy == nullptr) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_get_nucleus_repulsion", "null pointer");
  const int64_t n = c->nucl_num;
  const double* x = c->nucl_coord.data();
  double e = 0.0;
  for (int64_t j = 1; j < n; ++j)
    for (int64_t i = 0; i < j; ++i) {
      const double dx = x[i] - x[j], dy = x[n + i] - x[n + j], dz = x[2 * n + i] - x[2 * n + j];
      e += c->nucl_charge[i] * c->nucl_charge[j] / sqrt(dx * dx + dy * dy + dz * dz);
    }
  *energy = e;
  return QMCKL_SUCCESS;
}

/* ------------------------------------------------------------------ electrons */
// org/qmckl_electron.org:197-212
bool qmckl_electron_provided(const qmckl_context context) {
  Context* c = ctx_of(context);
  return c ? c->elec_provided : false;
}

// org/qmckl_electron.org:256-290
qmckl_exit_code qmckl_set_electron_num(qmckl_context context, const int64_t up_num, const int64_t down_num) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (!(c->elec_uninit & 1u)) return fail(context, QMCKL_ALREADY_SET, "qmckl_set_electron_*", nullptr);
  if (up_num <= 0) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_set_electron_num", "up_num <= 0");
  if (down_num < 0) return fail(context, QMCKL_INVALID_ARG_3, "qmckl_set_electron_num", "down_num < 0");
  c->up_num = up_num;
  c->down_num = down_num;
  c->elec_num = up_num + down_num;
  c->elec_uninit &= ~1u;
  c->elec_provided = (c->elec_uninit == 0);
  touch_const(c);
  return QMCKL_SUCCESS;
}

// org/qmckl_electron.org:315-375: the coordinates of all walkers become the points.
qmckl_exit_code qmckl_set_electron_coord(qmckl_context context, const char transp, const int64_t walk_num,
                                         const double* coord, const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (transp != 'N' && transp != 'T')
    return fail(context, QMCKL_INVALID_ARG_2, "qmckl_set_electron_coord", "transp should be 'N' or 'T'");
  if (walk_num <= 0) return fail(context, QMCKL_INVALID_ARG_3, "qmckl_set_electron_coord", "walk_num <= 0");
  if (coord == nullptr) return fail(context, QMCKL_INVALID_ARG_4, "qmckl_set_electron_coord", "coord is a null pointer");
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (c->elec_num == 0) return fail(context, QMCKL_FAILURE, "qmckl_set_electron_coord", "elec_num is not set");
  qmckl_exit_code rc = qmckl_set_point(context, transp, walk_num * c->elec_num, coord, size_max);
  if (rc != QMCKL_SUCCESS) return rc;
  c->walk_num = walk_num;
  return QMCKL_SUCCESS;
}

// org/qmckl_electron.org:406-470
qmckl_exit_code qmckl_get_electron_num(const qmckl_context context, int64_t* const num) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (num == nullptr) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_get_electron_num", "num is a null pointer");
  if (c->elec_uninit & 1u) return fail(context, QMCKL_NOT_PROVIDED, "qmckl_get_electron_num", "elec_num is not provided");
  *num = c->elec_num;
  return QMCKL_SUCCESS;
}
qmckl_exit_code qmckl_get_electron_up_num(const qmckl_context context, int64_t* const up_num) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (up_num == nullptr) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_get_electron_up_num", "up_num is a null pointer");
  if (c->elec_uninit & 1u) return fail(context, QMCKL_NOT_PROVIDED, "qmckl_get_electron_up_num", "up_num is not provided");
  *up_num = c->up_num;
  return QMCKL_SUCCESS;
}
qmckl_exit_code qmckl_get_electron_down_num(const qmckl_context context, int64_t* const down_num) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (down_num == nullptr) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_get_electron_down_num", "down_num is a null pointer");
  if (c->elec_uninit & 1u) return fail(context, QMCKL_NOT_PROVIDED, "qmckl_get_electron_down_num", "down_num is not provided");
  *down_num = c->down_num;
  return QMCKL_SUCCESS;
}

// org/qmckl_electron.org:504-540
qmckl_exit_code qmckl_get_electron_walk_num(const qmckl_context context, int64_t* const walk_num) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (walk_num == nullptr) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_get_electron_walk_num", "walk_num is a null pointer");
  *walk_num = c->walk_num;
  return QMCKL_SUCCESS;
}

// org/qmckl_electron.org:543-600
qmckl_exit_code qmckl_get_electron_coord(const qmckl_context context, const char transp, double* const coord,
                                         const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (transp != 'N' && transp != 'T')
    return fail(context, QMCKL_INVALID_ARG_2, "qmckl_get_electron_coord", "transp should be 'N' or 'T'");
  if (coord == nullptr) return fail(context, QMCKL_INVALID_ARG_3, "qmckl_get_electron_coord", "coord is a null pointer");
  if (!c->elec_provided) return fail(context, QMCKL_NOT_PROVIDED, "qmckl_get_electron_coord", nullptr);
  if (c->walk_num <= 0) return fail(context, QMCKL_NOT_PROVIDED, "qmckl_get_electron_coord", "coordinates not set");
  return qmckl_get_point(context, transp, coord, size_max);
}

static qmckl_exit_code copy_out(Context* c, qmckl_context context, const void* d_src, double* dst, int64_t n,
                                const char* fn) {
  const hipError_t e = copy_to_host(c, dst, d_src, sizeof(double) * n);
  if (e != hipSuccess) return hip_fail(context, e, fn);
  return QMCKL_SUCCESS;
}

// org/qmckl_electron.org:776-960: ee_distance[walk][elec][elec]
qmckl_exit_code qmckl_get_electron_ee_distance(qmckl_context context, double* const distance, const int64_t size_max) {
  const char* fn = "qmckl_get_electron_ee_distance";
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (distance == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "null pointer");
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (!c->elec_provided || c->walk_num <= 0) return fail(context, QMCKL_NOT_PROVIDED, fn, nullptr);
  const int64_t n = c->walk_num * c->elec_num * c->elec_num;
  if (size_max < n) return fail(context, QMCKL_INVALID_ARG_3, fn, "Array too small");
  qmckl_exit_code rc = provide_ee_distance(context, c, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  return copy_out(c, context, c->d_ee_dist.ptr, distance, n, fn);
}

// org/qmckl_electron.org:1203-1390: en_distance[walk*elec][nucl]
qmckl_exit_code qmckl_get_electron_en_distance(qmckl_context context, double* const distance, const int64_t size_max) {
  const char* fn = "qmckl_get_electron_en_distance";
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (distance == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "null pointer");
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (!c->elec_provided || c->walk_num <= 0 || !c->nucl_provided) return fail(context, QMCKL_NOT_PROVIDED, fn, nullptr);
  const int64_t n = c->point_num * c->nucl_num;
  if (size_max < n) return fail(context, QMCKL_INVALID_ARG_3, fn, "Array too small");
  qmckl_exit_code rc = provide_en_distance(context, c, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  return copy_out(c, context, c->d_en_dist.ptr, distance, n, fn);
}

// org/qmckl_electron.org:1022-1192: V_ee[walk] = sum_{i<j, r_ij > 1e-5} 1/r_ij, and
// org/qmckl_electron.org:1477-1658: V_en[walk] = -sum_{i,a, r_ia > 1e-6} Z_a / r_ia
// (coincident particles left out, :1172, :1638), one workgroup per walker on the device.
static qmckl_exit_code get_potential(qmckl_context context, double* out, int which, const char* fn) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (out == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "null pointer");
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (!c->elec_provided || c->walk_num <= 0 || (which == 4 && !c->nucl_provided))
    return fail(context, QMCKL_NOT_PROVIDED, fn, nullptr);
  qmckl_exit_code rc = provide_energy(c, context, fn, false);
  if (rc != QMCKL_SUCCESS) return rc;
  const hipError_t e = copy_to_host(c, out, c->d_energy.as<double>() + which * c->walk_num, sizeof(double) * c->walk_num);
  if (e != hipSuccess) return hip_fail(context, e, fn);
  return QMCKL_SUCCESS;
}

qmckl_exit_code qmckl_get_electron_ee_potential(qmckl_context context, double* const ee_potential) {
  return get_potential(context, ee_potential, 3, "qmckl_get_electron_ee_potential");
}

qmckl_exit_code qmckl_get_electron_en_potential(qmckl_context context, double* const en_potential) {
  return get_potential(context, en_potential, 4, "qmckl_get_electron_en_potential");
}

/* ------------------------------------------------------------------ device extensions */
qmckl_exit_code qmckl_amd_set_stream(qmckl_context context, void* stream) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  c->stream = stream ? static_cast<hipStream_t>(stream) : nullptr;
  return QMCKL_SUCCESS;
}

qmckl_exit_code qmckl_amd_synchronize(qmckl_context context) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (c->stream == nullptr) return QMCKL_SUCCESS;
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_fail(context, e, "qmckl_amd_synchronize");
  return QMCKL_SUCCESS;
}

qmckl_exit_code qmckl_amd_set_profiling(qmckl_context context, const int32_t on) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  qmckl_exit_code rc = device_ready(c, context, "qmckl_amd_set_profiling");
  if (rc != QMCKL_SUCCESS) return rc;
  c->profiling = on != 0;
  return QMCKL_SUCCESS;
}

qmckl_exit_code qmckl_amd_get_kernel_times(qmckl_context context, double* const ms, int64_t* const counts,
                                           const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (ms == nullptr) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_amd_get_kernel_times", "NULL pointer");
  if (counts == nullptr) return fail(context, QMCKL_INVALID_ARG_3, "qmckl_amd_get_kernel_times", "NULL pointer");
  if (size_max < QMCKL_AMD_KERNEL_SLOTS)
    return fail(context, QMCKL_INVALID_ARG_4, "qmckl_amd_get_kernel_times", "Array too small");
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  drain_profile(c);
  for (int i = 0; i < QMCKL_AMD_KERNEL_SLOTS; ++i) {
    ms[i] = c->prof_ms[i];
    counts[i] = c->prof_count[i];
    c->prof_ms[i] = 0.0;
    c->prof_count[i] = 0;
  }
  return QMCKL_SUCCESS;
}

qmckl_exit_code qmckl_amd_set_electron_coord_device(qmckl_context context, const char transp,
                                                    const int64_t walk_num, const double* d_coord,
                                                    const int64_t size_max) {
  const char* fn = "qmckl_amd_set_electron_coord_device";
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (transp != 'T')
    return fail(context, QMCKL_INVALID_ARG_2, fn, "device coordinates must be given as [3][walk_num*elec_num] ('T')");
  if (walk_num <= 0) return fail(context, QMCKL_INVALID_ARG_3, fn, "walk_num <= 0");
  if (d_coord == nullptr) return fail(context, QMCKL_INVALID_ARG_4, fn, "coord is a null pointer");
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (c->elec_num == 0) return fail(context, QMCKL_FAILURE, fn, "elec_num is not set");
  const int64_t num = walk_num * c->elec_num;
  if (size_max < 3 * num) return fail(context, QMCKL_INVALID_ARG_5, fn, "Array too small");
  qmckl_exit_code rc = device_ready(c, context, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  hipError_t e = c->d_coord.reserve(sizeof(double) * 3 * num);
  if (e != hipSuccess) return hip_fail(context, e, fn);
  e = hipMemcpyAsync(c->d_coord.ptr, d_coord, sizeof(double) * 3 * num, hipMemcpyDeviceToDevice, c->stream);
  if (e != hipSuccess) return hip_fail(context, e, fn);
  c->point_num = num;
  c->walk_num = walk_num;
  c->point_on_device = true;
  c->date += 1;  // coordinates only
  c->point_date = c->date;
  c->d_coord_date = c->date;
  return QMCKL_SUCCESS;
}

}  // extern "C"

namespace qmckl_amd {

qmckl_exit_code provide_ee_distance(qmckl_context handle, Context* c, const char* fn) {
  if (c->ee_dist_date >= c->date) return QMCKL_SUCCESS;
  qmckl_exit_code rc = upload_coords(c, handle, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  const int64_t W = c->walk_num, Ne = c->elec_num;
  hipError_t e = c->d_ee_dist.reserve(sizeof(double) * W * Ne * Ne);
  if (e == hipSuccess) e = launch_ee_distance(c->d_coord.as<double>(), W, Ne, c->d_ee_dist.as<double>(), c->stream);
  if (e != hipSuccess) return hip_fail(handle, e, fn);
  c->ee_dist_date = c->date;
  return QMCKL_SUCCESS;
}

qmckl_exit_code provide_en_distance(qmckl_context handle, Context* c, const char* fn) {
  if (c->en_dist_date >= c->date) return QMCKL_SUCCESS;
  qmckl_exit_code rc = upload_coords(c, handle, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  const int64_t P = c->point_num, Nn = c->nucl_num;
  hipError_t e = c->d_en_dist.reserve(sizeof(double) * P * Nn);
  if (e == hipSuccess) e = c->dc.nucl_coord.reserve(sizeof(double) * 3 * Nn);
  if (e == hipSuccess)
    e = hipMemcpyAsync(c->dc.nucl_coord.ptr, c->nucl_coord.data(), sizeof(double) * 3 * Nn,
                       hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess)
    e = launch_en_distance(c->d_coord.as<double>(), P, P, c->dc.nucl_coord.as<double>(), Nn,
                           c->d_en_dist.as<double>(), c->stream);
  if (e != hipSuccess) return hip_fail(handle, e, fn);
  c->en_dist_date = c->date;
  return QMCKL_SUCCESS;
}

}  // namespace qmckl_amd
