// Jastrow value and GL of e^J that allocates no memory: the variant of
// qmckl_amd_get_jastrow_champ_value_device / _gl_device for callers that cannot call
// hipMalloc (code under HIP graph capture, tight serving loops, harnesses that forbid
// allocation).  No counterpart in the reference.
//
// The caller passes the coordinates and the two outputs in device memory.  Every
// intermediate of the evaluation (the e-e, e-n and e-e-n factors and their per-electron
// GLs, and the kernels' partial sums) goes to `scratch`, a file-scope device array of
// kScratchBytes that each GPU holds until the process ends; its device address is taken
// once per device, on the first call there, with that device current.  Every call on a GPU
// shares that GPU's array: calls that overlap on one GPU (other streams, other threads)
// race.  The kernels write every scratch element before reading it, so nothing is cleared.
#include <hip/hip_runtime.h>

#include <mutex>

#include "internal.hpp"

namespace qmckl_amd {
namespace {

constexpr int64_t kScratchBytes = int64_t(64) << 20;  // 64 MiB per GPU
constexpr int64_t kScratchDoubles = kScratchBytes / 8;
constexpr int kMaxDevices = 64;

__device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_mutex;
double* g_scratch[kMaxDevices] = {};  // device address of `scratch` on each device

// The address of `scratch` on device dev (current), looked up on the first call there.
hipError_t scratch_on(int dev, double** out) {
  std::lock_guard<std::mutex> g(g_mutex);
  if (g_scratch[dev] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[dev] = static_cast<double*>(p);
  }
  *out = g_scratch[dev];
  return hipSuccess;
}

int64_t up256(int64_t n) { return (n + 31) & ~int64_t(31); }  // doubles, to 256 bytes

}  // namespace
}  // namespace qmckl_amd

using namespace qmckl_amd;

extern "C" {

qmckl_exit_code qmckl_amd_get_jastrow_champ_device_noalloc(qmckl_context context, const int64_t walk_num,
                                                           const double* d_coord, const int64_t coord_size,
                                                           double* const d_value, double* const d_gl,
                                                           const int64_t gl_size) {
  const char* fn = "qmckl_amd_get_jastrow_champ_device_noalloc";
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (!c->jas_provided) return fail(context, QMCKL_NOT_PROVIDED, fn, nullptr);
  if (walk_num <= 0) return fail(context, QMCKL_INVALID_ARG_2, fn, "walk_num should be > 0");
  const int64_t W = walk_num, Ne = c->elec_num;
  if (d_coord == nullptr) return fail(context, QMCKL_INVALID_ARG_3, fn, "NULL pointer");
  if (coord_size < 3 * W * Ne) return fail(context, QMCKL_INVALID_ARG_4, fn, "Array too small: 3*walk_num*elec_num");
  if (d_value == nullptr) return fail(context, QMCKL_INVALID_ARG_5, fn, "NULL pointer");
  if (d_gl == nullptr) return fail(context, QMCKL_INVALID_ARG_6, fn, "NULL pointer");
  if (gl_size < 4 * W * Ne) return fail(context, QMCKL_INVALID_ARG_7, fn, "Array too small: 4*walk_num*elec_num");
  JasDevView v;
  if (!jastrow_view_resident(c, W, d_coord, v))
    return fail(context, QMCKL_NOT_PROVIDED, fn,
                "the context's Jastrow tables are not on the device (or are older than its parameters): "
                "evaluate the Jastrow once through a device getter first");
  // scratch: factors [3][W], factor GLs [3][W][4][Ne], then the kernels' partial sums
  const int64_t off_gl = up256(3 * W), off_s = up256(off_gl + 12 * W * Ne);
  const int64_t need = off_s + jastrow_scratch_doubles(v);
  if (need > kScratchDoubles)
    return fail(context, QMCKL_INVALID_ARG_2, fn,
                "walk_num too large: the evaluation needs more than the library's 64 MiB static scratch");
  if (c->device >= kMaxDevices) return fail(context, QMCKL_FAILURE, fn, "device index beyond the scratch table");
  hipError_t e = hipSetDevice(c->device);
  double* s = nullptr;
  if (e == hipSuccess) e = scratch_on(c->device, &s);
  if (e != hipSuccess) return hip_fail(context, e, fn);
  JasOut o;
  o.fee = s;
  o.fen = s + W;
  o.feen = s + 2 * W;
  o.fee_gl = s + off_gl;
  o.fen_gl = o.fee_gl + 4 * W * Ne;
  o.feen_gl = o.fen_gl + 4 * W * Ne;
  o.value = d_value;
  o.gl = d_gl;
  o.scratch = s + off_s;
  e = launch_jastrow_pair(v, o, c->stream);
  if (e == hipSuccess) e = launch_jastrow_een(v, o, c->stream);
  if (e == hipSuccess) e = launch_jastrow_total(v, o, c->stream);
  if (e != hipSuccess) return hip_fail(context, e, fn);
  return QMCKL_SUCCESS;
}

}  // extern "C"
