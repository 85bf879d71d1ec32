// Batched evaluation of several contexts (molecules of different sizes) replayed from one
// HIP graph.  No counterpart in the reference: it composes the device entry points of each
// context (coordinates in, MO VGL, Jastrow value and GL out) into one graph so that a batch
// of molecules costs one launch.
//
// Recording runs every context's step once eagerly (which sizes and allocates all of its
// device buffers and uploads its constant tables), then captures the same steps.  Molecules
// are spread over up to kBranches streams that fork from and join back to the caller's
// stream, so independent molecules can run concurrently inside one replay.  All streams keep
// the default priority.  Each context's stream is left set to its branch stream.
#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "internal.hpp"

using namespace qmckl_amd;

namespace {

constexpr int kBranches = 4;

// What a recorded graph depends on in one context: its sizes, its constant data and the
// address and size of every device buffer its kernels read or write.  A replay is refused
// when any of these has changed since the recording (a setter, a new walker count, a
// buffer that grew and moved).
struct Snapshot {
  int64_t walk_num = 0, elec_num = 0, mo_num = 0;
  uint64_t const_date = 0;
  int32_t mo_precision = 0;
  std::vector<std::pair<void*, size_t>> buffers;

  static Snapshot of(Context* c) {
    Snapshot s;
    s.walk_num = c->walk_num;
    s.elec_num = c->elec_num;
    s.mo_num = c->mo_num;
    s.const_date = c->const_date;
    s.mo_precision = c->mo_precision;
    for (DevBuf* b : device_buffers(c)) s.buffers.emplace_back(b->ptr, b->bytes);
    return s;
  }
  bool operator==(const Snapshot& o) const {
    return walk_num == o.walk_num && elec_num == o.elec_num && mo_num == o.mo_num &&
           const_date == o.const_date && mo_precision == o.mo_precision && buffers == o.buffers;
  }
};

struct Batch {
  hipStream_t origin = nullptr;
  bool own_origin = false;  // created here when the caller passes no stream
  hipStream_t branch[kBranches] = {};
  hipEvent_t fork = nullptr, join[kBranches] = {};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int nbranch = 0;
  std::vector<qmckl_context> contexts;
  std::vector<Snapshot> recorded;  // per context, taken after the capture
};

std::mutex g_batch_mutex;
std::map<int64_t, Batch*> g_batches;
int64_t g_next_batch = 1;

void destroy_batch(Batch* b) {
  if (!b) return;
  // the contexts go back to their own streams before the branch streams are destroyed
  for (qmckl_context ctx : b->contexts)
    if (ctx_of(ctx)) (void)qmckl_amd_set_stream(ctx, nullptr);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (int i = 0; i < kBranches; ++i) {
    if (b->join[i]) (void)hipEventDestroy(b->join[i]);
    if (b->branch[i]) (void)hipStreamDestroy(b->branch[i]);
  }
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->own_origin && b->origin) (void)hipStreamDestroy(b->origin);
  delete b;
}

// One device step of context i, as the four device entry points.
qmckl_exit_code step(qmckl_context ctx, const double* d_coord, double* d_mo, double* d_value, double* d_gl) {
  Context* c = ctx_of(ctx);
  const int64_t W = c->walk_num, Ne = c->elec_num, P = W * Ne;
  qmckl_exit_code rc = qmckl_amd_set_electron_coord_device(ctx, 'T', W, d_coord, 3 * P);
  if (rc == QMCKL_SUCCESS && d_mo != nullptr)
    rc = qmckl_amd_get_mo_basis_mo_vgl_device(ctx, d_mo, P * 5 * c->mo_num);
  if (rc == QMCKL_SUCCESS && d_value != nullptr) rc = qmckl_amd_get_jastrow_champ_value_device(ctx, d_value, W);
  if (rc == QMCKL_SUCCESS && d_gl != nullptr) rc = qmckl_amd_get_jastrow_champ_gl_device(ctx, d_gl, W * 4 * Ne);
  return rc;
}

}  // namespace

extern "C" {

qmckl_exit_code qmckl_amd_batch_record(const qmckl_context* contexts, const int64_t n, void* stream,
                                       const double* const* d_coord, double* const* d_mo_vgl,
                                       double* const* d_value, double* const* d_gl, int64_t* const batch) {
  const char* fn = "qmckl_amd_batch_record";
  if (contexts == nullptr || n <= 0) return QMCKL_INVALID_ARG_1;
  if (d_coord == nullptr) return QMCKL_INVALID_ARG_4;
  if (batch == nullptr) return QMCKL_INVALID_ARG_8;
  for (int64_t i = 0; i < n; ++i) {
    Context* c = ctx_of(contexts[i]);
    if (!c) return QMCKL_INVALID_CONTEXT;
    if (c->walk_num <= 0 || c->elec_num <= 0)
      return fail(contexts[i], QMCKL_NOT_PROVIDED, fn, "electron coordinates of the context are not set");
    if (d_coord[i] == nullptr) return fail(contexts[i], QMCKL_INVALID_ARG_4, fn, "null coordinate pointer");
    if (c->profiling) return fail(contexts[i], QMCKL_FAILURE, fn, "kernel profiling must be off while recording");
  }
  Batch* b = new Batch();
  b->origin = static_cast<hipStream_t>(stream);
  b->contexts.assign(contexts, contexts + n);
  b->nbranch = (int)std::min<int64_t>(n, kBranches);
  hipError_t e = hipSuccess;
  if (b->origin == nullptr) {  // a graph cannot be captured on the null stream
    e = hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking);
    b->own_origin = e == hipSuccess;
  }
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
  for (int i = 0; i < b->nbranch && e == hipSuccess; ++i) {
    e = hipStreamCreateWithFlags(&b->branch[i], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming);
  }
  if (e != hipSuccess) {
    destroy_batch(b);
    return hip_fail(contexts[0], e, fn);
  }
  auto d_or_null = [](double* const* a, int64_t i) { return a ? a[i] : nullptr; };
  // eager pass: sizes every buffer and uploads every table outside the capture
  qmckl_exit_code rc = QMCKL_SUCCESS;
  for (int64_t i = 0; i < n && rc == QMCKL_SUCCESS; ++i) {
    rc = qmckl_amd_set_stream(contexts[i], b->branch[i % b->nbranch]);
    if (rc == QMCKL_SUCCESS) rc = step(contexts[i], d_coord[i], d_or_null(d_mo_vgl, i), d_or_null(d_value, i),
                                       d_or_null(d_gl, i));
  }
  for (int i = 0; i < b->nbranch && e == hipSuccess; ++i) e = hipStreamSynchronize(b->branch[i]);
  if (rc != QMCKL_SUCCESS || e != hipSuccess) {
    destroy_batch(b);
    return rc != QMCKL_SUCCESS ? rc : hip_fail(contexts[0], e, fn);
  }
  // capture: fork the branches from the origin stream, record every step, join back
  e = hipStreamBeginCapture(b->origin, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) e = hipEventRecord(b->fork, b->origin);
  for (int i = 0; i < b->nbranch && e == hipSuccess; ++i) e = hipStreamWaitEvent(b->branch[i], b->fork, 0);
  for (int64_t i = 0; i < n && e == hipSuccess && rc == QMCKL_SUCCESS; ++i)
    rc = step(contexts[i], d_coord[i], d_or_null(d_mo_vgl, i), d_or_null(d_value, i), d_or_null(d_gl, i));
  for (int i = 0; i < b->nbranch && e == hipSuccess; ++i) {
    e = hipEventRecord(b->join[i], b->branch[i]);
    if (e == hipSuccess) e = hipStreamWaitEvent(b->origin, b->join[i], 0);
  }
  hipGraph_t graph = nullptr;
  const hipError_t ee = hipStreamEndCapture(b->origin, &graph);
  b->graph = graph;
  if (e == hipSuccess) e = ee;
  if (e == hipSuccess && rc == QMCKL_SUCCESS) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
  if (rc != QMCKL_SUCCESS || e != hipSuccess) {
    destroy_batch(b);
    return rc != QMCKL_SUCCESS ? rc : hip_fail(contexts[0], e, fn);
  }
  for (int64_t i = 0; i < n; ++i) b->recorded.push_back(Snapshot::of(ctx_of(contexts[i])));
  std::lock_guard<std::mutex> g(g_batch_mutex);
  const int64_t h = g_next_batch++;
  g_batches[h] = b;
  *batch = h;
  return QMCKL_SUCCESS;
}

qmckl_exit_code qmckl_amd_batch_replay(const int64_t batch) {
  Batch* b = nullptr;
  {
    std::lock_guard<std::mutex> g(g_batch_mutex);
    auto it = g_batches.find(batch);
    if (it == g_batches.end()) return QMCKL_INVALID_ARG_1;
    b = it->second;
  }
  // The graph reads and writes the contexts' device buffers with the sizes and parameters
  // of the recording: they must all still exist, unchanged.
  for (size_t i = 0; i < b->contexts.size(); ++i) {
    Context* c = ctx_of(b->contexts[i]);
    if (!c) return QMCKL_INVALID_CONTEXT;
    std::lock_guard<std::recursive_mutex> g(c->mutex);
    if (!(Snapshot::of(c) == b->recorded[i]))
      return fail(b->contexts[i], QMCKL_FAILURE, "qmckl_amd_batch_replay",
                  "the context changed since the batch was recorded (sizes, parameters or device "
                  "buffers): record the batch again");
  }
  return hipGraphLaunch(b->exec, b->origin) == hipSuccess ? QMCKL_SUCCESS : QMCKL_FAILURE;
}

qmckl_exit_code qmckl_amd_batch_destroy(const int64_t batch) {
  Batch* b = nullptr;
  {
    std::lock_guard<std::mutex> g(g_batch_mutex);
    auto it = g_batches.find(batch);
    if (it == g_batches.end()) return QMCKL_INVALID_ARG_1;
    b = it->second;
    g_batches.erase(it);
  }
  (void)hipStreamSynchronize(b->origin);
  destroy_batch(b);
  return QMCKL_SUCCESS;
}

qmckl_exit_code qmckl_amd_batch_synchronize(const int64_t batch) {
  Batch* b = nullptr;
  {
    std::lock_guard<std::mutex> g(g_batch_mutex);
    auto it = g_batches.find(batch);
    if (it == g_batches.end()) return QMCKL_INVALID_ARG_1;
    b = it->second;
  }
  return hipStreamSynchronize(b->origin) == hipSuccess ? QMCKL_SUCCESS : QMCKL_FAILURE;
}

}  // extern "C"
