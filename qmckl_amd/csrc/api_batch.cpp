// Batched evaluation of several contexts (molecules of different sizes) replayed from one
// HIP graph.  No counterpart in the reference: it composes the device step of each context
// (coordinates in, MO VGL, Jastrow value and GL out) into one graph so that a batch of
// molecules costs one launch.
//
// Recording runs every context's step once eagerly through the device entry points (which
// sizes and allocates all of its device buffers and uploads its constant tables), then
// builds one grouped launch per kernel over all the molecules, each workgroup serving the
// molecule whose block range holds it:
//   1. one copy kernel: every molecule's coordinates into its context;
//   2. on two branches, the fused AO->MO products, one launch per product shape (precision,
//      column tiles per pass, waves per workgroup) dealt to the branches by cost, one
//      workgroup per (molecule, point tile, column group) item, the costliest molecules'
//      items first; then the MO cusp correction of the contexts that have one;
//   3. on a third branch, the Jastrow kernels, one launch per kernel and cord: the e-e /
//      e-n pair kernel and the pass-based e-e-n kernel of the molecules too large for the
//      resident e-e-n kernel; on a fourth, the resident e-e-n kernel over all the molecules
//      it serves (whole walkers per workgroup, workgroups shared in proportion to the
//      molecules' work); then, on the third, the totals and one copy kernel of every
//      molecule's value and GL into the caller's buffers;
// and captures those launches.  All streams keep the default priority.  After recording,
// every context's stream is the batch's stream, so that a context's own calls are ordered
// after the replays.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "internal.hpp"

using namespace qmckl_amd;

namespace {

constexpr int kBranches = 4;  // AO->MO products (two); Jastrow pair and pass-based e-e-n; resident e-e-n

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
    s.mo_precision = mo_mode(c);
    for (DevBuf* b : device_buffers(c)) s.buffers.emplace_back(b->ptr, b->bytes);
    return s;
  }
  bool operator==(const Snapshot& o) const {
    return walk_num == o.walk_num && elec_num == o.elec_num && mo_num == o.mo_num &&
           const_date == o.const_date && mo_precision == o.mo_precision && buffers == o.buffers;
  }
};

// Copies of many device arrays in one launch: block b copies chunk b - starts[s] of segment
// s, the one whose block range holds b.
struct Segment {
  const double* src;
  double* dst;
  int64_t n;
};
constexpr int kCopyChunk = 4096;  // doubles per block

__global__ void __launch_bounds__(256) k_copy_segments(const Segment* __restrict__ seg,
                                                       const int* __restrict__ starts, int nseg) {
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (starts[mid] <= (int)blockIdx.x)
      lo = mid;
    else
      hi = mid - 1;
  }
  const Segment sg = seg[lo];
  const int64_t b0 = (int64_t)((int)blockIdx.x - starts[lo]) * kCopyChunk;
  const int64_t b1 = b0 + kCopyChunk < sg.n ? b0 + kCopyChunk : sg.n;
  for (int64_t t = b0 + threadIdx.x; t < b1; t += 256) sg.dst[t] = sg.src[t];
}

struct CopyPlan {
  Segment* d_seg = nullptr;
  int* d_starts = nullptr;
  int nseg = 0, nblocks = 0;

  hipError_t build(const std::vector<Segment>& segs) {
    std::vector<Segment> v;
    for (const Segment& s : segs)
      if (s.n > 0) v.push_back(s);
    std::vector<int> starts(v.size() + 1, 0);
    int64_t acc = 0;
    for (size_t i = 0; i < v.size(); ++i) {
      starts[i] = (int)acc;
      acc += (v[i].n + kCopyChunk - 1) / kCopyChunk;
    }
    if (acc > INT32_MAX) return hipErrorInvalidValue;
    starts[v.size()] = (int)acc;
    nseg = (int)v.size();
    nblocks = (int)acc;
    if (nseg == 0) return hipSuccess;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&d_seg), sizeof(Segment) * v.size());
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&d_starts), sizeof(int) * starts.size());
    if (e == hipSuccess) e = hipMemcpy(d_seg, v.data(), sizeof(Segment) * v.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_starts, starts.data(), sizeof(int) * starts.size(), hipMemcpyHostToDevice);
    return e;
  }
  hipError_t launch(hipStream_t s) const {
    if (nblocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_copy_segments, dim3((unsigned)nblocks), dim3(256), 0, s, d_seg, d_starts, nseg);
    return hipGetLastError();
  }
  void free() {
    if (d_seg) (void)hipFree(d_seg);
    if (d_starts) (void)hipFree(d_starts);
    d_seg = nullptr;
    d_starts = nullptr;
    nseg = nblocks = 0;
  }
};

struct Batch {
  hipStream_t origin = nullptr;
  bool own_origin = false;  // created here when the caller passes no stream
  hipStream_t branch[kBranches] = {};
  hipEvent_t fork = nullptr, join[kBranches] = {}, res_done = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<qmckl_context> contexts;
  std::vector<Snapshot> recorded;  // per context, taken after the capture
  CopyPlan copy_in, copy_out;
  MoBatchPlan mo;
  JasBatchPlan jas;
  std::vector<std::pair<qmckl_context, double*>> cusp;  // MO outputs to correct, per context
};

std::mutex g_batch_mutex;
std::map<int64_t, Batch*> g_batches;
int64_t g_next_batch = 1;

void destroy_batch(Batch* b) {
  if (!b) return;
  // the contexts go back to their own streams before the batch's streams are destroyed
  for (qmckl_context ctx : b->contexts)
    if (ctx_of(ctx)) (void)qmckl_amd_set_stream(ctx, nullptr);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->copy_in.free();
  b->copy_out.free();
  mo_batch_free(b->mo);
  jastrow_batch_free(b->jas);
  for (int i = 0; i < kBranches; ++i) {
    if (b->join[i]) (void)hipEventDestroy(b->join[i]);
    if (b->branch[i]) (void)hipStreamDestroy(b->branch[i]);
  }
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->res_done) (void)hipEventDestroy(b->res_done);
  if (b->own_origin && b->origin) (void)hipStreamDestroy(b->origin);
  delete b;
}

// One device step of a context, as the four device entry points (the eager pass).
qmckl_exit_code step(qmckl_context ctx, const double* d_coord, double* d_mo, double* d_value, double* d_gl) {
  Context* c = ctx_of(ctx);
  const int64_t W = c->walk_num, Ne = c->elec_num, P = W * Ne;
  qmckl_exit_code rc = qmckl_amd_set_electron_coord_device(ctx, 'T', W, d_coord, 3 * P);
  if (rc == QMCKL_SUCCESS && d_mo != nullptr)
    rc = qmckl_amd_get_mo_basis_mo_vgl_device(ctx, d_mo, P * 5 * c->mo_num);
  // the full Jastrow evaluation, as the recorded graph runs it (with a value alone asked for,
  // the value getter could run the values-only one, which leaves the GL buffers the graph
  // writes unallocated)
  if (rc == QMCKL_SUCCESS && (d_value != nullptr || d_gl != nullptr)) {
    std::lock_guard<std::recursive_mutex> g(c->mutex);
    rc = provide_jastrow(ctx, c, "qmckl_amd_batch_record");
  }
  if (rc == QMCKL_SUCCESS && d_value != nullptr) rc = qmckl_amd_get_jastrow_champ_value_device(ctx, d_value, W);
  if (rc == QMCKL_SUCCESS && d_gl != nullptr) rc = qmckl_amd_get_jastrow_champ_gl_device(ctx, d_gl, W * 4 * Ne);
  return rc;
}

// The grouped launches of one step, in capture order.
hipError_t enqueue_step(Batch* b) {
  hipError_t e = b->copy_in.launch(b->origin);
  if (e == hipSuccess) e = hipEventRecord(b->fork, b->origin);
  for (int i = 0; i < kBranches && e == hipSuccess; ++i) e = hipStreamWaitEvent(b->branch[i], b->fork, 0);
  if (e == hipSuccess) e = mo_batch_launch(b->mo, b->branch, 2);
  if (!b->cusp.empty() && e == hipSuccess) {  // both product branches done before the corrections
    e = hipEventRecord(b->join[1], b->branch[1]);
    if (e == hipSuccess) e = hipStreamWaitEvent(b->branch[0], b->join[1], 0);
  }
  for (const auto& cc : b->cusp) {
    if (e != hipSuccess) break;
    Context* c = ctx_of(cc.first);
    c->stream = b->branch[0];
    if (apply_mo_cusp(c, cc.first, cc.second, 0, "qmckl_amd_batch_record") != QMCKL_SUCCESS) e = hipErrorUnknown;
    c->stream = b->origin;
  }
  if (e == hipSuccess) e = jastrow_batch_launch(b->jas, b->branch[2], b->branch[3], b->res_done);
  if (e == hipSuccess) e = b->copy_out.launch(b->branch[2]);
  for (int i = 0; i < kBranches && e == hipSuccess; ++i) {
    e = hipEventRecord(b->join[i], b->branch[i]);
    if (e == hipSuccess) e = hipStreamWaitEvent(b->origin, b->join[i], 0);
  }
  return e;
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
    for (int64_t k = 0; k < i; ++k)
      if (contexts[k] == contexts[i]) return fail(contexts[i], QMCKL_INVALID_ARG_1, fn, "a context appears twice");
  }
  Batch* b = new Batch();
  b->origin = static_cast<hipStream_t>(stream);
  b->contexts.assign(contexts, contexts + n);
  hipError_t e = hipSuccess;
  if (b->origin == nullptr) {  // a graph cannot be captured on the null stream
    e = hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking);
    b->own_origin = e == hipSuccess;
  }
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->res_done, hipEventDisableTiming);
  for (int i = 0; i < kBranches && e == hipSuccess; ++i) {
    e = hipStreamCreateWithFlags(&b->branch[i], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming);
  }
  if (e != hipSuccess) {
    destroy_batch(b);
    return hip_fail(contexts[0], e, fn);
  }
  auto d_or_null = [](double* const* a, int64_t i) { return a ? a[i] : nullptr; };
  // eager pass on the batch's stream: sizes every buffer and uploads every table
  qmckl_exit_code rc = QMCKL_SUCCESS;
  for (int64_t i = 0; i < n && rc == QMCKL_SUCCESS; ++i) {
    rc = qmckl_amd_set_stream(contexts[i], b->origin);
    if (rc == QMCKL_SUCCESS) rc = step(contexts[i], d_coord[i], d_or_null(d_mo_vgl, i), d_or_null(d_value, i),
                                       d_or_null(d_gl, i));
  }
  if (rc == QMCKL_SUCCESS) e = hipStreamSynchronize(b->origin);
  // the grouped launches
  std::vector<Segment> in, out;
  std::vector<std::pair<AoMoArgs, int>> mo;
  std::vector<std::pair<JasDevView, JasOut>> jas;
  for (int64_t i = 0; i < n && rc == QMCKL_SUCCESS && e == hipSuccess; ++i) {
    Context* c = ctx_of(contexts[i]);
    std::lock_guard<std::recursive_mutex> g(c->mutex);
    const int64_t W = c->walk_num, Ne = c->elec_num, P = W * Ne;
    in.push_back({d_coord[i], c->d_coord.as<double>(), 3 * P});
    if (d_mo_vgl != nullptr && d_mo_vgl[i] != nullptr) {
      AoMoArgs a;
      int mode = 0;
      rc = mo_product_args(contexts[i], c, d_mo_vgl[i], fn, a, mode);
      if (rc != QMCKL_SUCCESS) break;
      mo.push_back({a, mode});
      if (!c->r_cusp.empty()) b->cusp.push_back({contexts[i], d_mo_vgl[i]});
    }
    const bool want_v = d_value != nullptr && d_value[i] != nullptr, want_g = d_gl != nullptr && d_gl[i] != nullptr;
    if (want_v || want_g) {
      JasDevView v;
      rc = jastrow_view_ready(c, contexts[i], fn, v);
      if (rc != QMCKL_SUCCESS) break;
      jas.push_back({v, jastrow_out(c)});
      if (want_v) out.push_back({c->d_jvalue.as<double>(), d_value[i], W});
      if (want_g) out.push_back({c->d_jgl.as<double>(), d_gl[i], 4 * P});
    }
  }
  if (rc == QMCKL_SUCCESS && e == hipSuccess) e = b->copy_in.build(in);
  if (rc == QMCKL_SUCCESS && e == hipSuccess) e = b->copy_out.build(out);
  if (rc == QMCKL_SUCCESS && e == hipSuccess) e = mo_batch_build(mo, b->mo);
  if (rc == QMCKL_SUCCESS && e == hipSuccess) e = jastrow_batch_build(jas, b->jas);
  if (rc != QMCKL_SUCCESS || e != hipSuccess) {
    destroy_batch(b);
    return rc != QMCKL_SUCCESS ? rc : hip_fail(contexts[0], e, fn);
  }
  // capture the grouped launches
  e = hipStreamBeginCapture(b->origin, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    e = enqueue_step(b);
    hipGraph_t graph = nullptr;
    const hipError_t ee = hipStreamEndCapture(b->origin, &graph);
    b->graph = graph;
    if (e == hipSuccess) e = ee;
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    destroy_batch(b);
    return hip_fail(contexts[0], e, fn);
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
