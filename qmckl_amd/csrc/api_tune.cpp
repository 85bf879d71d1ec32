// Tile-shape autotuner of the fused AO->MO product (no counterpart in the reference).
//
// The product's best shape -- column tiles of 16 MOs per pass over the AOs, 4 or 8 waves per
// workgroup -- depends on the system: more tiles per pass evaluate the AOs fewer times but
// hold more accumulators per wave (fewer waves per SIMD), 8 waves share the AO evaluation of
// a tile between more waves but leave some without a column tile.  The tuner groups the
// contexts into size buckets, times every built shape on the first context of a bucket with
// HIP events on that context's stream (its current points, one warm-up launch then two timed
// ones, the faster kept), and sets the fastest shape on every context of the bucket.  The
// choice is kept per process and device: a bucket is timed once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "internal.hpp"

using namespace qmckl_amd;

namespace {

// bucket: device, MO precision, column tiles, 32-AO chunks and point tiles rounded up to
// powers of two, shell type
typedef std::tuple<int, int, int64_t, int64_t, int64_t, int> Bucket;

std::mutex g_tune_mutex;
std::map<Bucket, std::pair<int, int>> g_tuned;

int64_t pow2_ceil(int64_t x) {
  int64_t p = 1;
  while (p < x) p <<= 1;
  return p;
}

// the built shapes worth timing for nct column tiles: 1 to 4 passes on 4 or 8 waves (4 only
// for Slater shells), and the widest passes of each
std::vector<std::pair<int, int>> candidates(int64_t nct, int slater) {
  std::vector<std::pair<int, int>> out;
  for (int nw : {4, 8})
    for (int64_t passes = 1; passes <= 4; ++passes) {
      const int G = (int)((nct + passes - 1) / passes);
      if (!ao_mo_variant_ok(0, G, nw, slater)) continue;
      if (std::find(out.begin(), out.end(), std::make_pair(G, nw)) == out.end()) out.push_back({G, nw});
    }
  // the widest passes, for any number of column tiles
  const std::pair<int, int> wide[2] = {{(int)std::min<int64_t>(nct, 8), 4}, {(int)std::min<int64_t>(nct, 16), 8}};
  for (const auto& w : wide)
    if (ao_mo_variant_ok(0, w.first, w.second, slater) && std::find(out.begin(), out.end(), w) == out.end())
      out.push_back(w);
  return out;
}

void set_shape(Context* c, int G, int nw) {
  if (c->mo_tile_g == G && c->mo_tile_nw == nw) return;
  c->mo_tile_g = G;
  c->mo_tile_nw = nw;
  touch_const(c);  // the product's operand layout depends on the shape
}

// Milliseconds of c's product with shape (G, nw) into out: the faster of two timed launches
// after a warm-up (which also rebuilds the operand layout).
qmckl_exit_code time_shape(qmckl_context h, Context* c, int G, int nw, double* out, const char* fn, float& ms) {
  set_shape(c, G, nw);
  AoMoArgs a;
  int mode = 0;
  qmckl_exit_code rc = mo_product_args(h, c, out, fn, a, mode);
  if (rc != QMCKL_SUCCESS) return rc;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = launch_ao_mo(a, mode, c->stream);
  ms = 1e30f;
  for (int r = 0; r < 2 && e == hipSuccess; ++r) {
    e = hipEventRecord(e0, c->stream);
    if (e == hipSuccess) e = launch_ao_mo(a, mode, c->stream);
    if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (e == hipSuccess) ms = std::min(ms, t);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  return e == hipSuccess ? QMCKL_SUCCESS : hip_fail(h, e, fn);
}

}  // namespace

extern "C" {

qmckl_exit_code qmckl_amd_autotune(const qmckl_context* contexts, const int64_t n, int32_t* const shapes) {
  const char* fn = "qmckl_amd_autotune";
  if (contexts == nullptr || n <= 0) return QMCKL_INVALID_ARG_1;
  std::vector<Bucket> key(n);
  for (int64_t i = 0; i < n; ++i) {
    Context* c = ctx_of(contexts[i]);
    if (!c) return QMCKL_INVALID_CONTEXT;
    std::lock_guard<std::recursive_mutex> g(c->mutex);
    if (!c->mo_provided) return fail(contexts[i], QMCKL_NOT_PROVIDED, fn, "MO basis not set");
    if (c->point_num <= 0) return fail(contexts[i], QMCKL_NOT_PROVIDED, fn, "no points set");
    qmckl_exit_code rc = device_ready(c, contexts[i], fn);
    if (rc != QMCKL_SUCCESS) return rc;
    const int64_t nct = (c->mo_num + 15) / 16, nchunk = (c->ao_num + AO_MO_KC - 1) / AO_MO_KC;
    key[i] = Bucket(c->device, mo_mode(c), nct, pow2_ceil(nchunk), pow2_ceil((c->point_num + 15) / 16),
                    c->tables.slater ? 1 : 0);
  }
  for (int64_t i = 0; i < n; ++i) {
    Context* c = ctx_of(contexts[i]);
    std::lock_guard<std::recursive_mutex> g(c->mutex);
    std::pair<int, int> best(0, 0);
    bool known = false;
    {
      std::lock_guard<std::mutex> t(g_tune_mutex);
      auto it = g_tuned.find(key[i]);
      if (it != g_tuned.end()) {
        best = it->second;
        known = true;
      }
    }
    if (!known) {
      const int64_t nct = std::get<2>(key[i]);
      const size_t bytes = sizeof(double) * (size_t)c->point_num * 5 * c->mo_num;
      hipError_t e = c->d_mo_vgl.reserve(bytes);
      if (e != hipSuccess) return hip_fail(contexts[i], e, fn);
      c->mo_vgl_date = 0;  // the timing runs overwrite the context's MO VGL
      float best_ms = 1e30f;
      for (const auto& cand : candidates(nct, c->tables.slater ? 1 : 0)) {
        float ms = 0.f;
        qmckl_exit_code rc = time_shape(contexts[i], c, cand.first, cand.second, c->d_mo_vgl.as<double>(), fn, ms);
        if (rc != QMCKL_SUCCESS) return rc;
        if (ms < best_ms) {
          best_ms = ms;
          best = cand;
        }
      }
      std::lock_guard<std::mutex> t(g_tune_mutex);
      g_tuned[key[i]] = best;
    }
    set_shape(c, best.first, best.second);
    if (shapes != nullptr) {
      int32_t G = 0, nw = 0, passes = 0;
      qmckl_exit_code rc = qmckl_amd_get_mo_tile(contexts[i], &G, &nw, &passes);
      if (rc != QMCKL_SUCCESS) return rc;
      shapes[3 * i] = G;
      shapes[3 * i + 1] = nw;
      shapes[3 * i + 2] = passes;
    }
  }
  return QMCKL_SUCCESS;
}

}  // extern "C"
