// Contracted-Gaussian AO basis and MO basis: setters, getters, basis finalisation and
// the providers of AO VGL and MO VGL.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "internal.hpp"

using namespace qmckl_amd;

namespace {

// ---- AO setter scaffolding, org/qmckl_ao.org:386-420 (mask bits and ALREADY_SET rule)
qmckl_exit_code ao_pre(Context* c, qmckl_context ctx, uint32_t mask) {
  if (mask != 0 && !(c->ao_uninit & mask)) return fail(ctx, QMCKL_ALREADY_SET, "qmckl_set_ao_*", nullptr);
  return QMCKL_SUCCESS;
}

// Builds the device-oriented tables once the 14 mandatory items are set;
// org/qmckl_ao.org:2965-3183 (qmckl_finalize_basis).
qmckl_exit_code finalize_basis(Context* c, qmckl_context ctx) {
  AoTables& t = c->tables;
  t = AoTables();
  const int64_t Nn = c->nucl_num;
  if (Nn <= 0) return fail(ctx, QMCKL_NOT_PROVIDED, "qmckl_finalize_basis", "nucleus data is not provided");
  if (c->ao_type != 'G' && c->ao_type != 'S')
    return fail(ctx, QMCKL_FAILURE, "qmckl_finalize_basis", "basis type must be 'G' or 'S'");
  if (c->r_power.empty()) c->r_power.assign(c->shell_num, 0);
  t.slater = c->ao_type == 'S';
  t.nucl_num = Nn;
  t.shell_num = c->shell_num;
  t.prim_num = c->prim_num;
  t.ao_num = c->ao_num;
  t.nucl_range.assign(Nn, 0.0);
  t.nucl_max_l.assign(Nn, 0);
  int64_t ao = 0;
  for (int64_t a = 0; a < Nn; ++a) {
    const int64_t s0 = c->nucleus_index[a], s1 = s0 + c->nucleus_shell_num[a];
    if (s0 < 0 || s1 > c->shell_num)
      return fail(ctx, QMCKL_FAILURE, "qmckl_finalize_basis", "nucleus_index/nucleus_shell_num out of range");
    for (int64_t s = s0; s < s1; ++s) {
      const int l = c->shell_ang_mom[s];
      if (l < 0 || l > 10) return fail(ctx, QMCKL_FAILURE, "qmckl_finalize_basis", "angular momentum outside [0,10]");
      const int64_t p0 = c->shell_prim_index[s], np = c->shell_prim_num[s];
      if (p0 < 0 || np < 0 || p0 + np > c->prim_num)
        return fail(ctx, QMCKL_FAILURE, "qmckl_finalize_basis", "shell primitives out of range");
      t.shell_nucl.push_back((int32_t)a);
      t.shell_l.push_back(l);
      t.shell_prim_start.push_back((int32_t)t.prim_expo.size());
      t.shell_prim_count.push_back((int32_t)np);
      t.shell_ao_start.push_back((int32_t)ao);
      t.shell_rpow.push_back(c->r_power[s]);
      for (int64_t p = p0; p < p0 + np; ++p) {
        t.prim_expo.push_back(c->exponent[p]);
        t.prim_coef.push_back(c->coefficient[p] * c->prim_factor[p] * c->shell_factor[s]);
        const double range = (double)(c->r_power[s] + 1) / c->exponent[p];
        t.nucl_range[a] = std::max(t.nucl_range[a], range);
      }
      t.nucl_max_l[a] = std::max(t.nucl_max_l[a], l);
      t.lmax = std::max(t.lmax, (int32_t)l);
      // Cartesian polynomials in the order of org/qmckl_ao.org:6270-6300:
      // a = l..0, b = l-a..0, c = l-a-b.
      for (int pa = l; pa >= 0; --pa)
        for (int pb = l - pa; pb >= 0; --pb) {
          const int pc = l - pa - pb;
          t.ao_shell.push_back((int32_t)(t.shell_l.size() - 1));
          t.ao_pow.push_back(pa | (pb << 8) | (pc << 16));
          ++ao;
        }
    }
  }
  if (ao != c->ao_num)
    return fail(ctx, QMCKL_FAILURE, "qmckl_finalize_basis", "ao_num does not match the shells' angular momenta");
  if ((int64_t)t.shell_l.size() != c->shell_num)
    return fail(ctx, QMCKL_FAILURE, "qmckl_finalize_basis", "the nuclei's shell ranges do not cover shell_num shells");
  for (int64_t a = 1; a < Nn; ++a)
    if (c->nucleus_index[a] != c->nucleus_index[a - 1] + c->nucleus_shell_num[a - 1])
      return fail(ctx, QMCKL_FAILURE, "qmckl_finalize_basis", "shells must be stored nucleus by nucleus");
  t.shell_num = (int64_t)t.shell_l.size();
  t.ao_factor = c->ao_factor;
  t.chunk_shells_max = 1;
  for (int64_t k0 = 0; k0 < ao; k0 += 32) {
    const int64_t k1 = std::min<int64_t>(k0 + 32, ao) - 1;
    t.chunk_shells_max = std::max<int32_t>(t.chunk_shells_max, t.ao_shell[k1] - t.ao_shell[k0] + 1);
  }
  touch_const(c);
  return QMCKL_SUCCESS;
}

qmckl_exit_code ao_post(Context* c, qmckl_context ctx, uint32_t mask) {
  c->ao_uninit &= ~mask;
  c->ao_provided = (c->ao_uninit == 0);
  touch_const(c);
  if (c->ao_provided) return finalize_basis(c, ctx);
  return QMCKL_SUCCESS;
}

template <class T>
qmckl_exit_code set_array(Context* c, qmckl_context ctx, uint32_t mask, const T* src, int64_t n,
                          int64_t size_max, std::vector<T>& dst, const char* fn, const char* what) {
  qmckl_exit_code rc = ao_pre(c, ctx, mask);
  if (rc != QMCKL_SUCCESS) return rc;
  if (n <= 0) return fail(ctx, QMCKL_FAILURE, fn, what);
  if (src == nullptr) return fail(ctx, QMCKL_INVALID_ARG_2, fn, "null pointer");
  if (size_max < n) return fail(ctx, QMCKL_INVALID_ARG_3, fn, "input array too small");
  dst.assign(src, src + n);
  return ao_post(c, ctx, mask);
}

template <class T>
qmckl_exit_code get_array(Context* c, qmckl_context ctx, uint32_t mask, T* dst, int64_t n, int64_t size_max,
                          const std::vector<T>& src, const char* fn) {
  if (dst == nullptr) return fail(ctx, QMCKL_INVALID_ARG_2, fn, "NULL pointer");
  if (c->ao_uninit & mask) return fail(ctx, QMCKL_NOT_PROVIDED, fn, nullptr);
  if (size_max < n) return fail(ctx, QMCKL_INVALID_ARG_3, fn, "input array too small");
  std::copy(src.begin(), src.begin() + n, dst);
  return QMCKL_SUCCESS;
}

}  // namespace

namespace qmckl_amd {

// Host copy of split3 (ao_device.hpp): x ~ hi + mid + lo, three bfloat16 terms, each the
// round-to-nearest-even of the float of the remainder, so host and device splits agree.
static uint16_t bf16_rne_host(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
static double bf16_value_host(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return (double)f;
}
static void split3_host(double x, uint16_t h[3]) {
  h[0] = bf16_rne_host((float)x);
  const double r1 = x - bf16_value_host(h[0]);
  h[1] = bf16_rne_host((float)r1);
  const double r2 = r1 - bf16_value_host(h[1]);
  h[2] = bf16_rne_host((float)r2);
}

// Order of the AOs in the fused AO->MO kernel's 32-AO chunks (perm[position] = AO index).
// The product sums over every AO, so any order gives the MO VGL (up to the order of the FP64
// sums across chunks).  The kernel skips the chunks whose shells are all screened at every
// point of a workgroup's tile, and a tile holds 16 points near one another (the same
// electron of 16 walkers): so shells that vanish at the same points go together.  A shell
// vanishes beyond r^2 = cutoff * min(range of its nucleus, 1 / smallest exponent); the
// shells are taken compact ones first (that radius at most the median), then diffuse ones,
// each group in the Morton order of their nuclei (10 bits per axis over the nuclei's box),
// whole shells, their AOs in the reference order.
static std::vector<int64_t> fused_ao_order(const Context* c) {
  const AoTables& t = c->tables;
  const int64_t S = t.shell_num, Nn = t.nucl_num;
  const double cutoff = -log(ldexp(1.0, 1 - c->precision));
  std::vector<double> r2(S);
  for (int64_t s = 0; s < S; ++s) {
    double amin = INFINITY;
    for (int q = t.shell_prim_start[s]; q < t.shell_prim_start[s] + t.shell_prim_count[s]; ++q)
      amin = std::min(amin, t.prim_expo[q]);
    r2[s] = cutoff * std::min(t.nucl_range[t.shell_nucl[s]], 1.0 / amin);
  }
  std::vector<double> sorted_r2(r2);
  std::nth_element(sorted_r2.begin(), sorted_r2.begin() + S / 2, sorted_r2.end());
  const double median = S ? sorted_r2[S / 2] : 0.0;
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t a = 0; a < Nn; ++a)
    for (int d3 = 0; d3 < 3; ++d3) {
      lo[d3] = std::min(lo[d3], c->nucl_coord[d3 * Nn + a]);
      hi[d3] = std::max(hi[d3], c->nucl_coord[d3 * Nn + a]);
    }
  std::vector<uint64_t> morton(Nn, 0);
  for (int64_t a = 0; a < Nn; ++a)
    for (int d3 = 0; d3 < 3; ++d3) {
      const double span = hi[d3] - lo[d3];
      const uint64_t qd = span > 0 ? (uint64_t)((c->nucl_coord[d3 * Nn + a] - lo[d3]) / span * 1023.0) : 0;
      for (int b = 0; b < 10; ++b) morton[a] |= ((qd >> b) & 1u) << (3 * b + d3);
    }
  std::vector<int64_t> order(S);
  for (int64_t s = 0; s < S; ++s) order[s] = s;
  std::stable_sort(order.begin(), order.end(), [&](int64_t x, int64_t y) {
    const bool cx = r2[x] <= median, cy = r2[y] <= median;
    if (cx != cy) return cx;
    return morton[t.shell_nucl[x]] < morton[t.shell_nucl[y]];
  });
  std::vector<int64_t> perm;
  perm.reserve(t.ao_num);
  for (int64_t s : order) {
    const int64_t n = (int64_t)(t.shell_l[s] + 1) * (t.shell_l[s] + 2) / 2;
    for (int64_t k = 0; k < n; ++k) perm.push_back(t.shell_ao_start[s] + k);
  }
  return perm;
}

// Shape of c's fused product for nct column tiles: the requested one (capped at nct tiles,
// the waves adjusted when the cap leaves the requested pair unbuilt), else the default; with
// Slater shells at most 8 tiles on 4 waves (the shapes built for them).
static void mo_tile_shape(const Context* c, int64_t nct, int& G, int& nw) {
  ao_mo_default_shape(nct, G, nw);
  if (c->mo_tile_g > 0) {
    G = (int)std::min<int64_t>(c->mo_tile_g, std::max<int64_t>(nct, 1));
    nw = ao_mo_variant_ok(0, G, c->mo_tile_nw) ? c->mo_tile_nw : ao_mo_waves(G);
  }
  if (c->tables.slater && nw != 4) {  // Slater shells: the 4-wave shapes only
    G = std::min(G, 8);
    nw = 4;
  }
}

// Uploads the AO tables and nuclear coordinates (once per context date change).
static qmckl_exit_code upload_ao(Context* c, qmckl_context handle, const char* fn) {
  qmckl_exit_code rc = device_ready(c, handle, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  if (c->dc.date >= c->const_date) return QMCKL_SUCCESS;
  const AoTables& t = c->tables;
  DevConst& d = c->dc;
  hipError_t e = hipSuccess;
  auto up = [&](DevBuf& b, const void* src, size_t bytes) {
    if (e != hipSuccess) return;
    e = b.reserve(bytes);
    if (e == hipSuccess && bytes) e = hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, c->stream);
  };
  up(d.shell_nucl, t.shell_nucl.data(), 4 * t.shell_nucl.size());
  up(d.shell_l, t.shell_l.data(), 4 * t.shell_l.size());
  up(d.shell_prim_start, t.shell_prim_start.data(), 4 * t.shell_prim_start.size());
  up(d.shell_prim_count, t.shell_prim_count.data(), 4 * t.shell_prim_count.size());
  up(d.shell_ao_start, t.shell_ao_start.data(), 4 * t.shell_ao_start.size());
  up(d.shell_rpow, t.shell_rpow.data(), 4 * t.shell_rpow.size());
  up(d.prim_expo, t.prim_expo.data(), 8 * t.prim_expo.size());
  up(d.prim_coef, t.prim_coef.data(), 8 * t.prim_coef.size());
  up(d.ao_shell, t.ao_shell.data(), 4 * t.ao_shell.size());
  up(d.ao_pow, t.ao_pow.data(), 4 * t.ao_pow.size());
  up(d.ao_factor, t.ao_factor.data(), 8 * t.ao_factor.size());
  up(d.nucl_range, t.nucl_range.data(), 8 * t.nucl_range.size());
  up(d.nucl_coord, c->nucl_coord.data(), 8 * c->nucl_coord.size());
  const std::vector<int64_t> perm = fused_ao_order(c);
  {
    // Descriptors of the 32-AO chunks for the fused AO->MO kernel (layout: AoMoArgs), the
    // AOs taken in the order `perm`: each chunk lists the shells of its AOs (in order of first
    // appearance), their nuclei and primitives, with chunk-local indices.
    const int64_t Nao = t.ao_num, nchunk = (Nao + AO_MO_KC - 1) / AO_MO_KC, Nn = t.nucl_num;
    std::vector<double> desc;
    std::vector<int64_t> off(1, 0);
    int dmax = 0, nsmax = 1;
    for (int64_t ch = 0; ch < nchunk; ++ch) {
      std::vector<int> shells, nuclei, shell_local(AO_MO_KC, 0);
      for (int kk = 0; kk < AO_MO_KC; ++kk) {
        const int64_t pos = ch * AO_MO_KC + kk;
        if (pos >= Nao) break;
        const int sh = t.ao_shell[perm[pos]];
        auto it = std::find(shells.begin(), shells.end(), sh);
        shell_local[kk] = (int)(it - shells.begin());
        if (it == shells.end()) shells.push_back(sh);
        if (std::find(nuclei.begin(), nuclei.end(), t.shell_nucl[sh]) == nuclei.end())
          nuclei.push_back(t.shell_nucl[sh]);
      }
      int np = 0;
      for (int sh : shells) np += t.shell_prim_count[sh];
      const size_t at = desc.size();
      desc.insert(desc.end(), {(double)shells.size(), (double)nuclei.size(), (double)np, 0.0});
      for (int a : nuclei)
        desc.insert(desc.end(), {c->nucl_coord[a], c->nucl_coord[Nn + a], c->nucl_coord[2 * Nn + a],
                                 t.nucl_range[a]});
      int p_first = 0;
      for (int sh : shells) {
        const int nl = (int)(std::find(nuclei.begin(), nuclei.end(), t.shell_nucl[sh]) - nuclei.begin());
        // shell record: chunk-local nucleus + 256 x primitives, first primitive + 65536 x
        // (r_power + 128)
        desc.insert(desc.end(), {(double)(nl + 256 * t.shell_prim_count[sh]),
                                 (double)p_first + 65536.0 * (double)(t.shell_rpow[sh] + 128)});
        p_first += t.shell_prim_count[sh];
      }
      for (int sh : shells)
        for (int q = t.shell_prim_start[sh]; q < t.shell_prim_start[sh] + t.shell_prim_count[sh]; ++q)
          desc.insert(desc.end(), {t.prim_expo[q], t.prim_coef[q]});
      for (int kk = 0; kk < AO_MO_KC; ++kk) {
        const int64_t pos = ch * AO_MO_KC + kk;
        if (pos < Nao)
          desc.insert(desc.end(), {(double)shell_local[kk] + 256.0 * (double)t.ao_pow[perm[pos]],
                                   t.ao_factor[perm[pos]]});
        else
          desc.insert(desc.end(), {0.0, 0.0});
      }
      if (desc.size() % 2) desc.push_back(0.0);
      dmax = std::max<int>(dmax, (int)(desc.size() - at));
      nsmax = std::max<int>(nsmax, (int)shells.size());
      off.push_back((int64_t)desc.size());
    }
    d.ao_nchunk = (int)nchunk;
    d.ao_desc_max = dmax;
    d.ao_ns_max = nsmax;
    up(d.ao_desc, desc.data(), 8 * desc.size());
    up(d.ao_desc_off, off.data(), 8 * off.size());
  }
  {
    // Nucleus of each primitive in the caller's order, org/qmckl_ao.org:3000-3015
    // (nucleus_prim_index from the first shell of each nucleus).
    std::vector<int32_t> pn(c->prim_num, 0);
    for (int64_t a = 0; a < c->nucl_num; ++a) {
      const int64_t p0 = c->shell_prim_index[c->nucleus_index[a]];
      const int64_t p1 = (a + 1 < c->nucl_num) ? c->shell_prim_index[c->nucleus_index[a + 1]] : c->prim_num;
      for (int64_t p = p0; p < p1 && p < c->prim_num; ++p) pn[p] = (int32_t)a;
    }
    up(d.prim_nucl_c, pn.data(), 4 * pn.size());
    up(d.prim_expo_c, c->exponent.data(), 8 * c->exponent.size());
  }
  if (e == hipSuccess && c->mo_provided) {
    // coefficient_t (org/qmckl_mo.org:316-375) zero-padded to k_pad = 32*ceil(ao/32) rows
    // and mo_pad = 16*ceil(mo/16) columns, so the product kernels need no guards on it.
    const int64_t Nao = c->ao_num, Nmo = c->mo_num;
    const int64_t pad = ((Nmo + 15) / 16) * 16, kpad = ((Nao + 31) / 32) * 32;
    std::vector<double> ct((size_t)kpad * pad, 0.0);
    for (int64_t k = 0; k < Nao; ++k)
      for (int64_t j = 0; j < Nmo; ++j) ct[k * pad + j] = c->mo_coef[j * Nao + k];
    d.mo_pad = pad;
    d.k_pad = kpad;
    up(d.coef_t, ct.data(), 8 * ct.size());
    // the operand layouts of the fused kernel: per (group of G column tiles, 32-AO chunk),
    // row k scaled by ao_factor[k] (the fused kernel's AO values leave the factor out:
    // sum_k (f_k P_k) C_kj = sum_k P_k (f_k C_kj))
    const int64_t nct = pad / 16, nchunk = kpad / AO_MO_KC;
    int G = 0, nw = 0;
    mo_tile_shape(c, nct, G, nw);
    const int64_t ngroup = (nct + G - 1) / G, NG16 = 16 * G;
    d.ao_mo_g = G;
    d.ao_mo_nw = nw;
    d.ao_mo_ngroup = (int)ngroup;
    std::vector<double> b64((size_t)(ngroup * nchunk) * AO_MO_KC * NG16, 0.0);
    std::vector<float> b32(b64.size(), 0.0f);
    std::vector<uint16_t> bbf((size_t)(ngroup * nchunk) * 3 * NG16 * AO_MO_BF_LD, 0);
    for (int64_t ch = 0; ch < nchunk; ++ch)
      for (int kk = 0; kk < AO_MO_KC; ++kk) {
        // row kk of chunk ch is AO perm[32 ch + kk] (zero past ao_num)
        const int64_t pos = ch * AO_MO_KC + kk;
        const int64_t k = pos < Nao ? perm[pos] : Nao;
        const double fk = pos < Nao ? t.ao_factor[k] : 0.0;
        for (int64_t col = 0; col < pad; ++col) {
          const double x = pos < Nao ? ct[k * pad + col] * fk : 0.0;
          const int64_t g = col / NG16, n = col - g * NG16;
          const size_t blk = (size_t)(g * nchunk + ch);
          b64[(blk * AO_MO_KC + kk) * NG16 + n] = x;
          b32[(blk * AO_MO_KC + kk) * NG16 + n] = (float)x;
          uint16_t h3[3];
          split3_host(x, h3);
          for (int pl = 0; pl < 3; ++pl) bbf[((blk * 3 + pl) * NG16 + n) * AO_MO_BF_LD + kk] = h3[pl];
        }
      }
    up(d.coef_ch64, b64.data(), 8 * b64.size());
    up(d.coef_ch32, b32.data(), 4 * b32.size());
    up(d.coef_chbf, bbf.data(), 2 * bbf.size());
  }
  if (e != hipSuccess) return hip_fail(handle, e, fn);
  d.date = c->const_date;
  return QMCKL_SUCCESS;
}

AoDevView ao_view(Context* c) {
  const DevConst& d = c->dc;
  AoDevView v;
  v.nucl_num = c->nucl_num;
  v.shell_num = c->tables.shell_num;
  v.ao_num = c->ao_num;
  v.shell_nucl = d.shell_nucl.as<int32_t>();
  v.shell_l = d.shell_l.as<int32_t>();
  v.shell_prim_start = d.shell_prim_start.as<int32_t>();
  v.shell_prim_count = d.shell_prim_count.as<int32_t>();
  v.shell_ao_start = d.shell_ao_start.as<int32_t>();
  v.prim_expo = d.prim_expo.as<double>();
  v.prim_coef = d.prim_coef.as<double>();
  v.ao_shell = d.ao_shell.as<int32_t>();
  v.ao_pow = d.ao_pow.as<int32_t>();
  v.ao_factor = d.ao_factor.as<double>();
  v.nucl_range = d.nucl_range.as<double>();
  v.nucl_coord = d.nucl_coord.as<double>();
  // cutoff = -log(epsilon), epsilon = 2^(1-p): org/qmckl_ao.org:118-120, :7811-7815
  v.cutoff = -log(ldexp(1.0, 1 - c->precision));
  v.slater = c->tables.slater ? 1 : 0;
  v.shell_rpow = d.shell_rpow.as<int32_t>();
  return v;
}

qmckl_exit_code ao_ready(Context* c, qmckl_context handle, const char* fn) {
  if (!c->ao_provided) return fail(handle, QMCKL_NOT_PROVIDED, fn, nullptr);
  if (c->point_num <= 0) return fail(handle, QMCKL_NOT_PROVIDED, fn, "no points set");
  qmckl_exit_code rc = upload_ao(c, handle, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  return upload_coords(c, handle, fn);
}

// AO VGL of all points into the context buffer; org/qmckl_ao.org:9022-9226.
qmckl_exit_code provide_ao_vgl(qmckl_context handle, Context* c, const char* fn) {
  if (c->ao_vgl_date >= c->date) return QMCKL_SUCCESS;
  qmckl_exit_code rc = ao_ready(c, handle, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  const int64_t P = c->point_num, Nao = c->ao_num;
  hipError_t e = c->d_ao_vgl.reserve(sizeof(double) * P * 5 * Nao);
  if (e == hipSuccess) {
    ProfScope ps(c, SLOT_AO);
    e = launch_ao_vgl(ao_view(c), c->d_coord.as<double>(), P, 0, P, c->d_ao_vgl.ptr, 0, Nao, 0, 0, c->stream);
  }
  if (e != hipSuccess) return hip_fail(handle, e, fn);
  c->ao_vgl_date = c->date;
  return QMCKL_SUCCESS;
}

void ao_mo_default_shape(int64_t nct, int& G, int& nw) {
  if (nct <= 8) {
    G = (int)std::max<int64_t>(nct, 1);
    nw = 4;
    return;
  }
  const int64_t ngroup = (nct + 15) / 16;
  G = (int)((nct + ngroup - 1) / ngroup);  // >= 5
  nw = 8;
}

qmckl_exit_code mo_product_args(qmckl_context handle, Context* c, double* d_out, const char* fn, AoMoArgs& a,
                                int& mode) {
  if (!c->mo_provided) return fail(handle, QMCKL_NOT_PROVIDED, fn, nullptr);
  qmckl_exit_code rc = ao_ready(c, handle, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  const int64_t P = c->point_num, Nmo = c->mo_num;
  const DevConst& d = c->dc;
  mode = mo_mode(c);
  a.coord = c->d_coord.as<double>();
  a.coord_ld = P;
  a.p0 = 0;
  a.np = P;
  // tiles of the same electron of consecutive walkers (points close to one another) when
  // the points are the walkers' electrons; else consecutive points
  const bool walkers = c->walk_num > 1 && c->walk_num * c->elec_num == P;
  a.tw = walkers ? c->walk_num : 1;
  a.te = walkers ? c->elec_num : P;
  a.desc = d.ao_desc.as<double>();
  a.desc_off = d.ao_desc_off.as<int64_t>();
  a.desc_max = d.ao_desc_max;
  a.ns_max = d.ao_ns_max;
  a.nchunk = d.ao_nchunk;
  a.G = d.ao_mo_g;
  a.nw = d.ao_mo_nw;
  a.ngroup = d.ao_mo_ngroup;
  a.B = mode == QMCKL_AMD_MO_FP64 ? d.coef_ch64.ptr : mode == QMCKL_AMD_MO_FP32 ? d.coef_ch32.ptr : d.coef_chbf.ptr;
  a.b_chunk_bytes = mode == QMCKL_AMD_MO_FP64   ? AO_MO_KC * 16 * a.G * 8
                    : mode == QMCKL_AMD_MO_FP32 ? AO_MO_KC * 16 * a.G * 4
                                                : 3 * 16 * a.G * AO_MO_BF_LD * 2;
  a.C = d_out;
  a.ldc = Nmo;
  a.N = Nmo;
  a.cutoff = ao_view(c).cutoff;
  a.slater = c->tables.slater ? 1 : 0;
  a.vo = 0;
  return QMCKL_SUCCESS;
}

// MO VGL into d_out (or into the context buffer when d_out is null); org/qmckl_mo.org:1749-1878.
// The AO VGL is evaluated inside the product kernel (it is never stored), unless the FP64
// AO VGL of these points is already resident, in which case it is multiplied directly.
qmckl_exit_code provide_mo_vgl(qmckl_context handle, Context* c, double* d_out, const char* fn) {
  if (!c->mo_provided) return fail(handle, QMCKL_NOT_PROVIDED, fn, nullptr);
  if (d_out == nullptr && c->mo_vgl_date >= c->date) return QMCKL_SUCCESS;
  qmckl_exit_code rc = ao_ready(c, handle, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  const int64_t P = c->point_num, Nao = c->ao_num, Nmo = c->mo_num;
  hipError_t e = hipSuccess;
  double* out = d_out;
  if (out == nullptr) {
    e = c->d_mo_vgl.reserve(sizeof(double) * P * 5 * Nmo);
    if (e != hipSuccess) return hip_fail(handle, e, fn);
    out = c->d_mo_vgl.as<double>();
  }
  const int mode = mo_mode(c);
  const int64_t pad = c->dc.mo_pad;
  ProfScope ps(c, SLOT_MO);
  if (mode == QMCKL_AMD_MO_FP64 && c->ao_vgl_date >= c->date) {
    // The FP64 AO VGL of these points is resident: plain GEMM on it.
    e = launch_gemm_f64(c->d_ao_vgl.as<double>(), Nao, c->dc.coef_t.as<double>(), pad, out, Nmo, 5 * P, Nmo,
                        Nao, c->stream);
  } else {
    AoMoArgs a;
    int m = 0;
    rc = mo_product_args(handle, c, out, fn, a, m);
    if (rc != QMCKL_SUCCESS) return rc;
    e = launch_ao_mo(a, m, c->stream);
  }
  if (e != hipSuccess) return hip_fail(handle, e, fn);
  if (!c->r_cusp.empty()) {
    rc = apply_mo_cusp(c, handle, out, 0, fn);
    if (rc != QMCKL_SUCCESS) return rc;
  }
  if (d_out == nullptr) c->mo_vgl_date = c->date;
  return QMCKL_SUCCESS;
}

// MO values [point][mo] into d_out (or into the context buffer when d_out is null);
// org/qmckl_mo.org:988-1100.  The value rows of the MO VGL when it is current (as the
// reference takes them), else the values-only product: AO values without derivatives and a
// fifth of the matrix work, the same sums bit for bit.  With the cusp correction, or the
// FP64 AO VGL resident, the full evaluation gives them.
qmckl_exit_code provide_mo_value(qmckl_context handle, Context* c, double* d_out, const char* fn) {
  if (!c->mo_provided) return fail(handle, QMCKL_NOT_PROVIDED, fn, nullptr);
  if (d_out == nullptr && c->mo_value_date >= c->date) return QMCKL_SUCCESS;
  const int64_t P = c->point_num, Nmo = c->mo_num;
  double* out = d_out;
  hipError_t e = hipSuccess;
  if (out == nullptr) {
    e = c->d_mo_value.reserve(sizeof(double) * std::max<int64_t>(P, 1) * Nmo);
    if (e != hipSuccess) return hip_fail(handle, e, fn);
    out = c->d_mo_value.as<double>();
  }
  const bool via_vgl = c->mo_vgl_date >= c->date || !c->r_cusp.empty() ||
                       (mo_mode(c) == QMCKL_AMD_MO_FP64 && c->ao_vgl_date >= c->date);
  if (via_vgl) {
    qmckl_exit_code rc = provide_mo_vgl(handle, c, nullptr, fn);
    if (rc != QMCKL_SUCCESS) return rc;
    e = hipMemcpy2DAsync(out, sizeof(double) * Nmo, c->d_mo_vgl.ptr, sizeof(double) * 5 * Nmo, sizeof(double) * Nmo,
                         P, hipMemcpyDeviceToDevice, c->stream);
  } else {
    AoMoArgs a;
    int mode = 0;
    qmckl_exit_code rc = mo_product_args(handle, c, out, fn, a, mode);
    if (rc != QMCKL_SUCCESS) return rc;
    a.vo = 1;
    ProfScope ps(c, SLOT_MO);
    e = launch_ao_mo(a, mode, c->stream);
  }
  if (e != hipSuccess) return hip_fail(handle, e, fn);
  if (d_out == nullptr) c->mo_value_date = c->date;
  return QMCKL_SUCCESS;
}

}  // namespace qmckl_amd

namespace {

qmckl_exit_code copy_out(Context* c, qmckl_context context, const void* d_src, void* dst, size_t bytes,
                         const char* fn) {
  const hipError_t e = copy_to_host(c, dst, d_src, bytes);
  if (e != hipSuccess) return hip_fail(context, e, fn);
  return QMCKL_SUCCESS;
}

}  // namespace

extern "C" {

/* ------------------------------------------------------------------ AO setters */
// org/qmckl_ao.org:422-450
qmckl_exit_code qmckl_set_ao_basis_type(qmckl_context context, const char basis_type) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  qmckl_exit_code rc = ao_pre(c, context, 1u);
  if (rc != QMCKL_SUCCESS) return rc;
  if (basis_type != 'G' && basis_type != 'S')
    return fail(context, QMCKL_INVALID_ARG_2, "qmckl_set_ao_basis_type", nullptr);
  c->ao_type = basis_type;
  return ao_post(c, context, 1u);
}

// org/qmckl_ao.org:451-488
qmckl_exit_code qmckl_set_ao_basis_shell_num(qmckl_context context, const int64_t shell_num) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  qmckl_exit_code rc = ao_pre(c, context, 1u << 1);
  if (rc != QMCKL_SUCCESS) return rc;
  if (shell_num <= 0) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_set_ao_basis_shell_num", "shell_num <= 0");
  if (0 < c->prim_num && c->prim_num < shell_num)
    return fail(context, QMCKL_INVALID_ARG_2, "qmckl_set_ao_basis_shell_num", "shell_num > prim_num");
  c->shell_num = shell_num;
  return ao_post(c, context, 1u << 1);
}

// org/qmckl_ao.org:489-532
qmckl_exit_code qmckl_set_ao_basis_prim_num(qmckl_context context, const int64_t prim_num) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  qmckl_exit_code rc = ao_pre(c, context, 1u << 2);
  if (rc != QMCKL_SUCCESS) return rc;
  if (prim_num <= 0) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_set_ao_basis_prim_num", "prim_num must be positive");
  if (c->shell_num <= 0) return fail(context, QMCKL_FAILURE, "qmckl_set_ao_basis_prim_num", "shell_num is not set");
  if (prim_num < c->shell_num) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_set_ao_basis_prim_num", "prim_num < shell_num");
  c->prim_num = prim_num;
  return ao_post(c, context, 1u << 2);
}

// org/qmckl_ao.org:534-595
qmckl_exit_code qmckl_set_ao_basis_nucleus_shell_num(qmckl_context context, const int64_t* nucleus_shell_num,
                                                     const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  return set_array(c, context, 1u << 3, nucleus_shell_num, c->nucl_num, size_max, c->nucleus_shell_num,
                   "qmckl_set_ao_basis_nucleus_shell_num", "nucl_num is not set");
}

// org/qmckl_ao.org:596-655
qmckl_exit_code qmckl_set_ao_basis_nucleus_index(qmckl_context context, const int64_t* nucleus_index,
                                                 const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  return set_array(c, context, 1u << 4, nucleus_index, c->nucl_num, size_max, c->nucleus_index,
                   "qmckl_set_ao_basis_nucleus_index", "nucl_num is not set");
}

// org/qmckl_ao.org:657-718
qmckl_exit_code qmckl_set_ao_basis_shell_ang_mom(qmckl_context context, const int32_t* shell_ang_mom,
                                                 const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  return set_array(c, context, 1u << 5, shell_ang_mom, c->shell_num, size_max, c->shell_ang_mom,
                   "qmckl_set_ao_basis_shell_ang_mom", "shell_num is not set");
}

// org/qmckl_ao.org:720-781
qmckl_exit_code qmckl_set_ao_basis_shell_prim_num(qmckl_context context, const int64_t* shell_prim_num,
                                                  const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  return set_array(c, context, 1u << 6, shell_prim_num, c->shell_num, size_max, c->shell_prim_num,
                   "qmckl_set_ao_basis_shell_prim_num", "shell_num is not set");
}

// org/qmckl_ao.org:783-843
qmckl_exit_code qmckl_set_ao_basis_shell_prim_index(qmckl_context context, const int64_t* shell_prim_index,
                                                    const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  return set_array(c, context, 1u << 7, shell_prim_index, c->shell_num, size_max, c->shell_prim_index,
                   "qmckl_set_ao_basis_shell_prim_index", "shell_num is not set");
}

// org/qmckl_ao.org:845-905
qmckl_exit_code qmckl_set_ao_basis_shell_factor(qmckl_context context, const double* shell_factor,
                                                const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  return set_array(c, context, 1u << 8, shell_factor, c->shell_num, size_max, c->shell_factor,
                   "qmckl_set_ao_basis_shell_factor", "shell_num is not set");
}

// org/qmckl_ao.org:907-967
qmckl_exit_code qmckl_set_ao_basis_exponent(qmckl_context context, const double* exponent, const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  return set_array(c, context, 1u << 9, exponent, c->prim_num, size_max, c->exponent, "qmckl_set_ao_basis_exponent",
                   "prim_num is not set");
}

// org/qmckl_ao.org:969-1029
qmckl_exit_code qmckl_set_ao_basis_coefficient(qmckl_context context, const double* coefficient,
                                               const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  return set_array(c, context, 1u << 10, coefficient, c->prim_num, size_max, c->coefficient,
                   "qmckl_set_ao_basis_coefficient", "prim_num is not set");
}

// org/qmckl_ao.org:1031-1091
qmckl_exit_code qmckl_set_ao_basis_prim_factor(qmckl_context context, const double* prim_factor,
                                               const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  return set_array(c, context, 1u << 11, prim_factor, c->prim_num, size_max, c->prim_factor,
                   "qmckl_set_ao_basis_prim_factor", "prim_num is not set");
}

// org/qmckl_ao.org:1093-1136
qmckl_exit_code qmckl_set_ao_basis_ao_num(qmckl_context context, const int64_t ao_num) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  qmckl_exit_code rc = ao_pre(c, context, 1u << 12);
  if (rc != QMCKL_SUCCESS) return rc;
  if (ao_num <= 0) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_set_ao_basis_ao_num", "ao_num must be positive");
  if (c->shell_num <= 0) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_set_ao_basis_ao_num", "shell_num is not set");
  if (ao_num < c->shell_num) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_set_ao_basis_ao_num", "ao_num < shell_num");
  c->ao_num = ao_num;
  return ao_post(c, context, 1u << 12);
}

// org/qmckl_ao.org:1138-1197
qmckl_exit_code qmckl_set_ao_basis_ao_factor(qmckl_context context, const double* ao_factor, const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  return set_array(c, context, 1u << 13, ao_factor, c->ao_num, size_max, c->ao_factor, "qmckl_set_ao_basis_ao_factor",
                   "ao_num is not set");
}

// org/qmckl_ao.org:1199-1262 (optional; Gaussian shells require zeros)
qmckl_exit_code qmckl_set_ao_basis_r_power(qmckl_context context, const int32_t* r_power, const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (c->shell_num == 0) return fail(context, QMCKL_FAILURE, "qmckl_set_ao_basis_r_power", "shell_num is not set");
  if (r_power == nullptr) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_set_ao_basis_r_power", "null pointer");
  if (size_max < c->shell_num) return fail(context, QMCKL_INVALID_ARG_3, "qmckl_set_ao_basis_r_power", "input array too small");
  c->r_power.assign(r_power, r_power + c->shell_num);
  return ao_post(c, context, 0u);
}

// org/qmckl_ao.org:1264-1283 (stored; only Cartesian polynomials are evaluated, as in the reference)
qmckl_exit_code qmckl_set_ao_basis_cartesian(qmckl_context context, const bool cartesian) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  c->ao_cartesian = cartesian;
  touch_const(c);
  return QMCKL_SUCCESS;
}

/* ------------------------------------------------------------------ AO getters, org/qmckl_ao.org:1483-2260 */
qmckl_exit_code qmckl_get_ao_basis_type(const qmckl_context context, char* const basis_type) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (basis_type == nullptr) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_get_ao_basis_type", "NULL pointer");
  if (c->ao_uninit & 1u) return fail(context, QMCKL_NOT_PROVIDED, "qmckl_get_ao_basis_type", nullptr);
  *basis_type = c->ao_type;
  return QMCKL_SUCCESS;
}
qmckl_exit_code qmckl_get_ao_basis_shell_num(const qmckl_context context, int64_t* const shell_num) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (shell_num == nullptr) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_get_ao_basis_shell_num", "NULL pointer");
  if (c->ao_uninit & (1u << 1)) return fail(context, QMCKL_NOT_PROVIDED, "qmckl_get_ao_basis_shell_num", nullptr);
  *shell_num = c->shell_num;
  return QMCKL_SUCCESS;
}
qmckl_exit_code qmckl_get_ao_basis_prim_num(const qmckl_context context, int64_t* const prim_num) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (prim_num == nullptr) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_get_ao_basis_prim_num", "NULL pointer");
  if (c->ao_uninit & (1u << 2)) return fail(context, QMCKL_NOT_PROVIDED, "qmckl_get_ao_basis_prim_num", nullptr);
  *prim_num = c->prim_num;
  return QMCKL_SUCCESS;
}
qmckl_exit_code qmckl_get_ao_basis_nucleus_shell_num(const qmckl_context context, int64_t* const v, const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  return get_array(c, context, 1u << 3, v, c->nucl_num, size_max, c->nucleus_shell_num, "qmckl_get_ao_basis_nucleus_shell_num");
}
qmckl_exit_code qmckl_get_ao_basis_nucleus_index(const qmckl_context context, int64_t* const v, const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  return get_array(c, context, 1u << 4, v, c->nucl_num, size_max, c->nucleus_index, "qmckl_get_ao_basis_nucleus_index");
}
qmckl_exit_code qmckl_get_ao_basis_r_power(const qmckl_context context, int32_t* const v, const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (v == nullptr) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_get_ao_basis_r_power", "NULL pointer");
  if (c->r_power.empty()) return fail(context, QMCKL_NOT_PROVIDED, "qmckl_get_ao_basis_r_power", nullptr);
  if (size_max < c->shell_num) return fail(context, QMCKL_INVALID_ARG_3, "qmckl_get_ao_basis_r_power", "input array too small");
  std::copy(c->r_power.begin(), c->r_power.end(), v);
  return QMCKL_SUCCESS;
}
qmckl_exit_code qmckl_get_ao_basis_shell_ang_mom(const qmckl_context context, int32_t* const v, const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  return get_array(c, context, 1u << 5, v, c->shell_num, size_max, c->shell_ang_mom, "qmckl_get_ao_basis_shell_ang_mom");
}
qmckl_exit_code qmckl_get_ao_basis_shell_prim_num(const qmckl_context context, int64_t* const v, const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  return get_array(c, context, 1u << 6, v, c->shell_num, size_max, c->shell_prim_num, "qmckl_get_ao_basis_shell_prim_num");
}
qmckl_exit_code qmckl_get_ao_basis_shell_prim_index(const qmckl_context context, int64_t* const v, const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  return get_array(c, context, 1u << 7, v, c->shell_num, size_max, c->shell_prim_index, "qmckl_get_ao_basis_shell_prim_index");
}
qmckl_exit_code qmckl_get_ao_basis_shell_factor(const qmckl_context context, double* const v, const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  return get_array(c, context, 1u << 8, v, c->shell_num, size_max, c->shell_factor, "qmckl_get_ao_basis_shell_factor");
}
qmckl_exit_code qmckl_get_ao_basis_exponent(const qmckl_context context, double* const v, const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  return get_array(c, context, 1u << 9, v, c->prim_num, size_max, c->exponent, "qmckl_get_ao_basis_exponent");
}
qmckl_exit_code qmckl_get_ao_basis_coefficient(const qmckl_context context, double* const v, const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  return get_array(c, context, 1u << 10, v, c->prim_num, size_max, c->coefficient, "qmckl_get_ao_basis_coefficient");
}
qmckl_exit_code qmckl_get_ao_basis_prim_factor(const qmckl_context context, double* const v, const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  return get_array(c, context, 1u << 11, v, c->prim_num, size_max, c->prim_factor, "qmckl_get_ao_basis_prim_factor");
}
qmckl_exit_code qmckl_get_ao_basis_ao_num(const qmckl_context context, int64_t* const ao_num) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (ao_num == nullptr) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_get_ao_basis_ao_num", "NULL pointer");
  if (c->ao_uninit & (1u << 12)) return fail(context, QMCKL_NOT_PROVIDED, "qmckl_get_ao_basis_ao_num", nullptr);
  *ao_num = c->ao_num;
  return QMCKL_SUCCESS;
}
qmckl_exit_code qmckl_get_ao_basis_ao_factor(const qmckl_context context, double* const v, const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  return get_array(c, context, 1u << 13, v, c->ao_num, size_max, c->ao_factor, "qmckl_get_ao_basis_ao_factor");
}
bool qmckl_ao_basis_provided(const qmckl_context context) {
  Context* c = ctx_of(context);
  return c ? c->ao_provided : false;
}

/* ------------------------------------------------------------------ AO values */
// org/qmckl_ao.org:3417-3480: primitive_vgl[point][5][prim_num]
qmckl_exit_code qmckl_get_ao_basis_primitive_vgl(qmckl_context context, double* const out, const int64_t size_max) {
  const char* fn = "qmckl_get_ao_basis_primitive_vgl";
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  qmckl_exit_code rc = ao_ready(c, context, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  const int64_t n = c->point_num * 5 * c->prim_num;
  if (out == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "NULL pointer");
  if (size_max < n) return fail(context, QMCKL_INVALID_ARG_3, fn, "input array too small");
  // Gaussian primitives only, as the reference (org/qmckl_ao.org:4596-4625)
  if (c->tables.slater) return fail(context, QMCKL_FAILURE, "compute_ao_basis_primitive_vgl", "Not yet implemented");
  hipError_t e = c->d_jint.reserve(sizeof(double) * n);
  if (e == hipSuccess)
    e = launch_prim_vgl(ao_view(c), c->prim_num, c->dc.prim_nucl_c.as<int32_t>(), c->dc.prim_expo_c.as<double>(),
                        c->d_coord.as<double>(), c->point_num, c->d_jint.as<double>(), c->stream);
  if (e != hipSuccess) return hip_fail(context, e, fn);
  c->jint_date = 0;
  return copy_out(c, context, c->d_jint.ptr, out, sizeof(double) * n, fn);
}

// org/qmckl_ao.org:3486-3545: shell_vgl[point][5][shell_num]
qmckl_exit_code qmckl_get_ao_basis_shell_vgl(qmckl_context context, double* const out, const int64_t size_max) {
  const char* fn = "qmckl_get_ao_basis_shell_vgl";
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  qmckl_exit_code rc = ao_ready(c, context, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  const int64_t n = c->point_num * 5 * c->shell_num;
  if (out == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "NULL pointer");
  if (size_max < n) return fail(context, QMCKL_INVALID_ARG_3, fn, "input array too small");
  hipError_t e = c->d_jint.reserve(sizeof(double) * n);
  if (e == hipSuccess) e = launch_shell_vgl(ao_view(c), c->d_coord.as<double>(), c->point_num, c->d_jint.as<double>(), c->stream);
  if (e != hipSuccess) return hip_fail(context, e, fn);
  c->jint_date = 0;
  return copy_out(c, context, c->d_jint.ptr, out, sizeof(double) * n, fn);
}

// org/qmckl_ao.org:3547-3600
qmckl_exit_code qmckl_get_ao_basis_ao_vgl(qmckl_context context, double* const ao_vgl, const int64_t size_max) {
  const char* fn = "qmckl_get_ao_basis_ao_vgl";
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (!c->ao_provided) return fail(context, QMCKL_NOT_PROVIDED, fn, nullptr);
  if (ao_vgl == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "NULL pointer");
  const int64_t n = c->point_num * 5 * c->ao_num;
  if (size_max < n) return fail(context, QMCKL_INVALID_ARG_3, fn, "input array too small");
  qmckl_exit_code rc = provide_ao_vgl(context, c, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  return copy_out(c, context, c->d_ao_vgl.ptr, ao_vgl, sizeof(double) * n, fn);
}

// org/qmckl_ao.org:3609-3670: same result, the caller's array is the output
qmckl_exit_code qmckl_get_ao_basis_ao_vgl_inplace(qmckl_context context, double* const ao_vgl, const int64_t size_max) {
  return qmckl_get_ao_basis_ao_vgl(context, ao_vgl, size_max);
}

// org/qmckl_ao.org:3675-3735: ao_value[point][ao_num]
qmckl_exit_code qmckl_get_ao_basis_ao_value(qmckl_context context, double* const ao_value, const int64_t size_max) {
  const char* fn = "qmckl_get_ao_basis_ao_value";
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (!c->ao_provided) return fail(context, QMCKL_NOT_PROVIDED, fn, nullptr);
  if (ao_value == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "NULL pointer");
  const int64_t n = c->point_num * c->ao_num;
  if (size_max < n) return fail(context, QMCKL_INVALID_ARG_3, fn, "input array too small");
  qmckl_exit_code rc = ao_ready(c, context, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  hipError_t e = c->d_jint.reserve(sizeof(double) * n);
  if (e == hipSuccess)
    e = launch_ao_vgl(ao_view(c), c->d_coord.as<double>(), c->point_num, 0, c->point_num, c->d_jint.ptr, 0,
                      c->ao_num, 0, 1, c->stream);
  if (e != hipSuccess) return hip_fail(context, e, fn);
  c->jint_date = 0;
  return copy_out(c, context, c->d_jint.ptr, ao_value, sizeof(double) * n, fn);
}

qmckl_exit_code qmckl_get_ao_basis_ao_value_inplace(qmckl_context context, double* const ao_value, const int64_t size_max) {
  return qmckl_get_ao_basis_ao_value(context, ao_value, size_max);
}

qmckl_exit_code qmckl_amd_get_ao_basis_ao_vgl_device(qmckl_context context, double* const d_ao_vgl, const int64_t size_max) {
  const char* fn = "qmckl_amd_get_ao_basis_ao_vgl_device";
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (!c->ao_provided) return fail(context, QMCKL_NOT_PROVIDED, fn, nullptr);
  if (d_ao_vgl == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "NULL pointer");
  const int64_t n = c->point_num * 5 * c->ao_num;
  if (size_max < n) return fail(context, QMCKL_INVALID_ARG_3, fn, "input array too small");
  qmckl_exit_code rc = ao_ready(c, context, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  hipError_t e = launch_ao_vgl(ao_view(c), c->d_coord.as<double>(), c->point_num, 0, c->point_num, d_ao_vgl, 0,
                               c->ao_num, 0, 0, c->stream);
  if (e != hipSuccess) return hip_fail(context, e, fn);
  return QMCKL_SUCCESS;
}

/* ------------------------------------------------------------------ MO basis */
// org/qmckl_mo.org:187-300
qmckl_exit_code qmckl_set_mo_basis_mo_num(qmckl_context context, const int64_t mo_num) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (!(c->mo_uninit & 1u)) return fail(context, QMCKL_ALREADY_SET, "qmckl_set_mo_basis_mo_num", nullptr);
  if (mo_num <= 0) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_set_mo_basis_mo_num", "mo_num <= 0");
  c->mo_num = mo_num;
  c->mo_uninit &= ~1u;
  c->mo_provided = (c->mo_uninit == 0);
  touch_const(c);
  return QMCKL_SUCCESS;
}

// org/qmckl_mo.org:230-300 (may be called again to replace the coefficients)
qmckl_exit_code qmckl_set_mo_basis_coefficient(qmckl_context context, const double* coefficient, const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (c->mo_num <= 0) return fail(context, QMCKL_FAILURE, "qmckl_set_mo_basis_coefficient", "mo_num is not set");
  if (c->ao_num <= 0) return fail(context, QMCKL_FAILURE, "qmckl_set_mo_basis_coefficient", "ao_num is not set");
  if (coefficient == nullptr) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_set_mo_basis_coefficient", "NULL pointer");
  const int64_t n = c->mo_num * c->ao_num;
  if (size_max < n) return fail(context, QMCKL_INVALID_ARG_3, "qmckl_set_mo_basis_coefficient", "input array too small");
  c->mo_coef.assign(coefficient, coefficient + n);
  c->mo_uninit &= ~2u;
  c->mo_provided = (c->mo_uninit == 0);
  touch_const(c);
  return QMCKL_SUCCESS;
}

// org/qmckl_mo.org:705-740
qmckl_exit_code qmckl_get_mo_basis_mo_num(const qmckl_context context, int64_t* mo_num) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (mo_num == nullptr) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_get_mo_basis_mo_num", "NULL pointer");
  if (c->mo_uninit & 1u) return fail(context, QMCKL_NOT_PROVIDED, "qmckl_get_mo_basis_mo_num", nullptr);
  *mo_num = c->mo_num;
  return QMCKL_SUCCESS;
}

// org/qmckl_mo.org:742-797
qmckl_exit_code qmckl_get_mo_basis_coefficient(const qmckl_context context, double* const coefficient, const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (coefficient == nullptr) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_get_mo_basis_coefficient", "NULL pointer");
  if (c->mo_uninit & 2u) return fail(context, QMCKL_NOT_PROVIDED, "qmckl_get_mo_basis_coefficient", nullptr);
  const int64_t n = c->mo_num * c->ao_num;
  if (size_max < n) return fail(context, QMCKL_INVALID_ARG_3, "qmckl_get_mo_basis_coefficient", "input array too small");
  std::copy(c->mo_coef.begin(), c->mo_coef.begin() + n, coefficient);
  return QMCKL_SUCCESS;
}

bool qmckl_mo_basis_provided(const qmckl_context context) {
  Context* c = ctx_of(context);
  return c ? c->mo_provided : false;
}

// org/qmckl_mo.org:882-985: keep the MOs with keep[i] != 0
qmckl_exit_code qmckl_mo_basis_select_mo(const qmckl_context context, const int32_t* keep, const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (!c->mo_provided) return fail(context, QMCKL_NOT_PROVIDED, "qmckl_get_mo_basis_select_mo", nullptr);
  if (keep == nullptr) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_get_mo_basis_select_mo", "NULL pointer");
  if (size_max < c->mo_num)
    return fail(context, QMCKL_INVALID_ARG_3, "qmckl_get_mo_basis_select_mo", "Array too small: expected mo_num.");
  const int64_t Nao = c->ao_num;
  std::vector<double> coef;
  int64_t k = 0;
  for (int64_t i = 0; i < c->mo_num; ++i)
    if (keep[i] != 0) {
      coef.insert(coef.end(), c->mo_coef.begin() + i * Nao, c->mo_coef.begin() + (i + 1) * Nao);
      ++k;
    }
  c->mo_coef.swap(coef);
  c->mo_num = k;
  touch_const(c);
  if (!c->r_cusp.empty()) {
    // the cusp fit is per MO: fitted again for the kept MOs (org/qmckl_mo.org:959-967)
    const std::vector<double> r(c->r_cusp);
    return qmckl_set_mo_basis_r_cusp(context, r.data(), (int64_t)r.size());
  }
  return QMCKL_SUCCESS;
}

// org/qmckl_mo.org:3472-3530: multiplies all MO coefficients by scaling_factor
qmckl_exit_code qmckl_mo_basis_rescale(qmckl_context context, const double scaling_factor) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (!c->mo_provided) return fail(context, QMCKL_NOT_PROVIDED, "qmckl_mo_basis_rescale", nullptr);
  for (double& x : c->mo_coef) x *= scaling_factor;
  touch_const(c);
  return QMCKL_SUCCESS;
}

// org/qmckl_mo.org:1626-1686
qmckl_exit_code qmckl_get_mo_basis_mo_vgl(qmckl_context context, double* const mo_vgl, const int64_t size_max) {
  const char* fn = "qmckl_get_mo_basis_mo_vgl";
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (!c->mo_provided) return fail(context, QMCKL_NOT_PROVIDED, fn, nullptr);
  if (mo_vgl == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "NULL pointer");
  const int64_t n = c->point_num * 5 * c->mo_num;
  if (size_max < n) return fail(context, QMCKL_INVALID_ARG_3, fn, "input array too small");
  qmckl_exit_code rc = provide_mo_vgl(context, c, nullptr, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  return copy_out(c, context, c->d_mo_vgl.ptr, mo_vgl, sizeof(double) * n, fn);
}

qmckl_exit_code qmckl_get_mo_basis_mo_vgl_inplace(qmckl_context context, double* const mo_vgl, const int64_t size_max) {
  return qmckl_get_mo_basis_mo_vgl(context, mo_vgl, size_max);
}

// org/qmckl_mo.org:988-1100: mo_value[point][mo_num] (the value row of the MO VGL)
qmckl_exit_code qmckl_get_mo_basis_mo_value(qmckl_context context, double* const mo_value, const int64_t size_max) {
  const char* fn = "qmckl_get_mo_basis_mo_value";
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (!c->mo_provided) return fail(context, QMCKL_NOT_PROVIDED, fn, nullptr);
  if (mo_value == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "NULL pointer");
  const int64_t n = c->point_num * c->mo_num;
  if (size_max < n) return fail(context, QMCKL_INVALID_ARG_3, fn, "input array too small");
  qmckl_exit_code rc = provide_mo_value(context, c, nullptr, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  return copy_out(c, context, c->d_mo_value.ptr, mo_value, sizeof(double) * n, fn);
}

qmckl_exit_code qmckl_get_mo_basis_mo_value_inplace(qmckl_context context, double* const mo_value, const int64_t size_max) {
  return qmckl_get_mo_basis_mo_value(context, mo_value, size_max);
}

qmckl_exit_code qmckl_amd_set_mo_precision(qmckl_context context, const int32_t mode) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (mode != QMCKL_AMD_MO_FP64 && mode != QMCKL_AMD_MO_FP32 && mode != QMCKL_AMD_MO_BF16X3)
    return fail(context, QMCKL_INVALID_ARG_2, "qmckl_amd_set_mo_precision", "unknown mode");
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  c->mo_precision = mode;
  c->mo_vgl_date = 0;
  c->mo_value_date = 0;
  return QMCKL_SUCCESS;
}

qmckl_exit_code qmckl_amd_set_mo_tile(qmckl_context context, const int32_t col_tiles, const int32_t waves) {
  const char* fn = "qmckl_amd_set_mo_tile";
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (col_tiles < 0 || col_tiles > AO_MO_MAX_G) return fail(context, QMCKL_INVALID_ARG_2, fn, "column tiles: 0 (default) or 1..16");
  if (col_tiles > 0 && !ao_mo_variant_ok(0, col_tiles, waves))
    return fail(context, QMCKL_INVALID_ARG_3, fn, "waves: 4 for 1..8 column tiles, 8 for 5..16");
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (c->mo_tile_g == col_tiles && c->mo_tile_nw == (col_tiles > 0 ? waves : 0)) return QMCKL_SUCCESS;
  c->mo_tile_g = col_tiles;
  c->mo_tile_nw = col_tiles > 0 ? waves : 0;
  touch_const(c);  // the operand layout of C_t depends on the shape
  return QMCKL_SUCCESS;
}

qmckl_exit_code qmckl_amd_get_mo_tile(qmckl_context context, int32_t* const col_tiles, int32_t* const waves,
                                      int32_t* const passes) {
  const char* fn = "qmckl_amd_get_mo_tile";
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (col_tiles == nullptr || waves == nullptr || passes == nullptr)
    return fail(context, QMCKL_INVALID_ARG_2, fn, "NULL pointer");
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (!c->mo_provided) return fail(context, QMCKL_NOT_PROVIDED, fn, nullptr);
  const int64_t nct = ((c->mo_num + 15) / 16);
  int G = 0, nw = 0;
  mo_tile_shape(c, nct, G, nw);
  *col_tiles = G;
  *waves = nw;
  *passes = (int32_t)((nct + G - 1) / G);
  return QMCKL_SUCCESS;
}

qmckl_exit_code qmckl_amd_get_mo_precision(qmckl_context context, int32_t* const mode) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (mode == nullptr) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_amd_get_mo_precision", "NULL pointer");
  *mode = mo_mode(c);
  return QMCKL_SUCCESS;
}

qmckl_exit_code qmckl_amd_get_mo_basis_mo_vgl_device(qmckl_context context, double* const d_mo_vgl, const int64_t size_max) {
  const char* fn = "qmckl_amd_get_mo_basis_mo_vgl_device";
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (!c->mo_provided) return fail(context, QMCKL_NOT_PROVIDED, fn, nullptr);
  if (d_mo_vgl == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "NULL pointer");
  const int64_t n = c->point_num * 5 * c->mo_num;
  if (size_max < n) return fail(context, QMCKL_INVALID_ARG_3, fn, "input array too small");
  return provide_mo_vgl(context, c, d_mo_vgl, fn);
}

qmckl_exit_code qmckl_amd_get_mo_basis_mo_value_device(qmckl_context context, double* const d_mo_value,
                                                       const int64_t size_max) {
  const char* fn = "qmckl_amd_get_mo_basis_mo_value_device";
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (!c->mo_provided) return fail(context, QMCKL_NOT_PROVIDED, fn, nullptr);
  if (d_mo_value == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "NULL pointer");
  const int64_t n = c->point_num * c->mo_num;
  if (size_max < n) return fail(context, QMCKL_INVALID_ARG_3, fn, "input array too small");
  return provide_mo_value(context, c, d_mo_value, fn);
}

}  // extern "C"
