// CHAMP Jastrow factor: parameter setters/getters, finalisation (asymptotic shifts),
// the provider of value and gradient/Laplacian, and the getters of its intermediates.
#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <vector>

#include "internal.hpp"

using namespace qmckl_amd;

namespace {

// org/qmckl_jastrow_champ.org:11365-11405 (qmckl_compute_dim_c_vector)
int64_t dim_c_of(int64_t cord) {
  int64_t n = 0;
  for (int64_t p = 2; p <= cord; ++p)
    for (int64_t k = p - 1; k >= 0; --k) {
      const int64_t lmax = (k != 0) ? p - k : p - k - 2;
      for (int64_t l = lmax; l >= 0; --l)
        if (((p - k - l) & 1) == 0) ++n;
    }
  return n;
}

// pre2 of org/qmckl_jastrow_champ.org:545-560
qmckl_exit_code jas_pre(Context* c, qmckl_context ctx, uint32_t mask) {
  if (mask != 0 && !(c->jas_uninit & mask)) return fail(ctx, QMCKL_ALREADY_SET, "qmckl_set_jastrow_champ_*", nullptr);
  return QMCKL_SUCCESS;
}

// org/qmckl_jastrow_champ.org:1036-1085 (qmckl_finalize_jastrow_champ): the asymptotic
// shifts asymp_jasa (:5830-5875) and asymp_jasb (:2271-2320).
qmckl_exit_code finalize_jastrow(Context* c, qmckl_context ctx) {
  if (!c->elec_provided) return fail(ctx, QMCKL_NOT_PROVIDED, "qmckl_electron", nullptr);
  if (!c->nucl_provided) return fail(ctx, QMCKL_NOT_PROVIDED, "qmckl_nucleus", nullptr);
  const double kinv = 1.0 / c->rescale_ee;
  const double* b = c->b_vector.data();
  const double asym_one = b[0] * kinv / (1.0 + b[1] * kinv);
  c->asymp_jasb[0] = (c->spin_independent == 1) ? asym_one : 0.5 * asym_one;
  c->asymp_jasb[1] = asym_one;
  double x = kinv;
  for (int64_t p = 2; p <= c->bord; ++p) {
    x *= kinv;
    c->asymp_jasb[0] += b[p] * x;
    c->asymp_jasb[1] += b[p] * x;
  }
  c->asymp_jasa.assign(c->type_nucl_num, 0.0);
  for (int64_t t = 0; t < c->type_nucl_num; ++t) {
    const double* a = c->a_vector.data() + t * (c->aord + 1);
    const double ki = 1.0 / c->rescale_en[t];
    double v = a[0] * ki / (1.0 + a[1] * ki);
    double y = ki;
    for (int64_t p = 2; p <= c->aord; ++p) {
      y *= ki;
      v += a[p] * y;
    }
    c->asymp_jasa[t] = v;
  }
  touch_const(c);
  return QMCKL_SUCCESS;
}

qmckl_exit_code jas_post(Context* c, qmckl_context ctx, uint32_t mask) {
  c->jas_uninit &= ~mask;
  c->jas_provided = (c->jas_uninit == 0);
  touch_const(c);
  if (c->jas_provided) return finalize_jastrow(c, ctx);
  return QMCKL_SUCCESS;
}

qmckl_exit_code upload_jastrow(Context* c, qmckl_context handle, const char* fn) {
  qmckl_exit_code rc = device_ready(c, handle, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  if (c->dc.jas_date >= c->const_date) return QMCKL_SUCCESS;
  DevConst& d = c->dc;
  hipError_t e = hipSuccess;
  auto up = [&](DevBuf& b, const void* src, size_t bytes) {
    if (e != hipSuccess) return;
    e = b.reserve(bytes);
    if (e == hipSuccess && bytes) e = hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, c->stream);
  };
  // c_vector_full [dim_c][nucl], org/qmckl_jastrow_champ.org:11865-11895
  const int64_t Nn = c->nucl_num, dim_c = c->dim_c;
  std::vector<double> cfull((size_t)std::max<int64_t>(dim_c, 1) * Nn, 0.0);
  for (int64_t n = 0; n < dim_c; ++n)
    for (int64_t a = 0; a < Nn; ++a) cfull[n * Nn + a] = c->c_vector[c->type_nucl_vector[a] * dim_c + n];
  up(d.jas_a, c->a_vector.data(), 8 * c->a_vector.size());
  up(d.jas_b, c->b_vector.data(), 8 * c->b_vector.size());
  up(d.jas_cfull, cfull.data(), 8 * cfull.size());
  // the same with the nuclei padded to whole 16-nucleus tiles (zeros), and the nucleus
  // coordinates and e-n rescaling factors in one table: the e-e-n kernel copies them to LDS
  const int64_t NA16 = 16 * ((Nn + 15) / 16);
  std::vector<double> cpad((size_t)std::max<int64_t>(dim_c, 1) * NA16, 0.0), npad(4 * (size_t)NA16, 0.0);
  for (int64_t n = 0; n < dim_c; ++n)
    for (int64_t a = 0; a < Nn; ++a) cpad[n * NA16 + a] = cfull[n * Nn + a];
  for (int64_t a = 0; a < Nn; ++a) {
    for (int d3 = 0; d3 < 3; ++d3) npad[d3 * NA16 + a] = c->nucl_coord[d3 * Nn + a];
    npad[3 * NA16 + a] = c->rescale_en[c->type_nucl_vector[a]];
  }
  up(d.jas_cfull_pad, cpad.data(), 8 * cpad.size());
  up(d.jas_nuc_pad, npad.data(), 8 * npad.size());
  up(d.jas_type, c->type_nucl_vector.data(), 8 * c->type_nucl_vector.size());
  up(d.jas_kappa_en, c->rescale_en.data(), 8 * c->rescale_en.size());
  up(d.jas_asymp_a, c->asymp_jasa.data(), 8 * c->asymp_jasa.size());
  up(d.nucl_coord, c->nucl_coord.data(), 8 * c->nucl_coord.size());
  up(d.jas_charge, c->nucl_charge.data(), 8 * c->nucl_charge.size());
  if (e != hipSuccess) return hip_fail(handle, e, fn);
  d.jas_date = c->const_date;
  return QMCKL_SUCCESS;
}

JasDevView jas_view(Context* c) {
  JasDevView v;
  v.walk_num = c->walk_num;
  v.elec_num = c->elec_num;
  v.up_num = c->up_num;
  v.nucl_num = c->nucl_num;
  v.aord = c->aord;
  v.bord = c->bord;
  v.cord = c->cord;
  v.dim_c = c->dim_c;
  v.type_num = c->type_nucl_num;
  v.spin_independent = c->spin_independent;
  v.rescale_ee = c->rescale_ee;
  v.coord = c->d_coord.as<double>();
  v.coord_ld = c->point_num;
  v.nucl_coord = c->dc.nucl_coord.as<double>();
  v.type_nucl = c->dc.jas_type.as<int64_t>();
  v.kappa_en = c->dc.jas_kappa_en.as<double>();
  v.a_vector = c->dc.jas_a.as<double>();
  v.b_vector = c->dc.jas_b.as<double>();
  v.cfull = c->dc.jas_cfull.as<double>();
  v.cfull_pad = c->dc.jas_cfull_pad.as<double>();
  v.nuc_pad = c->dc.jas_nuc_pad.as<double>();
  v.asymp_jasb0 = c->asymp_jasb[0];
  v.asymp_jasb1 = c->asymp_jasb[1];
  v.asymp_jasa = c->dc.jas_asymp_a.as<double>();
  return v;
}

qmckl_exit_code jas_ready(Context* c, qmckl_context handle, const char* fn) {
  if (!c->jas_provided) return fail(handle, QMCKL_NOT_PROVIDED, fn, nullptr);
  if (!c->elec_provided || c->walk_num <= 0) return fail(handle, QMCKL_NOT_PROVIDED, fn, "electron coordinates not set");
  qmckl_exit_code rc = upload_jastrow(c, handle, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  return upload_coords(c, handle, fn);
}

qmckl_exit_code copy_out(Context* c, qmckl_context context, const void* d_src, double* dst, int64_t n, const char* fn) {
  const hipError_t e = copy_to_host(c, dst, d_src, sizeof(double) * n);
  if (e != hipSuccess) return hip_fail(context, e, fn);
  return QMCKL_SUCCESS;
}

enum Which { F_EE, F_EN, F_EEN, F_EE_GL, F_EN_GL, F_EEN_GL, J_VALUE, J_GL };

qmckl_exit_code get_result(qmckl_context context, Which which, double* out, int64_t size_max, const char* fn) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (out == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "NULL pointer");
  if (!c->jas_provided) return fail(context, QMCKL_NOT_PROVIDED, fn, nullptr);
  const int64_t W = c->walk_num, Ne = c->elec_num;
  const bool is_gl = which == F_EE_GL || which == F_EN_GL || which == F_EEN_GL || which == J_GL;
  const int64_t n = is_gl ? W * 4 * Ne : W;
  if (size_max < n) return fail(context, QMCKL_INVALID_ARG_3, fn, "Array too small");
  qmckl_exit_code rc = is_gl ? provide_jastrow(context, c, fn) : provide_jastrow_value(context, c, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  const DevBuf* src[] = {&c->d_fee, &c->d_fen, &c->d_feen, &c->d_fee_gl, &c->d_fen_gl, &c->d_feen_gl, &c->d_jvalue, &c->d_jgl};
  return copy_out(c, context, src[which]->ptr, out, n, fn);
}

qmckl_exit_code get_inter(qmckl_context context, int kind, double* out, int64_t size_max, const char* fn) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (out == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "NULL pointer");
  qmckl_exit_code rc = jas_ready(c, context, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  const JasDevView v = jas_view(c);
  const int64_t n = jastrow_inter_size(v, kind);
  if (size_max < n) return fail(context, QMCKL_INVALID_ARG_3, fn, "Array too small");
  if (!(c->jint_date >= c->date && c->jint_kind == kind)) {
    hipError_t e = c->d_jint.reserve(sizeof(double) * std::max<int64_t>(n, 1));
    if (e == hipSuccess) e = launch_jastrow_inter(v, kind, c->d_jint.as<double>(), c->stream);
    if (e != hipSuccess) return hip_fail(context, e, fn);
    c->jint_date = c->date;
    c->jint_kind = kind;
  }
  return copy_out(c, context, c->d_jint.ptr, out, n, fn);
}

// ---- parameter derivatives and nuclear forces ------------------------------------------
// Buffers of Context::d_deriv: 0..9 the JasInter intermediates, then:
enum DerivBuf { DB_EDIST = 10, DB_FTMP_C, DB_FDTMP_C, DB_FEEN_N, DB_PART, DB_OUT, DB_LKPM, DB_AJB, DB_AJA };

// (l, k, p, m) of the c-coefficients in the order of
// org/qmckl_jastrow_champ.org:12000-12040 (lkpm_combined_index), as [4][dim].
std::vector<int64_t> lkpm_table(int64_t cord) {
  const int64_t dim = dim_c_of(cord);
  std::vector<int64_t> t(4 * std::max<int64_t>(dim, 1), 0);
  int64_t n = 0;
  for (int64_t p = 2; p <= cord; ++p)
    for (int64_t k = p - 1; k >= 0; --k) {
      const int64_t lmax = (k != 0) ? p - k : p - k - 2;
      for (int64_t l = lmax; l >= 0; --l) {
        if (((p - k - l) & 1) == 1) continue;
        t[n] = l;
        t[dim + n] = k;
        t[2 * dim + n] = p;
        t[3 * dim + n] = (p - k - l) / 2;
        ++n;
      }
    }
  return t;
}

// org/qmckl_jastrow_champ.org:4948-4994 (asymp_jasb_pderiv): [2][bord+1], entry [s][p] is
// asymp_jasb_pderiv(p+1, s+1) of the reference.
std::vector<double> asymp_jasb_pderiv_of(const Context* c) {
  const int64_t n = c->bord + 1;
  std::vector<double> o(2 * n, 0.0);
  const double ki = 1.0 / c->rescale_ee, b0 = c->b_vector[0], b1 = c->bord >= 1 ? c->b_vector[1] : 0.0;
  const double one = ki / (1.0 + b1 * ki), two = -b0 * ki * ki / ((1.0 + b1 * ki) * (1.0 + b1 * ki));
  const bool si = c->spin_independent == 1;
  o[0] = si ? one : 0.5 * one;
  o[n] = one;
  if (c->bord >= 1) {
    o[1] = si ? two : 0.5 * two;
    o[n + 1] = two;
  }
  double x = ki;
  for (int64_t p = 2; p <= c->bord; ++p) {
    x *= ki;
    o[p] = x;
    o[n + p] = x;
  }
  return o;
}

// org/qmckl_jastrow_champ.org:8192-8215 (asymp_jasa_pderiv): [type][aord+1].
std::vector<double> asymp_jasa_pderiv_of(const Context* c) {
  const int64_t n = c->aord + 1;
  std::vector<double> o(c->type_nucl_num * n, 0.0);
  for (int64_t t = 0; t < c->type_nucl_num; ++t) {
    const double* av = c->a_vector.data() + t * n;
    const double ki = 1.0 / c->rescale_en[t];
    o[t * n] = ki / (1.0 + av[1] * ki);
    o[t * n + 1] = -av[0] * ki * ki / ((1.0 + av[1] * ki) * (1.0 + av[1] * ki));
    double x = ki;
    for (int64_t p = 2; p <= c->aord; ++p) {
      x *= ki;
      o[t * n + p] = x;
    }
  }
  return o;
}

hipError_t upload_table(Context* c, int slot, const void* src, size_t bytes) {
  hipError_t e = c->d_deriv[slot].reserve(std::max<size_t>(bytes, 8));
  if (e == hipSuccess) e = hipMemcpy(c->d_deriv[slot].ptr, src, bytes, hipMemcpyHostToDevice);
  return e;
}

// A reference-layout intermediate, computed at most once per context date.
hipError_t need_inter(Context* c, const JasDevView& v, int kind) {
  if (c->deriv_date[kind] >= c->date && c->d_deriv[kind].ptr) return hipSuccess;
  const int64_t n = jastrow_inter_size(v, kind);
  hipError_t e = c->d_deriv[kind].reserve(sizeof(double) * std::max<int64_t>(n, 1));
  if (e == hipSuccess) e = launch_jastrow_inter(v, kind, c->d_deriv[kind].as<double>(), c->stream);
  if (e == hipSuccess) c->deriv_date[kind] = c->date;
  return e;
}

// A force intermediate (forces_tmp_c, forces_dtmp_c, forces_een_rescaled_n_gl, or the e-n
// distances), computed at most once per context date.
hipError_t need_force_inter(Context* c, const JasDevView& v, DerivArgs& d, int slot) {
  if (c->deriv_date[slot] >= c->date && c->d_deriv[slot].ptr) return hipSuccess;
  hipError_t e = hipSuccess;
  int64_t n = 0;
  if (slot == DB_EDIST) {
    n = v.walk_num * v.elec_num * v.nucl_num;
    e = c->d_deriv[slot].reserve(sizeof(double) * std::max<int64_t>(n, 1));
    if (e == hipSuccess)
      e = launch_en_distance(v.coord, v.walk_num * v.elec_num, v.coord_ld, v.nucl_coord, v.nucl_num,
                             c->d_deriv[slot].as<double>(), c->stream);
  } else {
    const int what = slot == DB_FTMP_C ? DV_FORCES_TMP_C : slot == DB_FDTMP_C ? DV_FORCES_DTMP_C : DV_FORCES_EEN_N;
    n = deriv_size(d, what);
    e = c->d_deriv[slot].reserve(sizeof(double) * std::max<int64_t>(n, 1));
    if (e == hipSuccess) e = launch_deriv(d, what, c->d_deriv[slot].as<double>(), c->stream);
  }
  if (e == hipSuccess) c->deriv_date[slot] = c->date;
  return e;
}

// Fills `d` and provides the intermediates `what` reads, for the walkers of the view v (all
// of c's walkers, or a range of them).  With `bytes` non-null nothing is computed: it
// receives the size of those intermediates.
hipError_t prepare_deriv(Context* c, const JasDevView& v, int what, DerivArgs& d, int64_t* bytes = nullptr) {
  d = DerivArgs{};
  d.W = v.walk_num;
  d.Ne = c->elec_num;
  d.Nn = c->nucl_num;
  d.up = c->up_num;
  d.C = c->cord;
  d.dim = c->dim_c;
  d.ntype = c->type_nucl_num;
  d.aord = c->aord;
  d.bord = c->bord;
  d.spin_indep = c->spin_independent;
  d.type = v.type_nucl;
  d.a = v.a_vector;
  d.b = v.b_vector;
  d.kappa = v.kappa_en;
  d.cfull = v.cfull;
  hipError_t e = hipSuccess;
  auto inter = [&](int kind) -> const double* {
    if (bytes) *bytes += 8 * jastrow_inter_size(v, kind);
    else if (e == hipSuccess) e = need_inter(c, v, kind);
    return c->d_deriv[kind].as<double>();
  };
  auto force = [&](int slot) -> const double* {
    if (bytes) {
      *bytes += 8 * (slot == DB_EDIST ? v.walk_num * v.elec_num * v.nucl_num
                                      : deriv_size(d, slot == DB_FTMP_C   ? DV_FORCES_TMP_C
                                                      : slot == DB_FDTMP_C ? DV_FORCES_DTMP_C
                                                                           : DV_FORCES_EEN_N));
    } else if (e == hipSuccess) {
      e = need_force_inter(c, v, d, slot);
    }
    return c->d_deriv[slot].as<double>();
  };
  if (what == DV_EE_PDERIV || what == DV_EE_GL_PDERIV) {
    d.ee = inter(J_EE_RESCALED);
    if (what == DV_EE_GL_PDERIV) d.eeg = inter(J_EE_RESCALED_GL);
    const std::vector<double> ajb = asymp_jasb_pderiv_of(c);
    if (e == hipSuccess && !bytes) e = upload_table(c, DB_AJB, ajb.data(), sizeof(double) * ajb.size());
    d.ajb = c->d_deriv[DB_AJB].as<double>();
    return e;
  }
  if (what == DV_EN_PDERIV || what == DV_EN_GL_PDERIV || what == DV_FORCES_EN || what == DV_FORCES_EN_G ||
      what == DV_FORCES_EN_L) {
    d.en = inter(J_EN_RESCALED);
    if (what != DV_EN_PDERIV) d.eng = inter(J_EN_RESCALED_GL);
    if (what == DV_FORCES_EN_G || what == DV_FORCES_EN_L) d.edist = force(DB_EDIST);
    if (what == DV_EN_PDERIV) {
      const std::vector<double> aja = asymp_jasa_pderiv_of(c);
      if (e == hipSuccess && !bytes) e = upload_table(c, DB_AJA, aja.data(), sizeof(double) * aja.size());
      d.aja = c->d_deriv[DB_AJA].as<double>();
    }
    return e;
  }
  // e-e-n quantities
  const std::vector<int64_t> lk = lkpm_table(c->cord);
  if (e == hipSuccess && !bytes) e = upload_table(c, DB_LKPM, lk.data(), sizeof(int64_t) * lk.size());
  d.lkpm = c->d_deriv[DB_LKPM].as<int64_t>();
  d.tmp_c = inter(J_TMP_C);
  d.een_n = inter(J_EEN_N);
  if (what == DV_EEN_PDERIV) return e;
  d.een_n_gl = inter(J_EEN_N_GL);
  if (what == DV_EEN_GL_PDERIV) {
    d.dtmp_c = inter(J_DTMP_C);
    return e;
  }
  d.een_e = inter(J_EEN_E);
  d.ftmp_c = force(DB_FTMP_C);
  if (what == DV_FORCES_EEN) return e;
  d.dtmp_c = inter(J_DTMP_C);
  d.een_e_gl = inter(J_EEN_E_GL);
  d.edist = force(DB_EDIST);
  d.fdtmp_c = force(DB_FDTMP_C);
  d.feen_n = force(DB_FEEN_N);
  return e;
}

// Device memory for the materialised intermediates of one derivative getter: 2 GiB, or
// QMCKL_AMD_DERIV_BUDGET_MB megabytes.
int64_t deriv_budget_bytes() {
  const char* s = std::getenv("QMCKL_AMD_DERIV_BUDGET_MB");
  const long long mb = s ? std::atoll(s) : 0;
  return (mb > 0 ? (int64_t)mb : 2048) << 20;
}

// Host getter of a derivative: `what` evaluated on the device, walker-summed quantities
// reduced over walkers (mean for the e-e and e-n parameter derivatives, sum for e-e-n, as
// the reference), then copied to the caller's buffer.
qmckl_exit_code get_deriv(qmckl_context context, int what, double* out, int64_t size_max, const char* fn) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (out == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "NULL pointer");
  qmckl_exit_code rc = jas_ready(c, context, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  const bool summed = what <= DV_EEN_GL_PDERIV;
  const bool een = what == DV_EEN_PDERIV || what == DV_EEN_GL_PDERIV || what >= DV_FORCES_TMP_C;
  DerivArgs probe{};
  probe.W = c->walk_num;
  probe.Ne = c->elec_num;
  probe.Nn = c->nucl_num;
  probe.C = c->cord;
  probe.dim = c->dim_c;
  probe.ntype = c->type_nucl_num;
  probe.aord = c->aord;
  probe.bord = c->bord;
  const int64_t n_all = deriv_size(probe, what);
  const int64_t n = summed ? n_all / std::max<int64_t>(c->walk_num, 1) : n_all;
  if (size_max < n) return fail(context, QMCKL_INVALID_ARG_3, fn, "Array too small");
  if (een && (c->cord < 2 || c->dim_c == 0)) {  // no e-e-n term: its derivatives vanish
    std::fill(out, out + n, 0.0);
    return QMCKL_SUCCESS;
  }
  // The intermediates (reference layouts, several MB per walker at a hundred electrons) are
  // materialised for ranges of walkers that fit the budget, each range's per-walker results
  // written in place; the walker sums run once over all walkers, in walker order, so the
  // results do not depend on the ranges.
  const JasDevView all = jas_view(c);
  const int64_t W = c->walk_num, per_walker = n_all / std::max<int64_t>(W, 1);
  int64_t wbytes = 0;
  DerivArgs d;
  {
    JasDevView one = all;
    one.walk_num = 1;
    (void)prepare_deriv(c, one, what, d, &wbytes);
  }
  const int64_t chunk = std::max<int64_t>(1, deriv_budget_bytes() / std::max<int64_t>(wbytes, 1));
  hipError_t e = c->d_deriv[DB_OUT].reserve(sizeof(double) * std::max<int64_t>(n, 1));
  if (e == hipSuccess && summed) e = c->d_deriv[DB_PART].reserve(sizeof(double) * std::max<int64_t>(n_all, 1));
  double* dst = summed ? c->d_deriv[DB_PART].as<double>() : c->d_deriv[DB_OUT].as<double>();
  for (int64_t w0 = 0; w0 < W && e == hipSuccess; w0 += chunk) {
    JasDevView v = all;
    if (chunk < W) {  // a range of walkers; nothing cached across ranges
      v.walk_num = std::min(chunk, W - w0);
      v.coord = all.coord + w0 * c->elec_num;
      for (uint64_t& dd : c->deriv_date) dd = 0;
    }
    e = prepare_deriv(c, v, what, d);
    if (e == hipSuccess) e = launch_deriv(d, what, dst + w0 * per_walker, c->stream);
  }
  if (chunk < W)
    for (uint64_t& dd : c->deriv_date) dd = 0;
  if (e == hipSuccess && summed) {
    const double scale = (what == DV_EEN_PDERIV || what == DV_EEN_GL_PDERIV) ? 1.0 : 1.0 / (double)W;
    e = launch_sum_walkers(c->d_deriv[DB_PART].as<double>(), W, n, scale, c->d_deriv[DB_OUT].as<double>(), c->stream);
  }
  if (e != hipSuccess) return hip_fail(context, e, fn);
  return copy_out(c, context, c->d_deriv[DB_OUT].ptr, out, n, fn);
}

}  // namespace

namespace qmckl_amd {

// The Jastrow view of c's walkers, its device tables and coordinates uploaded.
qmckl_exit_code jastrow_view_ready(Context* c, qmckl_context handle, const char* fn, JasDevView& v) {
  qmckl_exit_code rc = jas_ready(c, handle, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  v = jas_view(c);
  return QMCKL_SUCCESS;
}

// The device buffers of c's factors, GLs and values, as the Jastrow kernels take them.
JasOut jastrow_out(Context* c) {
  JasOut o;
  o.fee = c->d_fee.as<double>();
  o.fen = c->d_fen.as<double>();
  o.feen = c->d_feen.as<double>();
  o.fee_gl = c->d_fee_gl.as<double>();
  o.fen_gl = c->d_fen_gl.as<double>();
  o.feen_gl = c->d_feen_gl.as<double>();
  o.value = c->d_jvalue.as<double>();
  o.gl = c->d_jgl.as<double>();
  o.scratch = c->d_jscratch.as<double>();
  return o;
}

bool jastrow_view_resident(Context* c, int64_t walk_num, const double* coord, JasDevView& v) {
  if (c->stream == nullptr || c->device < 0 || c->dc.jas_date < c->const_date) return false;
  v = jas_view(c);
  v.walk_num = walk_num;
  v.coord = coord;
  v.coord_ld = walk_num * c->elec_num;
  return true;
}

// Value, factors and their gradients/Laplacians of all walkers in one pass over the
// device; org/qmckl_jastrow_champ.org:16481-16560 (value) and :16887-16960 (gl).
namespace {
qmckl_exit_code provide_jastrow_full(qmckl_context handle, Context* c, const char* fn);
}

qmckl_exit_code provide_jastrow(qmckl_context handle, Context* c, const char* fn) {
  c->jgl_asked_date = c->date;
  return provide_jastrow_full(handle, c, fn);
}

namespace {
// Factors, e^J and every GL in one evaluation (pair, e-e-n and total kernels).
qmckl_exit_code provide_jastrow_full(qmckl_context handle, Context* c, const char* fn) {
  if (c->jastrow_date >= c->date) return QMCKL_SUCCESS;
  qmckl_exit_code rc = jas_ready(c, handle, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  const int64_t W = c->walk_num, Ne = c->elec_num;
  const JasDevView v = jas_view(c);
  hipError_t e = hipSuccess;
  DevBuf* w1[] = {&c->d_fee, &c->d_fen, &c->d_feen, &c->d_jvalue};
  DevBuf* w4[] = {&c->d_fee_gl, &c->d_fen_gl, &c->d_feen_gl, &c->d_jgl};
  for (DevBuf* b : w1)
    if (e == hipSuccess) e = b->reserve(sizeof(double) * W);
  for (DevBuf* b : w4)
    if (e == hipSuccess) e = b->reserve(sizeof(double) * W * 4 * Ne);
  if (e == hipSuccess) e = c->d_jscratch.reserve(sizeof(double) * std::max<int64_t>(jastrow_scratch_doubles(v), 1));
  if (e != hipSuccess) return hip_fail(handle, e, fn);
  JasOut o;
  o.fee = c->d_fee.as<double>();
  o.fen = c->d_fen.as<double>();
  o.feen = c->d_feen.as<double>();
  o.fee_gl = c->d_fee_gl.as<double>();
  o.fen_gl = c->d_fen_gl.as<double>();
  o.feen_gl = c->d_feen_gl.as<double>();
  o.value = c->d_jvalue.as<double>();
  o.gl = c->d_jgl.as<double>();
  o.scratch = c->d_jscratch.as<double>();
  {
    ProfScope ps(c, SLOT_JPAIR);
    e = launch_jastrow_pair(v, o, c->stream);
  }
  if (e == hipSuccess) {
    ProfScope ps(c, SLOT_JEEN);
    e = launch_jastrow_een(v, o, c->stream);
  }
  if (e == hipSuccess) {
    ProfScope ps(c, SLOT_JTOTAL);
    e = launch_jastrow_total(v, o, c->stream);
  }
  if (e != hipSuccess) return hip_fail(handle, e, fn);
  c->jastrow_date = c->date;
  return QMCKL_SUCCESS;
}
}  // namespace

// The factors ee, en, een and e^J alone, as the reference's value providers give them
// (org/qmckl_jastrow_champ.org:16700-16820: the value from the three factors, no GL): one
// resident e-e-n launch on the value component of its products only, bit-identical to the
// values of the full evaluation.  Where that kernel does not apply, and when the caller asked
// for the GL after its previous value request (it will most likely again: a QMC step reads
// both), the full evaluation, so that the GL request that follows finds it done.
qmckl_exit_code provide_jastrow_value(qmckl_context handle, Context* c, const char* fn) {
  if (c->jastrow_date >= c->date || c->jvalue_date >= c->date) return QMCKL_SUCCESS;
  const bool gl_too = c->jgl_asked_date != 0 && c->jgl_asked_date >= c->jval_asked_date;
  c->jval_asked_date = c->date;
  qmckl_exit_code rc = jas_ready(c, handle, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  const JasDevView v = jas_view(c);
  if (gl_too || !jastrow_values_only_ok(v)) return provide_jastrow_full(handle, c, fn);
  const int64_t W = c->walk_num, Ne = c->elec_num;
  hipError_t e = hipSuccess;
  DevBuf* w1[] = {&c->d_fee, &c->d_fen, &c->d_feen, &c->d_jvalue};
  for (DevBuf* b : w1)
    if (e == hipSuccess) e = b->reserve(sizeof(double) * W);
  // the pair kernel, which runs where the resident e-e-n kernel does not, writes the e-e and
  // e-n GL with the values
  const bool pair_gl = !een_fused(v);
  if (pair_gl && e == hipSuccess) e = c->d_fee_gl.reserve(sizeof(double) * W * 4 * Ne);
  if (pair_gl && e == hipSuccess) e = c->d_fen_gl.reserve(sizeof(double) * W * 4 * Ne);
  if (e == hipSuccess) e = c->d_jscratch.reserve(sizeof(double) * std::max<int64_t>(jastrow_scratch_doubles(v), 1));
  if (e != hipSuccess) return hip_fail(handle, e, fn);
  JasOut o;
  o.fee = c->d_fee.as<double>();
  o.fen = c->d_fen.as<double>();
  o.feen = c->d_feen.as<double>();
  o.fee_gl = pair_gl ? c->d_fee_gl.as<double>() : nullptr;
  o.fen_gl = pair_gl ? c->d_fen_gl.as<double>() : nullptr;
  o.feen_gl = o.gl = nullptr;  // not touched by the values-only kernels
  o.value = c->d_jvalue.as<double>();
  o.scratch = c->d_jscratch.as<double>();
  {
    ProfScope ps(c, SLOT_JEEN);
    e = launch_jastrow_values(v, o, c->stream);
  }
  if (e != hipSuccess) return hip_fail(handle, e, fn);
  c->jvalue_date = c->date;
  return QMCKL_SUCCESS;
}

}  // namespace qmckl_amd

extern "C" {

/* ---- setters, org/qmckl_jastrow_champ.org:566-1000 */
qmckl_exit_code qmckl_set_jastrow_champ_aord_num(qmckl_context context, const int64_t aord_num) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  qmckl_exit_code rc = jas_pre(c, context, 1u << 0);
  if (rc != QMCKL_SUCCESS) return rc;
  if (aord_num < 0) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_set_jastrow_champ_aord_num", "aord_num < 0");
  c->aord = aord_num;
  c->jas_uninit |= (1u << 5);
  return jas_post(c, context, 1u << 0);
}

qmckl_exit_code qmckl_set_jastrow_champ_bord_num(qmckl_context context, const int64_t bord_num) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  qmckl_exit_code rc = jas_pre(c, context, 1u << 1);
  if (rc != QMCKL_SUCCESS) return rc;
  if (bord_num < 0) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_set_jastrow_champ_bord_num", "bord_num < 0");
  c->bord = bord_num;
  c->jas_uninit |= (1u << 6);
  return jas_post(c, context, 1u << 1);
}

qmckl_exit_code qmckl_set_jastrow_champ_cord_num(qmckl_context context, const int64_t cord_num) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  qmckl_exit_code rc = jas_pre(c, context, 1u << 2);
  if (rc != QMCKL_SUCCESS) return rc;
  if (cord_num < 0) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_set_jastrow_champ_cord_num", "cord_num < 0");
  if (cord_num > 8)
    return fail(context, QMCKL_INVALID_ARG_2, "qmckl_set_jastrow_champ_cord_num", "cord_num > 8 is not supported");
  c->cord = cord_num;
  c->dim_c = dim_c_of(cord_num);
  if (cord_num > 0)
    c->jas_uninit |= (1u << 7);
  else
    c->jas_uninit &= ~(1u << 7);
  return jas_post(c, context, 1u << 2);
}

qmckl_exit_code qmckl_set_jastrow_champ_type_nucl_num(qmckl_context context, const int64_t type_nucl_num) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  qmckl_exit_code rc = jas_pre(c, context, 1u << 3);
  if (rc != QMCKL_SUCCESS) return rc;
  if (type_nucl_num <= 0)
    return fail(context, QMCKL_INVALID_ARG_2, "qmckl_set_jastrow_champ_type_nucl_num", "type_nucl_num < 0");
  c->type_nucl_num = type_nucl_num;
  return jas_post(c, context, 1u << 3);
}

qmckl_exit_code qmckl_set_jastrow_champ_type_nucl_vector(qmckl_context context, const int64_t* type_nucl_vector,
                                                         const int64_t size_max) {
  const char* fn = "qmckl_set_jastrow_champ_type_nucl_vector";
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  qmckl_exit_code rc = jas_pre(c, context, 1u << 4);
  if (rc != QMCKL_SUCCESS) return rc;
  const int64_t Nn = c->nucl_num;
  if (c->type_nucl_num <= 0) return fail(context, QMCKL_NOT_PROVIDED, fn, "type_nucl_num not initialized");
  if (type_nucl_vector == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "type_nucl_vector = NULL");
  if (Nn <= 0) return fail(context, QMCKL_NOT_PROVIDED, fn, "nucl_num not initialized");
  if (size_max < Nn) return fail(context, QMCKL_INVALID_ARG_3, fn, "Array too small");
  for (int64_t i = 0; i < Nn; ++i) {
    if (type_nucl_vector[i] < 0) return fail(context, QMCKL_INVALID_ARG_2, fn, "Inconsistent values of type_nucl_vector (<0)");
    if (type_nucl_vector[i] >= c->type_nucl_num)
      return fail(context, QMCKL_INVALID_ARG_2, fn, "Inconsistent values of type_nucl_vector (>=type_nucl_num)");
  }
  c->type_nucl_vector.assign(type_nucl_vector, type_nucl_vector + Nn);
  return jas_post(c, context, 1u << 4);
}

qmckl_exit_code qmckl_set_jastrow_champ_a_vector(qmckl_context context, const double* a_vector, const int64_t size_max) {
  const char* fn = "qmckl_set_jastrow_champ_a_vector";
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  qmckl_exit_code rc = jas_pre(c, context, 1u << 5);
  if (rc != QMCKL_SUCCESS) return rc;
  if (c->aord < 0) return fail(context, QMCKL_NOT_PROVIDED, fn, "aord_num not set");
  if (c->type_nucl_num <= 0) return fail(context, QMCKL_NOT_PROVIDED, fn, "type_nucl_num not set");
  if (a_vector == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "a_vector = NULL");
  const int64_t n = (c->aord + 1) * c->type_nucl_num;
  if (size_max < n) return fail(context, QMCKL_INVALID_ARG_3, fn, "Array too small. Expected (aord_num+1)*type_nucl_num");
  c->a_vector.assign(a_vector, a_vector + n);
  return jas_post(c, context, 1u << 5);
}

qmckl_exit_code qmckl_set_jastrow_champ_b_vector(qmckl_context context, const double* b_vector, const int64_t size_max) {
  const char* fn = "qmckl_set_jastrow_champ_b_vector";
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  qmckl_exit_code rc = jas_pre(c, context, 1u << 6);
  if (rc != QMCKL_SUCCESS) return rc;
  if (c->bord < 0) return fail(context, QMCKL_NOT_PROVIDED, fn, "bord_num not set");
  if (b_vector == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "b_vector = NULL");
  if (size_max < c->bord + 1) return fail(context, QMCKL_INVALID_ARG_3, fn, "Array too small. Expected (bord_num+1)");
  c->b_vector.assign(b_vector, b_vector + c->bord + 1);
  return jas_post(c, context, 1u << 6);
}

qmckl_exit_code qmckl_set_jastrow_champ_c_vector(qmckl_context context, const double* c_vector, const int64_t size_max) {
  const char* fn = "qmckl_set_jastrow_champ_c_vector";
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  qmckl_exit_code rc = jas_pre(c, context, 1u << 7);
  if (rc != QMCKL_SUCCESS) return rc;
  if (c->type_nucl_num <= 0) return fail(context, QMCKL_NOT_PROVIDED, fn, "type_nucl_num not set");
  if (c->dim_c < 0) return fail(context, QMCKL_NOT_PROVIDED, fn, "cord_num not set");
  if (c_vector == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "c_vector = NULL");
  const int64_t n = c->dim_c * c->type_nucl_num;
  if (size_max < n) {
    char msg[256];
    snprintf(msg, sizeof msg, "Array too small. Expected dim_c_vector*type_nucl_num = %ld", (long)n);
    return fail(context, QMCKL_INVALID_ARG_3, fn, msg);
  }
  c->c_vector.assign(c_vector, c_vector + n);
  return jas_post(c, context, 1u << 7);
}

qmckl_exit_code qmckl_set_jastrow_champ_rescale_factor_ee(qmckl_context context, const double kappa_ee) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  qmckl_exit_code rc = jas_pre(c, context, 1u << 8);
  if (rc != QMCKL_SUCCESS) return rc;
  if (kappa_ee <= 0.0)
    return fail(context, QMCKL_INVALID_ARG_2, "qmckl_set_jastrow_champ_rescale_factor_ee", "rescale_factor_ee <= 0.0");
  c->rescale_ee = kappa_ee;
  return jas_post(c, context, 1u << 8);
}

qmckl_exit_code qmckl_set_jastrow_champ_rescale_factor_en(qmckl_context context, const double* kappa_en,
                                                          const int64_t size_max) {
  const char* fn = "qmckl_set_jastrow_champ_rescale_factor_en";
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  qmckl_exit_code rc = jas_pre(c, context, 1u << 9);
  if (rc != QMCKL_SUCCESS) return rc;
  if (c->type_nucl_num <= 0) return fail(context, QMCKL_NOT_PROVIDED, fn, "type_nucl_num not set");
  if (kappa_en == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "Null pointer");
  if (size_max < c->type_nucl_num) return fail(context, QMCKL_INVALID_ARG_3, fn, "Array too small. Expected type_nucl_num.");
  for (int64_t i = 0; i < c->type_nucl_num; ++i)
    if (kappa_en[i] <= 0.0) return fail(context, QMCKL_INVALID_ARG_2, fn, "rescale_factor_en <= 0.0");
  c->rescale_en.assign(kappa_en, kappa_en + c->type_nucl_num);
  return jas_post(c, context, 1u << 9);
}

qmckl_exit_code qmckl_set_jastrow_champ_spin_independent(qmckl_context context, const int32_t spin_independent) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  qmckl_exit_code rc = jas_pre(c, context, 1u << 10);
  if (rc != QMCKL_SUCCESS) return rc;
  c->spin_independent = spin_independent;
  return jas_post(c, context, 1u << 10);
}

/* ---- parameter getters, org/qmckl_jastrow_champ.org:1437-1901 */
#define JAS_GET_SCALAR(NAME, FIELD, BIT, T)                                                       \
  qmckl_exit_code NAME(qmckl_context context, T* const out) {                                     \
    Context* c = ctx_of(context);                                                                 \
    if (!c) return QMCKL_INVALID_CONTEXT;                                                         \
    if (out == nullptr) return fail(context, QMCKL_INVALID_ARG_2, #NAME, "NULL pointer");          \
    if (c->jas_uninit & (1u << BIT)) return fail(context, QMCKL_NOT_PROVIDED, #NAME, nullptr);     \
    *out = (T)c->FIELD;                                                                           \
    return QMCKL_SUCCESS;                                                                         \
  }
JAS_GET_SCALAR(qmckl_get_jastrow_champ_aord_num, aord, 0, int64_t)
JAS_GET_SCALAR(qmckl_get_jastrow_champ_bord_num, bord, 1, int64_t)
JAS_GET_SCALAR(qmckl_get_jastrow_champ_cord_num, cord, 2, int64_t)
JAS_GET_SCALAR(qmckl_get_jastrow_champ_type_nucl_num, type_nucl_num, 3, int64_t)
JAS_GET_SCALAR(qmckl_get_jastrow_champ_spin_independent, spin_independent, 10, int32_t)
#undef JAS_GET_SCALAR

qmckl_exit_code qmckl_get_jastrow_champ_rescale_factor_ee(const qmckl_context context, double* const out) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (out == nullptr) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_get_jastrow_champ_rescale_factor_ee", "NULL pointer");
  if (c->jas_uninit & (1u << 8)) return fail(context, QMCKL_NOT_PROVIDED, "qmckl_get_jastrow_champ_rescale_factor_ee", nullptr);
  *out = c->rescale_ee;
  return QMCKL_SUCCESS;
}

qmckl_exit_code qmckl_get_jastrow_champ_dim_c_vector(qmckl_context context, int64_t* const out) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (out == nullptr) return fail(context, QMCKL_INVALID_ARG_2, "qmckl_get_jastrow_champ_dim_c_vector", "NULL pointer");
  *out = c->dim_c;
  return QMCKL_SUCCESS;
}

static qmckl_exit_code jas_get_vec(qmckl_context context, uint32_t bit, const std::vector<double>& v, double* out,
                                   int64_t size_max, const char* fn) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  if (out == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "NULL pointer");
  if (c->jas_uninit & (1u << bit)) return fail(context, QMCKL_NOT_PROVIDED, fn, nullptr);
  if (size_max < (int64_t)v.size()) return fail(context, QMCKL_INVALID_ARG_3, fn, "Array too small");
  std::copy(v.begin(), v.end(), out);
  return QMCKL_SUCCESS;
}

qmckl_exit_code qmckl_get_jastrow_champ_a_vector(qmckl_context context, double* const a, const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  return jas_get_vec(context, 5, c->a_vector, a, size_max, "qmckl_get_jastrow_champ_a_vector");
}
qmckl_exit_code qmckl_get_jastrow_champ_b_vector(qmckl_context context, double* const b, const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  return jas_get_vec(context, 6, c->b_vector, b, size_max, "qmckl_get_jastrow_champ_b_vector");
}
qmckl_exit_code qmckl_get_jastrow_champ_c_vector(qmckl_context context, double* const cv, const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  return jas_get_vec(context, 7, c->c_vector, cv, size_max, "qmckl_get_jastrow_champ_c_vector");
}
qmckl_exit_code qmckl_get_jastrow_champ_rescale_factor_en(const qmckl_context context, double* const k, const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  return jas_get_vec(context, 9, c->rescale_en, k, size_max, "qmckl_get_jastrow_champ_rescale_factor_en");
}
qmckl_exit_code qmckl_get_jastrow_champ_type_nucl_vector(qmckl_context context, int64_t* const t, const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  const char* fn = "qmckl_get_jastrow_champ_type_nucl_vector";
  if (t == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "NULL pointer");
  if (c->jas_uninit & (1u << 4)) return fail(context, QMCKL_NOT_PROVIDED, fn, nullptr);
  if (size_max < (int64_t)c->type_nucl_vector.size()) return fail(context, QMCKL_INVALID_ARG_3, fn, "Array too small");
  std::copy(c->type_nucl_vector.begin(), c->type_nucl_vector.end(), t);
  return QMCKL_SUCCESS;
}

bool qmckl_jastrow_champ_provided(const qmckl_context context) {
  Context* c = ctx_of(context);
  return c ? c->jas_provided : false;
}

// org/qmckl_jastrow_champ.org:2143-2185
qmckl_exit_code qmckl_get_jastrow_champ_asymp_jasb(qmckl_context context, double* const asymp_jasb, const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  const char* fn = "qmckl_get_jastrow_champ_asymp_jasb";
  if (!c->jas_provided) return fail(context, QMCKL_NOT_PROVIDED, fn, nullptr);
  if (asymp_jasb == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "NULL pointer");
  if (size_max < 2) return fail(context, QMCKL_INVALID_ARG_3, fn, "Array too small. Expected 2");
  asymp_jasb[0] = c->asymp_jasb[0];
  asymp_jasb[1] = c->asymp_jasb[1];
  return QMCKL_SUCCESS;
}

// org/qmckl_jastrow_champ.org:5677-5720
qmckl_exit_code qmckl_get_jastrow_champ_asymp_jasa(qmckl_context context, double* const asymp_jasa, const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  const char* fn = "qmckl_get_jastrow_champ_asymp_jasa";
  if (!c->jas_provided) return fail(context, QMCKL_NOT_PROVIDED, fn, nullptr);
  if (asymp_jasa == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "NULL pointer");
  if (size_max < c->type_nucl_num) return fail(context, QMCKL_INVALID_ARG_3, fn, "Array too small. Expected type_nucl_num");
  std::copy(c->asymp_jasa.begin(), c->asymp_jasa.end(), asymp_jasa);
  return QMCKL_SUCCESS;
}

/* ---- intermediates (materialised on demand) */
qmckl_exit_code qmckl_get_jastrow_champ_ee_distance_rescaled(qmckl_context ctx, double* const out, int64_t const n) {
  return get_inter(ctx, J_EE_RESCALED, out, n, "qmckl_get_jastrow_champ_ee_distance_rescaled");
}
qmckl_exit_code qmckl_get_jastrow_champ_ee_distance_rescaled_gl(qmckl_context ctx, double* const out, const int64_t n) {
  return get_inter(ctx, J_EE_RESCALED_GL, out, n, "qmckl_get_jastrow_champ_ee_distance_rescaled_gl");
}
qmckl_exit_code qmckl_get_jastrow_champ_en_distance_rescaled(qmckl_context ctx, double* const out, const int64_t n) {
  return get_inter(ctx, J_EN_RESCALED, out, n, "qmckl_get_jastrow_champ_en_distance_rescaled");
}
qmckl_exit_code qmckl_get_jastrow_champ_en_distance_rescaled_gl(qmckl_context ctx, double* const out, const int64_t n) {
  return get_inter(ctx, J_EN_RESCALED_GL, out, n, "qmckl_get_jastrow_champ_en_distance_rescaled_gl");
}
qmckl_exit_code qmckl_get_jastrow_champ_een_rescaled_e(qmckl_context ctx, double* const out, const int64_t n) {
  return get_inter(ctx, J_EEN_E, out, n, "qmckl_get_jastrow_champ_een_rescaled_e");
}
qmckl_exit_code qmckl_get_jastrow_champ_een_rescaled_e_gl(qmckl_context ctx, double* const out, const int64_t n) {
  return get_inter(ctx, J_EEN_E_GL, out, n, "qmckl_get_jastrow_champ_een_rescaled_e_gl");
}
qmckl_exit_code qmckl_get_jastrow_champ_een_rescaled_n(qmckl_context ctx, double* const out, const int64_t n) {
  return get_inter(ctx, J_EEN_N, out, n, "qmckl_get_jastrow_champ_een_rescaled_n");
}
qmckl_exit_code qmckl_get_jastrow_champ_een_rescaled_n_gl(qmckl_context ctx, double* const out, const int64_t n) {
  return get_inter(ctx, J_EEN_N_GL, out, n, "qmckl_get_jastrow_champ_een_rescaled_n_gl");
}
qmckl_exit_code qmckl_get_jastrow_champ_tmp_c(qmckl_context ctx, double* const out, const int64_t n) {
  return get_inter(ctx, J_TMP_C, out, n, "qmckl_get_jastrow_champ_tmp_c");
}
qmckl_exit_code qmckl_get_jastrow_champ_dtmp_c(qmckl_context ctx, double* const out, const int64_t n) {
  return get_inter(ctx, J_DTMP_C, out, n, "qmckl_get_jastrow_champ_dtmp_c");
}

/* ---- results, org/qmckl_jastrow_champ.org:3405, 4013, 6839, 7323, 13836, 14394, 16409, 16767 */
qmckl_exit_code qmckl_get_jastrow_champ_factor_ee(qmckl_context ctx, double* const out, const int64_t n) {
  return get_result(ctx, F_EE, out, n, "qmckl_get_jastrow_champ_factor_ee");
}
qmckl_exit_code qmckl_get_jastrow_champ_factor_ee_gl(qmckl_context ctx, double* const out, const int64_t n) {
  return get_result(ctx, F_EE_GL, out, n, "qmckl_get_jastrow_champ_factor_ee_gl");
}
qmckl_exit_code qmckl_get_jastrow_champ_factor_en(qmckl_context ctx, double* const out, const int64_t n) {
  return get_result(ctx, F_EN, out, n, "qmckl_get_jastrow_champ_factor_en");
}
qmckl_exit_code qmckl_get_jastrow_champ_factor_en_gl(qmckl_context ctx, double* const out, const int64_t n) {
  return get_result(ctx, F_EN_GL, out, n, "qmckl_get_jastrow_champ_factor_en_gl");
}
qmckl_exit_code qmckl_get_jastrow_champ_factor_een(qmckl_context ctx, double* const out, const int64_t n) {
  return get_result(ctx, F_EEN, out, n, "qmckl_get_jastrow_champ_factor_een");
}
qmckl_exit_code qmckl_get_jastrow_champ_factor_een_gl(qmckl_context ctx, double* const out, const int64_t n) {
  return get_result(ctx, F_EEN_GL, out, n, "qmckl_get_jastrow_champ_factor_een_gl");
}
qmckl_exit_code qmckl_get_jastrow_champ_value(qmckl_context ctx, double* const out, const int64_t n) {
  return get_result(ctx, J_VALUE, out, n, "qmckl_get_jastrow_champ_value");
}
qmckl_exit_code qmckl_get_jastrow_champ_gl(qmckl_context ctx, double* const out, const int64_t n) {
  return get_result(ctx, J_GL, out, n, "qmckl_get_jastrow_champ_gl");
}

// org/qmckl_jastrow_champ.org:16823-16880 and :17314-17350: grad[walk][3][elec] = e^J grad J
qmckl_exit_code qmckl_get_jastrow_champ_grad(qmckl_context context, double* const grad, const int64_t size_max) {
  const char* fn = "qmckl_get_jastrow_champ_grad";
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (grad == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "NULL pointer");
  if (!c->jas_provided) return fail(context, QMCKL_NOT_PROVIDED, fn, nullptr);
  const int64_t W = c->walk_num, Ne = c->elec_num;
  if (size_max < W * 3 * Ne) return fail(context, QMCKL_INVALID_ARG_3, fn, "Array too small");
  qmckl_exit_code rc = provide_jastrow(context, c, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  // The first three rows of each walker's GL block are the gradient.
  hipError_t e = hipMemcpy2DAsync(grad, sizeof(double) * 3 * Ne, c->d_jgl.ptr, sizeof(double) * 4 * Ne,
                                  sizeof(double) * 3 * Ne, W, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_fail(context, e, fn);
  return QMCKL_SUCCESS;
}

// org/qmckl_jastrow_champ.org:14403, 14448-14475: factor_een_grad[walk][3][elec], the
// gradient rows of factor_een_gl (the same e-e-n sums, without the Laplacian)
qmckl_exit_code qmckl_get_jastrow_champ_factor_een_grad(qmckl_context context, double* const out,
                                                        const int64_t size_max) {
  const char* fn = "qmckl_get_jastrow_champ_factor_een_grad";
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (out == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "Null pointer");
  if (!c->jas_provided) return fail(context, QMCKL_NOT_PROVIDED, fn, nullptr);
  const int64_t W = c->walk_num, Ne = c->elec_num;
  if (size_max < W * 3 * Ne) return fail(context, QMCKL_INVALID_ARG_3, fn, "Array too small. Expected 3*walk_num*elec_num");
  qmckl_exit_code rc = provide_jastrow(context, c, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  hipError_t e = hipMemcpy2DAsync(out, sizeof(double) * 3 * Ne, c->d_feen_gl.ptr, sizeof(double) * 4 * Ne,
                                  sizeof(double) * 3 * Ne, W, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_fail(context, e, fn);
  return QMCKL_SUCCESS;
}

// Gradient of L = sum_w [wJ ln Psi_J + wL E_L^surr] (E_L^surr = -1/2 sum_i (lap_i J +
// |grad_i J|^2) + V_ee + V_en + V_nn) in the Jastrow parameters and nuclear coordinates, by
// the adjoint kernel of kernels_adjoint.hip on the forward pass's factor GLs.  No
// counterpart in the reference.
qmckl_exit_code qmckl_amd_get_jastrow_champ_adjoint(qmckl_context context, const double weight_j,
                                                    const double weight_el, double* const value,
                                                    double* const grad_a, const int64_t size_a,
                                                    double* const grad_b, const int64_t size_b,
                                                    double* const grad_c, const int64_t size_c,
                                                    double* const grad_nucl, const int64_t size_nucl) {
  const char* fn = "qmckl_amd_get_jastrow_champ_adjoint";
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (!c->jas_provided) return fail(context, QMCKL_NOT_PROVIDED, fn, nullptr);
  const int64_t T = c->type_nucl_num, Nn = c->nucl_num, W = c->walk_num;
  const int64_t na = T * (c->aord + 1), nb = c->bord + 1, nc = c->cord >= 2 ? T * c->dim_c : 0;
  if (value == nullptr) return fail(context, QMCKL_INVALID_ARG_4, fn, "NULL pointer");
  if (grad_a == nullptr || size_a < na) return fail(context, QMCKL_INVALID_ARG_6, fn, "grad_a: type_nucl_num*(aord_num+1)");
  if (grad_b == nullptr || size_b < nb) return fail(context, QMCKL_INVALID_ARG_8, fn, "grad_b: bord_num+1");
  if ((nc > 0 && grad_c == nullptr) || size_c < nc)
    return fail(context, QMCKL_INVALID_ARG_10, fn, "grad_c: type_nucl_num*dim_c_vector");
  if (grad_nucl == nullptr || size_nucl < 3 * Nn) return fail(context, QMCKL_INVALID_ARG_12, fn, "grad_nucl: 3*nucl_num");
  if (c->aord < 1 || c->bord < 1 || c->aord > 8 || c->bord > 8)
    return fail(context, QMCKL_FAILURE, fn, "the adjoint needs 1 <= aord_num, bord_num <= 8");
  qmckl_exit_code rc = provide_jastrow(context, c, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  AdjArgs a;
  a.v = jas_view(c);
  a.ntype = (int)T;
  a.fee = c->d_fee.as<double>();
  a.fen = c->d_fen.as<double>();
  a.feen = c->d_feen.as<double>();
  a.fee_gl = c->d_fee_gl.as<double>();
  a.fen_gl = c->d_fen_gl.as<double>();
  a.feen_gl = c->d_feen_gl.as<double>();
  a.charge = c->dc.jas_charge.as<double>();
  a.wJ = weight_j;
  a.wL = weight_el;
  a.nout = adjoint_outputs(a.v, a.ntype);
  hipError_t e = c->d_deriv[DB_PART].reserve(sizeof(double) * W * a.nout);
  if (e == hipSuccess) e = c->d_deriv[DB_OUT].reserve(sizeof(double) * a.nout);
  if (e != hipSuccess) return hip_fail(context, e, fn);
  c->deriv_date[DB_PART] = c->deriv_date[DB_OUT] = 0;  // no longer the derivative getters' results
  a.out = c->d_deriv[DB_PART].as<double>();
  e = launch_jastrow_adjoint(a, c->stream);
  if (e == hipSuccess)
    e = launch_sum_walkers(a.out, W, a.nout, 1.0, c->d_deriv[DB_OUT].as<double>(), c->stream);
  std::vector<double> h(a.nout);
  if (e == hipSuccess) e = copy_to_host(c, h.data(), c->d_deriv[DB_OUT].ptr, sizeof(double) * a.nout);
  if (e != hipSuccess) return hip_fail(context, e, fn);
  // V_nn = sum_{a<b} Z_a Z_b / R_ab (org/qmckl_nucleus.org:1040-1076), once per walker
  double vnn = 0.0;
  std::vector<double> dvnn(3 * Nn, 0.0);
  for (int64_t p = 0; p < Nn; ++p)
    for (int64_t q = 0; q < p; ++q) {
      double d[3], r2 = 0.0;
      for (int k = 0; k < 3; ++k) {
        d[k] = c->nucl_coord[k * Nn + p] - c->nucl_coord[k * Nn + q];
        r2 += d[k] * d[k];
      }
      const double r = std::sqrt(r2), zz = c->nucl_charge[p] * c->nucl_charge[q];
      vnn += zz / r;
      for (int k = 0; k < 3; ++k) {
        dvnn[3 * p + k] -= zz * d[k] / (r2 * r);
        dvnn[3 * q + k] += zz * d[k] / (r2 * r);
      }
    }
  const double wn = weight_el * (double)W;
  *value = h[0] + wn * vnn;
  int64_t o = 1;
  for (int64_t k = 0; k < na; ++k) grad_a[k] = h[o++];
  for (int64_t k = 0; k < nb; ++k) grad_b[k] = h[o++];
  for (int64_t k = 0; k < nc; ++k) grad_c[k] = h[o++];
  for (int64_t k = 0; k < 3 * Nn; ++k) grad_nucl[k] = h[o++] + wn * dvnn[k];
  return QMCKL_SUCCESS;
}

/* ---- device-resident results */
qmckl_exit_code qmckl_amd_get_jastrow_champ_value_device(qmckl_context context, double* const d_value, const int64_t size_max) {
  const char* fn = "qmckl_amd_get_jastrow_champ_value_device";
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (d_value == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "NULL pointer");
  if (!c->jas_provided) return fail(context, QMCKL_NOT_PROVIDED, fn, nullptr);
  if (size_max < c->walk_num) return fail(context, QMCKL_INVALID_ARG_3, fn, "Array too small");
  qmckl_exit_code rc = provide_jastrow_value(context, c, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  hipError_t e = hipMemcpyAsync(d_value, c->d_jvalue.ptr, sizeof(double) * c->walk_num, hipMemcpyDeviceToDevice, c->stream);
  if (e != hipSuccess) return hip_fail(context, e, fn);
  return QMCKL_SUCCESS;
}

qmckl_exit_code qmckl_amd_get_jastrow_champ_gl_device(qmckl_context context, double* const d_gl, const int64_t size_max) {
  const char* fn = "qmckl_amd_get_jastrow_champ_gl_device";
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (d_gl == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "NULL pointer");
  if (!c->jas_provided) return fail(context, QMCKL_NOT_PROVIDED, fn, nullptr);
  const int64_t n = c->walk_num * 4 * c->elec_num;
  if (size_max < n) return fail(context, QMCKL_INVALID_ARG_3, fn, "Array too small");
  qmckl_exit_code rc = provide_jastrow(context, c, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  hipError_t e = hipMemcpyAsync(d_gl, c->d_jgl.ptr, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream);
  if (e != hipSuccess) return hip_fail(context, e, fn);
  return QMCKL_SUCCESS;
}


/* ---- parameter derivatives (org/qmckl_jastrow_champ.org) and nuclear forces
 *      (org/qmckl_forces.org) */
// org/qmckl_jastrow_champ.org:4807-4850
qmckl_exit_code qmckl_get_jastrow_champ_asymp_jasb_pderiv(qmckl_context context, double* const out,
                                                          const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  const char* fn = "qmckl_get_jastrow_champ_asymp_jasb_pderiv";
  if (!c->jas_provided) return fail(context, QMCKL_NOT_PROVIDED, fn, nullptr);
  if (out == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "NULL pointer");
  const std::vector<double> v = asymp_jasb_pderiv_of(c);
  if (size_max < (int64_t)v.size()) return fail(context, QMCKL_INVALID_ARG_3, fn, "Array too small. Expected (bord_num+1)*2");
  std::copy(v.begin(), v.end(), out);
  return QMCKL_SUCCESS;
}
// org/qmckl_jastrow_champ.org:8041-8090
qmckl_exit_code qmckl_get_jastrow_champ_asymp_jasa_pderiv(qmckl_context context, double* const out,
                                                          const int64_t size_max) {
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  const char* fn = "qmckl_get_jastrow_champ_asymp_jasa_pderiv";
  if (!c->jas_provided) return fail(context, QMCKL_NOT_PROVIDED, fn, nullptr);
  if (out == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "NULL pointer");
  const std::vector<double> v = asymp_jasa_pderiv_of(c);
  if (size_max < (int64_t)v.size()) return fail(context, QMCKL_INVALID_ARG_3, fn, "Array too small");
  std::copy(v.begin(), v.end(), out);
  return QMCKL_SUCCESS;
}
// org/qmckl_jastrow_champ.org:5031-5075
qmckl_exit_code qmckl_get_jastrow_champ_factor_ee_pderiv(qmckl_context ctx, double* const out, const int64_t n) {
  return get_deriv(ctx, DV_EE_PDERIV, out, n, "qmckl_get_jastrow_champ_factor_ee_pderiv");
}
// org/qmckl_jastrow_champ.org:5332-5376
qmckl_exit_code qmckl_get_jastrow_champ_factor_ee_gl_pderiv(qmckl_context ctx, double* const out, const int64_t n) {
  return get_deriv(ctx, DV_EE_GL_PDERIV, out, n, "qmckl_get_jastrow_champ_factor_ee_gl_pderiv");
}
// org/qmckl_jastrow_champ.org:8269-8316
qmckl_exit_code qmckl_get_jastrow_champ_factor_en_pderiv(qmckl_context ctx, double* const out, const int64_t n) {
  return get_deriv(ctx, DV_EN_PDERIV, out, n, "qmckl_get_jastrow_champ_factor_en_pderiv");
}
// org/qmckl_jastrow_champ.org:8573-8617
qmckl_exit_code qmckl_get_jastrow_champ_factor_en_gl_pderiv(qmckl_context ctx, double* const out, const int64_t n) {
  return get_deriv(ctx, DV_EN_GL_PDERIV, out, n, "qmckl_get_jastrow_champ_factor_en_gl_pderiv");
}
// org/qmckl_jastrow_champ.org:15751-15795
qmckl_exit_code qmckl_get_jastrow_champ_factor_een_pderiv(qmckl_context ctx, double* const out, const int64_t n) {
  return get_deriv(ctx, DV_EEN_PDERIV, out, n, "qmckl_get_jastrow_champ_factor_een_pderiv");
}
// org/qmckl_jastrow_champ.org:16055-16100
qmckl_exit_code qmckl_get_jastrow_champ_factor_een_gl_pderiv(qmckl_context ctx, double* const out, const int64_t n) {
  return get_deriv(ctx, DV_EEN_GL_PDERIV, out, n, "qmckl_get_jastrow_champ_factor_een_gl_pderiv");
}
// org/qmckl_forces.org:566-612
qmckl_exit_code qmckl_get_forces_jastrow_en(qmckl_context ctx, double* const out, const int64_t n) {
  return get_deriv(ctx, DV_FORCES_EN, out, n, "qmckl_get_forces_jastrow_en");
}
// org/qmckl_forces.org:947-993
qmckl_exit_code qmckl_get_forces_jastrow_en_g(qmckl_context ctx, double* const out, const int64_t n) {
  return get_deriv(ctx, DV_FORCES_EN_G, out, n, "qmckl_get_forces_jastrow_en_g");
}
// org/qmckl_forces.org:1372-1418
qmckl_exit_code qmckl_get_forces_jastrow_en_l(qmckl_context ctx, double* const out, const int64_t n) {
  return get_deriv(ctx, DV_FORCES_EN_L, out, n, "qmckl_get_forces_jastrow_en_l");
}
// org/qmckl_forces.org:2984-3030
qmckl_exit_code qmckl_get_forces_jastrow_een(qmckl_context ctx, double* const out, const int64_t n) {
  return get_deriv(ctx, DV_FORCES_EEN, out, n, "qmckl_get_forces_jastrow_een");
}
// org/qmckl_forces.org:3715-3760
qmckl_exit_code qmckl_get_forces_jastrow_een_g(qmckl_context ctx, double* const out, const int64_t n) {
  return get_deriv(ctx, DV_FORCES_EEN_G, out, n, "qmckl_get_forces_jastrow_een_g");
}
// org/qmckl_forces.org:4079-4125
qmckl_exit_code qmckl_get_forces_jastrow_een_l(qmckl_context ctx, double* const out, const int64_t n) {
  return get_deriv(ctx, DV_FORCES_EEN_L, out, n, "qmckl_get_forces_jastrow_een_l");
}

}  // extern "C"
