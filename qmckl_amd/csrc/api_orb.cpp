// Orbital Hessians, the forces on the orbitals and the MO cusp correction: the C-ABI of
//   qmckl_get_ao_basis_shell_hessian, qmckl_get_ao_basis_ao_hessian (org/qmckl_ao.org:5257-5435,
//   9397-9556), qmckl_get_mo_basis_mo_hessian (org/qmckl_mo.org:2362-2536),
//   qmckl_set_mo_basis_r_cusp (org/qmckl_mo.org:479-700),
//   qmckl_get_forces_ao_value, _mo_value(_inplace), _mo_g(_inplace), _mo_l
//   (org/qmckl_forces.org:5122-6878)
// over the kernels of kernels_orb.hip.
//
// The forces on the MOs are contractions of the AO gradients or Hessians of one nucleus' AOs
// with the MO coefficients (one nucleus block per output block, sign -1); the MO Hessian is the
// same contraction over all AOs.  With r_cusp set, every MO value/VGL/Hessian getter applies
// the cusp correction: at a point closer than r_cusp[A] to nucleus A the s AOs of A are taken
// out and the fitted cubic in r added (the reference's _cusp computations).  One departure:
// the reference leaves the s AOs out for r <= r_cusp in the VGL and for r < r_cusp in the
// values, and adds the cubic for r < r_cusp; here both use r < r_cusp.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.hpp"

using namespace qmckl_amd;

namespace {

// First AO of each nucleus [Nn + 1], then 0 and ao_num (the whole basis as one block), on the
// device.  AOs are numbered nucleus by nucleus (org/qmckl_ao.org:8470-8480).
qmckl_exit_code nucl_ao_ready(Context* c, qmckl_context h, const char* fn) {
  if (c->nucl_ao_date == c->const_date && c->d_nucl_ao.ptr) return QMCKL_SUCCESS;
  const AoTables& t = c->tables;
  const int64_t Nn = c->nucl_num;
  std::vector<int64_t> ks(Nn + 3, 0);
  for (int64_t k = 0; k < c->ao_num; ++k) ks[t.shell_nucl[t.ao_shell[k]] + 1] += 1;
  for (int64_t a = 0; a < Nn; ++a) ks[a + 1] += ks[a];
  ks[Nn + 1] = 0;
  ks[Nn + 2] = c->ao_num;
  for (int64_t k = 0; k < c->ao_num; ++k) {
    const int a = t.shell_nucl[t.ao_shell[k]];
    if (k < ks[a] || k >= ks[a + 1]) return fail(h, QMCKL_FAILURE, fn, "the AOs of a nucleus are not contiguous");
  }
  hipError_t e = c->d_nucl_ao.reserve(sizeof(int64_t) * ks.size());
  if (e == hipSuccess)
    e = hipMemcpyAsync(c->d_nucl_ao.ptr, ks.data(), sizeof(int64_t) * ks.size(), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_fail(h, e, fn);
  c->nucl_ao_date = c->const_date;
  return QMCKL_SUCCESS;
}

qmckl_exit_code provide_ao_hessian(Context* c, qmckl_context h, const char* fn) {
  qmckl_exit_code rc = ao_ready(c, h, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  if (c->tables.slater) return fail(h, QMCKL_FAILURE, fn, "Hessians of Slater shells are not implemented");
  if (c->ao_hess_date == c->date) return QMCKL_SUCCESS;
  const int64_t P = c->point_num;
  hipError_t e = c->d_ao_hess.reserve(sizeof(double) * 12 * P * c->ao_num);
  if (e == hipSuccess) e = launch_ao_hessian(ao_view(c), c->d_coord.as<double>(), P, c->d_ao_hess.as<double>(), c->stream);
  if (e != hipSuccess) return hip_fail(h, e, fn);
  c->ao_hess_date = c->date;
  return QMCKL_SUCCESS;
}

qmckl_exit_code copy_out(Context* c, qmckl_context h, void* dst, const void* src, size_t bytes, const char* fn) {
  const hipError_t e = copy_to_host(c, dst, src, bytes);
  if (e != hipSuccess) return hip_fail(h, e, fn);
  return QMCKL_SUCCESS;
}

// Common checks of the getters: live context, non-null output, size_max >= n.
Context* getter_ctx(qmckl_context context, const double* out, int64_t size_max, int64_t n, const char* fn,
                    qmckl_exit_code& rc) {
  rc = QMCKL_SUCCESS;
  Context* c = ctx_of(context);
  if (!c) {
    rc = QMCKL_INVALID_CONTEXT;
    return nullptr;
  }
  if (out == nullptr) rc = fail(context, QMCKL_INVALID_ARG_2, fn, "NULL pointer");
  else if (size_max < n) rc = fail(context, QMCKL_INVALID_ARG_3, fn, "Array too small");
  return rc == QMCKL_SUCCESS ? c : nullptr;
}

// Contraction of the device array src with the FP64 C^T over nucleus blocks (na = Nn) or the
// whole basis (na = 1), into c->d_orb_out.
qmckl_exit_code contract(Context* c, qmckl_context h, const char* fn, const double* src, int64_t base, int64_t U,
                         int64_t su, int64_t V, int64_t sv, int64_t oa, int64_t ou, int64_t ov, bool per_nucleus,
                         double sign, int64_t n_out) {
  qmckl_exit_code rc = nucl_ao_ready(c, h, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  hipError_t e = c->d_orb_out.reserve(sizeof(double) * std::max<int64_t>(n_out, 1));
  if (e != hipSuccess) return hip_fail(h, e, fn);
  ContractArgs a;
  a.src = src;
  a.base = base;
  a.U = U;
  a.su = su;
  a.V = V;
  a.sv = sv;
  a.ct = c->dc.coef_t.as<double>();
  a.ldct = c->dc.mo_pad;
  a.mo = c->mo_num;
  a.k_start = c->d_nucl_ao.as<int64_t>() + (per_nucleus ? 0 : c->nucl_num + 1);
  a.na = per_nucleus ? c->nucl_num : 1;
  a.out = c->d_orb_out.as<double>();
  a.oa = oa;
  a.ou = ou;
  a.ov = ov;
  a.sign = sign;
  e = launch_orb_contract(a, c->stream);
  if (e != hipSuccess) return hip_fail(h, e, fn);
  return QMCKL_SUCCESS;
}

// The cusp tables on the device: r_cusp [Nn], cusp_param [Nn][4][mo] (doubles), then
// s_start [Nn + 1] and s_ao (int32): the s AOs of each nucleus.
qmckl_exit_code cusp_upload(Context* c, qmckl_context h, const char* fn) {
  const AoTables& t = c->tables;
  const int64_t Nn = c->nucl_num;
  std::vector<int32_t> start(Nn + 1, 0), sao;
  for (int64_t a = 0; a < Nn; ++a) {
    for (int64_t k = 0; k < c->ao_num; ++k)
      if (t.shell_nucl[t.ao_shell[k]] == a && t.shell_l[t.ao_shell[k]] == 0) sao.push_back((int32_t)k);
    start[a + 1] = (int32_t)sao.size();
    if (start[a + 1] - start[a] > cusp_max_s_ao())
      return fail(h, QMCKL_FAILURE, fn, "more s AOs on one nucleus than the cusp kernel stages");
  }
  const size_t nd = c->r_cusp.size() + c->cusp_param.size();
  const size_t bytes = sizeof(double) * nd + sizeof(int32_t) * (start.size() + sao.size() + 1);
  hipError_t e = c->d_cusp.reserve(bytes);
  char* d = static_cast<char*>(c->d_cusp.ptr);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d, c->r_cusp.data(), sizeof(double) * c->r_cusp.size(), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d + sizeof(double) * c->r_cusp.size(), c->cusp_param.data(),
                       sizeof(double) * c->cusp_param.size(), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d + sizeof(double) * nd, start.data(), sizeof(int32_t) * start.size(), hipMemcpyHostToDevice,
                       c->stream);
  if (e == hipSuccess && !sao.empty())
    e = hipMemcpyAsync(d + sizeof(double) * nd + sizeof(int32_t) * start.size(), sao.data(),
                       sizeof(int32_t) * sao.size(), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_fail(h, e, fn);
  c->cusp_s_total = (int64_t)sao.size();
  c->cusp_uploaded = true;
  return QMCKL_SUCCESS;
}

}  // namespace

namespace qmckl_amd {

// The cusp correction of MO results on the device: out is [P][5][mo] (hessian = 0) or
// [3][P][3][mo] (hessian = 1) for the context's points.
qmckl_exit_code apply_mo_cusp(Context* c, qmckl_context h, double* out, int hessian, const char* fn) {
  if (c->r_cusp.empty()) return QMCKL_SUCCESS;
  if ((int64_t)c->cusp_param.size() != c->nucl_num * 4 * c->mo_num)
    return fail(h, QMCKL_FAILURE, fn, "cusp parameters do not match the MOs");
  if (!c->cusp_uploaded) {
    qmckl_exit_code rc = cusp_upload(c, h, fn);
    if (rc != QMCKL_SUCCESS) return rc;
  }
  const int64_t Nn = c->nucl_num;
  const size_t nd = c->r_cusp.size() + c->cusp_param.size();
  CuspArgs a;
  a.coord = c->d_coord.as<double>();
  a.P = c->point_num;
  a.mo = c->mo_num;
  a.nucl_num = Nn;
  a.r_cusp = c->d_cusp.as<double>();
  a.param = a.r_cusp + Nn;
  a.s_start = reinterpret_cast<const int32_t*>(c->d_cusp.as<double>() + nd);
  a.s_ao = a.s_start + Nn + 1;
  a.ct = c->dc.coef_t.as<double>();
  a.ldct = c->dc.mo_pad;
  a.v = ao_view(c);
  a.out = out;
  a.hessian = hessian;
  const hipError_t e = launch_mo_cusp(a, c->stream);
  if (e != hipSuccess) return hip_fail(h, e, fn);
  return QMCKL_SUCCESS;
}

}  // namespace qmckl_amd

extern "C" {

// org/qmckl_ao.org:5268-5305: [point][3][4][shell]
qmckl_exit_code qmckl_get_ao_basis_shell_hessian(qmckl_context context, double* const shell_hessian,
                                                 const int64_t size_max) {
  const char* fn = "qmckl_get_ao_basis_shell_hessian";
  qmckl_exit_code rc;
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  const int64_t n = c->point_num * 12 * c->shell_num;
  if (!getter_ctx(context, shell_hessian, size_max, n, fn, rc)) return rc;
  rc = ao_ready(c, context, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  if (c->tables.slater) return fail(context, QMCKL_FAILURE, fn, "Hessians of Slater shells are not implemented");
  hipError_t e = c->d_orb_out.reserve(sizeof(double) * std::max<int64_t>(n, 1));
  if (e == hipSuccess)
    e = launch_shell_hessian(ao_view(c), c->d_coord.as<double>(), c->point_num, c->d_orb_out.as<double>(), c->stream);
  if (e != hipSuccess) return hip_fail(context, e, fn);
  return copy_out(c, context, shell_hessian, c->d_orb_out.ptr, sizeof(double) * n, fn);
}

// org/qmckl_ao.org:9397-9435: [3][point][4][ao]
qmckl_exit_code qmckl_get_ao_basis_ao_hessian(qmckl_context context, double* const ao_hessian, const int64_t size_max) {
  const char* fn = "qmckl_get_ao_basis_ao_hessian";
  qmckl_exit_code rc;
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  const int64_t n = 12 * c->point_num * c->ao_num;
  if (!getter_ctx(context, ao_hessian, size_max, n, fn, rc)) return rc;
  rc = provide_ao_hessian(c, context, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  return copy_out(c, context, ao_hessian, c->d_ao_hess.ptr, sizeof(double) * n, fn);
}

// org/qmckl_mo.org:2366-2420: [3][point][3][mo]
qmckl_exit_code qmckl_get_mo_basis_mo_hessian(qmckl_context context, double* const mo_hessian, const int64_t size_max) {
  const char* fn = "qmckl_get_mo_basis_mo_hessian";
  qmckl_exit_code rc;
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (!c->mo_provided) return fail(context, QMCKL_NOT_PROVIDED, fn, nullptr);
  const int64_t P = c->point_num, ao = c->ao_num, mo = c->mo_num;
  const int64_t n = 9 * P * mo;
  if (!getter_ctx(context, mo_hessian, size_max, n, fn, rc)) return rc;
  rc = provide_ao_hessian(c, context, fn);
  if (rc == QMCKL_SUCCESS)
    rc = contract(c, context, fn, c->d_ao_hess.as<double>(), 0, 3 * P, 4 * ao, 3, ao, 0, 3 * mo, mo, false, 1.0, n);
  if (rc == QMCKL_SUCCESS) rc = apply_mo_cusp(c, context, c->d_orb_out.as<double>(), 1, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  return copy_out(c, context, mo_hessian, c->d_orb_out.ptr, sizeof(double) * n, fn);
}

// org/qmckl_forces.org:5127-5198: [nucl][point][3][ao]
qmckl_exit_code qmckl_get_forces_ao_value(qmckl_context context, double* const forces_ao_value, const int64_t size_max) {
  const char* fn = "qmckl_get_forces_ao_value";
  qmckl_exit_code rc;
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  const int64_t n = c->nucl_num * c->point_num * 3 * c->ao_num;
  if (!getter_ctx(context, forces_ao_value, size_max, n, fn, rc)) return rc;
  rc = provide_ao_vgl(context, c, fn);
  if (rc != QMCKL_SUCCESS) return rc;
  hipError_t e = c->d_orb_out.reserve(sizeof(double) * std::max<int64_t>(n, 1));
  if (e == hipSuccess)
    e = launch_forces_ao_value(ao_view(c), c->d_ao_vgl.as<double>(), c->point_num, c->d_orb_out.as<double>(), c->stream);
  if (e != hipSuccess) return hip_fail(context, e, fn);
  return copy_out(c, context, forces_ao_value, c->d_orb_out.ptr, sizeof(double) * n, fn);
}

// org/qmckl_forces.org:5473-5530: [nucl][point][3][mo]
qmckl_exit_code qmckl_get_forces_mo_value(qmckl_context context, double* const forces_mo_value, const int64_t size_max) {
  const char* fn = "qmckl_get_forces_mo_value";
  qmckl_exit_code rc;
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (!c->mo_provided) return fail(context, QMCKL_NOT_PROVIDED, fn, nullptr);
  const int64_t P = c->point_num, ao = c->ao_num, mo = c->mo_num;
  const int64_t n = c->nucl_num * P * 3 * mo;
  if (!getter_ctx(context, forces_mo_value, size_max, n, fn, rc)) return rc;
  rc = provide_ao_vgl(context, c, fn);
  if (rc == QMCKL_SUCCESS)
    rc = contract(c, context, fn, c->d_ao_vgl.as<double>(), ao, P, 5 * ao, 3, ao, P * 3 * mo, 3 * mo, mo, true, -1.0, n);
  if (rc != QMCKL_SUCCESS) return rc;
  return copy_out(c, context, forces_mo_value, c->d_orb_out.ptr, sizeof(double) * n, fn);
}

qmckl_exit_code qmckl_get_forces_mo_value_inplace(qmckl_context context, double* const forces_mo_value,
                                                  const int64_t size_max) {
  return qmckl_get_forces_mo_value(context, forces_mo_value, size_max);
}

// org/qmckl_forces.org:5917-5935: [nucl][3][point][3][mo]
qmckl_exit_code qmckl_get_forces_mo_g(qmckl_context context, double* const forces_mo_g, const int64_t size_max) {
  const char* fn = "qmckl_get_forces_mo_g";
  qmckl_exit_code rc;
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (!c->mo_provided) return fail(context, QMCKL_NOT_PROVIDED, fn, nullptr);
  const int64_t P = c->point_num, ao = c->ao_num, mo = c->mo_num;
  const int64_t n = c->nucl_num * 9 * P * mo;
  if (!getter_ctx(context, forces_mo_g, size_max, n, fn, rc)) return rc;
  rc = provide_ao_hessian(c, context, fn);
  if (rc == QMCKL_SUCCESS)
    rc = contract(c, context, fn, c->d_ao_hess.as<double>(), 0, 3 * P, 4 * ao, 3, ao, 9 * P * mo, 3 * mo, mo, true, -1.0,
                  n);
  if (rc != QMCKL_SUCCESS) return rc;
  return copy_out(c, context, forces_mo_g, c->d_orb_out.ptr, sizeof(double) * n, fn);
}

qmckl_exit_code qmckl_get_forces_mo_g_inplace(qmckl_context context, double* const forces_mo_g, const int64_t size_max) {
  return qmckl_get_forces_mo_g(context, forces_mo_g, size_max);
}

// org/qmckl_forces.org:6497-6535: [nucl][3][point][mo]
qmckl_exit_code qmckl_get_forces_mo_l(qmckl_context context, double* const forces_mo_l, const int64_t size_max) {
  const char* fn = "qmckl_get_forces_mo_l";
  qmckl_exit_code rc;
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (!c->mo_provided) return fail(context, QMCKL_NOT_PROVIDED, fn, nullptr);
  const int64_t P = c->point_num, ao = c->ao_num, mo = c->mo_num;
  const int64_t n = c->nucl_num * 3 * P * mo;
  if (!getter_ctx(context, forces_mo_l, size_max, n, fn, rc)) return rc;
  rc = provide_ao_hessian(c, context, fn);
  if (rc == QMCKL_SUCCESS)
    rc = contract(c, context, fn, c->d_ao_hess.as<double>(), 3 * ao, 3, P * 4 * ao, P, 4 * ao, 3 * P * mo, P * mo, mo,
                  true, -1.0, n);
  if (rc != QMCKL_SUCCESS) return rc;
  return copy_out(c, context, forces_mo_l, c->d_orb_out.ptr, sizeof(double) * n, fn);
}

// org/qmckl_mo.org:479-700.  The parameters are fitted here, once: the MO values at each
// nucleus without the s AOs inside their r_cusp, and the nucleus' own s AOs at R + r_cusp z,
// both from the device getters on temporary points (the caller's points are restored).
qmckl_exit_code qmckl_set_mo_basis_r_cusp(qmckl_context context, const double* r_cusp, const int64_t size_max) {
  const char* fn = "qmckl_set_mo_basis_r_cusp";
  Context* c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (r_cusp == nullptr) return fail(context, QMCKL_INVALID_ARG_2, fn, "r_cusp: Null pointer");
  if (size_max < c->nucl_num) return fail(context, QMCKL_INVALID_ARG_3, fn, "Array too small");
  if (!c->mo_provided || !c->nucl_provided) return fail(context, QMCKL_NOT_PROVIDED, fn, nullptr);
  const int64_t Nn = c->nucl_num, mo = c->mo_num, ao = c->ao_num;
  for (int64_t a = 0; a < Nn; ++a) {
    const double Z = c->nucl_charge[a], R = r_cusp[a];
    if (Z < 0.1) continue;  // dummy atoms are not fitted
    if (!(R > 0.0) || std::fabs(2. * R * Z - 6.) <= 1.e-6 || std::fabs(R * R * Z - 3. * R) <= 1.e-6 ||
        std::fabs(R * R * R * Z - 3. * R * R) <= 1.e-6)
      return fail(context, QMCKL_INVALID_ARG_2, fn, "r_cusp and the nuclear charge make the fit singular");
  }
  // the caller's points
  const int64_t old_num = c->point_num;
  std::vector<double> old(3 * std::max<int64_t>(old_num, 0));
  qmckl_exit_code rc = QMCKL_SUCCESS;
  if (old_num > 0) rc = qmckl_get_point(context, 'T', old.data(), 3 * old_num);
  if (rc != QMCKL_SUCCESS) return rc;
  c->r_cusp.assign(r_cusp, r_cusp + Nn);
  c->cusp_param.assign(Nn * 4 * mo, 0.0);
  c->cusp_uploaded = false;
  touch_const(c);
  // MO values at the nuclei, s AOs inside r_cusp left out (cubic zero)
  std::vector<double> eta(Nn * mo), pts(c->nucl_coord), aov(Nn * 5 * ao);
  rc = qmckl_set_point(context, 'T', Nn, pts.data(), 3 * Nn);
  if (rc == QMCKL_SUCCESS) rc = qmckl_get_mo_basis_mo_value(context, eta.data(), Nn * mo);
  // the AOs at R + r_cusp z
  for (int64_t a = 0; a < Nn && rc == QMCKL_SUCCESS; ++a) pts[2 * Nn + a] += r_cusp[a];
  if (rc == QMCKL_SUCCESS) rc = qmckl_set_point(context, 'T', Nn, pts.data(), 3 * Nn);
  if (rc == QMCKL_SUCCESS) rc = qmckl_get_ao_basis_ao_vgl(context, aov.data(), Nn * 5 * ao);
  if (rc != QMCKL_SUCCESS) {
    c->r_cusp.clear();
    c->cusp_param.clear();
    return rc;
  }
  const AoTables& t = c->tables;
  for (int64_t a = 0; a < Nn; ++a) {
    const double Z = c->nucl_charge[a], R = r_cusp[a];
    if (Z < 0.1) continue;
    for (int64_t i = 0; i < mo; ++i) {
      double phi = 0.0, grad = 0.0, lap = 0.0;
      for (int64_t k = 0; k < ao; ++k) {
        if (t.shell_nucl[t.ao_shell[k]] != a || t.shell_l[t.ao_shell[k]] != 0) continue;
        const double ck = c->mo_coef[i * ao + k];
        const double* v = aov.data() + a * 5 * ao;
        phi += ck * v[k];
        grad += ck * v[3 * ao + k];
        lap += ck * v[4 * ao + k];
      }
      const double e = eta[a * mo + i];
      double* cp = c->cusp_param.data() + a * 4 * mo;
      cp[i] = -(R * (2. * e * Z - 6. * grad) + lap * R * R + 6. * phi) / (2. * R * Z - 6.);
      cp[mo + i] = (lap * R * R * Z - 6. * grad * R * Z + 6. * phi * Z + 6. * e * Z) / (2. * R * Z - 6.);
      cp[2 * mo + i] = -(R * (-5. * grad * Z - 1.5 * lap) + lap * R * R * Z + 3. * phi * Z + 3. * e * Z + 6. * grad) /
                       (R * R * Z - 3. * R);
      cp[3 * mo + i] = (R * (-2. * grad * Z - lap) + 0.5 * lap * R * R * Z + phi * Z + e * Z + 3. * grad) /
                       (R * R * R * Z - 3. * R * R);
    }
  }
  c->cusp_uploaded = false;
  touch_const(c);
  if (old_num > 0) rc = qmckl_set_point(context, 'T', old_num, old.data(), 3 * old_num);
  return rc;
}

}  // extern "C"
