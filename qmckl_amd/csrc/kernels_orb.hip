// Second and third derivatives of the orbitals, the forces on the orbitals and the MO cusp
// correction:
//   shell Hessian [P][3][4][shell] and AO Hessian [3][P][4][ao] (org/qmckl_ao.org:5257-5585,
//   9397-9718); the nucleus-block contraction behind the MO Hessian (org/qmckl_mo.org:
//   2362-2776) and the forces on the MO values, gradients and Laplacians
//   (org/qmckl_forces.org:5467-6878); the forces on the AO values (:5122-5466); the cusp
//   correction of the MO VGL and Hessian inside r_cusp of a nucleus (org/qmckl_mo.org:
//   2777-3470).
//
// The AO Hessian is one thread per (point, AO): the radial part S = sum_p c_p exp(-a_p r^2) of
// the AO's shell with its first, second and third derivatives, the cartesian monomial with
// its own, combined by the product rule (d_i d_j (M S) and d_i lap (M S)).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "internal.hpp"

namespace qmckl_amd {
namespace {

constexpr int kOrbBlock = 256;

// Radial part of shell s at r (r2 = |r|^2): value, gradient, Laplacian in S[5], second
// derivatives in H[3][3], gradient of the Laplacian in T[3].  Same screening as the AO VGL
// (primitives with a r^2 > cutoff left out).
__device__ void radial_derivs(const AoDevView& v, int s, const double r[3], double r2, double S[5], double H[3][3],
                              double T[3]) {
#pragma unroll
  for (int i = 0; i < 5; ++i) S[i] = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    T[i] = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) H[i][j] = 0.0;
  }
  const int p0 = v.shell_prim_start[s], np = v.shell_prim_count[s];
  for (int q = p0; q < p0 + np; ++q) {
    const double a = v.prim_expo[q];
    const double ar2 = a * r2;
    if (ar2 > v.cutoff) continue;
    const double e = v.prim_coef[q] * exp(-ar2);
    const double ta = 2.0 * a;
    S[0] += e;
    S[4] += (ta * ta * r2 - 3.0 * ta) * e;
    const double t3 = (5.0 * ta * ta - ta * ta * ta * r2) * e;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      S[1 + i] -= ta * r[i] * e;
      T[i] += t3 * r[i];
#pragma unroll
      for (int j = 0; j < 3; ++j) H[i][j] += (ta * ta * r[i] * r[j] - (i == j ? ta : 0.0)) * e;
    }
  }
}

__device__ inline double ipow_nn(double y, int e) {
  if (e < 0) return 0.0;
  double r = 1.0;
  for (int i = 0; i < e; ++i) r *= y;
  return r;
}

// one-dimensional grids: block b covers items (b % nb) * kOrbBlock + thread of point b / nb
__global__ void __launch_bounds__(kOrbBlock) k_ao_hessian(AoDevView v, const double* __restrict__ coord, int64_t P,
                                                         double* __restrict__ out) {
  const int64_t nb = (v.ao_num + kOrbBlock - 1) / kOrbBlock;
  const int64_t k = (blockIdx.x % nb) * kOrbBlock + threadIdx.x;
  const int64_t p = blockIdx.x / nb;
  if (k >= v.ao_num) return;
  const int s = v.ao_shell[k];
  const int a = v.shell_nucl[s];
  const double r[3] = {coord[p] - v.nucl_coord[a], coord[P + p] - v.nucl_coord[v.nucl_num + a],
                       coord[2 * P + p] - v.nucl_coord[2 * v.nucl_num + a]};
  const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
  double o[3][4];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) o[i][j] = 0.0;
  if (r2 <= v.cutoff * v.nucl_range[a]) {
    double S[5], Sh[3][3], St[3];
    radial_derivs(v, s, r, r2, S, Sh, St);
    const int pw = v.ao_pow[k];
    const int n[3] = {pw & 255, (pw >> 8) & 255, (pw >> 16) & 255};
    double p0[3], d1[3], d2[3], d3[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double e = n[d];
      p0[d] = ipow_nn(r[d], n[d]);
      d1[d] = e * ipow_nn(r[d], n[d] - 1);
      d2[d] = e * (e - 1.0) * ipow_nn(r[d], n[d] - 2);
      d3[d] = e * (e - 1.0) * (e - 2.0) * ipow_nn(r[d], n[d] - 3);
    }
    double Mg[3], Mh[3][3], Mt[3];
    const double M = p0[0] * p0[1] * p0[2];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const int j = (i + 1) % 3, q = (i + 2) % 3;
      Mg[i] = d1[i] * p0[j] * p0[q];
      Mh[i][i] = d2[i] * p0[j] * p0[q];
      Mh[i][j] = Mh[j][i] = d1[i] * d1[j] * p0[q];
      Mt[i] = d3[i] * p0[j] * p0[q] + d1[i] * (d2[j] * p0[q] + p0[j] * d2[q]);
    }
    const double Ml = Mh[0][0] + Mh[1][1] + Mh[2][2];
    const double f = v.ao_factor[k];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
      for (int j = 0; j < 3; ++j) o[i][j] = (Mh[i][j] * S[0] + M * Sh[i][j] + Mg[j] * S[1 + i] + Mg[i] * S[1 + j]) * f;
      double cross = 0.0;
#pragma unroll
      for (int j = 0; j < 3; ++j) cross += Mh[j][i] * S[1 + j] + Mg[j] * Sh[j][i];
      o[i][3] = (Mt[i] * S[0] + M * St[i] + Ml * S[1 + i] + Mg[i] * S[4] + 2.0 * cross) * f;
    }
  }
  const int64_t ao = v.ao_num;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) out[((i * P + p) * 4 + j) * ao + k] = o[i][j];
}

__global__ void __launch_bounds__(kOrbBlock) k_shell_hessian(AoDevView v, const double* __restrict__ coord,
                                                            int64_t P, double* __restrict__ out) {
  const int64_t nb = (v.shell_num + kOrbBlock - 1) / kOrbBlock;
  const int64_t s = (blockIdx.x % nb) * kOrbBlock + threadIdx.x;
  const int64_t p = blockIdx.x / nb;
  if (s >= v.shell_num) return;
  const int a = v.shell_nucl[s];
  const double r[3] = {coord[p] - v.nucl_coord[a], coord[P + p] - v.nucl_coord[v.nucl_num + a],
                       coord[2 * P + p] - v.nucl_coord[2 * v.nucl_num + a]};
  const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
  double S[5], H[3][3], T[3];
  if (r2 <= v.cutoff * v.nucl_range[a]) {
    radial_derivs(v, (int)s, r, r2, S, H, T);
  } else {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      T[i] = 0.0;
#pragma unroll
      for (int j = 0; j < 3; ++j) H[i][j] = 0.0;
    }
  }
  const int64_t ns = v.shell_num;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j) out[((p * 3 + i) * 4 + j) * ns + s] = H[i][j];
    out[((p * 3 + i) * 4 + 3) * ns + s] = T[i];
  }
}

// out[a][u][v][i] (strides oa, ou, ov) = sign * sum_{k0[a] <= k < k0[a+1]} src[base + u su +
// v sv + k] * ct[k][i]; one workgroup per (a, u, v) row, the threads over the MOs.
__global__ void __launch_bounds__(kOrbBlock) k_orb_contract(ContractArgs c) {
  const int64_t row = blockIdx.x;
  const int64_t v = row % c.V;
  const int64_t u = (row / c.V) % c.U;
  const int64_t a = row / (c.V * c.U);
  const int64_t k0 = c.k_start[a], k1 = c.k_start[a + 1];
  const double* src = c.src + c.base + u * c.su + v * c.sv;
  double* out = c.out + a * c.oa + u * c.ou + v * c.ov;
  for (int64_t i = threadIdx.x; i < c.mo; i += kOrbBlock) {
    double acc = 0.0;
    for (int64_t k = k0; k < k1; ++k) acc += src[k] * c.ct[k * c.ldct + i];
    out[i] = c.sign * acc;
  }
}

// out[a][p][d][k] = -ao_vgl[p][1 + d][k] for the AOs k on nucleus a, 0 for the others
__global__ void __launch_bounds__(kOrbBlock) k_forces_ao_value(AoDevView v, const double* __restrict__ ao_vgl,
                                                              int64_t P, double* __restrict__ out) {
  const int64_t nb = (v.ao_num + kOrbBlock - 1) / kOrbBlock;
  const int64_t k = (blockIdx.x % nb) * kOrbBlock + threadIdx.x;
  const int64_t p = (blockIdx.x / nb) % P;
  const int64_t a = blockIdx.x / (nb * P);
  if (k >= v.ao_num) return;
  const int64_t ao = v.ao_num;
  const bool on = v.shell_nucl[v.ao_shell[k]] == a;
#pragma unroll
  for (int d = 0; d < 3; ++d)
    out[((a * P + p) * 3 + d) * ao + k] = on ? -ao_vgl[(p * 5 + 1 + d) * ao + k] : 0.0;
}

constexpr int kCuspMaxS = 64;  // s AOs per nucleus staged in LDS

// For each point and each nucleus closer than its r_cusp: the nucleus' s AOs taken out of the
// MOs and c0 + c1 r + c2 r^2 + c3 r^3 added (value, gradient, Laplacian; or the Hessian).
__global__ void __launch_bounds__(kOrbBlock) k_mo_cusp(CuspArgs c) {
  __shared__ double sv[kCuspMaxS][9];
  __shared__ int sk[kCuspMaxS];
  const int64_t p = blockIdx.x;
  const AoDevView& v = c.v;
  const double x = c.coord[p], y = c.coord[c.P + p], z = c.coord[2 * c.P + p];
  for (int64_t A = 0; A < c.nucl_num; ++A) {
    const double r3[3] = {x - v.nucl_coord[A], y - v.nucl_coord[v.nucl_num + A], z - v.nucl_coord[2 * v.nucl_num + A]};
    const double r2 = r3[0] * r3[0] + r3[1] * r3[1] + r3[2] * r3[2];
    const double r = sqrt(r2);
    if (!(r < c.r_cusp[A])) continue;  // uniform over the workgroup
    const int s0 = c.s_start[A], ns = c.s_start[A + 1] - s0;
    const bool screened = r2 > v.cutoff * v.nucl_range[A];
    for (int t = threadIdx.x; t < ns; t += kOrbBlock) {
      const int k = c.s_ao[s0 + t];
      double S[5], H[3][3], T[3];
      if (screened) {
        for (int q = 0; q < 5; ++q) S[q] = 0.0;
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) H[i][j] = 0.0;
      } else {
        radial_derivs(v, v.ao_shell[k], r3, r2, S, H, T);
      }
      const double f = v.ao_factor[k];
      sk[t] = k;
      if (c.hessian) {
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) sv[t][3 * i + j] = f * H[i][j];
      } else {
        for (int q = 0; q < 5; ++q) sv[t][q] = f * S[q];
      }
    }
    __syncthreads();
    const double ir = 1.0 / r;
    for (int64_t i = threadIdx.x; i < c.mo; i += kOrbBlock) {
      const double* cp = c.param + A * 4 * c.mo;
      const double c0 = cp[i], c1 = cp[c.mo + i], c2 = cp[2 * c.mo + i], c3 = cp[3 * c.mo + i];
      if (c.hessian) {
        for (int n = 0; n < 3; ++n)
          for (int m = 0; m < 3; ++m) {
            double s = 0.0;
            for (int t = 0; t < ns; ++t) s += c.ct[sk[t] * c.ldct + i] * sv[t][3 * n + m];
            const double d = n == m ? 1.0 : 0.0;
            const double poly = c1 * (d * ir - r3[n] * r3[m] * ir * ir * ir) + 2.0 * c2 * d +
                                3.0 * c3 * (r3[n] * r3[m] * ir + d * r);
            double* o = c.out + ((n * c.P + p) * 3 + m) * c.mo + i;
            *o = *o - s + poly;
          }
      } else {
        double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        for (int t = 0; t < ns; ++t) {
          const double w = c.ct[sk[t] * c.ldct + i];
          for (int q = 0; q < 5; ++q) s[q] += w * sv[t][q];
        }
        const double g = c1 * ir + 2.0 * c2 + 3.0 * c3 * r;
        double* o = c.out + p * 5 * c.mo + i;
        o[0] = o[0] - s[0] + (c0 + r * (c1 + r * (c2 + r * c3)));
        for (int d = 0; d < 3; ++d) o[(1 + d) * c.mo] = o[(1 + d) * c.mo] - s[1 + d] + r3[d] * g;
        o[4 * c.mo] = o[4 * c.mo] - s[4] + (2.0 * c1 * ir + 6.0 * c2 + 12.0 * c3 * r);
      }
    }
    __syncthreads();
  }
}

}  // namespace

hipError_t launch_ao_hessian(const AoDevView& v, const double* coord, int64_t P, double* out, hipStream_t s) {
  if (P <= 0 || v.ao_num <= 0) return hipSuccess;
  const int64_t nb = (v.ao_num + kOrbBlock - 1) / kOrbBlock;
  k_ao_hessian<<<dim3((unsigned)(nb * P)), dim3(kOrbBlock), 0, s>>>(v, coord, P, out);
  return hipGetLastError();
}

hipError_t launch_shell_hessian(const AoDevView& v, const double* coord, int64_t P, double* out, hipStream_t s) {
  if (P <= 0 || v.shell_num <= 0) return hipSuccess;
  const int64_t nb = (v.shell_num + kOrbBlock - 1) / kOrbBlock;
  k_shell_hessian<<<dim3((unsigned)(nb * P)), dim3(kOrbBlock), 0, s>>>(v, coord, P, out);
  return hipGetLastError();
}

hipError_t launch_orb_contract(const ContractArgs& c, hipStream_t s) {
  const int64_t rows = c.na * c.U * c.V;
  if (rows <= 0 || c.mo <= 0) return hipSuccess;
  k_orb_contract<<<dim3((unsigned)rows), dim3(kOrbBlock), 0, s>>>(c);
  return hipGetLastError();
}

hipError_t launch_forces_ao_value(const AoDevView& v, const double* ao_vgl, int64_t P, double* out, hipStream_t s) {
  if (P <= 0 || v.ao_num <= 0 || v.nucl_num <= 0) return hipSuccess;
  const int64_t nb = (v.ao_num + kOrbBlock - 1) / kOrbBlock;
  k_forces_ao_value<<<dim3((unsigned)(nb * P * v.nucl_num)), dim3(kOrbBlock), 0, s>>>(v, ao_vgl, P, out);
  return hipGetLastError();
}

int cusp_max_s_ao() { return kCuspMaxS; }

hipError_t launch_mo_cusp(const CuspArgs& c, hipStream_t s) {
  if (c.P <= 0 || c.mo <= 0) return hipSuccess;
  k_mo_cusp<<<dim3((unsigned)c.P), dim3(kOrbBlock), 0, s>>>(c);
  return hipGetLastError();
}

}  // namespace qmckl_amd
