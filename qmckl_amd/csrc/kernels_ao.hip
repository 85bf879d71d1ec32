// Atomic-orbital kernels for gfx950: screened contracted-Gaussian AO values, gradients
// and Laplacians (AO VGL), plus the primitive and shell VGL getters.
//
// Layout.  A workgroup of 256 lanes owns a tile of TP points.  The radial part of every
// shell of the tile (value, gradient, Laplacian: 5 doubles) is staged in LDS, then the
// lanes sweep the AOs of each point so that the 5 output rows [point][c][ao] are
// written with consecutive lanes on consecutive AOs (coalesced 64-lane stores).
//
// Arithmetic follows org/qmckl_ao.org (doc path):
//   shell radial part  :4769-4868  nucleus skipped if r^2 > cutoff*range,
//                                  primitive skipped if alpha r^2 > cutoff
//   polynomials        :6233-6331  Y = (x - X) + 1e-20, x^a y^b z^c and derivatives
//   AO assembly        :8438-8572  value, product-rule gradient,
//                                  Laplacian = P_l s1 + P s5 + 2 grad P . grad s
#include <hip/hip_runtime.h>

#include "internal.hpp"
#include "ao_device.hpp"

namespace qmckl_amd {

namespace {

constexpr int kBlock = 256;

// Stages per-(point, nucleus) displacement and per-(point, shell) radial VGL in LDS.
__device__ void stage_shells(const AoDevView& v, const double* __restrict__ coord, int64_t coord_ld, int64_t pbase,
                             int tp, double* s_nuc, double* s_shell) {
  const int Nn = (int)v.nucl_num, Ns = (int)v.shell_num;
  for (int t = threadIdx.x; t < tp * Nn; t += blockDim.x) {
    const int q = t / Nn, a = t - q * Nn;
    const int64_t p = pbase + q;
    const double x = coord[p] - v.nucl_coord[a];
    const double y = coord[p + coord_ld] - v.nucl_coord[a + Nn];
    const double z = coord[p + 2 * coord_ld] - v.nucl_coord[a + 2 * Nn];
    double* o = s_nuc + 4 * t;
    o[0] = x;
    o[1] = y;
    o[2] = z;
    o[3] = x * x + y * y + z * z;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < tp * Ns; t += blockDim.x) {
    const int q = t / Ns, s = t - q * Ns;
    const int a = v.shell_nucl[s];
    const double* d = s_nuc + 4 * (q * Nn + a);
    const double x = d[0], y = d[1], z = d[2], r2 = d[3];
    double r[5];
    shell_radial(v, s, a, x, y, z, r2, r);
    const double v0 = r[0], v1 = r[1], v2 = r[2], v3 = r[3], v4 = r[4];
    double* o = s_shell + 5 * t;
    o[0] = v0;
    o[1] = v1;
    o[2] = v2;
    o[3] = v3;
    o[4] = v4;
  }
  __syncthreads();
}

template <int OUT>
__global__ void __launch_bounds__(kBlock) k_ao_vgl(AoDevView v, const double* __restrict__ coord, int64_t coord_ld,
                                                   int64_t p0, int64_t np, void* __restrict__ out, int64_t ld_out,
                                                   int64_t plane, int values_only, int tp) {
  extern __shared__ double lds[];
  const int Nn = (int)v.nucl_num, Ns = (int)v.shell_num, Nao = (int)v.ao_num;
  double* s_nuc = lds;
  double* s_shell = lds + 4 * tp * Nn;
  const int64_t tile0 = (int64_t)blockIdx.x * tp;  // local point index (relative to p0)
  const int ntp = (int)min((int64_t)tp, np - tile0);
  if (ntp <= 0) return;
  stage_shells(v, coord, coord_ld, p0 + tile0, ntp, s_nuc, s_shell);
  const double cutoff = v.cutoff;
  for (int t = threadIdx.x; t < ntp * Nao; t += blockDim.x) {
    const int q = t / Nao, k = t - q * Nao;
    const int s = v.ao_shell[k];
    const int a = v.shell_nucl[s];
    const double* d = s_nuc + 4 * (q * Nn + a);
    double r[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (!(d[3] > cutoff * v.nucl_range[a])) {
      const int pw = v.ao_pow[k];
      double pv[5];
      monomial_vgl(d[0] + 1.e-20, d[1] + 1.e-20, d[2] + 1.e-20, pw & 0xff, (pw >> 8) & 0xff, (pw >> 16) & 0xff, pv);
      const double* sh = s_shell + 5 * (q * Ns + s);
      const double f = v.ao_factor[k];
      r[0] = pv[0] * sh[0] * f;
      r[1] = (pv[1] * sh[0] + pv[0] * sh[1]) * f;
      r[2] = (pv[2] * sh[0] + pv[0] * sh[2]) * f;
      r[3] = (pv[3] * sh[0] + pv[0] * sh[3]) * f;
      r[4] = (pv[4] * sh[0] + pv[0] * sh[4] + 2.0 * (pv[1] * sh[1] + pv[2] * sh[2] + pv[3] * sh[3])) * f;
    }
    const int64_t lp = tile0 + q;  // point index within the chunk
    if (OUT == 0) {
      double* o = static_cast<double*>(out);
      if (values_only) {
        o[lp * ld_out + k] = r[0];
      } else {
#pragma unroll
        for (int c = 0; c < 5; ++c) o[(lp * 5 + c) * ld_out + k] = r[c];
      }
    } else if (OUT == 1) {
      float* o = static_cast<float*>(out);
#pragma unroll
      for (int c = 0; c < 5; ++c) o[(lp * 5 + c) * ld_out + k] = (float)r[c];
    } else {
      uint16_t* o = static_cast<uint16_t*>(out);
#pragma unroll
      for (int c = 0; c < 5; ++c) {
        uint16_t h, m, l;
        split3(r[c], h, m, l);
        const int64_t idx = (lp * 5 + c) * ld_out + k;
        o[idx] = h;
        o[idx + plane] = m;
        o[idx + 2 * plane] = l;
      }
    }
  }
}

__global__ void __launch_bounds__(kBlock) k_shell_vgl(AoDevView v, const double* __restrict__ coord, int64_t P,
                                                      double* __restrict__ out, int tp) {
  extern __shared__ double lds[];
  const int Nn = (int)v.nucl_num, Ns = (int)v.shell_num;
  double* s_nuc = lds;
  double* s_shell = lds + 4 * tp * Nn;
  const int64_t tile0 = (int64_t)blockIdx.x * tp;
  const int ntp = (int)min((int64_t)tp, P - tile0);
  if (ntp <= 0) return;
  stage_shells(v, coord, P, tile0, ntp, s_nuc, s_shell);
  for (int t = threadIdx.x; t < ntp * Ns; t += blockDim.x) {
    const int q = t / Ns, s = t - q * Ns;
#pragma unroll
    for (int c = 0; c < 5; ++c) out[((tile0 + q) * 5 + c) * Ns + s] = s_shell[5 * t + c];
  }
}

// org/qmckl_ao.org:4463-4520: un-contracted Gaussians exp(-alpha r^2), no coefficient.
__global__ void k_prim_vgl(const double* __restrict__ nucl_coord, int64_t Nn, const int32_t* __restrict__ prim_nucl,
                           const double* __restrict__ expo, int64_t prim_num, double cutoff,
                           const double* __restrict__ coord, int64_t P, double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P * prim_num) return;
  const int64_t p = t / prim_num, i = t - p * prim_num;
  const int a = prim_nucl[i];
  const double x = coord[p] - nucl_coord[a];
  const double y = coord[p + P] - nucl_coord[a + Nn];
  const double z = coord[p + 2 * P] - nucl_coord[a + 2 * Nn];
  const double r2 = x * x + y * y + z * z;
  const double ar2 = expo[i] * r2;
  double r[5] = {0, 0, 0, 0, 0};
  if (!(ar2 > cutoff)) {
    const double e = exp_neg(-ar2);
    const double two_a = -2.0 * expo[i] * e;
    r[0] = e;
    r[1] = two_a * x;
    r[2] = two_a * y;
    r[3] = two_a * z;
    r[4] = two_a * (3.0 - 2.0 * ar2);
  }
#pragma unroll
  for (int c = 0; c < 5; ++c) out[(p * 5 + c) * prim_num + i] = r[c];
}

__global__ void k_to_f32(const double* __restrict__ src, int64_t n, float* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (float)src[i];
}

__global__ void k_split_bf16x3(const double* __restrict__ src, int64_t n, uint16_t* __restrict__ dst, int64_t plane) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint16_t h, m, l;
  split3(src[i], h, m, l);
  dst[i] = h;
  dst[i + plane] = m;
  dst[i + 2 * plane] = l;
}

// Points per workgroup so that the staged shells fit in 48 KiB of LDS.
int points_per_tile(const AoDevView& v) {
  const size_t per_point = sizeof(double) * (4 * v.nucl_num + 5 * v.shell_num);
  int tp = (int)((48 * 1024) / per_point);
  if (tp > 8) tp = 8;
  if (tp < 1) tp = 1;
  return tp;
}

}  // namespace

hipError_t launch_ao_vgl(const AoDevView& v, const double* coord, int64_t coord_ld, int64_t p0, int64_t np, void* out,
                         int out_kind, int64_t ld_out, int64_t plane_stride, int values_only, hipStream_t s) {
  if (np <= 0) return hipSuccess;
  const int tp = points_per_tile(v);
  const size_t lds = sizeof(double) * tp * (4 * v.nucl_num + 5 * v.shell_num);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  const dim3 grid((unsigned)((np + tp - 1) / tp));
  if (out_kind == 0)
    hipLaunchKernelGGL(k_ao_vgl<0>, grid, dim3(kBlock), lds, s, v, coord, coord_ld, p0, np, out, ld_out, plane_stride,
                       values_only, tp);
  else if (out_kind == 1)
    hipLaunchKernelGGL(k_ao_vgl<1>, grid, dim3(kBlock), lds, s, v, coord, coord_ld, p0, np, out, ld_out, plane_stride,
                       0, tp);
  else
    hipLaunchKernelGGL(k_ao_vgl<2>, grid, dim3(kBlock), lds, s, v, coord, coord_ld, p0, np, out, ld_out, plane_stride,
                       0, tp);
  return hipGetLastError();
}

hipError_t launch_shell_vgl(const AoDevView& v, const double* coord, int64_t P, double* out, hipStream_t s) {
  const int tp = points_per_tile(v);
  const size_t lds = sizeof(double) * tp * (4 * v.nucl_num + 5 * v.shell_num);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_shell_vgl, dim3((unsigned)((P + tp - 1) / tp)), dim3(kBlock), lds, s, v, coord, P, out, tp);
  return hipGetLastError();
}

hipError_t launch_prim_vgl(const AoDevView& v, int64_t prim_num, const int32_t* prim_nucl, const double* prim_expo,
                           const double* coord, int64_t P, double* out, hipStream_t s) {
  const int64_t n = P * prim_num;
  hipLaunchKernelGGL(k_prim_vgl, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, v.nucl_coord, v.nucl_num,
                     prim_nucl, prim_expo, prim_num, v.cutoff, coord, P, out);
  return hipGetLastError();
}

hipError_t launch_to_f32(const double* src, int64_t n, float* dst, hipStream_t s) {
  hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, n, dst);
  return hipGetLastError();
}

hipError_t launch_split_bf16x3(const double* src, int64_t n, uint16_t* dst, int64_t plane, hipStream_t s) {
  hipLaunchKernelGGL(k_split_bf16x3, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, n, dst, plane);
  return hipGetLastError();
}

}  // namespace qmckl_amd
