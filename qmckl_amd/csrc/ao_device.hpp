// Device helpers shared by the AO kernels and the fused AO->MO kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "exp_neg.hpp"
#include "internal.hpp"

namespace qmckl_amd {

__device__ __forceinline__ uint16_t bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

// Three-term bfloat16 split by truncation of f = float(x): hi keeps the top 8 significand
// bits of f, mid the next 8 of the (exact) remainder and lo the rest, so that
// hi + mid + lo == f exactly; the only rounding is x -> float(x) (relative 2^-24).
__device__ __forceinline__ void split3_trunc(double x, uint16_t& h, uint16_t& m, uint16_t& l) {
  const float f = (float)x;
  const uint32_t uf = __float_as_uint(f);
  h = (uint16_t)(uf >> 16);
  const float r1 = f - __uint_as_float(uf & 0xFFFF0000u);
  const uint32_t u1 = __float_as_uint(r1);
  m = (uint16_t)(u1 >> 16);
  const float r2 = r1 - __uint_as_float(u1 & 0xFFFF0000u);
  l = (uint16_t)(__float_as_uint(r2) >> 16);
}

// Three-term bfloat16 split of a double: x ~ hi + mid + lo, each term rounded to nearest.
__device__ __forceinline__ void split3(double x, uint16_t& h, uint16_t& m, uint16_t& l) {
  h = bf16_rne((float)x);
  const double r1 = x - (double)bf16_to_f(h);
  m = bf16_rne((float)r1);
  const double r2 = r1 - (double)bf16_to_f(m);
  l = bf16_rne((float)r2);
}

// Value, gradient and Laplacian of x^a y^b z^c at Y = (x,y,z) + 1e-20, with the
// multiplication order of qmckl_ao_polynomial_vgl_doc.
// y^n, y^(n-1) and y^(n-2) by repeated multiplication (negative powers are returned as 1:
// they are always multiplied by a zero factor).
__device__ __forceinline__ void powers3(double y, int n, double& p0, double& p1, double& p2) {
  if (n <= 3) {
    // s to f shells without a loop (the lanes of a wave hold different powers): y^n as
    // f1 f2 f3 with f_k = y when n >= k, else 1, three selects; the same products as the
    // loop ((y y) y), so the same values
    const double f1 = n >= 1 ? y : 1.0, f2 = n >= 2 ? y : 1.0, f3 = n >= 3 ? y : 1.0;
    const double f23 = f2 * f3;
    p0 = f1 * f23;
    p1 = f23;
    p2 = f3;
    return;
  }
  p0 = 1.0;
  p1 = 1.0;
  p2 = 1.0;
  for (int i = 0; i < n; ++i) {
    p2 = p1;
    p1 = p0;
    p0 = p0 * y;
  }
}

__device__ __forceinline__ void monomial_vgl(double yx, double yy, double yz, int a, int b, int c, double p[5]) {
  double xa, xa1, xa2, yb, yb1, yb2, zc, zc1, zc2;
  powers3(yx, a, xa, xa1, xa2);
  powers3(yy, b, yb, yb1, yb2);
  powers3(yz, c, zc, zc1, zc2);
  const double da = a, db = b, dc = c;
  double xy = xa * yb;
  double yzv = yb * zc;
  double xz = xa * zc;
  p[0] = xy * zc;
  xy = dc * xy;
  xz = db * xz;
  yzv = da * yzv;
  p[1] = xa1 * yzv;
  p[2] = yb1 * xz;
  p[3] = zc1 * xy;
  p[4] = (da - 1.0) * xa2 * yzv + (db - 1.0) * yb2 * xz + (dc - 1.0) * zc2 * xy;
}

// r^n for an integer n by repeated multiplication (n < 0: 1 / r^|n|), as Fortran's r**n.
__device__ __forceinline__ double ipow_r(double r, int n) {
  double p = 1.0;
  for (int i = 0; i < (n < 0 ? -n : n); ++i) p *= r;
  return n < 0 ? 1.0 / p : p;
}

// One Slater primitive c r^n e^(-alpha r) at distance r (1/r = r_inv), org/qmckl_ao.org:5004-5034:
// skipped when alpha r > cutoff; adds its value to v0, (n/r - alpha) v to g (the gradient is
// g (x, y, z) / r) and (n^2 - 2 n alpha r + alpha r (alpha r - 2) + n) v / r^2 to lap.
template <int DEG = 13>
__device__ __forceinline__ void slater_prim(double al, double coef, int n, double r, double r_inv, double cutoff,
                                            double& v0, double& g, double& lap) {
  const double ar = al * r;
  if (ar > cutoff) return;
  const double v = coef * ipow_r(r, n) * exp_neg_deg<DEG>(-ar);
  const double dn = (double)n;
  v0 += v;
  g += (dn * r_inv - al) * v;
  lap += r_inv * r_inv * (dn * dn - 2.0 * dn * ar + ar * (ar - 2.0) + dn) * v;
}

// Radial value, gradient and Laplacian of shell s (on nucleus a) at displacement
// (x, y, z) = r - R_a, r2 = |r - R_a|^2; org/qmckl_ao.org:4769-4868: the nucleus is skipped
// when r2 > cutoff * range, a primitive when alpha r2 > cutoff.
// Slater shells (v.slater), org/qmckl_ao.org:4930-5040: the nucleus is skipped when
// r > cutoff * range; at r < 1e-15 a shell of r_power 0 has the sum of its coefficients as
// value, others 0; else the primitives of slater_prim.
__device__ __forceinline__ void shell_radial(const AoDevView& v, int s, int a, double x, double y, double z,
                                             double r2, double r[5]) {
  const double cutoff = v.cutoff;
  double v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0;
  if (v.slater) {
    const double rr = sqrt(r2);
    const int p0 = v.shell_prim_start[s], np = v.shell_prim_count[s], n = v.shell_rpow[s];
    if (!(rr > cutoff * v.nucl_range[a])) {
      if (rr < 1.e-15) {
        if (n == 0)
          for (int i = 0; i < np; ++i) v0 += v.prim_coef[p0 + i];
      } else {
        const double r_inv = 1.0 / rr;
        double g = 0.0;
        for (int i = 0; i < np; ++i)
          slater_prim(v.prim_expo[p0 + i], v.prim_coef[p0 + i], n, rr, r_inv, cutoff, v0, g, v4);
        v1 = g * x * r_inv;
        v2 = g * y * r_inv;
        v3 = g * z * r_inv;
      }
    }
  } else if (!(r2 > cutoff * v.nucl_range[a])) {
    const int p0 = v.shell_prim_start[s], np = v.shell_prim_count[s];
    for (int i = 0; i < np; ++i) {
      const double al = v.prim_expo[p0 + i];
      const double ar2 = al * r2;
      if (ar2 > cutoff) continue;
      const double e = v.prim_coef[p0 + i] * exp_neg(-ar2);
      const double two_a = -2.0 * al * e;
      v0 += e;
      v1 += two_a * x;
      v2 += two_a * y;
      v3 += two_a * z;
      v4 += two_a * (3.0 - 2.0 * ar2);
    }
  }
  r[0] = v0;
  r[1] = v1;
  r[2] = v2;
  r[3] = v3;
  r[4] = v4;
}

// AO VGL from the monomial of AO k at displacement d = (x, y, z) and its shell's radial
// VGL sh[5]; org/qmckl_ao.org:8530-8565.
__device__ __forceinline__ void ao_from_shell(const AoDevView& v, int k, double x, double y, double z,
                                              const double* sh, double r[5]) {
  const int pw = v.ao_pow[k];
  double pv[5];
  monomial_vgl(x + 1.e-20, y + 1.e-20, z + 1.e-20, pw & 0xff, (pw >> 8) & 0xff, (pw >> 16) & 0xff, pv);
  const double f = v.ao_factor[k];
  r[0] = pv[0] * sh[0] * f;
  r[1] = (pv[1] * sh[0] + pv[0] * sh[1]) * f;
  r[2] = (pv[2] * sh[0] + pv[0] * sh[2]) * f;
  r[3] = (pv[3] * sh[0] + pv[0] * sh[3]) * f;
  r[4] = (pv[4] * sh[0] + pv[0] * sh[4] + 2.0 * (pv[1] * sh[1] + pv[2] * sh[2] + pv[3] * sh[3])) * f;
}

}  // namespace qmckl_amd
