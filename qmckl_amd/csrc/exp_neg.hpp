// e^x for -708 <= x <= 0 in FP64, as used by the radial parts of the AO kernels (the
// screening keeps x = -alpha r^2 in [-cutoff, 0], cutoff <= 52 ln 2).
//
// Method: Cody-Waite argument reduction (W. J. Cody and W. Waite, Software Manual for the
// Elementary Functions, 1980): n = rint(x / ln 2), r = x - n ln2_hi - n ln2_lo with ln2_hi
// = ln 2 rounded to 32 significant bits (n ln2_hi is exact for |n| < 2^21) and ln2_lo the
// double nearest ln 2 - ln2_hi, so |r| <= ln(2)/2; then e^r by its Taylor series to degree
// DEG in Horner form with fused multiply-adds, and e^x = 2^n e^r.  Every operation is an
// exactly rounded FP64 operation, so the host and the device compute the same bits.
//   DEG = 13 (exp_neg): truncation error |r|^14 / 14! < 4.2e-18, within 1 ulp of e^x
//            over the domain (measured maximum 0.80 ulp).
//   DEG = 9 (exp_neg_f32): truncation error |r|^10 / 10! < 7.1e-12 relative, for values
//            that are rounded to FP32 (2^-24 = 6e-8) before they are used.
// Below -708 the result leaves the normal range (gradual underflow through ldexp, 0 below
// about -745).
// REMAT: on the device the constants pass through an empty asm statement at each call, so
// that they are rematerialised in scalar registers there instead of being hoisted out of
// the caller's loops into vector registers for the whole kernel.
#pragma once

#include <hip/hip_runtime.h>

namespace qmckl_amd {

template <bool REMAT>
__host__ __device__ inline double exp_const(double c) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (REMAT) asm volatile("" : "+s"(c));
#endif
  return c;
}

template <int DEG, bool REMAT = false>
__host__ __device__ inline double exp_neg_deg(double x) {
  static_assert(DEG == 9 || DEG == 13, "degree");
  const double inv_ln2 = exp_const<REMAT>(0x1.71547652b82fep+0);  // 1 / ln 2
  const double ln2_hi = exp_const<REMAT>(0x1.62e42ff000000p-1);   // ln 2 to 32 significant bits
  const double ln2_lo = exp_const<REMAT>(-0x1.718432a1b0e26p-35);  // ln 2 - ln2_hi
  const double n = __builtin_rint(x * inv_ln2);
  const double r = __builtin_fma(n, -ln2_lo, __builtin_fma(n, -ln2_hi, x));
  double p;
  if constexpr (DEG == 13) {
    p = exp_const<REMAT>(0x1.6124613a86d09p-33);                       // 1/13!
    p = __builtin_fma(p, r, exp_const<REMAT>(0x1.1eed8eff8d898p-29));  // 1/12!
    p = __builtin_fma(p, r, exp_const<REMAT>(0x1.ae64567f544e4p-26));  // 1/11!
    p = __builtin_fma(p, r, exp_const<REMAT>(0x1.27e4fb7789f5cp-22));  // 1/10!
    p = __builtin_fma(p, r, exp_const<REMAT>(0x1.71de3a556c734p-19));  // 1/9!
  } else {
    p = exp_const<REMAT>(0x1.71de3a556c734p-19);  // 1/9!
  }
  p = __builtin_fma(p, r, exp_const<REMAT>(0x1.a01a01a01a01ap-16));  // 1/8!
  p = __builtin_fma(p, r, exp_const<REMAT>(0x1.a01a01a01a01ap-13));  // 1/7!
  p = __builtin_fma(p, r, exp_const<REMAT>(0x1.6c16c16c16c17p-10));  // 1/6!
  p = __builtin_fma(p, r, exp_const<REMAT>(0x1.1111111111111p-7));   // 1/5!
  p = __builtin_fma(p, r, exp_const<REMAT>(0x1.5555555555555p-5));   // 1/4!
  p = __builtin_fma(p, r, exp_const<REMAT>(0x1.5555555555555p-3));   // 1/3!
  p = __builtin_fma(p, r, 0.5);
  p = __builtin_fma(p, r, 1.0);
  p = __builtin_fma(p, r, 1.0);
  return __builtin_ldexp(p, (int)n);
}

__host__ __device__ inline double exp_neg(double x) { return exp_neg_deg<13>(x); }
__host__ __device__ inline double exp_neg_f32(double x) { return exp_neg_deg<9>(x); }

}  // namespace qmckl_amd
