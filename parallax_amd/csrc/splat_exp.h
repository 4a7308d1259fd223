// Deterministic single-precision exponential.
//
// Every discrete decision of the splat path -- the opacity cull, the 3-sigma
// radius that decides tile membership, the `alpha < min_alpha` skip and the
// transmittance cut-off of the compositing loop -- depends on an exponential.
// The hardware `v_exp_f32` and glibc's `expf` round differently in the last
// bit, so a kernel and a CPU checker built on them would disagree on a few
// pixels per frame. This function is built from IEEE-754 operations that
// round identically on the host and on the GPU (multiply, fused multiply-add,
// round to nearest even, exact scaling by a power of two), so the host checker
// and the kernels get the same bits by construction.
//
// Method: n = x/ln2 rounded to the nearest integer by one fused multiply-add
// with the shifter 1.5 * 2^23 (the sum's low mantissa bits then hold n);
// Cody-Waite range reduction x = n*ln2 + r with a two-part ln2 and fused
// multiply-adds, |r| <= ln2/2 (+2e-6 where the float log2(e) rounds n the
// other way); e^r by a degree-6 polynomial, Horner's rule in fused
// multiply-adds; then 2^n added into the exponent field with one integer add
// (p is in [0.70, 1.42] and n in [-116, 127], so the result is always a
// normal float and the scaling is exact). The polynomial is 1 + r q(r), q
// interpolated at 200 Chebyshev nodes of |r| <= ln2/2 + 2e-5 in double
// precision, coefficients rounded to float: |p/e^r - 1| <= 1e-8. Over every
// float in (-80, 88] the result is within 0.97 ulp of the correctly rounded
// e^x (the test suite checks a sample against the double-precision libm exp).
// Inputs at or below -80 return 0 and inputs above 88 return +inf; NaN
// propagates.
//
// The host side needs a hardware or correctly rounded fmaf (x86-64 with FMA3
// when built with -mfma; glibc's software fmaf otherwise); the device side
// uses v_fma_f32. Compile the including translation unit with
// -ffp-contract=off so the compiler does not fuse other multiplies and adds.
#pragma once

#include <stdint.h>

#include <cmath>
#include <cstring>

#if defined(__HIPCC__)
#define SPLAT_EXP_QUAL __host__ __device__ __forceinline__
#else
#define SPLAT_EXP_QUAL inline
#endif

namespace splat_det {

SPLAT_EXP_QUAL float bits_to_float(uint32_t u) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __uint_as_float(u);
#else
  float f;
  std::memcpy(&f, &u, sizeof f);
  return f;
#endif
}

SPLAT_EXP_QUAL uint32_t float_to_bits(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __float_as_uint(f);
#else
  uint32_t u;
  std::memcpy(&u, &f, sizeof u);
  return u;
#endif
}

SPLAT_EXP_QUAL float fma(float a, float b, float c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_fmaf(a, b, c);
#else
  return std::fma(a, b, c);
#endif
}

// e^x for x in (-80, 88], finite: the core shared by exp() and exp_nonpos().
SPLAT_EXP_QUAL float exp_core(float x) {
  // t = 1.5 * 2^23 + n, n = x * log2(e) rounded to nearest even, in [-116, 127]
  const float t = fma(x, 1.44269502e+00f, 12582912.0f);
  const float fn = t - 12582912.0f;  // exact
  // ln2 = 6.93147182e-01 (float nearest) + -1.90465421e-09
  float r = fma(-fn, 6.93147182e-01f, x);
  r = fma(-fn, -1.90465421e-09f, r);
  float p = fma(1.394111663e-03f, r, 8.369155228e-03f);
  p = fma(p, r, 4.166635126e-02f);
  p = fma(p, r, 1.666650474e-01f);
  p = fma(p, r, 5.0e-01f);
  p = fma(p, r, 1.0f);
  p = fma(p, r, 1.0f);
  // t's bits are those of 1.5 * 2^23 plus n; shifted left by 23 only n << 23
  // remains (mod 2^32), which adds n to p's exponent
  return bits_to_float(float_to_bits(p) + (float_to_bits(t) << 23));
}

SPLAT_EXP_QUAL float exp(float x) {
  if (x != x) return x;
  if (x <= -80.0f) return 0.0f;
  if (x > 88.0f) return bits_to_float(0x7f800000u);
  return exp_core(x);
}

/// exp(x) for x <= 0 and not NaN (the compositing loop's Gaussian falloff):
/// the same bits as exp() on that domain, without its other range checks and
/// without a branch (the core runs on a clamped argument, the result is
/// selected).
SPLAT_EXP_QUAL float exp_nonpos(float x) {
  const float e = exp_core(x > -80.0f ? x : -80.0f);
  return x <= -80.0f ? 0.0f : e;
}

/// Logistic function written with the same exponential.
SPLAT_EXP_QUAL float sigmoid(float x) { return 1.0f / (1.0f + exp(-x)); }

}  // namespace splat_det
