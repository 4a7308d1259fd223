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
// Method: Cody-Waite range reduction x = n*ln2 + r with a two-part ln2 and
// fused multiply-adds (r is exact to within the low part's rounding),
// |r| <= ln2/2, then the degree-7 Taylor polynomial of e^r by Horner's rule
// in fused multiply-adds, then scaling by 2^n through the exponent field.
// Accuracy is within a few ulp of the correctly rounded e^x over [-80, 88]
// (checked against the double-precision libm exp by the test suite). Inputs
// at or below -80 return 0 and inputs above 88 return +inf; NaN propagates.
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

SPLAT_EXP_QUAL float round_even(float v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_rintf(v);
#else
  return std::nearbyint(v);
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
  // n = nearest integer to x / ln2, in [-116, 127]
  const float fn = round_even(x * 1.44269502e+00f);
  // ln2 = 6.93147182e-01 (float nearest) + -1.90465421e-09
  float r = fma(-fn, 6.93147182e-01f, x);
  r = fma(-fn, -1.90465421e-09f, r);
  float p = fma(1.98412698e-04f, r, 1.38888889e-03f);  // 1/7!, 1/6!
  p = fma(p, r, 8.33333377e-03f);  // 1/5!
  p = fma(p, r, 4.16666679e-02f);  // 1/4!
  p = fma(p, r, 1.66666672e-01f);  // 1/3!
  p = fma(p, r, 5.00000000e-01f);
  p = fma(p, r, 1.00000000e+00f);
  p = fma(p, r, 1.00000000e+00f);
  const int n = static_cast<int>(fn);
  return p * bits_to_float(static_cast<uint32_t>(n + 127) << 23);
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
