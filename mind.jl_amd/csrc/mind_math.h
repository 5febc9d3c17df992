// FP64 elementary functions shared bit-for-bit by the HIP kernels and the host.
//
// Spike timing depends on every ulp of the voltage trajectory, so the device
// kernels and the CPU restatement used to check them must evaluate exp/log with
// the same sequence of IEEE-754 double operations. Library exp/log (libm on the
// host, the device math library on the GPU) differ in their last bits, so both
// sides call these functions instead. Everything here is plain double
// arithmetic and bit manipulation; it must be compiled with -ffp-contract=off.
//
// Methods (published, freely redistributable):
//  - mind_exp: argument reduction x = k*ln2 + r, |r| <= ln2/2, with ln2 split
//    into a 32-bit head and a tail, then the Remez rational form
//    exp(r) = 1 + 2r / (2 - r + r*R(r^2)) of Sun's fdlibm e_exp.c (error < 1 ulp).
//  - mind_log: x = 2^k (1+f) with sqrt(2)/2 <= 1+f < sqrt(2), s = f/(2+f),
//    log(1+f) = f - s(f - R(s^2)) with the Remez polynomial of fdlibm e_log.c
//    (error < 1 ulp).
// The reference evaluates these through its GPU vendor math library
// (src/compiler/codegen.jl:91, :626); that library is not reproduced here.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define MIND_HD __host__ __device__ inline
#else
#define MIND_HD inline
#endif

MIND_HD uint64_t mind_bits(double x) { return __builtin_bit_cast(uint64_t, x); }
MIND_HD double mind_from_bits(uint64_t b) { return __builtin_bit_cast(double, b); }

// 2^k for -1022 <= k <= 1023, built from its exponent field.
MIND_HD double mind_pow2i(int k) { return mind_from_bits((uint64_t)(k + 1023) << 52); }

MIND_HD double mind_exp(double x) {
  const double ln2_hi = 6.93147180369123816490e-01;
  const double ln2_lo = 1.90821492927058770002e-10;
  const double inv_ln2 = 1.44269504088896338700e+00;
  const double P1 = 1.66666666666666019037e-01;
  const double P2 = -2.77777777770155933842e-03;
  const double P3 = 6.61375632143793436117e-05;
  const double P4 = -1.65339022054652515390e-06;
  const double P5 = 4.13813679705723846039e-08;
  const double overflow = 7.09782712893383973096e+02;
  const double underflow = -7.45133219101941108420e+02;
  if (x != x) return x + x;
  if (x > overflow) return mind_from_bits(0x7ff0000000000000ull);
  if (x < underflow) return 0.0;
  const double ax = x < 0.0 ? -x : x;
  double hi = 0.0, lo = 0.0, r;
  int k = 0;
  if (ax > 0.5 * 0.6931471805599453) {
    if (ax < 1.5 * 0.6931471805599453) {
      k = x < 0.0 ? -1 : 1;
      hi = x < 0.0 ? x + ln2_hi : x - ln2_hi;
      lo = x < 0.0 ? -ln2_lo : ln2_lo;
    } else {
      k = (int)(inv_ln2 * x + (x < 0.0 ? -0.5 : 0.5));
      const double t = (double)k;
      hi = x - t * ln2_hi;
      lo = t * ln2_lo;
    }
    r = hi - lo;
  } else if (ax < 3.7252902984619140625e-09) {  // 2^-28: exp(x) rounds to 1 + x
    return 1.0 + x;
  } else {
    r = x;
  }
  const double t = r * r;
  const double c = r - t * (P1 + t * (P2 + t * (P3 + t * (P4 + t * P5))));
  if (k == 0) return 1.0 - ((r * c) / (c - 2.0) - r);
  const double y = 1.0 - ((lo - (r * c) / (2.0 - c)) - hi);
  if (k >= -1021) {
    if (k == 1024) return y * 2.0 * mind_pow2i(1023);
    return y * mind_pow2i(k);
  }
  return y * mind_pow2i(k + 1000) * mind_pow2i(-1000);
}

MIND_HD double mind_log(double x) {
  const double ln2_hi = 6.93147180369123816490e-01;
  const double ln2_lo = 1.90821492927058770002e-10;
  const double two54 = 1.80143985094819840000e+16;
  const double Lg1 = 6.666666666666735130e-01;
  const double Lg2 = 3.999999999940941908e-01;
  const double Lg3 = 2.857142874366239149e-01;
  const double Lg4 = 2.222219843214978396e-01;
  const double Lg5 = 1.818357216161805012e-01;
  const double Lg6 = 1.531383769920937332e-01;
  const double Lg7 = 1.479819860511658591e-01;
  uint64_t bits = mind_bits(x);
  int32_t hx = (int32_t)(bits >> 32);
  const uint32_t lx = (uint32_t)bits;
  int k = 0;
  if (hx < 0x00100000) {  // zero, negative or subnormal
    if (((hx & 0x7fffffff) | (int32_t)lx) == 0) return -mind_from_bits(0x7ff0000000000000ull);
    if (hx < 0) return mind_from_bits(0x7ff8000000000000ull);
    k -= 54;
    x *= two54;
    bits = mind_bits(x);
    hx = (int32_t)(bits >> 32);
  }
  if (hx >= 0x7ff00000) return x + x;
  k += (hx >> 20) - 1023;
  hx &= 0x000fffff;
  const int32_t i0 = (hx + 0x95f64) & 0x100000;
  // Normalise the mantissa into [sqrt(2)/2, sqrt(2)).
  bits = (bits & 0x00000000ffffffffull) | ((uint64_t)(uint32_t)(hx | (i0 ^ 0x3ff00000)) << 32);
  x = mind_from_bits(bits);
  k += (i0 >> 20);
  const double f = x - 1.0;
  const double dk = (double)k;
  if ((0x000fffff & (2 + hx)) < 3) {  // |f| < 2^-20
    if (f == 0.0) {
      if (k == 0) return 0.0;
      return dk * ln2_hi + dk * ln2_lo;
    }
    const double R = f * f * (0.5 - 0.33333333333333333 * f);
    if (k == 0) return f - R;
    return dk * ln2_hi - ((R - dk * ln2_lo) - f);
  }
  const double s = f / (2.0 + f);
  const double z = s * s;
  int32_t i = hx - 0x6147a;
  const double w = z * z;
  const int32_t j = 0x6b851 - hx;
  const double t1 = w * (Lg2 + w * (Lg4 + w * Lg6));
  const double t2 = z * (Lg1 + w * (Lg3 + w * (Lg5 + w * Lg7)));
  i |= j;
  const double R = t2 + t1;
  if (i > 0) {
    const double hfsq = 0.5 * f * f;
    if (k == 0) return f - (hfsq - s * (hfsq + R));
    return dk * ln2_hi - ((hfsq - (s * (hfsq + R) + dk * ln2_lo)) - f);
  }
  if (k == 0) return f - s * (f - R);
  return dk * ln2_hi - ((s * (f - R) - dk * ln2_lo) - f);
}

// x^y for the temperature coefficient q10 = 3^((celsius - 6.3)/10). Exact for
// y == 0 (the default 6.3 degC); otherwise exp(y*log(x)), a few ulp, positive x.
MIND_HD double mind_pow(double x, double y) {
  if (y == 0.0) return 1.0;
  if (y == 1.0) return x;
  return mind_exp(y * mind_log(x));
}

MIND_HD double mind_fabs(double x) { return mind_from_bits(mind_bits(x) & 0x7fffffffffffffffull); }

MIND_HD int mind_isfinite(double x) {
  return (mind_bits(x) & 0x7ff0000000000000ull) != 0x7ff0000000000000ull;
}

// Nernst potential (mV) of an ion of valence z at concentrations ci, co (mM),
// and the GHK current factor (mA/cm2 per cm/s) at voltage v (mV), in the
// operation order of src/model/micro/utils/biophysics.jl:43-70 (CODATA 2018
// exact k, N_A and e). The small-argument branch x/(exp(x) - 1) ~ 1 - x/2 for
// |x| < 1e-4 is NEURON's.
MIND_HD double mind_nernst(double ci, double co, double z, double celsius) {
  if (z == 0.0) return 0.0;
  if (ci <= 0.0) return 1.0e6;
  if (co <= 0.0) return -1.0e6;
  const double gas = 1.380649e-23 * 6.02214076e23;
  const double faraday = 1.602176634e-19 * 6.02214076e23;
  return (1000.0 * gas * (celsius + 273.15) / faraday) / z * mind_log(co / ci);
}
MIND_HD double mind_ghk_quotient(double x) {
  return mind_fabs(x) < 1.0e-4 ? 1.0 - x / 2.0 : x / (mind_exp(x) - 1.0);
}
MIND_HD double mind_ghk(double v, double ci, double co, double z, double celsius) {
  const double gas = 1.380649e-23 * 6.02214076e23;
  const double faraday = 1.602176634e-19 * 6.02214076e23;
  const double thermal = 1000.0 * gas * (celsius + 273.15) / faraday;
  const double scaled = z * v / thermal;
  return 0.001 * z * faraday * (ci * mind_ghk_quotient(-scaled) - co * mind_ghk_quotient(scaled));
}
