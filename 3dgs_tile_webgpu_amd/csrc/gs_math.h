// gs_math.h - deterministic single-precision elementary functions.
//
// The reference pipeline evaluates WGSL builtins (exp, log, sin, cos, atan2,
// exp2, unpack2x16float) whose precision is implementation-defined. This
// project fixes one implementation of each, written with IEEE-754 binary32
// additions, multiplications and divisions only, so that the gfx950 kernels
// and the host code that checks them round identically. Every caller must
// compile with floating-point contraction disabled (-ffp-contract=off);
// division and square root are the correctly rounded IEEE operations on both
// sides (hipcc's default for HIP, and SSE on x86-64).
//
// Polynomials and range reductions follow the published single-precision
// Cephes library (S. L. Moshier, "Methods and Programs for Mathematical
// Functions", 1989): expf, logf, sinf/cosf (Cody-Waite pi/4 reduction with
// three-part pi/4), atanf. Accuracy is checked against libm in double
// precision by the test suite (<= 2 ulp over the ranges used).
#ifndef GS_MATH_H
#define GS_MATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define GS_HD __host__ __device__
#else
#define GS_HD
#endif

#define GS_INLINE GS_HD static inline

GS_INLINE uint32_t gs_f2b(float x) { return __builtin_bit_cast(uint32_t, x); }
GS_INLINE float gs_b2f(uint32_t b) { return __builtin_bit_cast(float, b); }

GS_INLINE float gs_floorf(float x) { return __builtin_floorf(x); }
GS_INLINE float gs_ceilf(float x) { return __builtin_ceilf(x); }
GS_INLINE float gs_fabsf(float x) { return gs_b2f(gs_f2b(x) & 0x7fffffffu); }
GS_INLINE float gs_sqrtf(float x) { return __builtin_sqrtf(x); }
GS_INLINE int gs_isnan(float x) { return x != x; }

// WGSL min/max/clamp on floats. Written as selects so both sides agree on
// NaN propagation: the non-NaN operand wins in min/max, matching IEEE
// minNum/maxNum, which is what GPUs implement.
GS_INLINE float gs_minf(float a, float b) {
  if (a != a) return b;
  if (b != b) return a;
  return a < b ? a : b;
}
GS_INLINE float gs_maxf(float a, float b) {
  if (a != a) return b;
  if (b != b) return a;
  return a > b ? a : b;
}
GS_INLINE float gs_clampf(float x, float lo, float hi) {
  return gs_minf(gs_maxf(x, lo), hi);
}

// 2^k as an exact binary32 for k in [-126, 127].
GS_INLINE float gs_pow2i_normal(int k) {
  return gs_b2f((uint32_t)(k + 127) << 23);
}

// x * 2^n with one rounding, for n in [-252, 254]: the first product is exact
// whenever |x| >= 2^-51, the second carries the only rounding.
GS_INLINE float gs_ldexpf(float x, int n) {
  int n1 = n / 2;
  int n2 = n - n1;
  return (x * gs_pow2i_normal(n1)) * gs_pow2i_normal(n2);
}

// b / d correctly rounded, for an integer b and a constant integer divisor
// d as the packed formats use them (b in [-128, 127] for d = 127, in
// [0, 1023] for d = 1023, in [0, 4095] for d = 4095): q1 = b * (1/d), then
// one correction with the exact remainder fma(-q1, d, b) (Markstein). For
// these operands it equals the IEEE quotient b / d for every b (checked over
// all of them by the test suite), at one product and two fused
// multiply-adds instead of a full division.
GS_INLINE float gs_div_small(float b, float d, float inv_d) {
  float q1 = b * inv_d;
  float r = __builtin_fmaf(-q1, d, b);
  return __builtin_fmaf(r, inv_d, q1);
}

// Exact 2^e for an integer exponent e in [-149, 128] (2^128 is +inf). Used by
// the rgb8e8 SH decode, where e = byte - 127 lies in [-127, 128].
GS_INLINE float gs_exp2i(int e) {
  if (e >= 128) return gs_b2f(0x7f800000u);
  if (e >= -126) return gs_pow2i_normal(e);
  if (e < -149) return 0.0f;
  return gs_b2f(1u << (e + 149));
}

// Cephes expf reduction and polynomial: e^x = y 2^n.
GS_INLINE float gs_expf_reduced(float x, int& n) {
  float z = gs_floorf(1.44269504088896341f * x + 0.5f);
  x = x - z * 0.693359375f;
  x = x - z * -2.12194440e-4f;
  n = (int)z;
  z = x * x;
  float y = 1.9875691500E-4f;
  y = y * x + 1.3981999507E-3f;
  y = y * x + 8.3334519073E-3f;
  y = y * x + 4.1665795894E-2f;
  y = y * x + 1.6666665459E-1f;
  y = y * x + 5.0000001201E-1f;
  y = y * z + x + 1.0f;
  return y;
}

// Cephes expf core for x in [-103.3, 88.7]; no special-value handling.
GS_INLINE float gs_expf_core(float x) {
  int n;
  const float y = gs_expf_reduced(x, n);
  return gs_ldexpf(y, n);
}

GS_INLINE float gs_expf(float x) {
  if (x != x) return x;
  if (x > 88.72283905206835f) return gs_b2f(0x7f800000u);
  if (x < -103.278929903431851103f) return 0.0f;
  return gs_expf_core(x);
}

// gs_expf for x in [-87, 0] or NaN, without the range branches: there n is
// in [-126, 0], 2^n a normal binary32, and y 2^n one product with the same
// single rounding as gs_ldexpf's two. The rasterizer's candidate powers lie
// in [-ln(255 a), 0], a the opacity (at most 65504 in binary16), so inside.
GS_INLINE float gs_expf_nonpos(float x) {
  const bool nan = x != x;
  int n;
  const float y = gs_expf_reduced(nan ? 0.0f : x, n);
  return nan ? x : y * gs_pow2i_normal(n);
}

// Cephes logf.
GS_INLINE float gs_logf(float x) {
  if (x != x) return x;
  if (x < 0.0f) return gs_b2f(0x7fc00000u);
  if (x == 0.0f) return gs_b2f(0xff800000u);
  uint32_t ix = gs_f2b(x);
  if (ix == 0x7f800000u) return x;
  int e = 0;
  if (ix < 0x00800000u) {  // subnormal: scale into the normal range
    x = x * 16777216.0f;
    ix = gs_f2b(x);
    e = -24;
  }
  e += (int)((ix >> 23) & 0xffu) - 126;
  x = gs_b2f((ix & 0x007fffffu) | 0x3f000000u);  // mantissa in [0.5, 1)
  if (x < 0.70710678118654752440f) {
    e -= 1;
    x = x + x - 1.0f;
  } else {
    x = x - 1.0f;
  }
  float z = x * x;
  float y = 7.0376836292E-2f;
  y = y * x - 1.1514610310E-1f;
  y = y * x + 1.1676998740E-1f;
  y = y * x - 1.2420140846E-1f;
  y = y * x + 1.4249322787E-1f;
  y = y * x - 1.6668057665E-1f;
  y = y * x + 2.0000714765E-1f;
  y = y * x - 2.4999993993E-1f;
  y = y * x + 3.3333331174E-1f;
  y = y * x * z;
  float fe = (float)e;
  y = y + -2.12194440e-4f * fe;
  y = y + -0.5f * z;
  z = x + y;
  z = z + 0.693359375f * fe;
  return z;
}

// Natural log of a positive normal binary32, without the special values:
// the Cephes logf reduction and polynomial above, evaluated with fused
// multiply-adds (correctly rounded on both sides, so host and device agree
// bit for bit) and no branches. Used by the learned-cull features only,
// whose arguments are positive and normal by construction; the reference's
// own arithmetic keeps gs_logf.
GS_INLINE float gs_logf_pos(float x) {
  uint32_t ix = gs_f2b(x);
  int e = (int)((ix >> 23) & 0xffu) - 126;
  float m = gs_b2f((ix & 0x007fffffu) | 0x3f000000u);  // in [0.5, 1)
  bool low = m < 0.70710678118654752440f;
  e -= low ? 1 : 0;
  m = low ? (m + m) - 1.0f : m - 1.0f;
  float z = m * m;
  float y = 7.0376836292E-2f;
  y = __builtin_fmaf(y, m, -1.1514610310E-1f);
  y = __builtin_fmaf(y, m, 1.1676998740E-1f);
  y = __builtin_fmaf(y, m, -1.2420140846E-1f);
  y = __builtin_fmaf(y, m, 1.4249322787E-1f);
  y = __builtin_fmaf(y, m, -1.6668057665E-1f);
  y = __builtin_fmaf(y, m, 2.0000714765E-1f);
  y = __builtin_fmaf(y, m, -2.4999993993E-1f);
  y = __builtin_fmaf(y, m, 3.3333331174E-1f);
  y = y * m * z;
  float fe = (float)e;
  y = __builtin_fmaf(-2.12194440e-4f, fe, y);
  y = __builtin_fmaf(-0.5f, z, y);
  return __builtin_fmaf(0.693359375f, fe, m + y);
}

// Cephes sinf/cosf share one reduction to [-pi/4, pi/4]; valid for
// |x| < 8192 (beyond that the reduction loses all bits; not used here).
GS_INLINE float gs_sincos_poly_sin(float x, float z) {
  float y = -1.9515295891E-4f;
  y = y * z + 8.3321608736E-3f;
  y = y * z - 1.6666654611E-1f;
  y = y * z * x;
  return y + x;
}
GS_INLINE float gs_sincos_poly_cos(float z) {
  float y = 2.443315711809948E-005f;
  y = y * z - 1.388731625493765E-003f;
  y = y * z + 4.166664568298827E-002f;
  y = y * z * z;
  y = y - 0.5f * z;
  return y + 1.0f;
}

GS_INLINE float gs_sinf(float xx) {
  if (xx != xx) return xx;
  int negative = 0;
  float x = xx;
  if (x < 0.0f) {
    negative = 1;
    x = -x;
  }
  int j = (int)(1.27323954473516f * x);
  float y = (float)j;
  if (j & 1) {
    j += 1;
    y += 1.0f;
  }
  j &= 7;
  if (j > 3) {
    negative = !negative;
    j -= 4;
  }
  x = ((x - y * 0.78515625f) - y * 2.4187564849853515625e-4f) -
      y * 3.77489497744594108e-8f;
  float z = x * x;
  float r = (j == 1 || j == 2) ? gs_sincos_poly_cos(z)
                               : gs_sincos_poly_sin(x, z);
  return negative ? -r : r;
}

GS_INLINE float gs_cosf(float xx) {
  if (xx != xx) return xx;
  float x = gs_fabsf(xx);
  int j = (int)(1.27323954473516f * x);
  float y = (float)j;
  if (j & 1) {
    j += 1;
    y += 1.0f;
  }
  j &= 7;
  int negative = 0;
  if (j > 3) {
    j -= 4;
    negative = !negative;
  }
  if (j > 1) negative = !negative;
  x = ((x - y * 0.78515625f) - y * 2.4187564849853515625e-4f) -
      y * 3.77489497744594108e-8f;
  float z = x * x;
  float r = (j == 1 || j == 2) ? gs_sincos_poly_sin(x, z)
                               : gs_sincos_poly_cos(z);
  return negative ? -r : r;
}

// Cephes atanf.
GS_INLINE float gs_atanf(float xx) {
  if (xx != xx) return xx;
  int negative = xx < 0.0f;
  float x = gs_fabsf(xx);
  float y;
  if (x > 2.414213562373095f) {
    y = 1.5707963267948966192f;
    x = -(1.0f / x);
  } else if (x > 0.4142135623730950f) {
    y = 0.7853981633974483096f;
    x = (x - 1.0f) / (x + 1.0f);
  } else {
    y = 0.0f;
  }
  float z = x * x;
  float p = 8.05374449538e-2f;
  p = p * z - 1.38776856032E-1f;
  p = p * z + 1.99777106478E-1f;
  p = p * z - 3.33329491539E-1f;
  y = y + (p * z * x + x);
  return negative ? -y : y;
}

// atan2 with the C99 conventions for signed zeros.
GS_INLINE float gs_atan2f(float y, float x) {
  if (x != x || y != y) return x + y;
  const float pi = 3.14159265358979323846f;
  const float pio2 = 1.5707963267948966192f;
  int xneg = (gs_f2b(x) >> 31) != 0;
  if (y == 0.0f) {
    if (xneg) return (gs_f2b(y) >> 31) ? -pi : pi;
    return y;
  }
  if (x == 0.0f) return y > 0.0f ? pio2 : -pio2;
  float z = gs_atanf(y / x);
  if (x < 0.0f) z = (y < 0.0f) ? z - pi : z + pi;
  return z;
}

// IEEE binary16 -> binary32 (exact), as WGSL unpack2x16float.
GS_INLINE float gs_half_to_float(uint32_t h) {
  uint32_t sign = (h & 0x8000u) << 16;
  uint32_t exponent = (h >> 10) & 0x1fu;
  uint32_t mantissa = h & 0x3ffu;
  if (exponent == 0x1fu)
    return gs_b2f(sign | 0x7f800000u | (mantissa << 13));
  if (exponent != 0)
    return gs_b2f(sign | ((exponent + 112u) << 23) | (mantissa << 13));
  // zero or subnormal: mantissa * 2^-24, exact in binary32
  float v = (float)mantissa * 5.9604644775390625e-8f;
  return (sign != 0) ? -v : v;
}

// binary32 -> binary16 with round-to-nearest-even (IEEE), as a textureStore
// of an rgba16float target and as the producer's f16::from_f32.
GS_INLINE uint32_t gs_float_to_half(float f) {
  uint32_t x = gs_f2b(f);
  uint32_t sign = (x >> 16) & 0x8000u;
  uint32_t exponent = (x >> 23) & 0xffu;
  uint32_t mantissa = x & 0x7fffffu;
  if (exponent == 0xffu)
    return sign | 0x7c00u | (mantissa != 0 ? (0x200u | (mantissa >> 13)) : 0u);
  int e = (int)exponent - 127 + 15;
  if (e >= 31) return sign | 0x7c00u;
  if (e <= 0) {
    if (e < -10) return sign;  // rounds to zero
    mantissa |= 0x800000u;
    uint32_t shift = (uint32_t)(14 - e);
    uint32_t half_mantissa = mantissa >> shift;
    uint32_t remainder = mantissa & ((1u << shift) - 1u);
    uint32_t halfway = 1u << (shift - 1u);
    if (remainder > halfway || (remainder == halfway && (half_mantissa & 1u)))
      half_mantissa += 1u;
    return sign | half_mantissa;
  }
  uint32_t half = sign | ((uint32_t)e << 10) | (mantissa >> 13);
  uint32_t remainder = mantissa & 0x1fffu;
  if (remainder > 0x1000u || (remainder == 0x1000u && (half & 1u))) half += 1u;
  return half;  // a carry out of the mantissa correctly bumps the exponent
}

// WGSL f32 -> i32 / u32 conversion: truncate toward zero, saturate, NaN -> 0.
GS_INLINE int32_t gs_f2i(float x) {
  if (x != x) return 0;
  if (x >= 2147483648.0f) return 2147483647;
  if (x <= -2147483648.0f) return (int32_t)0x80000000u;
  return (int32_t)x;
}
GS_INLINE uint32_t gs_f2u(float x) {
  if (x != x || x <= 0.0f) return 0u;
  if (x >= 4294967296.0f) return 0xffffffffu;
  return (uint32_t)x;
}

#endif  // GS_MATH_H
