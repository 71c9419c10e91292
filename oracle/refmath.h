// Oracle (test infrastructure only): single-precision exp/log used by the CPU
// restatement of the DIMCV-SPH step.
//
// The reference evaluates ti.exp / ti.pow in f32 (dimcv_sph.py:298-303,
// :655-656, :703-705).  Their last-ulp behaviour depends on the Taichi
// backend and cannot be pinned here, so the oracle fixes one definition:
//   exp: Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, degree-8 Taylor
//        polynomial for e^r (Horner, no fused operations), scaling by 2^k
//        through exponent-bit construction;
//   log: x = m 2^e with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s),
//        s = (m-1)/(m+1), odd series to s^9;
//   pow(x, y) = exp(y log x)  (x >= 1 on every call site).
// Every operation is a single IEEE-754 f32 operation; the file must be built
// with -ffp-contract=off.  tests/test_oracle_math.py bounds the error against
// the C library (<= 2 ulp over the ranges the step uses).
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace refmath {

static inline float bits_to_f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
static inline uint32_t f_to_bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

// 2^k for k in [-126, 127] as an exact f32.
static inline float pow2i(int k) { return bits_to_f((uint32_t)(k + 127) << 23); }

static inline float expf_c(float x) {
  if (x != x) return x;
  if (x > 88.72283935546875f) return INFINITY;
  if (x < -103.97208404541015625f) return 0.0f;
  const float log2e = 1.44269502162933349609375f;
  const float ln2_hi = 0.693145751953125f;          // 0x3f317200, 15 significant bits
  const float ln2_lo = 1.428606765330187045e-06f;   // ln2 - ln2_hi
  float kf = std::floor(x * log2e + 0.5f);
  float r = (x - kf * ln2_hi) - kf * ln2_lo;
  float p = 2.4801587301587302e-05f;                 // 1/8!
  p = p * r + 1.9841269841269841e-04f;               // 1/7!
  p = p * r + 1.3888888888888889e-03f;               // 1/6!
  p = p * r + 8.3333333333333332e-03f;               // 1/5!
  p = p * r + 4.1666666666666664e-02f;               // 1/4!
  p = p * r + 1.6666666666666666e-01f;               // 1/3!
  p = p * r + 0.5f;
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  int k = (int)kf;
  if (k > 127) return (p * pow2i(127)) * pow2i(k - 127);
  if (k < -126) return (p * pow2i(k + 64)) * pow2i(-64);
  return p * pow2i(k);
}

static inline float logf_c(float x) {
  if (x != x) return x;
  if (x < 0.0f) return NAN;
  if (x == 0.0f) return -INFINITY;
  if (x == INFINITY) return INFINITY;
  int e = 0;
  if (x < 1.17549435e-38f) { x = x * 8388608.0f; e = -23; }   // lift subnormals
  uint32_t u = f_to_bits(x);
  e += (int)((u >> 23) & 0xffu) - 127;
  float m = bits_to_f((u & 0x007fffffu) | 0x3f800000u);       // [1, 2)
  if (m > 1.41421353816986083984375f) { m = m * 0.5f; e += 1; }
  float s = (m - 1.0f) / (m + 1.0f);
  float s2 = s * s;
  float t = 0.1111111111111111f;                               // 1/9
  t = t * s2 + 0.14285714285714285f;                           // 1/7
  t = t * s2 + 0.2f;
  t = t * s2 + 0.33333333333333331f;
  t = t * s2 + 1.0f;
  float lm = (2.0f * s) * t;
  float ef = (float)e;
  return ef * 0.693145751953125f + (ef * 1.428606765330187045e-06f + lm);
}

static inline float powf_c(float x, float y) { return expf_c(y * logf_c(x)); }

}  // namespace refmath
