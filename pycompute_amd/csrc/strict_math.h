// Reproducible exp / log / pow for the strict build.
//
// These are written so that the GPU (hiprtc, gfx950) and a host compiler
// (g++, x86-64 SSE) produce bit-identical results: only IEEE-754 operations
// that both sides round the same way are used (+ - * / with round-to-nearest,
// rint, ldexp, frexp), no fused multiply-add (the strict build compiles with
// -ffp-contract=off on both sides), and the same operation order.
//
// Method (Cody and Waite, "Software Manual for the Elementary Functions",
// 1980): argument reduction by a two-part ln 2 whose high part has trailing
// zero bits, so k*ln2_hi is exact; then
//   exp(r), |r| <= ln2/2 :  Taylor series sum r^n/n! in Horner form
//                          (degree 7 for float, 13 for double)
//   log(m), m in [sqrt(1/2), sqrt(2)) : 2 atanh(s), s = (m-1)/(m+1),
//                          odd series s^(2k+1)/(2k+1) (to s^13 / s^27)
//   pow(x, y) = exp(y log x) for x > 0.
// Accuracy is a few ulp, which is all the strict build needs: its purpose is
// an oracle-reproducible trajectory, not the last ulp.
#pragma once

#if defined(__HIPCC__) || defined(__HIPCC_RTC__)
#define PCA_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define PCA_HD inline
#endif

PCA_HD float pca_exp_strict(float x) {
  if (x != x) return x;
  if (x > 89.0f) return __builtin_inff();
  if (x < -104.0f) return 0.0f;
  const float k = rintf(x * 0x1.715476p+0f);  // x log2(e)
  const float r = (x - k * 0x1.62e4p-1f) - k * 0x1.7f7d1cp-20f;
  float p = 0x1.a01a02p-13f;  // 1/7!
  p = p * r + 0x1.6c16c2p-10f;  // 1/6!
  p = p * r + 0x1.111112p-7f;   // 1/5!
  p = p * r + 0x1.555556p-5f;   // 1/4!
  p = p * r + 0x1.555556p-3f;   // 1/3!
  p = p * r + 0.5f;
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  return ldexpf(p, (int)k);
}

PCA_HD double pca_exp_strict(double x) {
  if (x != x) return x;
  if (x > 710.0) return __builtin_inf();
  if (x < -746.0) return 0.0;
  const double k = rint(x * 0x1.71547652b82fep+0);
  const double r = (x - k * 0x1.62e42feep-1) - k * 0x1.a39ef35793c76p-33;
  // 1/n! for n = 13 .. 2
  const double c[12] = {1.0 / 6227020800.0, 1.0 / 479001600.0, 1.0 / 39916800.0,
                        1.0 / 3628800.0,    1.0 / 362880.0,    1.0 / 40320.0,
                        1.0 / 5040.0,       1.0 / 720.0,       1.0 / 120.0,
                        1.0 / 24.0,         1.0 / 6.0,         0.5};
  double p = c[0];
  for (int i = 1; i < 12; ++i) p = p * r + c[i];
  p = p * r + 1.0;
  p = p * r + 1.0;
  return ldexp(p, (int)k);
}

PCA_HD float pca_log_strict(float x) {
  if (x != x || x < 0.0f) return __builtin_nanf("");
  if (x == 0.0f) return -__builtin_inff();
  if (x == __builtin_inff()) return x;
  int e;
  float m = frexpf(x, &e);
  if (m < 0x1.6a09e6p-1f) {  // sqrt(1/2)
    m = m * 2.0f;
    e = e - 1;
  }
  const float s = (m - 1.0f) / (m + 1.0f);
  const float z = s * s;
  float q = 1.0f / 13.0f;
  q = q * z + 1.0f / 11.0f;
  q = q * z + 1.0f / 9.0f;
  q = q * z + 1.0f / 7.0f;
  q = q * z + 1.0f / 5.0f;
  q = q * z + 1.0f / 3.0f;
  q = q * z + 1.0f;
  const float lm = 2.0f * (s * q);
  const float fe = (float)e;
  return fe * 0x1.62e4p-1f + (fe * 0x1.7f7d1cp-20f + lm);
}

PCA_HD double pca_log_strict(double x) {
  if (x != x || x < 0.0) return __builtin_nan("");
  if (x == 0.0) return -__builtin_inf();
  if (x == __builtin_inf()) return x;
  int e;
  double m = frexp(x, &e);
  if (m < 0x1.6a09e667f3bcdp-1) {
    m = m * 2.0;
    e = e - 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double z = s * s;
  double q = 1.0 / 27.0;
  for (int k = 25; k >= 3; k -= 2) q = q * z + 1.0 / (double)k;
  q = q * z + 1.0;
  const double lm = 2.0 * (s * q);
  const double fe = (double)e;
  return fe * 0x1.62e42feep-1 + (fe * 0x1.a39ef35793c76p-33 + lm);
}

PCA_HD float pca_pow_strict(float x, float y) {
  if (x != x || y != y) return x + y;
  if (x == 0.0f) return y > 0.0f ? 0.0f : (y == 0.0f ? 1.0f : __builtin_inff());
  if (x < 0.0f) return __builtin_nanf("");
  return pca_exp_strict(y * pca_log_strict(x));
}

PCA_HD double pca_pow_strict(double x, double y) {
  if (x != x || y != y) return x + y;
  if (x == 0.0) return y > 0.0 ? 0.0 : (y == 0.0 ? 1.0 : __builtin_inf());
  if (x < 0.0) return __builtin_nan("");
  return pca_exp_strict(y * pca_log_strict(x));
}
