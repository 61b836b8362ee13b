// Per-voxel log-linear initial guess, shared by pca_loglinear_init
// (init_lib.hip) and the fallback of the fused MLP inference (mlp_lib.hip).
//
// c = P log y with a voxel-uniform K x M matrix P (double), then the guess of
// the model kind:
//   kind 0  p = c                              (DKI: log-linear OLS)
//   kind 1  p = [exp(c0), 1/c1]                (T2:  log S = log S0 - TE/T2)
//   kind 2  p = [exp(c0), c1, 1]               (stretched exp, alpha = 1)
//   kind 3  segmented IVIM of test/ivim/test_ivim.ipynb cell 6 (:146-165):
//           rows (c0_lo, c1_lo, c0_hi, c1_hi) of the low-b and high-b fits,
//           p = [exp(c0_lo), 1 - exp(c0_hi)/exp(c0_lo), c1_lo, c1_hi]
// Callers apply nan_to_num and the bounds.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace pca_init {

template <typename T>
__device__ __forceinline__ T nan_to_num(T a) {
  return (__builtin_isnan(a) || __builtin_isinf(a)) ? (T)0 : a;
}

template <typename T>
__device__ __forceinline__ T t_log(T a);
template <>
__device__ __forceinline__ float t_log<float>(float a) { return logf(a); }
template <>
__device__ __forceinline__ double t_log<double>(double a) { return log(a); }
template <typename T>
__device__ __forceinline__ T t_exp(T a);
template <>
__device__ __forceinline__ float t_exp<float>(float a) { return expf(a); }
template <>
__device__ __forceinline__ double t_exp<double>(double a) { return exp(a); }

// p[0..KA) of voxel v for measurements y[i * N + v], i < M; rows k < K of P
// are read (K <= KA). The host pairs kind 2 with KA >= 3 and kind 3 with
// KA >= 4, so the size guards only remove branches a caller never takes.
template <typename T, int KA>
__device__ __forceinline__ void loglinear_guess(const double* __restrict__ P, int K, int M,
                                                const T* __restrict__ y, long long N,
                                                long long v, int kind, T (&p)[KA]) {
  double c[KA];
#pragma unroll
  for (int k = 0; k < KA; ++k) c[k] = 0.0;
  for (int i = 0; i < M; ++i) {
    const double ly = (double)t_log(y[(long long)i * N + v]);
#pragma unroll
    for (int k = 0; k < KA; ++k)
      if (k < K) c[k] += P[k * M + i] * ly;
  }
  if (kind == 0) {
#pragma unroll
    for (int k = 0; k < KA; ++k) p[k] = (T)c[k];
  } else if (kind == 1) {
    p[0] = t_exp((T)c[0]);
    p[1] = (T)1 / (T)c[1];
  } else if (kind == 2) {
    if constexpr (KA >= 3) {
      p[0] = t_exp((T)c[0]);
      p[1] = (T)c[1];
      p[2] = (T)1;
    }
  } else {
    if constexpr (KA >= 4) {
      const T s_lo = t_exp((T)c[0]);
      const T s_hi = t_exp((T)c[2]);
      p[0] = s_lo;
      p[1] = (T)1 - s_hi / s_lo;
      p[2] = (T)c[1];
      p[3] = (T)c[3];
    }
  }
}

}  // namespace pca_init
