// Initial-guess kernels, compiled into libpycompute_amd.so for gfx950.
//
// pca_bpolyfit: batched polynomial least squares of pycompute/cuda/lsq.py:22-30
//   out[:, v] = argmin_c || A_v c - y_v ||,  A_v = [1, x_v, ..., x_v^deg]
// The reference forms pinv(A) with a batched SVD; for the full-rank systems it
// is used on, that equals the least-squares solution computed here from the
// normal equations in double precision (Cholesky, deg <= 7).
//
// pca_loglinear_init: out = T(P log y) per voxel with a voxel-uniform matrix P
// (K x M, double). With P the pseudo-inverse of a shared design this is the
// same least-squares fit as bpolyfit on replicated x, and covers the default
// initial guesses of the built-in models:
//   kind 0  p = P log y                       (DKI: log-linear OLS)
//   kind 1  p = [exp(c0), 1/c1]               (T2:  log S = log S0 - TE/T2)
//   kind 2  p = [exp(c0), c1, 1]              (stretched exp, alpha = 1)
//   kind 3  segmented IVIM of test/ivim/test_ivim.ipynb cell 6 (:146-165):
//           rows (c0_lo, c1_lo, c0_hi, c1_hi) of the low-b and high-b fits,
//           p = [exp(c0_lo), 1 - exp(c0_hi)/exp(c0_lo), c1_lo, c1_hi]
// followed by nan_to_num (NaN and +-Inf to 0) and clipping to [lb, ub], as the
// notebook does.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "pycompute_amd.h"
#include "internal.h"
#include "loglinear.h"

namespace {

constexpr int kMaxDeg = 7;
constexpr int kMaxK = 32;

template <typename T>
__global__ void __launch_bounds__(256)
k_bpolyfit(const T* __restrict__ x, int x_per_voxel, const T* __restrict__ y, int m, int deg,
           T* __restrict__ out, long long N) {
  const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= N) return;
  const int K = deg + 1;
  double G[(kMaxDeg + 1) * (kMaxDeg + 1)];
  double r[kMaxDeg + 1];
  for (int a = 0; a < K * K; ++a) G[a] = 0.0;
  for (int a = 0; a < K; ++a) r[a] = 0.0;
  for (int i = 0; i < m; ++i) {
    const double xi = (double)(x_per_voxel ? x[(long long)i * N + v] : x[i]);
    const double yi = (double)y[(long long)i * N + v];
    double pa = 1.0;
    for (int a = 0; a < K; ++a) {
      r[a] += pa * yi;
      double pb = 1.0;
      for (int b = 0; b <= a; ++b) {
        G[a * K + b] += pa * pb;
        pb *= xi;
      }
      pa *= xi;
    }
  }
  // Cholesky G = L L^T (lower), then two triangular solves.
  for (int j = 0; j < K; ++j) {
    double d = G[j * K + j];
    for (int q = 0; q < j; ++q) d -= G[j * K + q] * G[j * K + q];
    d = sqrt(d);
    G[j * K + j] = d;
    for (int i = j + 1; i < K; ++i) {
      double t = G[i * K + j];
      for (int q = 0; q < j; ++q) t -= G[i * K + q] * G[j * K + q];
      G[i * K + j] = t / d;
    }
  }
  for (int i = 0; i < K; ++i) {
    double t = r[i];
    for (int q = 0; q < i; ++q) t -= G[i * K + q] * r[q];
    r[i] = t / G[i * K + i];
  }
  for (int i = K - 1; i >= 0; --i) {
    double t = r[i];
    for (int q = i + 1; q < K; ++q) t -= G[q * K + i] * r[q];
    r[i] = t / G[i * K + i];
  }
  for (int a = 0; a < K; ++a) out[(long long)a * N + v] = (T)r[a];
}

// KC: compile-time size of the per-voxel arrays, max(rows of P, outputs)
// (2, 3 or 4 for the built-in models; 0 = generic, up to kMaxK). The rows of
// P actually read are Kr.
template <typename T, int KC>
__global__ void __launch_bounds__(256)
k_loglinear_init(const double* __restrict__ P, int Kr, int M, const T* __restrict__ y, int kind,
                 int n, const T* __restrict__ lb, const T* __restrict__ ub, T* __restrict__ out,
                 long long N) {
  const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= N) return;
  constexpr int KA = KC > 0 ? KC : kMaxK;
  T p[KA];
  pca_init::loglinear_guess<T, KA>(P, Kr, M, y, N, v, kind, p);
#pragma unroll
  for (int j = 0; j < KA; ++j) {
    if (j < n) {
      T a = pca_init::nan_to_num(p[j]);
      const T lo = lb[(long long)j * N + v], hi = ub[(long long)j * N + v];
      a = a < lo ? lo : a;
      a = a > hi ? hi : a;
      out[(long long)j * N + v] = a;
    }
  }
}

// Canonical IVIM labelling (test/ivim/test_ivim.ipynb:441-446): the model is
// symmetric under (f, D_1, D_2) -> (1 - f, D_2, D_1); where D_1 < D_2 the two
// diffusivities are swapped and f becomes 1 - f, so D_1 (pseudo-diffusion) >=
// D_2 everywhere. params (4, N): S0, f, D_1, D_2.
template <typename T>
__global__ void __launch_bounds__(256) k_ivim_canonicalize(T* p, long long N) {
  const long long v = (long long)blockIdx.x * 256 + threadIdx.x;
  if (v >= N) return;
  const T d1 = p[2 * N + v], d2 = p[3 * N + v];
  if (d1 < d2) {
    p[2 * N + v] = d2;
    p[3 * N + v] = d1;
    p[N + v] = (T)1 - p[N + v];
  }
}

}  // namespace

extern "C" int pca_ivim_canonicalize(void* params, int64_t N, int32_t is_double, void* stream) {
  if (N == 0) return PCA_OK;   // no voxels: empty arrays may have no address
  if (!params || N < 0) return pca_fail(PCA_ERR_ARG, "pca_ivim_canonicalize: bad argument");
  const dim3 grid((unsigned)((N + 255) / 256)), block(256);
  if (is_double)
    hipLaunchKernelGGL(k_ivim_canonicalize<double>, grid, block, 0, (hipStream_t)stream,
                       (double*)params, (long long)N);
  else
    hipLaunchKernelGGL(k_ivim_canonicalize<float>, grid, block, 0, (hipStream_t)stream,
                       (float*)params, (long long)N);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? PCA_OK : pca_fail(PCA_ERR_HIP, hipGetErrorString(e));
}

extern "C" int pca_bpolyfit(const void* x, int32_t x_per_voxel, const void* y, int32_t m,
                            int32_t deg, void* out, int64_t N, int32_t is_double, void* stream) {
  if (m <= 0 || deg < 0 || deg > kMaxDeg || N < 0)
    return pca_fail(PCA_ERR_ARG, "pca_bpolyfit: bad argument (deg must be 0..7)");
  if (N == 0) return PCA_OK;   // no voxels: empty arrays may have no address
  if (!x || !y || !out) return pca_fail(PCA_ERR_ARG, "pca_bpolyfit: NULL argument");
  const dim3 grid((unsigned)((N + 255) / 256)), block(256);
  if (is_double)
    hipLaunchKernelGGL(k_bpolyfit<double>, grid, block, 0, (hipStream_t)stream,
                       (const double*)x, x_per_voxel, (const double*)y, m, deg, (double*)out,
                       (long long)N);
  else
    hipLaunchKernelGGL(k_bpolyfit<float>, grid, block, 0, (hipStream_t)stream, (const float*)x,
                       x_per_voxel, (const float*)y, m, deg, (float*)out, (long long)N);
  {
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? PCA_OK : pca_fail(PCA_ERR_HIP, hipGetErrorString(e));
  }
}

extern "C" int pca_loglinear_init(const double* P, int32_t K, int32_t M, const void* y,
                                  int32_t kind, int32_t n, const void* lb, const void* ub,
                                  void* out, int64_t N, int32_t is_double, void* stream) {
  if (!P || K <= 0 || K > kMaxK || M <= 0 || n <= 0 || n > kMaxK || kind < 0 || kind > 3 ||
      N < 0)
    return pca_fail(PCA_ERR_ARG, "pca_loglinear_init: bad argument");
  if ((kind == 1 && (K != 2 || n != 2)) || (kind == 2 && (K != 2 || n != 3)) ||
      (kind == 3 && (K != 4 || n != 4)) || (kind == 0 && K != n))
    return pca_fail(PCA_ERR_ARG, "pca_loglinear_init: K/n do not match the init kind");
  if (N == 0) return PCA_OK;   // no voxels: empty arrays may have no address
  if (!y || !lb || !ub || !out) return pca_fail(PCA_ERR_ARG, "pca_loglinear_init: NULL argument");
  const dim3 grid((unsigned)((N + 255) / 256)), block(256);
  hipStream_t st = (hipStream_t)stream;
#define PCA_INIT(TT, KK)                                                                      \
  hipLaunchKernelGGL((k_loglinear_init<TT, KK>), grid, block, 0, st, P, K, M, (const TT*)y,  \
                     kind, n, (const TT*)lb, (const TT*)ub, (TT*)out, (long long)N)
  // kind 2 (stretched exponential) writes 3 parameters from 2 rows of P
  const int KS = kind == 2 ? 3 : K;
  if (is_double) {
    if (KS == 4) PCA_INIT(double, 4);
    else if (KS == 3) PCA_INIT(double, 3);
    else if (KS == 2) PCA_INIT(double, 2);
    else PCA_INIT(double, 0);
  } else {
    if (KS == 4) PCA_INIT(float, 4);
    else if (KS == 3) PCA_INIT(float, 3);
    else if (KS == 2) PCA_INIT(float, 2);
    else PCA_INIT(float, 0);
  }
#undef PCA_INIT
  {
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? PCA_OK : pca_fail(PCA_ERR_HIP, hipGetErrorString(e));
  }
}
