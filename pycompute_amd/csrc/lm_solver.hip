// Shared LM pieces: the argument block of a fit and, in namespace pca, the
// reference's linear solve (pivot, scaling, GMW81, LDL^T substitution) for a
// compile-time parameter count, and the settings conversion. Included by the
// generated-model kernels (lm_kernel.hip) and the plugin kernel
// (plugin_kernel.hip).

struct pca_lm_args {
  const real* data;      // (M, N)
  const real* consts;    // (NC*M, N) when per-voxel, else unused
  const real* lb;        // (NP, N)
  const real* ub;        // (NP, N)
  real* params;          // (NP, N) in/out
  real* lam;             // (N) in/out
  signed char* step_type;  // (N) in/out
  real* f;               // (N) in/out: SSR at the start of the last executed iteration
  real* first_f;         // (N) out or null
  real* last_f;          // (N) out or null
  int* n_iter;           // (N) += executed iterations, or null
  real* g_out;           // (NP, N) or null
  real* h_out;           // (NH, N) or null
  real* hl_out;          // (NH, N) or null
  real* step_out;        // (NP, N) or null
  long long N;
  int iters;
  int reserved;
  double tol, mu, eta, acc, dec;
};

#ifndef PCA_UC_STRUCT_DEFINED
#define PCA_UC_STRUCT_DEFINED
struct pca_uconsts {
  real c[PCA_UC_MAX];
};
#endif

namespace pca {

__host__ __device__ constexpr int tri(int i, int j) { return i * (i + 1) / 2 + j; }

// The reference's diagonal pivot never updates its running maximum
// (pycompute/cuda/permute.py:50-56), so for finite diagonals it always yields
// the rotation perm = [n-1, 0, 1, ..., n-2]. It is applied here at compile time.
// With a NaN diagonal every entry of the reference's solution is NaN whatever
// the pivot, and the NaN scrub zeroes the step either way.
__host__ __device__ constexpr int perm(int n, int i) { return i == 0 ? n - 1 : i - 1; }

// Solve hl s = g as pycompute/cuda/solver.py:428-488 does (pivot, diagonal
// scaling, GMW81 modified Cholesky solver.py:20-115, LDL^T solve
// solver.py:295-401, unscale, unpivot). hl is given as h (off-diagonal,
// bit-identical to the reference's separately accumulated hl) and hld (its
// damped diagonal). All indices are compile-time after unrolling, so the
// matrix lives in VGPRs.
template <int n>
__device__ __forceinline__ void gmw81_solve(const real* h, const real* hld, const real* g,
                                            real* s) {
  // PCA_FAST_SOLVE (fast builds): every division becomes a multiplication by
  // a hardware reciprocal computed once per divisor (1/sc_i = rsqrt|B_ii|,
  // 1/d_j, 1/beta^2); results equal up to rounding.
#ifdef PCA_FAST_SOLVE
  constexpr bool FAST = true;
#else
  constexpr bool FAST = false;
#endif
  constexpr int nh = n * (n + 1) / 2;
  real B[nh], b[n], sc[n];
#pragma unroll
  for (int i = 0; i < n; ++i) {
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      const int pi = perm(n, i), pj = perm(n, j);
      B[tri(i, j)] = (pi == pj) ? hld[pi] : (pi > pj ? h[tri(pi, pj)] : h[tri(pj, pi)]);
    }
    b[i] = g[perm(n, i)];
  }
#pragma unroll
  for (int i = 0; i < n; ++i) {
    if constexpr (FAST) {
      sc[i] = pca_rcp_fast(pca_sqrt_fast(pca_abs(B[tri(i, i)])));   // 1 / scale
      b[i] *= sc[i];
    } else {
      sc[i] = pca_sqrt(pca_abs(B[tri(i, i)]));
      b[i] /= sc[i];
    }
  }
#pragma unroll
  for (int i = 0; i < n; ++i)
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      if constexpr (FAST)
        B[tri(i, j)] *= (sc[i] * sc[j]);
      else
        B[tri(i, j)] /= (sc[i] * sc[j]);
    }

  // GMW81. The running maxima are fmax: every candidate is an absolute value
  // and the maxima start at 0, so fmax equals the reference's compare-and-
  // assign bit for bit, a NaN candidate included (both keep the old maximum).
  real gamma = 0, xi = 0;
#pragma unroll
  for (int i = 0; i < n; ++i)
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      const real t = pca_abs(B[tri(i, j)]);
      if (i == j)
        gamma = pca_fmax(gamma, t);
      else
        xi = pca_fmax(xi, t);
    }
  if (n > 1) {
    if constexpr (FAST)
      xi *= (real)1 / pca_sqrt((real)(n * n - 1));
    else
      xi /= pca_sqrt((real)(n * n - 1));
  }
  real beta2 = PCA_EPS_GMW, delta = PCA_EPS_GMW;
  beta2 = pca_fmax(pca_fmax(beta2, gamma), xi);
  if (gamma + xi > (real)1) delta *= (gamma + xi);
  real rbeta2 = 0, rd[n];
  if constexpr (FAST) rbeta2 = pca_rcp_fast(beta2);

#pragma unroll
  for (int j = 0; j < n; ++j) {
#pragma unroll
    for (int q = 0; q < j; ++q) {
      if constexpr (FAST)
        B[tri(j, q)] *= rd[q];
      else
        B[tri(j, q)] /= B[tri(q, q)];
    }
#pragma unroll
    for (int i = j + 1; i < n; ++i) {
      real t = B[tri(i, j)];
#pragma unroll
      for (int q = 0; q < j; ++q) t -= B[tri(j, q)] * B[tri(i, q)];
      B[tri(i, j)] = t;
    }
    real theta = 0;
#pragma unroll
    for (int i = j + 1; i < n; ++i) theta = pca_fmax(theta, pca_abs(B[tri(i, j)]));
    theta *= theta;
    real d = pca_abs(B[tri(j, j)]);
    if (d < delta) d = delta;
    const real t0 = FAST ? theta * rbeta2 : theta / beta2;
    if (d < t0) d = t0;
    B[tri(j, j)] = d;
    if constexpr (FAST) rd[j] = pca_rcp_fast(d);
#pragma unroll
    for (int i = j + 1; i < n; ++i) {
      const real t = B[tri(i, j)];
      if constexpr (FAST)
        B[tri(i, i)] -= t * t * rd[j];
      else
        B[tri(i, i)] -= t * t / d;
    }
  }

  // L D z = b, then L^T w = z
  real z[n];
#pragma unroll
  for (int i = 0; i < n; ++i) {
    real t = b[i];
#pragma unroll
    for (int j = 0; j < i; ++j) t -= B[tri(i, j)] * B[tri(j, j)] * z[j];
    if constexpr (FAST)
      z[i] = t * rd[i];
    else
      z[i] = t / B[tri(i, i)];
  }
#pragma unroll
  for (int i = n - 1; i >= 0; --i) {
    real t = z[i];
#pragma unroll
    for (int j = i + 1; j < n; ++j) t -= B[tri(j, i)] * z[j];
    z[i] = t;
  }
#pragma unroll
  for (int i = 0; i < n; ++i) {
    if constexpr (FAST)
      s[perm(n, i)] = z[i] * sc[i];
    else
      s[perm(n, i)] = z[i] / sc[i];
  }
}

struct Settings {
  real tol, mu, eta, acc, dec;
};

__device__ __forceinline__ Settings settings(const pca_lm_args& a) {
  return Settings{(real)a.tol, (real)a.mu, (real)a.eta, (real)a.acc, (real)a.dec};
}

}  // namespace pca
