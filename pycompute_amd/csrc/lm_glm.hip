// Wide-model LM fit (parameter counts above what one lane's registers hold)
// for generalised linear models f = phi(eta), eta = a0(c) + sum_k A_k(c) p_k,
// phi = exp (PCA_GLM_IDENTITY: phi = identity), e.g. the 22-parameter
// diffusion-kurtosis model.
//
// Same iteration as pca_lm_fit (pycompute/cuda/lsqnonlin.py:820-947): fghhl,
// GMW81 solve, NaN scrub, projected-gradient convergence, trial SSR, gain
// ratio, clamp; state carried between launches exactly the same way. What
// changes is how the n(n+1)/2-entry Hessian is accumulated and where it lives:
//
//  * For a GLM, J_k = phi'(eta) A_k and d2f_jk = phi''(eta) A_j A_k, so the
//    reference's h_jk += d2f_jk r + J_j J_k is h_jk += (w A_j) A_k with
//    w = phi'^2 + phi'' r: one FMA per packed entry. A (per measurement:
//    A_k, A_k^2, a0) is a consts-only table built once per protocol and read
//    with wave-uniform scalar loads.
//  * The packed triangle is accumulated in row tiles of at most PCA_TILE_MAX
//    entries, one pass over the measurements per tile (eta and phi are
//    recomputed per pass; the data row stays in L2). A finished tile is parked
//    in LDS (one column per lane, conflict-free) so only one tile is live in
//    VGPRs during a pass; after the last pass the 22x22 system is gathered
//    from LDS and registers and solved per lane.
//  * f, g and the damping diagonal sum_i max(lam J_j^2, 1e-4) are accumulated
//    in the first pass.
//  * The gain ratio's s^T h s is sum_i w_i (A_i . s)^2, accumulated in the
//    trial-point pass, so h is consumed by the factorisation.
//
// One wavefront per workgroup (64 voxels), the LDS park is per workgroup.
// Kernel arguments: the fit (pca_lm_args) and the table, (M, PCA_TS) rows of
// A_0..A_{n-1}, A_0^2..A_{n-1}^2, a0, padding; the table is a separate
// __restrict__ argument so its wave-uniform reads compile to scalar loads.
// Concatenated after lm_solver.hip with PCA_NP, PCA_M, PCA_TS (table row
// stride) defined.

#ifndef PCA_TILE_MAX
#define PCA_TILE_MAX 136
#endif
#define PCA_GLM_WAVE 64

namespace pca {
namespace glm {

constexpr int NP = PCA_NP, NH = NP * (NP + 1) / 2, M = PCA_M, TS = PCA_TS;

// First row of tile t: tiles are runs of whole rows of the packed lower
// triangle with at most PCA_TILE_MAX entries (at least one row each).
constexpr int tile_row(int t) {
  int r = 0;
  for (int q = 0; q < t && r < NP; ++q) {
    int e = 0;
    do {
      e += r + 1;
      ++r;
    } while (r < NP && e + r + 1 <= PCA_TILE_MAX);
  }
  return r;
}
constexpr int count_tiles() {
  int t = 0;
  while (tile_row(t) < NP) ++t;
  return t;
}
constexpr int NT = count_tiles();
constexpr int NPARK = tri(tile_row(NT - 1), 0);   // entries of all tiles but the last
constexpr int NPARK_ALLOC = NPARK > 0 ? NPARK : 1;

__device__ __forceinline__ real eta_of(const real* __restrict__ A, const real* p) {
  real eta = A[2 * NP];
#pragma unroll
  for (int k = 0; k < NP; ++k) eta += A[k] * p[k];
  return eta;
}

__device__ __forceinline__ real phi(real eta) {
#ifdef PCA_GLM_IDENTITY
  return eta;
#else
  return PCA_EXP(eta);
#endif
}

// One pass over the measurements accumulating the entries of tile T (and, in
// the first pass, f, g and the damping diagonal hd).
template <int T, int NE>
__device__ __forceinline__ void tile_pass(const pca_lm_args& a, const real* __restrict__ tab,
                                          long long v, const real* p, real lam,
                                          real (&acc)[NE], real& f, real* g, real* hd) {
  constexpr int r0 = tile_row(T), r1 = tile_row(T + 1), E0 = tri(r0, 0);
#pragma unroll
  for (int e = 0; e < NE; ++e) acc[e] = 0;
  const long long N = a.N;
  const real* __restrict__ y = a.data + v;
  real ynext = y[0];
  for (int i = 0; i < M; ++i) {
    const real yi = ynext;
    if (i + 1 < M) ynext = y[(long long)(i + 1) * N];
    const real* __restrict__ A = tab + i * TS;
    const real e = phi(eta_of(A, p));
#ifdef PCA_GLM_IDENTITY
    const real d1 = 1, r = e - yi, w = 1;
#else
    const real d1 = e, r = e - yi, w = e * e + e * r;
#endif
    if constexpr (T == 0) {
      f += r * r;
      const real dr = d1 * r, ld = lam * (d1 * d1);
#pragma unroll
      for (int j = 0; j < NP; ++j) {
        g[j] += dr * A[j];
        hd[j] += pca_fmax(ld * A[NP + j], (real)PCA_DAMP_FLOOR);
      }
    }
#pragma unroll
    for (int j = r0; j < r1; ++j) {
      const real vj = w * A[j];
#pragma unroll
      for (int k = 0; k <= j; ++k) acc[tri(j, k) - E0] += vj * A[k];
    }
  }
}

template <int T>
__device__ __forceinline__ void run_tiles(const pca_lm_args& a, const real* __restrict__ tab,
                                          long long v, int lane, const real* p, real lam,
                                          real* h, real& f, real* g, real* hd,
                                          real (*park)[PCA_GLM_WAVE]) {
  if constexpr (T < NT) {
    constexpr int E0 = tri(tile_row(T), 0), NE = tri(tile_row(T + 1), 0) - E0;
    real acc[NE];
    tile_pass<T, NE>(a, tab, v, p, lam, acc, f, g, hd);
    if constexpr (T + 1 < NT) {
#pragma unroll
      for (int e = 0; e < NE; ++e) park[E0 + e][lane] = acc[e];
    } else {
#pragma unroll
      for (int e = 0; e < NE; ++e) h[E0 + e] = acc[e];
    }
    run_tiles<T + 1>(a, tab, v, lane, p, lam, h, f, g, hd, park);
  }
}

}  // namespace glm
}  // namespace pca

extern "C" __global__ void __launch_bounds__(PCA_GLM_WAVE)
pca_lm_glm(pca_lm_args a, const real* __restrict__ tab) {
  using namespace pca;
  using namespace pca::glm;
  __shared__ real park[NPARK_ALLOC][PCA_GLM_WAVE];
  const long long N = a.N;
  const int lane = threadIdx.x;
  const long long v = (long long)blockIdx.x * PCA_GLM_WAVE + lane;
  if (v >= N) return;   // lanes never share LDS columns: no block barrier below
  const Settings S = settings(a);
  real p[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) p[j] = a.params[(long long)j * N + v];
  real lam = a.lam[v], f = a.f[v], f_first = f;
  int st = a.step_type[v];
  int cnt = 0;
  real g[NP], s[NP];
  for (int it = 0; it < a.iters && st != 0; ++it) {
    ++cnt;
    real hd[NP], h[NH];
    f = 0;
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      g[j] = 0;
      hd[j] = 0;
    }
    run_tiles<0>(a, tab, v, lane, p, lam, h, f, g, hd, park);
#pragma unroll
    for (int l = 0; l < NPARK; ++l) h[l] = park[l][lane];
    real hld[NP];
#pragma unroll
    for (int j = 0; j < NP; ++j) hld[j] = h[tri(j, j)] + hd[j];

    gmw81_solve<NP>(h, hld, g, s);   // h is dead afterwards
#pragma unroll
    for (int j = 0; j < NP; ++j)
      if (pca_nonfinite(s[j])) s[j] = 0;

    // projected-gradient convergence (lsqnonlin.py:691-729)
    bool clamped = false;
    real norm = 0;
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      const real lo = a.lb[(long long)j * N + v], hi = a.ub[(long long)j * N + v];
      real t = p[j] - g[j];
      if (t > hi) {
        clamped = true;
        t = hi;
      } else if (t < lo) {
        clamped = true;
        t = lo;
      }
      t = p[j] - t;
      norm += t * t;
    }
    if (it == 0) f_first = f;
    if (norm < S.tol * ((real)1 + f) && ((st & 1) || clamped)) {
      st = 0;
      break;
    }

    // trial point and its SSR (unclamped, lsqnonlin.py:902-911), and the
    // model-Hessian quadratic form of the gain ratio: for a GLM
    // s^T h s = sum_i w_i (A_i . s)^2 with w_i = phi'^2 + phi'' r at p, so h
    // need not outlive the in-place factorisation.
    real ptp[NP];
#pragma unroll
    for (int j = 0; j < NP; ++j) ptp[j] = p[j] - s[j];
    real ftp = 0, shs = 0;
    {
      const real* __restrict__ y = a.data + v;
      real ynext = y[0];
      for (int i = 0; i < M; ++i) {
        const real yi = ynext;
        if (i + 1 < M) ynext = y[(long long)(i + 1) * N];
        const real* __restrict__ A = tab + i * TS;
        const real rt = phi(eta_of(A, ptp)) - yi;
        ftp += rt * rt;
        real as = 0;
#pragma unroll
        for (int k = 0; k < NP; ++k) as += A[k] * s[k];
#ifdef PCA_GLM_IDENTITY
        shs += as * as;
#else
        const real e = phi(eta_of(A, p));
        shs += (e * e + e * (e - yi)) * (as * as);
#endif
      }
    }

    // gain ratio and step type (lsqnonlin.py:432-498)
    const real actual = (real)0.5 * (f - ftp);
    real pred = (real)-0.5 * shs;
#pragma unroll
    for (int i = 0; i < NP; ++i) pred += s[i] * g[i];
    const real rho = actual / pred;
    if (rho > S.mu && actual > 0) {
#pragma unroll
      for (int j = 0; j < NP; ++j) p[j] = ptp[j];
      if (rho > S.eta) {
        lam *= S.acc;
        st = 1;
      } else {
        st = 2;
      }
    } else {
      lam *= S.dec;
      st = 4;
    }
    if (pred < 0) {
      lam *= S.dec;
      st |= 8;
    }

    // clamp (lsqnonlin.py:584-607)
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      const real lo = a.lb[(long long)j * N + v], hi = a.ub[(long long)j * N + v];
      if (p[j] > hi)
        p[j] = hi;
      else if (p[j] < lo)
        p[j] = lo;
    }
  }

#pragma unroll
  for (int j = 0; j < NP; ++j) a.params[(long long)j * N + v] = p[j];
  a.lam[v] = lam;
  a.step_type[v] = (signed char)st;
  a.f[v] = f;
  if (a.first_f) a.first_f[v] = f_first;
  if (a.last_f) a.last_f[v] = f;
  if (a.n_iter) a.n_iter[v] += cnt;
  if (cnt > 0) {
    if (a.g_out) {
#pragma unroll
      for (int j = 0; j < NP; ++j) a.g_out[(long long)j * N + v] = g[j];
    }
    if (a.step_out) {
#pragma unroll
      for (int j = 0; j < NP; ++j) a.step_out[(long long)j * N + v] = s[j];
    }
  }
}
