// Wide-model LM fit for any generated model with more parameters than one
// lane's registers hold (n > 8; generalised linear models have the faster
// lm_glm.hip).
//
// Same iteration as pca_lm_fit (pycompute/cuda/lsqnonlin.py:820-947), in the
// reference's order and arithmetic: fghhl, GMW81 solve, NaN scrub,
// projected-gradient convergence, trial SSR, gain ratio, clamp; state carried
// between launches the same way. What changes is where the n(n+1)/2-entry
// Hessian lives:
//
//  * It is accumulated in row tiles of at most PCA_TILE_MAX entries, one pass
//    over the measurements per tile; each pass evaluates the model, Jacobian
//    and the tile's second derivatives (the generated fjh(), whose other
//    outputs the compiler drops), so only one tile is live in VGPRs.
//  * A finished tile is written to LDS, one column per lane (conflict-free);
//    the whole packed triangle stays there. The solve reads it into
//    registers, and the gain ratio's s^T h s reads it again afterwards, in
//    the reference's order (lsqnonlin.py:432-498).
//  * f, g and the damped diagonal are accumulated in the first pass with the
//    reference's per-measurement update (or, in the reassociated fast build,
//    the damping terms only, the diagonal of h added after the passes).
//
// One wavefront per workgroup (64 voxels); lanes never share LDS columns.
// Concatenated after the model struct (Model) and lm_solver.hip, with PCA_M,
// PCA_UC_MAX and PCA_TILE_MAX defined; consts voxel-uniform (kernel-argument
// block) or per voxel (PCA_CONSTS_PER_VOXEL, (k*M+i)*N+v).

#ifndef PCA_TILE_MAX
#define PCA_TILE_MAX 100
#endif
#define PCA_WIDE_WAVE 64

namespace pca {
namespace wide {

constexpr int NP = Model::NP, NH = Model::NH, NC = Model::NC, M = PCA_M;

// First row of tile t: runs of whole rows of the packed lower triangle with
// at most PCA_TILE_MAX entries (at least one row each).
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

// One pass over the measurements accumulating the entries of tile T into
// LDS; the first pass also accumulates f, g and the damped diagonal hld.
template <int T, class CF>
__device__ __forceinline__ void tile_pass(const pca_lm_args& a, const CF& cf, long long v,
                                          int lane, const real* p, real lam, real& f, real* g,
                                          real* hld, real (*hs)[PCA_WIDE_WAVE]) {
  constexpr int r0 = tile_row(T), r1 = tile_row(T + 1), E0 = tri(r0, 0);
  constexpr int NE = tri(r1, 0) - E0;
  real acc[NE];
#pragma unroll
  for (int e = 0; e < NE; ++e) acc[e] = 0;
  const long long N = a.N;
  for (int i = 0; i < M; ++i) {
    real c[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) c[k] = cf(k, i);
    const real yi = a.data[(long long)i * N + v];
    real m, J[NP], d2[NH];
    Model::fjh(p, c, m, J, d2);
    const real r = m - yi;
    if constexpr (T == 0) {
      f += r * r;
#pragma unroll
      for (int j = 0; j < NP; ++j) {
        const real jt = J[j] * J[j];
        if constexpr (Model::HLD_DAMP_ONLY) {
          hld[j] += pca_fmax(lam * jt, (real)PCA_DAMP_FLOOR);
        } else {
          const real hj = Model::HNZ[tri(j, j)] ? d2[tri(j, j)] * r + jt : jt;
          hld[j] += hj + pca_fmax(lam * jt, (real)PCA_DAMP_FLOOR);
        }
        if constexpr (Model::HLD_DAMP_ONLY)
          g[j] = pca_fma(J[j], r, g[j]);
        else
          g[j] += J[j] * r;
      }
    }
#pragma unroll
    for (int j = r0; j < r1; ++j)
#pragma unroll
      for (int k = 0; k <= j; ++k) {
        const int l = tri(j, k);
        if constexpr (Model::HLD_DAMP_ONLY) {
          const real t = Model::HNZ[l] ? pca_fma(d2[l], r, acc[l - E0]) : acc[l - E0];
          acc[l - E0] = pca_fma(J[j], J[k], t);
        } else {
          const real jt = J[j] * J[k];
          const real hj = Model::HNZ[l] ? d2[l] * r + jt : jt;
          acc[l - E0] += hj;
        }
      }
  }
#pragma unroll
  for (int e = 0; e < NE; ++e) hs[E0 + e][lane] = acc[e];
}

template <int T, class CF>
__device__ __forceinline__ void run_tiles(const pca_lm_args& a, const CF& cf, long long v,
                                          int lane, const real* p, real lam, real& f, real* g,
                                          real* hld, real (*hs)[PCA_WIDE_WAVE]) {
  if constexpr (T < NT) {
    tile_pass<T>(a, cf, v, lane, p, lam, f, g, hld, hs);
    run_tiles<T + 1>(a, cf, v, lane, p, lam, f, g, hld, hs);
  }
}

}  // namespace wide
}  // namespace pca

extern "C" __global__ void __launch_bounds__(PCA_WIDE_WAVE)
pca_lm_wide(pca_lm_args a, pca_uconsts uc) {
  using namespace pca;
  using namespace pca::wide;
  __shared__ real hs[NH][PCA_WIDE_WAVE];
  const long long N = a.N;
  const int lane = threadIdx.x;
  const long long v = (long long)blockIdx.x * PCA_WIDE_WAVE + lane;
  if (v >= N) return;   // lanes never share LDS columns: no block barrier below
#ifdef PCA_CONSTS_PER_VOXEL
  const real* __restrict__ cg = a.consts;
  auto cf = [&](int k, int i) { return cg[(long long)(k * M + i) * N + v]; };
#else
  auto cf = [&](int k, int i) { return uc.c[k * M + i]; };
#endif
  const Settings S = settings(a);
  real p[NP], lo[NP], hi[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    p[j] = a.params[(long long)j * N + v];
    lo[j] = a.lb[(long long)j * N + v];
    hi[j] = a.ub[(long long)j * N + v];
  }
  real lam = a.lam[v], f = a.f[v], f_first = f;
  int st = a.step_type[v];
  int cnt = 0;
  real g[NP], s[NP];
  for (int it = 0; it < a.iters && st != 0; ++it) {
    ++cnt;
    real hld[NP];
    f = 0;
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      g[j] = 0;
      hld[j] = 0;
    }
    run_tiles<0>(a, cf, v, lane, p, lam, f, g, hld, hs);
    if (it == 0) f_first = f;
    {
      real h[NH];
#pragma unroll
      for (int l = 0; l < NH; ++l) h[l] = hs[l][lane];
      if constexpr (Model::HLD_DAMP_ONLY) {
#pragma unroll
        for (int j = 0; j < NP; ++j) hld[j] += h[tri(j, j)];
      }
      gmw81_solve<NP>(h, hld, g, s);
    }
#pragma unroll
    for (int j = 0; j < NP; ++j)
      if (pca_nonfinite(s[j])) s[j] = 0;

    // projected-gradient convergence (lsqnonlin.py:691-729)
    bool clamped = false;
    real norm = 0;
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      real t = p[j] - g[j];
      const bool above = t > hi[j], below = t < lo[j];
      clamped = clamped | above | below;
      t = above ? hi[j] : (below ? lo[j] : t);
      t = p[j] - t;
      norm += t * t;
    }
    if (norm < S.tol * ((real)1 + f) && ((st & 1) || clamped)) {
      st = 0;
      break;
    }

    // trial point and its SSR (unclamped, lsqnonlin.py:902-911)
    real ptp[NP];
#pragma unroll
    for (int j = 0; j < NP; ++j) ptp[j] = p[j] - s[j];
    real ftp = 0;
    for (int i = 0; i < M; ++i) {
      real c[NC];
#pragma unroll
      for (int k = 0; k < NC; ++k) c[k] = cf(k, i);
      const real r = Model::resid(ptp, c, a.data[(long long)i * N + v]);
      ftp += r * r;
    }

    // gain ratio and step type (lsqnonlin.py:432-498), s^T h s from the
    // undamped h kept in LDS
    const real actual = (real)0.5 * (f - ftp);
    real pred = 0;
#pragma unroll
    for (int i = 0; i < NP; ++i)
#pragma unroll
      for (int j = 0; j <= i; ++j) {
        const real e = hs[tri(i, j)][lane] * s[i] * s[j];
        if (i == j)
          pred -= e;
        else
          pred -= (real)2 * e;
      }
    pred *= (real)0.5;
#pragma unroll
    for (int i = 0; i < NP; ++i) pred += s[i] * g[i];
    const real rho = actual / pred;
    const bool accept = rho > S.mu && actual > 0;
    const bool good = accept && rho > S.eta;
    const bool nondesc = pred < 0;
    real nl = good ? lam * S.acc : (accept ? lam : lam * S.dec);
    lam = nondesc ? nl * S.dec : nl;
    st = (good ? 1 : (accept ? 2 : 4)) | (nondesc ? 8 : 0);

    // take the trial point and clamp it (lsqnonlin.py:584-607)
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      const real q = accept ? ptp[j] : p[j];
      p[j] = q > hi[j] ? hi[j] : (q < lo[j] ? lo[j] : q);
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
