// Fused batched Levenberg-Marquardt fit: every LM iteration of a voxel runs in
// one lane, in registers, within a single launch.
//
// Template: the code generator (pycompute_amd/sym/hipgen.py) concatenates
// prelude.hip, a model struct typedef'd to `Model` (NP, NC, NH, resid, accum,
// fjh) and this file, with PCA_M (measurements) and PCA_BLOCK (threads per
// block) defined; hiprtc compiles it for gfx950.
//
// One launch of pca_lm_fit performs `iters` launches of the reference kernel
// k_second_order_levenberg_marquardt (pycompute/cuda/lsqnonlin.py:949-1073,
// emitted test/ivim/second_order_levenberg_marquardt_4_1_21_f_f7def86f0a03.cu:416-535).
// Voxel v iterates while step_type != 0, in the reference's order:
//   zero -> fghhl -> GMW81 solve -> NaN/Inf scrub -> convergence (may stop)
//   -> trial point -> f(trial) -> gain ratio -> clamp.
// State (params, lambda, step_type, f) stays in VGPRs between iterations and
// is read and written once per launch, so calling run() again resumes.
//
// Layout: per-voxel arrays are structure-of-arrays (rows, N), voxel index
// fastest, so each wavefront's loads and stores are unit-stride. Voxel-uniform
// consts arrive as a kernel argument block (scalar registers after unrolling);
// per-voxel consts are read from (NC*M, N).

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

#ifndef PCA_UNROLL_M
#define PCA_UNROLL_M PCA_M
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
    sc[i] = pca_sqrt(pca_abs(B[tri(i, i)]));
    b[i] /= sc[i];
  }
#pragma unroll
  for (int i = 0; i < n; ++i)
#pragma unroll
    for (int j = 0; j <= i; ++j) B[tri(i, j)] /= (sc[i] * sc[j]);

  // GMW81
  real gamma = 0, xi = 0;
#pragma unroll
  for (int i = 0; i < n; ++i)
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      const real t = pca_abs(B[tri(i, j)]);
      if (i == j) {
        if (t > gamma) gamma = t;
      } else {
        if (t > xi) xi = t;
      }
    }
  if (n > 1) xi /= pca_sqrt((real)(n * n - 1));
  real beta2 = PCA_EPS_GMW, delta = PCA_EPS_GMW;
  if (beta2 < gamma) beta2 = gamma;
  if (beta2 < xi) beta2 = xi;
  if (gamma + xi > (real)1) delta *= (gamma + xi);

#pragma unroll
  for (int j = 0; j < n; ++j) {
#pragma unroll
    for (int q = 0; q < j; ++q) B[tri(j, q)] /= B[tri(q, q)];
#pragma unroll
    for (int i = j + 1; i < n; ++i) {
      real t = B[tri(i, j)];
#pragma unroll
      for (int q = 0; q < j; ++q) t -= B[tri(j, q)] * B[tri(i, q)];
      B[tri(i, j)] = t;
    }
    real theta = 0;
#pragma unroll
    for (int i = j + 1; i < n; ++i) {
      const real t = pca_abs(B[tri(i, j)]);
      if (theta < t) theta = t;
    }
    theta *= theta;
    real d = pca_abs(B[tri(j, j)]);
    if (d < delta) d = delta;
    const real t0 = theta / beta2;
    if (d < t0) d = t0;
    B[tri(j, j)] = d;
#pragma unroll
    for (int i = j + 1; i < n; ++i) {
      const real t = B[tri(i, j)];
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
  for (int i = 0; i < n; ++i) s[perm(n, i)] = z[i] / sc[i];
}

}  // namespace pca

static_assert(Model::NC * PCA_M <= PCA_UC_MAX, "consts exceed the kernel-argument block");

extern "C" __global__ void __launch_bounds__(PCA_BLOCK)
pca_lm_fit(pca_lm_args a, pca_uconsts uc) {
  constexpr int NP = Model::NP, NC = Model::NC, NH = Model::NH, M = PCA_M;
  const long long N = a.N;
  const long long v = (long long)blockIdx.x * PCA_BLOCK + threadIdx.x;
  if (v >= N) return;

  const real tol = (real)a.tol, mu = (real)a.mu, eta = (real)a.eta;
  const real acc = (real)a.acc, dec = (real)a.dec;

  real y[M];
#pragma unroll
  for (int i = 0; i < M; ++i) y[i] = a.data[(long long)i * N + v];
  real p[NP], lo[NP], hi[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    p[j] = a.params[(long long)j * N + v];
    lo[j] = a.lb[(long long)j * N + v];
    hi[j] = a.ub[(long long)j * N + v];
  }
#ifdef PCA_CONSTS_PER_VOXEL
  real cv[NC * M];
#pragma unroll
  for (int q = 0; q < NC * M; ++q) cv[q] = a.consts[(long long)q * N + v];
#define PCA_CONST(k, i) cv[(k) * M + (i)]
#else
#define PCA_CONST(k, i) uc.c[(k) * M + (i)]
#endif

  real lam = a.lam[v];
  int st = a.step_type[v];
  real f = a.f[v];
  real f_first = f;
  real g[NP], h[NH], hld[NP], s[NP];
  int cnt = 0;

  for (int it = 0; it < a.iters && st != 0; ++it) {
    ++cnt;
    // fghhl (lsqnonlin.py:206-346)
    f = 0;
#pragma unroll
    for (int j = 0; j < NP; ++j) { g[j] = 0; hld[j] = 0; }
#pragma unroll
    for (int l = 0; l < NH; ++l) h[l] = 0;
#pragma unroll PCA_UNROLL_M
    for (int i = 0; i < M; ++i) {
      real c[NC];
#pragma unroll
      for (int k = 0; k < NC; ++k) c[k] = PCA_CONST(k, i);
      Model::accum(p, c, y[i], lam, f, g, h, hld);
    }
    if (it == 0) f_first = f;

    // step: hl s = g, then scrub NaN/Inf (lsqnonlin.py:843-891)
    pca::gmw81_solve<NP>(h, hld, g, s);
#pragma unroll
    for (int j = 0; j < NP; ++j)
      if (pca_nonfinite(s[j])) s[j] = 0;

    // projected-gradient convergence (lsqnonlin.py:691-729)
    {
      bool clamped = false;
      real norm = 0;
#pragma unroll
      for (int j = 0; j < NP; ++j) {
        real t = p[j] - g[j];
        if (t > hi[j]) {
          clamped = true;
          t = hi[j];
        } else if (t < lo[j]) {
          clamped = true;
          t = lo[j];
        }
        t = p[j] - t;
        norm += t * t;
      }
      if (norm < tol * ((real)1 + f) && ((st & 1) || clamped)) {
        st = 0;
        break;
      }
    }

    // trial point and its SSR (unclamped, lsqnonlin.py:902-911)
    real ptp[NP];
#pragma unroll
    for (int j = 0; j < NP; ++j) ptp[j] = p[j] - s[j];
    real ftp = 0;
#pragma unroll PCA_UNROLL_M
    for (int i = 0; i < M; ++i) {
      real c[NC];
#pragma unroll
      for (int k = 0; k < NC; ++k) c[k] = PCA_CONST(k, i);
      const real r = Model::resid(ptp, c, y[i]);
      ftp += r * r;
    }

    // gain ratio and step type (lsqnonlin.py:432-498)
    const real actual = (real)0.5 * (f - ftp);
    real pred = 0;
#pragma unroll
    for (int i = 0; i < NP; ++i)
#pragma unroll
      for (int j = 0; j <= i; ++j) {
        const real e = h[pca::tri(i, j)] * s[i] * s[j];
        if (i == j)
          pred -= e;
        else
          pred -= (real)2 * e;
      }
    pred *= (real)0.5;
#pragma unroll
    for (int i = 0; i < NP; ++i) pred += s[i] * g[i];
    const real rho = actual / pred;
    if (rho > mu && actual > 0) {
#pragma unroll
      for (int j = 0; j < NP; ++j) p[j] = ptp[j];
      if (rho > eta) {
        lam *= acc;
        st = 1;
      } else {
        st = 2;
      }
    } else {
      lam *= dec;
      st = 4;
    }
    if (pred < 0) {
      lam *= dec;
      st |= 8;
    }

    // clamp (lsqnonlin.py:584-607)
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      if (p[j] > hi[j])
        p[j] = hi[j];
      else if (p[j] < lo[j])
        p[j] = lo[j];
    }
  }
#undef PCA_CONST

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
    if (a.h_out || a.hl_out) {
#pragma unroll
      for (int j = 0; j < NP; ++j)
#pragma unroll
        for (int k = 0; k <= j; ++k) {
          const int l = pca::tri(j, k);
          if (a.h_out) a.h_out[(long long)l * N + v] = h[l];
          if (a.hl_out) a.hl_out[(long long)l * N + v] = (j == k) ? hld[j] : h[l];
        }
    }
  }
}
