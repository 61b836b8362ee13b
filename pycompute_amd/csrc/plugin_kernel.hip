// LM fit over user-supplied model code (the F_GradF plugin hook of
// pycompute/cuda/lsqnonlin.py:32-38, 814-816, 939-941).
//
// Template: concatenated after prelude.hip, the user's device code (linked by
// code_gen_walking) and lm_kernel.hip's solver, with PCA_GRADF / PCA_FOBJ naming
// the user's functions and PCA_NP, PCA_M set. The user's functions keep the
// reference signatures and global-memory (rows, N) layout:
//   PCA_GRADF(params, consts, data, lam, f, g, h, hl, tid, Nprobs)
//       accumulates f = sum r^2, g = J^T r, h (packed lower) and the damped hl
//       into the caller-zeroed arrays (the fghhl contract, lsqnonlin.py:206-346)
//   PCA_FOBJ(params, consts, data, f, tid, Nprobs)
//       accumulates f += sum r^2 (lsqnonlin.py:46-117)
// so the state lives in global memory as in the reference kernel
// (.cu:416-517); the iteration loop runs inside the launch.

struct pca_plugin_args {
  pca_lm_args a;
  real* g;          // (NP, N)
  real* h;          // (NH, N)
  real* hl;         // (NH, N)
  real* ftp;        // (N)
  real* params_tp;  // (NP, N)
  real* step;       // (NP, N)
};

extern "C" __global__ void __launch_bounds__(PCA_BLOCK) pca_lm_plugin(pca_plugin_args P) {
  constexpr int NP = PCA_NP, NH = PCA_NP * (PCA_NP + 1) / 2;
  const pca_lm_args& a = P.a;
  const long long N = a.N;
  const long long v = (long long)blockIdx.x * PCA_BLOCK + threadIdx.x;
  if (v >= N) return;
  const int tid = (int)v, Nprobs = (int)N;
  const pca::Settings S = pca::settings(a);
  real lo[NP], hi[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    lo[j] = a.lb[(long long)j * N + v];
    hi[j] = a.ub[(long long)j * N + v];
  }
  int st = a.step_type[v];
  real f_first = a.f[v];
  int cnt = 0;
  for (int it = 0; it < a.iters && st != 0; ++it) {
    ++cnt;
    a.f[v] = 0;
    P.ftp[v] = 0;
#pragma unroll
    for (int j = 0; j < NP; ++j) P.g[(long long)j * N + v] = 0;
#pragma unroll
    for (int l = 0; l < NH; ++l) {
      P.h[(long long)l * N + v] = 0;
      P.hl[(long long)l * N + v] = 0;
    }
    PCA_GRADF(a.params, a.consts, a.data, a.lam, a.f, P.g, P.h, P.hl, tid, Nprobs);
    const real f = a.f[v];
    if (it == 0) f_first = f;
    real p[NP], g[NP], h[NH], hl[NH], hld[NP], s[NP];
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      p[j] = a.params[(long long)j * N + v];
      g[j] = P.g[(long long)j * N + v];
    }
#pragma unroll
    for (int l = 0; l < NH; ++l) {
      h[l] = P.h[(long long)l * N + v];
      hl[l] = P.hl[(long long)l * N + v];
    }
#pragma unroll
    for (int j = 0; j < NP; ++j) hld[j] = hl[pca::tri(j, j)];
    pca::gmw81_solve<NP>(hl, hld, g, s);
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      if (pca_nonfinite(s[j])) s[j] = 0;
      P.step[(long long)j * N + v] = s[j];
    }
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
      if (norm < S.tol * ((real)1 + f) && ((st & 1) || clamped)) {
        st = 0;
        break;
      }
    }
    real ptp[NP];
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      ptp[j] = p[j] - s[j];
      P.params_tp[(long long)j * N + v] = ptp[j];
    }
    PCA_FOBJ(P.params_tp, a.consts, a.data, P.ftp, tid, Nprobs);
    const real ftp = P.ftp[v];
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
    real lam = a.lam[v];
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
    a.lam[v] = lam;
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      if (p[j] > hi[j])
        p[j] = hi[j];
      else if (p[j] < lo[j])
        p[j] = lo[j];
      a.params[(long long)j * N + v] = p[j];
    }
  }
  a.step_type[v] = (signed char)st;
  if (a.first_f) a.first_f[v] = f_first;
  if (a.last_f) a.last_f[v] = a.f[v];
  if (a.n_iter) a.n_iter[v] += cnt;
}
