// Fused batched Levenberg-Marquardt fit: every LM iteration of a voxel runs in
// one lane, in registers.
//
// Template: the code generator (pycompute_amd/lm.py) concatenates
// strict_math.h, prelude.hip, the generated model structs (Model0 ..
// Model{PCA_NMODELS-1}, with Model = Model0) and this file, with PCA_M
// (measurements of Model0), PCA_M0.. and PCA_BLOCK defined; hiprtc compiles
// the result for gfx950.
//
// pca_lm_fit: one launch performs `iters` launches of the reference kernel
// k_second_order_levenberg_marquardt (pycompute/cuda/lsqnonlin.py:949-1073,
// emitted test/ivim/second_order_levenberg_marquardt_4_1_21_f_f7def86f0a03.cu:416-535).
// Voxel v iterates while step_type != 0, in the reference's order:
//   zero -> fghhl -> GMW81 solve -> NaN/Inf scrub -> convergence (may stop)
//   -> trial point -> f(trial) -> gain ratio -> clamp.
// State (params, lambda, step_type, f) stays in VGPRs between iterations and
// is read and written once per launch, so calling run() again resumes.
//
// pca_lm_chunk: the same iteration, run as a sequence of chunks of k
// iterations over compacted work queues (one per model). After each chunk a
// wavefront appends its still-active voxels to the next queue with one atomic
// per wave; converged voxels write their results and leave (a queue entry
// holds the state that changes -- parameters, lambda, f, step type, iteration
// count -- and the voxel index; data and bounds travel with it or are read
// from the caller's arrays through the index, see pca_ws). Waves therefore
// stay full of active voxels instead of idling behind the slowest lane, and
// several models (different n, M) share one persistent launch. Per-voxel
// results are identical to pca_lm_fit.
//
// Layout: per-voxel arrays are structure-of-arrays (rows, N), voxel index
// fastest, so wavefront loads and stores are unit-stride; the working sets of
// the chunked path are packed the same way.

#ifndef PCA_UNROLL_M
#define PCA_UNROLL_M PCA_M
#endif
// Occupancy hint (waves per SIMD) chosen by the autotuner; caps VGPRs per lane.
#ifdef PCA_WAVES_PER_EU
#define PCA_KATTR __attribute__((amdgpu_waves_per_eu(PCA_WAVES_PER_EU, PCA_WAVES_PER_EU)))
#else
#define PCA_KATTR
#endif
// Register fence every PCA_SCHED_GROUP measurements of the unrolled loops:
// the running sums pass through an empty asm statement, then the parameters
// do. The next group's arithmetic reads the fenced parameters, so it cannot be
// scheduled ahead of this group's sums; the number of measurements in flight,
// and the registers they hold, stay bounded.
#ifndef PCA_SCHED_GROUP
#define PCA_SCHED_GROUP 0
#endif
// Factored accumulation (the model struct's prep/accf/finish, rprep/residf):
// parameter-only factors of the Jacobian and second derivatives are applied
// once per iteration instead of once per measurement, and const-only
// subexpressions are read as consts Mdl::NC .. Mdl::NCT-1 of the measurement
// (rows of the voxel-uniform consts table filled by the host).
#ifndef PCA_FACTORED
#define PCA_FACTORED 0
#endif
// Exponential recurrence (PCA_EXP_REC = R > 0, factored builds whose
// exponential factor rows are arithmetic progressions over the measurements):
// exp(k_i x) is evaluated directly every R measurements and by one multiply
// in between (pca_exp_rec in the prelude); the host checks the rows. The
// default PCA_EXP_REC 0 is set in the prelude.
//
// Voxels per lane of the chunked kernel (1 or 2): a lane iterates PCA_VPL
// queue entries in lockstep, so their independent dependency chains
// interleave (64 * PCA_VPL voxels per wavefront).
#ifndef PCA_VPL
#define PCA_VPL 1
#endif
namespace pca {
__device__ __forceinline__ void reg_fence(real& x) { asm volatile("" : "+v"(x)); }
template <int K>
__device__ __forceinline__ void reg_fence(real (&x)[K]) {
#pragma unroll
  for (int k = 0; k < K; ++k) reg_fence(x[k]);
}
}  // namespace pca
#ifndef PCA_NMODELS
#define PCA_NMODELS 1
typedef Model Model0;
#define PCA_M0 PCA_M
#define PCA_UNROLL_M0 PCA_UNROLL_M
#endif

namespace pca {

// One voxel's state in registers.
template <class Mdl, int M>
struct Voxel {
  static constexpr int NP = Mdl::NP, NH = Mdl::NH;
  real y[M];
  real p[NP], lo[NP], hi[NP];
  real lam, f, f_first;
  int st;
  real g[NP], h[NH], hld[NP], s[NP];
};

// One reference launch of the LM iteration for a voxel with st != 0
// (lsqnonlin.py:820-947). cf(k, i) returns const k of measurement i.
// Returns after the convergence test when the voxel converges (st = 0).
template <class Mdl, int M, int UNROLL, class CF>
__device__ __forceinline__ void iterate(Voxel<Mdl, M>& V, const Settings& S, const CF& cf) {
  constexpr int NP = Mdl::NP, NC = Mdl::NC, NH = Mdl::NH;
  constexpr bool factored = PCA_FACTORED && Mdl::FACTORED;
  // fghhl (lsqnonlin.py:206-346)
  V.f = 0;
  if constexpr (factored) {
    // the same sums with the parameter-only factors applied after the loop
    typename Mdl::Pre P;
    Mdl::prep(V.p, V.lam, P);
    real s[Mdl::NS], q[NH], dmp[NP], ex[Mdl::NE], qx[Mdl::NE];
#pragma unroll
    for (int k = 0; k < Mdl::NS; ++k) s[k] = 0;
#pragma unroll
    for (int l = 0; l < NH; ++l) q[l] = 0;
#pragma unroll
    for (int j = 0; j < NP; ++j) dmp[j] = 0;
#pragma unroll UNROLL
    for (int i = 0; i < M; ++i) {
      real c[Mdl::NCT];
#pragma unroll
      for (int k = 0; k < Mdl::NCT; ++k) c[k] = cf(k, i);
      Mdl::accf(P, c, V.y[i], V.f, s, q, dmp, i, ex, qx);
      if constexpr (PCA_SCHED_GROUP > 0) {
        if (i % PCA_SCHED_GROUP == PCA_SCHED_GROUP - 1 && i + 1 < M) {
          reg_fence(V.f);
          reg_fence(s);
          reg_fence(q);
          reg_fence(dmp);
          reg_fence(P.a);
          reg_fence(P.L);
          if constexpr (Mdl::EXP_REC) {
            reg_fence(ex);
            reg_fence(qx);
          }
        }
      }
    }
    Mdl::finish(V.p, P.L, s, q, dmp, V.g, V.h, V.hld);
  } else {
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      V.g[j] = 0;
      V.hld[j] = 0;
    }
#pragma unroll
    for (int l = 0; l < NH; ++l) V.h[l] = 0;
#pragma unroll UNROLL
    for (int i = 0; i < M; ++i) {
      real c[NC];
#pragma unroll
      for (int k = 0; k < NC; ++k) c[k] = cf(k, i);
      Mdl::accum(V.p, c, V.y[i], V.lam, V.f, V.g, V.h, V.hld);
      if constexpr (PCA_SCHED_GROUP > 0) {
        if (i % PCA_SCHED_GROUP == PCA_SCHED_GROUP - 1 && i + 1 < M) {
          reg_fence(V.f);
          reg_fence(V.g);
          reg_fence(V.h);
          reg_fence(V.hld);
          reg_fence(V.p);
        }
      }
    }
  }
  if constexpr (Mdl::HLD_DAMP_ONLY) {   // the fast build accumulates damping only
#pragma unroll
    for (int j = 0; j < NP; ++j) V.hld[j] += V.h[tri(j, j)];
  }

  // step: hl s = g, then scrub NaN/Inf (lsqnonlin.py:843-891)
  gmw81_solve<NP>(V.h, V.hld, V.g, V.s);
#pragma unroll
  for (int j = 0; j < NP; ++j)
    if (pca_nonfinite(V.s[j])) V.s[j] = 0;

  // projected-gradient convergence (lsqnonlin.py:691-729); selects rather
  // than branches (a NaN compares false both ways and stays unclamped)
  {
    bool clamped = false;
    real norm = 0;
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      real t = V.p[j] - V.g[j];
      const bool above = t > V.hi[j], below = t < V.lo[j];
      clamped = clamped | above | below;
      t = above ? V.hi[j] : (below ? V.lo[j] : t);
      t = V.p[j] - t;
      norm += t * t;
    }
    if (norm < S.tol * ((real)1 + V.f) && ((V.st & 1) || clamped)) {
      V.st = 0;
      return;
    }
  }

  // trial point and its SSR (unclamped, lsqnonlin.py:902-911)
  real ptp[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) ptp[j] = V.p[j] - V.s[j];
  real ftp = 0;
  if constexpr (factored) {
    typename Mdl::RPre R;
    Mdl::rprep(ptp, R);
    real ex[Mdl::NE], qx[Mdl::NE];
#pragma unroll UNROLL
    for (int i = 0; i < M; ++i) {
      real c[Mdl::NCT];
#pragma unroll
      for (int k = 0; k < Mdl::NCT; ++k) c[k] = cf(k, i);
      const real r = Mdl::residf(R, c, V.y[i], i, ex, qx);
      ftp += r * r;
      if constexpr (PCA_SCHED_GROUP > 0) {
        if (i % PCA_SCHED_GROUP == PCA_SCHED_GROUP - 1 && i + 1 < M) {
          reg_fence(ftp);
          reg_fence(R.a);
          if constexpr (Mdl::EXP_REC) {
            reg_fence(ex);
            reg_fence(qx);
          }
        }
      }
    }
  } else {
#pragma unroll UNROLL
    for (int i = 0; i < M; ++i) {
      real c[NC];
#pragma unroll
      for (int k = 0; k < NC; ++k) c[k] = cf(k, i);
      const real r = Mdl::resid(ptp, c, V.y[i]);
      ftp += r * r;
      if constexpr (PCA_SCHED_GROUP > 0) {
        if (i % PCA_SCHED_GROUP == PCA_SCHED_GROUP - 1 && i + 1 < M) {
          reg_fence(ftp);
          reg_fence(ptp);
        }
      }
    }
  }

  // gain ratio and step type (lsqnonlin.py:432-498)
  const real actual = (real)0.5 * (V.f - ftp);
  real pred = 0;
#pragma unroll
  for (int i = 0; i < NP; ++i)
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      const real e = V.h[tri(i, j)] * V.s[i] * V.s[j];
      if (i == j)
        pred -= e;
      else
        pred -= (real)2 * e;
    }
  pred *= (real)0.5;
#pragma unroll
  for (int i = 0; i < NP; ++i) pred += V.s[i] * V.g[i];
  const real rho = actual / pred;
  // accept / reject and the damping update, as selects
  const bool accept = rho > S.mu && actual > 0;
  const bool good = accept && rho > S.eta;
  const bool nondesc = pred < 0;
  real lam = good ? V.lam * S.acc : (accept ? V.lam : V.lam * S.dec);
  lam = nondesc ? lam * S.dec : lam;
  V.lam = lam;
  V.st = (good ? 1 : (accept ? 2 : 4)) | (nondesc ? 8 : 0);

  // take the trial point and clamp it (lsqnonlin.py:584-607)
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const real q = accept ? ptp[j] : V.p[j];
    V.p[j] = q > V.hi[j] ? V.hi[j] : (q < V.lo[j] ? V.lo[j] : q);
  }
}

// The iteration of iterate() for the VPL voxels of one lane together
// (VPL > 1). Every step is written for all of them at once, so the compiler
// interleaves their independent chains. There is no early return: a voxel that
// converges in this iteration, or had already stopped (st == 0), runs the
// rest of the iteration too and keeps its state through selects. Its f is
// kept as it was when it had already stopped; g, h, hld and s are recomputed,
// and since a converged voxel's p and lambda never change afterwards they come
// out equal to its converging iteration's (they are only stored for a voxel
// that iterated in this launch). cf(u, k, i): const k of measurement i for
// voxel u. Per-voxel results equal iterate()'s.
template <class Mdl, int M, int UNROLL, int VPL, class CF>
__device__ __forceinline__ void iterate_n(Voxel<Mdl, M> (&V)[VPL], const Settings& S,
                                          const CF& cf) {
  constexpr int NP = Mdl::NP, NC = Mdl::NC, NH = Mdl::NH;
  constexpr bool factored = PCA_FACTORED && Mdl::FACTORED;
  bool active[VPL];
  real f[VPL];
#pragma unroll
  for (int u = 0; u < VPL; ++u) {
    active[u] = V[u].st != 0;
    f[u] = 0;
  }
  // fghhl (lsqnonlin.py:206-346)
  if constexpr (factored) {
    typename Mdl::Pre P[VPL];
    real s[VPL][Mdl::NS], q[VPL][NH], dmp[VPL][NP], ex[VPL][Mdl::NE], qx[VPL][Mdl::NE];
#pragma unroll
    for (int u = 0; u < VPL; ++u) {
      Mdl::prep(V[u].p, V[u].lam, P[u]);
#pragma unroll
      for (int k = 0; k < Mdl::NS; ++k) s[u][k] = 0;
#pragma unroll
      for (int l = 0; l < NH; ++l) q[u][l] = 0;
#pragma unroll
      for (int j = 0; j < NP; ++j) dmp[u][j] = 0;
    }
#pragma unroll UNROLL
    for (int i = 0; i < M; ++i) {
#pragma unroll
      for (int u = 0; u < VPL; ++u) {
        real c[Mdl::NCT];
#pragma unroll
        for (int k = 0; k < Mdl::NCT; ++k) c[k] = cf(u, k, i);
        Mdl::accf(P[u], c, V[u].y[i], f[u], s[u], q[u], dmp[u], i, ex[u], qx[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < VPL; ++u)
      Mdl::finish(V[u].p, P[u].L, s[u], q[u], dmp[u], V[u].g, V[u].h, V[u].hld);
  } else {
#pragma unroll
    for (int u = 0; u < VPL; ++u) {
#pragma unroll
      for (int j = 0; j < NP; ++j) {
        V[u].g[j] = 0;
        V[u].hld[j] = 0;
      }
#pragma unroll
      for (int l = 0; l < NH; ++l) V[u].h[l] = 0;
    }
#pragma unroll UNROLL
    for (int i = 0; i < M; ++i) {
#pragma unroll
      for (int u = 0; u < VPL; ++u) {
        real c[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) c[k] = cf(u, k, i);
        Mdl::accum(V[u].p, c, V[u].y[i], V[u].lam, f[u], V[u].g, V[u].h, V[u].hld);
      }
    }
  }
  bool go[VPL];
#pragma unroll
  for (int u = 0; u < VPL; ++u) {
    if constexpr (Mdl::HLD_DAMP_ONLY) {
#pragma unroll
      for (int j = 0; j < NP; ++j) V[u].hld[j] += V[u].h[tri(j, j)];
    }
    // step and NaN/Inf scrub (lsqnonlin.py:843-891)
    gmw81_solve<NP>(V[u].h, V[u].hld, V[u].g, V[u].s);
#pragma unroll
    for (int j = 0; j < NP; ++j)
      if (pca_nonfinite(V[u].s[j])) V[u].s[j] = 0;
    // projected-gradient convergence (lsqnonlin.py:691-729)
    bool clamped = false;
    real norm = 0;
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      real t = V[u].p[j] - V[u].g[j];
      const bool above = t > V[u].hi[j], below = t < V[u].lo[j];
      clamped = clamped | above | below;
      t = above ? V[u].hi[j] : (below ? V[u].lo[j] : t);
      t = V[u].p[j] - t;
      norm += t * t;
    }
    const bool conv = norm < S.tol * ((real)1 + f[u]) && ((V[u].st & 1) || clamped);
    go[u] = active[u] && !conv;
    V[u].f = active[u] ? f[u] : V[u].f;
  }

  // trial points and their SSR (unclamped, lsqnonlin.py:902-911)
  real ptp[VPL][NP], ftp[VPL];
#pragma unroll
  for (int u = 0; u < VPL; ++u) {
    ftp[u] = 0;
#pragma unroll
    for (int j = 0; j < NP; ++j) ptp[u][j] = V[u].p[j] - V[u].s[j];
  }
  if constexpr (factored) {
    typename Mdl::RPre R[VPL];
    real ex[VPL][Mdl::NE], qx[VPL][Mdl::NE];
#pragma unroll
    for (int u = 0; u < VPL; ++u) Mdl::rprep(ptp[u], R[u]);
#pragma unroll UNROLL
    for (int i = 0; i < M; ++i) {
#pragma unroll
      for (int u = 0; u < VPL; ++u) {
        real c[Mdl::NCT];
#pragma unroll
        for (int k = 0; k < Mdl::NCT; ++k) c[k] = cf(u, k, i);
        const real r = Mdl::residf(R[u], c, V[u].y[i], i, ex[u], qx[u]);
        ftp[u] += r * r;
      }
    }
  } else {
#pragma unroll UNROLL
    for (int i = 0; i < M; ++i) {
#pragma unroll
      for (int u = 0; u < VPL; ++u) {
        real c[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) c[k] = cf(u, k, i);
        const real r = Mdl::resid(ptp[u], c, V[u].y[i]);
        ftp[u] += r * r;
      }
    }
  }

  // gain ratio, step type and clamp (lsqnonlin.py:432-498, 584-607), kept
  // only by voxels that iterate on
#pragma unroll
  for (int u = 0; u < VPL; ++u) {
    const real actual = (real)0.5 * (V[u].f - ftp[u]);
    real pred = 0;
#pragma unroll
    for (int i = 0; i < NP; ++i)
#pragma unroll
      for (int j = 0; j <= i; ++j) {
        const real e = V[u].h[tri(i, j)] * V[u].s[i] * V[u].s[j];
        if (i == j)
          pred -= e;
        else
          pred -= (real)2 * e;
      }
    pred *= (real)0.5;
#pragma unroll
    for (int i = 0; i < NP; ++i) pred += V[u].s[i] * V[u].g[i];
    const real rho = actual / pred;
    const bool accept = rho > S.mu && actual > 0;
    const bool good = accept && rho > S.eta;
    const bool nondesc = pred < 0;
    real lam = good ? V[u].lam * S.acc : (accept ? V[u].lam : V[u].lam * S.dec);
    lam = nondesc ? lam * S.dec : lam;
    V[u].lam = go[u] ? lam : V[u].lam;
    V[u].st = go[u] ? ((good ? 1 : (accept ? 2 : 4)) | (nondesc ? 8 : 0)) : 0;
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      const real q = accept ? ptp[u][j] : V[u].p[j];
      const real c = q > V[u].hi[j] ? V[u].hi[j] : (q < V[u].lo[j] ? V[u].lo[j] : q);
      V[u].p[j] = go[u] ? c : V[u].p[j];
    }
  }
}

// Write a voxel's results to the caller's arrays (original index v).
template <class Mdl, int M>
__device__ __forceinline__ void store_result(const pca_lm_args& a, long long v,
                                             const Voxel<Mdl, M>& V, int cnt_total,
                                             bool have_workspace) {
  constexpr int NP = Mdl::NP;
  const long long N = a.N;
#pragma unroll
  for (int j = 0; j < NP; ++j) a.params[(long long)j * N + v] = V.p[j];
  a.lam[v] = V.lam;
  a.step_type[v] = (signed char)V.st;
  a.f[v] = V.f;
  if (a.first_f) a.first_f[v] = V.f_first;
  if (a.last_f) a.last_f[v] = V.f;
  if (a.n_iter) a.n_iter[v] += cnt_total;
  if (have_workspace) {
    if (a.g_out) {
#pragma unroll
      for (int j = 0; j < NP; ++j) a.g_out[(long long)j * N + v] = V.g[j];
    }
    if (a.step_out) {
#pragma unroll
      for (int j = 0; j < NP; ++j) a.step_out[(long long)j * N + v] = V.s[j];
    }
    if (a.h_out || a.hl_out) {
#pragma unroll
      for (int j = 0; j < NP; ++j)
#pragma unroll
        for (int k = 0; k <= j; ++k) {
          const int l = tri(j, k);
          if (a.h_out) a.h_out[(long long)l * N + v] = V.h[l];
          if (a.hl_out) a.hl_out[(long long)l * N + v] = (j == k) ? V.hld[j] : V.h[l];
        }
    }
  }
}

template <class Mdl, int M>
__device__ __forceinline__ void load_user(const pca_lm_args& a, long long v, Voxel<Mdl, M>& V) {
  const long long N = a.N;
#pragma unroll
  for (int i = 0; i < M; ++i) V.y[i] = a.data[(long long)i * N + v];
#pragma unroll
  for (int j = 0; j < Mdl::NP; ++j) {
    V.p[j] = a.params[(long long)j * N + v];
    V.lo[j] = a.lb[(long long)j * N + v];
    V.hi[j] = a.ub[(long long)j * N + v];
  }
  V.lam = a.lam[v];
  V.st = a.step_type[v];
  V.f = a.f[v];
  V.f_first = V.f;
}

}  // namespace pca

#ifndef PCA_CONSTS_PER_VOXEL
// NCT - NC rows of derived consts follow the model's own (factored builds)
static_assert(Model::NCT * PCA_M <= PCA_UC_MAX, "consts exceed the kernel-argument block");
#else
static_assert(Model::NCT == Model::NC, "derived consts need voxel-uniform consts");
#endif

// Per-voxel consts: kept in registers when there are at most this many.
#define PCA_CV_REGS 64

extern "C" __global__ void __launch_bounds__(PCA_BLOCK) PCA_KATTR
pca_lm_fit(pca_lm_args a, pca_uconsts uc) {
  constexpr int NC = Model::NC, M = PCA_M;
  const long long N = a.N;
  const long long v = (long long)blockIdx.x * PCA_BLOCK + threadIdx.x;
  if (v >= N) return;
  const pca::Settings S = pca::settings(a);
  pca::Voxel<Model, M> V;
  pca::load_user(a, v, V);
#ifdef PCA_CONSTS_PER_VOXEL
  constexpr int NCV = (NC * M <= PCA_CV_REGS) ? NC * M : 1;
  real cv[NCV];
  if constexpr (NC * M <= PCA_CV_REGS) {
#pragma unroll
    for (int q = 0; q < NCV; ++q) cv[q] = a.consts[(long long)q * N + v];
  }
  const real* __restrict__ cg = a.consts;
  auto cf = [&](int k, int i) {
    if constexpr (NC * M <= PCA_CV_REGS)
      return cv[(k * M + i) % NCV];
    else
      return cg[(long long)(k * M + i) * N + v];
  };
#else
  auto cf = [&](int k, int i) { return uc.c[k * M + i]; };
#endif
  int cnt = 0;
  for (int it = 0; it < a.iters && V.st != 0; ++it) {
    ++cnt;
    pca::iterate<Model, M, PCA_UNROLL_M>(V, S, cf);
    if (it == 0) V.f_first = V.f;
  }
  pca::store_result(a, v, V, cnt, cnt > 0);
}

// ---------------------------------------------------------------------------
// chunked execution with convergence compaction
// ---------------------------------------------------------------------------
// Per-voxel consts travel with the working set only when they fit in
// registers (at most PCA_CV_REGS values); the generator defines PCA_NO_CHUNK
// for larger per-voxel layouts, which pca_lm_fit serves alone.
#ifndef PCA_NO_CHUNK

struct pca_ws {          // packed queue of the voxels still iterating, entry e of cap:
  real* y;               // (M, cap) measurements, or null
  real* cv;              // (NC*M, cap) per-voxel consts, or null
  real* lo;              // (NP, cap) or null
  real* hi;              // (NP, cap) or null
  real* p;               // (NP, cap) parameters
  real* lam;             // (cap)
  real* f;               // (cap)
  real* f_first;         // (cap)
  int* idx;              // (cap) original voxel index
  int* cnt;              // (cap) iterations executed in this run
  signed char* st;       // (cap)
  long long cap;
};
// Two queue layouts, chosen per call by y: carried (y, lo, hi and per-voxel
// consts travel with the entry: coalesced reads, about 3.6x the queue memory
// for IVIM) or compact (y null: they are read from the caller's arrays
// through idx; compaction by per-wave atomics scrambles the queue order, so
// later chunks gather from scattered voxels and fetch more bytes).

struct pca_model_job {
  pca_lm_args a;         // the caller's arrays for this model
  const real* uc;        // voxel-uniform consts (device), or null for per-voxel consts
  pca_ws ws[2];          // ping-pong queues
};

struct pca_chunk_args {
  pca_model_job job[PCA_NMODELS];
  int* counts;           // (n_chunks + 1, NMODELS): queue length read by chunk c
  int chunk;             // chunk c reads queue c & 1 (counts row c), appends to the other
  int it0;               // iteration index of the chunk's first iteration
  int k;                 // iterations in this chunk
  int last;              // 1: no chunk follows; every voxel stores its result
  int fresh;             // 1: chunk 0 starts every voxel from the setup() state
                         // (lambda 5, step_type 1, f 0, n_iter 0) instead of
                         // reading it: a captured fit needs no reset kernel
};

#if defined(PCA_UC_LDS) && !defined(PCA_CONSTS_PER_VOXEL)
// The voxel-uniform consts tables of every model, copied to LDS when the
// chunk kernel starts (PCA_UC_LDS): the measurement loop reads them at LDS
// latency instead of the vector cache's.
__shared__ real pca_uc_lds[PCA_NMODELS][PCA_UC_MAX];
#endif

namespace pca {

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

template <int MI, class Mdl, int M, int UNROLL>
__device__ __forceinline__ void chunk_wave(const pca_chunk_args& A, long long e0, long long count) {
  constexpr int NP = Mdl::NP, NC = Mdl::NC, VPL = PCA_VPL;
  const pca_model_job& J = A.job[MI];
  const pca_lm_args& a = J.a;
  const pca_ws& in = J.ws[A.chunk & 1];
  const pca_ws& out = J.ws[(A.chunk + 1) & 1];
  const Settings S = settings(a);
  const bool from_user = (A.chunk == 0);
  const bool carried = in.y != nullptr;   // uniform: the queue layout of this call
#ifdef PCA_CONSTS_PER_VOXEL
  constexpr bool per_voxel = true;
  static_assert(NC * M <= PCA_CV_REGS, "per-voxel consts too large for the chunked kernel");
#else
  constexpr bool per_voxel = false;
#endif

  // slot u of this lane: queue entry e0 + 64 u (each load of a wavefront
  // stays one contiguous row segment)
  Voxel<Mdl, M> V[VPL];
  long long v[VPL];
  bool valid[VPL];
  int cnt_total[VPL];
  constexpr int NCV = (NC * M <= PCA_CV_REGS) ? NC * M : 1;
  real cv[VPL][NCV];
#pragma unroll
  for (int u = 0; u < VPL; ++u) {
    const long long e = e0 + 64 * u;
    valid[u] = e < count;
    v[u] = 0;
    cnt_total[u] = 0;
    if constexpr (VPL > 1) {
      if (!valid[u]) {   // an empty slot computes on zeros and is never stored
#pragma unroll
        for (int i = 0; i < M; ++i) V[u].y[i] = 0;
#pragma unroll
        for (int j = 0; j < NP; ++j) V[u].p[j] = V[u].lo[j] = V[u].hi[j] = 0;
#pragma unroll
        for (int q = 0; q < NCV; ++q) cv[u][q] = 0;
        V[u].lam = V[u].f = V[u].f_first = 0;
        V[u].st = 0;
      }
    }
    if (!valid[u]) continue;
    if (from_user) {
      v[u] = e;
      load_user(a, e, V[u]);
      if (A.fresh) {
        V[u].lam = (real)5;
        V[u].st = 1;
        V[u].f = V[u].f_first = (real)0;
        if (a.n_iter) a.n_iter[e] = 0;
      }
      if constexpr (per_voxel) {
#pragma unroll
        for (int q = 0; q < NCV; ++q) cv[u][q] = a.consts[(long long)q * a.N + e];
      }
    } else {
      const long long cap = in.cap, N = a.N;
      const long long vv = in.idx[e];
      v[u] = vv;
      if (carried) {
#pragma unroll
        for (int i = 0; i < M; ++i) V[u].y[i] = in.y[(long long)i * cap + e];
#pragma unroll
        for (int j = 0; j < NP; ++j) {
          V[u].lo[j] = in.lo[(long long)j * cap + e];
          V[u].hi[j] = in.hi[(long long)j * cap + e];
        }
        if constexpr (per_voxel) {
#pragma unroll
          for (int q = 0; q < NCV; ++q) cv[u][q] = in.cv[(long long)q * cap + e];
        }
      } else {
#pragma unroll
        for (int i = 0; i < M; ++i) V[u].y[i] = a.data[(long long)i * N + vv];
#pragma unroll
        for (int j = 0; j < NP; ++j) {
          V[u].lo[j] = a.lb[(long long)j * N + vv];
          V[u].hi[j] = a.ub[(long long)j * N + vv];
        }
        if constexpr (per_voxel) {
#pragma unroll
          for (int q = 0; q < NCV; ++q) cv[u][q] = a.consts[(long long)q * N + vv];
        }
      }
#pragma unroll
      for (int j = 0; j < NP; ++j) V[u].p[j] = in.p[(long long)j * cap + e];
      V[u].lam = in.lam[e];
      V[u].f = in.f[e];
      V[u].f_first = in.f_first[e];
      V[u].st = in.st[e];
      cnt_total[u] = in.cnt[e];
    }
  }
  int cnt[VPL];
#pragma unroll
  for (int u = 0; u < VPL; ++u) cnt[u] = 0;
#ifdef PCA_CONSTS_PER_VOXEL
  auto cfu = [&](int u, int k, int i) { return cv[u][(k * M + i) % NCV]; };
#elif defined(PCA_UC_LDS)
  auto cfu = [&](int u, int k, int i) { return pca_uc_lds[MI][k * M + i]; };
#else
  const real* __restrict__ uc = J.uc;
  auto cfu = [&](int u, int k, int i) { return uc[k * M + i]; };
#endif
  if constexpr (VPL == 1) {
    if (valid[0]) {
      auto cf = [&](int k, int i) { return cfu(0, k, i); };
      for (int j = 0; j < A.k && V[0].st != 0; ++j) {
        iterate<Mdl, M, UNROLL>(V[0], S, cf);
        if (A.it0 + j == 0) V[0].f_first = V[0].f;
        ++cnt[0];
      }
    }
  } else {
    for (int j = 0; j < A.k; ++j) {
      bool act[VPL], any = false;
#pragma unroll
      for (int u = 0; u < VPL; ++u) {
        act[u] = V[u].st != 0;   // empty slots hold st = 0
        any = any || act[u];
      }
      if (!any) break;
      iterate_n<Mdl, M, UNROLL, VPL>(V, S, cfu);
#pragma unroll
      for (int u = 0; u < VPL; ++u) {
        if (A.it0 + j == 0) V[u].f_first = V[u].f;
        cnt[u] += act[u] ? 1 : 0;
      }
    }
  }
  // one atomic per wavefront reserves the slots of its surviving voxels
  bool keep[VPL];
  unsigned long long mask[VPL];
  int n_keep = 0;
#pragma unroll
  for (int u = 0; u < VPL; ++u) {
    cnt_total[u] += cnt[u];
    keep[u] = valid[u] && V[u].st != 0 && !A.last;
    mask[u] = __ballot(keep[u]);
    n_keep += __popcll(mask[u]);
  }
  int base = 0;
  if (lane_id() == 0 && n_keep > 0)
    base = atomicAdd(&A.counts[(A.chunk + 1) * PCA_NMODELS + MI], n_keep);
  base = __shfl(base, 0);
  const unsigned long long below_lane = (lane_id() == 0) ? 0ull : (~0ull >> (64 - lane_id()));
#pragma unroll
  for (int u = 0; u < VPL; ++u) {
    if (keep[u]) {
      const long long pos = base + __popcll(mask[u] & below_lane);
      const long long cap = out.cap;
      if (carried) {
#pragma unroll
        for (int i = 0; i < M; ++i) out.y[(long long)i * cap + pos] = V[u].y[i];
#pragma unroll
        for (int j = 0; j < NP; ++j) {
          out.lo[(long long)j * cap + pos] = V[u].lo[j];
          out.hi[(long long)j * cap + pos] = V[u].hi[j];
        }
        if constexpr (per_voxel) {
#pragma unroll
          for (int q = 0; q < NCV; ++q) out.cv[(long long)q * cap + pos] = cv[u][q];
        }
      }
#pragma unroll
      for (int j = 0; j < NP; ++j) out.p[(long long)j * cap + pos] = V[u].p[j];
      out.lam[pos] = V[u].lam;
      out.f[pos] = V[u].f;
      out.f_first[pos] = V[u].f_first;
      out.st[pos] = (signed char)V[u].st;
      out.cnt[pos] = cnt_total[u];
      out.idx[pos] = (int)v[u];
    } else if (valid[u]) {
      store_result(a, v[u], V[u], cnt_total[u], cnt[u] > 0);
    }
    base += __popcll(mask[u]);
  }
}

template <int MI>
struct ModelAt;
#define PCA_MODEL_AT(I)                                  \
  template <>                                            \
  struct ModelAt<I> {                                    \
    typedef Model##I type;                               \
    static constexpr int M = PCA_M##I;                   \
    static constexpr int UNROLL = PCA_UNROLL_M##I;       \
  };
PCA_MODEL_AT(0)
#if PCA_NMODELS > 1
PCA_MODEL_AT(1)
#endif
#if PCA_NMODELS > 2
PCA_MODEL_AT(2)
#endif
#if PCA_NMODELS > 3
PCA_MODEL_AT(3)
#endif

#if defined(PCA_UC_LDS) && !defined(PCA_CONSTS_PER_VOXEL)
// copies model MI's (and the following models') consts tables into LDS; the
// caller synchronises the workgroup afterwards
template <int MI>
__device__ __forceinline__ void load_uc_lds(const pca_chunk_args& A) {
  if constexpr (MI < PCA_NMODELS) {
    constexpr int n = ModelAt<MI>::type::NCT * ModelAt<MI>::M;
    const real* __restrict__ uc = A.job[MI].uc;
    for (int q = threadIdx.x; q < n; q += PCA_BLOCK) pca_uc_lds[MI][q] = uc[q];
    load_uc_lds<MI + 1>(A);
  }
}
#endif

template <int MI>
__device__ __forceinline__ void dispatch(int m, const pca_chunk_args& A, long long e0,
                                         long long count) {
  if constexpr (MI < PCA_NMODELS) {
    if (m == MI) {
      chunk_wave<MI, typename ModelAt<MI>::type, ModelAt<MI>::M, ModelAt<MI>::UNROLL>(A, e0, count);
    } else {
      dispatch<MI + 1>(m, A, e0, count);
    }
  }
}

}  // namespace pca

// Grid-stride over wave tasks of 64 * PCA_VPL queue slots; model m's queue
// occupies ceil(count_m / (64 * PCA_VPL)) consecutive tasks, so every
// wavefront works on one model.
extern "C" __global__ void __launch_bounds__(PCA_BLOCK) PCA_KATTR
pca_lm_chunk(pca_chunk_args A) {
  // queue lengths: chunk 0 reads every model's voxels, later chunks the counts
  // the previous chunk appended
  int cnt[PCA_NMODELS];
#pragma unroll
  for (int m = 0; m < PCA_NMODELS; ++m)
    cnt[m] = A.chunk == 0 ? (int)A.job[m].a.N : A.counts[A.chunk * PCA_NMODELS + m];
  long long tasks[PCA_NMODELS + 1];
  tasks[0] = 0;
#pragma unroll
  for (int m = 0; m < PCA_NMODELS; ++m)
    tasks[m + 1] = tasks[m] + ((long long)cnt[m] + 64 * PCA_VPL - 1) / (64 * PCA_VPL);
  const long long total = tasks[PCA_NMODELS];
#if defined(PCA_UC_LDS) && !defined(PCA_CONSTS_PER_VOXEL)
  pca::load_uc_lds<0>(A);
  __syncthreads();
#endif
  const long long waves_per_block = PCA_BLOCK / 64;
  const long long nwaves = (long long)gridDim.x * waves_per_block;
  for (long long t = (long long)blockIdx.x * waves_per_block + threadIdx.x / 64; t < total;
       t += nwaves) {
    int m = 0;
#pragma unroll
    for (int q = 1; q < PCA_NMODELS; ++q)
      if (t >= tasks[q]) m = q;
    const long long e0 = (t - tasks[m]) * (64 * PCA_VPL) + pca::lane_id();
    pca::dispatch<0>(m, A, e0, cnt[m]);
  }
}

#endif  // PCA_NO_CHUNK

// counts[0:NMODELS] = first[0:NMODELS], the rest 0 (no memset in a captured graph)
struct pca_counts_init {
  int first[8];
};
extern "C" __global__ void pca_reset_counts(int* counts, int n, pca_counts_init init) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    counts[i] = i < PCA_NMODELS ? init.first[i] : 0;
}

// The state setup() establishes (lsqnonlin.py:1036-1056): lambda = 5,
// step_type = 1, plus f = 0 and n_iter = 0; a kernel so that a captured graph
// can restart a fit without host work.
extern "C" __global__ void pca_lm_reset_state(pca_lm_args a) {
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < a.N;
       v += (long long)gridDim.x * blockDim.x) {
    a.lam[v] = (real)5;
    a.step_type[v] = 1;
    a.f[v] = (real)0;
    if (a.n_iter) a.n_iter[v] = 0;
  }
}
