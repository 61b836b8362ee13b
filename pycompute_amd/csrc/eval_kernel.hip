// Standalone model kernels over the same generated `Model` struct as
// lm_kernel.hip (concatenated after prelude.hip and the model struct).
//
// pca_ssr: the reference's F kernel (pycompute/cuda/lsqnonlin.py:46-199,
// emitted .cu:75-93): f[v] += sum_i (model(p_v; c_i) - y_i)^2, accumulated
// onto the existing f[v] in measurement order; voxels whose step_type is 0
// are skipped (the kernel wrapper at lsqnonlin.py:172-186). step_type may be
// null (no skipping).
//
// pca_eval: model value, Jacobian and packed second derivatives per
// (measurement, voxel): m (M, N), J (NP*M, N) rows j*M+i, d2 (NH*M, N) rows
// l*M+i. Any output may be null. It is the forward model used to generate
// training signals and the checkable surface of the code generator.

struct pca_ssr_args {
  const real* params;            // (NP, N)
  const real* data;              // (M, N)
  const real* consts;            // (NC*M, N) when per-voxel
  const signed char* step_type;  // (N) or null
  real* f;                       // (N) accumulated
  long long N;
};

struct pca_eval_args {
  const real* params;  // (NP, N)
  const real* consts;  // (NC*M, N) when per-voxel
  real* m;             // (M, N) or null
  real* J;             // (NP*M, N) or null
  real* d2;            // (NH*M, N) or null
  long long N;
};

#ifndef PCA_UC_STRUCT_DEFINED
#define PCA_UC_STRUCT_DEFINED
struct pca_uconsts {
  real c[PCA_UC_MAX];
};
#endif

#ifdef PCA_CONSTS_PER_VOXEL
#define PCA_ECONST(k, i) a.consts[((long long)(k) * PCA_M + (i)) * N + v]
#else
#define PCA_ECONST(k, i) uc.c[(k) * PCA_M + (i)]
#endif

extern "C" __global__ void __launch_bounds__(PCA_BLOCK)
pca_ssr(pca_ssr_args a, pca_uconsts uc) {
  constexpr int NP = Model::NP, NC = Model::NC;
  const long long N = a.N;
  const long long v = (long long)blockIdx.x * PCA_BLOCK + threadIdx.x;
  if (v >= N) return;
  if (a.step_type && a.step_type[v] == 0) return;
  real p[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) p[j] = a.params[(long long)j * N + v];
  real f = a.f[v];
  for (int i = 0; i < PCA_M; ++i) {
    real c[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) c[k] = PCA_ECONST(k, i);
    const real r = Model::resid(p, c, a.data[(long long)i * N + v]);
    f += r * r;
  }
  a.f[v] = f;
}

extern "C" __global__ void __launch_bounds__(PCA_BLOCK)
pca_eval(pca_eval_args a, pca_uconsts uc) {
  constexpr int NP = Model::NP, NC = Model::NC, NH = Model::NH;
  const long long N = a.N;
  const long long v = (long long)blockIdx.x * PCA_BLOCK + threadIdx.x;
  if (v >= N) return;
  real p[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) p[j] = a.params[(long long)j * N + v];
  for (int i = 0; i < PCA_M; ++i) {
    real c[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) c[k] = PCA_ECONST(k, i);
    real m, J[NP], d2[NH];
    Model::fjh(p, c, m, J, d2);
    if (a.m) a.m[(long long)i * N + v] = m;
    if (a.J) {
#pragma unroll
      for (int j = 0; j < NP; ++j) a.J[((long long)j * PCA_M + i) * N + v] = J[j];
    }
    if (a.d2) {
#pragma unroll
      for (int l = 0; l < NH; ++l) a.d2[((long long)l * PCA_M + i) * N + v] = d2[l];
    }
  }
}

#ifndef PCA_CONSTS_PER_VOXEL
// pca_pick_init: of two initial guesses keep, per voxel, the one with the
// lower SSR: a (in/out) is replaced by b unless SSR(a) <= SSR(b) (so a
// non-finite SSR of a always yields b). Used to guard a learned initial guess
// with the default one. Voxel-uniform consts.
struct pca_pick_args {
  real* a;            // (NP, N) in/out
  const real* b;      // (NP, N)
  const real* data;   // (M, N)
  long long N;
};

extern "C" __global__ void __launch_bounds__(PCA_BLOCK)
pca_pick_init(pca_pick_args a, pca_uconsts uc) {
  constexpr int NP = Model::NP, NC = Model::NC;
  const long long N = a.N;
  const long long v = (long long)blockIdx.x * PCA_BLOCK + threadIdx.x;
  if (v >= N) return;
  real pa[NP], pb[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    pa[j] = a.a[(long long)j * N + v];
    pb[j] = a.b[(long long)j * N + v];
  }
  real fa = 0, fb = 0;
  for (int i = 0; i < PCA_M; ++i) {
    real c[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) c[k] = PCA_ECONST(k, i);
    const real y = a.data[(long long)i * N + v];
    const real ra = Model::resid(pa, c, y), rb = Model::resid(pb, c, y);
    fa += ra * ra;
    fb += rb * rb;
  }
  if (!(fa <= fb)) {
#pragma unroll
    for (int j = 0; j < NP; ++j) a.a[(long long)j * N + v] = pb[j];
  }
}
#endif
#undef PCA_ECONST

// ---------------------------------------------------------------------------
// Synthetic training signals from the model itself: parameters drawn
// uniformly in [lo, hi] (or taken from truth_in), model evaluated at every
// measurement, Rician noise y = |m + sigma (n1 + i n2)| with
// sigma = m(measurement 0) / SNR and SNR uniform in [snr_lo, snr_hi].
// Random numbers: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC 2011) keyed by the seed, counter = (sample
// index, draw index); normals by the Box-Muller transform.
// ---------------------------------------------------------------------------

struct pca_synth_args {
  real* data;               // (M, N) out
  real* truth;              // (NP, N) out (sampled or copied)
  const real* truth_in;     // (NP, N) or null: use these parameters
  const real* lo;           // (NP) sampling box
  const real* hi;           // (NP)
  double snr_lo, snr_hi;
  unsigned long long seed;
  long long N;
  long long offset;         // sample index of entry 0 (counter base)
};

namespace pca {

struct U4 {
  unsigned x, y, z, w;
};

__device__ __forceinline__ U4 philox4x32_10(U4 c, unsigned k0, unsigned k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const unsigned long long p0 = (unsigned long long)0xD2511F53u * c.x;
    const unsigned long long p1 = (unsigned long long)0xCD9E8D57u * c.z;
    const unsigned hi0 = (unsigned)(p0 >> 32), lo0 = (unsigned)p0;
    const unsigned hi1 = (unsigned)(p1 >> 32), lo1 = (unsigned)p1;
    c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// uniform in (0, 1)
__device__ __forceinline__ double u01(unsigned a) { return ((double)a + 0.5) * 2.3283064365386963e-10; }

}  // namespace pca

extern "C" __global__ void __launch_bounds__(PCA_BLOCK)
pca_synth(pca_synth_args a, pca_uconsts uc) {
  constexpr int NP = Model::NP, NC = Model::NC;
  const long long v = (long long)blockIdx.x * PCA_BLOCK + threadIdx.x;
  if (v >= a.N) return;
  const unsigned long long id = (unsigned long long)(a.offset + v);
  const unsigned k0 = (unsigned)a.seed, k1 = (unsigned)(a.seed >> 32);
  unsigned draw = 0;
  auto next4 = [&]() {
    pca::U4 c{(unsigned)id, (unsigned)(id >> 32), draw++, 0x5eedu};
    return pca::philox4x32_10(c, k0, k1);
  };
  real p[NP];
  if (a.truth_in) {
#pragma unroll
    for (int j = 0; j < NP; ++j) p[j] = a.truth_in[(long long)j * a.N + v];
  } else {
    for (int j = 0; j < NP; j += 4) {
      const pca::U4 r = next4();
      const unsigned rr[4] = {r.x, r.y, r.z, r.w};
      for (int q = 0; q < 4 && j + q < NP; ++q)
        p[j + q] = (real)((double)a.lo[j + q] + ((double)a.hi[j + q] - (double)a.lo[j + q]) * pca::u01(rr[q]));
    }
  }
#pragma unroll
  for (int j = 0; j < NP; ++j) a.truth[(long long)j * a.N + v] = p[j];
  const pca::U4 rs = next4();
  const double snr = a.snr_lo + (a.snr_hi - a.snr_lo) * pca::u01(rs.x);
  double sigma = 0.0;
  for (int i = 0; i < PCA_M; ++i) {
    real c[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) c[k] = uc.c[k * PCA_M + i];
    const double m = (double)Model::resid(p, c, (real)0);
    if (i == 0) sigma = m / snr;
    const pca::U4 r = next4();
    const double rad = sqrt(-2.0 * log(pca::u01(r.x)));
    const double th = 6.283185307179586 * pca::u01(r.y);
    const double n1 = rad * cos(th), n2 = rad * sin(th);
    const double re = m + sigma * n1, im = sigma * n2;
    a.data[(long long)i * a.N + v] = (real)sqrt(re * re + im * im);
  }
}

// ---------------------------------------------------------------------------
// Random search (pycompute/cuda/lsqnonlin.py:1082-1242: SearchStep,
// RandomSearch, SecondOrderRandomSearch). Candidates are uniform in
// [lb, ub] per voxel; candidate k of voxel v takes its NP uniforms from
// Philox4x32-10 with key = seed and counter (v, v >> 32, k, 0x7a110000 + j/4)
// (u = (x + 0.5) 2^-32), so a search can be continued and replayed. A
// candidate's score is its own SSR (the reference accumulates every
// candidate's SSR into one never-cleared f, defect D11).
// ---------------------------------------------------------------------------

#ifdef PCA_CONSTS_PER_VOXEL
#define PCA_RCONST(k, i) a.consts[((long long)(k) * PCA_M + (i)) * N + v]
#else
#define PCA_RCONST(k, i) uc.c[(k) * PCA_M + (i)]
#endif

namespace pca {

// the NP uniforms of candidate k of voxel v, scaled to [lo, hi] as
// lo + (hi - lo) u with separate roundings (no FMA contraction, whatever the
// build allows), so the host can replay the candidates bit for bit
template <int NP>
__device__ __forceinline__ void random_candidate(unsigned long long seed, long long v,
                                                 long long k, const real* lo, const real* hi,
                                                 real* p) {
#pragma clang fp contract(off)
  const unsigned k0 = (unsigned)seed, k1 = (unsigned)(seed >> 32);
#pragma unroll
  for (int j0 = 0; j0 < NP; j0 += 4) {
    const U4 c{(unsigned)v, (unsigned)((unsigned long long)v >> 32), (unsigned)k,
               0x7a110000u + (unsigned)(j0 / 4)};
    const U4 r = philox4x32_10(c, k0, k1);
    const unsigned w[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (j0 + q < NP) p[j0 + q] = lo[j0 + q] + (hi[j0 + q] - lo[j0 + q]) * (real)u01(w[q]);
  }
}

template <class CF>
__device__ __forceinline__ real ssr_at(const real* p, const real* data, long long N, long long v,
                                       const CF& cf) {
  constexpr int NC = Model::NC;
  real f = 0;
  for (int i = 0; i < PCA_M; ++i) {
    real c[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) c[k] = cf(k, i);
    const real r = Model::resid(p, c, data[(long long)i * N + v]);
    f += r * r;
  }
  return f;
}

}  // namespace pca

struct pca_rsearch_args {
  const real* data;     // (M, N)
  const real* consts;   // (NC*M, N) when per-voxel
  const real* lb;       // (NP, N)
  const real* ub;       // (NP, N)
  real* best_p;         // (NP, N) in/out
  real* best_f;         // (N) in/out
  unsigned long long seed;
  long long N;
  long long draw0;      // index of the first candidate of this call
  int ncand;            // candidates per voxel in this call
  int reserved;
};

// RandomSearch.run: ncand candidates per voxel in one launch, best kept.
extern "C" __global__ void __launch_bounds__(PCA_BLOCK)
pca_random_search(pca_rsearch_args a, pca_uconsts uc) {
  constexpr int NP = Model::NP;
  const long long N = a.N;
  const long long v = (long long)blockIdx.x * PCA_BLOCK + threadIdx.x;
  if (v >= N) return;
  real lo[NP], hi[NP], bp[NP], p[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    lo[j] = a.lb[(long long)j * N + v];
    hi[j] = a.ub[(long long)j * N + v];
    bp[j] = a.best_p[(long long)j * N + v];
  }
  real bf = a.best_f[v];
  auto cf = [&](int k, int i) { return PCA_RCONST(k, i); };
  for (int k = 0; k < a.ncand; ++k) {
    pca::random_candidate<NP>(a.seed, v, a.draw0 + k, lo, hi, p);
    const real f = pca::ssr_at(p, a.data, N, v, cf);
    if (f < bf) {
      bf = f;
#pragma unroll
      for (int j = 0; j < NP; ++j) bp[j] = p[j];
    }
  }
#pragma unroll
  for (int j = 0; j < NP; ++j) a.best_p[(long long)j * N + v] = bp[j];
  a.best_f[v] = bf;
}

struct pca_rparams_args {
  const real* lb;       // (NP, N)
  const real* ub;       // (NP, N)
  real* p;              // (NP, N) out: candidate `draw` of every voxel
  unsigned long long seed;
  long long N;
  long long draw;
};

// The random starting points of SecondOrderRandomSearch (the same stream as
// pca_random_search).
extern "C" __global__ void __launch_bounds__(PCA_BLOCK) pca_random_params(pca_rparams_args a) {
  constexpr int NP = Model::NP;
  const long long N = a.N;
  const long long v = (long long)blockIdx.x * PCA_BLOCK + threadIdx.x;
  if (v >= N) return;
  real lo[NP], hi[NP], p[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    lo[j] = a.lb[(long long)j * N + v];
    hi[j] = a.ub[(long long)j * N + v];
  }
  pca::random_candidate<NP>(a.seed, v, a.draw, lo, hi, p);
#pragma unroll
  for (int j = 0; j < NP; ++j) a.p[(long long)j * N + v] = p[j];
}

struct pca_keep_args {
  const real* cand;     // (NP, N) candidate parameters (an LM result)
  const real* data;     // (M, N)
  const real* consts;   // (NC*M, N) when per-voxel
  real* best_p;         // (NP, N) in/out
  real* best_f;         // (N) in/out
  long long N;
};

// SearchStep after an LM run: keep the candidate when its SSR is lower.
extern "C" __global__ void __launch_bounds__(PCA_BLOCK)
pca_keep_best(pca_keep_args a, pca_uconsts uc) {
  constexpr int NP = Model::NP;
  const long long N = a.N;
  const long long v = (long long)blockIdx.x * PCA_BLOCK + threadIdx.x;
  if (v >= N) return;
  real p[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) p[j] = a.cand[(long long)j * N + v];
  auto cf = [&](int k, int i) { return PCA_RCONST(k, i); };
  const real f = pca::ssr_at(p, a.data, N, v, cf);
  if (f < a.best_f[v]) {
    a.best_f[v] = f;
#pragma unroll
    for (int j = 0; j < NP; ++j) a.best_p[(long long)j * N + v] = p[j];
  }
}
#undef PCA_RCONST
