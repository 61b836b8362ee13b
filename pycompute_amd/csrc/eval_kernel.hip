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
#undef PCA_ECONST
