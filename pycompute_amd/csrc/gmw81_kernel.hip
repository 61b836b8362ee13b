// Batched linear solve hl s = g with the LM step's solver (pivot rotation,
// diagonal scaling, GMW81, LDL^T substitution), one system per lane: the
// standalone solver of pycompute/cuda/solver.py:425-595 (GMW81Solver).
// Packed lower-triangle mat (NH, N) in the order l = i(i+1)/2 + j, rhs and sol
// (NP, N); voxel index fastest. Concatenated after lm_solver.hip with PCA_NP
// and PCA_BLOCK defined.

struct pca_gmw81_args {
  const real* mat;   // (NH, N)
  const real* rhs;   // (NP, N)
  real* sol;         // (NP, N)
  long long N;
};

extern "C" __global__ void __launch_bounds__(PCA_BLOCK) pca_gmw81_batch(pca_gmw81_args a) {
  constexpr int n = PCA_NP, nh = n * (n + 1) / 2;
  const long long v = (long long)blockIdx.x * PCA_BLOCK + threadIdx.x;
  if (v >= a.N) return;
  real h[nh], hd[n], g[n], s[n];
#pragma unroll
  for (int l = 0; l < nh; ++l) h[l] = a.mat[(long long)l * a.N + v];
#pragma unroll
  for (int j = 0; j < n; ++j) {
    hd[j] = h[pca::tri(j, j)];
    g[j] = a.rhs[(long long)j * a.N + v];
  }
  pca::gmw81_solve<n>(h, hd, g, s);
#pragma unroll
  for (int j = 0; j < n; ++j) a.sol[(long long)j * a.N + v] = s[j];
}
