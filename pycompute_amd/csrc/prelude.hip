// Common prelude of every generated module: value type, constants and math
// helpers shared by the model struct (emitted by pycompute_amd/sym/hipgen.py)
// and the kernels (lm_kernel.hip, eval_kernel.hip). The code generator
// concatenates: prelude, model struct(s), kernel template(s); hiprtc compiles
// the result for gfx950.

#ifndef PCA_REAL_DOUBLE
typedef float real;
#define PCA_EPS_GMW ((float)2e-7)
#define PCA_DAMP_FLOOR 1e-4f
#else
typedef double real;
#define PCA_EPS_GMW 4e-16
#define PCA_DAMP_FLOOR 1e-4
#endif

#ifndef PCA_UC_MAX
#define PCA_UC_MAX 256
#endif

#ifdef PCA_FAST_EXP
// v_exp_f32 on x*log2(e): relative error about |x| * 6e-8.
#define PCA_EXPF(x) __expf(x)
#define PCA_LOGF(x) __logf(x)
#else
#define PCA_EXPF(x) expf(x)
#define PCA_LOGF(x) logf(x)
#endif

__device__ __forceinline__ float pca_fmax(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ double pca_fmax(double a, double b) { return fmax(a, b); }
__device__ __forceinline__ float pca_abs(float a) { return fabsf(a); }
__device__ __forceinline__ double pca_abs(double a) { return fabs(a); }
__device__ __forceinline__ float pca_sqrt(float a) { return sqrtf(a); }
__device__ __forceinline__ double pca_sqrt(double a) { return sqrt(a); }
__device__ __forceinline__ bool pca_nonfinite(float a) { return __builtin_isnan(a) || __builtin_isinf(a); }
__device__ __forceinline__ bool pca_nonfinite(double a) { return __builtin_isnan(a) || __builtin_isinf(a); }
