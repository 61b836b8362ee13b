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

// Transcendentals used by the generated model code. The strict build uses the
// reproducible implementations of strict_math.h (concatenated before this
// file) and is compiled with -ffp-contract=off, so the CPU oracle can replay
// it bit for bit. The fast-exp build uses v_exp_f32 on x*log2(e) for float
// (relative error about |x| * 6e-8), and v_log_f32 for log and pow.
__device__ __forceinline__ float pca_exp(float x) { return expf(x); }
__device__ __forceinline__ double pca_exp(double x) { return exp(x); }
__device__ __forceinline__ float pca_exp_fast(float x) { return __expf(x); }
__device__ __forceinline__ double pca_exp_fast(double x) { return exp(x); }
__device__ __forceinline__ float pca_log(float x) { return logf(x); }
// Fast-exp build: v_log_f32 (log2) for log, and pow as 2^(y log2 x) with
// pow(x, 0) = 1; for float, errors of a few ulp times |y log2 x|. Double
// keeps the library functions.
__device__ __forceinline__ float pca_log_fast(float x) {
  return __builtin_amdgcn_logf(x) * 0.69314718055994531f;
}
__device__ __forceinline__ double pca_log_fast(double x) { return log(x); }
__device__ __forceinline__ float pca_pow_fast(float x, float y) {
  return y == 0.0f ? 1.0f : __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
}
__device__ __forceinline__ double pca_pow_fast(double x, double y) { return pow(x, y); }
__device__ __forceinline__ double pca_log(double x) { return log(x); }
__device__ __forceinline__ float pca_pow(float x, float y) { return powf(x, y); }
__device__ __forceinline__ double pca_pow(double x, double y) { return pow(x, y); }

#if defined(PCA_STRICT)
#define PCA_EXP(x) pca_exp_strict(x)
#define PCA_LOG(x) pca_log_strict(x)
#define PCA_POW(x, y) pca_pow_strict(x, y)
#elif defined(PCA_FAST_EXP)
#define PCA_EXP(x) pca_exp_fast(x)
#define PCA_LOG(x) pca_log_fast(x)
#define PCA_POW(x, y) pca_pow_fast(x, y)
#else
#define PCA_EXP(x) pca_exp(x)
#define PCA_LOG(x) pca_log(x)
#define PCA_POW(x, y) pca_pow(x, y)
#endif

// exp(k*x) for a loop-invariant k (the factored accumulation): the fast-exp
// float build folds log2(e) into k, so the loop pays one multiply.
#if defined(PCA_FAST_EXP) && !defined(PCA_REAL_DOUBLE)
#define PCA_EXPK(k, x) __builtin_amdgcn_exp2f(((k) * 1.44269504088896341f) * (x))
#else
#define PCA_EXPK(k, x) PCA_EXP((k) * (x))
#endif
// exp(k*x) with k read from the derived consts, already scaled by log2(e)
// on the host when the fast-exp float build evaluates exp through exp2.
#if defined(PCA_FAST_EXP) && !defined(PCA_REAL_DOUBLE)
#define PCA_EXPKD(ks, x) __builtin_amdgcn_exp2f((ks) * (x))
#else
#define PCA_EXPKD(ks, x) PCA_EXP((ks) * (x))
#endif

__device__ __forceinline__ float pca_fmax(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ double pca_fmax(double a, double b) { return fmax(a, b); }
__device__ __forceinline__ float pca_abs(float a) { return fabsf(a); }
__device__ __forceinline__ double pca_abs(double a) { return fabs(a); }
__device__ __forceinline__ float pca_sqrt(float a) { return sqrtf(a); }
__device__ __forceinline__ double pca_sqrt(double a) { return sqrt(a); }
__device__ __forceinline__ float pca_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double pca_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }
// Hardware reciprocal and square root (v_rcp_f32, v_sqrt_f32: about 1 ulp,
// no denormal scaling), used by the fast solve (PCA_FAST_SOLVE).
__device__ __forceinline__ float pca_rcp_fast(float a) { return __builtin_amdgcn_rcpf(a); }
__device__ __forceinline__ double pca_rcp_fast(double a) { return 1.0 / a; }
__device__ __forceinline__ float pca_sqrt_fast(float a) { return __builtin_amdgcn_sqrtf(a); }
__device__ __forceinline__ double pca_sqrt_fast(double a) { return sqrt(a); }
// A wave-uniform value held in a VGPR. On gfx950 a VALU instruction that
// reads an SGPR operand issues at about half the rate of its all-VGPR (or
// literal) form, so constants that a hot loop reads many times are pinned
// to a VGPR once (the empty asm makes the value opaque, hence per-lane).
__device__ __forceinline__ float pca_vgpr(float x) {
  asm volatile("" : "+v"(x));
  return x;
}
__device__ __forceinline__ double pca_vgpr(double x) {
  asm volatile("" : "+v"(x));
  return x;
}
// max(L*uu, fl) - L*uu = max(fl - L*uu, 0), fl the damping floor 1e-4: the
// floor's excess over L*uu (factored accumulation). For float one FMA with
// the clamp modifier: the value lies below 1e-4, so clamping to [0, 1] is the
// max. fl comes from a VGPR (pca_vgpr), keeping the FMA at the full rate.
__device__ __forceinline__ float pca_floor_excess(float L, float uu, float fl) {
  return __builtin_amdgcn_fmed3f(__builtin_fmaf(-L, uu, fl), 0.0f, 1.0f);
}
__device__ __forceinline__ double pca_floor_excess(double L, double uu, double fl) {
  return fmax(__builtin_fma(-L, uu, fl), 0.0);
}
// exp(k_i x) at measurement i of an exponential whose factor row k is an
// arithmetic progression, k_i = k_0 + i d (exponential recurrence, factored
// builds with PCA_EXP_REC = R > 0): every R-th measurement evaluates it
// directly, the ones between multiply the previous value by q = exp(d x),
// which measurement 0 evaluates. ks = d (scaled like k). With the measurement
// loop unrolled, i is a constant and the branches fold away. Relative error:
// at most R - 1 rounded products and the error of q times R - 1.
#ifndef PCA_EXP_REC
#define PCA_EXP_REC 0
#endif
template <int R>
__device__ __forceinline__ real pca_exp_rec(int i, real& e, real& q, real k, real ks, real x) {
  static_assert(R > 0, "the exponential recurrence needs PCA_EXP_REC > 0");
  if (i == 0) q = PCA_EXPKD(ks, x);
  if (i % R == 0)
    e = PCA_EXPKD(k, x);
  else
    e = e * q;
  return e;
}
__device__ __forceinline__ bool pca_nonfinite(float a) { return __builtin_isnan(a) || __builtin_isinf(a); }
__device__ __forceinline__ bool pca_nonfinite(double a) { return __builtin_isnan(a) || __builtin_isinf(a); }
