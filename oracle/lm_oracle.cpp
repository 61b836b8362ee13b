// CPU oracle for the batched per-voxel second-order Levenberg-Marquardt fitter.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker for the HIP fitter in
// pycompute_amd/: only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load it. The product path never links or calls it.
//
// It restates, operation by operation, the reference LM iteration as emitted
// by pycompute.cuda (turbotage/pycompute), templated on the value type so the
// same code is the FP32 restatement (T=float) and the FP64 oracle (T=double):
//
//   * one LM iteration  ... test/ivim/second_order_levenberg_marquardt_4_1_21_f_f7def86f0a03.cu:416-517
//                           (template: pycompute/cuda/lsqnonlin.py:820-947)
//   * fghhl accumulation .. .cu:2-73        (lsqnonlin.py:206-346)
//   * f (SSR)            .. .cu:75-93       (lsqnonlin.py:46-117)
//   * diagonal pivot     .. .cu:95-150      (pycompute/cuda/permute.py:44-204); the running max
//                           is never updated (permute.py:50-56), so the pivot is the rotation
//                           [n-1, 0, ..., n-2] unless a diagonal is NaN -- restated literally here
//   * GMW81              .. .cu:152-224     (pycompute/cuda/solver.py:20-115)
//   * LDL^T solve        .. .cu:234-263     (solver.py:295-401)
//   * gmw81 solver       .. .cu:273-305     (solver.py:428-488)
//   * gain ratio         .. .cu:307-365     (lsqnonlin.py:432-498), without the printf
//   * clamp              .. .cu:367-382     (lsqnonlin.py:584-607)
//   * convergence        .. .cu:384-414     (lsqnonlin.py:691-729)
//   * host loop          .. lsqnonlin.py:1036-1073 (lambda0=5, step_type0=1, first_f/last_f)
//
// Deliberate differences from the reference, each a defect fix:
//   * consts index (k*M+i)*N+v instead of k*nconst*N+i*N+v (only differs for nconst>1)
//   * f loads n parameters, not ndata (out-of-bounds in the reference)
//   * the FP64 instantiation is FP64 throughout (the reference hard-codes some floats)
//
// Models: IVIM follows the reference's own emitted CSE (.cu:19-47) literally.
// Mono-exponential T2, stretched exponential and the 22-parameter DKI
// exp-linear model are hand-written closed forms, independent of the project's
// symbolic code generator, so that the generator is checked against them.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <omp.h>

namespace {

constexpr int kMaxNP = 22;
constexpr int kMaxNH = kMaxNP * (kMaxNP + 1) / 2;

enum ModelId { kIVIM = 0, kT2 = 1, kSE = 2, kDKI = 3 };

template <typename T> inline T t_exp(T x);
template <> inline float t_exp<float>(float x) { return expf(x); }
template <> inline double t_exp<double>(double x) { return exp(x); }
template <typename T> inline T t_log(T x);
template <> inline float t_log<float>(float x) { return logf(x); }
template <> inline double t_log<double>(double x) { return log(x); }
template <typename T> inline T t_pow(T x, T y);
template <> inline float t_pow<float>(float x, float y) { return powf(x, y); }
template <> inline double t_pow<double>(double x, double y) { return pow(x, y); }
template <typename T> inline T t_abs(T x) { return std::fabs(x); }
template <typename T> inline T t_sqrt(T x);
template <> inline float t_sqrt<float>(float x) { return sqrtf(x); }
template <> inline double t_sqrt<double>(double x) { return sqrt(x); }
template <typename T> inline T t_fmax(T a, T b);
template <> inline float t_fmax<float>(float a, float b) { return fmaxf(a, b); }
template <> inline double t_fmax<double>(double a, double b) { return fmax(a, b); }

// Machine epsilon constants of gmw81 (solver.py:102-109): float '2e-7', double '4e-16'.
template <typename T> inline T gmw_eps();
template <> inline float gmw_eps<float>() { return (float)2e-7; }
template <> inline double gmw_eps<double>() { return 4e-16; }

// Damping floor of fghhl (lsqnonlin.py:271-276): 1e-4f / 1e-4.
template <typename T> inline T damp_floor();
template <> inline float damp_floor<float>() { return 1e-4f; }
template <> inline double damp_floor<double>() { return 1e-4; }

// ---------------------------------------------------------------------------
// Models. c points at the consts of one measurement: c[k] is const k.
// model_f:   the model value (the reference's F kernel evaluates it without
//            CSE substitutions: lsqnonlin.py:46-117).
// model_fjh: model value m, Jacobian J[n] and packed second derivatives
//            d2[l], l = j(j+1)/2 + k (k <= j), in the CSE order of
//            lsqnonlin.py:206-346.
// ---------------------------------------------------------------------------

// DKI direction monomials: 6 diffusion and 15 kurtosis weights per direction.
template <typename T>
inline void dki_weights(const T* c, T* a /*6*/, T* w /*15*/) {
  const T gx = c[1], gy = c[2], gz = c[3];
  a[0] = gx * gx; a[1] = gy * gy; a[2] = gz * gz;
  a[3] = T(2) * gx * gy; a[4] = T(2) * gx * gz; a[5] = T(2) * gy * gz;
  const T x2 = gx * gx, y2 = gy * gy, z2 = gz * gz;
  w[0] = x2 * x2; w[1] = y2 * y2; w[2] = z2 * z2;
  w[3] = T(4) * x2 * gx * gy; w[4] = T(4) * x2 * gx * gz;
  w[5] = T(4) * gx * y2 * gy; w[6] = T(4) * y2 * gy * gz;
  w[7] = T(4) * gx * z2 * gz; w[8] = T(4) * gy * z2 * gz;
  w[9] = T(6) * x2 * y2; w[10] = T(6) * x2 * z2; w[11] = T(6) * y2 * z2;
  w[12] = T(12) * x2 * gy * gz; w[13] = T(12) * gx * y2 * gz; w[14] = T(12) * gx * gy * z2;
}

// d log(signal) / d p_k for the DKI exp-linear model.
template <typename T>
inline void dki_dlin(const T* c, T* x /*22*/) {
  T a[6], w[15];
  dki_weights(c, a, w);
  const T b = c[0];
  const T bb6 = b * b / T(6);
  x[0] = T(1);
  for (int k = 0; k < 6; ++k) x[1 + k] = -b * a[k];
  for (int k = 0; k < 15; ++k) x[7 + k] = bb6 * w[k];
}

template <typename T>
inline T model_f(int model, const T* p, const T* c) {
  switch (model) {
    case kIVIM:  // S0*(exp(-D_1*b)*f + exp(-D_2*b)*(1-f))  (.cu:88)
      return p[0] * (t_exp(-p[2] * c[0]) * p[1] + t_exp(-p[3] * c[0]) * (T(1) - p[1]));
    case kT2:  // S0*exp(-TE/T2)
      return p[0] * t_exp(-c[0] / p[1]);
    case kSE:  // S0*exp(-(b*D)**alpha)
      return p[0] * t_exp(-t_pow(c[0] * p[1], p[2]));
    case kDKI: {
      T x[22];
      dki_dlin(c, x);
      T lin = p[0];
      for (int k = 1; k < 22; ++k) lin += x[k] * p[k];
      return t_exp(lin);
    }
  }
  return T(0);
}

template <typename T>
inline void model_fjh(int model, const T* p, const T* c, T& m, T* J, T* d2) {
  switch (model) {
    case kIVIM: {
      // Literal order of the reference's emitted CSE, .cu:19-47.
      const T b = c[0];
      T x0 = t_exp(-p[2] * b);
      T x1 = x0 * p[1];
      T x2 = t_exp(-p[3] * b);
      T x3 = x2 * (T(1) - p[1]);
      T x4 = x1 + x3;
      T x5 = x0 - x2;
      T x6 = p[0] * b;
      T x7 = x3 * b;
      T x8 = p[0] * (b * b);
      m = x4 * p[0];
      J[0] = x4;
      J[1] = x5 * p[0];
      J[2] = -x1 * x6;
      J[3] = -x7 * p[0];
      d2[0] = T(0);
      d2[1] = x5;
      d2[2] = T(0);
      d2[3] = -x1 * b;
      d2[4] = -x0 * x6;
      d2[5] = x1 * x8;
      d2[6] = -x7;
      d2[7] = x2 * x6;
      d2[8] = T(0);
      d2[9] = x3 * x8;
      return;
    }
    case kT2: {
      // m = S0*e, e = exp(-TE/T2); dm/dT2 = S0*e*TE/T2^2
      const T te = c[0];
      T inv = T(1) / p[1];
      T q = te * inv * inv;           // TE/T2^2
      T e = t_exp(-te * inv);
      m = p[0] * e;
      J[0] = e;
      J[1] = p[0] * e * q;
      d2[0] = T(0);
      d2[1] = e * q;
      // d/dT2 (S0 e TE/T2^2) = S0 e (TE^2/T2^4 - 2 TE/T2^3)
      d2[2] = p[0] * e * (q * q - T(2) * q * inv);
      return;
    }
    case kSE: {
      // u = b*D, v = u^alpha, e = exp(-v), m = S0 e
      const T b = c[0];
      T u = b * p[1];
      T v = t_pow(u, p[2]);
      T e = t_exp(-v);
      T lu = t_log(u);
      T dvdD = p[2] * v / p[1];      // alpha u^alpha / D
      T dvda = v * lu;               // u^alpha ln u
      m = p[0] * e;
      J[0] = e;
      J[1] = -p[0] * e * dvdD;
      J[2] = -p[0] * e * dvda;
      d2[0] = T(0);
      d2[1] = -e * dvdD;                                            // (D,S0)
      d2[2] = p[0] * e * (dvdD * dvdD - p[2] * (p[2] - T(1)) * v / (p[1] * p[1]));  // (D,D)
      d2[3] = -e * dvda;                                            // (alpha,S0)
      // d/dalpha (-S0 e alpha v / D) = -S0 e [ v/D + alpha v lu / D - alpha v dvda / D ]
      d2[4] = -p[0] * e * (v / p[1] + dvdD * lu - dvdD * dvda);     // (alpha,D)
      d2[5] = p[0] * e * (dvda * dvda - v * lu * lu);               // (alpha,alpha)
      return;
    }
    case kDKI: {
      T x[22];
      dki_dlin(c, x);
      T lin = p[0];
      for (int k = 1; k < 22; ++k) lin += x[k] * p[k];
      m = t_exp(lin);
      int l = 0;
      for (int j = 0; j < 22; ++j) {
        J[j] = m * x[j];
        for (int k = 0; k <= j; ++k) d2[l++] = J[j] * x[k];
      }
      return;
    }
  }
}

inline int model_nparam(int model) {
  switch (model) {
    case kIVIM: return 4;
    case kT2: return 2;
    case kSE: return 3;
    case kDKI: return 22;
  }
  return -1;
}

inline int model_nconst(int model) { return model == kDKI ? 4 : 1; }

// ---------------------------------------------------------------------------
// Dense per-voxel solver: permute.py / solver.py restated.
// ---------------------------------------------------------------------------

// permute.py:44-66: the running max is never updated, so this returns the last
// index in [offset, n) whose |diagonal| compares greater than -1 (not NaN).
template <typename T>
int max_diag_abs(int n, const T* mat, int offset) {
  T max_abs = T(-1);
  int max_index = 0;
  for (int i = offset; i < n; ++i)
    if (t_abs(mat[i * n + i]) > max_abs) max_index = i;
  return max_index;
}

// permute.py:88-204
template <typename T>
void diag_pivot(int n, T* mat, int* perm) {
  for (int i = 0; i < n; ++i) perm[i] = i;
  for (int i = 0; i < n; ++i) {
    int mi = max_diag_abs(n, mat, i);
    for (int k = 0; k < n; ++k) std::swap(mat[i * n + k], mat[mi * n + k]);
    for (int k = 0; k < n; ++k) std::swap(mat[k * n + i], mat[k * n + mi]);
    std::swap(perm[i], perm[mi]);
  }
}

// solver.py:20-115, in place on the lower triangle of a full row-major matrix.
template <typename T>
void gmw81(int n, T* mat) {
  T t0;
  T t1 = T(0);  // gamma
  T t2 = T(0);  // xi
  T beta2 = gmw_eps<T>();
  T delta = gmw_eps<T>();
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j <= i; ++j) {
      t0 = t_abs(mat[i * n + j]);
      if (i == j) {
        if (t0 > t1) t1 = t0;
      } else {
        if (t0 > t2) t2 = t0;
      }
    }
  }
  if (n > 1) t2 /= t_sqrt(T(n * n - 1));
  if (beta2 < t1) beta2 = t1;
  if (beta2 < t2) beta2 = t2;
  t0 = t1 + t2;
  if (t0 > T(1)) delta *= t0;

  for (int j = 0; j < n; ++j) {
    for (int s = 0; s < j; ++s) mat[j * n + s] /= mat[s * n + s];
    for (int i = j + 1; i < n; ++i) {
      t0 = mat[i * n + j];
      for (int s = 0; s < j; ++s) t0 -= mat[j * n + s] * mat[i * n + s];
      mat[i * n + j] = t0;
    }
    t1 = T(0);
    for (int i = j + 1; i < n; ++i) {
      t0 = t_abs(mat[i * n + j]);
      if (t1 < t0) t1 = t0;
    }
    t1 *= t1;
    t2 = t_abs(mat[j * n + j]);
    if (t2 < delta) t2 = delta;
    t0 = t1 / beta2;
    if (t2 < t0) t2 = t0;
    mat[j * n + j] = t2;
    for (int i = j + 1; i < n; ++i) {
      t0 = mat[i * n + j];
      mat[i * n + i] -= t0 * t0 / t2;
    }
  }
}

// solver.py:295-401
template <typename T>
void ldl_solve(int n, const T* mat, const T* rhs, T* sol) {
  T arr[kMaxNP];
  for (int i = 0; i < n; ++i) {
    arr[i] = rhs[i];
    for (int j = 0; j < i; ++j) arr[i] -= mat[i * n + j] * mat[j * n + j] * arr[j];
    arr[i] /= mat[i * n + i];
  }
  for (int i = n - 1; i >= 0; --i) {
    sol[i] = arr[i];
    for (int j = i + 1; j < n; ++j) sol[i] -= mat[j * n + i] * sol[j];
  }
}

// solver.py:428-488: pivot, scale, GMW81, LDL solve, unscale, unpivot.
template <typename T>
void gmw81_solver(int n, T* mat, const T* rhs, T* sol) {
  int perm[kMaxNP];
  T arr1[kMaxNP], arr2[kMaxNP], scale[kMaxNP];
  diag_pivot(n, mat, perm);
  for (int i = 0; i < n; ++i) arr1[i] = rhs[perm[i]];
  for (int i = 0; i < n; ++i) {
    scale[i] = t_sqrt(t_abs(mat[i * n + i]));
    arr1[i] /= scale[i];
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j <= i; ++j) mat[i * n + j] /= (scale[i] * scale[j]);
  gmw81(n, mat);
  ldl_solve(n, mat, arr1, arr2);
  for (int i = 0; i < n; ++i) arr2[i] /= scale[i];
  for (int i = 0; i < n; ++i) sol[perm[i]] = arr2[i];
}

// ---------------------------------------------------------------------------
// One voxel's LM state and iteration (.cu:416-517).
// ---------------------------------------------------------------------------

// A generated model (the project's code generator compiled for the host; see
// oracle/genmodel.py): one measurement of the fghhl accumulation and the
// residual, with the same straight-line code the GPU runs. Used by the strict
// build's bit-exact replay; the built-in models above are independent of it.
template <typename T>
using GenAccum = void (*)(const T* p, const T* c, T y, T lam, T* f, T* g, T* h, T* hld);
template <typename T>
using GenResid = T (*)(const T* p, const T* c, T y);

template <typename T>
struct Problem {
  int model, n, nc, M;
  long N;
  const T* consts;
  int consts_per_voxel;
  const T* data;
  const T* lb;
  const T* ub;
  T tol, mu, eta, acc, dec;
  GenAccum<T> gen_accum = nullptr;
  GenResid<T> gen_resid = nullptr;
};

template <typename T>
inline void load_consts(const Problem<T>& P, long v, int i, T* c) {
  for (int k = 0; k < P.nc; ++k) {
    long idx = (long)k * P.M + i;
    c[k] = P.consts_per_voxel ? P.consts[idx * P.N + v] : P.consts[idx];
  }
}

template <typename T>
T ssr(const Problem<T>& P, long v, const T* p) {
  T f = T(0);
  T c[8];
  for (int i = 0; i < P.M; ++i) {
    load_consts(P, v, i, c);
    T res = P.gen_resid ? P.gen_resid(p, c, P.data[(long)i * P.N + v])
                        : model_f(P.model, p, c) - P.data[(long)i * P.N + v];
    f += res * res;
  }
  return f;
}

// One launch's worth of work for voxel v. Returns true when the voxel ran.
// On exit st holds the new step type; f holds the SSR at the params at the
// start of the iteration (the reference's f_t).
template <typename T>
bool lm_iteration(const Problem<T>& P, long v, T* p, T& lam, int8_t& st, T& f) {
  if (st == 0) return false;
  const int n = P.n;
  const int nh = n * (n + 1) / 2;
  T g[kMaxNP], h[kMaxNH], hl[kMaxNH], J[kMaxNP], d2[kMaxNH], hes[kMaxNH];
  T c[8];
  f = T(0);
  for (int j = 0; j < n; ++j) g[j] = T(0);
  for (int l = 0; l < nh; ++l) { h[l] = T(0); hl[l] = T(0); }

  if (P.gen_accum) {
    // generated accumulation: h (undamped, off-diagonal equal to hl) and the
    // damped diagonal hld
    T hld[kMaxNP];
    for (int j = 0; j < n; ++j) hld[j] = T(0);
    for (int i = 0; i < P.M; ++i) {
      load_consts(P, v, i, c);
      P.gen_accum(p, c, P.data[(long)i * P.N + v], lam, &f, g, h, hld);
    }
    int l = 0;
    for (int j = 0; j < n; ++j)
      for (int k = 0; k <= j; ++k, ++l) hl[l] = (j == k) ? hld[j] : h[l];
  }

  // fghhl (.cu:17-72)
  for (int i = 0; i < (P.gen_accum ? 0 : P.M); ++i) {
    load_consts(P, v, i, c);
    T m = T(0);
    model_fjh(P.model, p, c, m, J, d2);
    T res = m - P.data[(long)i * P.N + v];
    for (int l = 0; l < nh; ++l) hes[l] = d2[l] * res;
    f += res * res;
    int l = 0;
    for (int j = 0; j < n; ++j) {
      for (int k = 0; k <= j; ++k) {
        T jt = J[j] * J[k];
        T hj = hes[l] + jt;
        h[l] += hj;
        if (j != k)
          hl[l] += hj;
        else
          hl[l] += hj + t_fmax(lam * jt, damp_floor<T>());
        ++l;
      }
      g[j] += J[j] * res;
    }
  }

  // Solve hl * s = g on a full symmetric copy (.cu:436-471)
  T mat[kMaxNP * kMaxNP], s[kMaxNP];
  {
    int k = 0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j <= i; ++j) {
        mat[i * n + j] = hl[k];
        if (i != j) mat[j * n + i] = hl[k];
        ++k;
      }
  }
  gmw81_solver(n, mat, g, s);
  for (int j = 0; j < n; ++j)
    if (std::isnan(s[j]) || std::isinf(s[j])) s[j] = T(0);

  // Convergence (.cu:384-414)
  {
    bool clamped = false;
    T norm = T(0);
    for (int i = 0; i < n; ++i) {
      T t1 = p[i];
      T t2 = t1 - g[i];
      T u = P.ub[(long)i * P.N + v], lo = P.lb[(long)i * P.N + v];
      if (t2 > u) {
        clamped = true;
        t2 = u;
      } else if (t2 < lo) {
        clamped = true;
        t2 = lo;
      }
      t2 = t1 - t2;
      norm += t2 * t2;
    }
    if (norm < P.tol * (T(1) + f)) {
      if ((st & 1) || clamped) {
        st = 0;
        return true;
      }
    }
  }

  // Trial point and its SSR (.cu:494-505)
  T ptp[kMaxNP];
  for (int j = 0; j < n; ++j) ptp[j] = p[j] - s[j];
  T ftp = ssr(P, v, ptp);

  // Gain ratio (.cu:307-365)
  T actual = T(0.5) * (f - ftp);
  T pred = T(0);
  {
    int k = 0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j <= i; ++j) {
        T e = h[k] * s[i] * s[j];
        if (i == j)
          pred -= e;
        else
          pred -= T(2) * e;
        ++k;
      }
  }
  pred *= T(0.5);
  for (int i = 0; i < n; ++i) pred += s[i] * g[i];
  T rho = actual / pred;
  if ((rho > P.mu) && (actual > T(0))) {
    for (int j = 0; j < n; ++j) p[j] = ptp[j];
    if (rho > P.eta) {
      lam *= P.acc;
      st = 1;
    } else {
      st = 2;
    }
  } else {
    lam *= P.dec;
    st = 4;
  }
  if (pred < T(0)) {
    lam *= P.dec;
    st |= 8;
  }

  // Clamp (.cu:367-382)
  for (int i = 0; i < n; ++i) {
    T u = P.ub[(long)i * P.N + v], lo = P.lb[(long)i * P.N + v];
    if (p[i] > u)
      p[i] = u;
    else if (p[i] < lo)
      p[i] = lo;
  }
  return true;
}

template <typename T>
int lm_run(const Problem<T>& P, T* params, T* lam, int8_t* step_type, T* f, T* first_f, T* last_f,
           int32_t* n_iter, int iters, T* trace_f, T* trace_lam, int8_t* trace_st, int nthreads) {
  if (!P.gen_accum && (P.n != model_nparam(P.model) || P.nc != model_nconst(P.model))) return -1;
  if (P.n < 1 || P.n > kMaxNP || P.nc < 1 || P.nc > 8) return -1;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
  for (long v = 0; v < P.N; ++v) {
    T p[kMaxNP];
    for (int j = 0; j < P.n; ++j) p[j] = params[(long)j * P.N + v];
    T lv = lam[v];
    int8_t sv = step_type[v];
    T fv = f ? f[v] : T(0);
    int32_t cnt = 0;
    for (int it = 0; it < iters; ++it) {
      if (lm_iteration(P, v, p, lv, sv, fv)) ++cnt;
      if (it == 0 && first_f) first_f[v] = fv;
      if (it == iters - 1 && last_f) last_f[v] = fv;
      if (trace_f) trace_f[(long)it * P.N + v] = fv;
      if (trace_lam) trace_lam[(long)it * P.N + v] = lv;
      if (trace_st) trace_st[(long)it * P.N + v] = sv;
    }
    for (int j = 0; j < P.n; ++j) params[(long)j * P.N + v] = p[j];
    lam[v] = lv;
    step_type[v] = sv;
    if (f) f[v] = fv;
    if (n_iter) n_iter[v] += cnt;
  }
  return 0;
}

template <typename T>
int ssr_run(const Problem<T>& P, const T* params, T* out, int nthreads) {
  if (!P.gen_resid && (P.n != model_nparam(P.model) || P.nc != model_nconst(P.model))) return -1;
  if (P.n < 1 || P.n > kMaxNP || P.nc < 1 || P.nc > 8) return -1;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
  for (long v = 0; v < P.N; ++v) {
    T p[kMaxNP];
    for (int j = 0; j < P.n; ++j) p[j] = params[(long)j * P.N + v];
    out[v] = ssr(P, v, p);
  }
  return 0;
}

template <typename T>
int model_eval(int model, int M, long N, const T* consts, int cpv, const T* params, T* out_m,
               T* out_J, T* out_d2) {
  Problem<T> P{};
  P.model = model;
  P.n = model_nparam(model);
  P.nc = model_nconst(model);
  P.M = M;
  P.N = N;
  P.consts = consts;
  P.consts_per_voxel = cpv;
  const int n = P.n, nh = n * (n + 1) / 2;
  for (long v = 0; v < N; ++v) {
    T p[kMaxNP], c[4], J[kMaxNP], d2[kMaxNH];
    for (int j = 0; j < n; ++j) p[j] = params[(long)j * N + v];
    for (int i = 0; i < M; ++i) {
      load_consts(P, v, i, c);
      T m = T(0);
      model_fjh(model, p, c, m, J, d2);
      long row = (long)i * N + v;
      if (out_m) out_m[row] = m;
      for (int j = 0; j < n; ++j)
        if (out_J) out_J[((long)j * M) * N + row] = J[j];
      for (int l = 0; l < nh; ++l)
        if (out_d2) out_d2[((long)l * M) * N + row] = d2[l];
    }
  }
  return 0;
}

}  // namespace

#define ORACLE_EXPORTS(T, SUF)                                                                   \
  extern "C" int oracle_lm_run_##SUF(int model, int n, int nc, int M, long N, const T* consts,    \
                                     int consts_per_voxel, const T* data, const T* lb,           \
                                     const T* ub, T* params, T* lam, int8_t* step_type, T* f,    \
                                     T* first_f, T* last_f, int32_t* n_iter, int iters, T tol,   \
                                     T mu, T eta, T acc, T dec, T* trace_f, T* trace_lam,        \
                                     int8_t* trace_st, int nthreads) {                            \
    Problem<T> P{model, n, nc, M, N, consts, consts_per_voxel, data, lb, ub, tol, mu, eta, acc,  \
                 dec};                                                                           \
    return lm_run(P, params, lam, step_type, f, first_f, last_f, n_iter, iters, trace_f,         \
                  trace_lam, trace_st, nthreads);                                                \
  }                                                                                              \
  extern "C" int oracle_ssr_##SUF(int model, int n, int nc, int M, long N, const T* consts,       \
                                  int consts_per_voxel, const T* data, const T* params, T* out,  \
                                  int nthreads) {                                                \
    Problem<T> P{model, n, nc, M, N, consts, consts_per_voxel, data, nullptr, nullptr,           \
                 T(0), T(0), T(0), T(0), T(0)};                                                  \
    return ssr_run(P, params, out, nthreads);                                                    \
  }                                                                                              \
  extern "C" int oracle_lm_run_gen_##SUF(                                                        \
      void* accum, void* resid, int n, int nc, int M, long N, const T* consts,                   \
      int consts_per_voxel, const T* data, const T* lb, const T* ub, T* params, T* lam,          \
      int8_t* step_type, T* f, T* first_f, T* last_f, int32_t* n_iter, int iters, T tol, T mu,   \
      T eta, T acc, T dec, int nthreads) {                                                       \
    Problem<T> P{-1, n, nc, M, N, consts, consts_per_voxel, data, lb, ub, tol, mu, eta, acc,     \
                 dec, (GenAccum<T>)accum, (GenResid<T>)resid};                                   \
    if (!accum || !resid) return -2;                                                             \
    return lm_run(P, params, lam, step_type, f, first_f, last_f, n_iter, iters, (T*)nullptr,     \
                  (T*)nullptr, (int8_t*)nullptr, nthreads);                                      \
  }                                                                                              \
  extern "C" int oracle_ssr_gen_##SUF(void* resid, int n, int nc, int M, long N,                 \
                                      const T* consts, int consts_per_voxel, const T* data,      \
                                      const T* params, T* out, int nthreads) {                   \
    Problem<T> P{-1, n, nc, M, N, consts, consts_per_voxel, data, nullptr, nullptr,              \
                 T(0), T(0), T(0), T(0), T(0), nullptr, (GenResid<T>)resid};                     \
    if (!resid) return -2;                                                                       \
    return ssr_run(P, params, out, nthreads);                                                    \
  }                                                                                              \
  extern "C" int oracle_model_eval_##SUF(int model, int M, long N, const T* consts, int cpv,      \
                                         const T* params, T* m, T* J, T* d2) {                   \
    return model_eval(model, M, N, consts, cpv, params, m, J, d2);                               \
  }                                                                                              \
  extern "C" void oracle_gmw81_##SUF(int n, T* mat) { gmw81(n, mat); }                            \
  extern "C" void oracle_diag_pivot_##SUF(int n, T* mat, int* perm) { diag_pivot(n, mat, perm); } \
  extern "C" void oracle_gmw81_solve_##SUF(int n, T* mat, const T* rhs, T* sol) {                 \
    gmw81_solver(n, mat, rhs, sol);                                                              \
  }

ORACLE_EXPORTS(float, f)
ORACLE_EXPORTS(double, d)
