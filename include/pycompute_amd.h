/*
 * pycompute_amd: C-ABI of the MI355X batched per-voxel Levenberg-Marquardt
 * fitter (libpycompute_amd.so, built with hipcc for gfx950).
 *
 * The reference (turbotage/pycompute) binds its GPU path from Python through
 * CuPy: the generated CUDA source goes to cp.RawModule (NVRTC) and the kernel
 * is launched with (grid, block, args) from a Python loop. These entry points
 * are what that Python layer binds here instead, through ctypes: compile the
 * generated HIP source with hiprtc, load the code object, launch the fused
 * kernels on the caller's HIP stream.
 *
 * Conventions
 *   - Every function returns 0 on success or a negative pca_status; the
 *     message of the last failure on the calling thread is pca_last_error().
 *   - Pointers are device pointers (hipMalloc / torch tensors) unless stated;
 *     the library never frees caller memory.
 *   - Arrays are structure-of-arrays (rows, N), voxel index fastest, with the
 *     element type the module was generated for (float or double).
 *   - Launches are asynchronous on `stream` (a hipStream_t, NULL = default).
 */
#ifndef PYCOMPUTE_AMD_H
#define PYCOMPUTE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum pca_status {
  PCA_OK = 0,
  PCA_ERR_ARG = -1,      /* invalid argument */
  PCA_ERR_COMPILE = -2,  /* hiprtc compile failed; log in pca_last_error() */
  PCA_ERR_HIP = -3,      /* HIP runtime error */
  PCA_ERR_LIMIT = -4     /* size beyond a documented maximum */
};

/* Maximum number of voxel-uniform consts (nconst * M) passed by value. */
#define PCA_UC_MAX 256

/* Arguments of one fused LM fit launch (all iterations).
 * Replaces the reference kernel ABI
 *   k_second_order_levenberg_marquardt_<n>_<c>_<M>_<t>_<hash>(consts, data,
 *     lower_bound, upper_bound, tol, mu, eta, acc, dec, params, params_tp,
 *     step, lam, step_type, f, ftp, g, h, hl, Nprobs)
 * (pycompute/cuda/lsqnonlin.py:987-1011; emitted
 *  test/ivim/second_order_levenberg_marquardt_4_1_21_f_f7def86f0a03.cu:519-523),
 * which the host launched once per iteration (lsqnonlin.py:1058-1073). The
 * scalars are double here and converted to the kernel's element type inside
 * the kernel (the reference passed Python floats to float parameters). */
typedef struct pca_lm_args {
  const void* data;      /* (M, N) measurements */
  const void* consts;    /* (nconst*M, N) per-voxel consts; unused by uniform-consts modules */
  const void* lb;        /* (n, N) lower bounds */
  const void* ub;        /* (n, N) upper bounds */
  void* params;          /* (n, N) initial guess in, fitted parameters out */
  void* lam;             /* (N) damping lambda, in/out (reference setup: 5) */
  int8_t* step_type;     /* (N) in/out: 0 converged, 1/2 accepted, 4 rejected, |8 non-descent */
  void* f;               /* (N) in/out: SSR at the start of the last executed iteration */
  void* first_f;         /* (N) out or NULL: f after the first iteration */
  void* last_f;          /* (N) out or NULL: f after the last iteration */
  int32_t* n_iter;       /* (N) incremented by executed iterations, or NULL */
  void* g_out;           /* (n, N) or NULL: gradient J^T r of the last iteration */
  void* h_out;           /* (n(n+1)/2, N) or NULL: Hessian, packed lower triangle */
  void* hl_out;          /* (n(n+1)/2, N) or NULL: damped Hessian */
  void* step_out;        /* (n, N) or NULL: last solved step */
  int64_t N;             /* voxels */
  int32_t iters;         /* iterations (the reference's launches) */
  int32_t reserved;      /* 0 */
  double tol, mu, eta, acc, dec;
} pca_lm_args;

/* Arguments of the standalone SSR kernel: f[v] += sum_i r_i^2, skipping
 * voxels whose step_type is 0. Replaces F.run(pars, consts, data, step_type, f)
 * (pycompute/cuda/lsqnonlin.py:137-145, kernel :157-177). */
typedef struct pca_ssr_args {
  const void* params;       /* (n, N) */
  const void* data;         /* (M, N) */
  const void* consts;       /* (nconst*M, N) per-voxel consts or unused */
  const int8_t* step_type;  /* (N) or NULL */
  void* f;                  /* (N) accumulated */
  int64_t N;
} pca_ssr_args;

/* Arguments of the model evaluation kernel (model, Jacobian, second
 * derivatives per measurement). Replaces the per-measurement evaluation of
 * fghhl (pycompute/cuda/lsqnonlin.py:206-346) as a standalone surface. */
typedef struct pca_eval_args {
  const void* params;  /* (n, N) */
  const void* consts;  /* (nconst*M, N) per-voxel consts or unused */
  void* m;             /* (M, N) or NULL */
  void* J;             /* (n*M, N) or NULL, row j*M+i */
  void* d2;            /* (n(n+1)/2*M, N) or NULL, row l*M+i */
  int64_t N;
} pca_eval_args;

/* Packed queue of the chunked (compacting) fit: the voxels still iterating,
 * entry e of `cap` entries, arrays (rows, cap) with the module's element
 * type. Two layouts, chosen by y: carried (y, lo, hi and, for per-voxel
 * consts, cv travel with the entry: coalesced reads) or compact (y, cv, lo
 * and hi all NULL: measurements, bounds and per-voxel consts are read from
 * the caller's arrays through idx; about 3.6x less queue memory for 4-
 * parameter, 16-measurement models, at the price of scattered reads in later
 * chunks, since compaction scrambles the queue order). */
typedef struct pca_ws {
  void* y;          /* (M, cap) measurements, or NULL */
  void* cv;         /* (nconst*M, cap) per-voxel consts, or NULL */
  void* lo;         /* (n, cap) or NULL */
  void* hi;         /* (n, cap) or NULL */
  void* p;          /* (n, cap) parameters */
  void* lam;        /* (cap) */
  void* f;          /* (cap) */
  void* f_first;    /* (cap) */
  int32_t* idx;     /* (cap) original voxel index */
  int32_t* cnt;     /* (cap) iterations executed in this run */
  int8_t* st;       /* (cap) */
  int64_t cap;
} pca_ws;

/* One model's part of a chunked launch: its caller arrays, its voxel-uniform
 * consts table (device pointer, laid out as pca_lm_launch's uconsts; NULL for
 * per-voxel consts) and two queues. */
typedef struct pca_model_job {
  pca_lm_args a;
  const void* uc;
  pca_ws ws[2];
} pca_model_job;

typedef struct pca_module_s* pca_module;
typedef struct pca_kernel_s* pca_kernel;
typedef struct pca_graph_s* pca_graph;

/* Message of the last failure on this thread ("" if none). */
const char* pca_last_error(void);

/* Library ABI version (bumped when a struct above changes). */
int32_t pca_abi_version(void);

/* Compile HIP source with hiprtc into a code object. Replaces
 * cp.RawModule(code=...) (pycompute/cuda/lsqnonlin.py:1013-1025,
 * solver.py:584). `options` are hiprtc options (e.g. --offload-arch=gfx950).
 * On success *code is a malloc'd buffer of *size bytes; release it with
 * pca_free(). On a compile error the log is in pca_last_error(). Needs no
 * GPU. */
int pca_compile(const char* source, const char* program_name, const char* const* options,
                int32_t n_options, void** code, size_t* size);

/* Free a buffer returned by pca_compile. */
void pca_free(void* p);

/* Load a code object on the current device. Replaces RawModule loading. */
int pca_module_load(const void* code, size_t size, pca_module* out);
int pca_module_unload(pca_module module);

/* Look up a kernel in a module. Replaces mod.get_function(name)
 * (pycompute/cuda/lsqnonlin.py:1020). */
int pca_module_kernel(pca_module module, const char* name, pca_kernel* out);

/* Launch the fused LM fit (kernel "pca_lm_fit" of a generated module) over
 * ceil(N/block) blocks; `block` must equal the block size the module was
 * generated for. `uconsts` holds the voxel-uniform consts table (uconsts_bytes <=
 * PCA_UC_MAX * element size; NULL for per-voxel modules): nconst*M values,
 * const k of measurement i at k*M+i, then, for modules of the factored fast
 * build, the rows of derived consts that the generated source lists in its
 * head comment (const-only subexpressions evaluated once per protocol). Replaces the
 * per-iteration launch loop of SecondOrderLevenbergMarquardt.run
 * (pycompute/cuda/lsqnonlin.py:1058-1073). */
int pca_lm_launch(pca_kernel kernel, const pca_lm_args* args, const void* uconsts,
                  size_t uconsts_bytes, uint32_t block, void* stream);

/* Chunked fit with convergence compaction over `nmodels` models (kernels
 * "pca_lm_chunk" and "pca_reset_counts" of one generated module): resets the
 * queue counters (device array of (n_chunks+1)*nmodels int32) with a kernel,
 * or, when reset_kernel is NULL, relies on the caller having zeroed rows 1..
 * n_chunks-1 on the same stream (a single chunk needs none: chunk 0 takes
 * each model's N from its job), then launches one chunk per entry of
 * chunk_iters (host array; their sum is
 * the iteration count) on a grid of `grid` blocks. Per-voxel results equal
 * those of pca_lm_launch with the same iteration count. Replaces the same
 * per-iteration launch loop (pycompute/cuda/lsqnonlin.py:1058-1073) and, for
 * several models, that many separate fits. Capturable in a graph.
 * fresh_state != 0: the first chunk starts every voxel from the state
 * SecondOrderLevenbergMarquardt.setup establishes (lambda 5, step_type 1,
 * f 0, n_iter 0; lsqnonlin.py:1036-1056) instead of reading lam, step_type
 * and f, so a captured whole fit needs no separate reset kernel; 0 resumes
 * from the arrays. */
int pca_lm_chunked_launch(pca_kernel chunk_kernel, pca_kernel reset_kernel,
                          const pca_model_job* jobs, int32_t nmodels, int32_t* counts,
                          const int32_t* chunk_iters, int32_t n_chunks, uint32_t grid,
                          uint32_t block, int32_t fresh_state, void* stream);

/* Bytes of the file-scope device array behind pca_lm_chunked_launch_static,
 * per device (96 MiB). */
#define PCA_STATIC_SCRATCH_BYTES ((size_t)96 << 20)
int64_t pca_static_scratch_bytes(void);

/* The chunked, compacting fit of pca_lm_chunked_launch with no allocation:
 * the queue counters and both compact work queues of every model are carved from a
 * file-scope device array of PCA_STATIC_SCRATCH_BYTES (per device, held until
 * the process ends) instead of caller memory, and the counters are cleared by
 * a kernel at the start of every call. No counterpart in the reference; it is
 * for callers that cannot allocate (code under graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). The caller passes its inputs and
 * outputs: one pca_lm_args per model (as in pca_model_job.a), the device
 * pointer of each model's voxel-uniform consts table (as pca_lm_launch's
 * uconsts, derived rows included; or NULL), each model's
 * parameter count, and is_double for the element type. Two queues of N
 * entries per model must fit: 2 * N * ((n + 3) * es + 9) bytes plus 256-byte
 * alignment per array (4-parameter float32 IVIM: about 1.36M voxels per
 * call); larger sizes return PCA_ERR_LIMIT before anything is launched.
 * Calls overlapping on one device share the array and race. Results equal
 * pca_lm_chunked_launch's with the same chunk schedule. */
int pca_lm_chunked_launch_static(pca_kernel chunk_kernel, const pca_lm_args* args,
                                 const void* const* uconsts, const int32_t* nparam,
                                 int32_t nmodels, int32_t is_double, const int32_t* chunk_iters,
                                 int32_t n_chunks, uint32_t grid, uint32_t block,
                                 int32_t fresh_state, void* stream);

/* hipGraph capture of everything launched on `stream` between begin and
 * end (no counterpart in the reference, which launches from a Python loop). */
int pca_graph_begin(void* stream);
int pca_graph_end(void* stream, pca_graph* out);
int pca_graph_launch(pca_graph graph, void* stream);
int pca_graph_destroy(pca_graph graph);

/* Launch the standalone SSR kernel ("pca_ssr"). Replaces F.run
 * (pycompute/cuda/lsqnonlin.py:137-145). */
int pca_ssr_launch(pca_kernel kernel, const pca_ssr_args* args, const void* uconsts,
                   size_t uconsts_bytes, uint32_t block, void* stream);

/* Launch the model evaluation kernel ("pca_eval"). */
int pca_eval_launch(pca_kernel kernel, const pca_eval_args* args, const void* uconsts,
                    size_t uconsts_bytes, uint32_t block, void* stream);

/* Generic launch of a generated kernel: `params` is the hipModuleLaunchKernel
 * kernelParams array (one pointer per kernel argument). Replaces
 * RawKernel.__call__(grid, block, args) for the generated helper kernels. */
int pca_launch(pca_kernel kernel, uint32_t gx, uint32_t gy, uint32_t gz, uint32_t bx, uint32_t by,
               uint32_t bz, uint32_t shared_bytes, void* stream, void** params);

/* Batched polynomial least squares: out (deg+1, N) fits y (m, N) on
 * [1, x, .., x^deg]; x is (m, N) when x_per_voxel, else (m,). deg <= 7.
 * Replaces bpolyfit (pycompute/cuda/lsq.py:22-30). is_double selects the
 * element type of x, y and out. */
int pca_bpolyfit(const void* x, int32_t x_per_voxel, const void* y, int32_t m, int32_t deg,
                 void* out, int64_t N, int32_t is_double, void* stream);

/* Canonical IVIM labelling of fitted parameters (4, N) = (S0, f, D_1, D_2),
 * in place: where D_1 < D_2, swap D_1 and D_2 and set f = 1 - f. Replaces the
 * post-fit cell of test/ivim/test_ivim.ipynb:441-446. is_double selects the
 * element type. */
int pca_ivim_canonicalize(void* params, int64_t N, int32_t is_double, void* stream);

/* Default initial guess: out (n, N) = clip(nan_to_num(T(P log y)), lb, ub)
 * with P a device (K, M) double matrix shared by all voxels and T chosen by
 * kind (0 identity, 1 mono-exponential, 2 stretched exponential, 3 segmented
 * IVIM). Kind 3 replaces the caller-side init of test/ivim/test_ivim.ipynb
 * cell 6 (lines 146-165: two bpolyfit calls, exp, f = 1 - S0_hi/S0_lo,
 * nan_to_num, clip). */
int pca_loglinear_init(const double* P, int32_t K, int32_t M, const void* y, int32_t kind,
                       int32_t n, const void* lb, const void* ub, void* out, int64_t N,
                       int32_t is_double, void* stream);

/* ---- learned initialiser (no counterpart in the reference; its initial
 * guesses are the segmented log-linear fit and random search,
 * test/ivim/test_ivim.ipynb cells 6 and 11-12). fp32 storage and
 * accumulation; training GEMMs take bf16 MFMA operands
 * (v_mfma_f32_16x16x16_bf16), inference fp16 (v_mfma_f32_16x16x16_f16). */

/* C (M x N) = op(A) . op(B): op(A)[m][k] = transA ? A[k*lda+m] : A[m*lda+k],
 * op(B)[k][n] = transB ? B[n*ldb+k] : B[k*ldb+n]. Epilogue epi: 0 store
 * (with splits > 1: K split over blocks, added atomically into a zeroed C),
 * 1 C = gelu(acc + bias[n]) and Z = acc + bias[n], 2 C = acc + bias[n],
 * 3 C = acc * gelu'(Zin). */
int pca_mlp_gemm(const float* A, int64_t lda, int32_t transA, const float* B, int64_t ldb,
                 int32_t transB, float* C, int64_t ldc, int32_t M, int32_t N, int32_t K,
                 int32_t epi, const float* bias, float* Z, const float* Zin, int32_t splits,
                 void* stream);

/* dY = 2 (Y - T) * scale; *loss += sum (Y - T)^2 * scale (loss may be NULL). */
int pca_mlp_mse_grad(const float* Y, const float* T, float* dY, float* loss, int64_t n,
                     float scale, void* stream);

/* out[n] += sum_m X[m*ld + n] (bias gradients). */
int pca_mlp_colsum(const float* X, int64_t ld, int32_t M, int32_t N, float* out, void* stream);

/* The same without atomics, so results repeat bit for bit: C = op(A) . op(B)
 * (plain store), the K splits' partial products written to workspace (at
 * least splits * M * N floats) and summed in split order; out[n] += the column
 * sums of 64-row chunks (at most 1024 chunks; workspace at least chunks * N
 * floats), summed in chunk order. */
int pca_mlp_gemm_ordered(const float* A, int64_t lda, int32_t transA, const float* B, int64_t ldb,
                         int32_t transB, float* C, int64_t ldc, int32_t M, int32_t N, int32_t K,
                         int32_t splits, float* workspace, int64_t workspace_floats,
                         void* stream);
int pca_mlp_colsum_ordered(const float* X, int64_t ld, int32_t M, int32_t N, float* out,
                           float* workspace, int64_t workspace_floats, void* stream);

/* One Adam step (bias-corrected, step >= 1) on a flat fp32 parameter vector. */
int pca_mlp_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, float b1,
                 float b2, float eps, int32_t step, float weight_decay, void* stream);

/* One training step's forward, MSE gradient and backward in one kernel:
 * workgroup k takes batch rows [64k, 64k + 64) and writes its partial
 * gradient of every parameter block to part[k * np + ...] (entries outside
 * the six blocks are left as they are) and its partial loss to loss_part[k].
 * X (in_padded, ldx) feature-major, its column 0 the batch's first row; T
 * (B, out) the batch's targets; offsets = the offsets of w1 (hidden,
 * in_padded), b1, w2 (hidden, hidden), b2, w3 (out, hidden), b3 in the flat
 * parameter vector w of np entries. B a positive multiple of 64; hidden 32 or
 * 64 with in_padded 16 and out <= 16, or hidden 128 with in_padded 16, 32 or
 * 64 and out <= 32. The same arithmetic as the GEMM path (bf16 operands, fp32
 * accumulation); loss scale and dY = 2 (Y - T) / (B out). */
int pca_mlp_train_step(const float* X, int64_t ldx, const float* T, int32_t B, const float* w,
                       int32_t hidden, int32_t in_padded, int32_t out, const int32_t* offsets,
                       int32_t np, float* part, float* loss_part, void* stream);

/* g[j] = sum over k < nwg of part[k * np + j], in a fixed order;
 * *loss += sum of loss_part (loss may be NULL). With w set, also one Adam
 * step of w, m, v (as pca_mlp_adam) on g; with w NULL, gradients only. */
int pca_mlp_grad_reduce(const float* part, const float* loss_part, int32_t nwg, int32_t np,
                        float* g, float* loss, float* w, float* m, float* v, float lr, float b1,
                        float b2, float eps, int32_t step, float weight_decay, void* stream);

/* p[0:n] = value; fp32 -> bf16, fp32 -> fp16 (round to nearest even). */
int pca_mlp_fill(float* p, int64_t n, float value, void* stream);
int pca_mlp_to_bf16(const float* x, int16_t* y, int64_t n, void* stream);
int pca_mlp_to_f16(const float* x, int16_t* y, int64_t n, void* stream);

/* Training inputs: X (in_padded, N) = y / y[norm_index] (zero rows past in),
 * T (N, out) = (transform(truth) - mu) / sd, the transform being: divide
 * parameter scale_param by s = y[norm_index], subtract log s from
 * logshift_param, take the log of the parameters in log_mask. */
int pca_mlp_prepare(const float* y, const float* truth, int64_t N, int32_t in, int32_t in_padded,
                    int32_t out, int32_t norm_index, int32_t scale_param, int32_t logshift_param,
                    int32_t log_mask, const float* mu, const float* sd, float* X, float* T,
                    void* stream);

/* Fused inference of the in -> hidden -> hidden -> out GELU MLP: reads the
 * signals (in, N), writes clip(inverse transform(mu + sd * y), lb, ub) into
 * params (out, N) for voxels whose outputs are finite (others keep what
 * params held: the fallback guess). w1 is (hidden, in_padded) with
 * in_padded = 16, 32 or 64 and zero columns past `in`. Weights are fp16
 * (pca_mlp_to_f16 of the fp32 masters); the activation is the tanh-form GELU
 * evaluated as a packed-fp16 polynomial (absolute error <= 3e-4 in exact
 * arithmetic, on top of fp16 rounding). */
typedef struct pca_mlp_infer_args {
  const void* data;      /* (in, N) fp32 */
  const int16_t* w1;     /* (hidden, in_padded) fp16 */
  const float* b1;       /* (hidden) */
  const int16_t* w2;     /* (hidden, hidden) fp16 */
  const float* b2;       /* (hidden) */
  const int16_t* w3;     /* (out, hidden) fp16 */
  const float* b3;       /* (out) */
  const float* mu;       /* (out) */
  const float* sd;       /* (out) */
  const void* lb;        /* (out, N) fp32 */
  const void* ub;        /* (out, N) fp32 */
  void* params;          /* (out, N) fp32, in: fallback guess, out: MLP guess */
  int64_t N;
  int32_t in, out, hidden;
  int32_t norm_index, scale_param, logshift_param, log_mask;
  int32_t reserved;
  /* optional fallback for voxels whose outputs are not finite: the
   * log-linear guess of pca_loglinear_init (P: K x in, double, on the device;
   * kind as there, with n = out), nan_to_num'd and clipped. NULL: such voxels
   * keep what params held. */
  const double* fallback_P;
  int32_t fallback_K, fallback_kind;
} pca_mlp_infer_args;
int pca_mlp_infer(const pca_mlp_infer_args* args, void* stream);

/* gcnArchName of `device` (e.g. "gfx950:sramecc+:xnack-"). */
int pca_device_arch(int32_t device, char* buf, int32_t len);

#ifdef __cplusplus
}
#endif

#endif /* PYCOMPUTE_AMD_H */
