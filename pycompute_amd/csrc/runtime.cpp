// Host side of libpycompute_amd.so: hiprtc compilation, code-object loading
// and kernel launches behind the C-ABI declared in include/pycompute_amd.h.
// Built with hipcc; no torch types cross this boundary.

#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "pycompute_amd.h"
#include "internal.h"

struct pca_module_s {
  hipModule_t mod;
};
struct pca_kernel_s {
  hipFunction_t fn;
};
struct pca_graph_s {
  hipGraphExec_t exec;
};

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(PCA_ERR_HIP, std::string(what) + ": " + hipGetErrorName(e) + " (" +
                               hipGetErrorString(e) + ")");
}

#define PCA_HIP(call)                                  \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return hip_fail(e_, #call);  \
  } while (0)

// The kernels take the voxel-uniform consts block by value; copy the caller's
// bytes into a zeroed block of the largest size either element type needs.
struct UConsts {
  alignas(16) unsigned char bytes[PCA_UC_MAX * sizeof(double)];
};

int fill_uconsts(UConsts& uc, const void* src, size_t n) {
  std::memset(uc.bytes, 0, sizeof(uc.bytes));
  if (n > sizeof(uc.bytes)) return fail(PCA_ERR_LIMIT, "uniform consts exceed PCA_UC_MAX");
  if (n && !src) return fail(PCA_ERR_ARG, "uniform consts pointer is NULL");
  if (n) std::memcpy(uc.bytes, src, n);
  return PCA_OK;
}

uint32_t grid_for(int64_t N, uint32_t block) {
  return (uint32_t)((N + block - 1) / block);
}

int launch2(pca_kernel k, int64_t N, uint32_t block, void* stream, void* a, UConsts* uc) {
  if (!k || block == 0 || block > 1024) return fail(PCA_ERR_ARG, "bad kernel or block size");
  if (N <= 0) return PCA_OK;
  const int64_t grid = (N + block - 1) / block;
  if (grid > 0x7fffffff) return fail(PCA_ERR_LIMIT, "too many voxels for one launch");
  void* params[2] = {a, uc};
  PCA_HIP(hipModuleLaunchKernel(k->fn, grid_for(N, block), 1, 1, block, 1, 1, 0,
                                (hipStream_t)stream, params, nullptr));
  return PCA_OK;
}

}  // namespace

int pca_fail(int code, const char* msg) { return fail(code, msg ? msg : ""); }

extern "C" {

const char* pca_last_error(void) { return g_err.c_str(); }

int32_t pca_abi_version(void) { return 4; }

int pca_compile(const char* source, const char* program_name, const char* const* options,
                int32_t n_options, void** code, size_t* size) {
  if (!source || !code || !size) return fail(PCA_ERR_ARG, "pca_compile: NULL argument");
  *code = nullptr;
  *size = 0;
  hiprtcProgram prog;
  hiprtcResult r = hiprtcCreateProgram(&prog, source, program_name ? program_name : "pca.hip", 0,
                                       nullptr, nullptr);
  if (r != HIPRTC_SUCCESS)
    return fail(PCA_ERR_COMPILE, std::string("hiprtcCreateProgram: ") + hiprtcGetErrorString(r));
  r = hiprtcCompileProgram(prog, n_options, const_cast<const char**>(options));
  size_t log_size = 0;
  hiprtcGetProgramLogSize(prog, &log_size);
  std::string log(log_size, '\0');
  if (log_size) hiprtcGetProgramLog(prog, &log[0]);
  if (r != HIPRTC_SUCCESS) {
    hiprtcDestroyProgram(&prog);
    return fail(PCA_ERR_COMPILE, std::string("hiprtc compile failed: ") +
                                     hiprtcGetErrorString(r) + "\n" + log);
  }
  size_t n = 0;
  r = hiprtcGetCodeSize(prog, &n);
  if (r != HIPRTC_SUCCESS || n == 0) {
    hiprtcDestroyProgram(&prog);
    return fail(PCA_ERR_COMPILE, "hiprtcGetCodeSize failed");
  }
  void* buf = std::malloc(n);
  if (!buf) {
    hiprtcDestroyProgram(&prog);
    return fail(PCA_ERR_LIMIT, "out of host memory");
  }
  r = hiprtcGetCode(prog, (char*)buf);
  hiprtcDestroyProgram(&prog);
  if (r != HIPRTC_SUCCESS) {
    std::free(buf);
    return fail(PCA_ERR_COMPILE, "hiprtcGetCode failed");
  }
  *code = buf;
  *size = n;
  g_err = log;  // warnings, if any
  return PCA_OK;
}

void pca_free(void* p) { std::free(p); }

int pca_module_load(const void* code, size_t size, pca_module* out) {
  if (!code || !size || !out) return fail(PCA_ERR_ARG, "pca_module_load: bad argument");
  hipModule_t m;
  PCA_HIP(hipModuleLoadData(&m, code));
  *out = new pca_module_s{m};
  return PCA_OK;
}

int pca_module_unload(pca_module module) {
  if (!module) return PCA_OK;
  hipError_t e = hipModuleUnload(module->mod);
  delete module;
  if (e != hipSuccess) return hip_fail(e, "hipModuleUnload");
  return PCA_OK;
}

int pca_module_kernel(pca_module module, const char* name, pca_kernel* out) {
  if (!module || !name || !out) return fail(PCA_ERR_ARG, "pca_module_kernel: bad argument");
  hipFunction_t fn;
  PCA_HIP(hipModuleGetFunction(&fn, module->mod, name));
  // Kernels are looked up once per module and kept for the module's life.
  *out = new pca_kernel_s{fn};
  return PCA_OK;
}

int pca_lm_launch(pca_kernel kernel, const pca_lm_args* args, const void* uconsts,
                  size_t uconsts_bytes, uint32_t block, void* stream) {
  // per-voxel arrays: empty ones (N = 0) may have no address
  if (!args || (args->N > 0 && (!args->data || !args->lb || !args->ub || !args->params ||
                                !args->lam || !args->step_type || !args->f)))
    return fail(PCA_ERR_ARG, "pca_lm_launch: a required array is NULL");
  if (args->iters < 0) return fail(PCA_ERR_ARG, "pca_lm_launch: iters < 0");
  UConsts uc;
  int rc = fill_uconsts(uc, uconsts, uconsts_bytes);
  if (rc) return rc;
  pca_lm_args a = *args;
  return launch2(kernel, a.N, block, stream, &a, &uc);
}

int pca_ssr_launch(pca_kernel kernel, const pca_ssr_args* args, const void* uconsts,
                   size_t uconsts_bytes, uint32_t block, void* stream) {
  if (!args || (args->N > 0 && (!args->params || !args->data || !args->f)))
    return fail(PCA_ERR_ARG, "pca_ssr_launch: a required array is NULL");
  UConsts uc;
  int rc = fill_uconsts(uc, uconsts, uconsts_bytes);
  if (rc) return rc;
  pca_ssr_args a = *args;
  return launch2(kernel, a.N, block, stream, &a, &uc);
}

int pca_eval_launch(pca_kernel kernel, const pca_eval_args* args, const void* uconsts,
                    size_t uconsts_bytes, uint32_t block, void* stream) {
  if (!args || (args->N > 0 && !args->params))
    return fail(PCA_ERR_ARG, "pca_eval_launch: params is NULL");
  UConsts uc;
  int rc = fill_uconsts(uc, uconsts, uconsts_bytes);
  if (rc) return rc;
  pca_eval_args a = *args;
  return launch2(kernel, a.N, block, stream, &a, &uc);
}

int pca_launch(pca_kernel kernel, uint32_t gx, uint32_t gy, uint32_t gz, uint32_t bx, uint32_t by,
               uint32_t bz, uint32_t shared_bytes, void* stream, void** params) {
  if (!kernel) return fail(PCA_ERR_ARG, "pca_launch: NULL kernel");
  if (!gx || !gy || !gz || !bx || !by || !bz || (uint64_t)bx * by * bz > 1024)
    return fail(PCA_ERR_ARG, "pca_launch: bad launch shape");
  PCA_HIP(hipModuleLaunchKernel(kernel->fn, gx, gy, gz, bx, by, bz, shared_bytes,
                                (hipStream_t)stream, params, nullptr));
  return PCA_OK;
}

int pca_lm_chunked_launch(pca_kernel chunk_kernel, pca_kernel reset_kernel,
                          const pca_model_job* jobs, int32_t nmodels, int32_t* counts,
                          const int32_t* chunk_iters, int32_t n_chunks, uint32_t grid,
                          uint32_t block, int32_t fresh_state, void* stream) {
  if (!chunk_kernel || !jobs || !counts || !chunk_iters)
    return fail(PCA_ERR_ARG, "pca_lm_chunked_launch: NULL argument");
  if (nmodels < 1 || nmodels > 4) return fail(PCA_ERR_ARG, "pca_lm_chunked_launch: 1..4 models");
  if (n_chunks < 1 || grid == 0 || block == 0 || block > 1024 || block % 64)
    return fail(PCA_ERR_ARG, "pca_lm_chunked_launch: bad chunk count or launch shape");
  for (int m = 0; m < nmodels; ++m) {
    const pca_model_job& j = jobs[m];
    if (j.a.N < 0 || j.a.N > 0x7fffffff) return fail(PCA_ERR_LIMIT, "voxels per model > 2^31-1");
    for (int q = 0; q < 2; ++q)
    {
      if (j.a.N > 0 && (j.ws[q].cap < j.a.N || !j.ws[q].p || !j.ws[q].lam || !j.ws[q].f ||
                        !j.ws[q].f_first || !j.ws[q].idx || !j.ws[q].cnt || !j.ws[q].st))
        return fail(PCA_ERR_ARG, "pca_lm_chunked_launch: working set smaller than N or NULL");
      // a carried queue has y, lo and hi (cv for per-voxel consts); a compact
      // one has none of them; both queues of a model use the same layout
      const bool carried = j.ws[q].y != nullptr;
      if (j.a.N > 0 && ((j.ws[q].lo != nullptr) != carried || (j.ws[q].hi != nullptr) != carried ||
                        (j.ws[0].y != nullptr) != (j.ws[1].y != nullptr) ||
                        (!carried && j.ws[q].cv != nullptr) ||
                        (carried && j.a.consts && !j.uc && !j.ws[q].cv)))
        return fail(PCA_ERR_ARG, "pca_lm_chunked_launch: queue arrays y, lo, hi (and cv for "
                                 "per-voxel consts) must be all set or all NULL");
    }
  }
  for (int c = 0; c < n_chunks; ++c)
    if (chunk_iters[c] < 1) return fail(PCA_ERR_ARG, "pca_lm_chunked_launch: chunk of < 1 iteration");

  // reset the queue counters: row 0 holds every model's voxel count (no
  // reset kernel: the caller has done it on this stream)
  struct {
    int32_t first[8];
  } init;
  std::memset(&init, 0, sizeof(init));
  for (int m = 0; m < nmodels; ++m) init.first[m] = (int32_t)jobs[m].a.N;
  int32_t total = (n_chunks + 1) * nmodels;
  void* rp[3] = {&counts, &total, &init};
  if (reset_kernel)
    PCA_HIP(hipModuleLaunchKernel(reset_kernel->fn, (uint32_t)((total + 255) / 256), 1, 1, 256,
                                  1, 1, 0, (hipStream_t)stream, rp, nullptr));

  // kernel argument block: jobs[nmodels], counts, chunk, it0, k, last, fresh
  // (zero padding to the struct's 8-byte alignment)
  const size_t jb = sizeof(pca_model_job) * (size_t)nmodels;
  std::vector<unsigned char> buf(jb + sizeof(int32_t*) + 6 * sizeof(int32_t));
  std::memcpy(buf.data(), jobs, jb);
  std::memcpy(buf.data() + jb, &counts, sizeof(counts));
  int32_t* tail = (int32_t*)(buf.data() + jb + sizeof(int32_t*));
  int32_t it0 = 0;
  for (int c = 0; c < n_chunks; ++c) {
    tail[0] = c;
    tail[1] = it0;
    tail[2] = chunk_iters[c];
    tail[3] = (c == n_chunks - 1) ? 1 : 0;
    tail[4] = (c == 0 && fresh_state) ? 1 : 0;
    void* kp[1] = {buf.data()};
    PCA_HIP(hipModuleLaunchKernel(chunk_kernel->fn, grid, 1, 1, block, 1, 1, 0,
                                  (hipStream_t)stream, kp, nullptr));
    it0 += chunk_iters[c];
  }
  return PCA_OK;
}

int pca_graph_begin(void* stream) {
  PCA_HIP(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
  return PCA_OK;
}

int pca_graph_end(void* stream, pca_graph* out) {
  if (!out) return fail(PCA_ERR_ARG, "pca_graph_end: NULL out");
  hipGraph_t g;
  PCA_HIP(hipStreamEndCapture((hipStream_t)stream, &g));
  hipGraphExec_t ge;
  hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  const hipError_t ed = hipGraphDestroy(g);
  if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
  if (ed != hipSuccess) {
    (void)hipGraphExecDestroy(ge);
    return hip_fail(ed, "hipGraphDestroy");
  }
  *out = new pca_graph_s{ge};
  return PCA_OK;
}

int pca_graph_launch(pca_graph graph, void* stream) {
  if (!graph) return fail(PCA_ERR_ARG, "pca_graph_launch: NULL graph");
  PCA_HIP(hipGraphLaunch(graph->exec, (hipStream_t)stream));
  return PCA_OK;
}

int pca_graph_destroy(pca_graph graph) {
  if (!graph) return PCA_OK;
  hipError_t e = hipGraphExecDestroy(graph->exec);
  delete graph;
  if (e != hipSuccess) return hip_fail(e, "hipGraphExecDestroy");
  return PCA_OK;
}

int pca_device_arch(int32_t device, char* buf, int32_t len) {
  if (!buf || len <= 0) return fail(PCA_ERR_ARG, "pca_device_arch: bad buffer");
  hipDeviceProp_t prop;
  PCA_HIP(hipGetDeviceProperties(&prop, device));
  std::snprintf(buf, (size_t)len, "%s", prop.gcnArchName);
  return PCA_OK;
}

}  // extern "C"
