// pca_lm_chunked_launch_static: the chunked, compacting LM fit of
// pca_lm_chunked_launch with its queue counters and both work queues of every
// model carved from one file-scope device array instead of caller memory, so
// a caller that cannot allocate (code under graph capture, tight serving
// loops) passes only its inputs and outputs. The array is
// PCA_STATIC_SCRATCH_BYTES per device, held until the process ends; every
// call on a device shares it and keeps the last call's contents, so calls
// that overlap on one device race. The queue counters, which must start at
// zero, are cleared by a kernel of this file at the start of every call.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "pycompute_amd.h"
#include "internal.h"

namespace {

constexpr size_t kCountsBytes = 64 << 10;          // queue counters, at offset 0
constexpr size_t kAlign = 256;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[PCA_STATIC_SCRATCH_BYTES];

// Device address of `scratch` per device (the array's name is a host-side
// stand-in; only these addresses may be handed to kernels).
void* g_scratch[kMaxDevices];
std::mutex g_mutex;

__global__ void static_reset_counts(int32_t* counts, int32_t n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    counts[i] = 0;
}

int scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return pca_fail(PCA_ERR_HIP, "pca_lm_chunked_launch_static: hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices)
    return pca_fail(PCA_ERR_LIMIT, "pca_lm_chunked_launch_static: device index >= 64");
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return pca_fail(PCA_ERR_HIP, "pca_lm_chunked_launch_static: hipGetSymbolAddress failed");
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return PCA_OK;
}

size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// Bytes of the two queues of one model of N voxels and n parameters.
size_t queue_bytes(int64_t N, int32_t n, size_t es) {
  const size_t cap = (size_t)(N > 0 ? N : 1);
  const size_t one = align_up(cap * es * n) + 3 * align_up(cap * es) + 2 * align_up(cap * 4) +
                     align_up(cap);
  return 2 * one;
}

// Carve one queue of capacity cap at base + off (device address arithmetic).
pca_ws carve(unsigned char* base, size_t& off, int64_t cap, int32_t n, size_t es) {
  pca_ws w;
  w.y = w.cv = w.lo = w.hi = nullptr;   // the compact layout
  auto take = [&](size_t bytes) {
    void* p = base + off;
    off += align_up(bytes);
    return p;
  };
  const size_t c = (size_t)cap;
  w.p = take(c * es * n);
  w.lam = take(c * es);
  w.f = take(c * es);
  w.f_first = take(c * es);
  w.idx = (int32_t*)take(c * 4);
  w.cnt = (int32_t*)take(c * 4);
  w.st = (int8_t*)take(c);
  w.cap = cap;
  return w;
}

}  // namespace

extern "C" {

int64_t pca_static_scratch_bytes(void) { return (int64_t)PCA_STATIC_SCRATCH_BYTES; }

int pca_lm_chunked_launch_static(pca_kernel chunk_kernel, const pca_lm_args* args,
                                 const void* const* uconsts, const int32_t* nparam,
                                 int32_t nmodels, int32_t is_double, const int32_t* chunk_iters,
                                 int32_t n_chunks, uint32_t grid, uint32_t block,
                                 int32_t fresh_state, void* stream) {
  if (!chunk_kernel || !args || !nparam || !chunk_iters)
    return pca_fail(PCA_ERR_ARG, "pca_lm_chunked_launch_static: NULL argument");
  if (nmodels < 1 || nmodels > 4)
    return pca_fail(PCA_ERR_ARG, "pca_lm_chunked_launch_static: 1..4 models");
  if (n_chunks < 1)
    return pca_fail(PCA_ERR_ARG, "pca_lm_chunked_launch_static: at least one chunk");
  const size_t es = is_double ? 8 : 4;
  const int64_t n_counts = (int64_t)(n_chunks + 1) * nmodels;
  if ((size_t)n_counts * 4 > kCountsBytes)
    return pca_fail(PCA_ERR_LIMIT, "pca_lm_chunked_launch_static: too many chunks for the "
                                   "counter area (at most 16383 chunk rows)");
  size_t need = kCountsBytes;
  for (int m = 0; m < nmodels; ++m) {
    if (args[m].N < 0 || args[m].N > 0x7fffffff)
      return pca_fail(PCA_ERR_LIMIT, "pca_lm_chunked_launch_static: voxels per model > 2^31-1");
    if (nparam[m] < 1 || nparam[m] > 64)
      return pca_fail(PCA_ERR_ARG, "pca_lm_chunked_launch_static: nparam out of range");
    need += queue_bytes(args[m].N, nparam[m], es);
  }
  if (need > (size_t)PCA_STATIC_SCRATCH_BYTES) {
    char msg[256];
    std::snprintf(msg, sizeof(msg),
                  "pca_lm_chunked_launch_static: the queues need %zu bytes, more than the %zu "
                  "bytes of static scratch; fit fewer voxels per call",
                  need, (size_t)PCA_STATIC_SCRATCH_BYTES);
    return pca_fail(PCA_ERR_LIMIT, msg);
  }
  void* sp = nullptr;
  int rc = scratch_address(&sp);
  if (rc) return rc;
  unsigned char* base = (unsigned char*)sp;
  pca_model_job jobs[4];
  std::memset(jobs, 0, sizeof(jobs));
  size_t off = kCountsBytes;
  for (int m = 0; m < nmodels; ++m) {
    jobs[m].a = args[m];
    jobs[m].uc = uconsts ? uconsts[m] : nullptr;
    const int64_t cap = args[m].N > 0 ? args[m].N : 1;
    jobs[m].ws[0] = carve(base, off, cap, nparam[m], es);
    jobs[m].ws[1] = carve(base, off, cap, nparam[m], es);
  }
  int32_t* counts = (int32_t*)base;
  // the counters start at zero on every call (the array keeps the last call's)
  hipLaunchKernelGGL(static_reset_counts, dim3(1), dim3(256), 0, (hipStream_t)stream, counts,
                     (int32_t)n_counts);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return pca_fail(PCA_ERR_HIP, "pca_lm_chunked_launch_static: counter reset launch failed");
  return pca_lm_chunked_launch(chunk_kernel, nullptr, jobs, nmodels, counts, chunk_iters,
                               n_chunks, grid, block, fresh_state, stream);
}

}  // extern "C"
