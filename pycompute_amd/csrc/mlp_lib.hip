// Learned initialiser: a small MLP (in -> H -> H -> out, GELU) mapping a
// voxel's normalised signal to standardised starting parameters. Its dense
// contractions run on the matrix cores with fp32 accumulation: training on
// v_mfma_f32_16x16x16_bf16 (bf16 operands), inference on
// v_mfma_f32_16x16x16_f16 (fp16 operands). Compiled into libpycompute_amd.so
// for gfx950.
//
// There is no counterpart in the reference (its initial guesses are the
// segmented log-linear fit and random search, test/ivim/test_ivim.ipynb
// cells 6, 11-12); LM refines from the MLP output and the log-linear guess is
// the fallback.
//
// Storage is fp32 (weights keep fp32 masters for Adam); training operands
// are rounded to bf16 (round to nearest even) as they are loaded into MFMA
// registers; inference reads fp16 copies of the weights (pca_mlp_to_f16).
//
// MFMA 16x16x16 (bf16 or f16) operand layout (wave64, lane l; 16x16x32 as
// noted at mfma32h):
//   A (16 x 16, M x K): row l % 16, k = 4 (l / 16) + 0..3
//   B (16 x 16, K x N): col l % 16, k = 4 (l / 16) + 0..3
//   C (16 x 16):        col l % 16, rows 4 (l / 16) + 0..3

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "internal.h"
#include "loglinear.h"
#include "pycompute_amd.h"

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// fp32 -> bf16, round to nearest even (v_cvt_pk_bf16_f32 on gfx950)
__device__ __forceinline__ short f2bf(float x) { return __builtin_bit_cast(short, (__bf16)x); }

__device__ __forceinline__ floatx4 mfma16(bf16x4 a, bf16x4 b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// GELU, tanh form: 0.5 x (1 + tanh u) = x sigmoid(2u), u = c (x + k x^3),
// c = sqrt(2/pi), k = 0.044715. With the constants folded into one exp2
// argument, x sigmoid(2u) = x / (1 + 2^(x (A + B x^2))), A = -2 c log2(e),
// B = A k: five VALU operations plus v_exp_f32 and v_rcp_f32 per activation.
constexpr float kGeluA = -2.0f * 0.7978845608028654f * 1.4426950408889634f;
constexpr float kGeluB = kGeluA * 0.044715f;
__device__ __forceinline__ float gelu_sigmoid(float x) {   // sigmoid(2u)
  const float t = x * __builtin_fmaf(kGeluB, x * x, kGeluA);
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(t));
}
__device__ __forceinline__ float gelu(float x) { return x * gelu_sigmoid(x); }
// d/dx x s(2u) = s + 2 x s (1 - s) u',  s = sigmoid(2u), u' = c (1 + 3 k x^2)
__device__ __forceinline__ float gelu_grad(float x) {
  const float s = gelu_sigmoid(x);
  const float du = 0.7978845608028654f * (1.0f + 3.0f * 0.044715f * x * x);
  return s + 2.0f * x * s * (1.0f - s) * du;
}

// Inference operands: fp16.
typedef _Float16 half4 __attribute__((ext_vector_type(4)));

typedef _Float16 half8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ floatx4 mfma16h(half4 a, half4 b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
}
// 16x16x32 (gfx950): A row l % 16, B column l % 16, k = 8 (l / 16) + 0..7
__device__ __forceinline__ floatx4 mfma32h(half8 a, half8 b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ half8 cat8(half4 a, half4 b) {
  return half8{a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3]};
}

// The same tanh-form GELU as a packed-fp16 polynomial for inference:
// gelu(x) = x s(x), s(x) = sigmoid(2u) = 1/2 + z P(z^2), z = clamp(x / 3.5,
// -1, 1), P of degree 5, fitted to s on |x| <= 3.5 by iteratively reweighted
// least squares on the error of x s(x) (a near-minimax fit). In exact
// arithmetic |x s_P(x) - gelu(x)| <= 2.8e-4 for |x| <= 3.5; beyond, s_P is
// held at its end value, within 2.6e-4 of 0 or 1. Two activations per
// instruction (v_pk_mul_f16 / v_pk_fma_f16), no transcendental instruction.
constexpr float kGeluP[6] = {1.390737289680657f, -2.74187719999597f, 4.43028147414536f,
                             -4.637580546145781f, 2.7523397819371955f, -0.694156068202283f};
constexpr float kGeluClamp = 3.5f;
// Four activations as one half4 expression: every step is two independent
// packed instructions, so neither waits on the other's result.
__device__ __forceinline__ half4 splat4(float c) {
  const _Float16 h = (_Float16)c;
  return half4{h, h, h, h};
}
__device__ __forceinline__ half4 gelu_pk4(floatx4 v) {
  const half4 x = {(_Float16)v[0], (_Float16)v[1], (_Float16)v[2], (_Float16)v[3]};
  // z = clamp(x / c, -1, 1) as 2w - 1 with w = clamp(x / 2c + 1/2, 0, 1): w is
  // one v_pk_fma_f16 with the output clamp, so z costs two instructions
  const half4 w = __builtin_elementwise_min(
      __builtin_elementwise_max(x * splat4(0.5f / kGeluClamp) + splat4(0.5f), splat4(0.0f)),
      splat4(1.0f));
  const half4 z = w * splat4(2.0f) - splat4(1.0f);
  const half4 t = z * z;
  half4 p = t * splat4(kGeluP[5]) + splat4(kGeluP[4]);
  p = p * t + splat4(kGeluP[3]);
  p = p * t + splat4(kGeluP[2]);
  p = p * t + splat4(kGeluP[1]);
  p = p * t + splat4(kGeluP[0]);
  return x * (z * p + splat4(0.5f));
}

// Copy an (R x K) fp16 matrix (row-major, rows past rows_valid read as zero)
// into LDS in MFMA fragment order: fragment f = (R/16 row tile, K/16 k tile)
// holds 64 lanes x 4 values, lane l = 16 (k within tile / 4) + row within
// tile, at dst[4 (64 f + l)].
template <int R, int K>
__device__ __forceinline__ void stage_fragments(short* dst, const short* src, int rows_valid) {
  constexpr int NL = (R / 16) * (K / 16) * 64;
  for (int i = threadIdx.x; i < NL; i += blockDim.x) {
    const int f = i >> 6, l = i & 63;
    const int row = (f / (K / 16)) * 16 + (l & 15), k = (f % (K / 16)) * 16 + (l >> 4) * 4;
    long long v = 0;
    if (row < rows_valid) v = *(const long long*)(src + (long long)row * K + k);
    *(long long*)(dst + 4 * i) = v;
  }
}

// The same for 16x16x32 fragments with the k permutation of the inference
// kernel: fragment f = (R/16 row tile, K/32 k block c), lane l = 16 g + row
// within tile holds W[row][32 c + 4 g + 0..3] then W[row][32 c + 16 + 4 g +
// 0..3], at dst[8 (64 f + l)].
template <int R, int K>
__device__ __forceinline__ void stage_fragments_k32(short* dst, const short* src, int rows_valid) {
  constexpr int NL = (R / 16) * (K / 32) * 64;
  for (int i = threadIdx.x; i < NL; i += blockDim.x) {
    const int f = i >> 6, l = i & 63;
    const int row = (f / (K / 32)) * 16 + (l & 15), k = (f % (K / 32)) * 32 + (l >> 4) * 4;
    long long v0 = 0, v1 = 0;
    if (row < rows_valid) {
      v0 = *(const long long*)(src + (long long)row * K + k);
      v1 = *(const long long*)(src + (long long)row * K + k + 16);
    }
    *(long long*)(dst + 8 * i) = v0;
    *(long long*)(dst + 8 * i + 4) = v1;
  }
}

enum Epi { EPI_STORE = 0, EPI_BIAS_GELU = 1, EPI_BIAS = 2, EPI_GELU_GRAD = 3 };

struct GemmArgs {
  const float* A;
  const float* B;
  float* C;
  const float* bias;  // EPI_BIAS*: (N)
  float* Z;           // EPI_BIAS_GELU: pre-activation out (M, ldc) or null
  const float* Zin;   // EPI_GELU_GRAD: pre-activation in (M, ldc)
  long long lda, ldb, ldc;
  int M, N, K;
  int transA, transB;
  int epi;
  int k_per_split;
};

// C (M x N) = op(A) (M x K) . op(B) (K x N), one 32 x 32 tile per wavefront
// (2 x 2 MFMA tiles), 4 wavefronts per 64 x 64 block tile. gridDim.z > 1
// splits K and adds partial tiles atomically (EPI_STORE only; C pre-zeroed).
__global__ void __launch_bounds__(256) k_gemm(GemmArgs g) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int m0 = blockIdx.x * 64 + (wave & 1) * 32;
  const int n0 = blockIdx.y * 64 + (wave >> 1) * 32;
  const int kb = blockIdx.z * g.k_per_split;
  const int ke = min(g.K, kb + g.k_per_split);
  const int r = lane & 15;
  const int kq = (lane >> 4) * 4;
  floatx4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = kb; k0 < ke; k0 += 16) {
    bf16x4 a[2], b[2];
    for (int t = 0; t < 2; ++t) {
      const int m = m0 + t * 16 + r;
      const int n = n0 + t * 16 + r;
      for (int q = 0; q < 4; ++q) {
        const int k = k0 + kq + q;
        float av = 0.f, bv = 0.f;
        if (m < g.M && k < ke) av = g.transA ? g.A[(long long)k * g.lda + m] : g.A[(long long)m * g.lda + k];
        if (n < g.N && k < ke) bv = g.transB ? g.B[(long long)n * g.ldb + k] : g.B[(long long)k * g.ldb + n];
        a[t][q] = f2bf(av);
        b[t][q] = f2bf(bv);
      }
    }
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) acc[i][j] = mfma16(a[i], b[j], acc[i][j]);
  }
  const bool split = gridDim.z > 1;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) {
      const int n = n0 + j * 16 + (lane & 15);
      if (n >= g.N) continue;
      for (int q = 0; q < 4; ++q) {
        const int m = m0 + i * 16 + (lane >> 4) * 4 + q;
        if (m >= g.M) continue;
        float v = acc[i][j][q];
        const long long o = (long long)m * g.ldc + n;
        if (split) {
          atomicAdd(&g.C[o], v);
          continue;
        }
        if (g.epi == EPI_BIAS_GELU) {
          v += g.bias[n];
          if (g.Z) g.Z[o] = v;
          v = gelu(v);
        } else if (g.epi == EPI_BIAS) {
          v += g.bias[n];
        } else if (g.epi == EPI_GELU_GRAD) {
          v *= gelu_grad(g.Zin[o]);
        }
        g.C[o] = v;
      }
    }
}

// dY = 2 (Y - T) / (B * O); loss += sum (Y - T)^2 / (B * O)
__global__ void k_mse_grad(const float* Y, const float* T, float* dY, float* loss, long long n,
                           float scale) {
  float acc = 0.f;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float d = Y[i] - T[i];
    dY[i] = 2.f * d * scale;
    acc += d * d * scale;
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o);
  if ((threadIdx.x & 63) == 0 && loss) atomicAdd(loss, acc);
}

// out[n] += sum_m X[m * ld + n] (bias gradients). Block (x, y) sums rows
// chunk y of the column tile [256 x, 256 x + nc): its 256 threads form
// 256 / nc row groups of nc adjacent columns (each group reads contiguous row
// segments), the groups' partial sums meet in LDS, and one atomic per column
// and block adds them to out. Every thread works whatever N is (4-wide
// output layers included).
__global__ void __launch_bounds__(256) k_colsum(const float* X, long long ld, int M, int N,
                                               float* out) {
  __shared__ float red[256];
  const int c0 = blockIdx.x * 256;
  const int nc = min(256, N - c0);
  const int groups = 256 / nc;
  const int t = threadIdx.x, g = t / nc, c = t % nc;
  const int rows = (M + gridDim.y - 1) / gridDim.y;
  const int mb = blockIdx.y * rows, me = min(M, mb + rows);
  float acc = 0.f;
  if (g < groups)
    for (int m = mb + g; m < me; m += groups) acc += X[(long long)m * ld + c0 + c];
  red[t] = acc;   // threads past groups * nc hold 0 and are not read
  __syncthreads();
  if (t < nc) {
    float sum = 0.f;
    for (int q = 0; q < groups; ++q) sum += red[q * nc + t];
    atomicAdd(&out[c0 + t], sum);
  }
}

// Adam (Kingma and Ba 2015) with bias correction, on a flat parameter vector.
__global__ void k_adam(float* w, const float* g, float* m, float* v, long long n, float lr,
                       float b1, float b2, float eps, float c1, float c2, float wd) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float gi = g[i] + wd * w[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

__global__ void k_fill(float* p, long long n, float value) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = value;
}

__global__ void k_to_bf16(const float* x, short* y, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    y[i] = f2bf(x[i]);
}

__global__ void k_to_f16(const float* x, short* y, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    y[i] = __builtin_bit_cast(short, (_Float16)x[i]);
}

// ---------------------------------------------------------------------------
// Fused inference: normalise, three layers, de-standardise, clip.
// ---------------------------------------------------------------------------

struct InferArgs {
  const float* data;     // (IN, N) measurements, voxel fastest
  const short* w1;       // (H, INP) fp16
  const float* b1;       // (H)
  const short* w2;       // (H, H) fp16
  const float* b2;       // (H)
  const short* w3;       // (OUT, H) fp16
  const float* b3;       // (OUT)
  const float* mu;       // (OUT) output de-standardisation: p = mu + sd * y
  const float* sd;       // (OUT)
  const float* lb;       // (OUT, N)
  const float* ub;       // (OUT, N)
  float* params;         // (OUT, N) out
  long long N;
  int IN, OUT;
  int norm_index;        // measurement whose value s normalises the signal
  int scale_param;       // parameter predicted relative to s (p *= s), -1: none
  int logshift_param;    // parameter predicted as a log relative to s (p += log s), -1: none
  int log_mask;          // bit j: parameter j is predicted as log(p_j)
  const double* fb_P;    // fallback log-linear guess (loglinear.h) for voxels whose
  int fb_K, fb_kind;     //   outputs are not finite; null: params keep what they held
};

// Register-resident inference in the transposed form: every layer computes
// out^T (units x voxels) = W (units x k) . in^T (k x voxels). The MFMA result
// layout (lane: column = voxel l%16, rows 4(l/16)+0..3 = units) is then the
// B-operand layout of the next layer, so activations stay in registers
// between layers, and each weight fragment (A operand) is reused for VT voxel
// tiles. Only the final OUT x 16VT outputs pass through LDS, to gather each
// voxel's parameters into one lane.
//
// Operands are fp16, accumulation fp32; every accumulator starts from its
// bias, and the activation is the packed-fp16 polynomial GELU (gelu_pk).
// Layer 1 uses v_mfma_f32_16x16x16_f16; layers 2 and 3 use
// v_mfma_f32_16x16x32_f16, whose B operand (k = 8 (l/16) + 0..7) is the
// concatenation of two consecutive 16-unit activation tiles: k slot 8g + i
// holds unit 32c + 4g + i (i < 4) or 32c + 16 + 4g + i - 4 (i >= 4) of k
// block c, and the weight fragments are gathered with the same permutation.
// LDSW stages weights once per workgroup into LDS in fragment order
// (conflict-free 8- and 16-byte reads) on a persistent grid: 1 = all three
// weight matrices (the 128-unit networks), 2 = W2 only (the 256-unit ones,
// whose 128 KB W2 takes a workgroup of its own per CU); 0 = none.
template <int H, int INP, int OUTP, int VT, int LDSW, int OUTC>
__global__ void __launch_bounds__(256) k_mlp_infer_t(InferArgs a) {
  // outputs: OUTC when fixed at compile time (the per-output guards fold away), else a.OUT
  const int nout = OUTC > 0 ? OUTC : a.OUT;
  constexpr int UT = H / 16;     // unit tiles per hidden layer
  constexpr int KC = H / 32;     // k blocks of 32 units
  constexpr int KT1 = INP / 16;  // k-steps of layer 1
  constexpr int OT = OUTP / 16;  // output tiles
  constexpr int V = 16 * VT;     // voxels per wavefront
  static_assert(V <= 64, "at most one voxel per lane in the epilogue");
  constexpr int W2OFF = H * INP, W3OFF = H * INP + H * H;
  constexpr int W2L = LDSW == 1 ? W2OFF : 0;   // offset of W2 in wl
  constexpr int WLN = LDSW == 1 ? W3OFF + OUTP * H : (LDSW == 2 ? H * H : 8);
  __shared__ float ybuf[4][OUTP * V];
  __shared__ __attribute__((aligned(16))) short wl[WLN];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int col = lane & 15;      // voxel within a tile
  const int kq = (lane >> 4) * 4; // k (or row) offset within a tile
  if constexpr (LDSW == 1) {
    stage_fragments<H, INP>(wl, a.w1, H);
    stage_fragments_k32<OUTP, H>(wl + W3OFF, a.w3, nout);
  }
  if constexpr (LDSW != 0) {
    stage_fragments_k32<H, H>(wl + W2L, a.w2, H);
    __syncthreads();
  }
  // weight fragments: w1 (row tile r, k tile k); w2, w3 (row tile r, k block c)
  auto w1f = [&](int r, int k) -> half4 {
    if constexpr (LDSW == 1) return *(const half4*)(wl + ((r * KT1 + k) * 64 + lane) * 4);
    else return *(const half4*)(a.w1 + (long long)(r * 16 + col) * INP + k * 16 + kq);
  };
  auto w2f = [&](int r, int c) -> half8 {
    if constexpr (LDSW != 0) {
      return *(const half8*)(wl + W2L + ((r * KC + c) * 64 + lane) * 8);
    } else {
      const short* p = a.w2 + (long long)(r * 16 + col) * H + c * 32 + kq;
      return cat8(*(const half4*)p, *(const half4*)(p + 16));
    }
  };
  auto w3f = [&](int r, int c) -> half8 {
    if constexpr (LDSW == 1) {
      return *(const half8*)(wl + W3OFF + ((r * KC + c) * 64 + lane) * 8);
    } else {
      half8 w = half8{0, 0, 0, 0, 0, 0, 0, 0};   // rows of W3 past OUT are zero
      if (r * 16 + col < nout) {
        const short* p = a.w3 + (long long)(r * 16 + col) * H + c * 32 + kq;
        w = cat8(*(const half4*)p, *(const half4*)(p + 16));
      }
      return w;
    }
  };
  float* yb = ybuf[wave];
  const long long nwaves = (long long)gridDim.x * 4;
  for (long long base = ((long long)blockIdx.x * 4 + wave) * V; base < a.N; base += nwaves * V) {
    // layer-1 B operands: normalised inputs x^T (k = feature, n = voxel); the
    // normaliser of tile t stays in sv[t] for the epilogue
    half4 xb[KT1][VT];
    float sv[VT];
#pragma unroll
    for (int t = 0; t < VT; ++t) {
      const long long v = base + t * 16 + col;
      const bool ok = v < a.N;
      float s = ok ? a.data[(long long)a.norm_index * a.N + v] : 1.f;
      float xr[KT1][4];
#pragma unroll
      for (int kt = 0; kt < KT1; ++kt)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int i = kt * 16 + kq + q;
          xr[kt][q] = (ok && i < a.IN) ? a.data[(long long)i * a.N + v] : 0.f;
        }
      if (!(s > 0.f)) s = 1.f;
      sv[t] = s;
      const float inv = 1.f / s;
#pragma unroll
      for (int kt = 0; kt < KT1; ++kt)
#pragma unroll
        for (int q = 0; q < 4; ++q) xb[kt][t][q] = (_Float16)(xr[kt][q] * inv);
    }
    // epilogue operands, loaded now so their latency hides under the layers:
    // lane l finishes voxel base + l
    const long long ve = base + lane;
    const bool mine = lane < V && ve < a.N;
    float lo[OUTP], hi[OUTP];
#pragma unroll
    for (int j = 0; j < OUTP; ++j)
      if (j < nout && mine) {
        lo[j] = a.lb[(long long)j * a.N + ve];
        hi[j] = a.ub[(long long)j * a.N + ve];
      }
    float se = sv[0];
#pragma unroll
    for (int t = 1; t < VT; ++t)
      if ((lane >> 4) == t) se = sv[t];
    // layer 1: h1^T = gelu(W1 x^T + b1)
    half4 h1[UT][VT];
#pragma unroll
    for (int u = 0; u < UT; ++u) {
      floatx4 bias;
#pragma unroll
      for (int q = 0; q < 4; ++q) bias[q] = a.b1[u * 16 + kq + q];
      floatx4 acc[VT];
#pragma unroll
      for (int t = 0; t < VT; ++t) acc[t] = bias;
#pragma unroll
      for (int kt = 0; kt < KT1; ++kt) {
        const half4 w = w1f(u, kt);
#pragma unroll
        for (int t = 0; t < VT; ++t) acc[t] = mfma16h(w, xb[kt][t], acc[t]);
      }
#pragma unroll
      for (int t = 0; t < VT; ++t) h1[u][t] = gelu_pk4(acc[t]);
    }
    // layers 2 and 3 fused: each pair of 16-unit tiles of h2^T = gelu(W2 h1^T
    // + b2) is consumed at once as one k block of y^T = W3 h2^T + b3, so h2
    // never needs to be held whole
    floatx4 acc3[OT][VT];
#pragma unroll
    for (int o = 0; o < OT; ++o) {
      floatx4 b3v;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int j = o * 16 + kq + q;
        b3v[q] = j < nout ? a.b3[j] : 0.f;
      }
#pragma unroll
      for (int t = 0; t < VT; ++t) acc3[o][t] = b3v;
    }
    // rolled over k blocks of h2: only the weight fragments of two unit tiles
    // are live at a time, and their accumulators give the matrix pipe 2 VT
    // independent chains
#pragma unroll 1
    for (int c = 0; c < KC; ++c) {
      floatx4 acc[2][VT];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        floatx4 bias;
#pragma unroll
        for (int q = 0; q < 4; ++q) bias[q] = a.b2[(2 * c + h) * 16 + kq + q];
#pragma unroll
        for (int t = 0; t < VT; ++t) acc[h][t] = bias;
      }
#pragma unroll
      for (int kc = 0; kc < KC; ++kc) {
        half8 w[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) w[h] = w2f(2 * c + h, kc);
#pragma unroll
        for (int t = 0; t < VT; ++t) {
          const half8 b = cat8(h1[2 * kc][t], h1[2 * kc + 1][t]);
#pragma unroll
          for (int h = 0; h < 2; ++h) acc[h][t] = mfma32h(w[h], b, acc[h][t]);
        }
      }
      half4 h2[2][VT];
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int t = 0; t < VT; ++t) h2[h][t] = gelu_pk4(acc[h][t]);
#pragma unroll
      for (int o = 0; o < OT; ++o) {
        const half8 w = w3f(o, c);
#pragma unroll
        for (int t = 0; t < VT; ++t) acc3[o][t] = mfma32h(w, cat8(h2[0][t], h2[1][t]), acc3[o][t]);
      }
    }
#pragma unroll
    for (int o = 0; o < OT; ++o)
#pragma unroll
      for (int t = 0; t < VT; ++t)
#pragma unroll
        for (int q = 0; q < 4; ++q) yb[(o * 16 + kq + q) * V + t * 16 + col] = acc3[o][t][q];
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    // one lane per voxel: inverse transform, finite check, clip, store
    if (mine) {
      const float s = se;
      bool finite = true;
      float p[OUTP];
#pragma unroll
      for (int j = 0; j < OUTP; ++j) {
        if (j < nout) {
          float t = a.mu[j] + a.sd[j] * yb[j * V + lane];
          if ((a.log_mask >> j) & 1) t = expf(t);
          if (j == a.scale_param) t *= s;
          if (j == a.logshift_param) t += logf(s);
          finite = finite && !(__builtin_isnan(t) || __builtin_isinf(t));
          p[j] = t;
        }
      }
      if (!finite && a.fb_P) {
        // rare: the log-linear guess instead, nan_to_num as pca_loglinear_init
        pca_init::loglinear_guess<float, OUTP>(a.fb_P, a.fb_K, a.IN, a.data, a.N, ve,
                                               a.fb_kind, p);
#pragma unroll
        for (int j = 0; j < OUTP; ++j) p[j] = pca_init::nan_to_num(p[j]);
        finite = true;
      }
      if (finite) {
#pragma unroll
        for (int j = 0; j < OUTP; ++j) {
          if (j < nout) {
            float t = p[j] < lo[j] ? lo[j] : p[j];
            t = t > hi[j] ? hi[j] : t;
            a.params[(long long)j * a.N + ve] = t;
          }
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
}

// Training inputs from signals and true parameters: X (INP, N) = y / y[norm]
// (rows >= IN zero), T (N, OUT) = (transform(p) - mu) / sd with the transform
// that k_mlp_infer_t inverts.
__global__ void k_prepare(const float* y, const float* truth, long long N, int IN, int INP,
                          int OUT, int norm_index, int scale_param, int logshift_param,
                          int log_mask, const float* mu, const float* sd, float* X, float* T) {
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < N;
       v += (long long)gridDim.x * blockDim.x) {
    float s = y[(long long)norm_index * N + v];
    if (!(s > 0.f)) s = 1.f;
    for (int i = 0; i < INP; ++i) X[(long long)i * N + v] = i < IN ? y[(long long)i * N + v] / s : 0.f;
    for (int j = 0; j < OUT; ++j) {
      float t = truth[(long long)j * N + v];
      if (j == scale_param) t /= s;
      if (j == logshift_param) t -= logf(s);
      if ((log_mask >> j) & 1) t = logf(t);
      T[v * OUT + j] = (t - mu[j]) / sd[j];
    }
  }
}

int grid_1d(long long n) { return (int)std::min<long long>((n + 255) / 256, 4096); }

int check_launch() {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? PCA_OK : pca_fail(PCA_ERR_HIP, hipGetErrorString(e));
}

}  // namespace

extern "C" int pca_mlp_gemm(const float* A, int64_t lda, int32_t transA, const float* B,
                            int64_t ldb, int32_t transB, float* C, int64_t ldc, int32_t M,
                            int32_t N, int32_t K, int32_t epi, const float* bias, float* Z,
                            const float* Zin, int32_t splits, void* stream) {
  if (!A || !B || !C || M <= 0 || N <= 0 || K <= 0 || epi < 0 || epi > 3 || splits < 1)
    return pca_fail(PCA_ERR_ARG, "pca_mlp_gemm: bad argument");
  if ((epi == EPI_BIAS_GELU || epi == EPI_BIAS) && !bias)
    return pca_fail(PCA_ERR_ARG, "pca_mlp_gemm: epilogue needs a bias");
  if (epi == EPI_GELU_GRAD && !Zin) return pca_fail(PCA_ERR_ARG, "pca_mlp_gemm: needs Zin");
  if (splits > 1 && epi != EPI_STORE)
    return pca_fail(PCA_ERR_ARG, "pca_mlp_gemm: split-K only with the plain epilogue");
  GemmArgs g{A, B, C, bias, Z, Zin, lda, ldb, ldc, M, N, K, transA, transB, epi, 0};
  int kps = (K + splits - 1) / splits;
  kps = (kps + 15) / 16 * 16;
  g.k_per_split = kps;
  const int zs = (K + kps - 1) / kps;
  dim3 grid((M + 63) / 64, (N + 63) / 64, zs);
  hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, (hipStream_t)stream, g);
  return check_launch();
}

extern "C" int pca_mlp_mse_grad(const float* Y, const float* T, float* dY, float* loss, int64_t n,
                                float scale, void* stream) {
  if (!Y || !T || !dY || n <= 0) return pca_fail(PCA_ERR_ARG, "pca_mlp_mse_grad: bad argument");
  hipLaunchKernelGGL(k_mse_grad, dim3(grid_1d(n)), dim3(256), 0, (hipStream_t)stream, Y, T, dY,
                     loss, (long long)n, scale);
  return check_launch();
}

extern "C" int pca_mlp_colsum(const float* X, int64_t ld, int32_t M, int32_t N, float* out,
                              void* stream) {
  if (!X || !out || M <= 0 || N <= 0) return pca_fail(PCA_ERR_ARG, "pca_mlp_colsum: bad argument");
  // chunks of 64 rows (at most 1024 chunks): a few dependent loads per thread
  dim3 grid((N + 255) / 256, std::min(1024, (M + 63) / 64));
  hipLaunchKernelGGL(k_colsum, grid, dim3(256), 0, (hipStream_t)stream, X, (long long)ld, M, N,
                     out);
  return check_launch();
}

extern "C" int pca_mlp_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr,
                            float b1, float b2, float eps, int32_t step, float weight_decay,
                            void* stream) {
  if (!w || !g || !m || !v || n <= 0 || step < 1)
    return pca_fail(PCA_ERR_ARG, "pca_mlp_adam: bad argument");
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(grid_1d(n)), dim3(256), 0, (hipStream_t)stream, w, g, m, v,
                     (long long)n, lr, b1, b2, eps, c1, c2, weight_decay);
  return check_launch();
}

extern "C" int pca_mlp_fill(float* p, int64_t n, float value, void* stream) {
  if (!p || n < 0) return pca_fail(PCA_ERR_ARG, "pca_mlp_fill: bad argument");
  if (n == 0) return PCA_OK;
  hipLaunchKernelGGL(k_fill, dim3(grid_1d(n)), dim3(256), 0, (hipStream_t)stream, p,
                     (long long)n, value);
  return check_launch();
}

extern "C" int pca_mlp_to_bf16(const float* x, int16_t* y, int64_t n, void* stream) {
  if (!x || !y || n <= 0) return pca_fail(PCA_ERR_ARG, "pca_mlp_to_bf16: bad argument");
  hipLaunchKernelGGL(k_to_bf16, dim3(grid_1d(n)), dim3(256), 0, (hipStream_t)stream, x,
                     (short*)y, (long long)n);
  return check_launch();
}

extern "C" int pca_mlp_to_f16(const float* x, int16_t* y, int64_t n, void* stream) {
  if (!x || !y || n <= 0) return pca_fail(PCA_ERR_ARG, "pca_mlp_to_f16: bad argument");
  hipLaunchKernelGGL(k_to_f16, dim3(grid_1d(n)), dim3(256), 0, (hipStream_t)stream, x,
                     (short*)y, (long long)n);
  return check_launch();
}

extern "C" int pca_mlp_infer(const pca_mlp_infer_args* args, void* stream) {
  if (!args || !args->w1 || !args->w2 || !args->w3 || !args->mu || !args->sd || !args->b1 ||
      !args->b2 || !args->b3)
    return pca_fail(PCA_ERR_ARG, "pca_mlp_infer: NULL argument");
  // per-voxel arrays: empty ones (N = 0) may have no address
  if (args->N > 0 && (!args->data || !args->params || !args->lb || !args->ub))
    return pca_fail(PCA_ERR_ARG, "pca_mlp_infer: NULL argument");
  const int H = args->hidden, IN = args->in, OUT = args->out;
  if (IN < 1 || IN > 64 || OUT < 1 || OUT > 32 ||
      (H != 32 && H != 64 && H != 128 && H != 256) || args->N < 0 || args->norm_index < 0 ||
      args->norm_index >= IN || (H < 128 && (IN > 16 || OUT > 16)))
    return pca_fail(PCA_ERR_ARG, "pca_mlp_infer: sizes must be in <= 64, out <= 32, hidden "
                                 "32|64|128|256 (32, 64: in <= 16, out <= 16)");
  if (args->fallback_P) {
    const int K = args->fallback_K, kind = args->fallback_kind;
    if ((kind == 0 && K != OUT) || (kind == 1 && (K != 2 || OUT != 2)) ||
        (kind == 2 && (K != 2 || OUT != 3)) || (kind == 3 && (K != 4 || OUT != 4)) ||
        kind < 0 || kind > 3)
      return pca_fail(PCA_ERR_ARG, "pca_mlp_infer: fallback K/kind do not match the outputs");
  }
  if (args->N == 0) return PCA_OK;
  InferArgs a{(const float*)args->data, (const short*)args->w1, args->b1, (const short*)args->w2,
              args->b2, (const short*)args->w3, args->b3, args->mu, args->sd,
              (const float*)args->lb, (const float*)args->ub, (float*)args->params, args->N, IN,
              OUT, args->norm_index, args->scale_param, args->logshift_param, args->log_mask,
              args->fallback_P, args->fallback_K, args->fallback_kind};
  // the input width is padded to a multiple of 16 and stored with that stride
  const int INP = IN <= 16 ? 16 : (IN <= 32 ? 32 : 64);
  const int OUTP = OUT <= 16 ? 16 : 32;
  const bool big = H == 256;
  const int V = big ? 32 : 64;  // voxels per wavefront
  const long long waves = (args->N + V - 1) / V;
  // weights in LDS when they fit next to the output buffer in 64 KB: then a
  // persistent grid of 512 workgroups (2 per CU) reuses the staged copy
  const bool ldsw = !big && OUTP == 16 && INP <= 32;
  // 256-unit networks stage W2 (128 KB) instead: one persistent workgroup per CU
  const int grid = (int)std::min<long long>((waves + 3) / 4, ldsw ? 512 : (big ? 256 : 8192));
  hipStream_t s = (hipStream_t)stream;
#define PCA_INFER(HH, II, OO, VT, L, OC) \
  hipLaunchKernelGGL((k_mlp_infer_t<HH, II, OO, VT, L, OC>), dim3(grid), dim3(256), 0, s, a)
  // the small-output networks of the built-in models get their output count
  // at compile time
#define PCA_INFER_OUT(HH, II, OO, VT, L)                                  \
  do {                                                                    \
    if (OUT == 2) PCA_INFER(HH, II, OO, VT, L, 2);                        \
    else if (OUT == 3) PCA_INFER(HH, II, OO, VT, L, 3);                   \
    else if (OUT == 4) PCA_INFER(HH, II, OO, VT, L, 4);                   \
    else PCA_INFER(HH, II, OO, VT, L, 0);                                 \
  } while (0)
  if (H == 32) {
    PCA_INFER_OUT(32, 16, 16, 4, 1);   // the narrow IVIM-sized networks
  } else if (H == 64) {
    PCA_INFER_OUT(64, 16, 16, 4, 1);
  } else if (!big) {
    if (INP == 16 && OUTP == 16) PCA_INFER_OUT(128, 16, 16, 4, 1);
    else if (INP == 16) PCA_INFER(128, 16, 32, 4, 0, 0);
    else if (INP == 32 && OUTP == 16) PCA_INFER_OUT(128, 32, 16, 4, 1);
    else if (INP == 32) PCA_INFER(128, 32, 32, 4, 0, 0);
    else if (OUTP == 16) PCA_INFER(128, 64, 16, 4, 0, 0);
    else PCA_INFER(128, 64, 32, 4, 0, 0);
  } else {
    if (INP == 16 && OUTP == 16) PCA_INFER(256, 16, 16, 2, 2, 0);
    else if (INP == 16) PCA_INFER(256, 16, 32, 2, 2, 0);
    else if (INP == 32 && OUTP == 16) PCA_INFER(256, 32, 16, 2, 2, 0);
    else if (INP == 32) PCA_INFER(256, 32, 32, 2, 2, 0);
    else if (OUTP == 16) PCA_INFER(256, 64, 16, 2, 2, 0);
    else PCA_INFER(256, 64, 32, 2, 2, 0);
  }
#undef PCA_INFER_OUT
#undef PCA_INFER
  return check_launch();
}

extern "C" int pca_mlp_prepare(const float* y, const float* truth, int64_t N, int32_t in,
                               int32_t in_padded, int32_t out, int32_t norm_index,
                               int32_t scale_param, int32_t logshift_param, int32_t log_mask,
                               const float* mu, const float* sd, float* X, float* T,
                               void* stream) {
  if (!y || !truth || !mu || !sd || !X || !T || N <= 0 || in < 1 || in_padded < in ||
      out < 1 || out > 32 || norm_index < 0 || norm_index >= in)
    return pca_fail(PCA_ERR_ARG, "pca_mlp_prepare: bad argument");
  hipLaunchKernelGGL(k_prepare, dim3(grid_1d(N)), dim3(256), 0, (hipStream_t)stream, y, truth,
                     (long long)N, in, in_padded, out, norm_index, scale_param, logshift_param,
                     log_mask, mu, sd, X, T);
  return check_launch();
}
