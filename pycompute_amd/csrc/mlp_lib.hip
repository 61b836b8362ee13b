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
  float* ws;          // split-K partials (gridDim.z, M, N), summed in order after; null: atomics
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
          if (g.ws)
            g.ws[((long long)blockIdx.z * g.M + m) * g.N + n] = v;
          else
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
                                               float* out, float* ws) {
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
    if (ws)
      ws[(long long)blockIdx.y * N + c0 + t] = sum;   // summed in order by k_sum_rows
    else
      atomicAdd(&out[c0 + t], sum);
  }
}

// out[m * ldo + n] (+)= sum over z < Z of ws[(z M + m) N + n], z in order:
// the fixed-order end of the split-K products and column sums.
__global__ void k_sum_rows(const float* ws, int Z, int M, int N, float* out, long long ldo,
                           int accumulate) {
  const long long MN = (long long)M * N;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < MN;
       i += (long long)gridDim.x * blockDim.x) {
    // 8 loads in flight, added in z order (the same bits as one at a time)
    float s = 0.f;
    for (int z0 = 0; z0 < Z; z0 += 8) {
      float v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = z0 + u < Z ? ws[(z0 + u) * MN + i] : 0.f;
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (z0 + u < Z) s += v[u];
    }
    const long long o = (i / N) * ldo + i % N;
    out[o] = accumulate ? out[o] + s : s;
  }
}

// Adam (Kingma and Ba 2015) with bias correction c1 = 1 - b1^t, c2 = 1 - b2^t.
struct AdamArgs {
  float* w;
  float* m;
  float* v;
  float lr, b1, b2, eps, c1, c2, wd;
};

__device__ __forceinline__ void adam_update(const AdamArgs& a, long long i, float g) {
  const float gi = g + a.wd * a.w[i];
  const float mi = a.b1 * a.m[i] + (1.f - a.b1) * gi;
  const float vi = a.b2 * a.v[i] + (1.f - a.b2) * gi * gi;
  a.m[i] = mi;
  a.v[i] = vi;
  a.w[i] -= a.lr * (mi / a.c1) / (sqrtf(vi / a.c2) + a.eps);
}

// Adam on a flat parameter vector.
__global__ void k_adam(AdamArgs a, const float* g, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    adam_update(a, i, g[i]);
}

// ---------------------------------------------------------------------------
// Fused training step. One workgroup takes 64 rows of the batch through the
// whole step -- forward (three layers), MSE gradient, backward -- with the
// weights, its inputs and every activation and gradient matrix held in LDS as
// bf16 (the MFMA operands; fp32 accumulation), and writes its partial
// gradient of every parameter and its partial loss to one row of a
// workspace. k_grad_reduce sums the rows in a fixed order (no atomics, so a
// step is deterministic) and applies Adam to the sums.
//
// LDS arrays of activations and gradients are [feature][batch row] (row
// length 64 + 4, the pad spreading 8-byte fragment reads over the banks), so
// a C fragment (4 consecutive rows of one column per lane) is one 8-byte
// store and the weight gradients (contractions over the batch) read both
// operands 8 bytes at a time; the forward and backward-data products read
// their batch-row operand with 4 two-byte reads. Weights keep their
// [out][in] layout. Rounding is that of the GEMM path: bf16 operands (the
// activations, weights, dY and the dZ), fp32 pre-activations, GELU
// derivatives, dY, dZ and bias sums.
// ---------------------------------------------------------------------------

constexpr int TRB = 64;   // batch rows per workgroup

struct TrainArgs {
  const float* X;     // (in_padded, ldx) feature-major; column 0 = first row of the batch
  const float* T;     // (B, O) targets of the batch
  const float* w;     // flat parameters
  float* part;        // (B / TRB, np) partial gradients (entries outside the blocks untouched)
  float* loss_part;   // (B / TRB) partial losses
  long long ldx;
  int o1, ob1, o2, ob2, o3, ob3;   // offsets of w1 b1 w2 b2 w3 b3 in w
  int np, O;
  float scale;                     // 1 / (B O)
};

__device__ __forceinline__ bf16x4 lds_row4(const short* p) { return *(const bf16x4*)p; }
__device__ __forceinline__ bf16x4 lds_col4(const short* p, int ld) {
  return bf16x4{p[0], p[ld], p[2 * ld], p[3 * ld]};
}

// acc[j] += A B over one wave's 16 batch rows: A (16 x 16 KS) from a
// [k][batch row] array (ld lda) at row row0; B (16 KS x 16 NT) from a
// [col][k] array (BT false) or a [k][col] array (BT true), ld ldb.
template <int NT, int KS, bool BT>
__device__ __forceinline__ void wave_rows(floatx4* acc, const short* A, int lda, int row0,
                                          const short* B, int ldb, int lane) {
  const int r = lane & 15, g = lane >> 4;
#pragma unroll
  for (int kk = 0; kk < KS; ++kk) {
    const int k = 16 * kk + 4 * g;
    const bf16x4 a = lds_col4(A + k * lda + row0 + r, lda);
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const int n = 16 * j + r;
      const bf16x4 b = BT ? lds_col4(B + k * ldb + n, ldb) : lds_row4(B + n * ldb + k);
      acc[j] = mfma16(a, b, acc[j]);
    }
  }
}

// One 16 x 16 tile of A^T B, contracted over the workgroup's batch rows: rows
// ra.. of A and rb.. of B, both [feature][batch row] arrays.
__device__ __forceinline__ floatx4 tile_tn(const short* A, int ra, const short* B, int rb,
                                           int lane) {
  constexpr int L = TRB + 4;
  const int r = lane & 15, g = lane >> 4;
  floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kk = 0; kk < TRB / 16; ++kk) {
    const int k = 16 * kk + 4 * g;
    acc = mfma16(lds_row4(A + (ra + r) * L + k), lds_row4(B + (rb + r) * L + k), acc);
  }
  return acc;
}

// A C fragment of tile (rows ra.., columns cb..) into the row-major p (row
// stride ld); rows from rmax on are dropped.
__device__ __forceinline__ void store_tile(float* p, int ld, int ra, int cb, int rmax, floatx4 v,
                                           int lane) {
  const int r = lane & 15, g = lane >> 4;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int row = ra + 4 * g + q;
    if (row < rmax) p[row * ld + cb + r] = v[q];
  }
}

// bf16 of the 4 rows a lane holds, into a [column][batch row] array.
__device__ __forceinline__ void put_rows(short* p, floatx4 v) {
  *(bf16x4*)p = bf16x4{f2bf(v[0]), f2bf(v[1]), f2bf(v[2]), f2bf(v[3])};
}

// Sum over the 16 rows of a wave's C fragment: lanes 0..15 hold column r.
__device__ __forceinline__ float col_sum(floatx4 v) {
  float s = (v[0] + v[1]) + (v[2] + v[3]);
  s += __shfl_xor(s, 16);
  s += __shfl_xor(s, 32);
  return s;
}

template <int H, int I, int OP>
__global__ void __launch_bounds__(256) k_train_step(TrainArgs a) {
  constexpr int L = TRB + 4, LI = I + 4, LH = H + 4;
  constexpr int NH = H / 16, NI = I / 16, NO = OP / 16, NR = 2 * H + OP + 1;
  __shared__ __attribute__((aligned(16))) short sX[I * L];     // X       [i][s]
  __shared__ __attribute__((aligned(16))) short sW1[H * LI];   // w1      [h][i]
  __shared__ __attribute__((aligned(16))) short sW2[H * LH];   // w2      [h'][h]
  __shared__ __attribute__((aligned(16))) short sW3[OP * LH];  // w3      [o][h], zero rows >= O
  __shared__ __attribute__((aligned(16))) short sH1[H * L];    // H1      [h][s]
  __shared__ __attribute__((aligned(16))) short sH2[H * L];    // H2, then dZ1
  __shared__ __attribute__((aligned(16))) short sZ2[H * L];    // dZ2
  __shared__ __attribute__((aligned(16))) short sD[OP * L];    // dY      [o][s]
  __shared__ float sred[4][NR];   // per wave: bias sums b1 | b2 | b3, loss
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6, r = lane & 15, g = lane >> 4;
  const long long s0 = (long long)blockIdx.x * TRB;
  const int row0 = 16 * wv;   // the wave's batch rows in the row-tiled products

  // staging: the trip counts are compile-time, so every thread's loads are in
  // flight together before the first conversion waits
  static_assert((H * I) % 256 == 0 && (H * H) % 256 == 0 && (OP * H) % 256 == 0 &&
                    (I * TRB) % 256 == 0, "staging loops take whole rounds of the block");
  {
    float v1[H * I / 256], v2[H * H / 256], v3[OP * H / 256], vx[I * TRB / 256];
#pragma unroll
    for (int k = 0; k < H * I / 256; ++k) v1[k] = a.w[a.o1 + t + 256 * k];
#pragma unroll
    for (int k = 0; k < H * H / 256; ++k) v2[k] = a.w[a.o2 + t + 256 * k];
#pragma unroll
    for (int k = 0; k < OP * H / 256; ++k) {
      const int e = t + 256 * k;
      v3[k] = e / H < a.O ? a.w[a.o3 + e] : 0.f;
    }
#pragma unroll
    for (int k = 0; k < I * TRB / 256; ++k) {
      const int e = t + 256 * k;
      vx[k] = a.X[(e / TRB) * a.ldx + s0 + e % TRB];
    }
#pragma unroll
    for (int k = 0; k < H * I / 256; ++k) {
      const int e = t + 256 * k;
      sW1[(e / I) * LI + e % I] = f2bf(v1[k]);
    }
#pragma unroll
    for (int k = 0; k < H * H / 256; ++k) {
      const int e = t + 256 * k;
      sW2[(e / H) * LH + e % H] = f2bf(v2[k]);
    }
#pragma unroll
    for (int k = 0; k < OP * H / 256; ++k) {
      const int e = t + 256 * k;
      sW3[(e / H) * LH + e % H] = f2bf(v3[k]);
    }
#pragma unroll
    for (int k = 0; k < I * TRB / 256; ++k) {
      const int e = t + 256 * k;
      sX[(e / TRB) * L + e % TRB] = f2bf(vx[k]);
    }
  }
  __syncthreads();

  floatx4 acc[NH];
  floatx4 gg1[NH], gg2[NH];   // gelu'(Z1), gelu'(Z2) of the wave's rows
  // H1 = gelu(X w1^T + b1)
#pragma unroll
  for (int j = 0; j < NH; ++j) acc[j] = floatx4{0.f, 0.f, 0.f, 0.f};
  wave_rows<NH, NI, false>(acc, sX, L, row0, sW1, LI, lane);
#pragma unroll
  for (int j = 0; j < NH; ++j) {
    const int h = 16 * j + r;
    const float b = a.w[a.ob1 + h];
    floatx4 hv;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float z = acc[j][q] + b;
      gg1[j][q] = gelu_grad(z);
      hv[q] = gelu(z);
    }
    put_rows(sH1 + h * L + row0 + 4 * g, hv);
  }
  __syncthreads();
  // H2 = gelu(H1 w2^T + b2)
#pragma unroll
  for (int j = 0; j < NH; ++j) acc[j] = floatx4{0.f, 0.f, 0.f, 0.f};
  wave_rows<NH, NH, false>(acc, sH1, L, row0, sW2, LH, lane);
#pragma unroll
  for (int j = 0; j < NH; ++j) {
    const int h = 16 * j + r;
    const float b = a.w[a.ob2 + h];
    floatx4 hv;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float z = acc[j][q] + b;
      gg2[j][q] = gelu_grad(z);
      hv[q] = gelu(z);
    }
    put_rows(sH2 + h * L + row0 + 4 * g, hv);
  }
  __syncthreads();
  // Y = H2 w3^T + b3; dY = 2 (Y - T) / (B O); loss = sum (Y - T)^2 / (B O)
  {
    floatx4 y[NO];
#pragma unroll
    for (int j = 0; j < NO; ++j) y[j] = floatx4{0.f, 0.f, 0.f, 0.f};
    wave_rows<NO, NH, false>(y, sH2, L, row0, sW3, LH, lane);
    float ls = 0.f;
#pragma unroll
    for (int j = 0; j < NO; ++j) {
      const int o = 16 * j + r;
      floatx4 d = {0.f, 0.f, 0.f, 0.f};
      if (o < a.O) {
        const float b = a.w[a.ob3 + o];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const float df = y[j][q] + b - a.T[(s0 + row0 + 4 * g + q) * a.O + o];
          ls += df * df * a.scale;
          d[q] = 2.f * df * a.scale;
        }
      }
      const float cs = col_sum(d);
      if (g == 0) sred[wv][2 * H + o] = cs;
      put_rows(sD + o * L + row0 + 4 * g, d);
    }
    for (int o = 32; o > 0; o >>= 1) ls += __shfl_xor(ls, o);
    if (lane == 0) sred[wv][NR - 1] = ls;
  }
  __syncthreads();
  float* P = a.part + (long long)blockIdx.x * a.np;
  // dw3 = dY^T H2
  for (int tt = wv; tt < NO * NH; tt += 4) {
    const int ro = tt / NH, co = tt % NH;
    store_tile(P + a.o3, H, 16 * ro, 16 * co, a.O, tile_tn(sD, 16 * ro, sH2, 16 * co, lane),
               lane);
  }
  // dZ2 = (dY w3) gelu'(Z2)
#pragma unroll
  for (int j = 0; j < NH; ++j) acc[j] = floatx4{0.f, 0.f, 0.f, 0.f};
  wave_rows<NH, NO, true>(acc, sD, L, row0, sW3, LH, lane);
#pragma unroll
  for (int j = 0; j < NH; ++j) {
    const int h = 16 * j + r;
    const floatx4 dz = acc[j] * gg2[j];
    const float cs = col_sum(dz);
    if (g == 0) sred[wv][H + h] = cs;
    put_rows(sZ2 + h * L + row0 + 4 * g, dz);
  }
  __syncthreads();
  // dw2 = dZ2^T H1
  for (int tt = wv; tt < NH * NH; tt += 4) {
    const int ro = tt / NH, co = tt % NH;
    store_tile(P + a.o2, H, 16 * ro, 16 * co, H, tile_tn(sZ2, 16 * ro, sH1, 16 * co, lane),
               lane);
  }
  // dZ1 = (dZ2 w2) gelu'(Z1), into H2's array (last read by dw3, before the barrier)
#pragma unroll
  for (int j = 0; j < NH; ++j) acc[j] = floatx4{0.f, 0.f, 0.f, 0.f};
  wave_rows<NH, NH, true>(acc, sZ2, L, row0, sW2, LH, lane);
#pragma unroll
  for (int j = 0; j < NH; ++j) {
    const int h = 16 * j + r;
    const floatx4 dz = acc[j] * gg1[j];
    const float cs = col_sum(dz);
    if (g == 0) sred[wv][h] = cs;
    put_rows(sH2 + h * L + row0 + 4 * g, dz);
  }
  __syncthreads();
  // dw1 = dZ1^T X
  for (int tt = wv; tt < NH * NI; tt += 4) {
    const int ro = tt / NI, co = tt % NI;
    store_tile(P + a.o1, I, 16 * ro, 16 * co, H, tile_tn(sH2, 16 * ro, sX, 16 * co, lane), lane);
  }
  for (int e = t; e < 2 * H + OP; e += 256) {
    const float s = (sred[0][e] + sred[1][e]) + (sred[2][e] + sred[3][e]);
    if (e < H)
      P[a.ob1 + e] = s;
    else if (e < 2 * H)
      P[a.ob2 + e - H] = s;
    else if (e - 2 * H < a.O)
      P[a.ob3 + e - 2 * H] = s;
  }
  if (t == 0)
    a.loss_part[blockIdx.x] =
        (sred[0][NR - 1] + sred[1][NR - 1]) + (sred[2][NR - 1] + sred[3][NR - 1]);
}

// g[j] = sum over workspace rows k of part[k][j], in a fixed order: a block
// takes 32 columns, its 8 thread groups sum rows q, q + 8, ... (rounds of 16
// rows whose loads are all in flight together, into four interleaved partial
// sums), and the 8 group sums meet in LDS. With ad.w set, Adam on the sums. Block 0 adds
// the partial losses to *loss.
__global__ void __launch_bounds__(256) k_grad_reduce(const float* part, const float* loss_part,
                                                    int nwg, int np, float* g, float* loss,
                                                    AdamArgs ad) {
  __shared__ float red[8][32];
  const int c = threadIdx.x & 31, q = threadIdx.x >> 5;
  const int j = blockIdx.x * 32 + c;
  float s[4] = {0.f, 0.f, 0.f, 0.f};
  if (j < np)
    for (int k0 = q; k0 < nwg; k0 += 128) {
      float v[16];
#pragma unroll
      for (int u = 0; u < 16; ++u)
        v[u] = k0 + 8 * u < nwg ? part[(long long)(k0 + 8 * u) * np + j] : 0.f;
#pragma unroll
      for (int u = 0; u < 16; ++u) s[u & 3] += v[u];
    }
  red[q][c] = (s[0] + s[1]) + (s[2] + s[3]);
  __syncthreads();
  if (q == 0 && j < np) {
    const float gj = ((red[0][c] + red[1][c]) + (red[2][c] + red[3][c])) +
                     ((red[4][c] + red[5][c]) + (red[6][c] + red[7][c]));
    g[j] = gj;
    if (ad.w) adam_update(ad, j, gj);
  }
  if (blockIdx.x == 0 && threadIdx.x < 64 && loss) {   // one wave, butterfly in a fixed order
    float l = 0.f;
    for (int k = threadIdx.x; k < nwg; k += 64) l += loss_part[k];
    for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
    if (threadIdx.x == 0) *loss += l;
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
__global__ void __launch_bounds__(256, H == 64 && LDSW == 1 && OUTC > 0 ? 3 : 1)
    k_mlp_infer_t(InferArgs a) {
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
  constexpr bool PF = (H == 64 || H == 128) && INP <= 32;
  __shared__ float ybuf[4][OUTP * V];
  __shared__ __attribute__((aligned(16))) short wl[WLN];
  // with the weights in LDS the biases go there too: b1 | b2 | b3 (zero past
  // the outputs), read with ds_read, so no global load inside the voxel loop
  // but the (prefetched) inputs, the bounds and the stores
  constexpr int BLN = LDSW == 1 ? 2 * H + OUTP : 4;
  __shared__ __attribute__((aligned(16))) float bl[BLN];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int col = lane & 15;      // voxel within a tile
  const int kq = (lane >> 4) * 4; // k (or row) offset within a tile
  if constexpr (LDSW == 1) {
    stage_fragments<H, INP>(wl, a.w1, H);
    stage_fragments_k32<OUTP, H>(wl + W3OFF, a.w3, nout);
    for (int i = threadIdx.x; i < BLN; i += blockDim.x)
      bl[i] = i < H ? a.b1[i] : (i < 2 * H ? a.b2[i - H] : (i - 2 * H < nout ? a.b3[i - 2 * H] : 0.f));
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
  // raw inputs of one pass: measurement rows kq..kq+3 of each k tile and the
  // normaliser, for voxel base + 16 t + col
  auto load_x = [&](long long b, float (&xr)[VT][KT1][4], float (&sr)[VT]) {
#pragma unroll
    for (int t = 0; t < VT; ++t) {
      const long long v = b + t * 16 + col;
      const bool ok = v < a.N;
      sr[t] = ok ? a.data[(long long)a.norm_index * a.N + v] : 1.f;
#pragma unroll
      for (int kt = 0; kt < KT1; ++kt)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int i = kt * 16 + kq + q;
          xr[t][kt][q] = (ok && i < a.IN) ? a.data[(long long)i * a.N + v] : 0.f;
        }
    }
  };
  const long long first = ((long long)blockIdx.x * 4 + wave) * V, stride = nwaves * V;
  // PF: the next pass's inputs are loaded before this pass's layers run, so
  // their latency hides under them (widths 64 and 128 with up to 32 inputs,
  // whose register budgets keep their occupancy, 2 and 1 waves per SIMD, and
  // spill nothing with the copy)
  float xn[VT][KT1][4], sn[VT];
  if (PF && first < a.N) load_x(first, xn, sn);
  for (long long base = first; base < a.N; base += stride) {
    // layer-1 B operands: normalised inputs x^T (k = feature, n = voxel); the
    // normaliser of tile t stays in sv[t] for the epilogue
    float xr[VT][KT1][4], sr[VT];
    if constexpr (PF) {
#pragma unroll
      for (int t = 0; t < VT; ++t) {
        sr[t] = sn[t];
#pragma unroll
        for (int kt = 0; kt < KT1; ++kt)
#pragma unroll
          for (int q = 0; q < 4; ++q) xr[t][kt][q] = xn[t][kt][q];
      }
      if (base + stride < a.N) load_x(base + stride, xn, sn);
    } else {
      load_x(base, xr, sr);
    }
    half4 xb[KT1][VT];
    float sv[VT];
#pragma unroll
    for (int t = 0; t < VT; ++t) {
      float s = sr[t];
      if (!(s > 0.f)) s = 1.f;
      sv[t] = s;
      const float inv = 1.f / s;
#pragma unroll
      for (int kt = 0; kt < KT1; ++kt)
#pragma unroll
        for (int q = 0; q < 4; ++q) xb[kt][t][q] = (_Float16)(xr[t][kt][q] * inv);
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
      for (int q = 0; q < 4; ++q) bias[q] = LDSW == 1 ? bl[u * 16 + kq + q] : a.b1[u * 16 + kq + q];
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
        b3v[q] = LDSW == 1 ? bl[2 * H + j] : (j < nout ? a.b3[j] : 0.f);
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
        for (int q = 0; q < 4; ++q)
          bias[q] = LDSW == 1 ? bl[H + (2 * c + h) * 16 + kq + q] : a.b2[(2 * c + h) * 16 + kq + q];
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

template <int H, int I, int OP>
void launch_train(const TrainArgs& a, int B, hipStream_t s) {
  hipLaunchKernelGGL((k_train_step<H, I, OP>), dim3(B / TRB), dim3(256), 0, s, a);
}

AdamArgs adam_args(float* w, float* m, float* v, float lr, float b1, float b2, float eps, int step,
                   float wd) {
  return AdamArgs{w, m, v, lr, b1, b2, eps, 1.f - powf(b1, (float)step),
                  1.f - powf(b2, (float)step), wd};
}

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
  GemmArgs g{A, B, C, nullptr, bias, Z, Zin, lda, ldb, ldc, M, N, K, transA, transB, epi, 0};
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
                     out, nullptr);
  return check_launch();
}

extern "C" int pca_mlp_gemm_ordered(const float* A, int64_t lda, int32_t transA, const float* B,
                                    int64_t ldb, int32_t transB, float* C, int64_t ldc, int32_t M,
                                    int32_t N, int32_t K, int32_t splits, float* workspace,
                                    int64_t workspace_floats, void* stream) {
  if (!A || !B || !C || !workspace || M <= 0 || N <= 0 || K <= 0 || splits < 1)
    return pca_fail(PCA_ERR_ARG, "pca_mlp_gemm_ordered: bad argument");
  int kps = (K + splits - 1) / splits;
  kps = (kps + 15) / 16 * 16;
  const int zs = (K + kps - 1) / kps;
  if (workspace_floats < (int64_t)zs * M * N)
    return pca_fail(PCA_ERR_ARG, "pca_mlp_gemm_ordered: workspace below splits * M * N floats");
  GemmArgs g{A, B, C, workspace, nullptr, nullptr, nullptr, lda, ldb, ldc, M, N, K, transA,
             transB, EPI_STORE, kps};
  if (zs == 1) g.ws = nullptr;   // one split: stored directly
  const hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_gemm, dim3((M + 63) / 64, (N + 63) / 64, zs), dim3(256), 0, s, g);
  if (zs > 1)
    hipLaunchKernelGGL(k_sum_rows, dim3(grid_1d((long long)M * N)), dim3(256), 0, s, workspace,
                       zs, M, N, C, (long long)ldc, 0);
  return check_launch();
}

extern "C" int pca_mlp_colsum_ordered(const float* X, int64_t ld, int32_t M, int32_t N, float* out,
                                      float* workspace, int64_t workspace_floats, void* stream) {
  if (!X || !out || !workspace || M <= 0 || N <= 0)
    return pca_fail(PCA_ERR_ARG, "pca_mlp_colsum_ordered: bad argument");
  const int ny = std::min(1024, (M + 63) / 64);
  if (workspace_floats < (int64_t)ny * N)
    return pca_fail(PCA_ERR_ARG, "pca_mlp_colsum_ordered: workspace below chunks * N floats");
  const hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_colsum, dim3((N + 255) / 256, ny), dim3(256), 0, s, X, (long long)ld, M, N,
                     out, workspace);
  hipLaunchKernelGGL(k_sum_rows, dim3(grid_1d(N)), dim3(256), 0, s, workspace, ny, 1, N, out,
                     (long long)N, 1);
  return check_launch();
}

extern "C" int pca_mlp_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr,
                            float b1, float b2, float eps, int32_t step, float weight_decay,
                            void* stream) {
  if (!w || !g || !m || !v || n <= 0 || step < 1)
    return pca_fail(PCA_ERR_ARG, "pca_mlp_adam: bad argument");
  const AdamArgs ad = adam_args(w, m, v, lr, b1, b2, eps, step, weight_decay);
  hipLaunchKernelGGL(k_adam, dim3(grid_1d(n)), dim3(256), 0, (hipStream_t)stream, ad, g,
                     (long long)n);
  return check_launch();
}

extern "C" int pca_mlp_train_step(const float* X, int64_t ldx, const float* T, int32_t B,
                                  const float* w, int32_t hidden, int32_t in_padded, int32_t out,
                                  const int32_t* offsets, int32_t np, float* part,
                                  float* loss_part, void* stream) {
  if (!X || !T || !w || !offsets || !part || !loss_part || B <= 0 || B % TRB != 0 || out < 1 ||
      np <= 0 || ldx < B)
    return pca_fail(PCA_ERR_ARG, "pca_mlp_train_step: bad argument (B a positive multiple of 64)");
  const int OP = out <= 16 ? 16 : 32;
  const bool narrow = hidden == 32 || hidden == 64;
  if (!((narrow && in_padded == 16 && OP == 16) ||
        (hidden == 128 && (in_padded == 16 || in_padded == 32 || in_padded == 64) && out <= 32)))
    return pca_fail(PCA_ERR_ARG,
                    "pca_mlp_train_step: hidden 32 or 64 (in_padded 16, out <= 16) or 128 "
                    "(in_padded 16, 32 or 64, out <= 32)");
  const long long sizes[6] = {(long long)hidden * in_padded, hidden, (long long)hidden * hidden,
                              hidden, (long long)out * hidden, out};
  for (int k = 0; k < 6; ++k)
    if (offsets[k] < 0 || offsets[k] + sizes[k] > np)
      return pca_fail(PCA_ERR_ARG, "pca_mlp_train_step: parameter block outside np");
  TrainArgs a{X, T, w, part, loss_part, (long long)ldx, offsets[0], offsets[1], offsets[2],
              offsets[3], offsets[4], offsets[5], np, out, 1.f / ((float)B * (float)out)};
  const hipStream_t s = (hipStream_t)stream;
  if (hidden == 32)
    launch_train<32, 16, 16>(a, B, s);
  else if (hidden == 64)
    launch_train<64, 16, 16>(a, B, s);
  else if (in_padded == 16)
    OP == 16 ? launch_train<128, 16, 16>(a, B, s) : launch_train<128, 16, 32>(a, B, s);
  else if (in_padded == 32)
    OP == 16 ? launch_train<128, 32, 16>(a, B, s) : launch_train<128, 32, 32>(a, B, s);
  else
    OP == 16 ? launch_train<128, 64, 16>(a, B, s) : launch_train<128, 64, 32>(a, B, s);
  return check_launch();
}

extern "C" int pca_mlp_grad_reduce(const float* part, const float* loss_part, int32_t nwg,
                                   int32_t np, float* g, float* loss, float* w, float* m, float* v,
                                   float lr, float b1, float b2, float eps, int32_t step,
                                   float weight_decay, void* stream) {
  if (!part || !loss_part || !g || nwg <= 0 || np <= 0 || (w && (!m || !v || step < 1)))
    return pca_fail(PCA_ERR_ARG, "pca_mlp_grad_reduce: bad argument");
  const AdamArgs ad = w ? adam_args(w, m, v, lr, b1, b2, eps, step, weight_decay)
                        : AdamArgs{nullptr, nullptr, nullptr, 0.f, 0.f, 0.f, 0.f, 1.f, 1.f, 0.f};
  hipLaunchKernelGGL(k_grad_reduce, dim3((np + 31) / 32), dim3(256), 0, (hipStream_t)stream, part,
                     loss_part, nwg, np, g, loss, ad);
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
  // persistent grid, as many workgroups per CU as the registers let run at
  // once (width 32: 4 waves per SIMD, 64: 3 with the output count fixed at
  // compile time, 128: 2), reuses the staged copy
  const bool ldsw = !big && OUTP == 16 && INP <= 32;
  const int per_cu = H == 32 ? 4 : (H == 64 && OUT <= 4 && OUT >= 2 ? 3 : 2);
  // 256-unit networks stage W2 (128 KB) instead: one persistent workgroup per CU
  const int grid =
      (int)std::min<long long>((waves + 3) / 4, ldsw ? 256 * per_cu : (big ? 256 : 8192));
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
