// Wide GLM fit with the packed Hessian accumulated on the matrix cores
// (float only; concatenated after lm_glm.hip, whose namespace pca::glm gives
// NP, NH, M, TS, eta_of and phi).
//
// The iteration is pca_lm_glm's (pycompute/cuda/lsqnonlin.py:820-947: fghhl,
// GMW81 solve, NaN scrub, projected-gradient convergence, trial SSR, gain
// ratio, clamp). For a GLM the Hessian update of measurement i is
//     h_e += w_i * (A_ij A_ik)       for every packed entry e = (j, k),
// a rank-one update shared by all voxels up to the per-voxel weight w_i.
// v_mfma_f32_4x4x1f32 runs 16 independent 4x4x1 products per wavefront:
// lane l = 4b + c supplies A[b][l % 4] and B[b][c] and holds D[b][0..3][c].
// With B[b][c] = w_i of lane l's own voxel and A[b][r] = A_ij A_ik of entry
// 4q + r (the same for every block), lane l's D accumulates entries 4q..4q+3
// of its own voxel: the Hessian builds up one voxel per lane, in registers,
// in the reference's measurement order, four entries per instruction. No
// data moves between lanes, so the per-lane GMW81 solve follows unchanged.
//
//  * The entry products (consts only) are evaluated once per protocol on the
//    host and staged in LDS by each workgroup: t2[i][r][q] = A_ij A_ik of
//    entry 4q + r, rows r padded so the four rows a wavefront reads start in
//    different banks.
//  * f, g and the damping diagonal sum_i max(lam J_j^2, 1e-4) stay on the
//    VALU, as in pca_lm_glm. The accumulators are split into two passes
//    over the measurements (half of the accumulation registers each, which
//    leaves the other half as spill space for the solve); the second pass
//    re-forms w_i from eta_i. Each lane's samples are staged in LDS once per
//    launch, so the measurement loops read no global memory; the trial pass
//    forms eta(p) = eta(p - s) + A_i . s for s^T h s = sum_i w_i (A_i . s)^2.
//  * The matrix instructions are issued from inline assembly with each
//    accumulator tied to its input (see mfma4): with the builtin the
//    compiler copies the loop-carried accumulators between registers.
//  * Four wavefronts per workgroup, one per SIMD (the Hessian and its solve
//    take all 512 registers of a lane; the tables fill the LDS). Every lane runs every iteration with
//    the whole wavefront active -- the matrix instructions read the other
//    lanes' table fragments -- and a voxel that has stopped, or lies past the
//    end, keeps its state through selects; a stopped voxel's p and lambda
//    never change, so its recomputed f, g and s equal the stored ones.

#define PCA_MFMA_WG 256

namespace pca {
namespace glmm {
constexpr int NQ = (glm::NH + 3) / 4;      // groups of four packed entries
constexpr int NQP = (NQ + 3) / 4 * 4;      // padded to whole 16-byte reads
constexpr int NQ1 = (NQP / 4 + 1) / 2 * 4; // groups of the first pass (whole reads)
constexpr int NQ2 = NQP - NQ1;             // groups of the second pass
constexpr int RS = NQP + 4;                // row stride: rows start 4 banks apart
typedef float f4 __attribute__((ext_vector_type(4)));

// Four rank-one updates D_q += A_q B (v_mfma_f32_4x4x1f32) on accumulators
// held in fixed accumulation registers (each tied to its own input, so the
// compiler neither copies nor rotates them between measurements). The
// leading s_nop covers the wait states between a VALU write of B or of an
// accumulator (its zeroing) and the matrix instruction reading it; fragment
// loads from LDS are waited for by the compiler (the operands are inputs).
__device__ __forceinline__ void mfma4(f4& c0, f4& c1, f4& c2, f4& c3, const f4 a, float b) {
  asm volatile(
      "s_nop 2\n\t"
      "v_mfma_f32_4x4x1_16b_f32 %0, %4, %8, %0\n\t"
      "v_mfma_f32_4x4x1_16b_f32 %1, %5, %8, %1\n\t"
      "v_mfma_f32_4x4x1_16b_f32 %2, %6, %8, %2\n\t"
      "v_mfma_f32_4x4x1_16b_f32 %3, %7, %8, %3"
      : "+a"(c0), "+a"(c1), "+a"(c2), "+a"(c3)
      : "v"(a.x), "v"(a.y), "v"(a.z), "v"(a.w), "v"(b));
}
// Wait states between the last matrix instruction and the first read of its
// results by the VALU (24, more than the latency of any f32 MFMA): ordered
// after every mfma4 (volatile) and before every read of the four
// accumulators passed (their values come out of this statement).
__device__ __forceinline__ void mfma_drain(f4& c0, f4& c1, f4& c2, f4& c3) {
  asm volatile("s_nop 7\n\ts_nop 7\n\ts_nop 7" : "+a"(c0), "+a"(c1), "+a"(c2), "+a"(c3));
}
__device__ __forceinline__ void mfma_order(f4& c0, f4& c1, f4& c2, f4& c3) {
  asm volatile("" : "+a"(c0), "+a"(c1), "+a"(c2), "+a"(c3));
}
}  // namespace glmm
}  // namespace pca

extern "C" __global__ void __launch_bounds__(PCA_MFMA_WG)
pca_lm_glm_mfma(pca_lm_args a, const real* __restrict__ tab, const float* __restrict__ tab2) {
  static_assert(sizeof(real) == 4, "the matrix-core Hessian is float only");
  using namespace pca;
  using namespace pca::glm;
  using glmm::NQ;
  using glmm::NQP;
  using glmm::NQ1;
  using glmm::NQ2;
  using glmm::RS;
  __shared__ glmm::f4 t2v[M * 4 * RS / 4];
  __shared__ float ypark[PCA_MFMA_WG / 64][M][64];   // the samples of every voxel
  __shared__ float gpark[PCA_MFMA_WG / 64][64][NP + 1];   // g over the solve (odd stride)
  {
    const glmm::f4* __restrict__ src = reinterpret_cast<const glmm::f4*>(tab2);
    for (int x = threadIdx.x; x < M * RS; x += PCA_MFMA_WG) t2v[x] = src[x];
  }
  __syncthreads();
  const float* t2 = reinterpret_cast<const float*>(t2v);
  const long long N = a.N;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const long long v0 = (long long)blockIdx.x * PCA_MFMA_WG + threadIdx.x;
  const bool live = v0 < N;
  if (__ballot(live) == 0) return;   // a wavefront wholly past the end; no barrier follows
  const long long v = live ? v0 : N - 1;
  const Settings S = settings(a);
  // p, the bounds and the g/s outputs live in the caller's arrays between
  // uses (p is written back after every iteration, then re-read after a
  // compiler memory barrier): only h, the damped diagonal and g are live
  // through the solve. Row j of a (rows, N) array is addressed by the 32-bit
  // element offset j N + v from the array's uniform base (the host checks
  // rows * N < 2^30); the lane's offset passes through an empty asm at the
  // top of every iteration, so the per-row addresses are formed where they
  // are used instead of being hoisted and held in registers.
  const unsigned Nu = (unsigned)N;
  unsigned vu = (unsigned)v;
  real lam = a.lam[v], f = a.f[v], f_first = f;
  int st = live ? (int)a.step_type[v] : 0;
  int cnt = 0;
  {
    // the samples stay in LDS for the whole fit: the measurement loops read
    // no global memory
    const real* __restrict__ y = a.data + v;
    real yb[M];
#pragma unroll
    for (int i = 0; i < M; ++i) yb[i] = y[(long long)i * N];
#pragma unroll
    for (int i = 0; i < M; ++i) ypark[wv][i][lane] = yb[i];
  }
  const float* __restrict__ frag = t2 + (lane & 3) * RS;   // row lane % 4 of the table

  for (int it = 0; it < a.iters; ++it) {
    const bool active = st != 0;
    if (__ballot(active) == 0) break;
    cnt += active ? 1 : 0;
    asm volatile("" : "+v"(vu));
    auto row = [&](int j) { return (unsigned)j * Nu + vu; };
    real p[NP];
#pragma unroll
    for (int j = 0; j < NP; ++j) p[j] = a.params[row(j)];

    // fghhl (lsqnonlin.py:206-346): f, g, damping on the VALU, h on the
    // matrix cores in two passes over the measurements, groups 0..NQ1-1 and
    // then the rest (groups past NQ update padding entries that are never
    // read): each pass holds half of the accumulation registers, which
    // leaves the other half to the solve. The first pass parks w_i in LDS;
    // the second reads it back.
    real h[NH];
    real fn = 0, gn[NP], hd[NP];
#pragma unroll
    for (int j = 0; j < NP; ++j) gn[j] = hd[j] = 0;
    {
      glmm::f4 acc[NQ1];
#pragma unroll
      for (int q = 0; q < NQ1; ++q) acc[q] = glmm::f4{0.f, 0.f, 0.f, 0.f};
      for (int i = 0; i < M; ++i) {
        const real yi = ypark[wv][i][lane];
        const real* __restrict__ A = tab + i * TS;
        const real e = phi(eta_of(A, p));
#ifdef PCA_GLM_IDENTITY
        const real d1 = 1, r = e - yi, w = 1;
#else
        const real d1 = e, r = e - yi, w = e * e + e * r;
#endif
        fn += r * r;
        const real dr = d1 * r, ld = lam * (d1 * d1);
#pragma unroll
        for (int j = 0; j < NP; ++j) {
          gn[j] += dr * A[j];
          hd[j] += pca_fmax(ld * A[NP + j], (real)PCA_DAMP_FLOOR);
        }
        const glmm::f4* __restrict__ fi = reinterpret_cast<const glmm::f4*>(frag + i * 4 * RS);
#pragma unroll
        for (int q = 0; q < NQ1; q += 4)
          glmm::mfma4(acc[q], acc[q + 1], acc[q + 2], acc[q + 3], fi[q / 4], w);
      }
      glmm::mfma_drain(acc[0], acc[1], acc[2], acc[3]);
#pragma unroll
      for (int q = 4; q < NQ1; q += 4) glmm::mfma_order(acc[q], acc[q + 1], acc[q + 2], acc[q + 3]);
#pragma unroll
      for (int e = 0; e < 4 * NQ1 && e < NH; ++e) h[e] = acc[e / 4][e % 4];
    }
    if constexpr (NQ2 > 0) {
      glmm::f4 acc[NQ2 > 0 ? NQ2 : 1];
#pragma unroll
      for (int q = 0; q < NQ2; ++q) acc[q] = glmm::f4{0.f, 0.f, 0.f, 0.f};
      for (int i = 0; i < M; ++i) {
#ifdef PCA_GLM_IDENTITY
        const real w = 1;
#else
        const real* __restrict__ A = tab + i * TS;
        const real e = phi(eta_of(A, p)), yi = ypark[wv][i][lane];
        const real w = e * e + e * (e - yi);
#endif
        const glmm::f4* __restrict__ fi =
            reinterpret_cast<const glmm::f4*>(frag + i * 4 * RS + NQ1);
#pragma unroll
        for (int q = 0; q < NQ2; q += 4)
          glmm::mfma4(acc[q], acc[q + 1], acc[q + 2], acc[q + 3], fi[q / 4], w);
      }
      glmm::mfma_drain(acc[0], acc[1], acc[2], acc[3]);
#pragma unroll
      for (int q = 4; q < NQ2; q += 4) glmm::mfma_order(acc[q], acc[q + 1], acc[q + 2], acc[q + 3]);
#pragma unroll
      for (int e = 4 * NQ1; e < NH; ++e) h[e] = acc[(e - 4 * NQ1) / 4][e % 4];
    }
    // projected-gradient convergence (lsqnonlin.py:691-729); it reads p, g,
    // the bounds and f only, so it is decided before the solve
    bool clamped = false;
    real norm = 0;
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      const real lo = a.lb[row(j)], hi = a.ub[row(j)];
      real t = p[j] - gn[j];
      const bool above = t > hi, below = t < lo;
      clamped = clamped | above | below;
      t = above ? hi : (below ? lo : t);
      t = p[j] - t;
      norm += t * t;
    }
    const bool conv = norm < S.tol * ((real)1 + fn) && ((st & 1) || clamped);
    const bool go = active && !conv;
    if (it == 0 && active) f_first = fn;
    f = active ? fn : f;
    float* gp = gpark[wv][lane];
#pragma unroll
    for (int j = 0; j < NP; ++j) gp[j] = gn[j];
    if (live && active && a.g_out) {   // g of the last executed iteration
#pragma unroll
      for (int j = 0; j < NP; ++j) a.g_out[row(j)] = gn[j];
    }

    // g enters the solve from LDS and p comes back from memory afterwards:
    // the barriers keep the compiler from carrying either through the solve
    // in registers
    real sn[NP];
    {
      real hld[NP];
#pragma unroll
      for (int j = 0; j < NP; ++j) hld[j] = h[tri(j, j)] + hd[j];
      asm volatile("" ::: "memory");
      gmw81_solve<NP>(h, hld, gp, sn);
    }
#pragma unroll
    for (int j = 0; j < NP; ++j)
      if (pca_nonfinite(sn[j])) sn[j] = 0;
    asm volatile("" ::: "memory");
    real sg = 0;
#pragma unroll
    for (int j = 0; j < NP; ++j) sg += sn[j] * gp[j];
    if (live && active && a.step_out) {
#pragma unroll
      for (int j = 0; j < NP; ++j) a.step_out[row(j)] = sn[j];
    }

    // trial point and its SSR (unclamped, lsqnonlin.py:902-911), and
    // s^T h s = sum_i w_i (A_i . s)^2 for the gain ratio
    real ptp[NP];
#pragma unroll
    for (int j = 0; j < NP; ++j) ptp[j] = a.params[row(j)] - sn[j];
    real ftp = 0, shs = 0;
    {
      for (int i = 0; i < M; ++i) {
        const real yi = ypark[wv][i][lane];
        const real* __restrict__ A = tab + i * TS;
        real as = 0;
#pragma unroll
        for (int k = 0; k < NP; ++k) as += A[k] * sn[k];
        const real eta_t = eta_of(A, ptp);
        const real rt = phi(eta_t) - yi;
        ftp += rt * rt;
#ifdef PCA_GLM_IDENTITY
        shs += as * as;
#else
        const real e = phi(eta_t + as);   // phi(eta(p)) up to rounding
        shs += (e * e + e * (e - yi)) * (as * as);
#endif
      }
    }

    // gain ratio, step type and clamp (lsqnonlin.py:432-498, 584-607), kept
    // only by voxels that iterate on
    const real actual = (real)0.5 * (fn - ftp);
    const real pred = (real)-0.5 * shs + sg;
    const real rho = actual / pred;
    const bool accept = rho > S.mu && actual > 0;
    const bool good = accept && rho > S.eta;
    const bool nondesc = pred < 0;
    real nl = good ? lam * S.acc : (accept ? lam : lam * S.dec);
    nl = nondesc ? nl * S.dec : nl;
    lam = go ? nl : lam;
    st = go ? ((good ? 1 : (accept ? 2 : 4)) | (nondesc ? 8 : 0)) : 0;
    if (live && go) {
#pragma unroll
      for (int j = 0; j < NP; ++j) {
        const real lo = a.lb[row(j)], hi = a.ub[row(j)], pj = a.params[row(j)];
        const real q = accept ? pj - sn[j] : pj;
        a.params[row(j)] = q > hi ? hi : (q < lo ? lo : q);
      }
    }
  }

  if (!live) return;
  a.lam[v] = lam;
  a.step_type[v] = (signed char)st;
  a.f[v] = f;
  if (a.first_f) a.first_f[v] = f_first;
  if (a.last_f) a.last_f[v] = f;
  if (a.n_iter) a.n_iter[v] += cnt;
}
