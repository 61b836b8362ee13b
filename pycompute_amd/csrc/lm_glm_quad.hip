// Wide-model LM fit for generalised linear models with PCA_LPV = 4 (or 2)
// lanes per voxel (pca_lm_glm_quad): the same iteration as pca_lm_glm
// (lm_glm.hip; fghhl, GMW81 solve, NaN scrub, projected-gradient
// convergence, trial SSR with the GLM form of s^T h s, gain ratio, clamp;
// pycompute/cuda/lsqnonlin.py:820-947) laid out for gfx950's issue rules
// instead of one voxel per lane:
//
//  * One voxel per group of L = PCA_LPV lanes (a quad, or half a quad; 64/L
//    voxels per wavefront). The solve works in the reference's pivoted order
//    (the fixed rotation of lm_solver.hip's perm()): row a of the pivoted
//    system is parameter perm(n, a), and lane r of the group owns rows
//    a = L t + r (t < NT). The host permutes the table's columns to that
//    order.
//  * Lane r keeps, for each of its rows, the packed entries of columns
//    0 .. L t + L - 1 (NR registers; columns past the row are unused) and the
//    diagonal, gradient, damping sum, parameter and step of the row. Every
//    other value a row update needs lives in another lane of the same group
//    and is read with a quad-permute DPP broadcast, so all the FMAs of the
//    accumulation and of the factorisation take VGPR operands (a VALU
//    instruction reading an SGPR issues at about half rate on gfx950), and
//    the lane needs about 1/L of the registers: several wavefronts per SIMD
//    instead of one.
//  * Sums over a voxel's rows (eta = a0 + A.p, A.s, the convergence norm,
//    s.g) are partial sums per lane closed by a butterfly over the group,
//    which leaves the bitwise same value in all its lanes; every decision of
//    the iteration is therefore taken identically by the group.
//  * The GMW81 factorisation follows lm_solver.hip's gmw81_solve (column j:
//    c_aj = b_aj - sum_q L_jq c_aq for the rows below, theta, d_j, then the
//    diagonal update), with L_jq = c_jq / d_q formed when row j's entries are
//    broadcast; the substitutions use c_aj = L_aj d_j directly.
//
// phi is exp, or the identity with PCA_GLM_IDENTITY (then w = 1, J = A).
// Per-measurement table in global memory, (M, L, TQ) floats: [i][r][t] =
// A_{perm(L t + r)} of measurement i (0 past the last row), [i][r][NT] = a0,
// zero padding to TQ = a multiple of 4.
// The workgroup copies it to LDS once. Float only (fast builds).
// Concatenated after lm_solver.hip with PCA_NP and PCA_M defined.

#define PCA_QUAD_BLOCK 256
#ifndef PCA_LPV
#define PCA_LPV 4
#endif

namespace pca {
namespace quad {

constexpr int L = PCA_LPV;                              // lanes per voxel
static_assert(L == 2 || L == 4, "2 or 4 lanes per voxel");
constexpr int NP = PCA_NP, M = PCA_M;
constexpr int NT = (NP + L - 1) / L;                    // row groups per lane
constexpr int KT(int t) { return L * t + L < NP ? L * t + L : NP; }   // columns kept
constexpr int OFF(int t) { return t == 0 ? 0 : OFF(t - 1) + KT(t - 1); }
constexpr int NR = OFF(NT);                             // packed entries per lane
constexpr int TQ = (NT + 1 + 3) / 4 * 4;                // table floats per (i, r)
constexpr int VPB = PCA_QUAD_BLOCK / L;                 // voxels per workgroup
static_assert(NT <= 15, "at most 15 row groups per lane");

// DPP quad_perm control that sends lane s of every group of L lanes to all of
// its lanes (L = 4: [s,s,s,s]; L = 2: [s,s,2+s,2+s])
constexpr int bcast_ctrl(int s) {
  return L == 4 ? (s | (s << 2) | (s << 4) | (s << 6))
                : (s | (s << 2) | ((2 + s) << 4) | ((2 + s) << 6));
}
template <int S>
__device__ __forceinline__ float qb(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), bcast_ctrl(S), 0xF, 0xF,
                                                 false));
}
// sum and max over the group (butterfly: quad_perm [1,0,3,2], then for four
// lanes [2,3,0,1])
__device__ __forceinline__ float qsum(float x) {
  x += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0xB1, 0xF, 0xF, false));
  if constexpr (L == 4)
    x += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x4E, 0xF, 0xF, false));
  return x;
}
__device__ __forceinline__ float qmax(float x) {
  x = pca_fmax(x, __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0xB1, 0xF, 0xF, false)));
  if constexpr (L == 4)
    x = pca_fmax(x, __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x4E, 0xF, 0xF, false)));
  return x;
}
// v of lane s of the group in every lane of the group; s is a constant once
// the caller's loops are unrolled
__device__ __forceinline__ float qbs(float v, int s) {
  if constexpr (L == 4) {
    switch (s) {
      case 0: return qb<0>(v);
      case 1: return qb<1>(v);
      case 2: return qb<2>(v);
      default: return qb<3>(v);
    }
  } else {
    return s == 0 ? qb<0>(v) : qb<1>(v);
  }
}
// value of row k of an own-row array x (slot k / L in lane k % L) in every
// lane of the group
__device__ __forceinline__ float rowb(const float* x, int k) { return qbs(x[k / L], k % L); }

}  // namespace quad
}  // namespace pca

// occupancy hint (wavefronts per SIMD, PCA_QUAD_WAVES > 0): caps the registers
// per lane, trading spills in the solve for latency hiding in the loops
#if defined(PCA_QUAD_WAVES) && PCA_QUAD_WAVES > 0
#define PCA_QUAD_ATTR __attribute__((amdgpu_waves_per_eu(PCA_QUAD_WAVES, PCA_QUAD_WAVES)))
#else
#define PCA_QUAD_ATTR
#endif

extern "C" __global__ void __launch_bounds__(PCA_QUAD_BLOCK) PCA_QUAD_ATTR
pca_lm_glm_quad(pca_lm_args a, const float* __restrict__ tq_g) {
  using namespace pca;
  using namespace pca::quad;
  constexpr int TQ4 = TQ / 4;   // float4 per (i, r)
  __shared__ float4 tq[M * L * TQ4];
  for (int q = threadIdx.x; q < M * L * TQ4; q += PCA_QUAD_BLOCK)
    tq[q] = reinterpret_cast<const float4*>(tq_g)[q];
  __syncthreads();

  const long long N = a.N;
  const int r = threadIdx.x % L;
  const long long v = ((long long)blockIdx.x * PCA_QUAD_BLOCK + threadIdx.x) / L;
  if (v >= N) return;   // whole groups leave together; no barrier follows
  const Settings S = settings(a);
  const float floor_v = pca_vgpr((float)PCA_DAMP_FLOOR);

  // own rows: a = L t + r, parameter perm(NP, a)
  bool valid[NT];
  long long prow[NT];
  float P[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int ra = L * t + r;
    valid[t] = ra < NP;
    prow[t] = valid[t] ? (long long)perm(NP, ra) * N + v : 0;
    P[t] = valid[t] ? a.params[prow[t]] : 0.0f;
  }
  float lam = a.lam[v], f = a.f[v], f_first = f;
  int st = a.step_type[v];
  int cnt = 0;
  float G[NT], Sx[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    G[t] = 0;
    Sx[t] = 0;
  }
  const float* __restrict__ y = a.data + v;

  for (int it = 0; it < a.iters && st != 0; ++it) {
    ++cnt;
    // ---- fghhl: f, g, the damping sums and the packed Hessian rows --------
    float R[NR], D[NT], HD[NT];
#pragma unroll
    for (int e = 0; e < NR; ++e) R[e] = 0;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      G[t] = 0;
      D[t] = 0;
      HD[t] = 0;
    }
    f = 0;
    float ynext = y[0];
    for (int i = 0; i < M; ++i) {
      const float yi = ynext;
      if (i + 1 < M) ynext = y[(long long)(i + 1) * N];
      float al[TQ];
#pragma unroll
      for (int u = 0; u < TQ4; ++u) {
        const float4 qv = tq[(i * L + r) * TQ4 + u];
        al[4 * u] = qv.x;
        al[4 * u + 1] = qv.y;
        al[4 * u + 2] = qv.z;
        al[4 * u + 3] = qv.w;
      }
      float ep = 0;
#pragma unroll
      for (int t = 0; t < NT; ++t) ep = pca_fma(al[t], P[t], ep);
      const float eta = qsum(ep) + al[NT];
#ifdef PCA_GLM_IDENTITY
      // phi = identity: J = A, no second derivative
      const float rr = eta - yi;
      f = pca_fma(rr, rr, f);
      const float w = 1.0f, er = rr, ld = lam;
#else
      // phi = exp: J = e A, d2f = e A A^T, w = e^2 + e r
      const float e = PCA_EXP(eta);
      const float rr = e - yi;
      f = pca_fma(rr, rr, f);
      const float w = pca_fma(e, e, e * rr), er = e * rr, ld = lam * (e * e);
#endif
      float vi[NT];
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        G[t] = pca_fma(er, al[t], G[t]);
        HD[t] += pca_fmax(ld * (al[t] * al[t]), floor_v);
        vi[t] = w * al[t];
        D[t] = pca_fma(vi[t], al[t], D[t]);
      }
      // column k of every own row: A_k comes from lane k % L
#pragma unroll
      for (int k = 0; k < NP; ++k) {
        const float Ak = rowb(al, k);
#pragma unroll
        for (int t = 0; t < NT; ++t)
          if (k < KT(t)) R[OFF(t) + k] = pca_fma(vi[t], Ak, R[OFF(t) + k]);
      }
    }

    // ---- solve hl s = g (lm_solver.hip gmw81_solve, pivoted order) --------
    // scaling: sc = 1/sqrt|hl_aa|
    float sc[NT], Dd[NT], b[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const float hl = D[t] + HD[t];
      sc[t] = valid[t] ? pca_rcp_fast(pca_sqrt_fast(pca_abs(hl))) : 0.0f;
      Dd[t] = hl * (sc[t] * sc[t]);
      b[t] = G[t] * sc[t];
    }
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      const float sk = rowb(sc, k);
#pragma unroll
      for (int t = 0; t < NT; ++t)
        if (k < KT(t)) R[OFF(t) + k] *= sc[t] * sk;
    }
    // gamma = max |B_aa|, xi = max |B_ab| (b < a)
    float gamma = 0, xi = 0;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      gamma = pca_fmax(gamma, pca_abs(Dd[t]));
#pragma unroll
      for (int k = 0; k < KT(t); ++k) {
        const float m = pca_abs(R[OFF(t) + k]);
        xi = (k < L * t + r) ? pca_fmax(xi, m) : xi;
      }
    }
    gamma = qmax(gamma);
    xi = qmax(xi);
    if (NP > 1) xi *= 1.0f / pca_sqrt((float)(NP * NP - 1));
    float beta2 = PCA_EPS_GMW, delta = PCA_EPS_GMW;
    beta2 = pca_fmax(pca_fmax(beta2, gamma), xi);
    if (gamma + xi > 1.0f) delta *= (gamma + xi);
    const float rbeta2 = pca_rcp_fast(beta2);
    // GMW81, column j: c_aj = B_aj - sum_q L_jq c_aq for the own rows a > j
    // (row j's L_jq = c_jq / d_q broadcast from lane j % L), theta_j, d_j,
    // then B_aa -= c_aj^2 / d_j. Entries of rows at or above j in column j
    // are left with values nothing reads.
    float rd[NP];
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      const int tj = j / L;
      float c[NT];
#pragma unroll
      for (int t = 0; t < NT; ++t)
        if (L * t + L - 1 > j) c[t] = R[OFF(t) + j];   // groups with a row below row j
#pragma unroll
      for (int q = 0; q < j; ++q) {
        const float Lq = qbs(R[OFF(tj) + q], j % L) * rd[q];
#pragma unroll
        for (int t = 0; t < NT; ++t)
          if (L * t + L - 1 > j) c[t] = pca_fma(-Lq, R[OFF(t) + q], c[t]);
      }
      float th = 0;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        if (L * t + L - 1 <= j) continue;
        R[OFF(t) + j] = c[t];
        th = (L * t + r > j) ? pca_fmax(th, pca_abs(c[t])) : th;
      }
      th = qmax(th);
      th *= th;
      float d = pca_abs(rowb(Dd, j));
      if (d < delta) d = delta;
      const float t0 = th * rbeta2;
      if (d < t0) d = t0;
      rd[j] = pca_rcp_fast(d);
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        if (L * t + L - 1 <= j) continue;
        const float c = R[OFF(t) + j];
        Dd[t] = (L * t + r > j) ? pca_fma(-c * c, rd[j], Dd[t]) : Dd[t];
      }
    }

    // L D z = b with c_aj = L_aj d_j: z_j = (b_j - sum_q c_jq z_q) / d_j,
    // column by column (rows at or above j collect values nothing reads)
    float zown[NT];
    {
      float acc[NT];
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[t] = b[t];
#pragma unroll
      for (int j = 0; j < NP; ++j) {
        const float zj = rowb(acc, j) * rd[j];
        zown[j / L] = (r == j % L) ? zj : zown[j / L];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          if (L * t + L - 1 <= j) continue;
          acc[t] = pca_fma(-R[OFF(t) + j], zj, acc[t]);
        }
      }
    }
    // L^T w = z: w_i = z_i - (sum_{a > i} c_ai w_a) / d_i; w of rows not yet
    // solved is 0, so the rows at or above i add nothing
    float wown[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) wown[t] = 0;
#pragma unroll
    for (int i = NP - 1; i >= 0; --i) {
      float part = 0;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        if (L * t + L - 1 <= i) continue;
        part = pca_fma(R[OFF(t) + i], wown[t], part);
      }
      const float wi = pca_fma(-qsum(part), rd[i], rowb(zown, i));
      wown[i / L] = (r == i % L) ? wi : wown[i / L];
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const float s = wown[t] * sc[t];
      Sx[t] = pca_nonfinite(s) ? 0.0f : s;
    }

    // ---- projected-gradient convergence (lsqnonlin.py:691-729) ------------
    float lo[NT], hi[NT];
    float norm = 0, clampf = 0;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      lo[t] = valid[t] ? a.lb[prow[t]] : 0.0f;
      hi[t] = valid[t] ? a.ub[prow[t]] : 0.0f;
      float tt = P[t] - G[t];
      const bool above = tt > hi[t], below = tt < lo[t];
      clampf = (above || below) ? 1.0f : clampf;
      tt = above ? hi[t] : (below ? lo[t] : tt);
      tt = P[t] - tt;
      norm = pca_fma(tt, tt, norm);
    }
    norm = qsum(norm);
    clampf = qmax(clampf);
    if (it == 0) f_first = f;
    if (norm < S.tol * (1.0f + f) && ((st & 1) || clampf > 0.0f)) {
      st = 0;
      break;
    }

    // ---- trial point, its SSR and s^T h s = sum_i w_i (A_i . s)^2 ----------
    float ptp[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) ptp[t] = P[t] - Sx[t];
    float ftp = 0, shs = 0;
    ynext = y[0];
    for (int i = 0; i < M; ++i) {
      const float yi = ynext;
      if (i + 1 < M) ynext = y[(long long)(i + 1) * N];
      float al[TQ];
#pragma unroll
      for (int u = 0; u < TQ4; ++u) {
        const float4 qv = tq[(i * L + r) * TQ4 + u];
        al[4 * u] = qv.x;
        al[4 * u + 1] = qv.y;
        al[4 * u + 2] = qv.z;
        al[4 * u + 3] = qv.w;
      }
      float et = 0, ep = 0, as = 0;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        et = pca_fma(al[t], ptp[t], et);
        ep = pca_fma(al[t], P[t], ep);
        as = pca_fma(al[t], Sx[t], as);
      }
      et = qsum(et) + al[NT];
      ep = qsum(ep) + al[NT];
      as = qsum(as);
#ifdef PCA_GLM_IDENTITY
      const float rt = et - yi;
      ftp = pca_fma(rt, rt, ftp);
      shs = pca_fma(as, as, shs);
      (void)ep;
#else
      const float rt = PCA_EXP(et) - yi;
      ftp = pca_fma(rt, rt, ftp);
      const float e = PCA_EXP(ep);
      shs = pca_fma(pca_fma(e, e, e * (e - yi)), as * as, shs);
#endif
    }

    // ---- gain ratio and step type (lsqnonlin.py:432-498) ------------------
    float sg = 0;
#pragma unroll
    for (int t = 0; t < NT; ++t) sg = pca_fma(Sx[t], G[t], sg);
    const float actual = 0.5f * (f - ftp);
    const float pred = -0.5f * shs + qsum(sg);
    const float rho = actual / pred;
    if (rho > S.mu && actual > 0) {
#pragma unroll
      for (int t = 0; t < NT; ++t) P[t] = ptp[t];
      if (rho > S.eta) {
        lam *= S.acc;
        st = 1;
      } else {
        st = 2;
      }
    } else {
      lam *= S.dec;
      st = 4;
    }
    if (pred < 0) {
      lam *= S.dec;
      st |= 8;
    }
    // clamp (lsqnonlin.py:584-607)
#pragma unroll
    for (int t = 0; t < NT; ++t) P[t] = P[t] > hi[t] ? hi[t] : (P[t] < lo[t] ? lo[t] : P[t]);
  }

#pragma unroll
  for (int t = 0; t < NT; ++t)
    if (valid[t]) a.params[prow[t]] = P[t];
  if (r == 0) {
    a.lam[v] = lam;
    a.step_type[v] = (signed char)st;
    a.f[v] = f;
    if (a.first_f) a.first_f[v] = f_first;
    if (a.last_f) a.last_f[v] = f;
    if (a.n_iter) a.n_iter[v] += cnt;
  }
  if (cnt > 0) {
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      if (!valid[t]) continue;
      if (a.g_out) a.g_out[prow[t]] = G[t];
      if (a.step_out) a.step_out[prow[t]] = Sx[t];
    }
  }
}
