// hk_step.hip — device-resident SQP row steps, simplex projections, EM normalisation and the
// Zhou-Alexander-Lange quasi-Newton extrapolation (gfx950, FP64).
//
// Reference behaviour (paths relative to the HaploADMIXTURE.jl repository):
//   update_q! QP stage      src/algorithms_inner.jl:135-168, step + projection :200-207
//   update_p! QP stage      src/algorithms_inner.jl:259-312 (block-diagonal assembly :272-283,
//                           default penalty rho on every (l, l2) block diagonal :295-302), step :344-351
//   em! normalisation       src/algorithms_inner.jl:66-76
//   project_p!              src/utils.jl:20-28 (per (k, j) 4-vector, pseudocount 1e-5)
//   ZAL update_UV!/QN!      src/algorithms_outer.jl:128-133
// The row QPs (OpenADMIXTURE.create_tableau!/quadratic_program!) are strictly convex box QPs with
// optional group-sum equalities; their optimum is unique.  Here each QP is solved by a primal
// active-set method on ONE wavefront: lane r owns row r of the (masked) curvature matrix in
// registers, pivots and columns are broadcast with v_readlane, and the two triangular solves use
// one LDS transpose.  Every reduction runs in a fixed order; the only atomic is the failure counter.
#include <mutex>

#include "hk_device.h"
#include "hk_internal.h"
#include "hk_qp.h"

#include <type_traits>

namespace hk {

// ------------------------------------------------------------------------------------------------
// update_q! row step: one wavefront per individual.
// res_q[i][NT]: Xtz (K) then packed lower XtX (K(K+1)/2), as produced by the pass (Xtz not negated).
// ------------------------------------------------------------------------------------------------
template <int K, bool PEN>
__global__ __launch_bounds__(64) void k_q_step(int64_t I, const double* __restrict__ res, const double* __restrict__ q,
                                               double* __restrict__ q_next, double rho, double eps,
                                               int* __restrict__ fails, double* __restrict__ dq, const int* __restrict__ gate) {
  if (gate_off(gate)) return;
  constexpr bool penalty = PEN;
  constexpr int NT = K + K * (K + 1) / 2;
  __shared__ double lds[K * K > 64 ? K * K : 64];
  const int64_t i = blockIdx.x;
  const int lane = threadIdx.x;
  const double* r = res + i * NT;
  const int rr = lane < K ? lane : 0;
  // Without the penalty the sum(dq) = 0 equality makes sigma 11' vanish on the feasible set; adding
  // it (sigma = largest diagonal) keeps the optimum and makes the curvature definite on free sets.
  const double sigma = wave_max(lane < K ? r[K + rr * (rr + 1) / 2 + rr] : 0.0);
  const double add = penalty ? rho : sigma;
  double Hrow[K];
#pragma unroll
  for (int c = 0; c < K; ++c) {
    const int a = rr > c ? rr : c, b = rr > c ? c : rr;
    Hrow[c] = (lane < K) ? r[K + a * (a + 1) / 2 + b] + add : 0.0;
  }
  const double qi = lane < K ? q[rr + (int64_t)K * i] : 0.0;
  const double g = lane < K ? -r[rr] : 0.0;
  const double lo = lane < K ? 0.0 - qi : 0.0, hi = lane < K ? 1.0 - qi : 0.0;
  double x = 0.0;
  int it;
  if constexpr (PEN) it = wave_qp<K, 0>(Hrow, g, lo, hi, 0, x, lds);
  else it = wave_qp<K, 1>(Hrow, g, lo, hi, 0, x, lds);
  if (it < 0) x = 0.0;
  if (it < 0 && lane == 0) atomicAdd(fails, 1);
  // iteration limit reached: the last (feasible) iterate is applied and counted in fails[1]
  if (it >= 4 * K + 20 && lane == 0) atomicAdd(fails + 1, 1);
  if (dq && lane < K) dq[lane + (int64_t)K * i] = x;  // raw step, for step halving
  // q_next = q + dq, then project_q! (lane 0 gathers the K values).
  double v[K];
#pragma unroll
  for (int c = 0; c < K; ++c) v[c] = rl(qi + x, c);
  if (lane == 0) {
    proj_simplex<K>(v, eps);
#pragma unroll
    for (int c = 0; c < K; ++c) q_next[c + (int64_t)K * i] = v[c];
  }
}

// Dense solves of the blocks listed in redo[1 .. redo[0]] (grid-stride over the list): the late list.
template <int K, bool PEN>
__global__ __launch_bounds__(64) void k_p_step_list(const double* __restrict__ res, const double* __restrict__ p,
                                                    double* __restrict__ p_next, double rho, double eps,
                                                    int* __restrict__ fails, double* __restrict__ dp,
                                                    const int* __restrict__ redo, const int* __restrict__ gate) {
  if (gate_off(gate)) return;
  __shared__ double lds[4 * K * 4 * K];
  const int n = redo[0];
  for (int idx = blockIdx.x; idx < n; idx += gridDim.x)
    p_step_dense<K, PEN>(redo[1 + idx], res, p, p_next, rho, eps, fails, dp, lds);
}

// ------------------------------------------------------------------------------------------------
// update_p! row step, four lanes per QP (block-structured active set)
//
// The P-side QP of block j has 4K variables x[l][k] (haplotype l, population k) and curvature
//     H = blockdiag(A_0, .., A_3) + a (J_4 (x) I_K),     A_l = XtX_p[:, :, 4j + l],
// a = rho in penalty mode (no equalities) or sigma = the largest diagonal in equality mode
// (sum_l x[l][k] = 0 for every k; the a-term vanishes on the feasible set).  Lane l of a quad owns
// block l.  Every active-set step of wave_qp is reproduced with the same ratio test, tie-breaking
// and multiplier rule, but the Newton system on the free set F is solved through the K-dimensional
// coupling: with B_F = blockdiag(A_l[F_l, F_l]), C_l = (A_l[F_l, F_l])^-1 zero-padded, y_l = C_l r_l,
//     penalty:   (B_F + rho E E')^-1 r = y - C t,     (I / rho + sum_l C_l) t = sum_l y_l   (Woodbury)
//     equality:  d = -(y + C lam),                    (sum_l C_l) lam = -sum_l y_l          (Schur)
// (sums over the quad by DPP).  A quad with a free block that is not positive definite (a haplotype
// compatible with too few individuals) is solved by the one-wave dense solver instead: the P-side
// finishing kernel lists those it can tell from the first active-set iteration (flags, early list, solved
// by the first workgroups of the same launch); any other lands on the late list, solved after it.
// ------------------------------------------------------------------------------------------------
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
  return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, false);
}
__device__ __forceinline__ double quad_max(double v) {
  v = fmax(v, dpp_f<0xB1>(v));
  return fmax(v, dpp_f<0x4E>(v));
}
// (value, index) extremum over a quad with ties to the smaller index; identical on the four lanes.
template <bool MIN>
__device__ __forceinline__ void quad_arg(double& v, int& idx) {
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const double ov = s == 0 ? dpp_f<0xB1>(v) : dpp_f<0x4E>(v);
    const int oi = s == 0 ? dpp_i<0xB1>(idx) : dpp_i<0x4E>(idx);
    if ((MIN ? ov < v : ov > v) || (ov == v && oi < idx)) {
      v = ov;
      idx = oi;
    }
  }
}

// Packed lower-triangle index of the symmetric entry (a, b).
__host__ __device__ constexpr int pk(int a, int b) { return a >= b ? a * (a + 1) / 2 + b : b * (b + 1) / 2 + a; }

// Cholesky solve of the K x K system S u = b (S symmetric positive definite, lower triangle used).
template <int K>
__device__ __forceinline__ void small_spd_solve(const double (&S)[K][K], double (&u)[K]) {
  double Lc[K][K], di[K];
#pragma unroll
  for (int c = 0; c < K; ++c) {
    double s = S[c][c];
#pragma unroll
    for (int t = 0; t < c; ++t) s -= Lc[c][t] * Lc[c][t];
    s = s > 0.0 ? s : 1e-300;
    const double sq = sqrt(s);
    di[c] = 1.0 / sq;
    Lc[c][c] = sq;
#pragma unroll
    for (int r = c + 1; r < K; ++r) {
      double v = S[r][c];
#pragma unroll
      for (int t = 0; t < c; ++t) v -= Lc[r][t] * Lc[c][t];
      Lc[r][c] = v * di[c];
    }
  }
#pragma unroll
  for (int r = 0; r < K; ++r) {
    double v = u[r];
#pragma unroll
    for (int t = 0; t < r; ++t) v -= Lc[r][t] * u[t];
    u[r] = v * di[r];
  }
#pragma unroll
  for (int r = K - 1; r >= 0; --r) {
    double v = u[r];
#pragma unroll
    for (int t = r + 1; t < K; ++t) v -= Lc[t][r] * u[t];
    u[r] = v * di[r];
  }
}

// Each 4-lane group (slot) solves the quads s, s + S, s + 2S, ... one after another (S = slots in the
// grid, which is capped at what the device holds at once), so lanes whose QP converged in a few
// active-set iterations go on to their next quad instead of idling until the slowest quad of their
// wavefront is done.  Every quad's result depends only on its own data, not on the slot or the order.
template <int K, bool PEN, bool MULTI>
__device__ __forceinline__ void p_step4_body(int64_t J, const double* __restrict__ res, const double* __restrict__ p,
                                             double* __restrict__ p_next, double rho, double eps,
                                             double* __restrict__ dp, const unsigned char* __restrict__ flags,
                                             const int* __restrict__ dense_list, int ndb, int* __restrict__ fails,
                                             int* __restrict__ late) {
  constexpr int NT = K + K * (K + 1) / 2;
  constexpr int maxit = 4 * (4 * K) + 20;  // as wave_qp with N = 4K
  constexpr int NKK = K * (K + 1) / 2;
  // The first ndb workgroups solve the early dense list, one QP per wavefront at a time, while the others
  // run the four-lane QPs: both start together in one launch.
  __shared__ double dlds[4][4 * K * 4 * K];
  if ((int)blockIdx.x < ndb) {
    const int w = threadIdx.x >> 6;
    const int n = dense_list[0];
    for (int idx = blockIdx.x * 4 + w; idx < n; idx += ndb * 4)
      p_step_dense<K, PEN>(dense_list[1 + idx], res, p, p_next, rho, eps, fails, dp, dlds[w]);
    return;
  }
  const int64_t blk = (int64_t)blockIdx.x - ndb;
  const int l = threadIdx.x & 3;
  const int64_t nslot = ((int64_t)(gridDim.x - ndb) * blockDim.x) >> 2;
  int64_t j = (blk * blockDim.x + threadIdx.x) >> 2;
  bool valid = j < J;
  double A[NKK];  // packed lower triangle, A[pk(k, k2)]
  double g[K], lo[K], hi[K], pv[K], x[K];
  int st[K];      // 0 free, -1 at lower, +1 at upper
  double lam[K];  // equality-mode multipliers of the last Newton step
  double a = 0.0, tol = 0.0, piv_min = 0.0;
  int it = 0;
  bool done = false, dense = false, skip = false;
  // quad-uniform: all four lanes of a quad take the same branches
  auto load = [&]() __attribute__((always_inline)) {
    skip = flags[j] != 0;  // on the early dense list: solved by the dense kernel
    dense = false;
    it = 0;
    done = skip;
    if (skip) return;
    const double* r = res + (4 * j + l) * NT;
    #pragma unroll
    for (int t = 0; t < NKK; ++t) A[t] = r[K + t];
    double gs = 0.0, bm = 0.0, dg = 0.0;
    #pragma unroll
    for (int k = 0; k < K; ++k) {
      pv[k] = p[k + (int64_t)K * (4 * j + l)];
      g[k] = -r[k];
      lo[k] = 0.0 - pv[k];
      hi[k] = 1.0 - pv[k];
      x[k] = 0.0;
      lam[k] = 0.0;
      st[k] = lo[k] >= 0.0 ? -1 : (hi[k] <= 0.0 ? 1 : 0);
      gs = fmax(gs, fabs(g[k]));
      bm = fmax(bm, fmax(fabs(lo[k]), fabs(hi[k])));
      dg = fmax(dg, A[pk(k, k)]);
    }
    a = PEN ? rho : quad_max(dg);
    double hm = a;  // the entries a of the off-diagonal blocks
    #pragma unroll
    for (int k = 0; k < K; ++k)
      #pragma unroll
      for (int k2 = 0; k2 <= k; ++k2) hm = fmax(hm, fabs(A[pk(k, k2)] + (k == k2 ? a : 0.0)));
    hm = quad_max(hm);
    gs = quad_max(gs);
    bm = quad_max(bm);
    tol = 1e-11 * (gs + hm * bm) + 1e-300;
    // pivots below this share of the block's largest diagonal send the QP to the dense solver
    piv_min = 1e-6 * dg + 1e-300;
  };
  // one active-set iteration (sets done when the QP is solved or handed to the dense solver)
  auto iterate = [&]() __attribute__((always_inline)) {
      bool fr[K];
      #pragma unroll
      for (int k = 0; k < K; ++k) fr[k] = st[k] == 0;
      // masked block Cholesky: A on F_l x F_l, identity elsewhere
      double L[K][K], di[K];
      bool pd = true;
      #pragma unroll
      for (int c = 0; c < K; ++c) {
        double s = fr[c] ? A[pk(c, c)] : 1.0;
        #pragma unroll
        for (int t = 0; t < c; ++t) s -= L[c][t] * L[c][t];
        const bool ok = !fr[c] || s > piv_min;
        pd = pd && ok;
        s = ok ? s : 1.0;
        const double sq = sqrt(s);
        di[c] = 1.0 / sq;
        L[c][c] = sq;
        #pragma unroll
        for (int rr = c + 1; rr < K; ++rr) {
          double v = (fr[rr] && fr[c]) ? A[pk(rr, c)] : 0.0;
          #pragma unroll
          for (int t = 0; t < c; ++t) v -= L[rr][t] * L[c][t];
          L[rr][c] = v * di[c];
        }
      }
      if (quad_max(pd ? 0.0 : 1.0) > 0.0) {
        dense = true;
        done = true;
        return;
      }
      // C = (A_l[F, F])^-1, zero outside F (packed lower triangle): column c solves L L' C[:, c] = e_c
      double C[NKK];
      #pragma unroll
      for (int c = 0; c < K; ++c) {
        double u[K];
        #pragma unroll
        for (int rr = 0; rr < K; ++rr) {
          double v = rr == c ? 1.0 : 0.0;
          #pragma unroll
          for (int t = 0; t < rr; ++t) v -= L[rr][t] * u[t];
          u[rr] = v * di[rr];
        }
        #pragma unroll
        for (int rr = K - 1; rr >= 0; --rr) {
          double v = u[rr];
          #pragma unroll
          for (int t = rr + 1; t < K; ++t) v -= L[t][rr] * u[t];
          u[rr] = v * di[rr];
        }
        #pragma unroll
        for (int rr = c; rr < K; ++rr) C[pk(rr, c)] = (fr[rr] && fr[c]) ? u[rr] : 0.0;
      }
      // gradient at x: r = A x + a sum_l x + g
      double xs[K], gr[K], y[K], ys[K], Cs[K][K];  // Cs: lower triangle used
      #pragma unroll
      for (int k = 0; k < K; ++k) xs[k] = quad_sum<double>(x[k]);
      #pragma unroll
      for (int k = 0; k < K; ++k) {
        double s = g[k];
        #pragma unroll
        for (int k2 = 0; k2 < K; ++k2) s = __builtin_fma(A[pk(k, k2)], x[k2], s);
        gr[k] = __builtin_fma(a, xs[k], s);
      }
      #pragma unroll
      for (int k = 0; k < K; ++k) {
        double s = 0.0;
        #pragma unroll
        for (int k2 = 0; k2 < K; ++k2) s = __builtin_fma(C[pk(k, k2)], fr[k2] ? gr[k2] : 0.0, s);
        y[k] = s;
      }
      #pragma unroll
      for (int k = 0; k < K; ++k) {
        ys[k] = quad_sum<double>(y[k]);
        #pragma unroll
        for (int k2 = 0; k2 <= k; ++k2) Cs[k][k2] = quad_sum<double>(C[pk(k, k2)]);
      }
      double d[K];
      if constexpr (PEN) {
        double t[K];
        #pragma unroll
        for (int k = 0; k < K; ++k) {
          t[k] = ys[k];
          Cs[k][k] += 1.0 / rho;
        }
        small_spd_solve<K>(Cs, t);
        #pragma unroll
        for (int k = 0; k < K; ++k) {
          double s = y[k];
          #pragma unroll
          for (int k2 = 0; k2 < K; ++k2) s = __builtin_fma(-C[pk(k, k2)], t[k2], s);
          d[k] = fr[k] ? -s : 0.0;
        }
      } else {
        // groups without a free member are already satisfied: identity rows, zero multiplier
        #pragma unroll
        for (int k = 0; k < K; ++k) {
          const bool has = Cs[k][k] > 0.0;
          lam[k] = has ? -ys[k] : 0.0;
          if (!has) {
            #pragma unroll
            for (int k2 = 0; k2 < K; ++k2) {
              if (k2 <= k) Cs[k][k2] = k2 == k ? 1.0 : 0.0;
              else Cs[k2][k] = 0.0;
            }
          }
        }
        small_spd_solve<K>(Cs, lam);
        #pragma unroll
        for (int k = 0; k < K; ++k) {
          double s = y[k];
          #pragma unroll
          for (int k2 = 0; k2 < K; ++k2) s = __builtin_fma(C[pk(k, k2)], lam[k2], s);
          // the lone free member of a group cannot move: its step is exactly 0
          const double nfree = quad_sum<double>(fr[k] ? 1.0 : 0.0);
          d[k] = (fr[k] && nfree > 1.5) ? -s : 0.0;
        }
      }
      // ratio test (smallest step to a bound; ties to the smaller variable index l K + k)
      double tmin = 2.0;
      int imin = l * K;
      #pragma unroll
      for (int k = 0; k < K; ++k) {
        double tc = 2.0;
        if (fr[k]) {
          if (d[k] < 0.0 && x[k] + d[k] < lo[k]) tc = (lo[k] - x[k]) / d[k];
          else if (d[k] > 0.0 && x[k] + d[k] > hi[k]) tc = (hi[k] - x[k]) / d[k];
        }
        if (tc < tmin) {
          tmin = tc;
          imin = l * K + k;
        }
      }
      quad_arg<true>(tmin, imin);
      const bool blocked = tmin < 1.0;
      const double alpha = blocked ? fmax(tmin, 0.0) : 1.0;
      #pragma unroll
      for (int k = 0; k < K; ++k)
        if (fr[k]) {
          const double v = x[k] + alpha * d[k];
          x[k] = v < lo[k] ? lo[k] : (v > hi[k] ? hi[k] : v);
        }
      if (blocked) {
        #pragma unroll
        for (int k = 0; k < K; ++k)
          if (imin == l * K + k) {
            const int side = d[k] < 0.0 ? -1 : 1;
            x[k] = side < 0 ? lo[k] : hi[k];
            st[k] = side;
          }
        return;
      }
      // at the minimiser of the working set: bound multipliers, release the most violated
      #pragma unroll
      for (int k = 0; k < K; ++k) xs[k] = quad_sum<double>(x[k]);
      double vmax = -1.0;
      int iv = l * K;
      #pragma unroll
      for (int k = 0; k < K; ++k) {
        double s = g[k];
        #pragma unroll
        for (int k2 = 0; k2 < K; ++k2) s = __builtin_fma(A[pk(k, k2)], x[k2], s);
        s = __builtin_fma(a, xs[k], s);
        if constexpr (!PEN) s += lam[k];
        const double v = st[k] != 0 ? (st[k] < 0 ? -s : s) : -1.0;
        if (v > vmax) {
          vmax = v;
          iv = l * K + k;
        }
      }
      quad_arg<false>(vmax, iv);
      if (!(vmax > tol)) {
        done = true;
        return;
      }
      #pragma unroll
      for (int k = 0; k < K; ++k)
        if (iv == l * K + k) st[k] = 0;
  };
  // result of a finished quad: the dense list, or p_next = proj(p + dp)
  auto finish = [&]() __attribute__((always_inline)) {
    if (skip) return;
    if (dense) {
      if (l == 0) late[1 + atomicAdd(late, 1)] = (int)j;
      return;
    }
    // stopped at the iteration limit (it == maxit; a solved quad stops below it): the last feasible
    // iterate is applied and counted in fails[1]
    if (it >= maxit && l == 0) atomicAdd(fails + 1, 1);
    if (dp) {
      #pragma unroll
      for (int k = 0; k < K; ++k) dp[k + (int64_t)K * (4 * j + l)] = x[k];
    }
    // p_next = p + dp, then project_p! on each (k, j) 4-vector (the quad's four lanes)
    #pragma unroll
    for (int k = 0; k < K; ++k) {
      const double nv = pv[k] + x[k];
      double v[4] = {quad_bcast<0>(nv), quad_bcast<1>(nv), quad_bcast<2>(nv), quad_bcast<3>(nv)};
      proj_simplex<4>(v, eps);
      p_next[k + (int64_t)K * (4 * j + l)] = l == 0 ? v[0] : (l == 1 ? v[1] : (l == 2 ? v[2] : v[3]));
    }
  };
  if (valid) load();
  if constexpr (!MULTI) {  // one quad per slot
    while (__ballot(valid && !done) != 0ull) {
      if (valid && !done) {
        iterate();
        if (!done && ++it >= maxit) done = true;
      }
    }
    if (valid) finish();
    return;
  }
  // A finished quad with a next one hands over at once; the others finish together once every quad of the
  // wavefront is done (one convergent pass instead of one per completion time).
  while (__ballot(valid) != 0ull) {
    if (__ballot(valid && !done) == 0ull) {
      if (valid) {
        finish();
        j += nslot;
        valid = j < J;
        if (valid) load();
      }
      continue;
    }
    if (valid && !done) {
      iterate();
      if (!done && ++it >= maxit) done = true;
    }
    if (valid && done && j + nslot < J) {
      finish();
      j += nslot;
      load();
    }
  }
}

// MULTI: the grid is capped at the resident workgroups and slots solve several quads; otherwise one quad
// per slot.  Two waves per SIMD (at most 256 registers) for the penalty form up to K = 6 and the equality
// form up to K = 5; the larger ones need more registers and run one wave per SIMD.
template <int K, bool PEN, bool MULTI>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_p_step4_w2(
    int64_t J, const double* __restrict__ res, const double* __restrict__ p, double* __restrict__ p_next, double rho,
    double eps, double* __restrict__ dp, const unsigned char* __restrict__ flags, const int* __restrict__ dense_list,
    int ndb, int* __restrict__ fails, int* __restrict__ late, const int* __restrict__ gate) {
  if (gate_off(gate)) return;
  p_step4_body<K, PEN, MULTI>(J, res, p, p_next, rho, eps, dp, flags, dense_list, ndb, fails, late);
}
template <int K, bool PEN, bool MULTI>
__global__ __launch_bounds__(256) void k_p_step4(int64_t J, const double* __restrict__ res,
                                                 const double* __restrict__ p, double* __restrict__ p_next,
                                                 double rho, double eps, double* __restrict__ dp,
                                                 const unsigned char* __restrict__ flags,
                                                 const int* __restrict__ dense_list, int ndb, int* __restrict__ fails,
                                                 int* __restrict__ late, const int* __restrict__ gate) {
  if (gate_off(gate)) return;
  p_step4_body<K, PEN, MULTI>(J, res, p, p_next, rho, eps, dp, flags, dense_list, ndb, fails, late);
}

// ------------------------------------------------------------------------------------------------
// Projections and EM normalisation (thread per column / per (k, j)).
// ------------------------------------------------------------------------------------------------
template <int K>
__global__ void k_project_q(int64_t I, double* __restrict__ q, double eps, const int* __restrict__ gate) {
  if (gate_off(gate)) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= I) return;
  double v[K];
#pragma unroll
  for (int c = 0; c < K; ++c) v[c] = q[c + (int64_t)K * i];
  proj_simplex<K>(v, eps);
#pragma unroll
  for (int c = 0; c < K; ++c) q[c + (int64_t)K * i] = v[c];
}

__global__ void k_project_p(int K, int64_t J, double* __restrict__ p, double eps, const int* __restrict__ gate) {
  if (gate_off(gate)) return;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)K * J) return;
  const int64_t j = idx / K;
  const int k = (int)(idx - j * K);
  double v[4];
#pragma unroll
  for (int h = 0; h < 4; ++h) v[h] = p[k + (int64_t)K * (4 * j + h)];
  proj_simplex<4>(v, eps);
#pragma unroll
  for (int h = 0; h < 4; ++h) p[k + (int64_t)K * (4 * j + h)] = v[h];
}

// em! (algorithms_inner.jl:66,75): q_next = q .* grad / (2 (J - dm_i)), then project_q!.
template <int K>
__global__ void k_em_q(int64_t I, int64_t J, const double* __restrict__ res, int NT, const int64_t* __restrict__ dm,
                       const double* __restrict__ q, double* __restrict__ q_next, double eps) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= I) return;
  const double den = (double)(2 * (J - dm[i]));
  double v[K];
#pragma unroll
  for (int c = 0; c < K; ++c) v[c] = q[c + (int64_t)K * i] * res[i * NT + c] / den;
  proj_simplex<K>(v, eps);
#pragma unroll
  for (int c = 0; c < K; ++c) q_next[c + (int64_t)K * i] = v[c];
}

// em! (algorithms_inner.jl:67-76): p_next = p .* grad normalised over the 4 haplotypes, then project.
__global__ void k_em_p(int K, int64_t J, const double* __restrict__ res, int NT, const double* __restrict__ p,
                       double* __restrict__ p_next, double eps) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)K * J) return;
  const int64_t j = idx / K;
  const int k = (int)(idx - j * K);
  double v[4];
#pragma unroll
  for (int h = 0; h < 4; ++h) v[h] = p[k + (int64_t)K * (4 * j + h)] * res[(4 * j + h) * NT + k];
  const double s = v[0] + v[1] + v[2] + v[3];
#pragma unroll
  for (int h = 0; h < 4; ++h) v[h] /= s;
  proj_simplex<4>(v, eps);
#pragma unroll
  for (int h = 0; h < 4; ++h) p_next[k + (int64_t)K * (4 * j + h)] = v[h];
}

// FP32 mirror of a (q, p) pair in one launch: dst = [q (nq) | p (np)].
__global__ void k_pair_to_f32(const double* __restrict__ q, int64_t nq, const double* __restrict__ p, int64_t np,
                              float* __restrict__ dst, const int* __restrict__ gate) {
  if (gate_off(gate)) return;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nq + np; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = i < nq ? (float)q[i] : (float)p[i - nq];
}

__global__ void k_fill(double* __restrict__ p, int64_t n, double v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Plain streaming copy (16 B per lane per iteration): HBM bandwidth and counter calibration.
__global__ __launch_bounds__(256) void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

__global__ void k_axpy(int64_t n, const double* __restrict__ a, const double* __restrict__ d, double f,
                       double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a[i] + f * d[i];
}

// ------------------------------------------------------------------------------------------------
// ZAL quasi-Newton: U[:,col] = x1 - x0, V[:,col] = x2 - x1; dots U'U, U'V, U'(x0 - x1);
// x_out = x1 - V (U'U - U'V)^{-1} U'(x0 - x1).  Qn <= 4.
// ------------------------------------------------------------------------------------------------
#define ZAL_QMAX 4
#define ZAL_NDOT (ZAL_QMAX * ZAL_QMAX * 3 + ZAL_QMAX * 2)  // U'U, U'V, V'V, U'r, V'r
#define ZAL_NB 512

// The QN iteration i (1-based) is read from device memory so that an iteration can be replayed from a
// captured graph: column (i-1) mod Qn receives the new secant pair, min(i, Qn) columns are used.
__device__ __forceinline__ void zal_cols(const int* it, int Qn, int& col, int& ncols) {
  const int i = *it;
  col = (i - 1) % Qn;
  ncols = i < Qn ? i : Qn;
}

__global__ __launch_bounds__(256) void k_zal_dots(int64_t N, const int* __restrict__ d_iter, int Qn,
                                                  double* __restrict__ U, double* __restrict__ V,
                                                  const double* __restrict__ x0, const double* __restrict__ x1,
                                                  const double* __restrict__ x2, double* __restrict__ part, const int* __restrict__ gate) {
  if (gate_off(gate)) return;
  __shared__ double sh[256];
  int col, ncols;
  zal_cols(d_iter, Qn, col, ncols);
  double acc[ZAL_NDOT];
#pragma unroll
  for (int d = 0; d < ZAL_NDOT; ++d) acc[d] = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < N; t += stride) {
    const double a0 = x0[t], a1 = x1[t], a2 = x2[t];
    const double u = a1 - a0, v = a2 - a1, rr = a0 - a1;
    U[t + N * col] = u;
    V[t + N * col] = v;
    double ua[ZAL_QMAX], va[ZAL_QMAX];
#pragma unroll
    for (int a = 0; a < ZAL_QMAX; ++a) {
      ua[a] = (a < ncols) ? (a == col ? u : U[t + N * a]) : 0.0;
      va[a] = (a < ncols) ? (a == col ? v : V[t + N * a]) : 0.0;
    }
    constexpr int Q2 = ZAL_QMAX * ZAL_QMAX;
#pragma unroll
    for (int a = 0; a < ZAL_QMAX; ++a) {
#pragma unroll
      for (int b = 0; b < ZAL_QMAX; ++b) {
        acc[a * ZAL_QMAX + b] = __builtin_fma(ua[a], ua[b], acc[a * ZAL_QMAX + b]);
        acc[Q2 + a * ZAL_QMAX + b] = __builtin_fma(ua[a], va[b], acc[Q2 + a * ZAL_QMAX + b]);
        acc[2 * Q2 + a * ZAL_QMAX + b] = __builtin_fma(va[a], va[b], acc[2 * Q2 + a * ZAL_QMAX + b]);
      }
      acc[3 * Q2 + a] = __builtin_fma(ua[a], rr, acc[3 * Q2 + a]);
      acc[3 * Q2 + ZAL_QMAX + a] = __builtin_fma(va[a], rr, acc[3 * Q2 + ZAL_QMAX + a]);
    }
  }
#pragma unroll
  for (int d = 0; d < ZAL_NDOT; ++d) {
    const double s = block_sum<256>(acc[d], sh);
    if (threadIdx.x == 0) part[blockIdx.x * ZAL_NDOT + d] = s;
    __syncthreads();
  }
}

// Coefficients of the quasi-Newton point x_qn = F(x) - V c (r = x - F(x), U / V the secant columns):
//   mode 0 (ZAL, Zhou-Alexander-Lange):     (U'U - U'V) c = U' r
//   mode 1 (LBQN, limited-memory Broyden):  (Y'Y) c = -Y' r,  Y = V - U
// Both are multisecant Broyden updates from H0 = -I of the inverse Jacobian of G(x) = F(x) - x: ZAL the
// first ("good") kind, LBQN the second ("bad") kind, H = -I + (U + Y)(Y'Y)^-1 Y', which satisfies the
// secants H Y = U; x_qn = x - H G(x).
// The ZAL_NB block partials of each dot product are summed by ZAL_NG thread groups (group g takes the
// partials g, g + ZAL_NG, ...; all loads issued together), then over the groups in order.
#define ZAL_NG 16
__global__ __launch_bounds__(64 * ZAL_NG) void k_zal_solve(const int* __restrict__ d_iter, int Qn,
                                                          const double* __restrict__ part,
                                                          double* __restrict__ coef, int mode, const int* __restrict__ gate) {
  if (gate_off(gate)) return;
  static_assert(ZAL_NDOT <= 64 && ZAL_NB % ZAL_NG == 0, "one dot product per lane of a group");
  __shared__ double red[ZAL_NG][64];
  __shared__ double dots[ZAL_NDOT];
  int col, ncols;
  zal_cols(d_iter, Qn, col, ncols);
  (void)col;
  const int d = threadIdx.x & 63, grp = threadIdx.x >> 6;
  if (d < ZAL_NDOT) {
    double v[ZAL_NB / ZAL_NG];
#pragma unroll
    for (int t = 0; t < ZAL_NB / ZAL_NG; ++t) v[t] = part[(grp + ZAL_NG * t) * ZAL_NDOT + d];
    double s = 0.0;
#pragma unroll
    for (int t = 0; t < ZAL_NB / ZAL_NG; ++t) s += v[t];
    red[grp][d] = s;
  }
  __syncthreads();
  if (threadIdx.x < ZAL_NDOT) {
    double s = 0.0;
#pragma unroll
    for (int g = 0; g < ZAL_NG; ++g) s += red[g][threadIdx.x];
    dots[threadIdx.x] = s;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  const int n = ncols;
  double A[ZAL_QMAX * ZAL_QMAX], y[ZAL_QMAX];
  constexpr int Q2 = ZAL_QMAX * ZAL_QMAX;
  for (int a = 0; a < n; ++a) {
    const double ur = dots[3 * Q2 + a], vr = dots[3 * Q2 + ZAL_QMAX + a];
    y[a] = mode == 0 ? ur : ur - vr;
    for (int b = 0; b < n; ++b) {
      const double uu = dots[a * ZAL_QMAX + b], uv = dots[Q2 + a * ZAL_QMAX + b];
      const double vu = dots[Q2 + b * ZAL_QMAX + a], vv = dots[2 * Q2 + a * ZAL_QMAX + b];
      A[a + n * b] = mode == 0 ? uu - uv : vv - uv - vu + uu;
    }
  }
  int rc = 0;
  // LU with partial pivoting (same elimination order as the CPU restatement)
  for (int c = 0; c < n && rc == 0; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (fabs(A[r + n * c]) > fabs(A[piv + n * c])) piv = r;
    if (A[piv + n * c] == 0.0) { rc = -1; break; }
    if (piv != c) {
      for (int t = 0; t < n; ++t) { const double tmp = A[c + n * t]; A[c + n * t] = A[piv + n * t]; A[piv + n * t] = tmp; }
      const double tmp = y[c]; y[c] = y[piv]; y[piv] = tmp;
    }
    for (int r = c + 1; r < n; ++r) {
      const double f = A[r + n * c] / A[c + n * c];
      for (int t = c; t < n; ++t) A[r + n * t] -= f * A[c + n * t];
      y[r] -= f * y[c];
    }
  }
  if (rc == 0)
    for (int r = n - 1; r >= 0; --r) {
      double s = y[r];
      for (int t = r + 1; t < n; ++t) s -= A[r + n * t] * y[t];
      y[r] = s / A[r + n * r];
    }
  for (int a = 0; a < ZAL_QMAX; ++a) coef[a] = (rc == 0 && a < n) ? y[a] : 0.0;
  coef[ZAL_QMAX] = (double)rc;
}

__global__ void k_zal_apply(int64_t N, const int* __restrict__ d_iter, int Qn, const double* __restrict__ V,
                            const double* __restrict__ x1, const double* __restrict__ coef, double* __restrict__ x_out, const int* __restrict__ gate) {
  if (gate_off(gate)) return;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
  int col, ncols;
  zal_cols(d_iter, Qn, col, ncols);
  (void)col;
  double s = x1[t];
  if (coef[ZAL_QMAX] == 0.0)
    for (int a = 0; a < ncols; ++a) s -= V[t + N * a] * coef[a];
  x_out[t] = s;
}

// ------------------------------------------------------------------------------------------------
// Device-side accept rule of admixture_qn! (algorithms_outer.jl:154-160): keep the QN point when its
// log-likelihood beats the plain double sweep.  ll[0] = ll_basic, ll[1] = ll_qn -> ll[2] = ll_new; the
// winning parameter vector is copied into x; block 0 / thread 0 then advances the iteration counter.
// ------------------------------------------------------------------------------------------------
__global__ void k_accept(int64_t N, double* __restrict__ ll, const double* __restrict__ x_basic,
                         const double* __restrict__ x_qn, double* __restrict__ x, int* __restrict__ d_iter, const int* __restrict__ gate) {
  if (gate_off(gate)) return;
  const bool take_qn = ll[0] < ll[1];
  const double* src = take_qn ? x_qn : x_basic;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < N; t += stride) x[t] = src[t];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ll[2] = take_qn ? ll[1] : ll[0];
    *d_iter += 1;
  }
}

// The stopping rule of admixture_qn! (algorithms_outer.jl:161-165) on the device: ll[2] holds the iteration's
// ll_new (k_accept); the same floating-point expression as the host's decides, so a batch graph stops each
// fit after exactly the iterations hk_fit would run.
__global__ void k_fit_control(const double* __restrict__ ll, FitCtl* __restrict__ ctl, double* __restrict__ trace,
                              int* __restrict__ gate) {
  if (threadIdx.x != 0 || blockIdx.x != 0 || *gate == 0) return;
  const double ll_new = ll[2];
  const int it = ctl->iter + 1;
  ctl->iter = it;
  if (trace) trace[it - 1] = ll_new;
  const double reldiff = fabs((ll_new - ctl->ll_prev) / ctl->ll_prev);
  ctl->ll_prev = ll_new;
  if (!isfinite(ll_new)) {
    ctl->status = 1;
    *gate = 0;
  } else if (reldiff < ctl->rtol) {
    ctl->converged = 1;
    *gate = 0;
  } else if (it >= ctl->n_iter) {
    *gate = 0;
  }
}

__global__ void k_copy_f64(int64_t n, const double* __restrict__ src, double* __restrict__ dst, const int* __restrict__ gate) {
  if (gate_off(gate)) return;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) dst[t] = src[t];
}

__global__ void k_set_int(int* __restrict__ p, int v, const int* __restrict__ gate) {
  if (gate_off(gate)) return;
  if (threadIdx.x == 0) *p = v;
}

// ------------------------------------------------------------------------------------------------
// Launchers
// ------------------------------------------------------------------------------------------------
// Calls f(std::integral_constant<int, K>) for the runtime K in 1..HK_MAXK.
template <int KK = 1, typename F>
static hipError_t with_k(int K, F&& f) {
  if constexpr (KK > HK_MAXK) {
    return hipErrorInvalidValue;
  } else {
    if (K == KK) {
      f(std::integral_constant<int, KK>{});
      return hipGetLastError();
    }
    return with_k<KK + 1>(K, f);
  }
}

hipError_t launch_q_step(int K, int64_t I, const double* res_q, const double* q, double* q_next, const StepOpts& o,
                         int* fails, hipStream_t s, double* dq) {
  return with_k(K, [&](auto kc) {
    constexpr int KT = decltype(kc)::value;
    if (o.penalty)
      hipLaunchKernelGGL((k_q_step<KT, true>), dim3((unsigned)I), dim3(64), 0, s, I, res_q, q, q_next, o.rho, o.eps,
                         fails, dq, o.gate);
    else
      hipLaunchKernelGGL((k_q_step<KT, false>), dim3((unsigned)I), dim3(64), 0, s, I, res_q, q, q_next, o.rho, o.eps,
                         fails, dq, o.gate);
  });
}

// Workgroups of k_p_step4 resident on the whole device at once (occupancy x compute units): the grid never
// exceeds it, each slot then solves several quads in turn.
template <typename F>
static unsigned resident_blocks(F kernel, int block) {
  int per_cu = 0, dev = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess || per_cu < 1) per_cu = 1;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus < 1)
    cus = 256;
  return (unsigned)(per_cu * cus);
}

// Four-lane P-step kernels of one K and constraint mode: (one quad per slot, several quads per slot, dense
// list kernel).  K = 6 takes the striding form for both (the one-quad-per-slot form would spill at 256
// registers); K <= 5 packs two quads per wave (the _w2 forms).
template <int KT, bool PEN>
struct PStepKernels {
  static auto single() {
    if constexpr (KT == 6 && PEN) return k_p_step4_w2<KT, PEN, true>;
    else if constexpr (KT <= 5) return k_p_step4_w2<KT, PEN, false>;
    else return k_p_step4<KT, PEN, false>;
  }
  static auto multi() {
    if constexpr (KT <= 6 && (KT <= 5 || PEN)) return k_p_step4_w2<KT, PEN, true>;
    else return k_p_step4<KT, PEN, true>;
  }
};

// Resident-slot caps per (constraint mode, K), queried by prime_p_step_caps at context set-up so that the
// launcher itself makes no runtime queries (it may run inside a stream capture).
static unsigned g_pcap[2][HK_MAXK + 1];
static std::mutex g_pcap_mu;

hipError_t prime_p_step_caps(int K) {
  if (K < 1 || K > HK_MAXK) return hipErrorInvalidValue;
  if (K > HK_MAXK_REG) return hipSuccess;  // wide K: every P QP takes the dense solver (launch_p_step_wide)
  std::lock_guard<std::mutex> lk(g_pcap_mu);
  return with_k(K, [&](auto kc) {
    constexpr int KT = decltype(kc)::value;
    if constexpr (KT <= HK_MAXK_REG) {
      if (!g_pcap[1][KT]) g_pcap[1][KT] = resident_blocks(PStepKernels<KT, true>::multi(), 256);
      if (!g_pcap[0][KT]) g_pcap[0][KT] = resident_blocks(PStepKernels<KT, false>::multi(), 256);
    }
  });
}

// One launch for the P row QPs: the first NDB workgroups run the dense one-wave solver over the early list
// (blocks launch_finish_p marked), the rest the four-lane QPs of all other blocks (several quads per slot
// when J exceeds the slots the device holds), so the dense solver's long active-set loops overlap the
// four-lane work.  Blocks the four-lane kernel finds not positive definite itself (rare: the early test is
// the same) go to the late list, solved after it.
hipError_t launch_p_step(int K, int64_t J, const double* res_p, const double* p, double* p_next, const StepOpts& o,
                         int* fails, const PStepLists& lists, hipStream_t s, double* dp) {
  if (K < 1 || K > HK_MAXK) return hipErrorInvalidValue;
  if (K > HK_MAXK_REG) return launch_p_step_wide(K, J, res_p, p, p_next, o, fails, s, dp);
  const unsigned cap = g_pcap[o.penalty ? 1 : 0][K];
  if (!cap) return hipErrorNotReady;  // prime_p_step_caps(K) was not called at set-up
  hipError_t e = launch_set_int(lists.late, 0, s, o.gate);
  if (e != hipSuccess) return e;
  constexpr int NDB = 64;  // 256 wavefronts for the early list (it holds a few blocks in the usual case)
  const unsigned nq = (unsigned)((4 * J + 255) / 256), nd = (unsigned)(J < 1024 ? J : 1024);
  return with_k(K, [&](auto kc) {
    constexpr int KT = decltype(kc)::value;
    if constexpr (KT <= HK_MAXK_REG) {
    auto run = [&](auto single, auto multi, auto list_kernel) {
      // the dense workgroups exit at once when their list is empty (the usual case), so the four-lane grid
      // keeps the whole device's worth of slots
      if (nq <= cap)
        hipLaunchKernelGGL(single, dim3(NDB + nq), dim3(256), 0, s, J, res_p, p, p_next, o.rho, o.eps, dp,
                           lists.flags, lists.dense, NDB, fails, lists.late, o.gate);
      else
        hipLaunchKernelGGL(multi, dim3(NDB + cap), dim3(256), 0, s, J, res_p, p, p_next, o.rho, o.eps, dp,
                           lists.flags, lists.dense, NDB, fails, lists.late, o.gate);
      hipLaunchKernelGGL(list_kernel, dim3(nd), dim3(64), 0, s, res_p, p, p_next, o.rho, o.eps, fails, dp,
                         lists.late, o.gate);
    };
    if (o.penalty)
      run(PStepKernels<KT, true>::single(), PStepKernels<KT, true>::multi(), k_p_step_list<KT, true>);
    else
      run(PStepKernels<KT, false>::single(), PStepKernels<KT, false>::multi(), k_p_step_list<KT, false>);
    }
  });
}

hipError_t launch_project_q(int K, int64_t I, double* q, double eps, hipStream_t s, const int* gate) {
  return with_k(K, [&](auto kc) {
    constexpr int KT = decltype(kc)::value;
    hipLaunchKernelGGL((k_project_q<KT>), dim3((unsigned)((I + 255) / 256)), dim3(256), 0, s, I, q, eps, gate);
  });
}

hipError_t launch_project_p(int K, int64_t J, double* p, double eps, hipStream_t s, const int* gate) {
  const int64_t n = (int64_t)K * J;
  hipLaunchKernelGGL(k_project_p, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, K, J, p, eps, gate);
  return hipGetLastError();
}

hipError_t launch_em_q(int K, int64_t I, int64_t J, const double* res_q, int NT, const int64_t* dm, const double* q,
                       double* q_next, double eps, hipStream_t s) {
  return with_k(K, [&](auto kc) {
    constexpr int KT = decltype(kc)::value;
    hipLaunchKernelGGL((k_em_q<KT>), dim3((unsigned)((I + 255) / 256)), dim3(256), 0, s, I, J, res_q, NT, dm, q,
                       q_next, eps);
  });
}

hipError_t launch_em_p(int K, int64_t J, const double* res_p, int NT, const double* p, double* p_next, double eps,
                       hipStream_t s) {
  const int64_t n = (int64_t)K * J;
  hipLaunchKernelGGL(k_em_p, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, K, J, res_p, NT, p, p_next, eps);
  return hipGetLastError();
}

hipError_t launch_pair_to_f32(const double* q, int64_t nq, const double* p, int64_t np, float* dst, hipStream_t s,
                              const int* gate) {
  hipLaunchKernelGGL(k_pair_to_f32, dim3(grid_1d(nq + np)), dim3(256), 0, s, q, nq, p, np, dst, gate);
  return hipGetLastError();
}

hipError_t launch_fill(double* p, int64_t n, double v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3(grid_1d(n)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_accept(int64_t N, double* ll, const double* x_basic, const double* x_qn, double* x, int* d_iter,
                         hipStream_t s, const int* gate) {
  hipLaunchKernelGGL(k_accept, dim3(1024), dim3(256), 0, s, N, ll, x_basic, x_qn, x, d_iter, gate);
  return hipGetLastError();
}

hipError_t launch_copy_f64(int64_t n, const double* src, double* dst, hipStream_t s, const int* gate) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy_f64, dim3(1024), dim3(256), 0, s, n, src, dst, gate);
  return hipGetLastError();
}

hipError_t launch_fit_control(const double* ll, FitCtl* ctl, double* trace, int* gate, hipStream_t s) {
  hipLaunchKernelGGL(k_fit_control, dim3(1), dim3(64), 0, s, ll, ctl, trace, gate);
  return hipGetLastError();
}

hipError_t launch_set_int(int* p, int v, hipStream_t s, const int* gate) {
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, s, p, v, gate);
  return hipGetLastError();
}

hipError_t launch_copy16(const void* src, void* dst, int64_t n16, hipStream_t s) {
  hipLaunchKernelGGL(k_copy16, dim3(4096), dim3(256), 0, s, (const uint4*)src, (uint4*)dst, n16);
  return hipGetLastError();
}

hipError_t launch_axpy_step(int64_t n, const double* a, const double* d, double f, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_axpy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, a, d, f, out);
  return hipGetLastError();
}

hipError_t launch_zal(int64_t N, int Qn, const int* d_iter, double* U, double* V, const double* x0, const double* x1,
                      const double* x2, double* x_out, double* dots_part, double* coef, hipStream_t s, int mode,
                      const int* gate) {
  if (Qn < 1 || Qn > ZAL_QMAX || mode < 0 || mode > 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_zal_dots, dim3(ZAL_NB), dim3(256), 0, s, N, d_iter, Qn, U, V, x0, x1, x2, dots_part, gate);
  hipLaunchKernelGGL(k_zal_solve, dim3(1), dim3(64 * ZAL_NG), 0, s, d_iter, Qn, dots_part, coef, mode, gate);
  hipLaunchKernelGGL(k_zal_apply, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, N, d_iter, Qn, V, x1, coef, x_out, gate);
  return hipGetLastError();
}

}  // namespace hk
