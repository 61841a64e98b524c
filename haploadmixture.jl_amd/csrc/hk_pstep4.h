// hk_pstep4.h — update_p! row-step kernels shared by the K <= 8 launchers (hk_step.hip) and the 9 <= K <= 16
// ones (hk_step_wide.hip): the dense list kernel and the four-lane block-structured solver, templated on K.
//
// Reference behaviour: update_p! QP stage src/algorithms_inner.jl:259-312 (block-diagonal assembly :272-283,
// default penalty rho on every (l, l2) block diagonal :295-302), step and projection :344-351.
#pragma once
#include "hk_device.h"
#include "hk_internal.h"
#include "hk_qp.h"

namespace hk {

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

// Workgroups of k_p_step4 resident on the whole device at once (occupancy x compute units): the grid never
// exceeds it, each slot then solves several quads in turn.
template <typename F>
inline unsigned resident_blocks(F kernel, int block) {
  int per_cu = 0, dev = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess || per_cu < 1) per_cu = 1;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus < 1)
    cus = 256;
  return (unsigned)(per_cu * cus);
}

}  // namespace hk
