// hk_qp.h — one-wave QP machinery shared by the row-step kernels (hk_step.hip, hk_step_wide.hip): wave
// reductions, the simplex projection, the one-wave active-set QP and the dense P-side row step.
#pragma once
#include "hk_device.h"
#include "hk_internal.h"

namespace hk {

// ------------------------------------------------------------------------------------------------
// Wave utilities
// ------------------------------------------------------------------------------------------------

__device__ __forceinline__ double rl(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)(unsigned)lo) | ((long long)hi << 32));
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}
// (value, index) minimum with ties to the smaller index; identical result on every lane.
__device__ __forceinline__ void wave_argmin(double& v, int& idx) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_xor(v, off, 64);
    const int oi = __shfl_xor(idx, off, 64);
    if (ov < v || (ov == v && oi < idx)) { v = ov; idx = oi; }
  }
}
__device__ __forceinline__ void wave_argmax(double& v, int& idx) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_xor(v, off, 64);
    const int oi = __shfl_xor(idx, off, 64);
    if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
  }
}

// ------------------------------------------------------------------------------------------------
// Simplex projection with pseudocount: argmin ||x - v|| s.t. x >= eps, sum x = 1.
// Sort (descending) by an odd-even transposition network, then the sorted-prefix threshold; the
// arithmetic sequence is the same as the CPU restatement's, so results agree bitwise.
// ------------------------------------------------------------------------------------------------
template <int N>
__device__ __forceinline__ void proj_simplex(double (&v)[N], double eps) {
  double u[N], s[N];
#pragma unroll
  for (int t = 0; t < N; ++t) { u[t] = v[t] - eps; s[t] = u[t]; }
#pragma unroll
  for (int r = 0; r < N; ++r)
#pragma unroll
    for (int t = r & 1; t + 1 < N; t += 2) {
      const double a = s[t], b = s[t + 1];
      s[t] = a < b ? b : a;
      s[t + 1] = a < b ? a : b;
    }
  const double target = 1.0 - N * eps;
  double cum = 0.0, tau = 0.0;
#pragma unroll
  for (int t = 0; t < N; ++t) {
    cum += s[t];
    const double tt = (cum - target) / (double)(t + 1);
    if (s[t] - tt > 0.0) tau = tt;
  }
#pragma unroll
  for (int t = 0; t < N; ++t) {
    const double y = u[t] - tau;
    v[t] = (y > 0.0 ? y : 0.0) + eps;
  }
}

// ------------------------------------------------------------------------------------------------
// One-wave active-set QP:  minimise 0.5 x'Hx + g'x  s.t.  lo <= x <= hi,  sum_{grp(i)=c} x_i = 0.
// Lane r < N holds Hrow = H[r][:], g, lo, hi and its group (0..M-1).  x = 0 must be feasible.
// Returns the iteration count (>= 0) or a negative code; x holds the solution on lanes < N.
// lds must hold N*N doubles.
// ------------------------------------------------------------------------------------------------
// Orders this wavefront's LDS accesses around the point (its lanes share a private buffer): the LDS
// executes one wavefront's operations in order, the fences keep the compiler from moving them.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int N, int M>
__device__ int wave_qp(const double (&Hrow)[N], double g, double lo, double hi, int grp, double& x_out,
                       double* lds) {
  const int lane = threadIdx.x & 63;
  const bool act = lane < N;
  double x = 0.0;
  int st = act ? (lo >= 0.0 ? -1 : (hi <= 0.0 ? 1 : 0)) : -1;
  double hm = 0.0;
#pragma unroll
  for (int c = 0; c < N; ++c) hm = fmax(hm, fabs(Hrow[c]));
  const double gs = wave_max(act ? fabs(g) : 0.0);
  const double bm = wave_max(act ? fmax(fabs(lo), fabs(hi)) : 0.0);
  hm = wave_max(act ? hm : 0.0);
  const double tol = 1e-11 * (gs + hm * bm) + 1e-300;
  double dmax = 0.0;
#pragma unroll
  for (int c = 0; c < N; ++c) dmax = (lane == c) ? fabs(Hrow[c]) : dmax;
  dmax = wave_max(act ? dmax : 0.0);

  const int maxit = 4 * N + 20;
  int it = 0;
  int code = 0;
  for (; it < maxit; ++it) {
    const bool fr = act && st == 0;
    const unsigned long long fmask = __ballot(fr);
    // Masked curvature: H on free x free, identity elsewhere.  A curvature that is semidefinite up to
    // round-off gets an escalating ridge (1e-12, 1e-9, 1e-6 of the largest diagonal).
    double R[N], dinv = 1.0;
    bool ok = true;
    for (int attempt = 0; attempt < 4; ++attempt) {
      const double ridge = attempt == 0 ? 0.0 : (attempt == 1 ? 1e-12 : attempt == 2 ? 1e-9 : 1e-6) * dmax + 1e-300;
#pragma unroll
      for (int c = 0; c < N; ++c) {
        const bool fc = (fmask >> c) & 1ull;
        R[c] = (fr && fc) ? Hrow[c] + (lane == c ? ridge : 0.0) : (lane == c ? 1.0 : 0.0);
      }
      ok = true;
#pragma unroll
      for (int k = 0; k < N; ++k) {
        double dkk = rl(R[k], k);
        ok = ok && (dkk > 0.0);  // no early exit: the loop must stay fully unrolled (R in registers)
        dkk = ok ? dkk : 1.0;
        const double sq = sqrt(dkk);
        const double isq = 1.0 / sq;
        R[k] = (lane == k) ? sq : (lane > k ? R[k] * isq : R[k]);
        dinv = (lane == k) ? isq : dinv;
#pragma unroll
        for (int c = k + 1; c < N; ++c) {
          const double lck = rl(R[k], c);
          R[c] = __builtin_fma(-R[k], lck, R[c]);
        }
      }
      if (ok) break;
    }
    if (!ok) { code = -2; break; }
    // Transpose L through LDS so lane r also holds column r (LT[k] = L[k][r]).  The buffer belongs to
    // this wavefront alone (a workgroup may run several QPs, one per wavefront).
    wave_lds_sync();
    if (act) {
#pragma unroll
      for (int c = 0; c < N; ++c)
        if (c <= lane) lds[lane * N + c] = R[c];
    }
    wave_lds_sync();
    double LT[N];
#pragma unroll
    for (int k = 0; k < N; ++k) LT[k] = (act && k >= lane) ? lds[k * N + lane] : 0.0;

    // Solve M y = b for b held one element per lane.
    auto solve = [&](double b) -> double {
#pragma unroll
      for (int k = 0; k < N; ++k) {
        const double yk = rl(b, k) * rl(dinv, k);
        b = (lane == k) ? yk : (lane > k ? __builtin_fma(-R[k], yk, b) : b);
      }
#pragma unroll
      for (int k = N - 1; k >= 0; --k) {
        const double xk = rl(b, k) * rl(dinv, k);
        b = (lane == k) ? xk : (lane < k ? __builtin_fma(-LT[k], xk, b) : b);
      }
      return b;
    };

    // Gradient at x: r = Hx + g.
    double r = g;
#pragma unroll
    for (int c = 0; c < N; ++c) r = __builtin_fma(Hrow[c], rl(x, c), r);
    double z = solve(fr ? r : 0.0);
    double lam_me = 0.0;  // multiplier of this lane's group
    bool lone = false;    // this lane is the only free member of its group (its step is exactly 0)
    if constexpr (M > 0) {
      double Y[M];
#pragma unroll
      for (int c = 0; c < M; ++c) Y[c] = solve((fr && grp == c) ? 1.0 : 0.0);
      double S[M][M], rhs[M];
      bool has[M];
#pragma unroll
      for (int c1 = 0; c1 < M; ++c1) {
        const bool mine = fr && grp == c1;
        rhs[c1] = -wave_sum(mine ? z : 0.0);
        const unsigned long long bm_c = __ballot(mine);
        has[c1] = bm_c != 0ull;
        lone = (grp == c1 && __popcll(bm_c) == 1) ? true : lone;
#pragma unroll
        for (int c2 = 0; c2 < M; ++c2) S[c1][c2] = wave_sum(mine ? Y[c2] : 0.0);
      }
      // Constraints without a free member are already satisfied: identity rows.
#pragma unroll
      for (int c1 = 0; c1 < M; ++c1)
        if (!has[c1]) {
#pragma unroll
          for (int c2 = 0; c2 < M; ++c2) { S[c1][c2] = (c1 == c2) ? 1.0 : 0.0; S[c2][c1] = (c1 == c2) ? 1.0 : 0.0; }
          rhs[c1] = 0.0;
        }
      // Cholesky solve of the small SPD system (uniform on every lane).
      bool sok = true;
#pragma unroll
      for (int k = 0; k < M; ++k) {
        double d = S[k][k];
#pragma unroll
        for (int t = 0; t < k; ++t) d -= S[k][t] * S[k][t];
        if (!(d > 0.0)) sok = false;
        d = sqrt(d);
        S[k][k] = d;
#pragma unroll
        for (int rr = k + 1; rr < M; ++rr) {
          double s = S[rr][k];
#pragma unroll
          for (int t = 0; t < k; ++t) s -= S[rr][t] * S[k][t];
          S[rr][k] = s / d;
        }
      }
      if (!sok) { code = -3; break; }
#pragma unroll
      for (int rr = 0; rr < M; ++rr) {
        double s = rhs[rr];
#pragma unroll
        for (int t = 0; t < rr; ++t) s -= S[rr][t] * rhs[t];
        rhs[rr] = s / S[rr][rr];
      }
#pragma unroll
      for (int rr = M - 1; rr >= 0; --rr) {
        double s = rhs[rr];
#pragma unroll
        for (int t = rr + 1; t < M; ++t) s -= S[t][rr] * rhs[t];
        rhs[rr] = s / S[rr][rr];
      }
#pragma unroll
      for (int c = 0; c < M; ++c) {
        z = __builtin_fma(Y[c], rhs[c], z);
        lam_me = (grp == c) ? rhs[c] : lam_me;
      }
    }
    const double d = (fr && !lone) ? -z : 0.0;
    // Ratio test.
    double tc = 2.0;
    int side = 0;
    if (fr) {
      if (d < 0.0 && x + d < lo) { tc = (lo - x) / d; side = -1; }
      else if (d > 0.0 && x + d > hi) { tc = (hi - x) / d; side = 1; }
    }
    double amin = tc;
    int blk = lane;
    wave_argmin(amin, blk);
    const bool blocked = amin < 1.0;
    const double alpha = blocked ? fmax(amin, 0.0) : 1.0;
    if (fr) {
      const double v = x + alpha * d;
      x = v < lo ? lo : (v > hi ? hi : v);
    }
    const int bside = __shfl(side, blk, 64);
    if (blocked) {
      if (lane == blk) { x = bside < 0 ? lo : hi; st = bside; }
      continue;
    }
    // At the minimiser of the working set: bound multipliers.
    double s = g;
#pragma unroll
    for (int c = 0; c < N; ++c) s = __builtin_fma(Hrow[c], rl(x, c), s);
    s += lam_me;
    double viol = (act && st != 0) ? (st < 0 ? -s : s) : -1.0;
    int worst = lane;
    wave_argmax(viol, worst);
    if (!(viol > tol)) break;
    if (lane == worst) st = 0;
  }
  x_out = x;
  return code < 0 ? code : it;
}

// ------------------------------------------------------------------------------------------------
// update_p! row step, dense: one wavefront per SNP block j; variable (l, k) at index l*K + k.
// ------------------------------------------------------------------------------------------------
template <int K, bool PEN>
__device__ __forceinline__ void p_step_dense(int64_t j, const double* __restrict__ res, const double* __restrict__ p,
                                             double* __restrict__ p_next, double rho, double eps,
                                             int* __restrict__ fails, double* __restrict__ dp, double* lds) {
  constexpr bool penalty = PEN;
  constexpr int NT = K + K * (K + 1) / 2;
  constexpr int N = 4 * K;
  const int lane = threadIdx.x & 63;
  const bool act = lane < N;
  const int l = act ? lane / K : 0, k = act ? lane % K : 0;
  const double* r = res + (4 * j + l) * NT;
  // penalty: rho on every (l, l2) block diagonal.  Without it the K equalities sum_l dp[l,k] = 0 make
  // sigma (11' x I_K) vanish on the feasible set; it is added for definiteness (sigma = largest diagonal).
  const double sigma = wave_max(act ? r[K + k * (k + 1) / 2 + k] : 0.0);
  const double add = penalty ? rho : sigma;
  double Hrow[N];
#pragma unroll
  for (int c = 0; c < N; ++c) {
    const int l2 = c / K, k2 = c % K;
    double h = 0.0;
    if (act && l2 == l) {
      const int a = k > k2 ? k : k2, b = k > k2 ? k2 : k;
      h = r[K + a * (a + 1) / 2 + b];
    }
    if (act && k2 == k) h += add;
    Hrow[c] = h;
  }
  const double pv = act ? p[k + (int64_t)K * (4 * j + l)] : 0.0;
  const double g = act ? -r[k] : 0.0;
  const double lo = act ? 0.0 - pv : 0.0, hi = act ? 1.0 - pv : 0.0;
  double x = 0.0;
  int it;
  if constexpr (PEN) it = wave_qp<N, 0>(Hrow, g, lo, hi, k, x, lds);
  else it = wave_qp<N, K>(Hrow, g, lo, hi, k, x, lds);
  if (it < 0) x = 0.0;
  if (it < 0 && lane == 0) atomicAdd(fails, 1);
  if (it >= 4 * N + 20 && lane == 0) atomicAdd(fails + 1, 1);  // iteration limit: last iterate applied
  if (dp && act) dp[k + (int64_t)K * (4 * j + l)] = x;  // raw step, for step halving
  // p_next = p + dp, then project_p! on each (k, j) 4-vector: lane k gathers haplotypes 0..3.
  const double nv = pv + x;
  double v[4];
#pragma unroll
  for (int h = 0; h < 4; ++h) v[h] = __shfl(nv, h * K + (lane < K ? lane : 0), 64);
  if (lane < K) {
    proj_simplex<4>(v, eps);
#pragma unroll
    for (int h = 0; h < 4; ++h) p_next[lane + (int64_t)K * (4 * j + h)] = v[h];
  }
}

}  // namespace hk
