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
#include "hk_pstep4.h"

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
  std::lock_guard<std::mutex> lk(g_pcap_mu);
  if (K > HK_MAXK_REG) {  // hk_step_wide.hip
    if (!g_pcap[1][K]) g_pcap[1][K] = p_step_cap_wide(K, true);
    if (!g_pcap[0][K]) g_pcap[0][K] = p_step_cap_wide(K, false);
    return g_pcap[0][K] && g_pcap[1][K] ? hipSuccess : hipErrorInvalidValue;
  }
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
  const unsigned cap = g_pcap[o.penalty ? 1 : 0][K];
  if (!cap) return hipErrorNotReady;  // prime_p_step_caps(K) was not called at set-up
  hipError_t e = launch_set_int(lists.late, 0, s, o.gate);
  if (e != hipSuccess) return e;
  constexpr int NDB = 64;  // 256 wavefronts for the early list (it holds a few blocks in the usual case)
  if (K > HK_MAXK_REG) return launch_p_step_wide(K, J, res_p, p, p_next, o, fails, lists, NDB, s, dp);
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
