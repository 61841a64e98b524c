// hk_pass.hip — genotype-pass kernels of the HaploADMIXTURE likelihood engine (gfx950).
//
// One "pair" is (individual i, SNP block j); block j = SNPs (2j, 2j+1).  For every pair the
// reference (src/loops.jl:15-155) forms the mixture haplotype frequencies
//     X_h = qp_h = sum_k q[k,i] p[k,4j+h],   h = 2a+b in {00,01,10,11},
// and from the two genotypes the ordered-pair likelihood L = sum over compatible ordered haplotype
// pairs of X_h1 X_h2 and the posterior expected counts c_h.  With the per-SNP compatibility
// matrices S1, S2 (hk_device.h) these have a closed form used here:
//     Y = S1 X S2 (2x2),  L = <X, Y>,  c_h = 2 X_h Y_h / L,
// so  c_h / qp_h = 2 Y_h / L  (the gradient weight w_h) and  c_h / qp_h^2 = w_h / X_h  (v_h).
// The reference's log(a1+a2) + log(b1+b2) (loops.jl:85-104) equals log L - log 2 * [a SNP is het];
// double-missing pairs contribute 0 to every sum (their @coefs are all zero).
//
// Passes (reference -> kernel):
//   loglikelihood_full_loop2 (loops.jl:236-258)  -> k_qside<MODE_LL>
//   em_loop! q part (loops.jl:260-295)           -> k_qside<MODE_GRAD>  (q_next = q .* grad)
//   em_loop!/em_p_loop! p part (:260-326)        -> k_pside<MODE_GRAD>  (p_next = p .* grad)
//   update_q_loop! (loops.jl:328-365)            -> k_qside<MODE_CURV>
//   update_p_loop! (loops.jl:367-403)            -> k_pside<MODE_CURV>
//
// Genotype streams in HBM (built once from the BED payload by k_repack_*), 4 bits per pair
// (bits 0-1 = BED code of SNP 2j, bits 2-3 = BED code of SNP 2j+1), 0.5 byte per pair:
//   gq[jw * Ip + i]  : uint32 = blocks 8jw..8jw+7 of individual i   (Q side: lane = individual)
//   gp[iw * Jp + j]  : uint32 = individuals 8iw..8iw+7 of block j   (P side: lane quad = block)
// Padding individuals (i >= I) and blocks (j >= J) carry nibble 0b0101 (double missing).
//
// Q side: one lane per individual, the block's 4K p values are wave-uniform (scalar loads);
//         the j range is split into chunks, partial sums [chunk][t][Ip] are reduced in order.
// P side: four lanes per block (one haplotype each), X_h exchanged with quad DPP broadcasts,
//         q[:, i] wave-uniform (scalar loads); optional i chunks reduced in order.
// Accumulators are FP64.  In MIXED mode the per-pair arithmetic is FP32 and the 8 pairs of a
// genotype word are summed in FP32 before being added to the FP64 accumulators.
#include "hk_device.h"
#include "hk_internal.h"

namespace hk {

// -----------------------------------------------------------------------------------------------
// Repacking and counting
// -----------------------------------------------------------------------------------------------

__global__ void k_repack_q(const uint8_t* __restrict__ raw, int64_t bps, int64_t I, int64_t J, int64_t Ip,
                           int64_t J8, uint32_t* __restrict__ gq) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= J8 * Ip) return;
  const int64_t jw = idx / Ip, i = idx - jw * Ip;
  uint32_t w = 0;
  for (int n = 0; n < 8; ++n) {
    const int64_t j = jw * 8 + n;
    uint32_t nib = HK_NIB_DM;
    if (i < I && j < J) {
      const unsigned sh = (unsigned)(i & 3) * 2u;
      const unsigned c1 = (raw[(2 * j) * bps + (i >> 2)] >> sh) & 3u;
      const unsigned c2 = (raw[(2 * j + 1) * bps + (i >> 2)] >> sh) & 3u;
      nib = c1 | (c2 << 2);
    }
    w |= nib << (4 * n);
  }
  gq[idx] = w;
}

__global__ void k_repack_p(const uint8_t* __restrict__ raw, int64_t bps, int64_t I, int64_t J, int64_t Jp,
                           int64_t I8, uint32_t* __restrict__ gp) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= I8 * Jp) return;
  const int64_t iw = idx / Jp, j = idx - iw * Jp;
  uint32_t w = 0;
  for (int n = 0; n < 8; ++n) {
    const int64_t i = iw * 8 + n;
    uint32_t nib = HK_NIB_DM;
    if (i < I && j < J) {
      const unsigned sh = (unsigned)(i & 3) * 2u;
      const unsigned c1 = (raw[(2 * j) * bps + (i >> 2)] >> sh) & 3u;
      const unsigned c2 = (raw[(2 * j + 1) * bps + (i >> 2)] >> sh) & 3u;
      nib = c1 | (c2 << 2);
    }
    w |= nib << (4 * n);
  }
  gp[idx] = w;
}

// count_double_missing! (src/utils.jl:1-18) over the first J blocks.
__global__ void k_count_dm(const uint32_t* __restrict__ gq, int64_t Ip, int64_t I, int64_t J, int64_t J8,
                           int64_t* __restrict__ dm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= I) return;
  int64_t n = 0;
  for (int64_t jw = 0; jw < J8; ++jw) {
    const uint32_t w = gq[jw * Ip + i];
    for (int t = 0; t < 8; ++t)
      if (jw * 8 + t < J && ((w >> (4 * t)) & 15u) == HK_NIB_DM) ++n;
  }
  dm[i] = n;
}

// Decoded genotype readback (bit-exact decode check): g[s * I + i] = BED code of SNP s.
__global__ void k_decode(const uint32_t* __restrict__ gq, int64_t Ip, int64_t I, int64_t J,
                         uint8_t* __restrict__ codes) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= I * J) return;
  const int64_t j = idx / I, i = idx - j * I;
  const uint32_t w = gq[(j >> 3) * Ip + i];
  const unsigned nib = (w >> (4 * (j & 7))) & 15u;
  codes[(2 * j) * I + i] = (uint8_t)(nib & 3u);
  codes[(2 * j + 1) * I + i] = (uint8_t)(nib >> 2);
}

// -----------------------------------------------------------------------------------------------
// Q side
// -----------------------------------------------------------------------------------------------

template <int K, int MODE>
struct NAcc {
  static constexpr int v = MODE == MODE_LL ? 1 : (MODE == MODE_GRAD ? K : K + K * (K + 1) / 2);
};

template <int K, typename TP, int MODE>
__global__ __launch_bounds__(256) void k_qside(const uint32_t* __restrict__ gq, int64_t Ip, int64_t I, int J8,
                                               int wpc, const TP* __restrict__ q, const TP* __restrict__ p,
                                               double* __restrict__ part) {
  constexpr int NT = NAcc<K, MODE>::v;
  constexpr bool F32 = sizeof(TP) == 4;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int chunk = blockIdx.y;
  const int jw0 = chunk * wpc;
  const int jw1 = min(J8, jw0 + wpc);
  const int64_t ic = i < I ? i : I - 1;
  TP qi[K];
#pragma unroll
  for (int k = 0; k < K; ++k) qi[k] = q[k + (int64_t)K * ic];

  double acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = 0.0;
  double llm = 1.0;  // FP64 log-likelihood: running product mantissa ...
  int lle = 0;       // ... and binary exponent

  for (int jw = jw0; jw < jw1; ++jw) {
    const uint32_t word = gq[(int64_t)jw * Ip + i];
    TP pa[NT];  // FP32 word partials (MIXED)
    if constexpr (F32) {
#pragma unroll
      for (int t = 0; t < NT; ++t) pa[t] = (TP)0;
    }
#pragma unroll
    for (int n = 0; n < 8; ++n) {
      const TP* pj = p + (int64_t)K * 4 * (jw * 8 + n);
      TP P[4][K];
#pragma unroll
      for (int h = 0; h < 4; ++h)
#pragma unroll
        for (int k = 0; k < K; ++k) P[h][k] = pj[h * K + k];
      const unsigned nib = (word >> (4 * n)) & 15u;
      const unsigned c1 = nib & 3u, c2 = nib >> 2;
      TP X[4];
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        TP s = (TP)0;
#pragma unroll
        for (int k = 0; k < K; ++k) s = __builtin_fma(qi[k], P[h][k], s);
        X[h] = s;
      }
      TP s1_00, s1_01, s1_11, s2_00, s2_01, s2_11;
      snp_s<TP>(c1, s1_00, s1_01, s1_11);
      snp_s<TP>(c2, s2_00, s2_01, s2_11);
      // Z = X S2 (rows a), Y = S1 Z
      const TP z00 = X[0] * s2_00 + X[1] * s2_01, z01 = X[0] * s2_01 + X[1] * s2_11;
      const TP z10 = X[2] * s2_00 + X[3] * s2_01, z11 = X[2] * s2_01 + X[3] * s2_11;
      TP Y[4];
      Y[0] = s1_00 * z00 + s1_01 * z10;
      Y[1] = s1_00 * z01 + s1_01 * z11;
      Y[2] = s1_01 * z00 + s1_11 * z10;
      Y[3] = s1_01 * z01 + s1_11 * z11;
      const TP L = X[0] * Y[0] + X[1] * Y[1] + X[2] * Y[2] + X[3] * Y[3];
      const bool dm = nib == HK_NIB_DM;
      if constexpr (MODE == MODE_LL) {
        const bool het = (c1 == 2u) | (c2 == 2u);
        TP Lh = het ? (TP)0.5 * L : L;
        Lh = dm ? (TP)1 : Lh;
        if constexpr (F32) {
          pa[0] += __builtin_amdgcn_logf(Lh);  // log2
        } else {
          llm *= Lh;
        }
      } else {
        const TP rl = dm ? (TP)0 : (TP)2 * hk_rcp(L);
        TP w[4];
#pragma unroll
        for (int h = 0; h < 4; ++h) w[h] = Y[h] * rl;
        TP* dst = nullptr;
        (void)dst;
#pragma unroll
        for (int k = 0; k < K; ++k) {
          TP g = w[0] * P[0][k];
          g = __builtin_fma(w[1], P[1][k], g);
          g = __builtin_fma(w[2], P[2][k], g);
          g = __builtin_fma(w[3], P[3][k], g);
          if constexpr (F32) pa[k] += g; else acc[k] += g;
        }
        if constexpr (MODE == MODE_CURV) {
          TP v[4];
#pragma unroll
          for (int h = 0; h < 4; ++h) v[h] = w[h] * hk_rcp(X[h]);
#pragma unroll
          for (int k = 0; k < K; ++k) {
            TP u0 = v[0] * P[0][k], u1 = v[1] * P[1][k], u2 = v[2] * P[2][k], u3 = v[3] * P[3][k];
#pragma unroll
            for (int k2 = 0; k2 <= k; ++k2) {
              const int t = K + k * (k + 1) / 2 + k2;
              if constexpr (F32) {
                TP s = pa[t];
                s = __builtin_fma(u0, P[0][k2], s);
                s = __builtin_fma(u1, P[1][k2], s);
                s = __builtin_fma(u2, P[2][k2], s);
                s = __builtin_fma(u3, P[3][k2], s);
                pa[t] = s;
              } else {
                double s = acc[t];
                s = __builtin_fma(u0, P[0][k2], s);
                s = __builtin_fma(u1, P[1][k2], s);
                s = __builtin_fma(u2, P[2][k2], s);
                s = __builtin_fma(u3, P[3][k2], s);
                acc[t] = s;
              }
            }
          }
        }
      }
    }
    if constexpr (F32) {
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[t] += (double)pa[t];
    } else if constexpr (MODE == MODE_LL) {
      int e;
      llm = frexp(llm, &e);
      lle += e;
    }
  }
  if constexpr (MODE == MODE_LL) {
    if constexpr (F32) acc[0] *= 0.69314718055994530942;  // log2 -> natural log
    else acc[0] = log(llm) + (double)lle * 0.69314718055994530942;
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) part[((int64_t)chunk * NT + t) * Ip + i] = acc[t];
}

// -----------------------------------------------------------------------------------------------
// P side
// -----------------------------------------------------------------------------------------------

template <int K, typename TP, int MODE>
__global__ __launch_bounds__(256) void k_pside(const uint32_t* __restrict__ gp, int64_t Jp, int64_t I, int I8,
                                               int wpc, const TP* __restrict__ q, const TP* __restrict__ p,
                                               double* __restrict__ out) {
  constexpr int NT = NAcc<K, MODE>::v;
  constexpr bool F32 = sizeof(TP) == 4;
  const int64_t jh = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // 4j + h
  const int64_t j = jh >> 2;
  const unsigned h = (unsigned)(jh & 3);
  const unsigned a = h >> 1, b = h & 1u;
  const int chunk = blockIdx.y;
  const int iw0 = chunk * wpc;
  const int iw1 = min(I8, iw0 + wpc);
  TP Pw[K];
#pragma unroll
  for (int k = 0; k < K; ++k) Pw[k] = p[k + (int64_t)K * jh];

  double acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = 0.0;

  for (int iw = iw0; iw < iw1; ++iw) {
    const uint32_t word = gp[(int64_t)iw * Jp + j];
    TP pa[NT];
    if constexpr (F32) {
#pragma unroll
      for (int t = 0; t < NT; ++t) pa[t] = (TP)0;
    }
#pragma unroll
    for (int n = 0; n < 8; ++n) {
      int64_t ii = (int64_t)iw * 8 + n;
      ii = ii < I ? ii : I - 1;
      const TP* qv = q + (int64_t)K * ii;
      TP Q[K];
#pragma unroll
      for (int k = 0; k < K; ++k) Q[k] = qv[k];
      const unsigned nib = (word >> (4 * n)) & 15u;
      const unsigned c1 = nib & 3u, c2 = nib >> 2;
      TP X = (TP)0;
#pragma unroll
      for (int k = 0; k < K; ++k) X = __builtin_fma(Q[k], Pw[k], X);
      const TP X0 = quad_bcast<0>(X), X1 = quad_bcast<1>(X), X2 = quad_bcast<2>(X), X3 = quad_bcast<3>(X);
      TP s1_00, s1_01, s1_11, s2_00, s2_01, s2_11;
      snp_s<TP>(c1, s1_00, s1_01, s1_11);
      snp_s<TP>(c2, s2_00, s2_01, s2_11);
      const TP r0 = a ? s1_01 : s1_00, r1 = a ? s1_11 : s1_01;
      const TP t0 = b ? s2_01 : s2_00, t1 = b ? s2_11 : s2_01;
      const TP Y = r0 * (t0 * X0 + t1 * X1) + r1 * (t0 * X2 + t1 * X3);
      const TP L = quad_sum<TP>(X * Y);
      const bool dm = nib == HK_NIB_DM;
      const TP w = dm ? (TP)0 : (TP)2 * Y * hk_rcp(L);
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if constexpr (F32) pa[k] = __builtin_fma(w, Q[k], pa[k]);
        else acc[k] = __builtin_fma(w, Q[k], acc[k]);
      }
      if constexpr (MODE == MODE_CURV) {
        const TP v = w * hk_rcp(X);
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const TP u = v * Q[k];
#pragma unroll
          for (int k2 = 0; k2 <= k; ++k2) {
            const int t = K + k * (k + 1) / 2 + k2;
            if constexpr (F32) pa[t] = __builtin_fma(u, Q[k2], pa[t]);
            else acc[t] = __builtin_fma(u, Q[k2], acc[t]);
          }
        }
      }
    }
    if constexpr (F32) {
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[t] += (double)pa[t];
    }
  }
  double* o = out + ((int64_t)chunk * 4 * Jp + jh) * NT;
#pragma unroll
  for (int t = 0; t < NT; ++t) o[t] = acc[t];
}

// -----------------------------------------------------------------------------------------------
// Reductions over chunks (fixed order)
// -----------------------------------------------------------------------------------------------

// part[c][t][Ip] -> res[i][NT] for i < I
__global__ void k_reduce_q(const double* __restrict__ part, int nchunk, int NT, int64_t Ip, int64_t I,
                           double* __restrict__ res) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= I) return;
  for (int t = 0; t < NT; ++t) {
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += part[((int64_t)c * NT + t) * Ip + i];
    res[i * NT + t] = s;
  }
}

// part[c][4Jp][NT] -> res[jh][NT] for jh < 4J
__global__ void k_reduce_p(const double* __restrict__ part, int nchunk, int NT, int64_t Jp, int64_t J,
                           double* __restrict__ res) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= 4 * J * NT) return;
  double s = 0.0;
  for (int c = 0; c < nchunk; ++c) s += part[(int64_t)c * 4 * Jp * NT + idx];
  res[idx] = s;
}

// Log-likelihood: rows[i] = sum_c part[c][0][i]; total over rows in a fixed tree order.
__global__ __launch_bounds__(1024) void k_reduce_ll(const double* __restrict__ part, int nchunk, int64_t Ip,
                                                    int64_t I, double* __restrict__ rows,
                                                    double* __restrict__ total) {
  __shared__ double sh[1024];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < I; i += 1024) {
    double r = 0.0;
    for (int c = 0; c < nchunk; ++c) r += part[(int64_t)c * Ip + i];
    if (rows) rows[i] = r;
    s += r;
  }
  s = block_sum<1024>(s, sh);
  if (threadIdx.x == 0) *total = s;
}

// res[n][NT] (Xtz first K, then packed lower XtX) -> reference layouts XtX (K,K,n), Xtz (K,n).
__global__ void k_expand(const double* __restrict__ res, int K, int NT, int64_t n, double* __restrict__ XtX,
                         double* __restrict__ Xtz) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const double* r = res + c * NT;
  if (Xtz)
    for (int k = 0; k < K; ++k) Xtz[k + (int64_t)K * c] = r[k];
  if (XtX)
    for (int k = 0; k < K; ++k)
      for (int k2 = 0; k2 <= k; ++k2) {
        const double v = r[K + k * (k + 1) / 2 + k2];
        XtX[k2 + (int64_t)K * (k + (int64_t)K * c)] = v;
        XtX[k + (int64_t)K * (k2 + (int64_t)K * c)] = v;
      }
}

// EM sums in the reference's form (em_loop!, loops.jl:286-291): out[k, c] = par[k, c] * grad[c][k].
__global__ void k_em_expand(const double* __restrict__ res, int K, int NT, int64_t n,
                            const double* __restrict__ par, double* __restrict__ out) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)K * n) return;
  const int64_t c = idx / K;
  const int k = (int)(idx - c * K);
  out[idx] = par[idx] * res[c * NT + k];
}

// -----------------------------------------------------------------------------------------------
// Dispatch
// -----------------------------------------------------------------------------------------------

template <int K, typename TP>
static hipError_t qside_k(int mode, const PassGeom& g, const void* q, const void* p, double* part,
                          hipStream_t s) {
  dim3 grid((unsigned)(g.Ip / 256), (unsigned)g.qchunks), block(256);
  const TP* qq = (const TP*)q;
  const TP* pp = (const TP*)p;
  if (mode == MODE_LL)
    hipLaunchKernelGGL((k_qside<K, TP, MODE_LL>), grid, block, 0, s, g.gq, g.Ip, g.I, (int)g.J8, g.qwpc, qq, pp, part);
  else if (mode == MODE_GRAD)
    hipLaunchKernelGGL((k_qside<K, TP, MODE_GRAD>), grid, block, 0, s, g.gq, g.Ip, g.I, (int)g.J8, g.qwpc, qq, pp, part);
  else
    hipLaunchKernelGGL((k_qside<K, TP, MODE_CURV>), grid, block, 0, s, g.gq, g.Ip, g.I, (int)g.J8, g.qwpc, qq, pp, part);
  return hipGetLastError();
}

template <int K, typename TP>
static hipError_t pside_k(int mode, const PassGeom& g, const void* q, const void* p, double* out,
                          hipStream_t s) {
  dim3 grid((unsigned)(4 * g.Jp / 256), (unsigned)g.pchunks), block(256);
  const TP* qq = (const TP*)q;
  const TP* pp = (const TP*)p;
  if (mode == MODE_GRAD)
    hipLaunchKernelGGL((k_pside<K, TP, MODE_GRAD>), grid, block, 0, s, g.gp, g.Jp, g.I, (int)g.I8, g.pwpc, qq, pp, out);
  else
    hipLaunchKernelGGL((k_pside<K, TP, MODE_CURV>), grid, block, 0, s, g.gp, g.Jp, g.I, (int)g.I8, g.pwpc, qq, pp, out);
  return hipGetLastError();
}

#define HK_K_SWITCH(FN, TP, ...)                 \
  switch (K) {                                   \
    case 1: return FN<1, TP>(__VA_ARGS__);       \
    case 2: return FN<2, TP>(__VA_ARGS__);       \
    case 3: return FN<3, TP>(__VA_ARGS__);       \
    case 4: return FN<4, TP>(__VA_ARGS__);       \
    case 5: return FN<5, TP>(__VA_ARGS__);       \
    case 6: return FN<6, TP>(__VA_ARGS__);       \
    case 7: return FN<7, TP>(__VA_ARGS__);       \
    case 8: return FN<8, TP>(__VA_ARGS__);       \
    default: return hipErrorInvalidValue;        \
  }

hipError_t launch_qside(int K, int prec, int mode, const PassGeom& g, const void* q, const void* p, double* part,
                        hipStream_t s) {
  if (prec == PREC_FP64) { HK_K_SWITCH(qside_k, double, mode, g, q, p, part, s) }
  HK_K_SWITCH(qside_k, float, mode, g, q, p, part, s)
}

hipError_t launch_pside(int K, int prec, int mode, const PassGeom& g, const void* q, const void* p, double* out,
                        hipStream_t s) {
  if (prec == PREC_FP64) { HK_K_SWITCH(pside_k, double, mode, g, q, p, out, s) }
  HK_K_SWITCH(pside_k, float, mode, g, q, p, out, s)
}

int nacc(int K, int mode) { return mode == MODE_LL ? 1 : (mode == MODE_GRAD ? K : K + K * (K + 1) / 2); }

hipError_t launch_repack(const uint8_t* raw, int64_t bps, const PassGeom& g, hipStream_t s) {
  int64_t nq = g.J8 * g.Ip, np = g.I8 * g.Jp;
  hipLaunchKernelGGL(k_repack_q, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, s, raw, bps, g.I, g.J, g.Ip,
                     g.J8, (uint32_t*)g.gq);
  hipLaunchKernelGGL(k_repack_p, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, raw, bps, g.I, g.J, g.Jp,
                     g.I8, (uint32_t*)g.gp);
  return hipGetLastError();
}

hipError_t launch_count_dm(const PassGeom& g, int64_t* dm, hipStream_t s) {
  hipLaunchKernelGGL(k_count_dm, dim3((unsigned)((g.I + 255) / 256)), dim3(256), 0, s, g.gq, g.Ip, g.I, g.J, g.J8, dm);
  return hipGetLastError();
}

hipError_t launch_decode(const PassGeom& g, uint8_t* codes, hipStream_t s) {
  int64_t n = g.I * g.J;
  hipLaunchKernelGGL(k_decode, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g.gq, g.Ip, g.I, g.J, codes);
  return hipGetLastError();
}

hipError_t launch_reduce_q(const double* part, int nchunk, int NT, const PassGeom& g, double* res, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_q, dim3((unsigned)((g.I + 255) / 256)), dim3(256), 0, s, part, nchunk, NT, g.Ip, g.I, res);
  return hipGetLastError();
}

hipError_t launch_reduce_p(const double* part, int nchunk, int NT, const PassGeom& g, double* res, hipStream_t s) {
  int64_t n = 4 * g.J * NT;
  hipLaunchKernelGGL(k_reduce_p, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, part, nchunk, NT, g.Jp, g.J, res);
  return hipGetLastError();
}

hipError_t launch_reduce_ll(const double* part, int nchunk, const PassGeom& g, double* rows, double* total,
                            hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_ll, dim3(1), dim3(1024), 0, s, part, nchunk, g.Ip, g.I, rows, total);
  return hipGetLastError();
}

hipError_t launch_expand(const double* res, int K, int NT, int64_t n, double* XtX, double* Xtz, hipStream_t s) {
  hipLaunchKernelGGL(k_expand, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, res, K, NT, n, XtX, Xtz);
  return hipGetLastError();
}

hipError_t launch_em_expand(const double* res, int K, int NT, int64_t n, const double* par, double* out,
                            hipStream_t s) {
  int64_t m = (int64_t)K * n;
  hipLaunchKernelGGL(k_em_expand, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, res, K, NT, n, par, out);
  return hipGetLastError();
}

}  // namespace hk
