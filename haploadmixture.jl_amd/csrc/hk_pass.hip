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
//   update_q_loop! (loops.jl:328-365)            -> k_qside_mm (matrix cores + VALU tail)
//   update_p_loop! (loops.jl:367-403)            -> k_pside_mm (matrix cores + VALU tail)
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
//
// The curvature passes accumulate XtX only: the reference's gradient follows from it exactly,
//     sum_k' XtX_q[k,k',i] q[k',i] = sum_{j,h} v_h p_kh X_h = sum_{j,h} w_h p_kh = Xtz_q[k,i]
// (likewise Xtz_p[:, jh] = XtX_p[:, :, jh] p[:, jh]), formed once per row by the reduce / finish
// kernels from the FP64 parameters.  All terms are non-negative, so no cancellation enters.
#include "hk_device.h"
#include "hk_internal.h"
#include "hk_pair.h"

namespace hk {

// -----------------------------------------------------------------------------------------------
// Repacking and counting
// -----------------------------------------------------------------------------------------------

// Element loops over sizes that can pass 2^32 (the genotype streams hold I J / 8 words: I = 80,000 x J = 540,000
// blocks is 2^32) stride over a capped grid (grid_1d): a launch of more than 2^32 work-items would be truncated.
__global__ void k_repack_q(const uint8_t* __restrict__ raw, int64_t bps, int64_t I, int64_t J, int64_t Ip,
                           int64_t J8, uint32_t* __restrict__ gq) {
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < J8 * Ip;
       idx += (int64_t)gridDim.x * blockDim.x) {
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
}

// Block-major stream from the individual-major one: an 8 x 8 nibble transpose per thread.  Thread (iw, jw)
// reads gq[jw][8 iw .. 8 iw + 7] (8 individuals x 8 blocks) and writes gp[iw][8 jw .. 8 jw + 7]; consecutive
// threads take consecutive jw, so the 32-byte writes are contiguous.  Padding nibbles carry over.
__global__ void k_repack_p(const uint32_t* __restrict__ gq, int64_t Ip, int64_t J8, int64_t Jp, int64_t I8,
                           uint32_t* __restrict__ gp) {
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < I8 * J8;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t iw = idx / J8, jw = idx - iw * J8;
    uint32_t in[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) in[t] = gq[jw * Ip + 8 * iw + t];
#pragma unroll
    for (int n = 0; n < 8; ++n) {
      uint32_t w = 0;
#pragma unroll
      for (int t = 0; t < 8; ++t) w |= ((in[t] >> (4 * n)) & 15u) << (4 * t);
      gp[iw * Jp + 8 * jw + n] = w;
    }
  }
}

// count_double_missing! (src/utils.jl:1-18) over the first J blocks: per individual, the double-missing
// nibbles (0b0101) of a chunk of DM_CW words, counted with bit operations (x = w ^ 0x55555555 has a zero
// nibble exactly there), added with integer atomics (exact in any order).  dm must start at zero.
constexpr int DM_CW = 64;
__global__ void k_count_dm(const uint32_t* __restrict__ gq, int64_t Ip, int64_t I, int64_t J, int64_t J8,
                           unsigned long long* __restrict__ dm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= I) return;
  const int64_t w0 = (int64_t)blockIdx.y * DM_CW, w1 = w0 + DM_CW < J8 ? w0 + DM_CW : J8;
  unsigned n = 0;
  for (int64_t jw = w0; jw < w1; ++jw) {
    uint32_t x = gq[jw * Ip + i] ^ 0x55555555u;
    x |= x >> 1;
    x |= x >> 2;
    uint32_t valid = 0x11111111u;  // nibbles of blocks < J
    if (jw * 8 + 8 > J) valid = J > jw * 8 ? (0x11111111u >> (4 * (int)(jw * 8 + 8 - J))) : 0u;
    n += (unsigned)__builtin_popcount(~x & valid);
  }
  if (n) atomicAdd(&dm[i], (unsigned long long)n);
}

// Double-missing counts of SNP pairs straight from BED columns (the part of a file past the fit's blocks):
// raw holds nb blocks = 2 nb SNP columns of bps bytes; thread (i, chunk of DM_CW blocks) adds its count to
// dm[i] with an integer atomic (exact in any order).
__global__ void k_count_dm_raw(const uint8_t* __restrict__ raw, int64_t bps, int64_t I, int64_t nb,
                               unsigned long long* __restrict__ dm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= I) return;
  const int64_t b0 = (int64_t)blockIdx.y * DM_CW, b1 = b0 + DM_CW < nb ? b0 + DM_CW : nb;
  const unsigned sh = (unsigned)(i & 3) * 2u;
  unsigned n = 0;
  for (int64_t b = b0; b < b1; ++b) {
    const unsigned c1 = (raw[(2 * b) * bps + (i >> 2)] >> sh) & 3u;
    const unsigned c2 = (raw[(2 * b + 1) * bps + (i >> 2)] >> sh) & 3u;
    n += (c1 == 1u && c2 == 1u) ? 1u : 0u;
  }
  if (n) atomicAdd(&dm[i], (unsigned long long)n);
}

// Decoded genotype readback (bit-exact decode check): g[s * I + i] = BED code of SNP s.
__global__ void k_decode(const uint32_t* __restrict__ gq, int64_t Ip, int64_t I, int64_t J,
                         uint8_t* __restrict__ codes) {
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < I * J;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = idx / I, i = idx - j * I;
    const uint32_t w = gq[(j >> 3) * Ip + i];
    const unsigned nib = (w >> (4 * (j & 7))) & 15u;
    codes[(2 * j) * I + i] = (uint8_t)(nib & 3u);
    codes[(2 * j + 1) * I + i] = (uint8_t)(nib >> 2);
  }
}

// -----------------------------------------------------------------------------------------------
// Q side (individual-parallel): log-likelihood and EM / gradient sums
// -----------------------------------------------------------------------------------------------

template <int K, int MODE>
struct NAcc {
  static constexpr int v = MODE == MODE_LL ? 1 : (MODE == MODE_GRAD ? K : K + K * (K + 1) / 2);
};

// Lane = individual; the block's 4K p values are wave-uniform (scalar loads).  MODE_LL: the per-pair
// likelihoods (>= ~1e-20) are multiplied into an FP64 mantissa that is renormalised once per genotype
// word (8 pairs, >= ~1e-160: no underflow) with an integer exponent; one log per lane at the end, so
// rounding grows like sqrt(J) ulps instead of J logs' worth.  MODE_GRAD: Xtz_q[k, i] = sum_{j,h} w_h p_kh.
// In FP64 at K <= 8 the likelihood mode takes the factored pair likelihood at 8 waves per SIMD (the pass
// waits on its scalar p loads; left to itself the compiler holds more loads in flight at half the
// occupancy, which is slower); above K = 8, and in FP32 (which would spill at that occupancy), the general
// form at the compiler's occupancy is faster.
template <int K, typename TP, int MODE>
struct QsideLik {
  static constexpr bool factored = MODE == MODE_LL && K <= 8 && sizeof(TP) == 8;
};

template <int K, typename TP, int MODE>
__global__ __launch_bounds__(256, (QsideLik<K, TP, MODE>::factored ? 8 : 1)) void k_qside(const uint32_t* __restrict__ gq, int64_t Ip, int64_t I, int J8,
                                               int wpc, const TP* __restrict__ q, const TP* __restrict__ p,
                                               double* __restrict__ part, int flush, const int* __restrict__ gate) {
  if (gate_off(gate)) return;
  constexpr int NT = NAcc<K, MODE>::v;
  constexpr bool F32 = sizeof(TP) == 4;
  constexpr bool FACT = QsideLik<K, TP, MODE>::factored;
  __shared__ __attribute__((aligned(16))) TP lut[FACT ? 160 : 128];
  if constexpr (FACT)
    fill_lik_lut<TP>(lut);
  else
    fill_pair_lut<TP>(lut);
  __syncthreads();
  unsigned bx, by;
  xcd_grouped_block(bx, by);  // the workgroups of a chunk read its p rows through one L2
  const int64_t i = (int64_t)bx * blockDim.x + threadIdx.x;
  const int chunk = by;
  const int jw0 = chunk * wpc;
  const int jw1 = min(J8, jw0 + wpc);
  const int64_t ic = i < I ? i : I - 1;
  TP qi[K];
#pragma unroll
  for (int k = 0; k < K; ++k) qi[k] = q[k + (int64_t)K * ic];

  double acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = 0.0;
  double llm = 1.0;
  int lle = 0, nfl = 0;
  TP pa[NT];  // FP32 partials (MIXED), flushed to acc every `flush` words
#pragma unroll
  for (int t = 0; t < NT; ++t) pa[t] = (TP)0;

  for (int jw = jw0; jw < jw1; ++jw) {
    const uint32_t word = gq[(int64_t)jw * Ip + i];
#pragma unroll
    for (int n = 0; n < 8; ++n) {
      const TP* pj = p + (int64_t)K * 4 * (jw * 8 + n);
      TP P[4][K];
#pragma unroll
      for (int h = 0; h < 4; ++h)
#pragma unroll
        for (int k = 0; k < K; ++k) P[h][k] = pj[h * K + k];
      TP X[4];
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        TP s = (TP)0;
#pragma unroll
        for (int k = 0; k < K; ++k) s = hk_fma(qi[k], P[h][k], s);
        X[h] = s;
      }
      const unsigned nib = (word >> (4 * n)) & 15u;
      if constexpr (MODE == MODE_LL) {
        // FP64 running product; renormalised by frexp per word
        llm *= (double)(FACT ? lik_factored<TP>(X, lut, nib) : pair_lik<TP>(X, lut, nib));
      } else {
        const PairForm<TP> f(X, lut, nib);
        const TP rl = hk_rcp(f.L);
        TP w[4];
#pragma unroll
        for (int h = 0; h < 4; ++h) w[h] = f.Y[h] * rl;
#pragma unroll
        for (int k = 0; k < K; ++k) {
          TP g = pa[k];
#pragma unroll
          for (int h = 0; h < 4; ++h) g = hk_fma(w[h], P[h][k], g);
          pa[k] = g;
        }
      }
    }
    if constexpr (MODE == MODE_LL) {
      int e;
      llm = frexp(llm, &e);
      lle += e;
    } else if constexpr (F32) {
      if (++nfl == flush) {
        nfl = 0;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          acc[t] += (double)pa[t];
          pa[t] = (TP)0;
        }
      }
    }
  }
  if constexpr (MODE == MODE_LL) {
    acc[0] = log(llm) + (double)lle * 0.69314718055994530942;
  } else {
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = 2.0 * (acc[t] + (double)pa[t]);  // w_h = 2 Y_h / L'
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) part[((int64_t)chunk * NT + t) * Ip + i] = acc[t];
}

// -----------------------------------------------------------------------------------------------
// P side (block-parallel): EM / gradient sums Xtz_p[k, 4j+h] = sum_i w_h q_ki
// -----------------------------------------------------------------------------------------------

// Lane = block with its 4K p values in registers; individual i's q column is wave-uniform (scalar
// loads); partial sums over an i chunk go to out[chunk][4Jp][K].
template <int K, typename TP>
__global__ __launch_bounds__(256) void k_pside_grad(const uint32_t* __restrict__ gp, int64_t Jp, int64_t I, int I8,
                                                    int wpc, const TP* __restrict__ q, const TP* __restrict__ p,
                                                    double* __restrict__ out, int flush, const int* __restrict__ gate) {
  if (gate_off(gate)) return;
  constexpr bool F32 = sizeof(TP) == 4;
  __shared__ __attribute__((aligned(16))) TP lut[128];
  fill_pair_lut<TP>(lut);
  __syncthreads();
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= Jp) return;  // the grid covers Jp (a multiple of 64) in workgroups of 256
  const int chunk = blockIdx.y;
  const int iw0 = chunk * wpc;
  const int iw1 = min(I8, iw0 + wpc);
  TP P[4][K];
#pragma unroll
  for (int h = 0; h < 4; ++h)
#pragma unroll
    for (int k = 0; k < K; ++k) P[h][k] = p[k + (int64_t)K * (4 * j + h)];
  double acc[4][K];
  TP pa[4][K];
  int nfl = 0;
#pragma unroll
  for (int h = 0; h < 4; ++h)
#pragma unroll
    for (int k = 0; k < K; ++k) {
      acc[h][k] = 0.0;
      pa[h][k] = (TP)0;
    }
  for (int iw = iw0; iw < iw1; ++iw) {
    const uint32_t word = gp[(int64_t)iw * Jp + j];
#pragma unroll
    for (int n = 0; n < 8; ++n) {
      int64_t ii = (int64_t)iw * 8 + n;  // padding individuals carry the double-missing nibble
      ii = ii < I ? ii : I - 1;
      const TP* qv = q + (int64_t)K * ii;
      TP Q[K];
#pragma unroll
      for (int k = 0; k < K; ++k) Q[k] = qv[k];
      TP X[4];
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        TP s = (TP)0;
#pragma unroll
        for (int k = 0; k < K; ++k) s = hk_fma(Q[k], P[h][k], s);
        X[h] = s;
      }
      const PairForm<TP> f(X, lut, (word >> (4 * n)) & 15u);
      const TP rl = hk_rcp(f.L);
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        const TP w = f.Y[h] * rl;
#pragma unroll
        for (int k = 0; k < K; ++k) pa[h][k] = hk_fma(w, Q[k], pa[h][k]);
      }
    }
    if constexpr (F32) {
      if (++nfl == flush) {
        nfl = 0;
#pragma unroll
        for (int h = 0; h < 4; ++h)
#pragma unroll
          for (int k = 0; k < K; ++k) {
            acc[h][k] += (double)pa[h][k];
            pa[h][k] = (TP)0;
          }
      }
    }
  }
#pragma unroll
  for (int h = 0; h < 4; ++h) {
    double* o = out + ((int64_t)chunk * 4 * Jp + 4 * j + h) * K;
#pragma unroll
    for (int k = 0; k < K; ++k) o[k] = 2.0 * (acc[h][k] + (double)pa[h][k]);  // w_h = 2 Y_h / L'
  }
}

// -----------------------------------------------------------------------------------------------
// Curvature passes on the matrix cores
//
// XtX is a GEMM with the per-pair weights as one operand:
//     XtX_p[jh, n] = sum_i  v[i, jh] * QQ[i, n]      QQ[i, n] = 2 q[k,i] q[k',i], n = k(k+1)/2 + k'
//     XtX_q[i, n]  = sum_jh v[i, jh] * PP[jh, n]     PP[jh, n] = 2 p[k,jh] p[k',jh]
// with v_h = Y_h / (L X_h) = c_h / (2 qp_h^2) (pair_v; the factor 2 is in the tables).  One lane
// computes one (i, j) pair with all four haplotype weights (no redundant per-pair work);
// v_mfma_{f32,f64}_16x16x4 takes A = v_h (16 rows x 4), B = QQ or PP (4 x 16 columns): one MFMA per
// haplotype and column tile.  gfx950 overlaps FP32 MFMA and VALU issue little, and one
// v_mfma_f32_16x16x4f32 holds the SIMD as long as ~12 FMAs, so a short tail of packed columns (<= 6,
// e.g. 16..20 for K = 6) is accumulated with VALU FMAs instead of a mostly empty extra tile (CurvShape).
//   A operand: lane l holds A[l % 16][l / 16];  B operand: lane l holds B[l / 16][l % 16];
//   result: lane l, register v holds D[row][l % 16], row = 4 (l/16) + v (f32) or l/16 + 4 v (f64).
// MIXED: FP32 products and FP32 accumulation over 64 individuals (P side) / 64 blocks (Q side), then
// added to FP64 accumulators.  FP64: FP64 throughout.
// -----------------------------------------------------------------------------------------------


// Column split of the packed curvature (NKK = K(K+1)/2 entries): NTILE MFMA tiles of 16, then NR VALU
// tail columns starting at TAIL0; QS = padded length of one packed row.  Staging tables (k_stage_rows):
//   P side, one row per individual: [2 q_k q_k' (QS) | q_k (K) | 0], RGP entries (a power of two);
//   Q side, one row per block:      [2 p_kh p_k'h (QS), h = 0..3 | p_kh (4K, h-major) | 0], RGQ entries.
// In LDS the rows are RLP / RLQ apart: = 16 (mod 64) words when MFMA B operands are read from them, so
// the four lane groups of a wave (16 consecutive words each, different rows) fall on distinct banks.
template <int K>
struct CurvShape {
  static constexpr int NKK = K * (K + 1) / 2;
  static constexpr int FULL = NKK / 16, REM = NKK - 16 * FULL;
  static constexpr bool VTAIL = REM > 0 && REM <= 6;
  static constexpr int NTILE = FULL + (REM > 0 && !VTAIL ? 1 : 0);
  static constexpr int NR = VTAIL ? REM : 0;
  static constexpr int TAIL0 = 16 * NTILE;
  static constexpr int QS = TAIL0 + (NR + 3) / 4 * 4;
  static constexpr int RGP = pow2_at_least(QS + K), RGQ = pow2_at_least(4 * QS + 4 * K);
  static constexpr int RLP = NTILE > 0 ? RGP + 16 : RGP, RLQ = NTILE > 0 ? RGQ + 16 : RGQ;
  static constexpr int PAD_ROWS = 32;  // rows past the last column (the last staging step reads up to 32)
};

// Sum of a value over the four lanes l, l ^ 16, l ^ 32, l ^ 48 (same l % 16); bitwise equal in all four.
__device__ __forceinline__ double sum_groups(double x) {
  x += __shfl_xor(x, 16);
  x += __shfl_xor(x, 32);
  return x;
}

// Reduce-scatter of FP32 tail partials over the four lanes of a block on the P side (lane l = 16 g + m
// holds A[h][n] summed over its individual group g): afterwards lane g adds the four-lane total of
// haplotype h = g to DA[n].  v_permlane32_swap exchanges the lane halves of two registers (g <-> g ^ 2),
// v_permlane16_swap the odd and even rows (g <-> g ^ 1); after each swap a lane holds its own and its
// partner's value of the haplotype it keeps, so one add per kept value and no selects.
template <int NR>
__device__ __forceinline__ void tail_flush_p(float (&A)[4][NR], double (&DA)[NR]) {
#pragma unroll
  for (int n = 0; n < NR; ++n) {
    const auto r02 = __builtin_amdgcn_permlane32_swap(__float_as_uint(A[0][n]), __float_as_uint(A[2][n]), false, false);
    const auto r13 = __builtin_amdgcn_permlane32_swap(__float_as_uint(A[1][n]), __float_as_uint(A[3][n]), false, false);
    const float b0 = __uint_as_float(r02[0]) + __uint_as_float(r02[1]);  // h = 2 (g >> 1)
    const float b1 = __uint_as_float(r13[0]) + __uint_as_float(r13[1]);  // h = 2 (g >> 1) + 1
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(b0), __float_as_uint(b1), false, false);
    DA[n] += (double)(__uint_as_float(r[0]) + __uint_as_float(r[1]));  // h = g
#pragma unroll
    for (int h = 0; h < 4; ++h) A[h][n] = 0.f;
  }
}

// P side: a wave owns 16 blocks (A rows); each MFMA step takes 4 individuals (A columns / B rows).
// The workgroup (64 blocks) stages q and the q q' products of 32 individuals (4 genotype words) in LDS.
// PF: staging double-buffered through registers (the global loads of stage s+1 -- q, q q' and the
// genotype words -- are issued before stage s is computed and written to LDS after it); without PF the
// loads of a stage are issued together at its start, which keeps fewer registers live (K >= 7).
template <int K, typename TP, bool PF>
__device__ __forceinline__ void pside_mm_body(const uint32_t* __restrict__ gp, int64_t Jp, int64_t I, int I8, int wpc,
                                              const TP* __restrict__ q, const TP* __restrict__ p,
                                              const TP* __restrict__ qq, double* __restrict__ out, int flush) {
  using S = CurvShape<K>;
  constexpr int NKK = S::NKK, NT = K + NKK, NTILE = S::NTILE, NR = S::NR, QS = S::QS, T0 = S::TAIL0;
  constexpr int RG = S::RGP, RL = S::RLP;
  constexpr int NTA = NTILE > 0 ? NTILE : 1, NRA = NR > 0 ? NR : 1;
  constexpr int WPS = 4;  // genotype words per LDS stage (32 individuals)
  constexpr int LQQ = 8 * WPS * RG / 256;
  constexpr bool F32 = sizeof(TP) == 4;
  using M = Mfma<TP>;
  using V4 = typename M::v4;
  __shared__ __attribute__((aligned(16))) TP qrow[8 * WPS * RL];  // staged table rows of 32 individuals
  __shared__ __attribute__((aligned(16))) TP lut[128];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int m = lane & 15, g = lane >> 4;
  const int64_t jbase = ((int64_t)blockIdx.x * 4 + wv) * 16;
  const int64_t j = jbase + m;
  const int chunk = blockIdx.y;
  const int iw0 = chunk * wpc, iw1 = min(I8, iw0 + wpc);
  fill_pair_lut<TP>(lut);
  TP P[4][K];
#pragma unroll
  for (int h = 0; h < 4; ++h)
#pragma unroll
    for (int k = 0; k < K; ++k) P[h][k] = p[k + (int64_t)K * (4 * j + h)];
  V4 C[4][NTA];
  double D[4][NTA][4];
  TP A[4][NRA];    // tail partials of this lane's individuals (FP64 engine: the full sums)
  double DA[NRA];  // MIXED: four-lane tail totals of haplotype h = g
#pragma unroll
  for (int h = 0; h < 4; ++h) {
#pragma unroll
    for (int t = 0; t < NTA; ++t) {
      C[h][t] = V4{0, 0, 0, 0};
#pragma unroll
      for (int v = 0; v < 4; ++v) D[h][t][v] = 0.0;
    }
#pragma unroll
    for (int n = 0; n < NRA; ++n) A[h][n] = (TP)0;
  }
#pragma unroll
  for (int n = 0; n < NRA; ++n) DA[n] = 0.0;
  // registers of the next stage: the 8 WPS table rows from individual 8 iws_ on (the table has
  // PAD_ROWS rows past Ip, so a chunk's last stage stays inside it)
  TP rqq[LQQ];
  uint32_t wnext[WPS] = {};
  auto fetch = [&](int iws_) {
    const TP* base = qq + (int64_t)iws_ * 8 * RG;
    const int nw_ = min(WPS, iw1 - iws_);
#pragma unroll
    for (int u = 0; u < LQQ; ++u) rqq[u] = base[threadIdx.x + 256 * u];
    if constexpr (PF) {
#pragma unroll
      for (int w = 0; w < WPS; ++w) wnext[w] = w < nw_ ? gp[(int64_t)(iws_ + w) * Jp + j] : 0u;
    }
  };
  if (PF && iw0 < iw1) fetch(iw0);
  int nword = 0;
  for (int iws = iw0; iws < iw1; iws += WPS) {
    const int nw = min(WPS, iw1 - iws);
    __syncthreads();
    if (!PF) fetch(iws);
#pragma unroll
    for (int u = 0; u < LQQ; ++u) {
      const int e = threadIdx.x + 256 * u;
      qrow[(e / RG) * RL + (e % RG)] = rqq[u];
    }
    uint32_t wcur[WPS];
#pragma unroll
    for (int w = 0; w < WPS; ++w) wcur[w] = wnext[w];
    __syncthreads();
    if (PF && iws + WPS < iw1) fetch(iws + WPS);
    auto do_word = [&](int w, uint32_t word) {
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const TP* row = qrow + (8 * w + 4 * s + g) * RL;  // staged individual of this lane's A column
        TP X[4];
#pragma unroll
        for (int h = 0; h < 4; ++h) {
          TP x = (TP)0;
#pragma unroll
          for (int k = 0; k < K; ++k) x = hk_fma(row[QS + k], P[h][k], x);
          X[h] = x;
        }
        TP v[4];
        pair_v<TP>(X, lut, nib_at(word, 4 * (4 * s + g)), v);
        if constexpr (NTILE > 0) {
          TP b[NTA];
#pragma unroll
          for (int t = 0; t < NTILE; ++t) b[t] = row[16 * t + m];
#pragma unroll
          for (int h = 0; h < 4; ++h)
#pragma unroll
            for (int t = 0; t < NTILE; ++t) C[h][t] = M::mma(v[h], b[t], C[h][t]);
        }
        if constexpr (NR > 0) {
          TP tq[NRA];
#pragma unroll
          for (int n = 0; n < NR; ++n) tq[n] = row[T0 + n];
#pragma unroll
          for (int h = 0; h < 4; ++h)
#pragma unroll
            for (int n = 0; n < NR; ++n) A[h][n] = hk_fma(v[h], tq[n], A[h][n]);
        }
      }
      if constexpr (F32) {
        if (++nword == flush) {
          nword = 0;
#pragma unroll
          for (int h = 0; h < 4; ++h)
#pragma unroll
            for (int t = 0; t < NTILE; ++t) {
#pragma unroll
              for (int v = 0; v < 4; ++v) D[h][t][v] += (double)C[h][t][v];
              C[h][t] = V4{0, 0, 0, 0};
            }
          if constexpr (NR > 0) tail_flush_p<NR>(A, DA);
        }
      }
    };
    if constexpr (PF) {
#pragma unroll
      for (int w = 0; w < WPS; ++w)
        if (w < nw) do_word(w, wcur[w]);
    } else {
      for (int w = 0; w < nw; ++w) do_word(w, gp[(int64_t)(iws + w) * Jp + j]);
    }
  }
  // gradient slots (formed by the finish kernel): lane (m, g) clears those of (block jbase + m, h = g)
  {
    double* o = out + ((int64_t)chunk * 4 * Jp + 4 * j + g) * NT;
#pragma unroll
    for (int k = 0; k < K; ++k) o[k] = 0.0;
  }
#pragma unroll
  for (int h = 0; h < 4; ++h)
#pragma unroll
    for (int t = 0; t < NTILE; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const double val = F32 ? D[h][t][v] + (double)C[h][t][v] : (double)C[h][t][v];
        const int n = 16 * t + m;
        const int64_t row = jbase + M::row(g, v);  // block of this result element
        if (n < NKK) out[((int64_t)chunk * 4 * Jp + 4 * row + h) * NT + K + n] = val;
      }
  // VALU tail: the four lanes of a block hold sums over different individuals; lane g writes h = g
  if constexpr (NR > 0) {
    double* o = out + ((int64_t)chunk * 4 * Jp + 4 * j + g) * NT + K + T0;
    if constexpr (F32) {
      tail_flush_p<NR>(A, DA);
#pragma unroll
      for (int n = 0; n < NR; ++n) o[n] = DA[n];
    } else {
#pragma unroll
      for (int h = 0; h < 4; ++h)
#pragma unroll
        for (int n = 0; n < NR; ++n) {
          const double tot = sum_groups((double)A[h][n]);
          if (g == h) o[n] = tot;
        }
    }
  }
}

// MIXED, K <= 6: prefetching variant held to 3 waves per SIMD (168 VGPRs); K >= 7 and FP64 need more
// registers and use the variant without prefetch.
template <int K, typename TP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_pside_mm_pf(
    const uint32_t* __restrict__ gp, int64_t Jp, int64_t I, int I8, int wpc, const TP* __restrict__ q,
    const TP* __restrict__ p, const TP* __restrict__ qq, double* __restrict__ out, int flush, const int* __restrict__ gate) {
  if (gate_off(gate)) return;
  pside_mm_body<K, TP, true>(gp, Jp, I, I8, wpc, q, p, qq, out, flush);
}
// waves_per_eu(2) keeps K = 8 at two waves per SIMD (and K = 7 at three) without spills.
template <int K, typename TP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_pside_mm(const uint32_t* __restrict__ gp, int64_t Jp, int64_t I, int I8,
                                                  int wpc, const TP* __restrict__ q, const TP* __restrict__ p,
                                                  const TP* __restrict__ qq, double* __restrict__ out, int flush, const int* __restrict__ gate) {
  if (gate_off(gate)) return;
  pside_mm_body<K, TP, false>(gp, Jp, I, I8, wpc, q, p, qq, out, flush);
}

// Q side: a wave owns 2 x 16 individuals (two A row tiles); each MFMA step takes 4 blocks (A columns /
// B rows).  The workgroup (128 individuals) stages p and the p p' products of 16 blocks (2 words) in LDS;
// PF as on the P side.
template <int K, typename TP, bool PF>
__device__ __forceinline__ void qside_mm_body(const uint32_t* __restrict__ gq, int64_t Ip, int64_t I, int J8, int wpc,
                                              const TP* __restrict__ q, const TP* __restrict__ p,
                                              const TP* __restrict__ pp, double* __restrict__ part, int flush) {
  using S = CurvShape<K>;
  constexpr int NKK = S::NKK, NT = K + NKK, NTILE = S::NTILE, NR = S::NR, QS = S::QS, T0 = S::TAIL0;
  constexpr int RG = S::RGQ, RL = S::RLQ;
  constexpr int NTA = NTILE > 0 ? NTILE : 1, NRA = NR > 0 ? NR : 1;
  constexpr int WPS = 2, R = 2;  // genotype words per LDS stage (16 blocks), row tiles per wave
  constexpr int LPP = 8 * WPS * RG / 256;
  constexpr bool F32 = sizeof(TP) == 4;
  using M = Mfma<TP>;
  using V4 = typename M::v4;
  __shared__ __attribute__((aligned(16))) TP prow[8 * WPS * RL];  // staged table rows of 16 blocks
  __shared__ __attribute__((aligned(16))) TP lut[128];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int m = lane & 15, g = lane >> 4;
  unsigned bx, by;
  xcd_grouped_block(bx, by);
  const int64_t ibase = ((int64_t)bx * 4 + wv) * 16 * R;
  const int chunk = by;
  const int jw0 = chunk * wpc, jw1 = min(J8, jw0 + wpc);
  fill_pair_lut<TP>(lut);
  TP qi[R][K];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    int64_t ic = ibase + 16 * r + m;
    ic = ic < I ? ic : I - 1;
#pragma unroll
    for (int k = 0; k < K; ++k) qi[r][k] = q[k + (int64_t)K * ic];
  }
  V4 C[R][NTA];
  double D[R][NTA][4];
  TP A[R][NRA];
  double DA[R][NRA];
#pragma unroll
  for (int r = 0; r < R; ++r) {
#pragma unroll
    for (int t = 0; t < NTA; ++t) {
      C[r][t] = V4{0, 0, 0, 0};
#pragma unroll
      for (int v = 0; v < 4; ++v) D[r][t][v] = 0.0;
    }
#pragma unroll
    for (int n = 0; n < NRA; ++n) {
      A[r][n] = (TP)0;
      DA[r][n] = 0.0;
    }
  }
  // registers of the next stage: the 8 WPS table rows from block 8 jws_ on (the table has PAD_ROWS
  // rows past Jp, so a chunk's last stage stays inside it)
  TP rpp[LPP];
  uint32_t wnext[WPS][R] = {};
  auto fetch = [&](int jws_) {
    const int nw_ = min(WPS, jw1 - jws_);
    const TP* base = pp + (int64_t)jws_ * 8 * RG;
#pragma unroll
    for (int u = 0; u < LPP; ++u) rpp[u] = base[threadIdx.x + 256 * u];
    if constexpr (PF) {
#pragma unroll
      for (int w = 0; w < WPS; ++w)
#pragma unroll
        for (int r = 0; r < R; ++r) wnext[w][r] = w < nw_ ? gq[(int64_t)(jws_ + w) * Ip + ibase + 16 * r + m] : 0u;
    }
  };
  if (PF && jw0 < jw1) fetch(jw0);
  int nword = 0;
  for (int jws = jw0; jws < jw1; jws += WPS) {
    const int nw = min(WPS, jw1 - jws);
    __syncthreads();
    if (!PF) fetch(jws);
#pragma unroll
    for (int u = 0; u < LPP; ++u) {
      const int e = threadIdx.x + 256 * u;
      prow[(e / RG) * RL + (e % RG)] = rpp[u];
    }
    uint32_t wcur[WPS][R];
#pragma unroll
    for (int w = 0; w < WPS; ++w)
#pragma unroll
      for (int r = 0; r < R; ++r) wcur[w][r] = wnext[w][r];
    __syncthreads();
    if (PF && jws + WPS < jw1) fetch(jws + WPS);
    auto do_word = [&](int w, const uint32_t (&word)[R]) {
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const TP* row = prow + (8 * w + 4 * s + g) * RL;  // staged block of this lane's A column
        TP b[4][NTA];
#pragma unroll
        for (int h = 0; h < 4; ++h)
#pragma unroll
          for (int t = 0; t < NTILE; ++t) b[h][t] = row[h * QS + 16 * t + m];
        TP tp[4][NRA];
#pragma unroll
        for (int h = 0; h < 4; ++h)
#pragma unroll
          for (int n = 0; n < NR; ++n) tp[h][n] = row[h * QS + T0 + n];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          TP X[4];
#pragma unroll
          for (int h = 0; h < 4; ++h) {
            TP x = (TP)0;
#pragma unroll
            for (int k = 0; k < K; ++k) x = hk_fma(qi[r][k], row[4 * QS + h * K + k], x);
            X[h] = x;
          }
          TP v[4];
          pair_v<TP>(X, lut, nib_at(word[r], 4 * (4 * s + g)), v);
#pragma unroll
          for (int h = 0; h < 4; ++h)
#pragma unroll
            for (int t = 0; t < NTILE; ++t) C[r][t] = M::mma(v[h], b[h][t], C[r][t]);
#pragma unroll
          for (int n = 0; n < NR; ++n) {
            TP a = A[r][n];
#pragma unroll
            for (int h = 0; h < 4; ++h) a = hk_fma(v[h], tp[h][n], a);
            A[r][n] = a;
          }
        }
      }
      if constexpr (F32) {
        if (++nword == flush) {
          nword = 0;
#pragma unroll
          for (int r = 0; r < R; ++r) {
#pragma unroll
            for (int t = 0; t < NTILE; ++t) {
#pragma unroll
              for (int v = 0; v < 4; ++v) D[r][t][v] += (double)C[r][t][v];
              C[r][t] = V4{0, 0, 0, 0};
            }
#pragma unroll
            for (int n = 0; n < NR; ++n) {
              DA[r][n] += (double)A[r][n];
              A[r][n] = (TP)0;
            }
          }
        }
      }
    };
    if constexpr (PF) {
#pragma unroll
      for (int w = 0; w < WPS; ++w)
        if (w < nw) do_word(w, wcur[w]);
    } else {
      for (int w = 0; w < nw; ++w) {
        uint32_t word[R];
#pragma unroll
        for (int r = 0; r < R; ++r) word[r] = gq[(int64_t)(jws + w) * Ip + ibase + 16 * r + m];
        do_word(w, word);
      }
    }
  }
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int t = 0; t < NTILE; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const double val = F32 ? D[r][t][v] + (double)C[r][t][v] : (double)C[r][t][v];
        const int n = 16 * t + m;
        const int64_t row = ibase + 16 * r + M::row(g, v);  // individual of this result element
        if (n < NKK) part[((int64_t)chunk * NT + K + n) * Ip + row] = val;
      }
  // VALU tail: the four lanes of an individual hold sums over different blocks
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int n = 0; n < NR; ++n) {
      const double tot = sum_groups(F32 ? DA[r][n] + (double)A[r][n] : (double)A[r][n]);
      if (g == (n & 3)) part[((int64_t)chunk * NT + K + T0 + n) * Ip + ibase + 16 * r + m] = tot;
    }
}

template <int K, typename TP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_qside_mm_pf(
    const uint32_t* __restrict__ gq, int64_t Ip, int64_t I, int J8, int wpc, const TP* __restrict__ q,
    const TP* __restrict__ p, const TP* __restrict__ pp, double* __restrict__ part, int flush, const int* __restrict__ gate) {
  if (gate_off(gate)) return;
  qside_mm_body<K, TP, true>(gq, Ip, I, J8, wpc, q, p, pp, part, flush);
}
template <int K, typename TP>
__global__ __launch_bounds__(256) void k_qside_mm(const uint32_t* __restrict__ gq, int64_t Ip, int64_t I, int J8,
                                                  int wpc, const TP* __restrict__ q, const TP* __restrict__ p,
                                                  const TP* __restrict__ pp, double* __restrict__ part, int flush, const int* __restrict__ gate) {
  if (gate_off(gate)) return;
  qside_mm_body<K, TP, false>(gq, Ip, I, J8, wpc, q, p, pp, part, flush);
}

// -----------------------------------------------------------------------------------------------
// Reductions over chunks (fixed order)
// -----------------------------------------------------------------------------------------------

// part[c][t][Ip] -> res[i][NT] for i < I
// Sum over chunks of slot t of individual i, chunks in order (several loads in flight).
__device__ __forceinline__ double chunk_sum(const double* __restrict__ part, int nchunk, int64_t stride, int64_t off) {
  double s = 0.0;
  int c = 0;
  for (; c + 4 <= nchunk; c += 4) {
    const double a0 = part[(int64_t)c * stride + off], a1 = part[(int64_t)(c + 1) * stride + off];
    const double a2 = part[(int64_t)(c + 2) * stride + off], a3 = part[(int64_t)(c + 3) * stride + off];
    s += a0;
    s += a1;
    s += a2;
    s += a3;
  }
  for (; c < nchunk; ++c) s += part[(int64_t)c * stride + off];
  return s;
}

// part[c][t][Ip] -> res[i][NT] for i < I, slots t0 + blockIdx.y (thread per (i, t)).
// acc32: the chunk partials are summed in FP32 (the reference's Float32 accumulators).
__global__ void k_reduce_q(const double* __restrict__ part, int nchunk, int NT, int t0, int64_t Ip, int64_t I,
                           double* __restrict__ res, int acc32, const int* __restrict__ gate) {
  if (gate_off(gate)) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int t = t0 + blockIdx.y;
  if (i >= I) return;
  const int64_t stride = (int64_t)NT * Ip, off = (int64_t)t * Ip + i;
  if (acc32) {
    float s = 0.f;
    for (int c = 0; c < nchunk; ++c) s += (float)part[(int64_t)c * stride + off];
    res[i * NT + t] = s;
  } else {
    res[i * NT + t] = chunk_sum(part, nchunk, stride, off);
  }
}

// Curvature pass: Xtz = XtX q (FP64 q) into slots 0..K-1 of res[i][NT], after k_reduce_q filled the XtX slots.
__global__ void k_xtz_q(int K, int64_t I, const double* __restrict__ q, double* __restrict__ res, const int* __restrict__ gate) {
  if (gate_off(gate)) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= I) return;
  const int NT = K + K * (K + 1) / 2;
  double* r = res + i * NT;
  for (int k = 0; k < K; ++k) {
    double s = 0.0;
    for (int k2 = 0; k2 < K; ++k2) {
      const int a = k > k2 ? k : k2, b = k > k2 ? k2 : k;
      s += r[K + a * (a + 1) / 2 + b] * q[k2 + (int64_t)K * i];
    }
    r[k] = s;
  }
}

// Curvature pass, P side: Xtz_p[:, jh] = XtX_p[:, :, jh] p[:, jh] into slots 0..K-1 of res[jh][NT].
// With flags: also marks the blocks whose P row QP the four-lane solver (k_p_step4) would hand to the
// dense one: a free haplotype block of the curvature with a Cholesky pivot below 1e-6 of its largest
// diagonal, the test k_p_step4 applies at its first active-set iteration, whose free set p fixes (and
// free subsets only raise the pivots).  They are listed in dense[1 .. dense[0]] so that the dense solver
// can run alongside the four-lane kernel instead of after it.  Lane = (block, haplotype): the four
// lanes of a block are adjacent (4J is a multiple of 4), so the quad's result is a DPP maximum.
template <int K>
__global__ void k_finish_p(int64_t J, const double* __restrict__ p, double* __restrict__ res,
                           unsigned char* __restrict__ flags, int* __restrict__ dense, const int* __restrict__ gate) {
  if (gate_off(gate)) return;
  const int64_t jh = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (jh >= 4 * J) return;
  constexpr int NT = K + K * (K + 1) / 2;
  double* r = res + jh * NT;
  double a[K * (K + 1) / 2], pv[K];
#pragma unroll
  for (int t = 0; t < K * (K + 1) / 2; ++t) a[t] = r[K + t];
#pragma unroll
  for (int k = 0; k < K; ++k) pv[k] = p[k + (int64_t)K * jh];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double s = 0.0;
#pragma unroll
    for (int k2 = 0; k2 < K; ++k2) {
      const int hi = k > k2 ? k : k2, lo = k > k2 ? k2 : k;
      s += a[hi * (hi + 1) / 2 + lo] * pv[k2];
    }
    r[k] = s;
  }
  if (flags == nullptr) return;
  bool fr[K];
  double dg = 0.0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double lo = 0.0 - pv[k], hi = 1.0 - pv[k];
    fr[k] = !(lo >= 0.0) && !(hi <= 0.0);
    dg = fmax(dg, a[k * (k + 1) / 2 + k]);
  }
  const double piv_min = 1e-6 * dg + 1e-300;
  double L[K][K], di[K];
  bool pd = true;
#pragma unroll
  for (int c = 0; c < K; ++c) {
    double s = fr[c] ? a[c * (c + 1) / 2 + c] : 1.0;
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
      double v = (fr[rr] && fr[c]) ? a[rr * (rr + 1) / 2 + c] : 0.0;
#pragma unroll
      for (int t = 0; t < c; ++t) v -= L[rr][t] * L[c][t];
      L[rr][c] = v * di[c];
    }
  }
  double bad = pd ? 0.0 : 1.0;
  bad = fmax(bad, dpp_f<0xB1>(bad));
  bad = fmax(bad, dpp_f<0x4E>(bad));
  if ((jh & 3) == 0) {
    const int64_t j = jh >> 2;
    flags[j] = bad > 0.0 ? 1 : 0;
    if (bad > 0.0) dense[1 + atomicAdd(dense, 1)] = (int)j;
  }
}

// part[c][4Jp][NT] -> res[jh][NT] for jh < 4J
__global__ void k_reduce_p(const double* __restrict__ part, int nchunk, int NT, int64_t Jp, int64_t J,
                           double* __restrict__ res, int acc32, const int* __restrict__ gate) {
  if (gate_off(gate)) return;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < 4 * J * NT;
       idx += (int64_t)gridDim.x * blockDim.x) {
    if (acc32) {
      float s = 0.f;
      for (int c = 0; c < nchunk; ++c) s += (float)part[(int64_t)c * 4 * Jp * NT + idx];
      res[idx] = s;
    } else {
      double s = 0.0;
      for (int c = 0; c < nchunk; ++c) s += part[(int64_t)c * 4 * Jp * NT + idx];
      res[idx] = s;
    }
  }
}

// Log-likelihood, stage 1: rows[i] = sum_c part[c][i]; bpart[b] = sum of the rows of workgroup b (tree).
__global__ __launch_bounds__(256) void k_reduce_ll_rows(const double* __restrict__ part, int nchunk, int64_t Ip,
                                                        int64_t I, double* __restrict__ rows,
                                                        double* __restrict__ bpart, const int* __restrict__ gate) {
  if (gate_off(gate)) return;
  __shared__ double sh[256];
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  double r = 0.0;
  if (i < I) {
    r = chunk_sum(part, nchunk, Ip, i);
    if (rows) rows[i] = r;
  }
  r = block_sum<256>(r, sh);
  if (threadIdx.x == 0) bpart[blockIdx.x] = r;
}

// Stage 2: total = sum of the nb workgroup sums (fixed order).
__global__ __launch_bounds__(256) void k_reduce_ll_total(const double* __restrict__ bpart, int nb,
                                                         double* __restrict__ total, const int* __restrict__ gate) {
  if (gate_off(gate)) return;
  __shared__ double sh[256];
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) s += bpart[b];
  s = block_sum<256>(s, sh);
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
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < (int64_t)K * n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = idx / K;
    const int k = (int)(idx - c * K);
    out[idx] = par[idx] * res[c * NT + k];
  }
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
    hipLaunchKernelGGL((k_qside<K, TP, MODE_LL>), grid, block, 0, s, g.gq, g.Ip, g.I, (int)g.J8, g.qwpc, qq, pp, part,
                       g.flush_grad, g.gate);
  else if (mode == MODE_GRAD)
    hipLaunchKernelGGL((k_qside<K, TP, MODE_GRAD>), grid, block, 0, s, g.gq, g.Ip, g.I, (int)g.J8, g.qwpc, qq, pp, part,
                       g.flush_grad, g.gate);
  else if constexpr (K > HK_MAXK_REG) {
    return launch_qside_curv_wide(K, sizeof(TP) == 8 ? PREC_FP64 : PREC_MIXED, g, q, part, s);
  } else {
    const dim3 gmm((unsigned)(g.Ip / 128), (unsigned)g.qchunks);
    // FP32: prefetching staging by default up to K = 6; g.qshape picks the other form (K <= 7)
    bool pf = false;
    if constexpr (sizeof(TP) == 4 && K <= 7) pf = (K <= 6) != (g.qshape != 0);
    if constexpr (sizeof(TP) == 4 && K <= 7) {
      if (pf)
        hipLaunchKernelGGL((k_qside_mm_pf<K, TP>), gmm, block, 0, s, g.gq, g.Ip, g.I, (int)g.J8, g.qwpc, qq, pp,
                           (const TP*)g.pp, part, g.flush_curv, g.gate);
    }
    if (!pf)
      hipLaunchKernelGGL((k_qside_mm<K, TP>), gmm, block, 0, s, g.gq, g.Ip, g.I, (int)g.J8, g.qwpc, qq, pp,
                         (const TP*)g.pp, part, g.flush_curv, g.gate);
  }
  return hipGetLastError();
}

template <int K, typename TP>
static hipError_t pside_k(int mode, const PassGeom& g, const void* q, const void* p, double* out,
                          hipStream_t s) {
  if constexpr (K > HK_MAXK_REG) {
    return launch_pside_wide(K, sizeof(TP) == 8 ? PREC_FP64 : PREC_MIXED, mode, g, q, p, out, s);
  } else {
    const dim3 block(256);
    const TP* qq = (const TP*)q;
    const TP* pp = (const TP*)p;
    if (mode == MODE_GRAD) {
      hipLaunchKernelGGL((k_pside_grad<K, TP>), dim3((unsigned)(g.Jp / 256 + (g.Jp % 256 ? 1 : 0)), (unsigned)g.pchunks),
                         block, 0, s, g.gp, g.Jp, g.I, (int)g.I8, g.pwpc, qq, pp, out, g.flush_grad, g.gate);
    } else {
      const dim3 gmm((unsigned)(g.Jp / 64), (unsigned)g.pchunks);
      // FP32 up to K = 6: prefetching staging by default, g.pshape picks the form without it
      bool pf = false;
      if constexpr (sizeof(TP) == 4 && K <= 6) {
        pf = g.pshape == 0;
        if (pf)
          hipLaunchKernelGGL((k_pside_mm_pf<K, TP>), gmm, block, 0, s, g.gp, g.Jp, g.I, (int)g.I8, g.pwpc, qq, pp,
                             (const TP*)g.qq, out, g.flush_curv, g.gate);
      }
      if (!pf)
        hipLaunchKernelGGL((k_pside_mm<K, TP>), gmm, block, 0, s, g.gp, g.Jp, g.I, (int)g.I8, g.pwpc, qq, pp,
                           (const TP*)g.qq, out, g.flush_curv, g.gate);
    }
    return hipGetLastError();
  }
}

// Staging tables of the curvature passes from the pass parameters: the Q-side pass reads the p table
// (rows per block), the P-side pass the q table (rows per individual).
template <int K, typename TP>
static hipError_t tables_k(int side, const PassGeom& g, const void* x, hipStream_t s) {
  if constexpr (K > HK_MAXK_REG) {
    return launch_curv_tables_wide(K, sizeof(TP) == 8 ? PREC_FP64 : PREC_MIXED, side, g, x, s);
  } else {
    using S = CurvShape<K>;
    if (side == 0) {
      const int64_t nrow = g.Jp + S::PAD_ROWS;
      hipLaunchKernelGGL((k_stage_rows<K, TP>), dim3(grid_1d(nrow * S::RGQ)), dim3(256), 0, s,
                         (const TP*)x, 4 * g.Jp, nrow, 4, S::QS, S::RGQ, (TP*)g.pp, g.gate);
    } else {
      const int64_t nrow = g.Ip + S::PAD_ROWS;
      hipLaunchKernelGGL((k_stage_rows<K, TP>), dim3(grid_1d(nrow * S::RGP)), dim3(256), 0, s,
                         (const TP*)x, g.I, nrow, 1, S::QS, S::RGP, (TP*)g.qq, g.gate);
    }
    return hipGetLastError();
  }
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
    case 9: return FN<9, TP>(__VA_ARGS__);       \
    case 10: return FN<10, TP>(__VA_ARGS__);     \
    case 11: return FN<11, TP>(__VA_ARGS__);     \
    case 12: return FN<12, TP>(__VA_ARGS__);     \
    case 13: return FN<13, TP>(__VA_ARGS__);     \
    case 14: return FN<14, TP>(__VA_ARGS__);     \
    case 15: return FN<15, TP>(__VA_ARGS__);     \
    case 16: return FN<16, TP>(__VA_ARGS__);     \
    default: return hipErrorInvalidValue;        \
  }

hipError_t launch_qside(int K, int prec, int mode, const PassGeom& g, const void* q, const void* p, double* part,
                        hipStream_t s) {
  if (prec == PREC_FP64) { HK_K_SWITCH(qside_k, double, mode, g, q, p, part, s) }
  HK_K_SWITCH(qside_k, float, mode, g, q, p, part, s)
}

hipError_t launch_curv_tables(int K, int prec, int side, const PassGeom& g, const void* q, const void* p,
                              hipStream_t s) {
  const void* x = side == 0 ? p : q;
  if (prec == PREC_FP64) { HK_K_SWITCH(tables_k, double, side, g, x, s) }
  HK_K_SWITCH(tables_k, float, side, g, x, s)
}

hipError_t launch_pside(int K, int prec, int mode, const PassGeom& g, const void* q, const void* p, double* out,
                        hipStream_t s) {
  if (prec == PREC_FP64) { HK_K_SWITCH(pside_k, double, mode, g, q, p, out, s) }
  HK_K_SWITCH(pside_k, float, mode, g, q, p, out, s)
}

template <int K>
static void stage_sizes_k(int64_t Ip, int64_t Jp, int64_t* nq, int64_t* np) {
  using S = CurvShape<K>;
  *nq = (Ip + S::PAD_ROWS) * S::RGP;
  *np = (Jp + S::PAD_ROWS) * S::RGQ;
}

void stage_sizes(int K, int64_t Ip, int64_t Jp, int64_t* nq, int64_t* np) {
  switch (K) {
    case 1: return stage_sizes_k<1>(Ip, Jp, nq, np);
    case 2: return stage_sizes_k<2>(Ip, Jp, nq, np);
    case 3: return stage_sizes_k<3>(Ip, Jp, nq, np);
    case 4: return stage_sizes_k<4>(Ip, Jp, nq, np);
    case 5: return stage_sizes_k<5>(Ip, Jp, nq, np);
    case 6: return stage_sizes_k<6>(Ip, Jp, nq, np);
    case 7: return stage_sizes_k<7>(Ip, Jp, nq, np);
    case 8: return stage_sizes_k<8>(Ip, Jp, nq, np);
    default: return stage_sizes_wide(K, Ip, Jp, nq, np);
  }
}

int nacc(int K, int mode) { return mode == MODE_LL ? 1 : (mode == MODE_GRAD ? K : K + K * (K + 1) / 2); }

hipError_t launch_repack(const uint8_t* raw, int64_t bps, const PassGeom& g, hipStream_t s) {
  int64_t nq = g.J8 * g.Ip, nt = g.I8 * g.J8;
  hipLaunchKernelGGL(k_repack_q, dim3(grid_1d(nq)), dim3(256), 0, s, raw, bps, g.I, g.J, g.Ip, g.J8, (uint32_t*)g.gq);
  hipLaunchKernelGGL(k_repack_p, dim3(grid_1d(nt)), dim3(256), 0, s, g.gq, g.Ip, g.J8, g.Jp, g.I8, (uint32_t*)g.gp);
  return hipGetLastError();
}

__global__ void k_zero_u64(unsigned long long* __restrict__ p, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0ull;
}

hipError_t launch_count_dm(const PassGeom& g, int64_t* dm, hipStream_t s) {
  unsigned long long* d = (unsigned long long*)dm;
  hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)((g.I + 255) / 256)), dim3(256), 0, s, d, g.I);
  hipLaunchKernelGGL(k_count_dm, dim3((unsigned)((g.I + 255) / 256), (unsigned)((g.J8 + DM_CW - 1) / DM_CW)), dim3(256), 0,
                     s, g.gq, g.Ip, g.I, g.J, g.J8, d);
  return hipGetLastError();
}

hipError_t launch_count_dm_raw(const uint8_t* raw, int64_t bps, int64_t I, int64_t nb, int64_t* dm, hipStream_t s) {
  if (nb <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_count_dm_raw, dim3((unsigned)((I + 255) / 256), (unsigned)((nb + DM_CW - 1) / DM_CW)), dim3(256),
                     0, s, raw, bps, I, nb, (unsigned long long*)dm);
  return hipGetLastError();
}

hipError_t launch_decode(const PassGeom& g, uint8_t* codes, hipStream_t s) {
  int64_t n = g.I * g.J;
  hipLaunchKernelGGL(k_decode, dim3(grid_1d(n)), dim3(256), 0, s, g.gq, g.Ip, g.I, g.J, codes);
  return hipGetLastError();
}

hipError_t launch_reduce_q(const double* part, int nchunk, int NT, const PassGeom& g, double* res, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_q, dim3((unsigned)((g.I + 255) / 256), (unsigned)NT), dim3(256), 0, s, part, nchunk, NT, 0,
                     g.Ip, g.I, res, g.acc32, g.gate);
  return hipGetLastError();
}

hipError_t launch_reduce_q_curv(const double* part, int nchunk, int K, const PassGeom& g, const double* q, double* res,
                               hipStream_t s) {
  const int NT = K + K * (K + 1) / 2;
  const unsigned nb = (unsigned)((g.I + 255) / 256);
  hipLaunchKernelGGL(k_reduce_q, dim3(nb, (unsigned)(NT - K)), dim3(256), 0, s, part, nchunk, NT, K, g.Ip, g.I, res,
                     g.acc32, g.gate);
  hipLaunchKernelGGL(k_xtz_q, dim3(nb), dim3(256), 0, s, K, g.I, q, res, g.gate);
  return hipGetLastError();
}

hipError_t launch_finish_p(int K, const PassGeom& g, const double* p, double* res, unsigned char* flags, int* dense,
                           hipStream_t s) {
  if (flags) {
    const hipError_t e = launch_set_int(dense, 0, s, g.gate);
    if (e != hipSuccess) return e;
  }
  const dim3 grid((unsigned)((4 * g.J + 255) / 256));
  switch (K) {
    case 1: hipLaunchKernelGGL(k_finish_p<1>, grid, dim3(256), 0, s, g.J, p, res, flags, dense, g.gate); break;
    case 2: hipLaunchKernelGGL(k_finish_p<2>, grid, dim3(256), 0, s, g.J, p, res, flags, dense, g.gate); break;
    case 3: hipLaunchKernelGGL(k_finish_p<3>, grid, dim3(256), 0, s, g.J, p, res, flags, dense, g.gate); break;
    case 4: hipLaunchKernelGGL(k_finish_p<4>, grid, dim3(256), 0, s, g.J, p, res, flags, dense, g.gate); break;
    case 5: hipLaunchKernelGGL(k_finish_p<5>, grid, dim3(256), 0, s, g.J, p, res, flags, dense, g.gate); break;
    case 6: hipLaunchKernelGGL(k_finish_p<6>, grid, dim3(256), 0, s, g.J, p, res, flags, dense, g.gate); break;
    case 7: hipLaunchKernelGGL(k_finish_p<7>, grid, dim3(256), 0, s, g.J, p, res, flags, dense, g.gate); break;
    case 8: hipLaunchKernelGGL(k_finish_p<8>, grid, dim3(256), 0, s, g.J, p, res, flags, dense, g.gate); break;
    case 9: hipLaunchKernelGGL(k_finish_p<9>, grid, dim3(256), 0, s, g.J, p, res, flags, dense, g.gate); break;
    case 10: hipLaunchKernelGGL(k_finish_p<10>, grid, dim3(256), 0, s, g.J, p, res, flags, dense, g.gate); break;
    case 11: hipLaunchKernelGGL(k_finish_p<11>, grid, dim3(256), 0, s, g.J, p, res, flags, dense, g.gate); break;
    case 12: hipLaunchKernelGGL(k_finish_p<12>, grid, dim3(256), 0, s, g.J, p, res, flags, dense, g.gate); break;
    case 13: hipLaunchKernelGGL(k_finish_p<13>, grid, dim3(256), 0, s, g.J, p, res, flags, dense, g.gate); break;
    case 14: hipLaunchKernelGGL(k_finish_p<14>, grid, dim3(256), 0, s, g.J, p, res, flags, dense, g.gate); break;
    case 15: hipLaunchKernelGGL(k_finish_p<15>, grid, dim3(256), 0, s, g.J, p, res, flags, dense, g.gate); break;
    case 16: hipLaunchKernelGGL(k_finish_p<16>, grid, dim3(256), 0, s, g.J, p, res, flags, dense, g.gate); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_reduce_p(const double* part, int nchunk, int NT, const PassGeom& g, double* res, hipStream_t s) {
  int64_t n = 4 * g.J * NT;
  hipLaunchKernelGGL(k_reduce_p, dim3(grid_1d(n)), dim3(256), 0, s, part, nchunk, NT, g.Jp, g.J, res,
                     g.acc32, g.gate);
  return hipGetLastError();
}

// The workgroup sums go to part + nchunk * Ip: the log-likelihood pass fills only the first nchunk * Ip
// entries of the partial buffer, which holds nchunk * NT * Ip (NT >= 2).
hipError_t launch_reduce_ll(const double* part, int nchunk, const PassGeom& g, double* rows, double* total,
                            hipStream_t s) {
  const int nb = (int)((g.I + 255) / 256);
  double* bpart = const_cast<double*>(part) + (int64_t)nchunk * g.Ip;
  hipLaunchKernelGGL(k_reduce_ll_rows, dim3((unsigned)nb), dim3(256), 0, s, part, nchunk, g.Ip, g.I, rows, bpart, g.gate);
  hipLaunchKernelGGL(k_reduce_ll_total, dim3(1), dim3(256), 0, s, bpart, nb, total, g.gate);
  return hipGetLastError();
}

hipError_t launch_expand(const double* res, int K, int NT, int64_t n, double* XtX, double* Xtz, hipStream_t s) {
  hipLaunchKernelGGL(k_expand, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, res, K, NT, n, XtX, Xtz);
  return hipGetLastError();
}

hipError_t launch_em_expand(const double* res, int K, int NT, int64_t n, const double* par, double* out,
                            hipStream_t s) {
  int64_t m = (int64_t)K * n;
  hipLaunchKernelGGL(k_em_expand, dim3(grid_1d(m)), dim3(256), 0, s, res, K, NT, n, par, out);
  return hipGetLastError();
}

}  // namespace hk
