// hk_ll.hip — log-likelihood pass of the MIXED engine on the matrix cores (loglikelihood_full_loop2,
// src/loops.jl:236-258).  Its own translation unit: it is built with the VGPR form of the matrix instructions
// (-mllvm -amdgpu-mfma-vgpr-form), so each word's 32 x 32 result lands in VGPRs where the per-pair arithmetic
// reads it, instead of accumulation registers that would need one move per value; the curvature passes keep
// their accumulators in AGPRs across many steps and stay in hk_pass.hip.
#include <algorithm>

#include "hk_internal.h"
#include "hk_pair.h"

namespace hk {

// -----------------------------------------------------------------------------------------------
// Log-likelihood on the matrix cores (MIXED engine), one or two parameter sets per genotype sweep
//
// Per genotype word (8 blocks) a wave of 32 individuals forms X = P^T Q, X_h = sum_k q_k p_kh, with
// v_mfma_f32_32x32x16_bf16: rows m = 4b + h (block b of the word, haplotype h), columns = the 32 individuals,
// and the reduction runs over (population k, BF16 term pair u).  Each FP32 operand enters as three BF16 terms
// (split3: x = x_h + x_m + x_l), and the six term pairs (p, q) = (h,h) (h,m) (m,h) (h,l) (m,m) (l,h) reproduce
// q_k p_kh to FP32 accuracy (the dropped pairs are below 2^-25 of it; every BF16 x BF16 product is exact in the
// FP32 accumulator).  Reduction slot v = 6k + u; instruction t, lane half hf, element e is slot v = 16t + 8hf + e,
// so NM = ceil(6K / 16) instructions per set (K = 6: 3, against ceil(K / 2) = 3 of the FP32 32 x 32 x 2
// instruction at twice its cycles; K = 8: 3 against 4).  Operand layout: lane l holds A[l % 32][slots of half
// l / 32] and B[slots of half l / 32][l % 32] (the same slots for both, whatever their order inside the
// instruction).  The A operands (p terms, per word) come pre-split from k_ll_split, NM 16-byte loads per lane and
// set (one contiguous kilobyte per wave and instruction), loaded one word ahead; the B operands (q terms) are
// split once per wave.
// Result register v of lane l holds D[8 (v / 4) + 4 (l / 32) + v % 4][l % 32], i.e. X_h (h = v % 4) of
// individual l % 32 at block 2 (v / 4) + l / 32 -- four pairs per lane, all four haplotypes of each in the
// lane.  The per-pair likelihood (lik_factored) then runs on the VALU; the FP64 running product is renormalised
// once per word.  NSET = 2 evaluates two parameter sets (the QN iteration's ll_basic and ll_qn) over one
// genotype load, nibble decode and class-table read per pair.
// part: set s, chunk c, individual i at part[s * pstride + c * Ip + i] (lanes l and l + 32 combined).
// -----------------------------------------------------------------------------------------------
__host__ __device__ constexpr int ll_nm(int K) { return (6 * K + 15) / 16; }
// BF16 term (0 = h, 1 = m, 2 = l) of p and of q in term pair u
__host__ __device__ constexpr int ll_term_p(int u) { return u == 2 || u == 4 ? 1 : (u == 5 ? 2 : 0); }
__host__ __device__ constexpr int ll_term_q(int u) { return u == 1 || u == 4 ? 1 : (u == 3 ? 2 : 0); }

size_t ll_stage_bytes(int K, int64_t J8) { return (size_t)J8 * 64 * (size_t)ll_nm(K) * 16; }

// A operands of k_ll_mm: for set blockIdx.y, stage[((jw NM + t) 64 + 32 hf + m) 8 + e] = BF16 term ll_term_p(u)
// of p[k + K (32 jw + m)] with v = 16 t + 8 hf + e, k = v / 6, u = v % 6 (0 for v >= 6K): the wave's 16-byte
// operands of instruction t are one contiguous kilobyte.  One thread per 8 terms.
template <int K>
__global__ void k_ll_split(const float* __restrict__ p1, const float* __restrict__ p2, int64_t J8,
                           uint16_t* __restrict__ stage, int64_t sstride, const int* __restrict__ gate) {
  if (gate_off(gate)) return;
  constexpr int NM = ll_nm(K);
  const float* p = blockIdx.y ? p2 : p1;
  uint16_t* out = stage + blockIdx.y * sstride;
  const int64_t n = J8 * 64 * NM;
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x) {
    const int ln = (int)(x & 63);  // lane 32 hf + m
    const int64_t jt = x >> 6;     // jw NM + t
    const int t = (int)(jt % NM), hf = ln >> 5;
    const float* pr = p + (int64_t)K * (32 * (jt / NM) + (ln & 31));
    uint32_t w[4];
#pragma unroll
    for (int e2 = 0; e2 < 4; ++e2) {
      uint16_t tv[2];
#pragma unroll
      for (int o = 0; o < 2; ++o) {
        const int v = 16 * t + 8 * hf + 2 * e2 + o;
        tv[o] = 0;
        if (v < 6 * K) {
          uint16_t h, m, l;
          split3(pr[v / 6], h, m, l);
          const int tp = ll_term_p(v % 6);
          tv[o] = tp == 0 ? h : (tp == 1 ? m : l);
        }
      }
      w[e2] = (uint32_t)tv[0] | ((uint32_t)tv[1] << 16);
    }
    reinterpret_cast<uint4*>(out)[x] = make_uint4(w[0], w[1], w[2], w[3]);
  }
}

template <int K, int NSET>
__global__ __launch_bounds__(256) void k_ll_mm(const uint32_t* __restrict__ gq, int64_t Ip, int64_t I, int J8, int wpc,
                                               const float* __restrict__ q1, const float* __restrict__ q2,
                                               const uint16_t* __restrict__ stage, int64_t sstride,
                                               double* __restrict__ part, int64_t pstride,
                                               const int* __restrict__ gate) {
  if (gate_off(gate)) return;
  typedef float v16 __attribute__((ext_vector_type(16)));
  typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
  typedef unsigned short u8v __attribute__((ext_vector_type(8)));
  constexpr int NM = ll_nm(K);
  __shared__ __attribute__((aligned(16))) float lut[160];
  fill_lik_lut<float>(lut);
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int col = lane & 31, half = lane >> 5;
  unsigned bx, by;
  xcd_grouped_block(bx, by);                               // a chunk's workgroups share one L2 for its p terms
  const int64_t i = (int64_t)bx * 128 + wv * 32 + col;   // < Ip: the grid covers Ip / 128 workgroups
  const int64_t ic = i < I ? i : I - 1;                  // padding individuals: double-missing nibbles
  const int chunk = by;
  const int jw0 = chunk * wpc, jw1 = min(J8, jw0 + wpc);
  // B operands: the BF16 terms of this individual's q for the lane's slots (both halves' candidates are
  // compile-time slots; the lane's half picks one)
  bf8 bq[NSET][NM];
#pragma unroll
  for (int st = 0; st < NSET; ++st) {
    const float* qi = (st ? q2 : q1) + (int64_t)K * ic;
    uint16_t tq[3][K];
#pragma unroll
    for (int k = 0; k < K; ++k) split3(qi[k], tq[0][k], tq[1][k], tq[2][k]);
#pragma unroll
    for (int t = 0; t < NM; ++t) {
      u8v b;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int v0 = 16 * t + e, v1 = v0 + 8;
        const unsigned short a0 = v0 < 6 * K ? tq[ll_term_q(v0 % 6)][v0 / 6] : (unsigned short)0;
        const unsigned short a1 = v1 < 6 * K ? tq[ll_term_q(v1 % 6)][v1 / 6] : (unsigned short)0;
        b[e] = half ? a1 : a0;
      }
      bq[st][t] = __builtin_bit_cast(bf8, b);
    }
  }
  double llm[NSET];
  int lle[NSET];
#pragma unroll
  for (int st = 0; st < NSET; ++st) {
    llm[st] = 1.0;
    lle[st] = 0;
  }
  // A operands (the p terms of the word's 32 rows) and the genotype word are loaded one word ahead into the other
  // of two register sets, so their latency is covered by the current word's arithmetic
  const uint4* sa[NSET];
#pragma unroll
  for (int st = 0; st < NSET; ++st)
    sa[st] = reinterpret_cast<const uint4*>(stage + st * sstride) + (int64_t)jw0 * NM * 64 + lane;
  const uint32_t* gw = gq + (int64_t)jw0 * Ip + i;
  auto fetch = [&](uint4 (&a)[NSET][NM], uint32_t& w) {
    w = *gw;
    gw += Ip;
#pragma unroll
    for (int st = 0; st < NSET; ++st) {
#pragma unroll
      for (int t = 0; t < NM; ++t) a[st][t] = sa[st][64 * t];
      sa[st] += 64 * NM;
    }
  };
  auto step = [&](const uint4 (&a)[NSET][NM], uint32_t word) {
    v16 X[NSET];
#pragma unroll
    for (int st = 0; st < NSET; ++st) {
      v16 acc = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int t = 0; t < NM; ++t)
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf8, a[st][t]), bq[st][t], acc, 0, 0, 0);
      X[st] = acc;
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const unsigned nib = nib_at(word, 4 * (2 * m + half));
#pragma unroll
      for (int st = 0; st < NSET; ++st) {
        const float Xh[4] = {X[st][4 * m], X[st][4 * m + 1], X[st][4 * m + 2], X[st][4 * m + 3]};
        llm[st] *= (double)lik_factored<float>(Xh, lut, nib);
      }
    }
#pragma unroll
    for (int st = 0; st < NSET; ++st) {
      int e;
      llm[st] = frexp(llm[st], &e);
      lle[st] += e;
    }
  };
  uint4 a0[NSET][NM], a1[NSET][NM];
  uint32_t w0 = 0, w1 = 0;
  const int n = jw1 - jw0;
  if (n > 0) fetch(a0, w0);
  int d = 0;
  for (; d + 1 < n; d += 2) {
    fetch(a1, w1);
    step(a0, w0);
    if (d + 2 < n) fetch(a0, w0);
    step(a1, w1);
  }
  if (d < n) step(a0, w0);
#pragma unroll
  for (int st = 0; st < NSET; ++st) {
    double lg = log(llm[st]) + (double)lle[st] * 0.69314718055994530942;
    lg += __shfl_xor(lg, 32);  // the other half of the lanes holds this individual's other four blocks
    if (half == 0) part[st * pstride + (int64_t)chunk * Ip + i] = lg;
  }
}

template <int K>
static hipError_t ll_mm_k(const PassGeom& g, const float* q1, const float* p1, const float* q2, const float* p2,
                          double* part, void* stage, hipStream_t s) {
  const int nset = q2 ? 2 : 1;
  const int64_t sstride = (int64_t)(ll_stage_bytes(K, g.J8) / sizeof(uint16_t));
  const int64_t nsplit = g.J8 * 64 * ll_nm(K);
  const unsigned sgrid = (unsigned)std::min<int64_t>((nsplit + 255) / 256, 4096);
  if ((uintptr_t)stage % 16) return hipErrorInvalidValue;  // the kernels' 16-byte term loads and stores
  hipLaunchKernelGGL((k_ll_split<K>), dim3(sgrid, nset), dim3(256), 0, s, p1, q2 ? p2 : p1, g.J8, (uint16_t*)stage,
                     sstride, g.gate);
  const dim3 grid((unsigned)(g.Ip / 128), (unsigned)g.qchunks), block(256);
  const int64_t pstride = (int64_t)g.qchunks * g.Ip;
  if (q2)
    hipLaunchKernelGGL((k_ll_mm<K, 2>), grid, block, 0, s, g.gq, g.Ip, g.I, (int)g.J8, g.qwpc, q1, q2,
                       (const uint16_t*)stage, sstride, part, pstride, g.gate);
  else
    hipLaunchKernelGGL((k_ll_mm<K, 1>), grid, block, 0, s, g.gq, g.Ip, g.I, (int)g.J8, g.qwpc, q1, q1,
                       (const uint16_t*)stage, sstride, part, pstride, g.gate);
  return hipGetLastError();
}

hipError_t launch_ll_mm(int K, const PassGeom& g, const float* q1, const float* p1, const float* q2, const float* p2,
                        double* part, void* stage, hipStream_t s) {
  switch (K) {
    case 1: return ll_mm_k<1>(g, q1, p1, q2, p2, part, stage, s);
    case 2: return ll_mm_k<2>(g, q1, p1, q2, p2, part, stage, s);
    case 3: return ll_mm_k<3>(g, q1, p1, q2, p2, part, stage, s);
    case 4: return ll_mm_k<4>(g, q1, p1, q2, p2, part, stage, s);
    case 5: return ll_mm_k<5>(g, q1, p1, q2, p2, part, stage, s);
    case 6: return ll_mm_k<6>(g, q1, p1, q2, p2, part, stage, s);
    case 7: return ll_mm_k<7>(g, q1, p1, q2, p2, part, stage, s);
    case 8: return ll_mm_k<8>(g, q1, p1, q2, p2, part, stage, s);
    case 9: return ll_mm_k<9>(g, q1, p1, q2, p2, part, stage, s);
    case 10: return ll_mm_k<10>(g, q1, p1, q2, p2, part, stage, s);
    case 11: return ll_mm_k<11>(g, q1, p1, q2, p2, part, stage, s);
    case 12: return ll_mm_k<12>(g, q1, p1, q2, p2, part, stage, s);
    case 13: return ll_mm_k<13>(g, q1, p1, q2, p2, part, stage, s);
    case 14: return ll_mm_k<14>(g, q1, p1, q2, p2, part, stage, s);
    case 15: return ll_mm_k<15>(g, q1, p1, q2, p2, part, stage, s);
    case 16: return ll_mm_k<16>(g, q1, p1, q2, p2, part, stage, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace hk
