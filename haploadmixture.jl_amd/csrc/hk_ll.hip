// hk_ll.hip — log-likelihood pass of the MIXED engine on the matrix cores (loglikelihood_full_loop2,
// src/loops.jl:236-258).  Its own translation unit: it is built with the VGPR form of the matrix instructions
// (-mllvm -amdgpu-mfma-vgpr-form), so each word's 32 x 32 result lands in VGPRs where the per-pair arithmetic
// reads it, instead of accumulation registers that would need one move per value; the curvature passes keep
// their accumulators in AGPRs across many steps and stay in hk_pass.hip.
#include "hk_internal.h"
#include "hk_pair.h"

namespace hk {

// -----------------------------------------------------------------------------------------------
// Log-likelihood on the matrix cores (MIXED engine), one or two parameter sets per genotype sweep
//
// Per genotype word (8 blocks) a wave of 32 individuals forms X = P^T Q with v_mfma_f32_32x32x2f32:
// A = p rows (32 = 8 blocks x 4 haplotypes, row 4b + h) x 2 populations, B = q (2 populations x 32
// individuals), ceil(K/2) = NK2 instructions (K = 6: 3, no padded half).  Operands: lane l holds A[l % 32][l / 32]
// and B[l / 32][l % 32]; instruction t contracts populations t (lanes 0-31) and NK2 + t (lanes 32-63, zero past
// K), so a lane's A operands over t are NK2 consecutive floats of one p row: one 16-byte load per set at K = 8
// and an 8-byte one at K = 4 (populations 2t and 2t + 1 per instruction put 32-byte strides between the lanes'
// single-float loads at K = 8, which held the texture addresser busy 2.6 times as long as at K = 7); K = 12: 8-byte
// loads, K = 16: two 16-byte ones.
// Result register v of lane l holds D[8 (v / 4) + 4 (l / 32) + v % 4][l % 32], i.e.
// X_h (h = v % 4) of individual l % 32 at block 2 (v / 4) + l / 32 -- four pairs per lane, all four
// haplotypes of each in the lane.  The per-pair likelihood (pair_lik) then runs on the VALU as in k_qside;
// the FP64 running product is renormalised once per word.  NSET = 2 evaluates two parameter sets (the
// QN iteration's ll_basic and ll_qn) over one genotype load, nibble decode and class-table read per pair.
// part: set s, chunk c, individual i at part[s * sstride + c * Ip + i] (lanes l and l + 32 combined).
// -----------------------------------------------------------------------------------------------
template <int K, int NSET>
__global__ __launch_bounds__(256) void k_ll_mm(const uint32_t* __restrict__ gq, int64_t Ip, int64_t I, int J8, int wpc,
                                               const float* __restrict__ q1, const float* __restrict__ p1,
                                               const float* __restrict__ q2, const float* __restrict__ p2,
                                               double* __restrict__ part, int64_t sstride,
                                               const int* __restrict__ gate) {
  if (gate_off(gate)) return;
  typedef float v16 __attribute__((ext_vector_type(16)));
  constexpr int NK2 = (K + 1) / 2;
  __shared__ __attribute__((aligned(16))) float lut[128];
  fill_pair_lut<float>(lut);
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int col = lane & 31, half = lane >> 5;
  unsigned bx, by;
  xcd_grouped_block(bx, by);                               // a chunk's workgroups share one L2 for its p rows
  const int64_t i = (int64_t)bx * 128 + wv * 32 + col;   // < Ip: the grid covers Ip / 128 workgroups
  const int64_t ic = i < I ? i : I - 1;                  // padding individuals: double-missing nibbles
  const int chunk = by;
  const int jw0 = chunk * wpc, jw1 = min(J8, jw0 + wpc);
  const float* qs[2] = {q1, q2};
  const float* ps[2] = {p1, p2};
  float bq[NSET][NK2];
#pragma unroll
  for (int st = 0; st < NSET; ++st)
#pragma unroll
    for (int t = 0; t < NK2; ++t) {
      const int k = half * NK2 + t;
      bq[st][t] = k < K ? qs[st][k + (int64_t)K * ic] : 0.f;
    }
  double llm[NSET];
  int lle[NSET];
#pragma unroll
  for (int st = 0; st < NSET; ++st) {
    llm[st] = 1.0;
    lle[st] = 0;
  }
  // A operands (p of the word's 32 rows) and the genotype word are loaded one word ahead, so their latency is
  // covered by the current word's arithmetic
  float an[NSET][NK2];
  uint32_t wn = 0;
  auto fetch = [&](int jw) {
    wn = gq[(int64_t)jw * Ip + i];
#pragma unroll
    for (int st = 0; st < NSET; ++st) {
      // p[k + K (4 (8 jw + b) + h)], col = 4b + h; this lane's populations half * NK2 + t
      const float* pb = ps[st] + (int64_t)K * (32 * (int64_t)jw + col) + half * NK2;
      if constexpr (K % 4 == 0 && NK2 % 4 == 0) {  // 16-byte aligned: checked by the launcher
#pragma unroll
        for (int u = 0; u < NK2 / 4; ++u) {
          const float4 v = reinterpret_cast<const float4*>(pb)[u];
          an[st][4 * u] = v.x, an[st][4 * u + 1] = v.y, an[st][4 * u + 2] = v.z, an[st][4 * u + 3] = v.w;
        }
      } else if constexpr (K % 2 == 0 && NK2 % 2 == 0) {  // 8-byte aligned
#pragma unroll
        for (int u = 0; u < NK2 / 2; ++u) {
          const float2 v = reinterpret_cast<const float2*>(pb)[u];
          an[st][2 * u] = v.x, an[st][2 * u + 1] = v.y;
        }
      } else {
#pragma unroll
        for (int t = 0; t < NK2; ++t) an[st][t] = half * NK2 + t < K ? pb[t] : 0.f;
      }
    }
  };
  if (jw0 < jw1) fetch(jw0);
  for (int jw = jw0; jw < jw1; ++jw) {
    const uint32_t word = wn;
    float ac[NSET][NK2];
#pragma unroll
    for (int st = 0; st < NSET; ++st)
#pragma unroll
      for (int t = 0; t < NK2; ++t) ac[st][t] = an[st][t];
    if (jw + 1 < jw1) fetch(jw + 1);
    v16 X[NSET];
#pragma unroll
    for (int st = 0; st < NSET; ++st) {
      v16 acc = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int t = 0; t < NK2; ++t) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(ac[st][t], bq[st][t], acc, 0, 0, 0);
      X[st] = acc;
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const unsigned nib = nib_at(word, 4 * (2 * m + half));
#pragma unroll
      for (int st = 0; st < NSET; ++st) {
        const float Xh[4] = {X[st][4 * m], X[st][4 * m + 1], X[st][4 * m + 2], X[st][4 * m + 3]};
        llm[st] *= (double)pair_lik<float>(Xh, lut, nib);
      }
    }
#pragma unroll
    for (int st = 0; st < NSET; ++st) {
      int e;
      llm[st] = frexp(llm[st], &e);
      lle[st] += e;
    }
  }
#pragma unroll
  for (int st = 0; st < NSET; ++st) {
    double lg = log(llm[st]) + (double)lle[st] * 0.69314718055994530942;
    lg += __shfl_xor(lg, 32);  // the other half of the lanes holds this individual's other four blocks
    if (half == 0) part[st * sstride + (int64_t)chunk * Ip + i] = lg;
  }
}

template <int K>
static hipError_t ll_mm_k(const PassGeom& g, const float* q1, const float* p1, const float* q2, const float* p2,
                          double* part, hipStream_t s) {
  const dim3 grid((unsigned)(g.Ip / 128), (unsigned)g.qchunks), block(256);
  const int64_t sstride = (int64_t)g.qchunks * g.Ip;
  constexpr int NK2 = (K + 1) / 2;
  constexpr uintptr_t al = (K % 4 == 0 && NK2 % 4 == 0) ? 16 : ((K % 2 == 0 && NK2 % 2 == 0) ? 8 : 4);
  if ((uintptr_t)p1 % al || (q2 && (uintptr_t)p2 % al)) return hipErrorInvalidValue;  // the kernel's p-row loads
  if (q2)
    hipLaunchKernelGGL((k_ll_mm<K, 2>), grid, block, 0, s, g.gq, g.Ip, g.I, (int)g.J8, g.qwpc, q1, p1, q2, p2, part,
                       sstride, g.gate);
  else
    hipLaunchKernelGGL((k_ll_mm<K, 1>), grid, block, 0, s, g.gq, g.Ip, g.I, (int)g.J8, g.qwpc, q1, p1, q1, p1, part,
                       sstride, g.gate);
  return hipGetLastError();
}

hipError_t launch_ll_mm(int K, const PassGeom& g, const float* q1, const float* p1, const float* q2, const float* p2,
                        double* part, hipStream_t s) {
  switch (K) {
    case 1: return ll_mm_k<1>(g, q1, p1, q2, p2, part, s);
    case 2: return ll_mm_k<2>(g, q1, p1, q2, p2, part, s);
    case 3: return ll_mm_k<3>(g, q1, p1, q2, p2, part, s);
    case 4: return ll_mm_k<4>(g, q1, p1, q2, p2, part, s);
    case 5: return ll_mm_k<5>(g, q1, p1, q2, p2, part, s);
    case 6: return ll_mm_k<6>(g, q1, p1, q2, p2, part, s);
    case 7: return ll_mm_k<7>(g, q1, p1, q2, p2, part, s);
    case 8: return ll_mm_k<8>(g, q1, p1, q2, p2, part, s);
    case 9: return ll_mm_k<9>(g, q1, p1, q2, p2, part, s);
    case 10: return ll_mm_k<10>(g, q1, p1, q2, p2, part, s);
    case 11: return ll_mm_k<11>(g, q1, p1, q2, p2, part, s);
    case 12: return ll_mm_k<12>(g, q1, p1, q2, p2, part, s);
    case 13: return ll_mm_k<13>(g, q1, p1, q2, p2, part, s);
    case 14: return ll_mm_k<14>(g, q1, p1, q2, p2, part, s);
    case 15: return ll_mm_k<15>(g, q1, p1, q2, p2, part, s);
    case 16: return ll_mm_k<16>(g, q1, p1, q2, p2, part, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace hk
