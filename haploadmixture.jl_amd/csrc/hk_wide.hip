// hk_wide.hip — curvature and P-side EM passes for 9 <= K <= 16 (update_q_loop!, loops.jl:328-365;
// update_p_loop!, loops.jl:367-403; em_p_loop!, loops.jl:297-326).
//
// The passes for K <= 8 (hk_pass.hip) keep every packed curvature column of a pair's K x K product in the
// accumulators of one workgroup.  At K = 16 that is 136 columns, more than the registers hold, so here the
// packed columns are split into windows of NTW tiles of 16 (grid.z = window; the tiles NTW does not divide go
// to one more launch with a smaller window): each workgroup recomputes the pair weights of its pairs and
// accumulates only its window's columns, reading the window's slice of the staging table.  The arithmetic per pair is the one of the K <= 8 kernels (pair_v / PairForm, MFMA with the
// weights as A and the table as B, MIXED: FP32 partials flushed into FP64); there is no VALU column tail, the
// last tile is zero-padded.  The P-side EM sums use the same kernel with the q values of each individual as
// the single B tile (Xtz_p[k, jh] = sum_i w_h q_ki).
#include "hk_internal.h"
#include "hk_pair.h"

namespace hk {

// Staging-table shape for wide K: NTILE tiles of 16 packed columns (the last one zero-padded), rows of
// [2 x_k x_k' (QS) | x_k (K)] per individual (P side) and [2 p_kh p_k'h (QS), h = 0..3 | p_kh (4K)] per
// block (Q side), as k_stage_rows writes them.
template <int K>
struct WideShape {
  static constexpr int NKK = K * (K + 1) / 2;
  static constexpr int NTILE = (NKK + 15) / 16;
  static constexpr int QS = 16 * NTILE;
  static constexpr int RGP = (QS + K + 3) / 4 * 4;
  static constexpr int RGQ = 4 * QS + 4 * K;
  static constexpr int PAD_ROWS = 32;  // the last staging step of a chunk reads up to 32 rows past its end
};

// LDS row pitch >= n words with pitch = 16 (mod 64): the four lane groups of a wave read 16 consecutive words
// of four different rows, which then fall on distinct banks.
__host__ __device__ constexpr int wide_pitch(int n) { return n <= 16 ? 16 : ((n - 16 + 63) / 64) * 64 + 16; }

// Packed-column tiles per window.  Fewer windows mean less repeated per-pair algebra (X and the pair weights are
// recomputed per window), more tiles per window more accumulators.  FP32 Q side: up to five tiles in one window
// (K <= 12), windows of four for seven or eight tiles (K = 14, 15) and of three otherwise (K = 13, 16), all at two
// waves per SIMD (six tiles, or five beside four, would drop to one).  FP64 Q side: three.  P side (whose lanes
// also hold a block's 4K p values): three for FP32 up to K = 12, two otherwise, and two for four tiles in all
// (K = 10), where two windows of two beat two of three, one of them half empty.
template <int K, typename TP, int SIDE>
constexpr int wide_tiles() {
  if constexpr (SIDE == 0 && sizeof(TP) == 4) {
    constexpr int NT = WideShape<K>::NTILE;
    return NT <= 5 ? NT : (NT == 7 || NT == 8 ? 4 : 3);
  }
  return (SIDE == 1 && (sizeof(TP) == 8 || K > 12)) || WideShape<K>::NTILE == 4 ? 2 : 3;
}

// Unit index of a staging copy, made opaque to the optimiser: its address arithmetic is then redone at every
// stage instead of being hoisted out of the genotype loop into registers held across it (which cost occupancy).
__device__ __forceinline__ unsigned stage_unit(unsigned e) {
  __asm__ volatile("" : "+v"(e));
  return e;
}

// ---------------------------------------------------------------------------------------------------------
// Q side: a wave owns 2 x 16 individuals; each MFMA step takes 4 blocks.  A stage holds 16 blocks (2 words):
// per block the window's 4 x 16 NTW table columns and the 4K p values.
// part[chunk][NT][Ip] receives the window's packed columns K + n.
// ---------------------------------------------------------------------------------------------------------
template <int K, typename TP, int NTW>
__global__ __launch_bounds__(256) void k_qside_mmw(const uint32_t* __restrict__ gq, int64_t Ip, int64_t I, int J8,
                                                   int wpc, const TP* __restrict__ q, const TP* __restrict__ pp,
                                                   double* __restrict__ part, int flush, int cbase,
                                                   const int* __restrict__ gate) {
  if (gate_off(gate)) return;
  using S = WideShape<K>;
  constexpr int NKK = S::NKK, NT = K + NKK, QS = S::QS, RG = S::RGQ;
  constexpr int WPS = 2, R = 2;
  constexpr int CW = 16 * NTW, RW = 4 * CW + 4 * K, RL = wide_pitch(RW);
  constexpr int U = 16 / (int)sizeof(TP), UR = RW / U, NU = 8 * WPS * UR, UPT = (NU + 255) / 256;
  static_assert(RW % U == 0 && CW % U == 0 && RG % U == 0 && QS % U == 0, "staged rows are whole 16-byte units");
  constexpr bool F32 = sizeof(TP) == 4;
  using M = Mfma<TP>;
  using V4 = typename M::v4;
  __shared__ __attribute__((aligned(16))) TP prow[8 * WPS * RL];
  __shared__ __attribute__((aligned(16))) TP lut[128];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int m = lane & 15, g = lane >> 4;
  unsigned bx, by;
  xcd_grouped_block(bx, by);
  const int64_t ibase = ((int64_t)bx * 4 + wv) * 16 * R;
  const int chunk = by;
  const int c0 = cbase + 16 * NTW * (int)blockIdx.z;  // first packed column of this window
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
  V4 C[R][NTW];
  double D[R][NTW][4];
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int t = 0; t < NTW; ++t) {
      C[r][t] = V4{0, 0, 0, 0};
#pragma unroll
      for (int v = 0; v < 4; ++v) D[r][t][v] = 0.0;
    }
  int nword = 0;
  for (int jws = jw0; jws < jw1; jws += WPS) {
    const int nw = min(WPS, jw1 - jws);
    // stage in 16-byte units (each thread's loads unrolled, so they are in flight together)
    const TP* base = pp + (int64_t)jws * 8 * RG;  // rows of blocks 8 jws .. 8 jws + 15 (the table is padded)
    __syncthreads();
#pragma unroll
    for (int u = 0; u < UPT; ++u) {
      const unsigned e = stage_unit(threadIdx.x + 256u * u);
      if (e < (unsigned)NU) {
        const unsigned b = e / UR, f = (e - b * UR) * U;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (f < 4u * CW) {
          const unsigned h = f / CW, col = c0 + (f - h * CW);
          if (col < (unsigned)QS) v = *reinterpret_cast<const uint4*>(base + (int64_t)b * RG + h * QS + col);
        } else {
          v = *reinterpret_cast<const uint4*>(base + (int64_t)b * RG + 4 * QS + (f - 4 * CW));
        }
        *reinterpret_cast<uint4*>(prow + b * RL + f) = v;  // columns past QS: zero padding
      }
    }
    __syncthreads();
    for (int w = 0; w < nw; ++w) {
      uint32_t word[R];
#pragma unroll
      for (int r = 0; r < R; ++r) word[r] = gq[(int64_t)(jws + w) * Ip + ibase + 16 * r + m];
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const TP* row = prow + (8 * w + 4 * s + g) * RL;  // staged block of this lane's A column
        TP b[4][NTW];
#pragma unroll
        for (int h = 0; h < 4; ++h)
#pragma unroll
          for (int t = 0; t < NTW; ++t) b[h][t] = row[h * CW + 16 * t + m];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          TP X[4];
#pragma unroll
          for (int h = 0; h < 4; ++h) {
            TP x = (TP)0;
#pragma unroll
            for (int k = 0; k < K; ++k) x = hk_fma(qi[r][k], row[4 * CW + h * K + k], x);
            X[h] = x;
          }
          TP v[4];
          pair_v<TP>(X, lut, nib_at(word[r], 4 * (4 * s + g)), v);
#pragma unroll
          for (int h = 0; h < 4; ++h)
#pragma unroll
            for (int t = 0; t < NTW; ++t) C[r][t] = M::mma(v[h], b[h][t], C[r][t]);
        }
      }
      if constexpr (F32) {
        if (++nword == flush) {
          nword = 0;
#pragma unroll
          for (int r = 0; r < R; ++r)
#pragma unroll
            for (int t = 0; t < NTW; ++t) {
#pragma unroll
              for (int v = 0; v < 4; ++v) D[r][t][v] += (double)C[r][t][v];
              C[r][t] = V4{0, 0, 0, 0};
            }
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int t = 0; t < NTW; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const double val = F32 ? D[r][t][v] + (double)C[r][t][v] : (double)C[r][t][v];
        const int n = c0 + 16 * t + m;
        const int64_t row = ibase + 16 * r + M::row(g, v);  // individual of this result element
        if (n < NKK) part[((int64_t)chunk * NT + K + n) * Ip + row] = val;
      }
}

// ---------------------------------------------------------------------------------------------------------
// P side: a wave owns 16 blocks (A rows, their 4K p values in registers); each MFMA step takes 4
// individuals.  A stage holds 32 individuals (4 words): MODE_CURV per individual the window's 16 NTW table
// columns and its K q values; MODE_GRAD its K q values, zero-padded to 16 (the single B tile, k = column).
// MODE_CURV: out[chunk][4Jp][NT] slots K + n (window 0 also clears the gradient slots 0..K-1, which the
// finishing kernel forms); MODE_GRAD: out[chunk][4Jp][K] = 2 sum_i (Y_h / L') q_ki.
// ---------------------------------------------------------------------------------------------------------
template <int K, typename TP, int NTW, int MODE>
__global__ __launch_bounds__(256) void k_pside_mmw(const uint32_t* __restrict__ gp, int64_t Jp, int64_t I, int I8,
                                                   int wpc, const TP* __restrict__ q, const TP* __restrict__ p,
                                                   const TP* __restrict__ qq, double* __restrict__ out, int flush,
                                                   int cbase, const int* __restrict__ gate) {
  if (gate_off(gate)) return;
  using S = WideShape<K>;
  constexpr bool CURV = MODE == MODE_CURV;
  constexpr int NKK = S::NKK, NT = CURV ? K + NKK : K, QS = S::QS, RG = S::RGP;
  constexpr int WPS = 4;
  constexpr int U = 16 / (int)sizeof(TP), KU = (K + U - 1) / U * U;  // q values staged in whole 16-byte units
  constexpr int CW = CURV ? 16 * NTW : 0, RW = CURV ? CW + KU : 16, RL = wide_pitch(RW), NST = 8 * WPS * RW;
  constexpr int UR = RW / U, NU = 8 * WPS * UR, UPT = (NU + 255) / 256;
  static_assert(!CURV || (RW % U == 0 && RG % U == 0 && QS % U == 0 && QS + KU <= RG), "whole 16-byte units");
  constexpr bool F32 = sizeof(TP) == 4;
  using M = Mfma<TP>;
  using V4 = typename M::v4;
  __shared__ __attribute__((aligned(16))) TP qrow[8 * WPS * RL];
  __shared__ __attribute__((aligned(16))) TP lut[128];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int m = lane & 15, g = lane >> 4;
  const int64_t jbase = ((int64_t)blockIdx.x * 4 + wv) * 16;
  const int64_t j = jbase + m;
  const int chunk = blockIdx.y;
  const int c0 = cbase + 16 * NTW * (int)blockIdx.z;
  const int iw0 = chunk * wpc, iw1 = min(I8, iw0 + wpc);
  fill_pair_lut<TP>(lut);
  TP P[4][K];
#pragma unroll
  for (int h = 0; h < 4; ++h)
#pragma unroll
    for (int k = 0; k < K; ++k) P[h][k] = p[k + (int64_t)K * (4 * j + h)];
  V4 C[4][NTW];
  double D[4][NTW][4];
#pragma unroll
  for (int h = 0; h < 4; ++h)
#pragma unroll
    for (int t = 0; t < NTW; ++t) {
      C[h][t] = V4{0, 0, 0, 0};
#pragma unroll
      for (int v = 0; v < 4; ++v) D[h][t][v] = 0.0;
    }
  int nword = 0;
  for (int iws = iw0; iws < iw1; iws += WPS) {
    const int nw = min(WPS, iw1 - iws);
    if constexpr (CURV) {
      // 16-byte units of the window's columns and the individual's q values (the table row is padded past
      // them; each thread's loads unrolled, so they are in flight together)
      __syncthreads();
#pragma unroll
      for (int u = 0; u < UPT; ++u) {
        const unsigned e = stage_unit(threadIdx.x + 256u * u);
        if (e < (unsigned)NU) {
          const unsigned b = e / UR, f = (e - b * UR) * U;
          const TP* src = qq + ((int64_t)iws * 8 + b) * RG;  // < Ip + 32 rows: the q table is padded
          uint4 v = make_uint4(0u, 0u, 0u, 0u);
          if (f < (unsigned)CW) {
            if (c0 + f < (unsigned)QS) v = *reinterpret_cast<const uint4*>(src + c0 + f);  // else zero
          } else {
            v = *reinterpret_cast<const uint4*>(src + QS + (f - CW));
          }
          *reinterpret_cast<uint4*>(qrow + b * RL + f) = v;
        }
      }
    } else {
      __syncthreads();
      for (int e = threadIdx.x; e < NST; e += 256) {
        const int b = e / RW, f = e - b * RW;
        const int64_t ind = (int64_t)iws * 8 + b;
        const int64_t ic = ind < I ? ind : I - 1;
        qrow[b * RL + f] = f < K ? q[f + (int64_t)K * ic] : (TP)0;
      }
    }
    __syncthreads();
    for (int w = 0; w < nw; ++w) {
      const uint32_t word = gp[(int64_t)(iws + w) * Jp + j];
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const TP* row = qrow + (8 * w + 4 * s + g) * RL;  // staged individual of this lane's A column
        TP X[4];
#pragma unroll
        for (int h = 0; h < 4; ++h) {
          TP x = (TP)0;
#pragma unroll
          for (int k = 0; k < K; ++k) x = hk_fma(row[CW + k], P[h][k], x);
          X[h] = x;
        }
        const unsigned nib = nib_at(word, 4 * (4 * s + g));
        TP a[4];
        if constexpr (CURV) {
          pair_v<TP>(X, lut, nib, a);
        } else {
          const PairForm<TP> pf(X, lut, nib);
          const TP rl = hk_rcp(pf.L);
#pragma unroll
          for (int h = 0; h < 4; ++h) a[h] = pf.Y[h] * rl;
        }
        TP b[NTW];
#pragma unroll
        for (int t = 0; t < NTW; ++t) b[t] = row[16 * t + m];
#pragma unroll
        for (int h = 0; h < 4; ++h)
#pragma unroll
          for (int t = 0; t < NTW; ++t) C[h][t] = M::mma(a[h], b[t], C[h][t]);
      }
      if constexpr (F32) {
        if (++nword == flush) {
          nword = 0;
#pragma unroll
          for (int h = 0; h < 4; ++h)
#pragma unroll
            for (int t = 0; t < NTW; ++t) {
#pragma unroll
              for (int v = 0; v < 4; ++v) D[h][t][v] += (double)C[h][t][v];
              C[h][t] = V4{0, 0, 0, 0};
            }
        }
      }
    }
  }
  if (CURV && c0 == 0) {  // gradient slots of (block j, haplotype g), formed by the finishing kernel
    double* o = out + ((int64_t)chunk * 4 * Jp + 4 * j + g) * NT;
#pragma unroll
    for (int k = 0; k < K; ++k) o[k] = 0.0;
  }
#pragma unroll
  for (int h = 0; h < 4; ++h)
#pragma unroll
    for (int t = 0; t < NTW; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const double val = F32 ? D[h][t][v] + (double)C[h][t][v] : (double)C[h][t][v];
        const int n = c0 + 16 * t + m;
        const int64_t row = jbase + M::row(g, v);  // block of this result element
        if constexpr (CURV) {
          if (n < NKK) out[((int64_t)chunk * 4 * Jp + 4 * row + h) * NT + K + n] = val;
        } else {
          if (n < K) out[((int64_t)chunk * 4 * Jp + 4 * row + h) * NT + n] = 2.0 * val;  // w_h = 2 Y_h / L'
        }
      }
}

// ---------------------------------------------------------------------------------------------------------
// Launchers (9 <= K <= 16)
// ---------------------------------------------------------------------------------------------------------

template <int K, typename TP>
static hipError_t wide_tables_k(int side, const PassGeom& g, const void* x, hipStream_t s) {
  using S = WideShape<K>;
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

// The packed columns' NTILE tiles go to NTILE / NTW windows of NTW tiles (one launch, grid.z = window) and, when
// NTW does not divide NTILE, one more launch for a window of the remaining tiles, so that no window carries
// all-padding tiles (each tile costs four matrix instructions per pair step).
template <int K, typename TP>
static hipError_t wide_qcurv_k(const PassGeom& g, const void* q, double* part, hipStream_t s) {
  constexpr int NTW = wide_tiles<K, TP, 0>(), NF = WideShape<K>::NTILE / NTW, REM = WideShape<K>::NTILE % NTW;
  static_assert(NF >= 1, "at least one full window (NTILE >= 3 for K >= 9)");
  hipLaunchKernelGGL((k_qside_mmw<K, TP, NTW>), dim3((unsigned)(g.Ip / 128), (unsigned)g.qchunks, (unsigned)NF),
                     dim3(256), 0, s, g.gq, g.Ip, g.I, (int)g.J8, g.qwpc, (const TP*)q, (const TP*)g.pp, part,
                     g.flush_curv, 0, g.gate);
  if constexpr (REM > 0)
    hipLaunchKernelGGL((k_qside_mmw<K, TP, REM>), dim3((unsigned)(g.Ip / 128), (unsigned)g.qchunks, 1u), dim3(256), 0,
                       s, g.gq, g.Ip, g.I, (int)g.J8, g.qwpc, (const TP*)q, (const TP*)g.pp, part, g.flush_curv,
                       16 * NTW * NF, g.gate);
  return hipGetLastError();
}

template <int K, typename TP>
static hipError_t wide_pside_k(int mode, const PassGeom& g, const void* q, const void* p, double* out,
                               hipStream_t s) {
  if (mode == MODE_CURV) {
    constexpr int NTW = wide_tiles<K, TP, 1>(), NF = WideShape<K>::NTILE / NTW, REM = WideShape<K>::NTILE % NTW;
    static_assert(NF >= 1, "at least one full window (NTILE >= 3 for K >= 9)");
    hipLaunchKernelGGL((k_pside_mmw<K, TP, NTW, MODE_CURV>), dim3((unsigned)(g.Jp / 64), (unsigned)g.pchunks, (unsigned)NF),
                       dim3(256), 0, s, g.gp, g.Jp, g.I, (int)g.I8, g.pwpc, (const TP*)q, (const TP*)p,
                       (const TP*)g.qq, out, g.flush_curv, 0, g.gate);
    if constexpr (REM > 0)
      hipLaunchKernelGGL((k_pside_mmw<K, TP, REM, MODE_CURV>), dim3((unsigned)(g.Jp / 64), (unsigned)g.pchunks, 1u),
                         dim3(256), 0, s, g.gp, g.Jp, g.I, (int)g.I8, g.pwpc, (const TP*)q, (const TP*)p,
                         (const TP*)g.qq, out, g.flush_curv, 16 * NTW * NF, g.gate);
  } else if (mode == MODE_GRAD) {
    const dim3 grid((unsigned)(g.Jp / 64), (unsigned)g.pchunks, 1);
    hipLaunchKernelGGL((k_pside_mmw<K, TP, 1, MODE_GRAD>), grid, dim3(256), 0, s, g.gp, g.Jp, g.I, (int)g.I8,
                       g.pwpc, (const TP*)q, (const TP*)p, (const TP*)nullptr, out, g.flush_grad, 0, g.gate);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

#define HK_WIDE_SWITCH(FN, TP, ...)            \
  switch (K) {                                 \
    case 9: return FN<9, TP>(__VA_ARGS__);     \
    case 10: return FN<10, TP>(__VA_ARGS__);   \
    case 11: return FN<11, TP>(__VA_ARGS__);   \
    case 12: return FN<12, TP>(__VA_ARGS__);   \
    case 13: return FN<13, TP>(__VA_ARGS__);   \
    case 14: return FN<14, TP>(__VA_ARGS__);   \
    case 15: return FN<15, TP>(__VA_ARGS__);   \
    case 16: return FN<16, TP>(__VA_ARGS__);   \
    default: return hipErrorInvalidValue;      \
  }

hipError_t launch_curv_tables_wide(int K, int prec, int side, const PassGeom& g, const void* x, hipStream_t s) {
  if (prec == PREC_FP64) { HK_WIDE_SWITCH(wide_tables_k, double, side, g, x, s) }
  HK_WIDE_SWITCH(wide_tables_k, float, side, g, x, s)
}

hipError_t launch_qside_curv_wide(int K, int prec, const PassGeom& g, const void* q, double* part, hipStream_t s) {
  if (prec == PREC_FP64) { HK_WIDE_SWITCH(wide_qcurv_k, double, g, q, part, s) }
  HK_WIDE_SWITCH(wide_qcurv_k, float, g, q, part, s)
}

hipError_t launch_pside_wide(int K, int prec, int mode, const PassGeom& g, const void* q, const void* p,
                             double* out, hipStream_t s) {
  if (prec == PREC_FP64) { HK_WIDE_SWITCH(wide_pside_k, double, mode, g, q, p, out, s) }
  HK_WIDE_SWITCH(wide_pside_k, float, mode, g, q, p, out, s)
}

template <int K>
static void wide_sizes_k(int64_t Ip, int64_t Jp, int64_t* nq, int64_t* np) {
  using S = WideShape<K>;
  *nq = (Ip + S::PAD_ROWS) * S::RGP;
  *np = (Jp + S::PAD_ROWS) * S::RGQ;
}

void stage_sizes_wide(int K, int64_t Ip, int64_t Jp, int64_t* nq, int64_t* np) {
  switch (K) {
    case 9: return wide_sizes_k<9>(Ip, Jp, nq, np);
    case 10: return wide_sizes_k<10>(Ip, Jp, nq, np);
    case 11: return wide_sizes_k<11>(Ip, Jp, nq, np);
    case 12: return wide_sizes_k<12>(Ip, Jp, nq, np);
    case 13: return wide_sizes_k<13>(Ip, Jp, nq, np);
    case 14: return wide_sizes_k<14>(Ip, Jp, nq, np);
    case 15: return wide_sizes_k<15>(Ip, Jp, nq, np);
    case 16: return wide_sizes_k<16>(Ip, Jp, nq, np);
    default: *nq = *np = 0;
  }
}

}  // namespace hk
