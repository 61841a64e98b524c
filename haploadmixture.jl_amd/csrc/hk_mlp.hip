// hk_mlp.hip — learned warm start for Q: genotype random projection and a 3-layer MLP trained and run
// with hand-written kernels (gfx950).
//
// Features (per individual i): f_i = S^{-1/2} sum_{s observed} (g_is - mu_s) R_s, with g the allele-2
// dosage of SNP s (first S = 2 * n_blocks SNPs of the context), mu_s its mean over observed individuals
// and R a seeded Gaussian S x D matrix (Box-Muller over a counter hash, D = 256).
//
// MLP: x (D) -> relu(W1 x + b1) (H) -> relu(W2 . + b2) (H) -> W3 . + b3 (K) -> softmax.  Weights are kept
// as FP32 masters (Adam state) and used in the matrix products as BF16 (round-to-nearest-even); the products
// run on v_mfma_f32_16x16x32_bf16 with FP32 accumulation (FP32 operands as three BF16 terms, k_gemm).  The loss
// is the cross-entropy against the converged Q of the plain method, with the clusters of each batch relabelled
// by the permutation that minimises it (cluster labels of an admixture fit are arbitrary).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "../../include/haplo_hip.h"
#include "hk_device.h"
#include "hk_internal.h"

namespace hk {

typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

// ---------------------------------------------------------------------------------------------------
// Random projection
// ---------------------------------------------------------------------------------------------------

// R[s * D + d] ~ N(0, 1): Box-Muller on two 32-bit halves of one counter hash.
__global__ void k_gauss(float* __restrict__ R, int64_t n, uint64_t seed) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint64_t h = mix64(mix64(seed) ^ (uint64_t)t);
  const float u1 = ((float)(uint32_t)(h >> 40) + 0.5f) * 0x1.0p-24f;
  const float u2 = ((float)(uint32_t)(h & 0xffffff) + 0.5f) * 0x1.0p-24f;
  R[t] = sqrtf(-2.0f * logf(u1)) * cosf(6.2831853f * u2);
}

// dosage of BED code c: 0 -> 0, 2 (het) -> 1, 3 -> 2; code 1 = missing.
__device__ __forceinline__ float dosage(unsigned c) { return c == 3u ? 2.0f : (c == 2u ? 1.0f : 0.0f); }

// mu[s] for s < 2 nb: mean dosage over observed individuals.  Integer dosage sums and observed counts per
// SNP from popcounts of the block-major stream (lane = block, grid.y = chunk of MU_CW individual words;
// with lo / hi the code bits of a SNP, dosage = hi + (hi & lo), missing = lo & ~hi), added with integer
// atomics (exact in any order); then mu = sum / n.
constexpr int MU_CW = 64;
__global__ void k_snp_counts(const uint32_t* __restrict__ gp, int64_t Jp, int nb, int64_t nw,
                             unsigned long long* __restrict__ cnt) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nb) return;
  const int64_t w0 = (int64_t)blockIdx.y * MU_CW, w1 = w0 + MU_CW < nw ? w0 + MU_CW : nw;
  constexpr uint32_t LOW = 0x11111111u;
  unsigned sum[2] = {0u, 0u}, obs[2] = {0u, 0u};
  for (int64_t iw = w0; iw < w1; ++iw) {
    const uint32_t w = gp[iw * Jp + j];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const uint32_t lo = (w >> (2 * a)) & LOW, hi = (w >> (2 * a + 1)) & LOW;
      sum[a] += __builtin_popcount(hi) + __builtin_popcount(hi & lo);
      obs[a] += 8u - __builtin_popcount(lo & ~hi);
    }
  }
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    atomicAdd(&cnt[2 * (2 * j + a)], (unsigned long long)sum[a]);
    atomicAdd(&cnt[2 * (2 * j + a) + 1], (unsigned long long)obs[a]);
  }
}

__global__ void k_fill_u64(unsigned long long* __restrict__ p, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0ull;
}

__global__ void k_snp_mu(const unsigned long long* __restrict__ cnt, int S, float* __restrict__ mu) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  const unsigned long long sum = cnt[2 * s], n = cnt[2 * s + 1];
  mu[s] = n ? (float)((double)sum / (double)n) : 0.0f;
}

// Partial projection over SNP range blockIdx.y (PROJ_SPLIT SNPs) on the matrix cores:
// part[(split * I + i) * D + d] = sum_{s in range} (x_is R_sd - o_is mu_s R_sd), x = allele-2 dosage (0 if
// missing), o = 1 if observed: F = X R - O (mu o R).  X and O are exact in BF16; R and W = -(mu o R) enter
// as BF16 pairs (hi = bf16(v), lo = bf16(v - hi), ~16 mantissa bits), so each 32-SNP slice is four
// v_mfma_f32_16x16x32_bf16 per 16 x 16 tile (X R_hi, X R_lo, O W_hi, O W_lo), accumulated in FP32.
// Workgroup = 64 individuals (one 16-row tile per wave) x 256 features (16 tiles per wave); the slice's
// four B operands (64 KB) are staged in LDS as [op][k group][feature][8 SNPs].
constexpr int PROJ_D = 256;
constexpr int PROJ_SPLIT = 2048;  // SNPs per partial (multiple of 32)
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

// Bg[((slice * 4 + op) * 4 + g) * D + d] (8 BF16: SNPs 32 slice + 8 g + e), op: R_hi, R_lo, W_hi, W_lo;
// SNPs >= S are zero.  Thread per (slice, g, d).
__global__ void k_proj_prep(const float* __restrict__ R, const float* __restrict__ mu, int S, int nslice,
                            u16x8* __restrict__ Bg) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nslice * 4 * PROJ_D) return;
  const int d = (int)(t % PROJ_D);
  const int64_t sg = t / PROJ_D;  // slice * 4 + g
  const int64_t slice = sg >> 2;
  const int g = (int)(sg & 3);
  u16x8 v[4];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int64_t sn = slice * 32 + 8 * g + e;
    const float r = sn < S ? R[sn * PROJ_D + d] : 0.0f;
    const float w = sn < S ? -(mu[sn] * r) : 0.0f;
    const uint16_t rh = f32_to_bf16(r), wh = f32_to_bf16(w);
    v[0][e] = rh;
    v[1][e] = f32_to_bf16(r - bf16_to_f32(rh));
    v[2][e] = wh;
    v[3][e] = f32_to_bf16(w - bf16_to_f32(wh));
  }
#pragma unroll
  for (int op = 0; op < 4; ++op) Bg[((slice * 4 + op) * 4 + g) * PROJ_D + d] = v[op];
}

__global__ __launch_bounds__(256) void k_project(const uint32_t* __restrict__ gq, int64_t Ip, int64_t I, int64_t J8,
                                                 int S, const u16x8* __restrict__ Bg, float* __restrict__ part) {
  __shared__ u16x8 Bs[4 * 4 * PROJ_D];  // 64 KB
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int m = lane & 15, g = lane >> 4;
  const int64_t ia = (int64_t)blockIdx.x * 64 + 16 * wv + m;  // A row of this lane
  const int nslice = (S + 31) / 32;
  const int sl0 = blockIdx.y * (PROJ_SPLIT / 32), sl1 = min(nslice, sl0 + PROJ_SPLIT / 32);
  typedef float v4f_t __attribute__((ext_vector_type(4)));
  v4f_t acc[16];
#pragma unroll
  for (int t = 0; t < 16; ++t) acc[t] = (v4f_t){0.f, 0.f, 0.f, 0.f};
  for (int sl = sl0; sl < sl1; ++sl) {
    // genotype codes of this lane's 8 SNPs (requested before the barrier)
    const int64_t wi = 2 * (int64_t)sl + (g >> 1);
    const uint32_t w = (ia < I && wi < J8) ? gq[wi * Ip + ia] : 0x55555555u;
    const uint32_t h16 = w >> (16 * (g & 1));
    __syncthreads();  // the previous slice has been consumed
    const u16x8* src = Bg + (int64_t)sl * 4 * 4 * PROJ_D;
#pragma unroll
    for (int x = 0; x < 16; ++x) Bs[tid + 256 * x] = src[tid + 256 * x];
    u16x8 ax, ao;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const unsigned c = (h16 >> (2 * e)) & 3u;  // 0b00 -> 0, 0b10 -> 1, 0b11 -> 2, 0b01 missing
      ax[e] = c == 3u ? (unsigned short)0x4000 : (c == 2u ? (unsigned short)0x3F80 : (unsigned short)0);
      ao[e] = c == 1u ? (unsigned short)0 : (unsigned short)0x3F80;
    }
    const bf16x8 aX = __builtin_bit_cast(bf16x8, ax), aO = __builtin_bit_cast(bf16x8, ao);
    __syncthreads();
#pragma unroll
    for (int op = 0; op < 4; ++op) {
      const bf16x8 a = op < 2 ? aX : aO;
#pragma unroll
      for (int t = 0; t < 16; ++t) {
        const bf16x8 b = __builtin_bit_cast(bf16x8, Bs[(op * 4 + g) * PROJ_D + 16 * t + m]);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
      }
    }
  }
  float* out = part + (int64_t)blockIdx.y * I * PROJ_D;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int64_t i = (int64_t)blockIdx.x * 64 + 16 * wv + 4 * g + r;  // D row of acc[t][r]
    if (i < I)
#pragma unroll
      for (int t = 0; t < 16; ++t) out[i * PROJ_D + 16 * t + m] = acc[t][r];
  }
}

// F = scale * (sum of the partials in range order).
__global__ void k_proj_sum(const float* __restrict__ part, int nsplit, int64_t n, float scale,
                           float* __restrict__ F) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  float v = 0.0f;
  for (int sp = 0; sp < nsplit; ++sp) v += part[(int64_t)sp * n + t];
  F[t] = scale * v;
}

// Projection scratch (Gaussian matrix, SNP counts / means, range partials), grown on demand; R is
// regenerated only when (seed, S) change.
struct ProjBuffers {
  float *R = nullptr, *mu = nullptr, *part = nullptr;
  unsigned long long* cnt = nullptr;
  u16x8* Bg = nullptr;  // BF16 operand slices
  int64_t capS = 0, capPart = 0;
  int S_gen = -1;
  uint64_t seed_gen = 0;
  ~ProjBuffers() { release(); }
  void release() {
    for (void* p : {(void*)R, (void*)mu, (void*)part, (void*)cnt, (void*)Bg})
      if (p) (void)hipFree(p);
    R = mu = part = nullptr;
    cnt = nullptr;
    Bg = nullptr;
    capS = capPart = 0;
    S_gen = -1;
  }
};


// ---------------------------------------------------------------------------------------------------
// MLP kernels
// ---------------------------------------------------------------------------------------------------

// C[M x N] (row stride ldc) = A(M x Kd) B(Kd x N) (+ C if accumulate): A FP32 with element strides
// (sam, sak), B BF16 or FP32 with strides (sbk, sbn), FP32 accumulation.
//
// On v_mfma_f32_16x16x32_bf16.  An FP32 operand enters as three BF16 terms x = x_hi + x_mid + x_lo (each the
// rounded remainder of the previous ones, so the three carry the 24 bits of x), and every BF16 x BF16 product
// is exact in the FP32 accumulator: A x BF16 weights takes three matrix instructions per tile and K-slice of 32,
// FP32 x FP32 six (the term pairs down to 2^-16 of the product; the dropped ones are below FP32's rounding).
// Workgroup tile 64 x 64, K-slices of 32 staged in LDS (any strides: each operand is read along its unit stride);
// wave w owns rows 16w .. 16w + 15 and the four 16-column tiles.  Operand layout of the instruction: lane l holds
// A[l % 16][8 (l / 16) .. + 8] and B[8 (l / 16) .. + 8][l % 16]; register r of the result is D[4 (l / 16) + r][l % 16].
constexpr int GT = 64, GK = 32, GP = GK + 8;  // tile edge, K-slice, LDS row pitch (16-byte aligned rows)

template <typename TB>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int Kd, const float* __restrict__ A, int64_t sam,
                                              int64_t sak, const TB* __restrict__ B, int64_t sbk, int64_t sbn,
                                              float* __restrict__ C, int64_t ldc, int accumulate) {
  constexpr bool BF = sizeof(TB) == 2;
  constexpr int NB = BF ? 1 : 3;  // BF16 terms of a B element
  __shared__ __attribute__((aligned(16))) uint16_t As[3][GT * GP];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[NB][GT * GP];  // [column][k]
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int m = lane & 15, g = lane >> 4;
  const int row0 = blockIdx.x * GT, col0 = blockIdx.y * GT;
  v4f acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = (v4f){0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < Kd; k0 += GK) {
    __syncthreads();  // the previous slice has been consumed
#pragma unroll
    for (int u = 0; u < GT * GK / 256; ++u) {
      const int e = tid + 256 * u;
      const int r = sak == 1 ? e / GK : e % GT, k = sak == 1 ? e % GK : e / GT;  // read along A's unit stride
      const int gr = row0 + r, gk = k0 + k;
      const float x = (gr < M && gk < Kd) ? A[(int64_t)gr * sam + (int64_t)gk * sak] : 0.f;
      uint16_t h, md, lo;
      split3(x, h, md, lo);
      As[0][r * GP + k] = h;
      As[1][r * GP + k] = md;
      As[2][r * GP + k] = lo;
    }
#pragma unroll
    for (int u = 0; u < GT * GK / 256; ++u) {
      const int e = tid + 256 * u;
      const int c = sbn == 1 ? e % GT : e / GK, k = sbn == 1 ? e / GT : e % GK;  // read along B's unit stride
      const int gc = col0 + c, gk = k0 + k;
      const bool in = gc < N && gk < Kd;
      if constexpr (BF) {
        Bs[0][c * GP + k] = in ? B[(int64_t)gk * sbk + (int64_t)gc * sbn] : (uint16_t)0;
      } else {
        uint16_t h, md, lo;
        split3(in ? B[(int64_t)gk * sbk + (int64_t)gc * sbn] : 0.f, h, md, lo);
        Bs[0][c * GP + k] = h;
        Bs[1][c * GP + k] = md;
        Bs[2][c * GP + k] = lo;
      }
    }
    __syncthreads();
    bf16x8 a[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) a[q] = *reinterpret_cast<const bf16x8*>(&As[q][(16 * wv + m) * GP + 8 * g]);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      bf16x8 b[NB];
#pragma unroll
      for (int q = 0; q < NB; ++q) b[q] = *reinterpret_cast<const bf16x8*>(&Bs[q][(16 * t + m) * GP + 8 * g]);
      if constexpr (BF) {
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], b[0], acc[t], 0, 0, 0);  // small terms first
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[0], acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[0], acc[t], 0, 0, 0);
      } else {
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], b[0], acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[2], acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[1], acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[0], acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[1], acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[0], acc[t], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = row0 + 16 * wv + 4 * g + v, c = col0 + 16 * t + m;
      if (r < M && c < N) {
        float* o = C + (int64_t)r * ldc + c;
        *o = accumulate ? *o + acc[t][v] : acc[t][v];
      }
    }
}

// Z[n x N] += bias; optionally H = relu(Z) in place.
__global__ void k_bias_act(float* __restrict__ Z, int64_t n, int N, const float* __restrict__ b, int relu) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * N) return;
  float v = Z[t] + b[t % N];
  Z[t] = relu ? fmaxf(v, 0.0f) : v;
}

// Row softmax of Z (n x K) into P.
__global__ void k_softmax(const float* __restrict__ Z, int64_t n, int K, float* __restrict__ P) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* z = Z + i * K;
  float mx = z[0];
  for (int k = 1; k < K; ++k) mx = fmaxf(mx, z[k]);
  float s = 0.f;
  for (int k = 0; k < K; ++k) s += expf(z[k] - mx);
  for (int k = 0; k < K; ++k) P[i * K + k] = expf(z[k] - mx) / s;
}

// Mx[k * K + k2] = sum_i T[i, k2] log P[i, k], in two fixed-order stages.  K <= 8: workgroup b sums rows b,
// b + nb, ... (four row groups of 64 threads, thread = entry (k, k2)) into part[b][64]; k_match_sum adds the nb
// partials.
constexpr int MATCH_BLOCKS = 64;
__global__ __launch_bounds__(256) void k_match(const float* __restrict__ P, const float* __restrict__ T, int64_t n,
                                               int K, double* __restrict__ part) {
  __shared__ double sh[4][64];
  const int e = threadIdx.x & 63, rg = threadIdx.x >> 6;
  double s = 0.0;
  if (e < K * K) {
    const int k = e / K, k2 = e - k * K;
    for (int64_t i = (int64_t)blockIdx.x * 4 + rg; i < n; i += (int64_t)gridDim.x * 4)
      s += (double)T[i * K + k2] * log(fmaxf(P[i * K + k], 1e-30f));
  }
  sh[rg][e] = s;
  __syncthreads();
  if (rg == 0) part[blockIdx.x * 64 + e] = ((sh[0][e] + sh[1][e]) + sh[2][e]) + sh[3][e];
}

// K = 9..16: one thread per entry (K * K <= 256), workgroup b sums rows b, b + nb, ... into part[b][256].
__global__ __launch_bounds__(256) void k_match_w(const float* __restrict__ P, const float* __restrict__ T, int64_t n,
                                                 int K, double* __restrict__ part) {
  const int e = threadIdx.x;
  double s = 0.0;
  if (e < K * K) {
    const int k = e / K, k2 = e - k * K;
    for (int64_t i = blockIdx.x; i < n; i += gridDim.x)
      s += (double)T[i * K + k2] * log(fmaxf(P[i * K + k], 1e-30f));
  }
  part[blockIdx.x * 256 + e] = s;
}

// Mx[e] = sum over b of part[b][e] (row stride 64 or 256), b ascending.
__global__ void k_match_sum(const double* __restrict__ part, int nb, int stride, double* __restrict__ Mx) {
  const int e = threadIdx.x;
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s += part[b * stride + e];
  Mx[e] = s;
}

// Minimum-cost assignment of rows to columns of the K x K matrix C (row-major): perm[k] = column of row k.
// Kuhn-Munkres with row/column potentials and shortest augmenting paths (O(K^3)): row r is added to the
// matching by a Dijkstra-like search over reduced costs C[r][c] - u[r] - v[c] >= 0, then the potentials are
// moved so the reduced costs stay non-negative and those on the matching stay zero.
static void assign_min_cost(const double* C, int K, int* perm) {
  constexpr double INF = 1e300;
  std::vector<double> u(K + 1, 0.0), v(K + 1, 0.0), minv(K + 1);
  std::vector<int> row_of(K + 1, 0), way(K + 1, 0);  // 1-based; column 0 is the virtual root
  std::vector<char> used(K + 1);
  for (int r = 1; r <= K; ++r) {
    row_of[0] = r;
    int c0 = 0;
    std::fill(minv.begin(), minv.end(), INF);
    std::fill(used.begin(), used.end(), 0);
    do {
      used[c0] = 1;
      const int r0 = row_of[c0];
      double delta = INF;
      int c1 = 0;
      for (int c = 1; c <= K; ++c) {
        if (used[c]) continue;
        const double cur = C[(r0 - 1) * K + (c - 1)] - u[r0] - v[c];
        if (cur < minv[c]) { minv[c] = cur; way[c] = c0; }
        if (minv[c] < delta) { delta = minv[c]; c1 = c; }
      }
      for (int c = 0; c <= K; ++c) {
        if (used[c]) { u[row_of[c]] += delta; v[c] -= delta; }
        else minv[c] -= delta;
      }
      c0 = c1;
    } while (row_of[c0] != 0);
    do {  // flip the augmenting path back to the root
      const int c1 = way[c0];
      row_of[c0] = row_of[c1];
      c0 = c1;
    } while (c0);
  }
  for (int c = 1; c <= K; ++c) perm[row_of[c] - 1] = c - 1;
}

// dZ[i, k] = (P[i, k] - T[i, perm[k]]) / n
__global__ void k_dlogits(const float* __restrict__ P, const float* __restrict__ T, int64_t n, int K,
                          const int* __restrict__ perm, float* __restrict__ dZ) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * K) return;
  const int64_t i = t / K;
  const int k = (int)(t - i * K);
  dZ[t] = (P[t] - T[i * K + perm[k]]) / (float)n;
}

// dZ *= [H > 0]
__global__ void k_relu_back(float* __restrict__ dZ, const float* __restrict__ H, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n && !(H[t] > 0.0f)) dZ[t] = 0.0f;
}

// db[c] = sum_i dZ[i, c]: workgroup = 16 columns x 16 row groups (rows i = rg, rg + 16, ...), the 16 partials of a
// column added in a fixed order.
__global__ __launch_bounds__(256) void k_colsum(const float* __restrict__ dZ, int64_t n, int N, float* __restrict__ db) {
  __shared__ float sh[16][17];
  const int cl = threadIdx.x & 15, rg = threadIdx.x >> 4;
  const int c = blockIdx.x * 16 + cl;
  float s = 0.f;
  if (c < N)
    for (int64_t i = rg; i < n; i += 16) s += dZ[i * N + c];
  sh[rg][cl] = s;
  __syncthreads();
  if (rg == 0 && c < N) {
    float t = 0.f;
#pragma unroll
    for (int r = 0; r < 16; ++r) t += sh[r][cl];
    db[c] = t;
  }
}

// Adam on the FP32 masters, then the BF16 copy of the weight matrices (mask[t] = 1 for matrix entries).
__global__ void k_adam(int64_t n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float gt = g[t];
  const float mt = b1 * m[t] + (1.f - b1) * gt;
  const float vt = b2 * v[t] + (1.f - b2) * gt * gt;
  m[t] = mt;
  v[t] = vt;
  w[t] -= lr * (mt / c1) / (sqrtf(vt / c2) + eps);
}

__global__ void k_to_bf16(const float* __restrict__ w, int64_t n, uint16_t* __restrict__ o) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) o[t] = f32_to_bf16(w[t]);
}

}  // namespace hk

using namespace hk;

// ---------------------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------------------

struct hk_mlp {
  int din = 0, width = 0, dout = 0;
  hipStream_t s = nullptr;
  // parameter block (FP32 masters), gradients, Adam moments; layout W1 b1 W2 b2 W3 b3
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  int64_t nparam = 0;
  int64_t off[6] = {0};
  uint16_t* wb = nullptr;  // BF16 copy of the parameter block
  // resident data and activations
  float *X = nullptr, *T = nullptr;
  int64_t ndata = 0;
  float *H1 = nullptr, *H2 = nullptr, *Z3 = nullptr, *P = nullptr, *dZ3 = nullptr, *dH2 = nullptr, *dH1 = nullptr;
  int64_t cap = 0;
  double* Mx = nullptr;
  int* perm = nullptr;
  // device-side prediction: features of a context, projection scratch
  float* Fp = nullptr;
  int64_t capFp = 0;
  hk::ProjBuffers* proj = nullptr;
  char err[256] = {0};
};

static int mlp_fail(hk_mlp* m, int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  if (m) vsnprintf(m->err, sizeof m->err, fmt, ap);
  va_end(ap);
  return code;
}

#define MH(call)                                                                                         \
  do {                                                                                                   \
    hipError_t e_ = (call);                                                                              \
    if (e_ != hipSuccess) return mlp_fail(m, HK_EHIP, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                                          __LINE__);                                                     \
  } while (0)

static unsigned nblk(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

template <typename TB>
static hipError_t gemm(hipStream_t s, int M, int N, int Kd, const float* A, int64_t sam, int64_t sak, const TB* B,
                       int64_t sbk, int64_t sbn, float* C, int64_t ldc, int acc) {
  if (M < 1 || N < 1) return hipSuccess;
  hipLaunchKernelGGL((k_gemm<TB>), dim3((unsigned)((M + GT - 1) / GT), (unsigned)((N + GT - 1) / GT)), dim3(256), 0, s,
                     M, N, Kd, A, sam, sak, B, sbk, sbn, C, ldc, acc);
  return hipGetLastError();
}

static int ensure_cap(hk_mlp* m, int64_t n) {
  if (n <= m->cap) return HK_OK;
  float** bufs[] = {&m->H1, &m->H2, &m->Z3, &m->P, &m->dZ3, &m->dH2, &m->dH1};
  for (float** b : bufs)
    if (*b) {
      MH(hipFree(*b));
      *b = nullptr;
    }
  MH(hipMalloc(&m->H1, sizeof(float) * n * m->width));
  MH(hipMalloc(&m->H2, sizeof(float) * n * m->width));
  MH(hipMalloc(&m->Z3, sizeof(float) * n * m->dout));
  MH(hipMalloc(&m->P, sizeof(float) * n * m->dout));
  MH(hipMalloc(&m->dZ3, sizeof(float) * n * m->dout));
  MH(hipMalloc(&m->dH2, sizeof(float) * n * m->width));
  MH(hipMalloc(&m->dH1, sizeof(float) * n * m->width));
  m->cap = n;
  return HK_OK;
}

// Forward on device rows X (n x din) -> P (softmax), keeping H1, H2 for the backward pass.
static int mlp_forward_dev(hk_mlp* m, const float* X, int64_t n) {
  const int D = m->din, H = m->width, K = m->dout;
  const uint16_t* W1 = m->wb + m->off[0];
  const uint16_t* W2 = m->wb + m->off[2];
  const uint16_t* W3 = m->wb + m->off[4];
  MH(gemm<uint16_t>(m->s, (int)n, H, D, X, D, 1, W1, H, 1, m->H1, H, 0));
  hipLaunchKernelGGL(k_bias_act, dim3(nblk(n * H)), dim3(256), 0, m->s, m->H1, n, H, m->w + m->off[1], 1);
  MH(gemm<uint16_t>(m->s, (int)n, H, H, m->H1, H, 1, W2, H, 1, m->H2, H, 0));
  hipLaunchKernelGGL(k_bias_act, dim3(nblk(n * H)), dim3(256), 0, m->s, m->H2, n, H, m->w + m->off[3], 1);
  MH(gemm<uint16_t>(m->s, (int)n, K, H, m->H2, H, 1, W3, K, 1, m->Z3, K, 0));
  hipLaunchKernelGGL(k_bias_act, dim3(nblk(n * K)), dim3(256), 0, m->s, m->Z3, n, K, m->w + m->off[5], 0);
  hipLaunchKernelGGL(k_softmax, dim3(nblk(n)), dim3(256), 0, m->s, m->Z3, n, K, m->P);
  MH(hipGetLastError());
  return HK_OK;
}

static int refresh_bf16(hk_mlp* m) {
  hipLaunchKernelGGL(k_to_bf16, dim3(nblk(m->nparam)), dim3(256), 0, m->s, m->w, m->nparam, m->wb);
  MH(hipGetLastError());
  return HK_OK;
}

extern "C" {

int hk_mlp_destroy(hk_mlp* m) {
  if (!m) return HK_OK;
  if (m->s) (void)hipStreamSynchronize(m->s);
  void* ptrs[] = {m->w, m->g, m->m, m->v, m->wb, m->X, m->T, m->H1, m->H2, m->Z3, m->P, m->dZ3, m->dH2, m->dH1,
                  m->Mx, m->perm};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (m->Fp) (void)hipFree(m->Fp);
  delete m->proj;
  if (m->s) (void)hipStreamDestroy(m->s);
  delete m;
  return HK_OK;
}

const char* hk_mlp_last_error(const hk_mlp* m) { return m ? m->err : ""; }

int hk_mlp_create(hk_mlp** out, int din, int width, int dout, uint64_t seed) {
  if (!out || din < 1 || width < 1 || dout < 2 || dout > HK_MAX_K) return HK_EINVAL;
  hk_mlp* m = new hk_mlp();
  m->din = din;
  m->width = width;
  m->dout = dout;
  const int64_t sizes[6] = {(int64_t)din * width, width, (int64_t)width * width, width, (int64_t)width * dout, dout};
  int64_t o = 0;
  for (int t = 0; t < 6; ++t) {
    m->off[t] = o;
    o += (sizes[t] + 63) / 64 * 64;
  }
  m->nparam = o;
  int rc = HK_OK;
  auto H = [&](hipError_t e) {
    if (e != hipSuccess && rc == HK_OK) rc = mlp_fail(m, HK_EHIP, "hk_mlp_create: %s", hipGetErrorString(e));
  };
  H(hipStreamCreateWithFlags(&m->s, hipStreamNonBlocking));
  H(hipMalloc(&m->w, sizeof(float) * o));
  H(hipMalloc(&m->g, sizeof(float) * o));
  H(hipMalloc(&m->m, sizeof(float) * o));
  H(hipMalloc(&m->v, sizeof(float) * o));
  H(hipMalloc(&m->wb, sizeof(uint16_t) * o));
  H(hipMalloc(&m->Mx, sizeof(double) * 256 * (1 + MATCH_BLOCKS)));  // result, then the stage-1 partials
  H(hipMalloc(&m->perm, sizeof(int) * HK_MAX_K));
  if (rc == HK_OK) {
    // He-uniform initialisation on the host (seeded), biases zero.
    std::vector<float> w(o, 0.0f);
    uint64_t st = seed * 0x9e3779b97f4a7c15ULL + 1;
    auto u01 = [&]() {
      st += 0x9e3779b97f4a7c15ULL;
      uint64_t z = st;
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
      z ^= z >> 31;
      return ((z >> 11) + 0.5) * 0x1.0p-53;
    };
    const int fan_in[3] = {din, width, width};
    for (int l = 0; l < 3; ++l) {
      const double a = sqrt(6.0 / fan_in[l]);
      for (int64_t t = 0; t < sizes[2 * l]; ++t) w[m->off[2 * l] + t] = (float)((2.0 * u01() - 1.0) * a);
    }
    H(hipMemcpyAsync(m->w, w.data(), sizeof(float) * o, hipMemcpyHostToDevice, m->s));
    H(hipMemsetAsync(m->m, 0, sizeof(float) * o, m->s));
    H(hipMemsetAsync(m->v, 0, sizeof(float) * o, m->s));
    H(hipMemsetAsync(m->g, 0, sizeof(float) * o, m->s));
    if (rc == HK_OK) rc = refresh_bf16(m);
    H(hipStreamSynchronize(m->s));
  }
  if (rc != HK_OK) {
    hk_mlp_destroy(m);
    return rc;
  }
  *out = m;
  return HK_OK;
}

int hk_mlp_num_params(hk_mlp* m, int64_t* n, int64_t* offsets) {
  if (!m || !n) return HK_EINVAL;
  *n = m->nparam;
  if (offsets)
    for (int t = 0; t < 6; ++t) offsets[t] = m->off[t];
  return HK_OK;
}

int hk_mlp_get_params(hk_mlp* m, float* w, float* grad) {
  if (!m) return HK_EINVAL;
  if (w) MH(hipMemcpyAsync(w, m->w, sizeof(float) * m->nparam, hipMemcpyDeviceToHost, m->s));
  if (grad) MH(hipMemcpyAsync(grad, m->g, sizeof(float) * m->nparam, hipMemcpyDeviceToHost, m->s));
  MH(hipStreamSynchronize(m->s));
  return HK_OK;
}

int hk_mlp_set_params(hk_mlp* m, const float* w) {
  if (!m || !w) return HK_EINVAL;
  MH(hipMemcpyAsync(m->w, w, sizeof(float) * m->nparam, hipMemcpyHostToDevice, m->s));
  int rc = refresh_bf16(m);
  if (rc != HK_OK) return rc;
  MH(hipStreamSynchronize(m->s));
  return HK_OK;
}

int hk_mlp_load_data(hk_mlp* m, const float* X, const float* T, int64_t n) {
  if (!m || !X || n < 1) return HK_EINVAL;
  if (m->X) {
    MH(hipFree(m->X));
    m->X = nullptr;
  }
  if (m->T) {
    MH(hipFree(m->T));
    m->T = nullptr;
  }
  MH(hipMalloc(&m->X, sizeof(float) * n * m->din));
  MH(hipMemcpyAsync(m->X, X, sizeof(float) * n * m->din, hipMemcpyHostToDevice, m->s));
  if (T) {
    MH(hipMalloc(&m->T, sizeof(float) * n * m->dout));
    MH(hipMemcpyAsync(m->T, T, sizeof(float) * n * m->dout, hipMemcpyHostToDevice, m->s));
  }
  MH(hipStreamSynchronize(m->s));
  m->ndata = n;
  return HK_OK;
}

int hk_mlp_forward(hk_mlp* m, const float* X, int64_t n, float* P) {
  if (!m || !X || !P || n < 1) return HK_EINVAL;
  float* Xd = nullptr;
  MH(hipMalloc(&Xd, sizeof(float) * n * m->din));
  int rc = HK_OK;
  hipError_t e = hipMemcpyAsync(Xd, X, sizeof(float) * n * m->din, hipMemcpyHostToDevice, m->s);
  if (e == hipSuccess) {
    rc = ensure_cap(m, n);
    if (rc == HK_OK) rc = mlp_forward_dev(m, Xd, n);
    if (rc == HK_OK) e = hipMemcpyAsync(P, m->P, sizeof(float) * n * m->dout, hipMemcpyDeviceToHost, m->s);
    if (e == hipSuccess) e = hipStreamSynchronize(m->s);
  }
  (void)hipFree(Xd);
  if (rc != HK_OK) return rc;
  MH(e);
  return HK_OK;
}

// Forward + backward on resident rows [row0, row0 + n).  match != 0: relabel the target columns by the
// permutation minimising the cross-entropy (chosen on the host from the K x K match matrix).
int hk_mlp_grad(hk_mlp* m, int64_t row0, int64_t n, int match, float* loss, int* perm_out) {
  if (!m || !m->X || !m->T || row0 < 0 || n < 1 || row0 + n > m->ndata) return HK_EINVAL;
  const int D = m->din, H = m->width, K = m->dout;
  const float* X = m->X + row0 * D;
  const float* T = m->T + row0 * K;
  int rc = ensure_cap(m, n);
  if (rc != HK_OK) return rc;
  rc = mlp_forward_dev(m, X, n);
  if (rc != HK_OK) return rc;
  const int nbm = (int)std::min<int64_t>(MATCH_BLOCKS, (n + 255) / 256);
  if (K <= 8) {
    hipLaunchKernelGGL(k_match, dim3(nbm), dim3(256), 0, m->s, m->P, T, n, K, m->Mx + 256);
    hipLaunchKernelGGL(k_match_sum, dim3(1), dim3(64), 0, m->s, m->Mx + 256, nbm, 64, m->Mx);
  } else {
    hipLaunchKernelGGL(k_match_w, dim3(nbm), dim3(256), 0, m->s, m->P, T, n, K, m->Mx + 256);
    hipLaunchKernelGGL(k_match_sum, dim3(1), dim3(256), 0, m->s, m->Mx + 256, nbm, 256, m->Mx);
  }
  double Mx[HK_MAX_K * HK_MAX_K];
  MH(hipMemcpyAsync(Mx, m->Mx, sizeof(double) * K * K, hipMemcpyDeviceToHost, m->s));
  MH(hipStreamSynchronize(m->s));
  int perm[HK_MAX_K], best[HK_MAX_K];
  for (int k = 0; k < K; ++k) perm[k] = best[k] = k;
  double best_v = 0.0;
  for (int k = 0; k < K; ++k) best_v -= Mx[k * K + k];
  if (match && K > 8) {
    // K! is out of reach: the assignment problem on the cost -Mx (exact; ties may resolve differently from
    // the enumeration's first minimum, which needs exactly equal sums)
    double C[HK_MAX_K * HK_MAX_K];
    for (int e = 0; e < K * K; ++e) C[e] = -Mx[e];
    assign_min_cost(C, K, best);
    double v = 0.0;
    for (int k = 0; k < K; ++k) v -= Mx[k * K + best[k]];
    if (v < best_v) best_v = v;
    else
      for (int k = 0; k < K; ++k) best[k] = k;  // the identity is optimal too: keep it, as the enumeration does
  } else if (match) {
    // enumerate permutations (Heap's algorithm order is irrelevant: keep the first minimum)
    std::vector<int> c(K, 0);
    int i = 1;
    while (i < K) {
      if (c[i] < i) {
        const int a = (i % 2 == 0) ? 0 : c[i];
        int tmp = perm[a]; perm[a] = perm[i]; perm[i] = tmp;
        double v = 0.0;
        for (int k = 0; k < K; ++k) v -= Mx[k * K + perm[k]];
        if (v < best_v) { best_v = v; for (int k = 0; k < K; ++k) best[k] = perm[k]; }
        ++c[i];
        i = 1;
      } else {
        c[i] = 0;
        ++i;
      }
    }
  }
  MH(hipMemcpyAsync(m->perm, best, sizeof(int) * K, hipMemcpyHostToDevice, m->s));
  if (loss) *loss = (float)(best_v / (double)n);
  if (perm_out)
    for (int k = 0; k < K; ++k) perm_out[k] = best[k];
  const uint16_t* W2 = m->wb + m->off[2];
  const uint16_t* W3 = m->wb + m->off[4];
  float* gW1 = m->g + m->off[0];
  float* gb1 = m->g + m->off[1];
  float* gW2 = m->g + m->off[2];
  float* gb2 = m->g + m->off[3];
  float* gW3 = m->g + m->off[4];
  float* gb3 = m->g + m->off[5];
  hipLaunchKernelGGL(k_dlogits, dim3(nblk(n * K)), dim3(256), 0, m->s, m->P, T, n, K, m->perm, m->dZ3);
  // layer 3: gW3 (H x K) = H2^T dZ3, gb3, dH2 = dZ3 W3^T
  MH(gemm<float>(m->s, H, K, (int)n, m->H2, 1, H, m->dZ3, K, 1, gW3, K, 0));
  hipLaunchKernelGGL(k_colsum, dim3(nblk(K, 16)), dim3(256), 0, m->s, m->dZ3, n, K, gb3);
  MH(gemm<uint16_t>(m->s, (int)n, H, K, m->dZ3, K, 1, W3, 1, K, m->dH2, H, 0));
  hipLaunchKernelGGL(k_relu_back, dim3(nblk(n * H)), dim3(256), 0, m->s, m->dH2, m->H2, n * H);
  // layer 2
  MH(gemm<float>(m->s, H, H, (int)n, m->H1, 1, H, m->dH2, H, 1, gW2, H, 0));
  hipLaunchKernelGGL(k_colsum, dim3(nblk(H, 16)), dim3(256), 0, m->s, m->dH2, n, H, gb2);
  MH(gemm<uint16_t>(m->s, (int)n, H, H, m->dH2, H, 1, W2, 1, H, m->dH1, H, 0));
  hipLaunchKernelGGL(k_relu_back, dim3(nblk(n * H)), dim3(256), 0, m->s, m->dH1, m->H1, n * H);
  // layer 1
  MH(gemm<float>(m->s, D, H, (int)n, X, 1, D, m->dH1, H, 1, gW1, H, 0));
  hipLaunchKernelGGL(k_colsum, dim3(nblk(H, 16)), dim3(256), 0, m->s, m->dH1, n, H, gb1);
  MH(hipGetLastError());
  MH(hipStreamSynchronize(m->s));
  return HK_OK;
}

int hk_mlp_adam(hk_mlp* m, float lr, float beta1, float beta2, float eps, int step) {
  if (!m || step < 1) return HK_EINVAL;
  const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(nblk(m->nparam)), dim3(256), 0, m->s, m->nparam, m->w, m->g, m->m, m->v, lr, beta1,
                     beta2, eps, c1, c2);
  MH(hipGetLastError());
  int rc = refresh_bf16(m);
  if (rc != HK_OK) return rc;
  MH(hipStreamSynchronize(m->s));
  return HK_OK;
}

}  // extern "C"

// P (I x dout, host) = softmax(MLP(features of the context's individuals)), the features computed on the
// device from the context's genotype streams (first n_blocks blocks) and fed to the network in place.
int hk::mlp_predict_from(hk_mlp* m, const PassGeom& g, int n_blocks, uint64_t seed, float* P) {
  if (!m || !P) return HK_EINVAL;
  if (m->din != PROJ_D) return mlp_fail(m, HK_EINVAL, "network input width %d, features have %d", m->din, PROJ_D);
  if (n_blocks < 1 || n_blocks > g.J)
    return mlp_fail(m, HK_EINVAL, "n_blocks %d outside 1..J=%lld", n_blocks, (long long)g.J);
  const int64_t n = g.I;
  if (m->capFp < n) {
    if (m->Fp) (void)hipFree(m->Fp);
    m->Fp = nullptr;
    m->capFp = 0;
    MH(hipMalloc(&m->Fp, sizeof(float) * n * PROJ_D));
    m->capFp = n;
  }
  if (!m->proj) m->proj = new ProjBuffers();
  MH(project_features_into(g, n_blocks, seed, m->Fp, *m->proj, m->s));
  int rc = ensure_cap(m, n);
  if (rc == HK_OK) rc = mlp_forward_dev(m, m->Fp, n);
  if (rc != HK_OK) return rc;
  MH(hipMemcpyAsync(P, m->P, sizeof(float) * n * m->dout, hipMemcpyDeviceToHost, m->s));
  MH(hipStreamSynchronize(m->s));
  return HK_OK;
}

// F_dev (I x 256, device) = the features of the context's individuals over its first n_blocks blocks.
hipError_t hk::project_features_into(const PassGeom& g, int n_blocks, uint64_t seed, float* F_dev, ProjBuffers& b,
                                     hipStream_t s) {
  const int S = 2 * n_blocks;
  const int64_t nw = (g.I + 7) / 8;
  const int nsplit = (S + PROJ_SPLIT - 1) / PROJ_SPLIT;
  const int64_t nF = g.I * PROJ_D;
  hipError_t e = hipSuccess;
  const int nslice = (S + 31) / 32;
  if (b.capS < S) {
    for (void* p : {(void*)b.R, (void*)b.mu, (void*)b.cnt, (void*)b.Bg})
      if (p) (void)hipFree(p);
    b.R = b.mu = nullptr;
    b.cnt = nullptr;
    b.Bg = nullptr;
    b.capS = 0;
    b.S_gen = -1;
    e = hipMalloc(&b.R, sizeof(float) * (int64_t)S * PROJ_D);
    if (e == hipSuccess) e = hipMalloc(&b.mu, sizeof(float) * S);
    if (e == hipSuccess) e = hipMalloc(&b.cnt, sizeof(unsigned long long) * 2 * S);
    if (e == hipSuccess) e = hipMalloc(&b.Bg, sizeof(u16x8) * (int64_t)nslice * 4 * 4 * PROJ_D);
    if (e != hipSuccess) return e;
    b.capS = S;
  }
  if (b.capPart < nF * nsplit) {
    if (b.part) (void)hipFree(b.part);
    b.part = nullptr;
    b.capPart = 0;
    e = hipMalloc(&b.part, sizeof(float) * nF * nsplit);
    if (e != hipSuccess) return e;
    b.capPart = nF * nsplit;
  }
  if (b.S_gen != S || b.seed_gen != seed) {
    hipLaunchKernelGGL(k_gauss, dim3(nblk((int64_t)S * PROJ_D)), dim3(256), 0, s, b.R, (int64_t)S * PROJ_D, seed);
    b.S_gen = S;
    b.seed_gen = seed;
  }
  hipLaunchKernelGGL(k_fill_u64, dim3(nblk(2 * (int64_t)S)), dim3(256), 0, s, b.cnt, 2 * (int64_t)S);
  hipLaunchKernelGGL(k_snp_counts, dim3(nblk(n_blocks), (unsigned)((nw + MU_CW - 1) / MU_CW)), dim3(256), 0, s, g.gp,
                     g.Jp, n_blocks, nw, b.cnt);
  hipLaunchKernelGGL(k_snp_mu, dim3(nblk(S)), dim3(256), 0, s, (const unsigned long long*)b.cnt, S, b.mu);
  hipLaunchKernelGGL(k_proj_prep, dim3(nblk((int64_t)nslice * 4 * PROJ_D)), dim3(256), 0, s, (const float*)b.R,
                     (const float*)b.mu, S, nslice, b.Bg);
  hipLaunchKernelGGL(k_project, dim3((unsigned)((g.I + 63) / 64), (unsigned)nsplit), dim3(256), 0, s, g.gq, g.Ip, g.I,
                     g.J8, S, (const u16x8*)b.Bg, b.part);
  hipLaunchKernelGGL(k_proj_sum, dim3(nblk(nF)), dim3(256), 0, s, (const float*)b.part, nsplit, nF,
                     1.0f / sqrtf((float)S), F_dev);
  return hipGetLastError();
}

int hk::project_features(const PassGeom& g, int n_blocks, uint64_t seed, float* F_host, hipStream_t s,
                         char* err, size_t errlen) {
  ProjBuffers b;
  float* F = nullptr;
  hipError_t e = hipMalloc(&F, sizeof(float) * g.I * PROJ_D);
  if (e == hipSuccess) e = project_features_into(g, n_blocks, seed, F, b, s);
  if (e == hipSuccess) e = hipMemcpyAsync(F_host, F, sizeof(float) * g.I * PROJ_D, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (F) (void)hipFree(F);
  if (e != hipSuccess) {
    snprintf(err, errlen, "feature projection: %s", hipGetErrorString(e));
    return HK_EHIP;
  }
  return HK_OK;
}
