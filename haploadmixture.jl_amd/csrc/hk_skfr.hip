// hk_skfr.hip — SKFR ancestry-informative-marker pre-selection on the genotype streams (gfx950).
//
// Reference: src/driver.jl:84-131 (_filter_SKFR) calls SparseKmeansFeatureRanking.sparsekmeans1
// (skfr_mode = :global) or sparsekmeans2 (:local), through sparsekmeans_repeat when tries > 1, on an
// ImputedSnpMatrix{Float64}(g, K; blocksize = 2) and keeps the SNPs of the selected blocks.  That package
// is not vendored with the reference; this is the published method (Zhang, Lange & Xu, "Simple and scalable
// sparse k-means clustering via feature ranking", NeurIPS 2020) with features grouped in SNP pairs:
//   z_ij   allele-2 dosage of individual i at SNP j standardised over the observed entries of SNP j,
//          (x_ij - m_j) / s_j, 0 for a constant SNP; a missing entry takes the center of the
//          individual's current cluster (0 while the starting centers are chosen);
//   start  k-means++ on all SNPs from a splitmix64 stream (or the k given individuals), then each
//          individual joins its nearest start;
//   repeat (at most max_iter times)
//          theta_cj = mean of the observed z_ij over the members of cluster c (0 if none);
//          score of block b = sum_c n_c (theta_c,2b^2 + theta_c,2b+1^2) (:global) or, per cluster c,
//          n_c (theta_c,2b^2 + theta_c,2b+1^2) (:local);
//          keep the `sparsity` best blocks (stable: the smaller block wins a tie), theta = 0 elsewhere;
//          reassign i to argmin_c sum over the kept SNPs of theta_cj^2 - 2 z_ij theta_cj (smaller c on a
//          tie; evaluated per partial of the kept list as sum theta^2 - sum (2 z) theta); stop when no
//          label changes.
//   tries  the try with the smallest within-cluster sum of squares over the kept SNPs wins.
//   output labels, and the kept blocks' SNP pairs in decreasing order of importance (docs/src/index.md:287):
//          the score sum_c n_c (theta_c,2b^2 + theta_c,2b+1^2) of the last center update, ties to the
//          smaller block.
// All arithmetic is FP64 with contraction off and every sum runs in a fixed order (SNPs ascending within
// a chunk of SK_CHUNK blocks, chunks ascending), so a CPU restatement reproduces labels and markers exactly.
// The three standardised values of each SNP, (d - m_j) / s_j for d = 0, 1, 2, are tabulated once.
//
// Layout: the fit's genotype streams, gq [J8][Ip] (8 blocks of one individual per word) for the
// individual-parallel kernels and gp [I8][Jp] (8 individuals of one block per word) for the
// block-parallel kernels; nibble = code(SNP 2b) | code(SNP 2b+1) << 2 with PLINK codes.
#include "hk_device.h"
#include "hk_internal.h"

#include <hipcub/hipcub.hpp>

#include <math.h>
#include <stdio.h>

#include <algorithm>
#include <vector>

#pragma clang fp contract(off)

namespace hk {

namespace {

constexpr int SK_MAXC = HK_MAXK;  // largest number of clusters (per-lane cluster arrays)
constexpr int SK_CHUNK = 1024;    // blocks per partial sum of the all-SNP distance kernel
constexpr int SK_SUB = 256;       // blocks (or kept-list entries) per LDS stage
constexpr int SK_CW = 128;        // genotype words (1024 individuals) per partial of the center counts
constexpr int SK_ACH_MIN = 256;   // kept blocks per partial of the reassignment (at least; <= 64 partials)

// Standardised value of a PLINK code at SNP column jj from the table ztab [2J][4]; `missing` for code 0b01.
__device__ __forceinline__ double zval(const double* __restrict__ ztab, int64_t jj, unsigned code, double missing) {
  const int d = code == 0u ? 0 : (code == 2u ? 1 : 2);
  return code == 1u ? missing : ztab[jj * 4 + d];
}

// Per-SNP mean and standard deviation over the observed entries (lane = block).  Integer sums, so the
// order does not matter; mean = s / n, var = ss / n - mean^2.
__global__ void k_sk_stats(const uint32_t* __restrict__ gp, int64_t Jp, int64_t J, int64_t I8,
                           double* __restrict__ mean, double* __restrict__ sd, double* __restrict__ ztab) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= J) return;
  int64_t n[2] = {0, 0}, s[2] = {0, 0}, ss[2] = {0, 0};
  // integer counts from popcounts (padding individuals decode as missing): with lo / hi the code bits of
  // an allele, dosage = hi + (hi & lo), dosage^2 = hi + 3 (hi & lo), missing = lo & ~hi
  constexpr uint32_t LOW = 0x11111111u;
  int64_t iw = 0;
  for (; iw < I8; iw += 4) {
    uint32_t w[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) w[u] = iw + u < I8 ? gp[(iw + u) * Jp + j] : 0x55555555u;
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const uint32_t lo = (w[u] >> (2 * a)) & LOW, hi = (w[u] >> (2 * a + 1)) & LOW;
        const int nh = __builtin_popcount(hi), nhl = __builtin_popcount(hi & lo);
        n[a] += 8 - __builtin_popcount(lo & ~hi);
        s[a] += nh + nhl;
        ss[a] += nh + 3 * nhl;
      }
  }
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    double m = 0.0, v = 0.0;
    if (n[a] > 0) {
      m = (double)s[a] / (double)n[a];
      v = (double)ss[a] / (double)n[a] - m * m;
    }
    const double sdv = v > 0.0 ? sqrt(v) : 0.0;
    mean[2 * j + a] = m;
    sd[2 * j + a] = sdv;
    for (int d = 0; d < 3; ++d) ztab[(2 * j + a) * 4 + d] = sdv > 0.0 ? ((double)d - m) / sdv : 0.0;
    ztab[(2 * j + a) * 4 + 3] = 0.0;
  }
}

// Row of individual r as a center, missing -> 0: row[2j + a].
__global__ void k_sk_row(const uint32_t* __restrict__ gq, int64_t Ip, int64_t J, int64_t r,
                         const double* __restrict__ ztab, double* __restrict__ row) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= J) return;
  const unsigned nib = (gq[(j >> 3) * Ip + r] >> (4 * (j & 7))) & 15u;
  for (int a = 0; a < 2; ++a) row[2 * j + a] = zval(ztab, 2 * j + a, (nib >> (2 * a)) & 3u, 0.0);
}

// Squared distances of every individual to nc center rows over the blocks of chunk blockIdx.y (missing z ->
// 0): part[(ch * nc + c) * I + i], SNPs ascending.  Grid: (ceil(I / 256), ceil(J / SK_CHUNK)); dynamic LDS
// SK_DIST_LDS bytes.  The chunk is staged dist_sub(nc) blocks at a time as a table of the squared
// differences e2[c][snp][code] = (z(code) - row_c,snp)^2 (code 0b01 = missing: z = 0), computed with the
// same operations as the element-wise form, so each element costs one LDS gather and one add per center.
constexpr size_t SK_DIST_LDS = 16384;  // bytes: 2048 doubles of e2 table per stage
// (NC x 4 x 2 dist_sub doubles = SK_DIST_LDS bytes at most: 2048 doubles for every NC up to 16)
__host__ __device__ constexpr int dist_sub(int nc) {
  return nc <= 1 ? 256 : (nc <= 2 ? 128 : (nc <= 4 ? 64 : (nc <= 8 ? 32 : 16)));
}

template <int NC>
__global__ void __launch_bounds__(256) k_sk_dist(const uint32_t* __restrict__ gq, int64_t Ip, int64_t I, int64_t J,
                                                 const double* __restrict__ ztab, const double* __restrict__ rows,
                                                 double* __restrict__ part) {
  extern __shared__ double sk_lds[];
  double* e2 = sk_lds;  // [NC][2 sub][4]
  constexpr int nc = NC;
  constexpr int sub = dist_sub(NC), snps = 2 * sub, NWQ = sub / 8;
  static_assert(NC * 4 * snps * sizeof(double) <= SK_DIST_LDS, "the e2 table of a stage exceeds SK_DIST_LDS");
  const int tid = threadIdx.x;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + tid;
  const bool act = i < I;
  const int64_t ch = blockIdx.y;
  const int64_t j0 = ch * SK_CHUNK;
  const int64_t j1 = j0 + SK_CHUNK < J ? j0 + SK_CHUNK : J;
  double acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = 0.0;
  for (int64_t s0 = j0; s0 < j1; s0 += sub) {
    const int64_t s1 = s0 + sub < j1 ? s0 + sub : j1;
    const int ns = (int)(2 * (s1 - s0));              // SNPs of this stage
    const int64_t w0 = s0 >> 3, nwq = ((s1 + 7) >> 3) - w0;  // s0 is a multiple of 8; nwq <= NWQ
    // the stage's genotype words are requested before the table is built, so their latency overlaps it
    uint32_t wv[NWQ];
#pragma unroll
    for (int q = 0; q < NWQ; ++q) wv[q] = (act && q < nwq) ? gq[(w0 + q) * Ip + i] : 0u;
    __syncthreads();  // the previous stage has been consumed
    for (int c = 0; c < nc; ++c)
      for (int r = tid; r < 4 * ns; r += 256) {  // r = 4 snp + code
        const int code = r & 3;
        const int64_t col = 2 * s0 + (r >> 2);
        const double z = code == 1 ? 0.0 : ztab[col * 4 + (code == 0 ? 0 : code - 1)];
        const double e = z - rows[(int64_t)c * 2 * J + col];
        e2[c * 4 * snps + r] = e * e;
      }
    __syncthreads();
    if (!act) continue;
    if (ns == snps) {  // whole stage: every word full
#pragma unroll
      for (int q = 0; q < NWQ; ++q) {
        const uint32_t w = wv[q];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          const int r = 4 * (16 * q + u) + (int)((w >> (2 * u)) & 3u);
#pragma unroll
          for (int c = 0; c < NC; ++c) acc[c] = acc[c] + e2[c * 4 * snps + r];
        }
      }
    } else {  // the last, partial stage of the SNP range
      for (int q = 0; q < nwq; ++q) {
        uint32_t w = 0u;
#pragma unroll
        for (int x = 0; x < NWQ; ++x)
          if (x == q) w = wv[x];
        const int nu = ns - 16 * q < 16 ? ns - 16 * q : 16;
        for (int u = 0; u < nu; ++u) {
          const int r = 4 * (16 * q + u) + (int)((w >> (2 * u)) & 3u);
#pragma unroll
          for (int c = 0; c < NC; ++c) acc[c] = acc[c] + e2[c * 4 * snps + r];
        }
      }
    }
  }
  if (!act) return;
#pragma unroll
  for (int c = 0; c < NC; ++c) part[(ch * NC + c) * I + i] = acc[c];
}

// Chunk partials summed in chunk order.  nearest == 0 (nc = 1): d2[i] = first ? d : min(d2[i], d);
// nearest == 1: labels[i] = argmin_c d_c (smaller c on a tie).
__global__ void k_sk_dist_finish(const double* __restrict__ part, int nch, int nc, int64_t I, int nearest, int first,
                                 double* __restrict__ d2, int* __restrict__ labels) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= I) return;
  double tot[SK_MAXC];
#pragma unroll
  for (int c = 0; c < SK_MAXC; ++c) tot[c] = 0.0;
  for (int ch = 0; ch < nch; ++ch)
#pragma unroll
    for (int c = 0; c < SK_MAXC; ++c)
      if (c < nc) tot[c] = tot[c] + part[((int64_t)ch * nc + c) * I + i];
  if (!nearest) {
    d2[i] = (first || tot[0] < d2[i]) ? tot[0] : d2[i];
    return;
  }
  int best = 0;
  double bv = tot[0];
#pragma unroll
  for (int c = 1; c < SK_MAXC; ++c)
    if (c < nc && tot[c] < bv) {
      bv = tot[c];
      best = c;
    }
  labels[i] = best;
}

// Cluster membership of the 8 individuals of word iw as nibble masks: memb[iw * 8 + c] has 0xF at the
// nibbles of the members of cluster c (padding individuals belong to none).
__global__ void k_sk_member(const int* __restrict__ labels, int64_t I, int64_t nw, uint32_t* __restrict__ memb) {
  const int64_t iw = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (iw >= nw) return;
  uint32_t M[SK_MAXC];
#pragma unroll
  for (int c = 0; c < SK_MAXC; ++c) M[c] = 0u;
  for (int t = 0; t < 8; ++t) {
    const int64_t i = iw * 8 + t;
    if (i >= I) break;
    const int lab = labels[i];
#pragma unroll
    for (int c = 0; c < SK_MAXC; ++c)
      if (c == lab) M[c] |= 0xFu << (4 * t);
  }
#pragma unroll
  for (int c = 0; c < SK_MAXC; ++c) memb[iw * SK_MAXC + c] = M[c];
}

// Per-cluster allele-2 dosage sums and missing counts of block j over the words of individual chunk
// blockIdx.y (SK_CW words), from popcounts: with lo / hi the low / high code bits of an allele at each
// individual's nibble, dosage = hi + (hi & lo) and missing = lo & ~hi.  Integer partials:
// ps / pm [chunk][k][2][J].
template <int KC>
__global__ void __launch_bounds__(256) k_sk_counts(const uint32_t* __restrict__ gp, int64_t Jp, int64_t J, int64_t nw,
                                                   const uint32_t* __restrict__ memb, int* __restrict__ ps,
                                                   int* __restrict__ pm) {
  constexpr int k = KC;
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= J) return;
  const int64_t ch = blockIdx.y;
  const int64_t w0 = ch * SK_CW;
  const int64_t w1 = w0 + SK_CW < nw ? w0 + SK_CW : nw;
  unsigned sa[SK_MAXC][2], ma[SK_MAXC][2];
#pragma unroll
  for (int c = 0; c < SK_MAXC; ++c) sa[c][0] = sa[c][1] = ma[c][0] = ma[c][1] = 0u;
  constexpr uint32_t LOW = 0x11111111u;
  int64_t iw = w0;
  for (; iw + 4 <= w1; iw += 4) {
    uint32_t w[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) w[u] = gp[(iw + u) * Jp + j];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const uint32_t lo = (w[u] >> (2 * a)) & LOW, hi = (w[u] >> (2 * a + 1)) & LOW;
        const uint32_t v = hi | ((hi & lo) << 1), mis = lo & ~hi;
#pragma unroll
        for (int c = 0; c < SK_MAXC; ++c)
          if (c < k) {
            const uint32_t M = memb[(iw + u) * SK_MAXC + c];
            sa[c][a] += __builtin_popcount(v & M);
            ma[c][a] += __builtin_popcount(mis & M);
          }
      }
    }
  }
  for (; iw < w1; ++iw) {
    const uint32_t w = gp[iw * Jp + j];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const uint32_t lo = (w >> (2 * a)) & LOW, hi = (w >> (2 * a + 1)) & LOW;
      const uint32_t v = hi | ((hi & lo) << 1), mis = lo & ~hi;
#pragma unroll
      for (int c = 0; c < SK_MAXC; ++c)
        if (c < k) {
          const uint32_t M = memb[iw * SK_MAXC + c];
          sa[c][a] += __builtin_popcount(v & M);
          ma[c][a] += __builtin_popcount(mis & M);
        }
    }
  }
#pragma unroll
  for (int c = 0; c < SK_MAXC; ++c)
    if (c < k)
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const int64_t o = ((ch * k + c) * 2 + a) * J + j;
        ps[o] = (int)sa[c][a];
        pm[o] = (int)ma[c][a];
      }
}

// Cluster centers over the observed entries (lane = block) from the integer partials, and block scores.
// theta [k][2J]; score rows [k][J] per cluster (:local only) and the row over all clusters at
// score + gr * J (gr = 0 for :global, k for :local).  Cluster sizes come from the host.
__global__ void k_sk_centers(const int* __restrict__ ps, const int* __restrict__ pm, int nch, int64_t J, int k,
                             const int64_t* __restrict__ csize, const double* __restrict__ mean,
                             const double* __restrict__ sd, double* __restrict__ theta, double* __restrict__ score,
                             int local) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= J) return;
  double g = 0.0;
  for (int c = 0; c < k; ++c) {
    double sc = 0.0;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      int64_t S = 0, Mi = 0;
      for (int ch = 0; ch < nch; ++ch) {
        const int64_t o = (((int64_t)ch * k + c) * 2 + a) * J + j;
        S += ps[o];
        Mi += pm[o];
      }
      const int64_t n = csize[c] - Mi;  // observed entries among the members
      const double m = mean[2 * j + a], sdv = sd[2 * j + a];
      double th = 0.0;
      if (n > 0 && sdv > 0.0) th = ((double)S / (double)n - m) / sdv;
      theta[(int64_t)c * 2 * J + 2 * j + a] = th;
      sc = sc + th * th;
    }
    const double v = (double)csize[c] * sc;
    if (local) score[(int64_t)c * J + j] = v;
    g = g + v;
  }
  score[(int64_t)(local ? k : 0) * J + j] = g;
}

// Sort key of the final ranking: the all-cluster score of a kept block, -1 for the others.
__global__ void k_sk_rank_key(int64_t J, const int* __restrict__ flags, const double* __restrict__ gscore,
                              double* __restrict__ key) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b < J) key[b] = flags[b] ? gscore[b] : -1.0;
}

// Bit c of sel[b] for the first s blocks of a score-sorted list.
__global__ void k_sk_mark(const int* __restrict__ sorted_idx, int64_t s, int c, unsigned* __restrict__ sel) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= s) return;
  atomicOr(&sel[sorted_idx[t]], 1u << c);
}

__global__ void k_sk_clear(unsigned* __restrict__ sel, int64_t n, int* __restrict__ counter) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) sel[t] = 0u;
  if (t == 0) *counter = 0;
}

// theta_c,2b+a = 0 where cluster c does not keep block b (:global: bit 0 for every c); flags[b] = kept by any.
__global__ void k_sk_zero(int64_t J, int k, int local, const unsigned* __restrict__ sel, double* __restrict__ theta,
                          int* __restrict__ flags) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= J) return;
  const unsigned m = sel[b];
  for (int c = 0; c < k; ++c) {
    const bool keep = local ? ((m >> c) & 1u) != 0u : (m & 1u) != 0u;
    if (!keep) {
      theta[(int64_t)c * 2 * J + 2 * b] = 0.0;
      theta[(int64_t)c * 2 * J + 2 * b + 1] = 0.0;
    }
  }
  flags[b] = m != 0u ? 1 : 0;
}

// Kept-list entries per LDS stage of k_sk_wss: SK_SUB up to 8 clusters, half of it above, so that the staging
// (value table 8 sub + theta 2 sub k doubles) stays within 48 KB: kept_lds_bytes(k).  The stage size does not
// change the order of the sums.
__host__ __device__ constexpr int kept_sub(int k) { return k <= 8 ? SK_SUB : SK_SUB / 2; }
__host__ __device__ constexpr size_t kept_lds_bytes(int k) {
  return (size_t)(8 * kept_sub(k) + 2 * kept_sub(k) * k) * 8;
}
static_assert(kept_lds_bytes(SK_MAXC) <= 48 * 1024 && kept_lds_bytes(8) <= 48 * 1024, "kept-list staging exceeds 48 KB");

// Stage kept-list entries [u0, u0 + nu) (nu <= sub) in LDS: zt[8 t + 4 a + d] = value table of SNP 2 b_t + a,
// th[c * 2 sub + 2 t + a] = theta of cluster c there.  Called by all threads between barriers.
__device__ __forceinline__ void stage_kept(const int* __restrict__ blocks, int u0, int nu, int sub, int64_t J, int k,
                                           const double* __restrict__ ztab, const double* __restrict__ theta,
                                           double* zt, double* th) {
  const int tid = threadIdx.x;
  for (int x = tid; x < 8 * nu; x += 256) zt[x] = ztab[8 * (int64_t)blocks[u0 + (x >> 3)] + (x & 7)];
  for (int x = tid; x < 2 * nu * k; x += 256) {
    const int c = x / (2 * nu), r = x - c * 2 * nu;
    th[c * 2 * sub + r] = theta[(int64_t)c * 2 * J + 2 * (int64_t)blocks[u0 + (r >> 1)] + (r & 1)];
  }
}

// Per-chunk constant of the reassignment cost: ca[ch * k + c] = sum over the SNPs of kept-list chunk ch
// (list order) of theta_cj^2.  One thread per (chunk, cluster).
__global__ void k_sk_chunk_a(const int* __restrict__ blocks, const int* __restrict__ nblocks, int ach, int nach,
                             int64_t J, int k, const double* __restrict__ theta, double* __restrict__ ca) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= nach * k) return;
  const int ch = x / k, c = x - ch * k;
  const int nb = *nblocks;
  const int t0 = ch * ach, t1 = t0 + ach < nb ? t0 + ach : nb;
  const double* th = theta + (int64_t)c * 2 * J;
  double a = 0.0;
  for (int t = t0; t < t1; t += 16) {
    // sixteen blocks' values fetched together, then added in list order
    double v[32];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int64_t b = blocks[t + u < t1 ? t + u : t];
      v[2 * u] = th[2 * b];
      v[2 * u + 1] = th[2 * b + 1];
    }
#pragma unroll
    for (int u = 0; u < 16; ++u)
      if (t + u < t1) {
        a = a + v[2 * u] * v[2 * u];
        a = a + v[2 * u + 1] * v[2 * u + 1];
      }
  }
  ca[x] = a;
}

// Reassignment cross term over chunk blockIdx.y of the kept-block list (ascending, nblocks[0] entries, `ach`
// per chunk): pc[(ch * k + c) * I + i] = sum over the chunk's SNPs of (2 z_ij) theta_cj, a missing z taken
// as theta of the current cluster; the cost of cluster c is sum_ch (ca[ch][c] - pc[ch][c][i]).
template <int KC>
__global__ void __launch_bounds__(256) k_sk_cost(const uint32_t* __restrict__ gq, int64_t Ip, int64_t I, int64_t J,
                                                 const int* __restrict__ blocks, const int* __restrict__ nblocks,
                                                 int ach, const double* __restrict__ ztab,
                                                 const double* __restrict__ theta, const int* __restrict__ labels,
                                                 double* __restrict__ pc) {
  constexpr int k = KC;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= I) return;
  const int nb = *nblocks;
  const int t0 = (int)blockIdx.y * ach;
  if (t0 >= nb) return;
  const int t1 = t0 + ach < nb ? t0 + ach : nb;
  const double* thc = theta + (int64_t)labels[i] * 2 * J;
  double cost[SK_MAXC];
#pragma unroll
  for (int c = 0; c < SK_MAXC; ++c) cost[c] = 0.0;
  for (int t = t0; t < t1; t += 4) {
    // four kept blocks per step: genotype words and table values fetched together, added in list order
    const int tn = t1 - t < 4 ? t1 - t : 4;
    int64_t b[4];
    unsigned nib[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) b[u] = blocks[u < tn ? t + u : t];
#pragma unroll
    for (int u = 0; u < 4; ++u) nib[u] = (gq[(b[u] >> 3) * Ip + i] >> (4 * (b[u] & 7))) & 15u;
    double z2[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int64_t jj = 2 * b[u >> 1] + (u & 1);
      z2[u] = 2.0 * zval(ztab, jj, (nib[u >> 1] >> (2 * (u & 1))) & 3u, thc[jj]);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if ((u >> 1) >= tn) break;
      const int64_t jj = 2 * b[u >> 1] + (u & 1);
#pragma unroll
      for (int c = 0; c < SK_MAXC; ++c)
        if (c < k) cost[c] = cost[c] + z2[u] * theta[(int64_t)c * 2 * J + jj];
    }
  }
#pragma unroll
  for (int c = 0; c < SK_MAXC; ++c)
    if (c < k) pc[((int64_t)blockIdx.y * k + c) * I + i] = cost[c];
}

// Chunk costs ca - pc summed in chunk order, argmin (smaller c on a tie) -> labels_out; *changed counts movers.
__global__ void k_sk_pick(const double* __restrict__ pc, const double* __restrict__ ca, const int* __restrict__ nblocks,
                          int ach, int64_t I, int k, const int* __restrict__ labels, int* __restrict__ labels_out,
                          int* __restrict__ changed) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= I) return;
  const int nch = (*nblocks + ach - 1) / ach;
  double tot[SK_MAXC];
#pragma unroll
  for (int c = 0; c < SK_MAXC; ++c) tot[c] = 0.0;
  for (int ch = 0; ch < nch; ++ch)
#pragma unroll
    for (int c = 0; c < SK_MAXC; ++c)
      if (c < k) tot[c] = tot[c] + (ca[ch * k + c] - pc[((int64_t)ch * k + c) * I + i]);
  int best = 0;
  double bv = tot[0];
#pragma unroll
  for (int c = 1; c < SK_MAXC; ++c)
    if (c < k && tot[c] < bv) {
      bv = tot[c];
      best = c;
    }
  labels_out[i] = best;
  if (best != labels[i]) atomicAdd(changed, 1);
}

// Within-cluster sum of squares of the new labels over chunk blockIdx.y of the kept list (a missing entry
// contributes 0): pw[ch * I + i].  Dynamic LDS kept_lds_bytes(k).
__global__ void __launch_bounds__(256) k_sk_wss(const uint32_t* __restrict__ gq, int64_t Ip, int64_t I, int64_t J,
                                                int k, const int* __restrict__ blocks, const int* __restrict__ nblocks,
                                                int ach, const double* __restrict__ ztab,
                                                const double* __restrict__ theta, const int* __restrict__ labels,
                                                double* __restrict__ pw) {
  extern __shared__ double sk_lds[];
  const int sub = kept_sub(k);
  double* zt = sk_lds;
  double* th = sk_lds + 8 * sub;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = i < I;
  const int nb = *nblocks;
  const int t0 = (int)blockIdx.y * ach;
  if (t0 >= nb) return;  // uniform over the workgroup
  const int t1 = t0 + ach < nb ? t0 + ach : nb;
  const int lab = act ? labels[i] : 0;
  double w = 0.0;
  for (int u0 = t0; u0 < t1; u0 += sub) {
    const int nu = t1 - u0 < sub ? t1 - u0 : sub;
    __syncthreads();
    stage_kept(blocks, u0, nu, sub, J, k, ztab, theta, zt, th);
    __syncthreads();
    if (!act) continue;
    for (int t = 0; t < nu; t += 4) {
      const int tn = nu - t < 4 ? nu - t : 4;
      unsigned nib[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int64_t b = blocks[u0 + t + (u < tn ? u : 0)];
        nib[u] = (gq[(b >> 3) * Ip + i] >> (4 * (b & 7))) & 15u;
      }
      double e[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int r = 2 * t + (u < 2 * tn ? u : 0);
        const double tb = th[lab * 2 * sub + r];
        e[u] = zval(zt, r, (nib[u >> 1] >> (2 * (u & 1))) & 3u, tb) - tb;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        if ((u >> 1) >= tn) break;
        w = w + e[u] * e[u];
      }
    }
  }
  if (act) pw[(int64_t)blockIdx.y * I + i] = w;
}

__global__ void k_sk_wss_sum(const double* __restrict__ pw, const int* __restrict__ nblocks, int ach, int64_t I,
                             double* __restrict__ wss) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= I) return;
  const int nch = (*nblocks + ach - 1) / ach;
  double w = 0.0;
  for (int ch = 0; ch < nch; ++ch) w = w + pw[(int64_t)ch * I + i];
  wss[i] = w;
}

__global__ void k_sk_iota(int* __restrict__ v, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) v[t] = (int)t;
}

// splitmix64 (Steele, Lea & Flood 2014): uniform doubles in [0, 1) from the top 53 bits.
struct SplitMix {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct DevBuf {
  std::vector<void*> ptrs;
  ~DevBuf() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  hipError_t get(T** out, size_t n) {
    void* p = nullptr;
    const hipError_t e = hipMalloc(&p, sizeof(T) * (n > 0 ? n : 1));
    if (e == hipSuccess) ptrs.push_back(p);
    *out = (T*)p;
    return e;
  }
};

inline unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

// Device time of one class of launches: start()/stop() bracket a launch on the stream, collect() after a
// stream synchronisation adds the bracketed time.  Inactive (no events) when profiling is off.
struct LaunchTimer {
  hipEvent_t a = nullptr, b = nullptr;
  bool pending = false;
  double ms = 0.0;
  int n = 0;
  hipError_t init(bool on) {
    if (!on) return hipSuccess;
    hipError_t e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    return e;
  }
  ~LaunchTimer() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
  void start(hipStream_t st) {
    if (a) (void)hipEventRecord(a, st);
  }
  void stop(hipStream_t st) {
    if (b) {
      (void)hipEventRecord(b, st);
      pending = true;
    }
  }
  void collect() {
    if (!pending) return;
    float t = 0.f;
    if (hipEventElapsedTime(&t, a, b) == hipSuccess) ms += t;
    ++n;
    pending = false;
  }
};

}  // namespace

// Launch a kernel template instantiated for the cluster count kc (1..SK_MAXC) as KC_.
#define SK_BY_K(kc, LAUNCH)                          \
  do {                                               \
    switch (kc) {                                    \
      case 1: { constexpr int KC_ = 1; LAUNCH; } break; \
      case 2: { constexpr int KC_ = 2; LAUNCH; } break; \
      case 3: { constexpr int KC_ = 3; LAUNCH; } break; \
      case 4: { constexpr int KC_ = 4; LAUNCH; } break; \
      case 5: { constexpr int KC_ = 5; LAUNCH; } break; \
      case 6: { constexpr int KC_ = 6; LAUNCH; } break; \
      case 7: { constexpr int KC_ = 7; LAUNCH; } break; \
      case 8: { constexpr int KC_ = 8; LAUNCH; } break; \
      case 9: { constexpr int KC_ = 9; LAUNCH; } break; \
      case 10: { constexpr int KC_ = 10; LAUNCH; } break; \
      case 11: { constexpr int KC_ = 11; LAUNCH; } break; \
      case 12: { constexpr int KC_ = 12; LAUNCH; } break; \
      case 13: { constexpr int KC_ = 13; LAUNCH; } break; \
      case 14: { constexpr int KC_ = 14; LAUNCH; } break; \
      case 15: { constexpr int KC_ = 15; LAUNCH; } break; \
      default: { constexpr int KC_ = 16; LAUNCH; } break; \
    }                                                \
  } while (0)

#define SKC(x)                                            \
  do {                                                    \
    const hipError_t e_ = (x);                            \
    if (e_ != hipSuccess) {                               \
      snprintf(err, errlen, "%s", hipGetErrorString(e_)); \
      return e_ == hipErrorOutOfMemory ? -3 : -2;         \
    }                                                     \
  } while (0)

int skfr_run(const PassGeom& g, int k, int sparsity, int local, int max_iter, int tries, uint64_t seed,
             const int32_t* init, const int32_t* init_labels, int32_t* clusters_out, int32_t* aims_out, int* n_aims,
             double* wss_out,
             int* iters_out, double* prof, hipStream_t st, char* err, size_t errlen) {
  const int64_t I = g.I, J = g.J;
  if (k < 1 || k > SK_MAXC || k > I || sparsity < 1 || sparsity > J || max_iter < 1 || tries < 1 ||
      (local != 0 && local != 1)) {
    snprintf(err, errlen,
             "invalid SKFR arguments k=%d (1..%d, <= I) sparsity=%d (1..J) max_iter=%d tries=%d mode=%d", k,
             SK_MAXC, sparsity, max_iter, tries, local);
    return -1;
  }
  if (init_labels)  // a label out of 0..k-1 would index past the per-cluster tables on the device
    for (int64_t i = 0; i < I; ++i)
      if (init_labels[i] < 0 || init_labels[i] >= k) {
        snprintf(err, errlen, "init_labels[%lld] = %d is not a cluster of 0..%d", (long long)i, init_labels[i], k - 1);
        return -1;
      }
  if (init)
    for (int c = 0; c < k; ++c)
      if (init[c] < 0 || init[c] >= I) {
        snprintf(err, errlen, "init[%d] = %d is not an individual index", c, init[c]);
        return -1;
      }
  const int nch = (int)((J + SK_CHUNK - 1) / SK_CHUNK);
  const int nsel = local ? k : 1;                 // score rows that select blocks
  const int64_t nw = (I + 7) / 8;                 // genotype words holding individuals
  const int nic = (int)((nw + SK_CW - 1) / SK_CW);  // individual chunks of the center counts
  const int64_t maxnb = local ? (J < (int64_t)k * sparsity ? J : (int64_t)k * sparsity) : sparsity;
  const int ach = (int)std::max<int64_t>(SK_ACH_MIN, (maxnb + 63) / 64);  // kept blocks per cost partial
  const int nach = (int)((maxnb + ach - 1) / ach);
  const int64_t gr = local ? k : 0;               // row of the all-cluster score
  DevBuf db;
  double *mean, *sd, *ztab, *rows, *part, *d2, *theta, *score, *score_sorted, *key, *wss;
  double *pc, *pw, *ca;
  int *labels, *labels2, *idx, *idx_sorted, *flags, *blocks, *nblocks, *changed, *ps, *pm;
  unsigned* sel;
  uint32_t* memb;
  int64_t* csize;
  SKC(db.get(&mean, 2 * J));
  SKC(db.get(&sd, 2 * J));
  SKC(db.get(&ztab, 8 * J));
  SKC(db.get(&rows, (size_t)k * 2 * J));
  SKC(db.get(&part, (size_t)nch * k * I));
  SKC(db.get(&d2, I));
  SKC(db.get(&theta, (size_t)k * 2 * J));
  SKC(db.get(&score, (size_t)(gr + 1) * J));
  SKC(db.get(&score_sorted, J));
  SKC(db.get(&key, J));
  SKC(db.get(&wss, I));
  SKC(db.get(&pc, (size_t)nach * k * I));
  SKC(db.get(&pw, (size_t)nach * I));
  SKC(db.get(&ca, (size_t)nach * k));
  SKC(db.get(&ps, (size_t)nic * k * 2 * J));
  SKC(db.get(&pm, (size_t)nic * k * 2 * J));
  SKC(db.get(&memb, (size_t)nw * SK_MAXC));
  SKC(db.get(&labels, I));
  SKC(db.get(&labels2, I));
  SKC(db.get(&idx, J));
  SKC(db.get(&idx_sorted, J));
  SKC(db.get(&flags, J));
  SKC(db.get(&blocks, J));
  SKC(db.get(&nblocks, 1));
  SKC(db.get(&changed, 1));
  SKC(db.get(&sel, J));
  SKC(db.get(&csize, k));
  size_t tmp_sort = 0, tmp_sel = 0;
  SKC(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, tmp_sort, score, score_sorted, idx, idx_sorted, (int)J, 0,
                                                    64, st));
  SKC(hipcub::DeviceSelect::Flagged(nullptr, tmp_sel, idx, flags, blocks, nblocks, (int)J, st));
  const size_t tmp_bytes = tmp_sort > tmp_sel ? tmp_sort : tmp_sel;
  char* tmp = nullptr;
  SKC(db.get(&tmp, tmp_bytes));


  // prof: device ms and launch counts of the center update (k_sk_member + k_sk_counts + k_sk_centers), the
  // all-SNP distances (k_sk_dist) and the reassignment (k_sk_cost + k_sk_pick + k_sk_wss + k_sk_wss_sum),
  // then the whole call's ms (the reassignment class includes k_sk_chunk_a)
  LaunchTimer t_cen, t_dist, t_asg, t_all;
  SKC(t_cen.init(prof != nullptr));
  SKC(t_dist.init(prof != nullptr));
  SKC(t_asg.init(prof != nullptr));
  SKC(t_all.init(prof != nullptr));
  auto sync = [&]() -> hipError_t {
    const hipError_t e = hipStreamSynchronize(st);
    t_cen.collect();
    t_dist.collect();
    t_asg.collect();
    return e;
  };
  t_all.start(st);
  hipLaunchKernelGGL(k_sk_stats, dim3(nblk(J)), dim3(256), 0, st, g.gp, g.Jp, J, g.I8, mean, sd, ztab);
  hipLaunchKernelGGL(k_sk_iota, dim3(nblk(J)), dim3(256), 0, st, idx, J);
  SKC(hipGetLastError());
  SplitMix rng{seed};
  std::vector<double> hd(I);
  std::vector<int> hlab(I), best_lab(I), best_blocks;
  std::vector<int64_t> cs(k);
  double best_wss = 0.0;
  int best_iters = 0;
  const dim3 gdist(nblk(I), (unsigned)nch);
  for (int tr = 0; tr < tries; ++tr) {
    const bool given = tr == 0 && init != nullptr;
    // a warm start (sparsekmeans_path): the first try begins from given labels -- the first center update
    // reads them -- instead of starting centers and a nearest-center assignment
    const bool from_labels = tr == 0 && init_labels != nullptr;
    if (from_labels) SKC(hipMemcpyAsync(labels, init_labels, sizeof(int) * I, hipMemcpyHostToDevice, st));
    // starting centers: the given individuals, or k-means++ (first uniformly, then proportional to the
    // squared distance to the nearest center chosen so far)
    for (int c = 0; c < k && !from_labels; ++c) {
      int64_t r;
      if (given) {
        r = init[c];
      } else if (c == 0) {
        r = (int64_t)(rng.uniform() * (double)I);
        if (r >= I) r = I - 1;
      } else {
        SKC(hipMemcpyAsync(hd.data(), d2, sizeof(double) * I, hipMemcpyDeviceToHost, st));
        SKC(sync());
        double tot = 0.0;
        for (int64_t i = 0; i < I; ++i) tot = tot + hd[i];
        const double u = rng.uniform() * tot;
        double cum = 0.0;
        r = I - 1;
        for (int64_t i = 0; i < I; ++i) {
          cum = cum + hd[i];
          if (cum > u) {
            r = i;
            break;
          }
        }
      }
      double* row = rows + (int64_t)c * 2 * J;
      hipLaunchKernelGGL(k_sk_row, dim3(nblk(J)), dim3(256), 0, st, g.gq, g.Ip, J, r, (const double*)ztab, row);
      if (!given && c + 1 < k) {
        t_dist.start(st);
        hipLaunchKernelGGL((k_sk_dist<1>), gdist, dim3(256), SK_DIST_LDS, st, g.gq, g.Ip, I, J, (const double*)ztab,
                           (const double*)row, part);
        t_dist.stop(st);
        hipLaunchKernelGGL(k_sk_dist_finish, dim3(nblk(I)), dim3(256), 0, st, (const double*)part, nch, 1, I, 0,
                           c == 0 ? 1 : 0, d2, labels);
      }
      SKC(hipGetLastError());
    }
    if (!from_labels) {
      t_dist.start(st);
      SK_BY_K(k, hipLaunchKernelGGL((k_sk_dist<KC_>), gdist, dim3(256), SK_DIST_LDS, st, g.gq, g.Ip, I, J,
                                    (const double*)ztab, (const double*)rows, part));
      t_dist.stop(st);
      hipLaunchKernelGGL(k_sk_dist_finish, dim3(nblk(I)), dim3(256), 0, st, (const double*)part, nch, k, I, 1, 0, d2,
                         labels);
    }
    SKC(hipGetLastError());
    int it = 0, moved = 1;
    double wsum = 0.0;
    for (it = 1; it <= max_iter; ++it) {
      SKC(hipMemcpyAsync(hlab.data(), labels, sizeof(int) * I, hipMemcpyDeviceToHost, st));
      SKC(sync());
      for (int c = 0; c < k; ++c) cs[c] = 0;
      for (int64_t i = 0; i < I; ++i) cs[hlab[i]] += 1;
      SKC(hipMemcpyAsync(csize, cs.data(), sizeof(int64_t) * k, hipMemcpyHostToDevice, st));
      t_cen.start(st);
      hipLaunchKernelGGL(k_sk_member, dim3(nblk(nw)), dim3(256), 0, st, (const int*)labels, I, nw, memb);
      SK_BY_K(k, hipLaunchKernelGGL((k_sk_counts<KC_>), dim3(nblk(J), (unsigned)nic), dim3(256), 0, st, g.gp, g.Jp,
                                    J, nw, (const uint32_t*)memb, ps, pm));
      hipLaunchKernelGGL(k_sk_centers, dim3(nblk(J)), dim3(256), 0, st, (const int*)ps, (const int*)pm, nic, J, k,
                         (const int64_t*)csize, (const double*)mean, (const double*)sd, theta, score, local);
      t_cen.stop(st);
      hipLaunchKernelGGL(k_sk_clear, dim3(nblk(J)), dim3(256), 0, st, sel, J, changed);
      for (int c = 0; c < nsel; ++c) {
        size_t tb = tmp_bytes;
        SKC(hipcub::DeviceRadixSort::SortPairsDescending((void*)tmp, tb, score + (int64_t)c * J, score_sorted, idx,
                                                          idx_sorted, (int)J, 0, 64, st));
        hipLaunchKernelGGL(k_sk_mark, dim3(nblk(sparsity)), dim3(256), 0, st, (const int*)idx_sorted,
                           (int64_t)sparsity, c, sel);
      }
      hipLaunchKernelGGL(k_sk_zero, dim3(nblk(J)), dim3(256), 0, st, J, k, local, (const unsigned*)sel, theta, flags);
      {
        size_t tb = tmp_bytes;
        SKC(hipcub::DeviceSelect::Flagged((void*)tmp, tb, idx, flags, blocks, nblocks, (int)J, st));
      }
      t_asg.start(st);
      hipLaunchKernelGGL(k_sk_chunk_a, dim3(nblk((int64_t)nach * k)), dim3(256), 0, st, (const int*)blocks,
                         (const int*)nblocks, ach, nach, J, k, (const double*)theta, ca);
      SK_BY_K(k, hipLaunchKernelGGL((k_sk_cost<KC_>), dim3(nblk(I), (unsigned)nach), dim3(256), 0, st, g.gq, g.Ip, I,
                                    J, (const int*)blocks, (const int*)nblocks, ach, (const double*)ztab,
                                    (const double*)theta, (const int*)labels, pc));
      hipLaunchKernelGGL(k_sk_pick, dim3(nblk(I)), dim3(256), 0, st, (const double*)pc, (const double*)ca,
                         (const int*)nblocks, ach, I, k, (const int*)labels, labels2, changed);
      hipLaunchKernelGGL(k_sk_wss, dim3(nblk(I), (unsigned)nach), dim3(256), kept_lds_bytes(k), st, g.gq, g.Ip, I, J, k,
                         (const int*)blocks, (const int*)nblocks, ach, (const double*)ztab, (const double*)theta,
                         (const int*)labels2, pw);
      hipLaunchKernelGGL(k_sk_wss_sum, dim3(nblk(I)), dim3(256), 0, st, (const double*)pw, (const int*)nblocks, ach, I,
                         wss);
      t_asg.stop(st);
      SKC(hipGetLastError());
      SKC(hipMemcpyAsync(&moved, changed, sizeof(int), hipMemcpyDeviceToHost, st));
      SKC(hipMemcpyAsync(hd.data(), wss, sizeof(double) * I, hipMemcpyDeviceToHost, st));
      SKC(sync());
      std::swap(labels, labels2);
      wsum = 0.0;
      for (int64_t i = 0; i < I; ++i) wsum = wsum + hd[i];
      if (moved == 0) break;
    }
    if (it > max_iter) it = max_iter;
    if (tr == 0 || wsum < best_wss) {
      best_wss = wsum;
      best_iters = it;
      int nb = 0;
      // kept blocks ranked by the all-cluster score of the last center update
      hipLaunchKernelGGL(k_sk_rank_key, dim3(nblk(J)), dim3(256), 0, st, J, (const int*)flags,
                         (const double*)(score + gr * J), key);
      size_t tb = tmp_bytes;
      SKC(hipcub::DeviceRadixSort::SortPairsDescending((void*)tmp, tb, key, score_sorted, idx, idx_sorted, (int)J, 0,
                                                        64, st));
      SKC(hipMemcpyAsync(best_lab.data(), labels, sizeof(int) * I, hipMemcpyDeviceToHost, st));
      SKC(hipMemcpyAsync(&nb, nblocks, sizeof(int), hipMemcpyDeviceToHost, st));
      SKC(sync());
      best_blocks.resize(nb);
      if (nb > 0) {
        SKC(hipMemcpyAsync(best_blocks.data(), idx_sorted, sizeof(int) * nb, hipMemcpyDeviceToHost, st));
        SKC(sync());
      }
    }
  }
  for (int64_t i = 0; i < I; ++i) clusters_out[i] = best_lab[i];
  const int na = (int)best_blocks.size();
  for (int t = 0; t < na; ++t) {
    aims_out[2 * t] = 2 * best_blocks[t];
    aims_out[2 * t + 1] = 2 * best_blocks[t] + 1;
  }
  *n_aims = 2 * na;
  if (wss_out) *wss_out = best_wss;
  if (iters_out) *iters_out = best_iters;
  if (prof) {
    t_all.stop(st);
    SKC(sync());
    t_all.collect();
    const double v[8] = {t_cen.ms, (double)t_cen.n, t_dist.ms, (double)t_dist.n, t_asg.ms, (double)t_asg.n, t_all.ms, 0.0};
    for (int t = 0; t < 8; ++t) prof[t] = v[t];
  }
  return 0;
}

#undef SKC
#undef SK_BY_K

}  // namespace hk
