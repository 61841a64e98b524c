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
//          tie); stop when no label changes.
//   tries  the try with the smallest within-cluster sum of squares over the kept SNPs wins.
// All arithmetic is FP64 with contraction off and every sum runs in a fixed order (SNPs ascending within
// a chunk of SK_CHUNK blocks, chunks ascending), so a CPU restatement reproduces labels and markers exactly.
//
// Layout: the fit's genotype streams, gq [J8][Ip] (8 blocks of one individual per word) for the
// individual-parallel kernels and gp [I8][Jp] (8 individuals of one block per word) for the
// block-parallel kernels; nibble = code(SNP 2b) | code(SNP 2b+1) << 2 with PLINK codes.
#include "hk_device.h"
#include "hk_internal.h"

#include <hipcub/hipcub.hpp>

#include <math.h>
#include <stdio.h>

#include <vector>

#pragma clang fp contract(off)

namespace hk {

namespace {

constexpr int SK_MAXC = HK_MAXK;  // largest number of clusters
constexpr int SK_CHUNK = 1024;    // blocks per partial sum of the all-SNP distance kernel

// Allele-2 dosage of a PLINK code (0b00 -> 0, 0b10 -> 1, 0b11 -> 2); -1 for missing (0b01).
__device__ __forceinline__ int dosage(unsigned code) { return code == 0u ? 0 : (code == 2u ? 1 : (code == 3u ? 2 : -1)); }

// Standardised dosage (0 for a constant SNP); `missing` for a missing entry.
__device__ __forceinline__ double zval(int d, double m, double s, double missing) {
  return d < 0 ? missing : (s > 0.0 ? ((double)d - m) / s : 0.0);
}

// Per-SNP mean and standard deviation over the observed entries (lane = block).  Integer sums, so the
// order does not matter; mean = s / n, var = ss / n - mean^2.
__global__ void k_sk_stats(const uint32_t* __restrict__ gp, int64_t Jp, int64_t J, int64_t I8,
                           double* __restrict__ mean, double* __restrict__ sd) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= J) return;
  int64_t n[2] = {0, 0}, s[2] = {0, 0}, ss[2] = {0, 0};
  for (int64_t iw = 0; iw < I8; ++iw) {
    const uint32_t w = gp[iw * Jp + j];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const unsigned nib = (w >> (4 * t)) & 15u;
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const int d = dosage((nib >> (2 * a)) & 3u);  // padding individuals decode as missing
        if (d >= 0) {
          n[a] += 1;
          s[a] += d;
          ss[a] += d * d;
        }
      }
    }
  }
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    double m = 0.0, v = 0.0;
    if (n[a] > 0) {
      m = (double)s[a] / (double)n[a];
      v = (double)ss[a] / (double)n[a] - m * m;
    }
    mean[2 * j + a] = m;
    sd[2 * j + a] = v > 0.0 ? sqrt(v) : 0.0;
  }
}

// Row of individual r as a center, missing -> 0: row[2j + a].
__global__ void k_sk_row(const uint32_t* __restrict__ gq, int64_t Ip, int64_t J, int64_t r,
                         const double* __restrict__ mean, const double* __restrict__ sd, double* __restrict__ row) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= J) return;
  const unsigned nib = (gq[(j >> 3) * Ip + r] >> (4 * (j & 7))) & 15u;
  for (int a = 0; a < 2; ++a)
    row[2 * j + a] = zval(dosage((nib >> (2 * a)) & 3u), mean[2 * j + a], sd[2 * j + a], 0.0);
}

// Squared distances of every individual to nc center rows over the blocks of chunk blockIdx.y (missing z ->
// 0): part[(ch * nc + c) * I + i], SNPs ascending.  Grid: (ceil(I / 256), ceil(J / SK_CHUNK)).
__global__ void __launch_bounds__(256) k_sk_dist(const uint32_t* __restrict__ gq, int64_t Ip, int64_t I, int64_t J,
                                                 int nc, const double* __restrict__ mean,
                                                 const double* __restrict__ sd, const double* __restrict__ rows,
                                                 double* __restrict__ part) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= I) return;
  const int64_t ch = blockIdx.y;
  const int64_t j0 = ch * SK_CHUNK;
  const int64_t j1 = j0 + SK_CHUNK < J ? j0 + SK_CHUNK : J;
  double acc[SK_MAXC];
#pragma unroll
  for (int c = 0; c < SK_MAXC; ++c) acc[c] = 0.0;
  for (int64_t jw = j0 >> 3; jw * 8 < j1; ++jw) {
    const uint32_t w = gq[jw * Ip + i];
    const int tn = j1 - jw * 8 < 8 ? (int)(j1 - jw * 8) : 8;
    for (int t = 0; t < tn; ++t) {
      const int64_t j = jw * 8 + t;
      const unsigned nib = (w >> (4 * t)) & 15u;
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const double z = zval(dosage((nib >> (2 * a)) & 3u), mean[2 * j + a], sd[2 * j + a], 0.0);
#pragma unroll
        for (int c = 0; c < SK_MAXC; ++c)
          if (c < nc) {
            const double e = z - rows[(int64_t)c * 2 * J + 2 * j + a];
            acc[c] = acc[c] + e * e;
          }
      }
    }
  }
#pragma unroll
  for (int c = 0; c < SK_MAXC; ++c)
    if (c < nc) part[(ch * nc + c) * I + i] = acc[c];
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

// Cluster centers over the observed entries (lane = block) and block scores.  theta [k][2J];
// score [J] (:global) or [k][J] (:local).  Cluster sizes come from the host.
__global__ void __launch_bounds__(256) k_sk_centers(const uint32_t* __restrict__ gp, int64_t Jp, int64_t J, int64_t I,
                                                    int k, const int* __restrict__ labels,
                                                    const int64_t* __restrict__ csize, const double* __restrict__ mean,
                                                    const double* __restrict__ sd, double* __restrict__ theta,
                                                    double* __restrict__ score, int local) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= J) return;
  int n[SK_MAXC][2], s[SK_MAXC][2];
#pragma unroll
  for (int c = 0; c < SK_MAXC; ++c) n[c][0] = n[c][1] = s[c][0] = s[c][1] = 0;
  const int64_t I8 = (I + 7) >> 3;
  for (int64_t iw = 0; iw < I8; ++iw) {
    const uint32_t w = gp[iw * Jp + j];
    const int tn = I - iw * 8 < 8 ? (int)(I - iw * 8) : 8;
    for (int t = 0; t < tn; ++t) {
      const int lab = labels[iw * 8 + t];
      const unsigned nib = (w >> (4 * t)) & 15u;
      // observed count and dosage of each allele of the pair: code 0 -> 0, 2 -> 1, 3 -> 2, 1 missing
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const unsigned code = (nib >> (2 * a)) & 3u;
        const int ob = code != 1u ? 1 : 0;
        const int d = code == 3u ? 2 : (code == 2u ? 1 : 0);
#pragma unroll
        for (int c = 0; c < SK_MAXC; ++c)
          if (c < k && c == lab) {
            n[c][a] += ob;
            s[c][a] += d;
          }
      }
    }
  }
  double g = 0.0;
#pragma unroll
  for (int c = 0; c < SK_MAXC; ++c) {
    if (c >= k) break;
    double sc = 0.0;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const double m = mean[2 * j + a], sdv = sd[2 * j + a];
      double th = 0.0;
      if (n[c][a] > 0 && sdv > 0.0) th = ((double)s[c][a] / (double)n[c][a] - m) / sdv;
      theta[(int64_t)c * 2 * J + 2 * j + a] = th;
      sc = sc + th * th;
    }
    const double v = (double)csize[c] * sc;
    if (local) score[(int64_t)c * J + j] = v;
    else g = g + v;
  }
  if (!local) score[j] = g;
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

// Reassignment over the kept blocks (ascending list, nblocks[0] of them): cost_c = sum (theta_cj^2 - 2 z theta_cj)
// with a missing z = theta of the current cluster; wss[i] = sum over the kept SNPs of (z - theta_new)^2 (a
// missing entry contributes 0); *changed counts the individuals that move.
__global__ void __launch_bounds__(256) k_sk_assign(const uint32_t* __restrict__ gq, int64_t Ip, int64_t I, int64_t J,
                                                   int k, const int* __restrict__ blocks,
                                                   const int* __restrict__ nblocks, const double* __restrict__ mean,
                                                   const double* __restrict__ sd, const double* __restrict__ theta,
                                                   const int* __restrict__ labels, int* __restrict__ labels_out,
                                                   double* __restrict__ wss, int* __restrict__ changed) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= I) return;
  const int cur = labels[i];
  const int nb = *nblocks;
  double cost[SK_MAXC];
#pragma unroll
  for (int c = 0; c < SK_MAXC; ++c) cost[c] = 0.0;
  for (int t = 0; t < nb; ++t) {
    const int64_t b = blocks[t];
    const unsigned nib = (gq[(b >> 3) * Ip + i] >> (4 * (b & 7))) & 15u;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const int64_t jj = 2 * b + a;
      const double z = zval(dosage((nib >> (2 * a)) & 3u), mean[jj], sd[jj], theta[(int64_t)cur * 2 * J + jj]);
#pragma unroll
      for (int c = 0; c < SK_MAXC; ++c)
        if (c < k) {
          const double th = theta[(int64_t)c * 2 * J + jj];
          cost[c] = cost[c] + (th * th - 2.0 * z * th);
        }
    }
  }
  int best = 0;
  double bv = cost[0];
#pragma unroll
  for (int c = 1; c < SK_MAXC; ++c)
    if (c < k && cost[c] < bv) {
      bv = cost[c];
      best = c;
    }
  labels_out[i] = best;
  if (best != cur) atomicAdd(changed, 1);
  double w = 0.0;
  for (int t = 0; t < nb; ++t) {
    const int64_t b = blocks[t];
    const unsigned nib = (gq[(b >> 3) * Ip + i] >> (4 * (b & 7))) & 15u;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const int64_t jj = 2 * b + a;
      const double th = theta[(int64_t)best * 2 * J + jj];
      const double e = zval(dosage((nib >> (2 * a)) & 3u), mean[jj], sd[jj], th) - th;
      w = w + e * e;
    }
  }
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

}  // namespace

#define SKC(x)                                            \
  do {                                                    \
    const hipError_t e_ = (x);                            \
    if (e_ != hipSuccess) {                               \
      snprintf(err, errlen, "%s", hipGetErrorString(e_)); \
      return e_ == hipErrorOutOfMemory ? -3 : -2;         \
    }                                                     \
  } while (0)

int skfr_run(const PassGeom& g, int k, int sparsity, int local, int max_iter, int tries, uint64_t seed,
             const int32_t* init, int32_t* clusters_out, int32_t* aims_out, int* n_aims, double* wss_out,
             int* iters_out, hipStream_t st, char* err, size_t errlen) {
  const int64_t I = g.I, J = g.J;
  if (k < 1 || k > SK_MAXC || k > I || sparsity < 1 || sparsity > J || max_iter < 1 || tries < 1 ||
      (local != 0 && local != 1)) {
    snprintf(err, errlen,
             "invalid SKFR arguments k=%d (1..%d, <= I) sparsity=%d (1..J) max_iter=%d tries=%d mode=%d", k,
             SK_MAXC, sparsity, max_iter, tries, local);
    return -1;
  }
  if (init)
    for (int c = 0; c < k; ++c)
      if (init[c] < 0 || init[c] >= I) {
        snprintf(err, errlen, "init[%d] = %d is not an individual index", c, init[c]);
        return -1;
      }
  const int nch = (int)((J + SK_CHUNK - 1) / SK_CHUNK);
  const int nscore = local ? k : 1;
  DevBuf db;
  double *mean, *sd, *rows, *part, *d2, *theta, *score, *score_sorted, *wss;
  int *labels, *labels2, *idx, *idx_sorted, *flags, *blocks, *nblocks, *changed;
  unsigned* sel;
  int64_t* csize;
  SKC(db.get(&mean, 2 * J));
  SKC(db.get(&sd, 2 * J));
  SKC(db.get(&rows, (size_t)k * 2 * J));
  SKC(db.get(&part, (size_t)nch * k * I));
  SKC(db.get(&d2, I));
  SKC(db.get(&theta, (size_t)k * 2 * J));
  SKC(db.get(&score, (size_t)nscore * J));
  SKC(db.get(&score_sorted, J));
  SKC(db.get(&wss, I));
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

  hipLaunchKernelGGL(k_sk_stats, dim3(nblk(J)), dim3(256), 0, st, g.gp, g.Jp, J, g.I8, mean, sd);
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
    // starting centers: the given individuals, or k-means++ (first uniformly, then proportional to the
    // squared distance to the nearest center chosen so far)
    for (int c = 0; c < k; ++c) {
      int64_t r;
      if (given) {
        r = init[c];
      } else if (c == 0) {
        r = (int64_t)(rng.uniform() * (double)I);
        if (r >= I) r = I - 1;
      } else {
        SKC(hipMemcpyAsync(hd.data(), d2, sizeof(double) * I, hipMemcpyDeviceToHost, st));
        SKC(hipStreamSynchronize(st));
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
      hipLaunchKernelGGL(k_sk_row, dim3(nblk(J)), dim3(256), 0, st, g.gq, g.Ip, J, r, mean, sd, row);
      if (!given && c + 1 < k) {
        hipLaunchKernelGGL(k_sk_dist, gdist, dim3(256), 0, st, g.gq, g.Ip, I, J, 1, mean, sd, (const double*)row, part);
        hipLaunchKernelGGL(k_sk_dist_finish, dim3(nblk(I)), dim3(256), 0, st, (const double*)part, nch, 1, I, 0,
                           c == 0 ? 1 : 0, d2, labels);
      }
      SKC(hipGetLastError());
    }
    hipLaunchKernelGGL(k_sk_dist, gdist, dim3(256), 0, st, g.gq, g.Ip, I, J, k, mean, sd, (const double*)rows, part);
    hipLaunchKernelGGL(k_sk_dist_finish, dim3(nblk(I)), dim3(256), 0, st, (const double*)part, nch, k, I, 1, 0, d2,
                       labels);
    SKC(hipGetLastError());
    int it = 0, moved = 1;
    double wsum = 0.0;
    for (it = 1; it <= max_iter; ++it) {
      SKC(hipMemcpyAsync(hlab.data(), labels, sizeof(int) * I, hipMemcpyDeviceToHost, st));
      SKC(hipStreamSynchronize(st));
      for (int c = 0; c < k; ++c) cs[c] = 0;
      for (int64_t i = 0; i < I; ++i) cs[hlab[i]] += 1;
      SKC(hipMemcpyAsync(csize, cs.data(), sizeof(int64_t) * k, hipMemcpyHostToDevice, st));
      hipLaunchKernelGGL(k_sk_centers, dim3(nblk(J)), dim3(256), 0, st, g.gp, g.Jp, J, I, k, (const int*)labels,
                         (const int64_t*)csize, mean, sd, theta, score, local);
      hipLaunchKernelGGL(k_sk_clear, dim3(nblk(J)), dim3(256), 0, st, sel, J, changed);
      for (int c = 0; c < nscore; ++c) {
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
      hipLaunchKernelGGL(k_sk_assign, dim3(nblk(I)), dim3(256), 0, st, g.gq, g.Ip, I, J, k, (const int*)blocks,
                         (const int*)nblocks, mean, sd, (const double*)theta, (const int*)labels, labels2, wss,
                         changed);
      SKC(hipGetLastError());
      SKC(hipMemcpyAsync(&moved, changed, sizeof(int), hipMemcpyDeviceToHost, st));
      SKC(hipMemcpyAsync(hd.data(), wss, sizeof(double) * I, hipMemcpyDeviceToHost, st));
      SKC(hipStreamSynchronize(st));
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
      SKC(hipMemcpyAsync(best_lab.data(), labels, sizeof(int) * I, hipMemcpyDeviceToHost, st));
      SKC(hipMemcpyAsync(&nb, nblocks, sizeof(int), hipMemcpyDeviceToHost, st));
      SKC(hipStreamSynchronize(st));
      best_blocks.resize(nb);
      if (nb > 0) {
        SKC(hipMemcpyAsync(best_blocks.data(), blocks, sizeof(int) * nb, hipMemcpyDeviceToHost, st));
        SKC(hipStreamSynchronize(st));
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
  return 0;
}

#undef SKC

}  // namespace hk
