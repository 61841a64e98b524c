/*
 * hk_sim.c — seeded synthetic PLINK BED generator for the HaploADMIXTURE likelihood engine.
 *
 * The reference ships no generator; its tests read SnpArrays' EUR_subset.bed
 * (test/runtests.jl:5).  This file writes data of the same format: PLINK 1 BED,
 * SNP-major, individuals packed 4 per byte LSB-first, 2-bit codes
 *   0b00 = homozygous allele 1, 0b01 = missing, 0b10 = heterozygous, 0b11 = homozygous allele 2
 * (decoded by src/loops.jl:55-83).  Block j is SNPs (2j, 2j+1) (0-based), haplotype index
 * h = 2a + b with a the allele at SNP 2j and b the allele at SNP 2j+1 (0 = allele 1).
 *
 * Model (haplotype admixture):
 *   Q_i ~ Dirichlet(alpha_q * 1_K)                                  ancestry proportions
 *   pi_j ~ Dirichlet(1,1,1,1)                                       ancestral haplotype freqs
 *   P_kj ~ Dirichlet(pi_j (1-F)/F)                                  Balding-Nichols drift, Fst = F
 *   each of the two haplotype copies of (i, j): z ~ Cat(Q_i), h ~ Cat(P_zj)
 *   genotype per SNP = number of allele-2 copies; each SNP genotype missing i.i.d. w.p. miss.
 *
 * Randomness is counter based (splitmix64 finaliser over (seed, stream, index)), so the output
 * does not depend on the thread count.  Built as a host-only shared library (gcc -fopenmp).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

static inline uint64_t mix64(uint64_t z) {
  /* splitmix64 output function (Steele, Lea & Flood, OOPSLA 2014). */
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

static inline uint64_t hash3(uint64_t seed, uint64_t stream, uint64_t idx) {
  return mix64(mix64(seed ^ (stream * 0xd1b54a32d192ed03ULL)) ^ idx);
}

/* Sequential generator for the small per-individual / per-block parameter draws. */
typedef struct { uint64_t s; } rng_t;
static inline uint64_t rng_next(rng_t* r) { r->s += 0x9e3779b97f4a7c15ULL; return mix64(r->s); }
static inline double rng_u01(rng_t* r) { return ((rng_next(r) >> 11) + 0.5) * 0x1.0p-53; }
static inline double rng_normal(rng_t* r) {
  double u1 = rng_u01(r), u2 = rng_u01(r);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}
/* Gamma(alpha, 1): Marsaglia & Tsang (2000); alpha < 1 via the U^(1/alpha) boost. */
static double rng_gamma(rng_t* r, double alpha) {
  if (alpha < 1.0) {
    double g = rng_gamma(r, alpha + 1.0);
    return g * pow(rng_u01(r), 1.0 / alpha);
  }
  double d = alpha - 1.0 / 3.0, c = 1.0 / sqrt(9.0 * d);
  for (;;) {
    double x = rng_normal(r), v = 1.0 + c * x;
    if (v <= 0.0) continue;
    v = v * v * v;
    double u = rng_u01(r);
    if (u < 1.0 - 0.0331 * x * x * x * x) return d * v;
    if (log(u) < 0.5 * x * x + d * (1.0 - v + log(v))) return d * v;
  }
}
static void rng_dirichlet(rng_t* r, const double* alpha, int n, double* out) {
  double s = 0.0;
  for (int t = 0; t < n; ++t) { out[t] = rng_gamma(r, alpha[t]); s += out[t]; }
  if (!(s > 0.0)) { for (int t = 0; t < n; ++t) out[t] = 1.0 / n; return; }
  for (int t = 0; t < n; ++t) out[t] /= s;
}

enum { ST_Q = 1, ST_P = 2, ST_G = 3, ST_INIT = 4 };

/*
 * hk_sim_bed: write the BED payload (no 3-byte magic) of I individuals x 2J SNPs.
 *   bed      : out, (2J) * ceil(I/4) bytes, SNP-major
 *   q_true   : out (may be NULL), K x I column-major
 *   p_true   : out (may be NULL), K x 4J column-major, p_true[k + K*(4j+h)]
 * Returns 0 on success, -1 on bad arguments.
 */
int hk_sim_bed(int64_t I, int64_t J, int K, double fst, double miss, double alpha_q, uint64_t seed,
               uint8_t* bed, double* q_true, double* p_true) {
  if (I <= 0 || J <= 0 || K <= 0 || K > 64 || !(fst > 0.0 && fst < 1.0) || miss < 0.0 || miss >= 1.0 ||
      !(alpha_q > 0.0) || !bed)
    return -1;
  const int64_t bps = (I + 3) / 4;
  /* Q: per-individual Dirichlet, cumulative sums kept for sampling. */
  double* qcum = (double*)malloc(sizeof(double) * (size_t)K * (size_t)I);
  if (!qcum) return -1;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < I; ++i) {
    rng_t r = {hash3(seed, ST_Q, (uint64_t)i)};
    double a[64], w[64];
    for (int k = 0; k < K; ++k) a[k] = alpha_q;
    rng_dirichlet(&r, a, K, w);
    double c = 0.0;
    for (int k = 0; k < K; ++k) {
      if (q_true) q_true[k + (int64_t)K * i] = w[k];
      c += w[k];
      qcum[k + (int64_t)K * i] = c;
    }
    qcum[K - 1 + (int64_t)K * i] = 1.0;
  }
  const double scale = (1.0 - fst) / fst;
  const uint32_t miss_thr = (uint32_t)(miss * 4294967296.0);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t j = 0; j < J; ++j) {
    rng_t r = {hash3(seed, ST_P, (uint64_t)j)};
    double one4[4] = {1.0, 1.0, 1.0, 1.0}, pi[4], al[4];
    double pcum[64][4];
    rng_dirichlet(&r, one4, 4, pi);
    for (int h = 0; h < 4; ++h) al[h] = pi[h] * scale > 1e-300 ? pi[h] * scale : 1e-300;
    for (int k = 0; k < K; ++k) {
      double w[4];
      rng_dirichlet(&r, al, 4, w);
      double c = 0.0;
      for (int h = 0; h < 4; ++h) {
        if (p_true) p_true[k + (int64_t)K * (4 * j + h)] = w[h];
        c += w[h];
        pcum[k][h] = c;
      }
      pcum[k][3] = 1.0;
    }
    uint8_t* col1 = bed + (2 * j) * bps;
    uint8_t* col2 = bed + (2 * j + 1) * bps;
    memset(col1, 0, (size_t)bps);
    memset(col2, 0, (size_t)bps);
    const uint64_t base = hash3(seed, ST_G, (uint64_t)j);
    for (int64_t i = 0; i < I; ++i) {
      uint64_t h1 = mix64(base + 2 * (uint64_t)i), h2 = mix64(base + 2 * (uint64_t)i + 1);
      double uz1 = (double)(uint32_t)h1 * 0x1.0p-32, uh1 = (double)(uint32_t)(h1 >> 32) * 0x1.0p-32;
      double uz2 = (double)(uint32_t)h2 * 0x1.0p-32, uh2 = (double)(uint32_t)(h2 >> 32) * 0x1.0p-32;
      const double* qc = qcum + (int64_t)K * i;
      int z1 = 0, z2 = 0;
      while (z1 < K - 1 && uz1 >= qc[z1]) ++z1;
      while (z2 < K - 1 && uz2 >= qc[z2]) ++z2;
      int hap1 = 0, hap2 = 0;
      while (hap1 < 3 && uh1 >= pcum[z1][hap1]) ++hap1;
      while (hap2 < 3 && uh2 >= pcum[z2][hap2]) ++hap2;
      int g1 = (hap1 >> 1) + (hap2 >> 1); /* allele-2 count at SNP 2j   */
      int g2 = (hap1 & 1) + (hap2 & 1);   /* allele-2 count at SNP 2j+1 */
      uint8_t c1 = (uint8_t)(g1 == 0 ? 0 : g1 == 1 ? 2 : 3);
      uint8_t c2 = (uint8_t)(g2 == 0 ? 0 : g2 == 1 ? 2 : 3);
      if (miss_thr) {
        uint64_t hm = mix64(base ^ (0x5bd1e995ULL + 2 * (uint64_t)i) * 0x2545f4914f6cdd1dULL);
        if ((uint32_t)hm < miss_thr) c1 = 1;
        if ((uint32_t)(hm >> 32) < miss_thr) c2 = 1;
      }
      const int sh = (int)(i & 3) * 2;
      col1[i >> 2] |= (uint8_t)(c1 << sh);
      col2[i >> 2] |= (uint8_t)(c2 << sh);
    }
  }
  free(qcum);
  return 0;
}

/*
 * hk_sim_init: random starting point in the reference's layout (src/structs.jl:155,178,189-199):
 * x = U(0,1) draws, q columns normalised to sum 1, each (k, j) 4-vector of p normalised to sum 1.
 * The reference draws from a user RNG (StableRNG in its tests); this uses the seeded counter RNG.
 */
int hk_sim_init(int64_t I, int64_t J, int K, uint64_t seed, double* q, double* p) {
  if (I <= 0 || J <= 0 || K <= 0 || !q || !p) return -1;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < I; ++i) {
    double s = 0.0;
    for (int k = 0; k < K; ++k) {
      uint64_t h = hash3(seed, ST_INIT, (uint64_t)(k + (int64_t)K * i));
      double u = ((h >> 11) + 0.5) * 0x1.0p-53;
      q[k + (int64_t)K * i] = u;
      s += u;
    }
    for (int k = 0; k < K; ++k) q[k + (int64_t)K * i] /= s;
  }
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < J; ++j) {
    for (int k = 0; k < K; ++k) {
      double s = 0.0, u[4];
      for (int h = 0; h < 4; ++h) {
        uint64_t c = (uint64_t)(k + (int64_t)K * (I + 4 * j + h));
        u[h] = ((hash3(seed, ST_INIT, c) >> 11) + 0.5) * 0x1.0p-53;
        s += u[h];
      }
      for (int h = 0; h < 4; ++h) p[k + (int64_t)K * (4 * j + h)] = u[h] / s;
    }
  }
  return 0;
}

/*
 * hk_sim_bed_world: datasets that share one population structure ("world"), for training and evaluating
 * a learned warm start.  Block j of every dataset is block j of a fixed panel: ancestral pi_j ~ Dir(1,1,1,1)
 * and a drift realisation D_kj ~ Dir(pi_j (1-F0)/F0), F0 = 0.15, are drawn from world_seed; a dataset with
 * Fst = F uses P_kj = (1 - lam) pi_j + lam D_kj with lam = sqrt(F / F0) (mean pi_j, variance ~ F pi(1-pi),
 * the Balding-Nichols moments), so population k means the same thing in every dataset.  Individuals (Q_i ~
 * Dir(alpha_q)), haplotype draws and missingness come from data_seed.
 */
int hk_sim_bed_world(int64_t I, int64_t J, int K, double fst, double miss, double alpha_q, uint64_t world_seed,
                     uint64_t data_seed, uint8_t* bed, double* q_true, double* p_true) {
  const double F0 = 0.15;
  if (I <= 0 || J <= 0 || K <= 0 || K > 64 || !(fst > 0.0 && fst <= F0) || miss < 0.0 || miss >= 1.0 ||
      !(alpha_q > 0.0) || !bed)
    return -1;
  const int64_t bps = (I + 3) / 4;
  const double lam = sqrt(fst / F0);
  double* qcum = (double*)malloc(sizeof(double) * (size_t)K * (size_t)I);
  if (!qcum) return -1;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < I; ++i) {
    rng_t r = {hash3(data_seed, ST_Q, (uint64_t)i)};
    double a[64], w[64];
    for (int k = 0; k < K; ++k) a[k] = alpha_q;
    rng_dirichlet(&r, a, K, w);
    double c = 0.0;
    for (int k = 0; k < K; ++k) {
      if (q_true) q_true[k + (int64_t)K * i] = w[k];
      c += w[k];
      qcum[k + (int64_t)K * i] = c;
    }
    qcum[K - 1 + (int64_t)K * i] = 1.0;
  }
  const double scale = (1.0 - F0) / F0;
  const uint32_t miss_thr = (uint32_t)(miss * 4294967296.0);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t j = 0; j < J; ++j) {
    rng_t r = {hash3(world_seed, ST_P, (uint64_t)j)};
    double one4[4] = {1.0, 1.0, 1.0, 1.0}, pi[4], al[4];
    double pcum[64][4];
    rng_dirichlet(&r, one4, 4, pi);
    for (int h = 0; h < 4; ++h) al[h] = pi[h] * scale > 1e-300 ? pi[h] * scale : 1e-300;
    for (int k = 0; k < K; ++k) {
      double d[4];
      rng_dirichlet(&r, al, 4, d);
      double c = 0.0;
      for (int h = 0; h < 4; ++h) {
        const double w = (1.0 - lam) * pi[h] + lam * d[h];
        if (p_true) p_true[k + (int64_t)K * (4 * j + h)] = w;
        c += w;
        pcum[k][h] = c;
      }
      pcum[k][3] = 1.0;
    }
    uint8_t* col1 = bed + (2 * j) * bps;
    uint8_t* col2 = bed + (2 * j + 1) * bps;
    memset(col1, 0, (size_t)bps);
    memset(col2, 0, (size_t)bps);
    const uint64_t base = hash3(data_seed, ST_G, (uint64_t)j);
    for (int64_t i = 0; i < I; ++i) {
      uint64_t h1 = mix64(base + 2 * (uint64_t)i), h2 = mix64(base + 2 * (uint64_t)i + 1);
      double uz1 = (double)(uint32_t)h1 * 0x1.0p-32, uh1 = (double)(uint32_t)(h1 >> 32) * 0x1.0p-32;
      double uz2 = (double)(uint32_t)h2 * 0x1.0p-32, uh2 = (double)(uint32_t)(h2 >> 32) * 0x1.0p-32;
      const double* qc = qcum + (int64_t)K * i;
      int z1 = 0, z2 = 0;
      while (z1 < K - 1 && uz1 >= qc[z1]) ++z1;
      while (z2 < K - 1 && uz2 >= qc[z2]) ++z2;
      int hap1 = 0, hap2 = 0;
      while (hap1 < 3 && uh1 >= pcum[z1][hap1]) ++hap1;
      while (hap2 < 3 && uh2 >= pcum[z2][hap2]) ++hap2;
      int g1 = (hap1 >> 1) + (hap2 >> 1);
      int g2 = (hap1 & 1) + (hap2 & 1);
      uint8_t c1 = (uint8_t)(g1 == 0 ? 0 : g1 == 1 ? 2 : 3);
      uint8_t c2 = (uint8_t)(g2 == 0 ? 0 : g2 == 1 ? 2 : 3);
      if (miss_thr) {
        uint64_t hm = mix64(base ^ (0x5bd1e995ULL + 2 * (uint64_t)i) * 0x2545f4914f6cdd1dULL);
        if ((uint32_t)hm < miss_thr) c1 = 1;
        if ((uint32_t)(hm >> 32) < miss_thr) c2 = 1;
      }
      const int sh = (int)(i & 3) * 2;
      col1[i >> 2] |= (uint8_t)(c1 << sh);
      col2[i >> 2] |= (uint8_t)(c2 << sh);
    }
  }
  free(qcum);
  return 0;
}

/* Number of threads the library was built to use. */
int hk_sim_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
