/*
 * hk_oracle.c — CPU ORACLE for the HaploADMIXTURE likelihood engine.  TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C FP64 restatement of the reference's per-(individual, SNP-pair) loops and of the
 * QN-accelerated SQP loop that drives them.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker or the timed CPU baseline; the product
 * path (libhaplo_hip.so) never calls it.
 *
 * Reference mapping (paths relative to the HaploADMIXTURE.jl repository):
 *   decode                 src/loops.jl:1-2, 55-83    (@gcoefs_snparray, g_map)
 *   qp                     src/loops.jl:15-36          (qp_block!)
 *   log-likelihood terms   src/loops.jl:85-104         (@coefs_likelihood), :236-258 (loop2)
 *   EM / gradient coefs    src/loops.jl:114-155        (@coefs)
 *   em_loop!, em_p_loop!   src/loops.jl:260-326
 *   update_q_loop!         src/loops.jl:328-365
 *   update_p_loop!         src/loops.jl:367-403
 *   count_double_missing!  src/utils.jl:1-18
 *   project_p!             src/utils.jl:20-28 (-> OpenADMIXTURE.project_q!, restated below)
 *   em!                    src/algorithms_inner.jl:29-84
 *   update_q!, update_p!   src/algorithms_inner.jl:97-211, 224-355
 *   init_em!, admixture_qn! src/algorithms_outer.jl:1-22, 38-172
 * Third-party pieces restated from their published algorithms (OpenADMIXTURE.jl 0.1.0, git-tree
 * a3f2f3e104c84384ca0424a52be197128318efdc, not present here):
 *   create_tableau!/quadratic_program!  -> dense primal active-set QP (unique optimum of a strictly
 *                                          convex QP; pinned by KKT and brute-force enumeration tests)
 *   project_q!                           -> Euclidean projection onto {x >= eps, sum x = 1}
 *                                          (sort-based, Held/Wolfe/Crowder; pinned by its KKT form)
 *   update_UV!/update_QN!                -> Zhou-Alexander-Lange (2011) quasi-Newton extrapolation
 *                                          x_qn = F(x) - V (U'U - U'V)^{-1} U'(x - F(x))
 *
 * Layouts are the reference's (src/structs.jl:14-139), 0-based here:
 *   bed payload  bed[s * bps + (i >> 2)], bps = ceil(I_file / 4), SNP s, 2 bits per individual LSB-first
 *   q            K x I   column-major, q[k + K*i]
 *   p            K x 4J  column-major, p[k + K*(4j + h)], h = 0..3 = haplotypes 00,01,10,11
 *   XtX_q        K x K x I,  XtX[k2 + K*(k + K*i)]        (reference XtX[k2, k, i])
 *   XtX_p        K x K x 4J, XtX[k2 + K*(k + K*(4j + h))]
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define HKO_MAXK 32
#define HKO_QP_MAXN 128
#define HKO_QP_MAXM 32

/* ---------------------------------------------------------------------------------------------- */
/* Per-pair coefficients, literal restatement                                                      */
/* ---------------------------------------------------------------------------------------------- */

typedef struct {
  double g1, g2;
  double g10, g11, g12, g1M, g1B;
  double g20, g21, g22, g2M, g2B;
} gco_t;

/* loops.jl:1-2 (g_map = [0, 3, 1, 2]) and :55-83 (@gcoefs_snparray). */
static inline double g_of_code(unsigned c) { return c == 0 ? 0.0 : c == 1 ? 3.0 : c == 2 ? 1.0 : 2.0; }

static inline void gcoefs(const uint8_t* bed, int64_t bps, int64_t i, int64_t j, gco_t* o) {
  uint8_t blk1 = bed[(2 * j) * bps + (i >> 2)];
  uint8_t blk2 = bed[(2 * j + 1) * bps + (i >> 2)];
  int re = (int)(i % 4);
  unsigned g1_pre = (unsigned)(blk1 >> (re << 1)) & 3u;
  unsigned g2_pre = (unsigned)(blk2 >> (re << 1)) & 3u;
  double g1 = g_of_code(g1_pre), g2 = g_of_code(g2_pre);
  o->g1 = g1;
  o->g2 = g2;
  o->g11 = (double)(g1 == 1);
  o->g12 = (double)(g1 == 2);
  o->g10 = (double)(g1 == 0);
  o->g1M = (double)(g1 == 3);
  o->g1B = (double)(g1 != 3);
  o->g21 = (double)(g2 == 1);
  o->g22 = (double)(g2 == 2);
  o->g20 = (double)(g2 == 0);
  o->g2M = (double)(g2 == 3);
  o->g2B = (double)(g2 != 3);
}

/* loops.jl:15-36 (qp_block!, one element): qp_h = sum_k q[k,i] p[k, 4j+h], accumulated in k order. */
static inline void qp_pair(const double* q, const double* p, int K, int64_t i, int64_t j, double qp[4]) {
  double a = 0.0, b = 0.0, c = 0.0, d = 0.0;
  const double* qi = q + (int64_t)K * i;
  const double* pj = p + (int64_t)K * 4 * j;
  for (int k = 0; k < K; ++k) {
    a += qi[k] * pj[k];
    b += qi[k] * pj[K + k];
    c += qi[k] * pj[2 * K + k];
    d += qi[k] * pj[3 * K + k];
  }
  qp[0] = a; qp[1] = b; qp[2] = c; qp[3] = d;
}

/* loops.jl:85-104 (@coefs_likelihood) and :256 (r += log(a1 + a2) + log(b1 + b2)). */
static inline double ll_pair(const gco_t* g, const double qp[4]) {
  const double qp00 = qp[0], qp01 = qp[1], qp10 = qp[2], qp11 = qp[3];
  const double g10 = g->g10, g11 = g->g11, g12 = g->g12, g1M = g->g1M;
  const double g20 = g->g20, g21 = g->g21, g22 = g->g22, g2M = g->g2M;
  double a1 = g1M * g2M + g11 * g21 * qp00 * qp11 +
              ((g1M + g10) * (g20 + g21) + g2M * (g10 + g11) + g11 * g20) * qp00 +
              ((g1M + g10 + g11) * g22) * qp01 + (g12 * (g2M + g20 + g21)) * qp10 + (g12 * g22) * qp11;
  double a2 = g11 * g21 * qp01 * qp10 + g1M * (g20 + g21) * qp10 + g2M * (g10 + g11) * qp01 +
              (g1M * g22 + g12 * g2M) * qp11;
  double b1 = g1M * g2M + g11 * g21 + (g10 * g20 + g1M * g20 + g10 * g2M) * qp00 +
              (g1M + g10) * (g21 + g22) * qp01 + (g11 + g12) * (g2M + g20) * qp10 +
              (g11 * g22 + g12 * (g21 + g22)) * qp11;
  double b2 = (g1M * g21 + g1M * g22 + g11 * g2M + g12 * g2M) * qp11 + g1M * g20 * qp10 + g10 * g2M * qp01;
  return log(a1 + a2) + log(b1 + b2);
}

/* loops.jl:114-155 (@coefs): posterior expected haplotype counts c00, c01, c10, c11. */
static inline void coefs(const gco_t* g, const double qp[4], double c[4]) {
  const double qp00 = qp[0], qp01 = qp[1], qp10 = qp[2], qp11 = qp[3];
  const double g1 = g->g1, g2 = g->g2;
  const double g10 = g->g10, g11 = g->g11, g12 = g->g12, g1M = g->g1M, g1B = g->g1B;
  const double g20 = g->g20, g21 = g->g21, g22 = g->g22, g2M = g->g2M, g2B = g->g2B;
  double a0001 = qp00 / (qp00 + qp01);
  double a0010 = qp00 / (qp00 + qp10);
  double a0100 = qp01 / (qp01 + qp00);
  double a0111 = qp01 / (qp01 + qp11);
  double a1000 = qp10 / (qp10 + qp00);
  double a1011 = qp10 / (qp10 + qp11);
  double a1101 = qp11 / (qp11 + qp01);
  double a1110 = qp11 / (qp11 + qp10);
  double ahet0011 = qp00 * qp11 / (qp00 * qp11 + qp01 * qp10);
  double ahet0110 = qp01 * qp10 / (qp00 * qp11 + qp01 * qp10);

  double c00 = 0.0;
  c00 += g11 * g21 * ahet0011;
  c00 += g1M * g2B * (2 - g2) * a0010;
  c00 += g1B * g2M * (2 - g1) * a0001;
  c00 += g10 * g2B * (2 - g2);
  c00 += g11 * g20;

  double c01 = 0.0;
  c01 += g11 * g21 * ahet0110;
  c01 += g1M * g2B * g2 * a0111;
  c01 += g1B * g2M * (2 - g1) * a0100;
  c01 += g10 * g2B * g2;
  c01 += g11 * g22;

  double c10 = 0.0;
  c10 += g11 * g21 * ahet0110;
  c10 += g1M * g2B * (2 - g2) * a1000;
  c10 += g1B * g2M * g1 * a1011;
  c10 += g12 * g2B * (2 - g2);
  c10 += g11 * g20;

  double c11 = 0.0;
  c11 += g11 * g21 * ahet0011;
  c11 += g1M * g2B * g2 * a1101;
  c11 += g1B * g2M * g1 * a1110;
  c11 += g12 * g2B * g2;
  c11 += g11 * g22;
  c[0] = c00; c[1] = c01; c[2] = c10; c[3] = c11;
}

/* ---------------------------------------------------------------------------------------------- */
/* The five passes (loops.jl:236-403).  Reductions run in a fixed order (per-i over j ascending,    */
/* per-(j,h) over i ascending), so results do not depend on the thread count.                      */
/* ---------------------------------------------------------------------------------------------- */

/* utils.jl:1-18, restricted to the first J blocks (the generator writes exactly 2J SNPs). */
void hko_count_double_missing(const uint8_t* bed, int64_t bps, int64_t I, int64_t J, int64_t* dm) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < I; ++i) {
    int64_t n = 0;
    for (int64_t j = 0; j < J; ++j) {
      unsigned c1 = (bed[(2 * j) * bps + (i >> 2)] >> ((i & 3) * 2)) & 3u;
      unsigned c2 = (bed[(2 * j + 1) * bps + (i >> 2)] >> ((i & 3) * 2)) & 3u;
      if (c1 == 1u && c2 == 1u) ++n;
    }
    dm[i] = n;
  }
}

/* Per-individual log-likelihood rows; loops.jl:236-258. */
void hko_loglik_rows(const uint8_t* bed, int64_t bps, int64_t I, int64_t J, int K, const double* q,
                     const double* p, double* ll_i) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < I; ++i) {
    double r = 0.0;
    for (int64_t j = 0; j < J; ++j) {
      gco_t g;
      double qp[4];
      gcoefs(bed, bps, i, j, &g);
      qp_pair(q, p, K, i, j, qp);
      r += ll_pair(&g, qp);
    }
    ll_i[i] = r;
  }
}

double hko_loglik(const uint8_t* bed, int64_t bps, int64_t I, int64_t J, int K, const double* q,
                  const double* p) {
  double* rows = (double*)malloc(sizeof(double) * (size_t)I);
  if (!rows) return NAN;
  hko_loglik_rows(bed, bps, I, J, K, q, p, rows);
  double r = 0.0;
  for (int64_t i = 0; i < I; ++i) r += rows[i];
  free(rows);
  return r;
}

/* em_loop! (loops.jl:260-295) / em_p_loop! (:297-326): unnormalised sums.  q_next is skipped when
 * fix_q != 0 (em_p_loop!), p_next when fix_p != 0. */
void hko_em(const uint8_t* bed, int64_t bps, int64_t I, int64_t J, int K, const double* q, const double* p,
            double* q_next, double* p_next, int fix_q, int fix_p) {
  if (!fix_q) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < I; ++i) {
      double* qn = q_next + (int64_t)K * i;
      const double* qi = q + (int64_t)K * i;
      for (int k = 0; k < K; ++k) qn[k] = 0.0;
      for (int64_t j = 0; j < J; ++j) {
        gco_t g;
        double qp[4], c[4];
        gcoefs(bed, bps, i, j, &g);
        qp_pair(q, p, K, i, j, qp);
        coefs(&g, qp, c);
        const double* pj = p + (int64_t)K * 4 * j;
        for (int k = 0; k < K; ++k) {
          qn[k] += c[0] * qi[k] * pj[k] / qp[0];
          qn[k] += c[1] * qi[k] * pj[K + k] / qp[1];
          qn[k] += c[2] * qi[k] * pj[2 * K + k] / qp[2];
          qn[k] += c[3] * qi[k] * pj[3 * K + k] / qp[3];
        }
      }
    }
  }
  if (!fix_p) {
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < J; ++j) {
      double* pn = p_next + (int64_t)K * 4 * j;
      const double* pj = p + (int64_t)K * 4 * j;
      for (int t = 0; t < 4 * K; ++t) pn[t] = 0.0;
      for (int64_t i = 0; i < I; ++i) {
        gco_t g;
        double qp[4], c[4];
        gcoefs(bed, bps, i, j, &g);
        qp_pair(q, p, K, i, j, qp);
        coefs(&g, qp, c);
        const double* qi = q + (int64_t)K * i;
        for (int k = 0; k < K; ++k) {
          pn[k] += c[0] * qi[k] * pj[k] / qp[0];
          pn[K + k] += c[1] * qi[k] * pj[K + k] / qp[1];
          pn[2 * K + k] += c[2] * qi[k] * pj[2 * K + k] / qp[2];
          pn[3 * K + k] += c[3] * qi[k] * pj[3 * K + k] / qp[3];
        }
      }
    }
  }
}

/* update_q_loop! (loops.jl:328-365). */
void hko_update_q(const uint8_t* bed, int64_t bps, int64_t I, int64_t J, int K, const double* q,
                  const double* p, double* XtX, double* Xtz) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < I; ++i) {
    double* A = XtX + (int64_t)K * K * i;
    double* z = Xtz + (int64_t)K * i;
    for (int t = 0; t < K * K; ++t) A[t] = 0.0;
    for (int k = 0; k < K; ++k) z[k] = 0.0;
    for (int64_t j = 0; j < J; ++j) {
      gco_t g;
      double qp[4], c[4];
      gcoefs(bed, bps, i, j, &g);
      qp_pair(q, p, K, i, j, qp);
      coefs(&g, qp, c);
      const double* pj = p + (int64_t)K * 4 * j;
      const double s0 = qp[0] * qp[0], s1 = qp[1] * qp[1], s2 = qp[2] * qp[2], s3 = qp[3] * qp[3];
      for (int k = 0; k < K; ++k) {
        z[k] += c[0] * pj[k] / qp[0];
        z[k] += c[1] * pj[K + k] / qp[1];
        z[k] += c[2] * pj[2 * K + k] / qp[2];
        z[k] += c[3] * pj[3 * K + k] / qp[3];
        for (int k2 = 0; k2 < K; ++k2) {
          double* a = A + k2 + K * k;
          *a += c[0] * pj[k] * pj[k2] / s0;
          *a += c[1] * pj[K + k] * pj[K + k2] / s1;
          *a += c[2] * pj[2 * K + k] * pj[2 * K + k2] / s2;
          *a += c[3] * pj[3 * K + k] * pj[3 * K + k2] / s3;
        }
      }
    }
  }
}

/* update_p_loop! (loops.jl:367-403). */
void hko_update_p(const uint8_t* bed, int64_t bps, int64_t I, int64_t J, int K, const double* q,
                  const double* p, double* XtX, double* Xtz) {
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < J; ++j) {
    double* A = XtX + (int64_t)K * K * 4 * j;
    double* z = Xtz + (int64_t)K * 4 * j;
    for (int t = 0; t < 4 * K * K; ++t) A[t] = 0.0;
    for (int t = 0; t < 4 * K; ++t) z[t] = 0.0;
    for (int64_t i = 0; i < I; ++i) {
      gco_t g;
      double qp[4], c[4];
      gcoefs(bed, bps, i, j, &g);
      qp_pair(q, p, K, i, j, qp);
      coefs(&g, qp, c);
      const double* qi = q + (int64_t)K * i;
      for (int h = 0; h < 4; ++h) {
        const double sq = qp[h] * qp[h];
        double* Ah = A + (int64_t)K * K * h;
        for (int k = 0; k < K; ++k) {
          z[h * K + k] += c[h] * qi[k] / qp[h];
          for (int k2 = 0; k2 < K; ++k2) Ah[k2 + K * k] += c[h] * qi[k] * qi[k2] / sq;
        }
      }
    }
  }
}

/* ---------------------------------------------------------------------------------------------- */
/* Simplex projection with pseudocount (OpenADMIXTURE.project_q!, called at algorithms_inner.jl:75, */
/* 178, 206; project_p! utils.jl:20-28).  x = argmin ||x - v|| s.t. x >= eps, sum x = 1.            */
/* ---------------------------------------------------------------------------------------------- */

void hko_project_simplex(double* v, int n, double eps) {
  double u[HKO_QP_MAXN];
  const double target = 1.0 - n * eps;
  for (int t = 0; t < n; ++t) u[t] = v[t] - eps;
  /* sort descending (insertion sort, n is small) */
  double s[HKO_QP_MAXN];
  for (int t = 0; t < n; ++t) {
    double x = u[t];
    int r = t - 1;
    while (r >= 0 && s[r] < x) { s[r + 1] = s[r]; --r; }
    s[r + 1] = x;
  }
  double cum = 0.0, tau = 0.0;
  for (int t = 0; t < n; ++t) {
    cum += s[t];
    double tt = (cum - target) / (t + 1);
    if (s[t] - tt > 0.0) tau = tt;
  }
  for (int t = 0; t < n; ++t) {
    double y = u[t] - tau;
    v[t] = (y > 0.0 ? y : 0.0) + eps;
  }
}

void hko_project_q(double* q, int K, int64_t I, double eps) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < I; ++i) hko_project_simplex(q + (int64_t)K * i, K, eps);
}

void hko_project_p(double* p, int K, int64_t J, double eps) {
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < J; ++j) {
    for (int k = 0; k < K; ++k) {
      double v[4];
      for (int h = 0; h < 4; ++h) v[h] = p[k + (int64_t)K * (4 * j + h)];
      hko_project_simplex(v, 4, eps);
      for (int h = 0; h < 4; ++h) p[k + (int64_t)K * (4 * j + h)] = v[h];
    }
  }
}

/* ---------------------------------------------------------------------------------------------- */
/* Dense primal active-set QP                                                                      */
/*   minimise 0.5 x'Hx + g'x  subject to lo <= x <= hi and E'x = 0 (E: n x m, disjoint 0/1 columns) */
/* x = 0 must be feasible.  This is the problem OpenADMIXTURE.quadratic_program! solves for the     */
/* update_q!/update_p! row steps (algorithms_inner.jl:141-168, 264-312); for a positive definite H  */
/* its optimum is unique, so any exact method returns the same point.                              */
/* ---------------------------------------------------------------------------------------------- */

static int chol_inplace(double* A, int n, int lda) {
  for (int k = 0; k < n; ++k) {
    double d = A[k + lda * k];
    for (int t = 0; t < k; ++t) d -= A[k + lda * t] * A[k + lda * t];
    if (!(d > 0.0)) return -1;
    d = sqrt(d);
    A[k + lda * k] = d;
    for (int r = k + 1; r < n; ++r) {
      double s = A[r + lda * k];
      for (int t = 0; t < k; ++t) s -= A[r + lda * t] * A[k + lda * t];
      A[r + lda * k] = s / d;
    }
  }
  return 0;
}

static void chol_solve(const double* L, int n, int lda, double* b) {
  for (int r = 0; r < n; ++r) {
    double s = b[r];
    for (int t = 0; t < r; ++t) s -= L[r + lda * t] * b[t];
    b[r] = s / L[r + lda * r];
  }
  for (int r = n - 1; r >= 0; --r) {
    double s = b[r];
    for (int t = r + 1; t < n; ++t) s -= L[t + lda * r] * b[t];
    b[r] = s / L[r + lda * r];
  }
}

int hko_qp_solve(int n, const double* H, const double* g, const double* lo, const double* hi, int m,
                 const int* grp, double* x_out) {
  /* grp[i] in [0, m) names the equality constraint variable i belongs to (sum over the group = 0),
   * or -1 for none; m = 0 means no equality constraints. */
  if (n <= 0 || n > HKO_QP_MAXN || m < 0 || m > HKO_QP_MAXM) return -1;
  double x[HKO_QP_MAXN], r[HKO_QP_MAXN], d[HKO_QP_MAXN];
  int st[HKO_QP_MAXN]; /* 0 free, -1 at lower, +1 at upper */
  static __thread double M[HKO_QP_MAXN * HKO_QP_MAXN];
  static __thread double Y[HKO_QP_MAXN * HKO_QP_MAXM];
  int fidx[HKO_QP_MAXN];
  double gscale = 0.0, hmax = 0.0, bmax = 0.0;
  for (int i = 0; i < n; ++i) {
    x[i] = 0.0;
    st[i] = (lo[i] >= 0.0) ? -1 : (hi[i] <= 0.0 ? 1 : 0);
    if (fabs(g[i]) > gscale) gscale = fabs(g[i]);
    if (fabs(lo[i]) > bmax) bmax = fabs(lo[i]);
    if (fabs(hi[i]) > bmax) bmax = fabs(hi[i]);
    for (int t = 0; t < n; ++t)
      if (fabs(H[i + n * t]) > hmax) hmax = fabs(H[i + n * t]);
  }
  const double tol = 1e-11 * (gscale + hmax * bmax) + 1e-300;
  const int maxit = 4 * n + 20;
  int it;
  for (it = 0; it < maxit; ++it) {
    int nf = 0;
    for (int i = 0; i < n; ++i)
      if (st[i] == 0) fidx[nf++] = i;
    for (int i = 0; i < n; ++i) {
      double s = g[i];
      for (int t = 0; t < n; ++t) s += H[i + n * t] * x[t];
      r[i] = s;
    }
    /* EQP on the free set: H_FF d + E_F lam = -r_F, E_F' d = 0. */
    for (int a = 0; a < nf; ++a)
      for (int b = 0; b < nf; ++b) M[a + nf * b] = H[fidx[a] + n * fidx[b]];
    if (nf > 0 && chol_inplace(M, nf, nf) != 0) {
      /* Curvature semidefinite up to round-off on the free set: escalating relative ridge. */
      double dmax = 0.0;
      for (int i = 0; i < n; ++i) dmax = fmax(dmax, fabs(H[i + n * i]));
      int ok = 0;
      const double ridges[3] = {1e-12, 1e-9, 1e-6};
      for (int t = 0; t < 3 && !ok; ++t) {
        for (int a = 0; a < nf; ++a)
          for (int b = 0; b < nf; ++b)
            M[a + nf * b] = H[fidx[a] + n * fidx[b]] + (a == b ? ridges[t] * dmax + 1e-300 : 0.0);
        ok = chol_inplace(M, nf, nf) == 0;
      }
      if (!ok) return -2;
    }
    double z[HKO_QP_MAXN];
    for (int a = 0; a < nf; ++a) z[a] = r[fidx[a]];
    if (nf > 0) chol_solve(M, nf, nf, z);
    double lam[HKO_QP_MAXM];
    for (int c = 0; c < m; ++c) lam[c] = 0.0;
    if (m > 0) {
      /* Constraints with a free member; Y = H_FF^{-1} E_F, S = E_F' Y. */
      int cidx[HKO_QP_MAXM], nc = 0, cmap[HKO_QP_MAXM];
      for (int c = 0; c < m; ++c) {
        cmap[c] = -1;
        for (int a = 0; a < nf; ++a)
          if (grp[fidx[a]] == c) { cmap[c] = nc; cidx[nc++] = c; break; }
      }
      if (nc > 0) {
        for (int cc = 0; cc < nc; ++cc) {
          double* y = Y + (int64_t)nf * cc;
          for (int a = 0; a < nf; ++a) y[a] = (grp[fidx[a]] == cidx[cc]) ? 1.0 : 0.0;
          chol_solve(M, nf, nf, y);
        }
        double S[HKO_QP_MAXM * HKO_QP_MAXM], rhs[HKO_QP_MAXM];
        for (int c1 = 0; c1 < nc; ++c1) {
          double s = 0.0;
          for (int a = 0; a < nf; ++a)
            if (grp[fidx[a]] == cidx[c1]) s += z[a];
          rhs[c1] = -s;
          for (int c2 = 0; c2 < nc; ++c2) {
            double t = 0.0;
            for (int a = 0; a < nf; ++a)
              if (grp[fidx[a]] == cidx[c1]) t += Y[a + (int64_t)nf * c2];
            S[c1 + nc * c2] = t;
          }
        }
        if (chol_inplace(S, nc, nc) != 0) return -3;
        chol_solve(S, nc, nc, rhs);
        for (int cc = 0; cc < nc; ++cc) lam[cidx[cc]] = rhs[cc];
        (void)cmap;
        for (int a = 0; a < nf; ++a) {
          double s = z[a];
          for (int cc = 0; cc < nc; ++cc) s += Y[a + (int64_t)nf * cc] * rhs[cc];
          z[a] = s;
        }
      }
    }
    for (int i = 0; i < n; ++i) d[i] = 0.0;
    for (int a = 0; a < nf; ++a) d[fidx[a]] = -z[a];
    /* The lone free member of an equality group cannot move (its group sum is fixed): its step is
     * exactly 0, which keeps round-off from re-blocking it at a bound it was just released from. */
    for (int c = 0; c < m; ++c) {
      int cnt = 0, last = -1;
      for (int a = 0; a < nf; ++a)
        if (grp[fidx[a]] == c) { ++cnt; last = fidx[a]; }
      if (cnt == 1) d[last] = 0.0;
    }
    /* Ratio test. */
    double alpha = 1.0;
    int blk = -1, bside = 0;
    for (int a = 0; a < nf; ++a) {
      int i = fidx[a];
      if (d[i] < 0.0 && x[i] + d[i] < lo[i]) {
        double t = (lo[i] - x[i]) / d[i];
        if (t < alpha) { alpha = t; blk = i; bside = -1; }
      } else if (d[i] > 0.0 && x[i] + d[i] > hi[i]) {
        double t = (hi[i] - x[i]) / d[i];
        if (t < alpha) { alpha = t; blk = i; bside = 1; }
      }
    }
    if (alpha < 0.0) alpha = 0.0;
    for (int a = 0; a < nf; ++a) {
      int i = fidx[a];
      double v = x[i] + alpha * d[i];
      x[i] = v < lo[i] ? lo[i] : (v > hi[i] ? hi[i] : v);
    }
    if (blk >= 0) {
      x[blk] = bside < 0 ? lo[blk] : hi[blk];
      st[blk] = bside;
      continue;
    }
    /* At the minimiser of the current working set: check the bound multipliers. */
    int worst = -1;
    double wv = tol;
    for (int i = 0; i < n; ++i) {
      if (st[i] == 0) continue;
      double s = g[i];
      for (int t = 0; t < n; ++t) s += H[i + n * t] * x[t];
      if (m > 0 && grp[i] >= 0) s += lam[grp[i]];
      double viol = st[i] < 0 ? -s : s;
      if (viol > wv) { wv = viol; worst = i; }
    }
    if (worst < 0) break;
    st[worst] = 0;
  }
  for (int i = 0; i < n; ++i) x_out[i] = x[i];
  return it;
}

/* ---------------------------------------------------------------------------------------------- */
/* Row steps of update_q! / update_p! (algorithms_inner.jl:135-207, 259-351).                       */
/* ---------------------------------------------------------------------------------------------- */

/* update_q! QP stage: XtX, Xtz as returned by the pass (Xtz NOT negated).  q_next = proj(q + dq).
 * penalty != 0 uses H = XtX_i + rho 11' with no equality (the intended reading of :158-162; the
 * reference adds rho to the whole XtX array once per row, which depends on thread scheduling). */
static int q_step_core(int K, int64_t I, const double* XtX, const double* Xtz, const double* q, double* q_next,
                       int penalty, double rho, double eps, double* dq);

int hko_q_step(int K, int64_t I, const double* XtX, const double* Xtz, const double* q, double* q_next,
               int penalty, double rho, double eps) {
  return q_step_core(K, I, XtX, Xtz, q, q_next, penalty, rho, eps, NULL);
}

int hko_q_step_ex(int K, int64_t I, const double* XtX, const double* Xtz, const double* q, double* q_next,
                  int penalty, double rho, double eps, double* dq) {
  return q_step_core(K, I, XtX, Xtz, q, q_next, penalty, rho, eps, dq);
}

/* dq (may be NULL) receives the raw step qdiff (d.q_tmp, algorithms_inner.jl:101). */
static int q_step_core(int K, int64_t I, const double* XtX, const double* Xtz, const double* q, double* q_next,
                       int penalty, double rho, double eps, double* dq) {
  int fails = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : fails)
  for (int64_t i = 0; i < I; ++i) {
    double H[HKO_MAXK * HKO_MAXK], g[HKO_MAXK], lo[HKO_MAXK], hi[HKO_MAXK], x[HKO_MAXK];
    int grp[HKO_MAXK];
    const double* A = XtX + (int64_t)K * K * i;
    /* With the sum(dq) = 0 equality, sigma 11' vanishes on the feasible set: adding it leaves the
     * optimum unchanged and makes a semidefinite curvature definite on every free set. */
    double sigma = 0.0;
    for (int a = 0; a < K; ++a) sigma = fmax(sigma, A[a + K * a]);
    for (int a = 0; a < K; ++a) {
      for (int b = 0; b < K; ++b) H[a + K * b] = A[a + K * b] + (penalty ? rho : sigma);
      g[a] = -Xtz[a + (int64_t)K * i];
      lo[a] = 0.0 - q[a + (int64_t)K * i];
      hi[a] = 1.0 - q[a + (int64_t)K * i];
      grp[a] = 0;
    }
    if (hko_qp_solve(K, H, g, lo, hi, penalty ? 0 : 1, grp, x) < 0) {
      ++fails;
      for (int a = 0; a < K; ++a) x[a] = 0.0;
    }
    for (int a = 0; a < K; ++a) q_next[a + (int64_t)K * i] = q[a + (int64_t)K * i] + x[a];
    if (dq)
      for (int a = 0; a < K; ++a) dq[a + (int64_t)K * i] = x[a];
  }
  hko_project_q(q_next, K, I, eps);
  return fails;
}

/* update_p! QP stage: 4K x 4K block-diagonal curvature (:272-283), default penalty rho on every
 * (l, l2) block diagonal (:295-302), box [0, 1] on p + dp; without penalty, K equalities
 * sum_l dp[l, k] = 0.  Variables are ordered (l-1)K + k as in pv[j] (structs.jl:218). */
static int p_step_core(int K, int64_t J, const double* XtX, const double* Xtz, const double* p, double* p_next,
                       int penalty, double rho, double eps, double* dp);

int hko_p_step(int K, int64_t J, const double* XtX, const double* Xtz, const double* p, double* p_next,
               int penalty, double rho, double eps) {
  return p_step_core(K, J, XtX, Xtz, p, p_next, penalty, rho, eps, NULL);
}

int hko_p_step_ex(int K, int64_t J, const double* XtX, const double* Xtz, const double* p, double* p_next,
                  int penalty, double rho, double eps, double* dp) {
  return p_step_core(K, J, XtX, Xtz, p, p_next, penalty, rho, eps, dp);
}

/* dp (may be NULL) receives the raw step pdiff (d.p_tmp, algorithms_inner.jl:228). */
static int p_step_core(int K, int64_t J, const double* XtX, const double* Xtz, const double* p, double* p_next,
                       int penalty, double rho, double eps, double* dp) {
  int fails = 0;
  const int n = 4 * K;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : fails)
  for (int64_t j = 0; j < J; ++j) {
    double H[4 * HKO_MAXK * 4 * HKO_MAXK], g[4 * HKO_MAXK], lo[4 * HKO_MAXK], hi[4 * HKO_MAXK],
        x[4 * HKO_MAXK];
    int grp[4 * HKO_MAXK];
    for (int t = 0; t < n * n; ++t) H[t] = 0.0;
    for (int l = 0; l < 4; ++l) {
      const double* A = XtX + (int64_t)K * K * (4 * j + l);
      for (int k = 0; k < K; ++k)
        for (int k2 = 0; k2 < K; ++k2) H[(l * K + k2) + n * (l * K + k)] = A[k2 + K * k];
    }
    /* penalty: rho on every (l, l2) block diagonal.  Without it, the K equalities sum_l dp[l,k] = 0
     * make sigma (11' x I_K) vanish on the feasible set; it is added for definiteness as in q_step. */
    double sigma = 0.0;
    for (int t = 0; t < n; ++t) sigma = fmax(sigma, H[t + n * t]);
    const double add = penalty ? rho : sigma;
    for (int l2 = 0; l2 < 4; ++l2)
      for (int l = 0; l < 4; ++l)
        for (int k = 0; k < K; ++k) H[(l * K + k) + n * (l2 * K + k)] += add;
    const double* pj = p + (int64_t)K * 4 * j;
    for (int t = 0; t < n; ++t) {
      g[t] = -Xtz[(int64_t)K * 4 * j + t];
      lo[t] = 0.0 - pj[t];
      hi[t] = 1.0 - pj[t];
      grp[t] = t % K;
    }
    if (hko_qp_solve(n, H, g, lo, hi, penalty ? 0 : K, grp, x) < 0) {
      ++fails;
      for (int t = 0; t < n; ++t) x[t] = 0.0;
    }
    for (int t = 0; t < n; ++t) p_next[(int64_t)K * 4 * j + t] = pj[t] + x[t];
    if (dp)
      for (int t = 0; t < n; ++t) dp[(int64_t)K * 4 * j + t] = x[t];
  }
  hko_project_p(p_next, K, J, eps);
  return fails;
}

/* em! normalisation and projections (algorithms_inner.jl:66-76). */
void hko_em_normalize(int K, int64_t I, int64_t J, const int64_t* dm, double* q_next, double* p_next,
                      int fix_q, int fix_p, double eps) {
  if (!fix_q) {
    for (int64_t i = 0; i < I; ++i)
      for (int k = 0; k < K; ++k) q_next[k + (int64_t)K * i] /= (double)(2 * (J - dm[i]));
  }
  for (int64_t j = 0; j < J; ++j)
    for (int k = 0; k < K; ++k) {
      double* v = p_next + (int64_t)K * 4 * j + k;
      double s = v[0] + v[K] + v[2 * K] + v[3 * K];
      for (int l = 0; l < 4; ++l) v[l * K] /= s;
    }
  hko_project_q(q_next, K, I, eps);
  hko_project_p(p_next, K, J, eps);
  (void)fix_p;
}

/* ---------------------------------------------------------------------------------------------- */
/* Zhou-Alexander-Lange QN (OpenADMIXTURE.update_UV!/update_QN!, algorithms_outer.jl:128-133)        */
/* ---------------------------------------------------------------------------------------------- */

/* Solve the small dense system A y = b (LU with partial pivoting).  Returns -1 if singular. */
static int lu_solve(double* A, int n, double* b) {
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (fabs(A[r + n * c]) > fabs(A[piv + n * c])) piv = r;
    if (A[piv + n * c] == 0.0) return -1;
    if (piv != c) {
      for (int t = 0; t < n; ++t) { double tmp = A[c + n * t]; A[c + n * t] = A[piv + n * t]; A[piv + n * t] = tmp; }
      double tmp = b[c]; b[c] = b[piv]; b[piv] = tmp;
    }
    for (int r = c + 1; r < n; ++r) {
      double f = A[r + n * c] / A[c + n * c];
      for (int t = c; t < n; ++t) A[r + n * t] -= f * A[c + n * t];
      b[r] -= f * b[c];
    }
  }
  for (int r = n - 1; r >= 0; --r) {
    double s = b[r];
    for (int t = r + 1; t < n; ++t) s -= A[r + n * t] * b[t];
    b[r] = s / A[r + n * r];
  }
  return 0;
}

/* U[:, col] = x1 - x0, V[:, col] = x2 - x1; x_out = x1 - V (U'U - U'V)^{-1} U'(x0 - x1) over the
 * first ncols columns. */
/* Quasi-Newton point of admixture_qn! from x0 = x, x1 = F(x), x2 = F(F(x)) (algorithms_outer.jl:128-141).
 * Secant columns (update_UV!'s cycling): U[:, col] = x1 - x0, V[:, col] = x2 - x1, col = (iter-1) mod Qn,
 * using the first min(iter, Qn) columns.  With r = x0 - x1:
 *   mode 0, ZAL (Zhou, Alexander & Lange 2011):  x_out = x1 - V c,  (U'U - U'V) c = U' r
 *   mode 1, LBQN (limited-memory Broyden, Agarwal & Xu):  x_out = x1 - V c,  (Y'Y) c = -Y' r,  Y = V - U
 * Both are multisecant Broyden updates from H0 = -I of the inverse Jacobian of G(x) = F(x) - x (ZAL of the
 * first kind; LBQN of the second kind, H = -I + (U + Y)(Y'Y)^-1 Y', H Y = U) with x_out = x0 - H G(x0).
 * OpenADMIXTURE's update_UV_LBQN!/update_QN_LBQN! are not available; mode 1 is the published method. */
int hko_qn(int64_t N, int Qn, int iter, int mode, double* U, double* V, const double* x0, const double* x1,
           const double* x2, double* x_out);

int hko_zal(int64_t N, int Qn, int iter, double* U, double* V, const double* x0, const double* x1,
            const double* x2, double* x_out) {
  return hko_qn(N, Qn, iter, 0, U, V, x0, x1, x2, x_out);
}

int hko_qn(int64_t N, int Qn, int iter, int mode, double* U, double* V, const double* x0, const double* x1,
           const double* x2, double* x_out) {
  const int col = (iter - 1) % Qn;
  const int ncols = iter < Qn ? iter : Qn;
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < N; ++t) {
    U[t + N * col] = x1[t] - x0[t];
    V[t + N * col] = x2[t] - x1[t];
  }
  double UtU[64], UtV[64], VtV[64], Utr[8], Vtr[8];
  for (int a = 0; a < ncols; ++a) {
    const double* ua = U + N * a;
    const double* va = V + N * a;
    double s = 0.0, sv = 0.0;
    for (int64_t t = 0; t < N; ++t) { s += ua[t] * (x0[t] - x1[t]); sv += va[t] * (x0[t] - x1[t]); }
    Utr[a] = s;
    Vtr[a] = sv;
    for (int b = 0; b < ncols; ++b) {
      const double* ub = U + N * b;
      const double* vb = V + N * b;
      double su = 0.0, suv = 0.0, svv = 0.0;
      for (int64_t t = 0; t < N; ++t) { su += ua[t] * ub[t]; suv += ua[t] * vb[t]; svv += va[t] * vb[t]; }
      UtU[a + ncols * b] = su;
      UtV[a + ncols * b] = suv;
      VtV[a + ncols * b] = svv;
    }
  }
  double A[64];
  for (int a = 0; a < ncols; ++a)
    for (int b = 0; b < ncols; ++b) {
      const int t = a + ncols * b;
      A[t] = mode == 0 ? UtU[t] - UtV[t] : VtV[t] - UtV[t] - UtV[b + ncols * a] + UtU[t];
    }
  if (mode != 0)
    for (int a = 0; a < ncols; ++a) Utr[a] -= Vtr[a];
  int rc = lu_solve(A, ncols, Utr);
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < N; ++t) {
    double s = x1[t];
    if (rc == 0)
      for (int a = 0; a < ncols; ++a) s -= V[t + N * a] * Utr[a];
    x_out[t] = s;
  }
  return rc;
}

/* ---------------------------------------------------------------------------------------------- */
/* Whole fit: init_em! + admixture_qn! (algorithms_outer.jl), ZAL mode, reference defaults          */
/* (p_penalty = true, q_penalty = false when penalty is nothing, :58-64).                          */
/* ---------------------------------------------------------------------------------------------- */

typedef struct {
  int em_iters;      /* admix_n_em_iters (driver.jl:47), 5 */
  int n_iter;        /* admix_n_iter, 1000 */
  double rtol;       /* admix_rtol, 1e-7 */
  int Qn;            /* QN secant count Q, 3 */
  int fix_q, fix_p;  /* driver.jl:54-55 */
  int q_penalty, p_penalty;
  double rho;        /* 1e7 */
  double eps;        /* pseudocount 1e-5 */
  int guarantee_increase; /* admixture_qn!(...; guarantee_increase), algorithms_outer.jl:41 */
  int qn_mode;            /* admixture_qn!(...; mode): 0 = :ZAL, 1 = :LBQN (algorithms_outer.jl:128-141) */
} hko_fit_opts;

/* Step-halving loop of update_q! / update_p! with guarantee_increase (algorithms_inner.jl:170-199 and
 * 314-343), after the QP stage left out = proj(base + delta) for factor 1:
 *   factor = 1; loop { out = proj(base + factor delta); ll = loglik(q_ll, p_ll);
 *                      if ll >= ll_tmp (Q) / ll > ll_tmp (P) { ll_tmp = ll; break }
 *                      factor /= 2; if factor < 1e-3 { factor = 0 }; if factor == 0 break }
 * so when no factor qualifies, out keeps the last one tried (2^-9) and ll_tmp is unchanged. */
static void halve_step(const uint8_t* bed, int64_t bps, int64_t I, int64_t J, int K, int is_q, const double* base,
                       const double* delta, double* out, const double* q_ll, const double* p_ll, double* ll_tmp,
                       double eps) {
  const int64_t n = is_q ? (int64_t)K * I : (int64_t)K * 4 * J;
  double factor = 1.0;
  for (;;) {
    if (factor < 1.0) {
      for (int64_t t = 0; t < n; ++t) out[t] = base[t] + factor * delta[t];
      if (is_q) hko_project_q(out, K, I, eps);
      else hko_project_p(out, K, J, eps);
    }
    const double ll = hko_loglik(bed, bps, I, J, K, q_ll, p_ll);
    if (is_q ? ll >= *ll_tmp : ll > *ll_tmp) {
      *ll_tmp = ll;
      return;
    }
    factor /= 2;
    if (factor < 1e-3) factor = 0.0;
    if (factor == 0.0) return;
  }
}

/* x is K x (I + 4J): q = x[:, 0:I], p = x[:, I:]; q and p are updated in place.
 * ll_trace (may be NULL) receives ll_new after each QN iteration (length n_iter).
 * Returns the number of QN iterations run; *ll_out = final ll_new. */
int hko_fit(const uint8_t* bed, int64_t bps, int64_t I, int64_t J, int K, const hko_fit_opts* o, double* x,
            double* ll_out, double* ll_trace) {
  const int64_t N = (int64_t)K * (I + 4 * J);
  double* q = x;
  double* p = x + (int64_t)K * I;
  double* x_next = (double*)malloc(sizeof(double) * N);
  double* x_next2 = (double*)malloc(sizeof(double) * N);
  double* x_tmp = (double*)malloc(sizeof(double) * N);
  double* U = (double*)calloc((size_t)N * o->Qn, sizeof(double));
  double* V = (double*)calloc((size_t)N * o->Qn, sizeof(double));
  double* XtXq = (double*)malloc(sizeof(double) * K * K * I);
  double* Xtzq = (double*)malloc(sizeof(double) * K * I);
  double* XtXp = (double*)malloc(sizeof(double) * K * K * 4 * J);
  double* Xtzp = (double*)malloc(sizeof(double) * K * 4 * J);
  int64_t* dm = (int64_t*)malloc(sizeof(int64_t) * I);
  if (!x_next || !x_next2 || !x_tmp || !U || !V || !XtXq || !Xtzq || !XtXp || !Xtzp || !dm) return -1;
  double* q_next = x_next;
  double* p_next = x_next + (int64_t)K * I;
  double* q_next2 = x_next2;
  double* p_next2 = x_next2 + (int64_t)K * I;
  double* q_tmp = x_tmp;
  double* p_tmp = x_tmp + (int64_t)K * I;
  hko_count_double_missing(bed, bps, I, J, dm);
  double ll_new = NAN;
  /* init_em! (algorithms_outer.jl:1-22) */
  /* driver.jl:182 calls init_em! without fix_q/fix_p, so the warm-up always updates both. */
  for (int it = 0; it < o->em_iters; ++it) {
    hko_em(bed, bps, I, J, K, q, p, q_next, p_next, 0, 0);
    hko_em_normalize(K, I, J, dm, q_next, p_next, 0, 0, o->eps);
    memcpy(p, p_next, sizeof(double) * K * 4 * J);
    memcpy(q, q_next, sizeof(double) * K * I);
    ll_new = hko_loglik(bed, bps, I, J, K, q, p);
    hko_project_p(p, K, J, o->eps);
    hko_project_q(q, K, I, o->eps);
  }
  if (isnan(ll_new)) ll_new = hko_loglik(bed, bps, I, J, K, q, p);
  int iters = 0;
  const int gi = o->guarantee_increase;
  double* dq = gi ? q_tmp : NULL; /* qdiff / pdiff live in x_tmp until ZAL overwrites it */
  double* dp = gi ? p_tmp : NULL;
  for (int i = 1; i <= o->n_iter; ++i) {
    iters = i;
    const double ll_prev = ll_new;
    double ll_tmp = ll_new;
    /* first sweep: F(x) */
    if (!o->fix_q) {
      hko_update_q(bed, bps, I, J, K, q, p, XtXq, Xtzq);
      q_step_core(K, I, XtXq, Xtzq, q, q_next, o->q_penalty, o->rho, o->eps, dq);
      if (gi) halve_step(bed, bps, I, J, K, 1, q, dq, q_next, q_next, p, &ll_tmp, o->eps);
    } else memcpy(q_next, q, sizeof(double) * K * I);
    if (!o->fix_p) {
      hko_update_p(bed, bps, I, J, K, q_next, p, XtXp, Xtzp);
      p_step_core(K, J, XtXp, Xtzp, p, p_next, o->p_penalty, o->rho, o->eps, dp);
      if (gi) halve_step(bed, bps, I, J, K, 0, p, dp, p_next, q_next, p_next, &ll_tmp, o->eps);
    } else memcpy(p_next, p, sizeof(double) * K * 4 * J);
    /* second sweep: F(F(x)) */
    if (!o->fix_q) {
      hko_update_q(bed, bps, I, J, K, q_next, p_next, XtXq, Xtzq);
      q_step_core(K, I, XtXq, Xtzq, q_next, q_next2, o->q_penalty, o->rho, o->eps, dq);
      if (gi) halve_step(bed, bps, I, J, K, 1, q_next, dq, q_next2, q_next2, p_next, &ll_tmp, o->eps);
    } else memcpy(q_next2, q, sizeof(double) * K * I);
    if (!o->fix_p) {
      hko_update_p(bed, bps, I, J, K, q_next2, p_next, XtXp, Xtzp);
      p_step_core(K, J, XtXp, Xtzp, p_next, p_next2, o->p_penalty, o->rho, o->eps, dp);
      if (gi) halve_step(bed, bps, I, J, K, 0, p_next, dp, p_next2, q_next2, p_next2, &ll_tmp, o->eps);
    } else memcpy(p_next2, p, sizeof(double) * K * 4 * J);
    const double ll_basic = hko_loglik(bed, bps, I, J, K, q_next2, p_next2);
    hko_qn(N, o->Qn, i, o->qn_mode, U, V, x, x_next, x_next2, x_tmp);
    hko_project_p(p_tmp, K, J, o->eps);
    hko_project_q(q_tmp, K, I, o->eps);
    const double ll_qn = hko_loglik(bed, bps, I, J, K, q_tmp, p_tmp);
    if (ll_basic < ll_qn) {
      memcpy(x, x_tmp, sizeof(double) * N);
      ll_new = ll_qn;
    } else {
      memcpy(x, x_next2, sizeof(double) * N);
      ll_new = ll_basic;
    }
    if (ll_trace) ll_trace[i - 1] = ll_new;
    const double reldiff = fabs((ll_new - ll_prev) / ll_prev);
    if (reldiff < o->rtol) break;
  }
  *ll_out = ll_new;
  free(x_next); free(x_next2); free(x_tmp); free(U); free(V);
  free(XtXq); free(Xtzq); free(XtXp); free(Xtzp); free(dm);
  return iters;
}

int hko_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
