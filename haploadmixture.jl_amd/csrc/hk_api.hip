// hk_api.hip — C ABI of libhaplo_hip.so: context, parameter transfer, the reference's GPU runner entry
// points and the device-resident QN/SQP fit loop.  Host code; the kernels live in hk_pass.hip and
// hk_step.hip.
#include <limits.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <utility>
#include <vector>

#include "../../include/haplo_hip.h"
#include "hk_internal.h"

using namespace hk;

static char g_create_err[512];

static int fail(hk_ctx* c, int code, const char* fmt, ...) {
  char* dst = c ? c->err : g_create_err;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(dst, 512, fmt, ap);
  va_end(ap);
  return code;
}

#define HIPC(ctx, call)                                                                          \
  do {                                                                                           \
    hipError_t e_ = (call);                                                                      \
    if (e_ != hipSuccess)                                                                        \
      return fail((ctx), e_ == hipErrorOutOfMemory ? HK_ENOMEM : HK_EHIP, "%s failed: %s (%s:%d)", \
                  #call, hipGetErrorString(e_), __FILE__, __LINE__);                             \
  } while (0)

#define CHK(x)                \
  do {                        \
    int rc_ = (x);            \
    if (rc_ != HK_OK) return rc_; \
  } while (0)

// The P row step's device lists; the early list is filled by the P-side curvature pass (run_ppass).
static PStepLists p_lists(hk_ctx* c) { return PStepLists{c->redo, c->pflag, c->redo_late}; }

static int64_t round_up(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

static int ntq(int K) { return K + K * (K + 1) / 2; }

// Chunking of the passes so that one launch has >= ~2048 workgroups (8 per CU on 256 CUs).
// Default chunking: enough workgroups for ~32 per CU on 256 CUs (the passes keep 2-4 resident per CU, so
// the last round is a small fraction), but at least 16 genotype words per chunk so that the per-chunk
// partial sums stay small next to the genotype stream.
static void plan_chunks(hk_ctx* c) {
  PassGeom& g = c->g;
  constexpr int64_t kTargetWG = 8192, kMinWords = 16;
  const int64_t wgq = g.Ip / 256, wgp = 4 * g.Jp / 256;
  // FP32 accumulation (the reference's Float32 sums) runs each sum over the whole i / j range
  const int64_t auto_q = g.acc32 ? 1 : (kTargetWG + wgq - 1) / wgq, auto_p = g.acc32 ? 1 : (kTargetWG + wgp - 1) / wgp;
  int64_t qc = c->qchunks_req > 0 ? c->qchunks_req : auto_q;
  int64_t pc = c->pchunks_req > 0 ? c->pchunks_req : auto_p;
  if (c->qchunks_req <= 0 && qc > g.J8 / kMinWords) qc = g.J8 / kMinWords;
  if (c->pchunks_req <= 0 && pc > g.I8 / kMinWords) pc = g.I8 / kMinWords;
  if (qc < 1) qc = 1;
  if (qc > g.J8) qc = g.J8;
  if (pc < 1) pc = 1;
  if (pc > g.I8) pc = g.I8;
  g.qwpc = (int)((g.J8 + qc - 1) / qc);
  g.qchunks = (int)((g.J8 + g.qwpc - 1) / g.qwpc);
  g.pwpc = (int)((g.I8 + pc - 1) / pc);
  g.pchunks = (int)((g.I8 + g.pwpc - 1) / g.pwpc);
}

static int alloc_parts(hk_ctx* c) {
  // the curvature slots; at least 3 rows so that the two-set log-likelihood sweep's partials (2 x qchunks x Ip)
  // and its workgroup sums fit
  const int NT = ntq(c->K) > 3 ? ntq(c->K) : 3;
  const int64_t lq = (int64_t)c->g.qchunks * NT * c->g.Ip;
  const int64_t lp = c->g.pchunks > 1 ? (int64_t)c->g.pchunks * 4 * c->g.Jp * NT : 0;
  if (lq > c->part_q_len) {
    if (c->part_q) HIPC(c, hipFree(c->part_q));
    c->part_q = nullptr;
    HIPC(c, hipMalloc(&c->part_q, sizeof(double) * lq));
    c->part_q_len = lq;
  }
  if (lp > c->part_p_len) {
    if (c->part_p) HIPC(c, hipFree(c->part_p));
    c->part_p = nullptr;
    HIPC(c, hipMalloc(&c->part_p, sizeof(double) * lp));
    c->part_p_len = lp;
  }
  return HK_OK;
}

static double* qof(hk_ctx* c, double* x) { return x; }
static double* pof(hk_ctx* c, double* x) { return x + (int64_t)c->K * c->I; }

// Pointer handed to a pass for parameters (q, p): FP64 directly, or the FP32 mirror.
static int pass_params(hk_ctx* c, const double* q, const double* p, const void** qq, const void** pp,
                       const int* gate = nullptr, float* mirror = nullptr) {
  if (c->prec == PREC_FP64) {
    *qq = q;
    *pp = p;
    return HK_OK;
  }
  float* m = mirror ? mirror : c->x32;
  const int64_t nq = (int64_t)c->K * c->I, np = (int64_t)c->K * 4 * c->g.Jp;
  HIPC(c, launch_pair_to_f32(q, nq, p, np, m, c->s, gate));
  *qq = m;
  *pp = m + nq;
  return HK_OK;
}

namespace {

struct Timer {
  bool on = false;
  hipStream_t s = nullptr;
  std::vector<hipEvent_t> pool;
  std::vector<std::pair<int, int>> marks;  // (category, event index of start)
  size_t used = 0;
  ~Timer() {
    for (auto e : pool) (void)hipEventDestroy(e);
  }
  hipEvent_t get() {
    if (used == pool.size()) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      pool.push_back(e);
    }
    return pool[used++];
  }
  int begin(int cat) {
    if (!on) return -1;
    hipEvent_t e = get();
    if (!e) return -1;
    (void)hipEventRecord(e, s);
    marks.push_back({cat, (int)(used - 1)});
    return 0;
  }
  void end() {
    if (!on || marks.empty()) return;
    hipEvent_t e = get();
    if (e) (void)hipEventRecord(e, s);
  }
  void collect(double* t, int* n) {
    for (size_t m = 0; m < marks.size(); ++m) {
      const int cat = marks[m].first, i0 = marks[m].second;
      if ((size_t)i0 + 1 >= used) continue;
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, pool[i0], pool[i0 + 1]) == hipSuccess) {
        t[cat] += ms;
        n[cat] += 1;
      }
    }
  }
};

enum { T_QPASS, T_PPASS, T_LL, T_QSTEP, T_PSTEP, T_EM, T_QN, T_NCAT };

}  // namespace

// Q-side pass (individual-parallel): MODE_LL -> ll_dev / ll_rows, else res_q.  gate: the fit's device gate
// when the pass is part of a batch graph (every kernel then returns at once after the fit stopped).
static int run_qpass(hk_ctx* c, int mode, const double* q, const double* p, Timer* tm = nullptr, int cat = 0,
                     double* ll_out = nullptr, const int* gate = nullptr) {
  PassGeom g = c->g;
  g.gate = gate;
  const void *qq, *pp;
  CHK(pass_params(c, q, p, &qq, &pp, gate));
  if (mode == MODE_CURV) HIPC(c, launch_curv_tables(c->K, c->prec, 0, g, qq, pp, c->s));
  if (tm) tm->begin(cat);
  if (mode == MODE_LL && c->prec != PREC_FP64)
    HIPC(c, launch_ll_mm(c->K, g, (const float*)qq, (const float*)pp, nullptr, nullptr, c->part_q, c->ll_stage,
                         c->s));
  else
    HIPC(c, launch_qside(c->K, c->prec, mode, g, qq, pp, c->part_q, c->s));
  if (tm) tm->end();
  if (mode == MODE_LL)
    HIPC(c, launch_reduce_ll(c->part_q, g.qchunks, g, c->ll_rows, ll_out ? ll_out : c->ll_dev, c->s));
  else if (mode == MODE_CURV) HIPC(c, launch_reduce_q_curv(c->part_q, g.qchunks, c->K, g, q, c->res_q, c->s));
  else HIPC(c, launch_reduce_q(c->part_q, g.qchunks, nacc(c->K, mode), g, c->res_q, c->s));
  return HK_OK;
}

// P-side pass (block-parallel) -> res_p [4Jp][NT].
static int run_ppass(hk_ctx* c, int mode, const double* q, const double* p, Timer* tm = nullptr, int cat = 0,
                     const int* gate = nullptr) {
  PassGeom g = c->g;
  g.gate = gate;
  const void *qq, *pp;
  CHK(pass_params(c, q, p, &qq, &pp, gate));
  const int NT = nacc(c->K, mode);
  if (mode == MODE_CURV) HIPC(c, launch_curv_tables(c->K, c->prec, 1, g, qq, pp, c->s));
  if (g.pchunks == 1) {
    if (tm) tm->begin(cat);
    HIPC(c, launch_pside(c->K, c->prec, mode, g, qq, pp, c->res_p, c->s));
    if (tm) tm->end();
  } else {
    if (tm) tm->begin(cat);
    HIPC(c, launch_pside(c->K, c->prec, mode, g, qq, pp, c->part_p, c->s));
    if (tm) tm->end();
    HIPC(c, launch_reduce_p(c->part_p, g.pchunks, NT, g, c->res_p, c->s));
  }
  if (mode == MODE_CURV) HIPC(c, launch_finish_p(c->K, g, p, c->res_p, c->pflag, c->redo, c->s));
  return HK_OK;
}

// Log-likelihoods of two parameter sets -> ll1 and ll2 (device).  MIXED: one genotype sweep of the matrix-core
// pass for both; FP64: two passes.
static int run_ll_pair(hk_ctx* c, const double* q1, const double* p1, const double* q2, const double* p2, double* ll1,
                       double* ll2, Timer* tm = nullptr, int cat = 0, const int* gate = nullptr) {
  if (c->prec == PREC_FP64) {
    CHK(run_qpass(c, MODE_LL, q1, p1, tm, cat, ll1, gate));
    CHK(run_qpass(c, MODE_LL, q2, p2, tm, cat, ll2, gate));
    return HK_OK;
  }
  PassGeom g = c->g;
  g.gate = gate;
  const void *qa, *pa, *qb, *pb;
  CHK(pass_params(c, q1, p1, &qa, &pa, gate, c->x32));
  CHK(pass_params(c, q2, p2, &qb, &pb, gate, c->x32b));
  if (tm) tm->begin(cat);
  HIPC(c, launch_ll_mm(c->K, g, (const float*)qa, (const float*)pa, (const float*)qb, (const float*)pb, c->part_q,
                       c->ll_stage, c->s));
  if (tm) tm->end();
  // set 2 first: set 1's workgroup sums go where set 2's row partials were
  HIPC(c, launch_reduce_ll(c->part_q + (int64_t)g.qchunks * g.Ip, g.qchunks, g, c->ll_rows, ll2, c->s));
  HIPC(c, launch_reduce_ll(c->part_q, g.qchunks, g, c->ll_rows, ll1, c->s));
  return HK_OK;
}

static int ensure_full(hk_ctx* c, int64_t n) {
  if (n <= c->full_len) return HK_OK;
  if (c->full_a) HIPC(c, hipFree(c->full_a));
  if (c->full_b) HIPC(c, hipFree(c->full_b));
  c->full_a = c->full_b = nullptr;
  HIPC(c, hipMalloc(&c->full_a, sizeof(double) * n));
  HIPC(c, hipMalloc(&c->full_b, sizeof(double) * n));
  c->full_len = n;
  return HK_OK;
}

extern "C" {

const char* hk_create_error(void) { return g_create_err; }

const char* hk_last_error(const hk_ctx* c) { return c ? c->err : g_create_err; }

int hk_ctx_destroy(hk_ctx* c) {
  if (!c) return HK_OK;
  if (c->s) (void)hipStreamSynchronize(c->s);
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  void* owned[] = {c->qq, c->pp, c->xb[0], c->xb[1], c->xb[2], c->xb[3], c->x32, c->x32b, c->ll_stage,
                   c->res_q, c->res_p, c->part_q, c->part_p, c->ll_dev, c->ll_rows, c->U, c->V, c->zal_part, c->zal_coef, c->fails,
                   c->redo, c->redo_late, c->pflag, c->full_a, c->full_b, c->d_ll, c->d_iter,
                   c->named[0], c->named[1], c->named[2], c->named[3]};
  for (void* p : owned)
    if (p) (void)hipFree(p);
  if (c->owns_geno) {
    void* geno[] = {c->gq, c->gp, c->dm};
    for (void* p : geno)
      if (p) (void)hipFree(p);
  }
  if (c->h_ll) (void)hipHostFree(c->h_ll);
  if (c->s) (void)hipStreamDestroy(c->s);
  delete c;
  return HK_OK;
}

// Per-fit device buffers (everything except the genotype streams and dm).
static int init_fit_buffers(hk_ctx* c) {
  const int K = c->K;
  const int64_t I = c->I;
  PassGeom& g = c->g;
  {
    int64_t nq = 0, np = 0;
    stage_sizes(K, g.Ip, g.Jp, &nq, &np);
    const size_t el = c->prec == HK_PREC_FP64 ? sizeof(double) : sizeof(float);
    HIPC(c, hipMalloc(&c->qq, el * nq));
    HIPC(c, hipMalloc(&c->pp, el * np));
  }
  g.qq = c->qq;
  g.pp = c->pp;
  c->xlen = (int64_t)K * (I + 4 * g.Jp);
  for (int b = 0; b < 4; ++b) {
    HIPC(c, hipMalloc(&c->xb[b], sizeof(double) * c->xlen));
    HIPC(c, launch_fill(c->xb[b], (int64_t)K * I, 1.0 / K, c->s));
    HIPC(c, launch_fill(c->xb[b] + (int64_t)K * I, (int64_t)K * 4 * g.Jp, 0.25, c->s));
  }
  if (c->prec == HK_PREC_MIXED) {
    HIPC(c, hipMalloc(&c->x32, sizeof(float) * c->xlen));
    HIPC(c, hipMalloc(&c->x32b, sizeof(float) * c->xlen));
    HIPC(c, hipMalloc(&c->ll_stage, 2 * ll_stage_bytes(K, g.J8)));
  }
  const int NT = ntq(K);
  HIPC(c, hipMalloc(&c->res_q, sizeof(double) * I * NT));
  HIPC(c, hipMalloc(&c->res_p, sizeof(double) * 4 * g.Jp * NT));
  HIPC(c, hipMalloc(&c->ll_dev, sizeof(double) * 4));
  HIPC(c, hipMalloc(&c->ll_rows, sizeof(double) * I));
  HIPC(c, hipMalloc(&c->fails, 2 * sizeof(int)));  // [QP failures, QPs stopped at the iteration limit]
  HIPC(c, hipMalloc(&c->redo, sizeof(int) * (c->J + 1)));
  HIPC(c, hipMalloc(&c->redo_late, sizeof(int) * (c->J + 1)));
  HIPC(c, hipMalloc(&c->pflag, c->J));
  HIPC(c, hipMalloc(&c->d_ll, sizeof(double) * 8));
  HIPC(c, hipMalloc(&c->d_iter, sizeof(int)));
  HIPC(c, hipHostMalloc((void**)&c->h_ll, sizeof(double) * 8, hipHostMallocDefault));
  HIPC(c, hipMalloc(&c->zal_part, sizeof(double) * 512 * 64));
  HIPC(c, hipMalloc(&c->zal_coef, sizeof(double) * 8));
  plan_chunks(c);
  CHK(alloc_parts(c));
  HIPC(c, prime_p_step_caps(K));  // runtime occupancy queries here, never inside a captured iteration
  HIPC(c, hipStreamSynchronize(c->s));
  return HK_OK;
}

static bool valid_precision(int p) { return p == HK_PREC_FP64 || p == HK_PREC_MIXED || p == HK_PREC_FP32; }

// Kernel precision (FP64 or FP32 per-pair arithmetic) and accumulation semantics of an ABI precision.
static void set_precision(hk_ctx* c, int precision) {
  c->prec = precision == HK_PREC_FP64 ? PREC_FP64 : PREC_MIXED;
  const bool acc32 = precision == HK_PREC_FP32;
  c->g.acc32 = acc32 ? 1 : 0;
  c->g.flush_curv = acc32 ? INT_MAX : 16;
  c->g.flush_grad = acc32 ? INT_MAX : 8;
}

int hk_ctx_create(hk_ctx** out, int64_t I, int64_t J, int K, const uint8_t* bed, int64_t bps, int64_t n_snp,
                  int precision) {
  if (!out) return fail(nullptr, HK_EINVAL, "out is NULL");
  *out = nullptr;
  if (I <= 0 || J <= 0 || K < 1 || K > HK_MAX_K || !bed || bps < (I + 3) / 4 || n_snp < 2 * J ||
      !valid_precision(precision))
    return fail(nullptr, HK_EINVAL,
                "invalid arguments: I=%lld J=%lld K=%d bytes_per_snp=%lld n_snp=%lld precision=%d (need K in 1..%d, "
                "bytes_per_snp >= ceil(I/4), n_snp >= 2J)",
                (long long)I, (long long)J, K, (long long)bps, (long long)n_snp, precision, HK_MAX_K);
  if (I > (int64_t)1 << 31 || J > (int64_t)1 << 29)
    return fail(nullptr, HK_EINVAL, "I or J too large (I <= 2^31, J <= 2^29)");
  hk_ctx* c = new hk_ctx();
  c->I = I;
  c->J = J;
  c->K = K;
  set_precision(c, precision);
  int rc = HK_OK;
  hipError_t e = hipGetDevice(&c->dev);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking);
  PassGeom& g = c->g;
  g.I = I;
  g.J = J;
  g.Ip = round_up(I, 256);
  g.Jp = round_up(J, 64);
  g.J8 = g.Jp / 8;
  g.I8 = g.Ip / 8;
  if (e == hipSuccess) e = hipMalloc(&c->gq, sizeof(uint32_t) * g.J8 * g.Ip);
  if (e == hipSuccess) e = hipMalloc(&c->gp, sizeof(uint32_t) * g.I8 * g.Jp);
  g.gq = c->gq;
  g.gp = c->gp;
  // Upload the first 2J SNP columns (each bps bytes) and repack into the two nibble streams.
  if (e == hipSuccess) {
    uint8_t* raw = nullptr;
    const int64_t nbytes = 2 * J * bps;
    e = hipMalloc(&raw, (size_t)nbytes);
    if (e == hipSuccess) e = hipMemcpyAsync(raw, bed, (size_t)nbytes, hipMemcpyHostToDevice, c->s);
    if (e == hipSuccess) e = launch_repack(raw, bps, g, c->s);
    if (e == hipSuccess) e = hipStreamSynchronize(c->s);
    if (raw) (void)hipFree(raw);
  }
  if (e == hipSuccess) e = hipMalloc(&c->dm, sizeof(int64_t) * I);
  if (e == hipSuccess) e = launch_count_dm(g, c->dm, c->s);
  if (e != hipSuccess) {
    rc = e == hipErrorOutOfMemory ? HK_ENOMEM : HK_EHIP;
    fail(nullptr, rc, "context setup failed: %s", hipGetErrorString(e));
  } else if ((rc = init_fit_buffers(c)) != HK_OK) {
    snprintf(g_create_err, sizeof g_create_err, "%s", c->err);
  }
  if (rc != HK_OK) {
    hk_ctx_destroy(c);
    return rc;
  }
  *out = c;
  return HK_OK;
}

int hk_ctx_create_shared(hk_ctx** out, hk_ctx* base, int K, int precision) {
  if (!out || !base) return fail(nullptr, HK_EINVAL, "null pointer");
  *out = nullptr;
  if (K < 1 || K > HK_MAX_K || !valid_precision(precision))
    return fail(nullptr, HK_EINVAL, "invalid K=%d or precision=%d", K, precision);
  hipError_t e = hipSetDevice(base->dev);
  if (e != hipSuccess) return fail(nullptr, HK_EHIP, "hipSetDevice: %s", hipGetErrorString(e));
  hk_ctx* c = new hk_ctx();
  c->I = base->I;
  c->J = base->J;
  c->K = K;
  c->dev = base->dev;
  c->owns_geno = false;
  c->g = base->g;
  set_precision(c, precision);
  c->gq = base->gq;
  c->gp = base->gp;
  c->dm = base->dm;
  c->qchunks_req = c->pchunks_req = 0;
  int rc = HK_OK;
  e = hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking);
  if (e != hipSuccess) {
    rc = HK_EHIP;
    fail(nullptr, rc, "hipStreamCreate: %s", hipGetErrorString(e));
  } else if ((rc = init_fit_buffers(c)) != HK_OK) {
    snprintf(g_create_err, sizeof g_create_err, "%s", c->err);
  }
  if (rc != HK_OK) {
    hk_ctx_destroy(c);
    return rc;
  }
  *out = c;
  return HK_OK;
}

static void drop_graph(hk_ctx* c);

int hk_set_tuning(hk_ctx* c, int qc, int pc) {
  if (!c || qc < 0 || pc < 0) return fail(c, HK_EINVAL, "invalid tuning");
  drop_graph(c);
  c->qchunks_req = qc;
  c->pchunks_req = pc;
  plan_chunks(c);
  return alloc_parts(c);
}

int hk_get_tuning(hk_ctx* c, int* qc, int* pc) {
  if (!c) return HK_EINVAL;
  if (qc) *qc = c->g.qchunks;
  if (pc) *pc = c->g.pchunks;
  return HK_OK;
}

int hk_set_pass_shape(hk_ctx* c, int q_shape, int p_shape) {
  if (!c || q_shape < 0 || q_shape > 1 || p_shape < 0 || p_shape > 1) return fail(c, HK_EINVAL, "invalid pass shape");
  if ((q_shape && !has_alt_shape(c->K, c->prec, 0)) || (p_shape && !has_alt_shape(c->K, c->prec, 1)))
    return fail(c, HK_EINVAL, "no second curvature-pass variant for K=%d in this precision", c->K);
  drop_graph(c);  // captured iterations hold the kernel choice
  c->g.qshape = q_shape;
  c->g.pshape = p_shape;
  return HK_OK;
}

int hk_get_pass_shape(hk_ctx* c, int* q_shape, int* p_shape) {
  if (!c) return HK_EINVAL;
  if (q_shape) *q_shape = c->g.qshape;
  if (p_shape) *p_shape = c->g.pshape;
  return HK_OK;
}

int hk_sync(hk_ctx* c) {
  if (!c) return HK_EINVAL;
  HIPC(c, hipStreamSynchronize(c->s));
  return HK_OK;
}

int hk_set_params(hk_ctx* c, const double* q, const double* p) {
  if (!c || !q || !p) return fail(c, HK_EINVAL, "null pointer");
  const int64_t nq = (int64_t)c->K * c->I, np = (int64_t)c->K * 4 * c->J;
  for (int64_t t = 0; t < nq; ++t)
    if (!isfinite(q[t])) return fail(c, HK_ENAN, "q has a non-finite value at %lld", (long long)t);
  for (int64_t t = 0; t < np; ++t)
    if (!isfinite(p[t])) return fail(c, HK_ENAN, "p has a non-finite value at %lld", (long long)t);
  HIPC(c, hipMemcpyAsync(qof(c, c->xb[0]), q, sizeof(double) * nq, hipMemcpyHostToDevice, c->s));
  HIPC(c, hipMemcpyAsync(pof(c, c->xb[0]), p, sizeof(double) * np, hipMemcpyHostToDevice, c->s));
  HIPC(c, hipStreamSynchronize(c->s));
  return HK_OK;
}

int hk_get_params(hk_ctx* c, double* q, double* p) {
  if (!c) return HK_EINVAL;
  if (q) HIPC(c, hipMemcpyAsync(q, qof(c, c->xb[0]), sizeof(double) * c->K * c->I, hipMemcpyDeviceToHost, c->s));
  if (p)
    HIPC(c, hipMemcpyAsync(p, pof(c, c->xb[0]), sizeof(double) * c->K * 4 * c->J, hipMemcpyDeviceToHost, c->s));
  HIPC(c, hipStreamSynchronize(c->s));
  return HK_OK;
}

int hk_loglik(hk_ctx* c, double* ll, double* rows) {
  if (!c || !ll) return fail(c, HK_EINVAL, "null pointer");
  CHK(run_qpass(c, MODE_LL, qof(c, c->xb[0]), pof(c, c->xb[0])));
  HIPC(c, hipMemcpyAsync(ll, c->ll_dev, sizeof(double), hipMemcpyDeviceToHost, c->s));
  if (rows) HIPC(c, hipMemcpyAsync(rows, c->ll_rows, sizeof(double) * c->I, hipMemcpyDeviceToHost, c->s));
  HIPC(c, hipStreamSynchronize(c->s));
  return HK_OK;
}

int hk_em(hk_ctx* c, int fix_q, double* q_next, double* p_next) {
  if (!c || (!fix_q && !q_next) || !p_next) return fail(c, HK_EINVAL, "null pointer");
  const int64_t nq = (int64_t)c->K * c->I, np = (int64_t)c->K * 4 * c->J;
  CHK(ensure_full(c, nq > np ? nq : np));
  double* q = qof(c, c->xb[0]);
  double* p = pof(c, c->xb[0]);
  if (!fix_q) {
    CHK(run_qpass(c, MODE_GRAD, q, p));
    HIPC(c, launch_em_expand(c->res_q, c->K, c->K, c->I, q, c->full_a, c->s));
    HIPC(c, hipMemcpyAsync(q_next, c->full_a, sizeof(double) * nq, hipMemcpyDeviceToHost, c->s));
  }
  CHK(run_ppass(c, MODE_GRAD, q, p));
  HIPC(c, launch_em_expand(c->res_p, c->K, c->K, 4 * c->J, p, c->full_b, c->s));
  HIPC(c, hipMemcpyAsync(p_next, c->full_b, sizeof(double) * np, hipMemcpyDeviceToHost, c->s));
  HIPC(c, hipStreamSynchronize(c->s));
  return HK_OK;
}

int hk_update_q(hk_ctx* c, double* XtX, double* Xtz) {
  if (!c || !XtX || !Xtz) return fail(c, HK_EINVAL, "null pointer");
  const int K = c->K;
  CHK(ensure_full(c, (int64_t)K * K * c->I));
  CHK(run_qpass(c, MODE_CURV, qof(c, c->xb[0]), pof(c, c->xb[0])));
  HIPC(c, launch_expand(c->res_q, K, ntq(K), c->I, c->full_a, c->full_b, c->s));
  HIPC(c, hipMemcpyAsync(XtX, c->full_a, sizeof(double) * K * K * c->I, hipMemcpyDeviceToHost, c->s));
  HIPC(c, hipMemcpyAsync(Xtz, c->full_b, sizeof(double) * K * c->I, hipMemcpyDeviceToHost, c->s));
  HIPC(c, hipStreamSynchronize(c->s));
  return HK_OK;
}

int hk_update_p(hk_ctx* c, double* XtX, double* Xtz) {
  if (!c || !XtX || !Xtz) return fail(c, HK_EINVAL, "null pointer");
  const int K = c->K;
  CHK(ensure_full(c, (int64_t)K * K * 4 * c->J));
  CHK(run_ppass(c, MODE_CURV, qof(c, c->xb[0]), pof(c, c->xb[0])));
  HIPC(c, launch_expand(c->res_p, K, ntq(K), 4 * c->J, c->full_a, c->full_b, c->s));
  HIPC(c, hipMemcpyAsync(XtX, c->full_a, sizeof(double) * K * K * 4 * c->J, hipMemcpyDeviceToHost, c->s));
  HIPC(c, hipMemcpyAsync(Xtz, c->full_b, sizeof(double) * K * 4 * c->J, hipMemcpyDeviceToHost, c->s));
  HIPC(c, hipStreamSynchronize(c->s));
  return HK_OK;
}

int hk_double_missing(hk_ctx* c, int64_t* dm) {
  if (!c || !dm) return fail(c, HK_EINVAL, "null pointer");
  HIPC(c, hipMemcpyAsync(dm, c->dm, sizeof(int64_t) * c->I, hipMemcpyDeviceToHost, c->s));
  HIPC(c, hipStreamSynchronize(c->s));
  return HK_OK;
}

int hk_decode(hk_ctx* c, uint8_t* codes) {
  if (!c || !codes) return fail(c, HK_EINVAL, "null pointer");
  uint8_t* d = nullptr;
  const int64_t n = 2 * c->I * c->J;
  HIPC(c, hipMalloc(&d, (size_t)n));
  hipError_t e = launch_decode(c->g, d, c->s);
  if (e == hipSuccess) e = hipMemcpyAsync(codes, d, (size_t)n, hipMemcpyDeviceToHost, c->s);
  if (e == hipSuccess) e = hipStreamSynchronize(c->s);
  (void)hipFree(d);
  HIPC(c, e);
  return HK_OK;
}

int hk_q_step(hk_ctx* c, int penalty, double rho, double* q_next) {
  if (!c || !q_next) return fail(c, HK_EINVAL, "null pointer");
  double* x = c->xb[0];
  double* xn = c->xb[1];
  CHK(run_qpass(c, MODE_CURV, qof(c, x), pof(c, x)));
  StepOpts so{penalty, rho, 1e-5};
  HIPC(c, launch_set_int(c->fails, 0, c->s));
  HIPC(c, launch_set_int(c->fails + 1, 0, c->s));
  HIPC(c, launch_q_step(c->K, c->I, c->res_q, qof(c, x), qof(c, xn), so, c->fails, c->s));
  int nf = 0;
  HIPC(c, hipMemcpyAsync(&nf, c->fails, sizeof(int), hipMemcpyDeviceToHost, c->s));
  HIPC(c, hipMemcpyAsync(q_next, qof(c, xn), sizeof(double) * c->K * c->I, hipMemcpyDeviceToHost, c->s));
  HIPC(c, hipStreamSynchronize(c->s));
  if (nf) return fail(c, HK_EQP, "%d row QPs failed", nf);
  return HK_OK;
}

int hk_p_step(hk_ctx* c, const double* q_pass, int penalty, double rho, double* p_next) {
  if (!c || !p_next) return fail(c, HK_EINVAL, "null pointer");
  double* x = c->xb[0];
  double* xn = c->xb[1];
  const double* qp = qof(c, x);
  if (q_pass) {
    HIPC(c, hipMemcpyAsync(qof(c, xn), q_pass, sizeof(double) * c->K * c->I, hipMemcpyHostToDevice, c->s));
    qp = qof(c, xn);
  }
  CHK(run_ppass(c, MODE_CURV, qp, pof(c, x)));
  StepOpts so{penalty, rho, 1e-5};
  HIPC(c, launch_set_int(c->fails, 0, c->s));
  HIPC(c, launch_set_int(c->fails + 1, 0, c->s));
  HIPC(c, launch_p_step(c->K, c->J, c->res_p, pof(c, x), pof(c, xn), so, c->fails, p_lists(c), c->s));
  int nf = 0;
  HIPC(c, hipMemcpyAsync(&nf, c->fails, sizeof(int), hipMemcpyDeviceToHost, c->s));
  HIPC(c, hipMemcpyAsync(p_next, pof(c, xn), sizeof(double) * c->K * 4 * c->J, hipMemcpyDeviceToHost, c->s));
  HIPC(c, hipStreamSynchronize(c->s));
  if (nf) return fail(c, HK_EQP, "%d row QPs failed", nf);
  return HK_OK;
}

// ---- runner level: the device arrays of CuAdmixData by name --------------------------------------

// Device address and length (doubles) of a named array; allocates XtX / Xtz arrays on first use.
static int named_buf(hk_ctx* c, int id, double** ptr, int64_t* len) {
  const int64_t K = c->K, I = c->I, J4 = 4 * c->J;
  switch (id) {
    case HK_BUF_Q: *ptr = qof(c, c->xb[0]); *len = K * I; return HK_OK;
    case HK_BUF_P: *ptr = pof(c, c->xb[0]); *len = K * J4; return HK_OK;
    case HK_BUF_Q_NEXT: *ptr = qof(c, c->xb[1]); *len = K * I; return HK_OK;
    case HK_BUF_P_NEXT: *ptr = pof(c, c->xb[1]); *len = K * J4; return HK_OK;
    default: break;
  }
  if (id < HK_BUF_XTX_Q || id > HK_BUF_XTZ_P) return fail(c, HK_EINVAL, "unknown device array %d", id);
  const int64_t lens[4] = {K * K * I, K * I, K * K * J4, K * J4};
  const int t = id - HK_BUF_XTX_Q;
  if (!c->named[t]) {
    HIPC(c, hipMalloc(&c->named[t], sizeof(double) * lens[t]));
    HIPC(c, launch_fill(c->named[t], lens[t], 0.0, c->s));
  }
  *ptr = c->named[t];
  *len = lens[t];
  return HK_OK;
}

// Same shape class: q-shaped (K x I), p-shaped (K x 4J), or the same curvature / gradient array kind.
static int buf_shape(int id) {
  switch (id) {
    case HK_BUF_Q: case HK_BUF_Q_NEXT: return 0;
    case HK_BUF_P: case HK_BUF_P_NEXT: return 1;
    default: return 2 + id;
  }
}

int hk_buf_upload(hk_ctx* c, int id, const double* host) {
  if (!c || !host) return fail(c, HK_EINVAL, "null pointer");
  double* d;
  int64_t n;
  CHK(named_buf(c, id, &d, &n));
  if (id <= HK_BUF_P_NEXT)
    for (int64_t t = 0; t < n; ++t)
      if (!isfinite(host[t])) return fail(c, HK_ENAN, "array %d has a non-finite value at %lld", id, (long long)t);
  HIPC(c, hipMemcpyAsync(d, host, sizeof(double) * n, hipMemcpyHostToDevice, c->s));
  HIPC(c, hipStreamSynchronize(c->s));
  return HK_OK;
}

int hk_buf_download(hk_ctx* c, int id, double* host) {
  if (!c || !host) return fail(c, HK_EINVAL, "null pointer");
  double* d;
  int64_t n;
  CHK(named_buf(c, id, &d, &n));
  HIPC(c, hipMemcpyAsync(host, d, sizeof(double) * n, hipMemcpyDeviceToHost, c->s));
  HIPC(c, hipStreamSynchronize(c->s));
  return HK_OK;
}

int hk_buf_copy(hk_ctx* c, int dst, int src) {
  if (!c) return HK_EINVAL;
  if (buf_shape(dst) != buf_shape(src)) return fail(c, HK_EINVAL, "arrays %d and %d differ in shape", dst, src);
  double *a, *b;
  int64_t na, nb;
  CHK(named_buf(c, dst, &a, &na));
  CHK(named_buf(c, src, &b, &nb));
  if (a != b) HIPC(c, launch_copy_f64(na, b, a, c->s));
  HIPC(c, hipStreamSynchronize(c->s));
  return HK_OK;
}

int hk_loglik_buf(hk_ctx* c, int qbuf, int pbuf, double* ll) {
  if (!c || !ll) return fail(c, HK_EINVAL, "null pointer");
  if (buf_shape(qbuf) != 0 || buf_shape(pbuf) != 1) return fail(c, HK_EINVAL, "loglikelihood needs a q and a p array");
  double *q, *p;
  int64_t n;
  CHK(named_buf(c, qbuf, &q, &n));
  CHK(named_buf(c, pbuf, &p, &n));
  CHK(run_qpass(c, MODE_LL, q, p));
  HIPC(c, hipMemcpyAsync(ll, c->ll_dev, sizeof(double), hipMemcpyDeviceToHost, c->s));
  HIPC(c, hipStreamSynchronize(c->s));
  return HK_OK;
}

int hk_em_buf(hk_ctx* c, int fix_q) {
  if (!c) return HK_EINVAL;
  double* q = qof(c, c->xb[0]);
  double* p = pof(c, c->xb[0]);
  if (!fix_q) {
    CHK(run_qpass(c, MODE_GRAD, q, p));
    HIPC(c, launch_em_expand(c->res_q, c->K, c->K, c->I, q, qof(c, c->xb[1]), c->s));
  }
  CHK(run_ppass(c, MODE_GRAD, q, p));
  HIPC(c, launch_em_expand(c->res_p, c->K, c->K, 4 * c->J, p, pof(c, c->xb[1]), c->s));
  HIPC(c, hipStreamSynchronize(c->s));
  return HK_OK;
}

int hk_update_q_buf(hk_ctx* c) {
  if (!c) return HK_EINVAL;
  double *XtX, *Xtz;
  int64_t n;
  CHK(named_buf(c, HK_BUF_XTX_Q, &XtX, &n));
  CHK(named_buf(c, HK_BUF_XTZ_Q, &Xtz, &n));
  CHK(run_qpass(c, MODE_CURV, qof(c, c->xb[0]), pof(c, c->xb[0])));
  HIPC(c, launch_expand(c->res_q, c->K, ntq(c->K), c->I, XtX, Xtz, c->s));
  HIPC(c, hipStreamSynchronize(c->s));
  return HK_OK;
}

int hk_update_p_buf(hk_ctx* c) {
  if (!c) return HK_EINVAL;
  double *XtX, *Xtz;
  int64_t n;
  CHK(named_buf(c, HK_BUF_XTX_P, &XtX, &n));
  CHK(named_buf(c, HK_BUF_XTZ_P, &Xtz, &n));
  CHK(run_ppass(c, MODE_CURV, qof(c, c->xb[0]), pof(c, c->xb[0])));
  HIPC(c, launch_expand(c->res_p, c->K, ntq(c->K), 4 * c->J, XtX, Xtz, c->s));
  HIPC(c, hipStreamSynchronize(c->s));
  return HK_OK;
}

// count_double_missing!(doublemissing, g) as src/structs.jl:169 calls it: over the SNP pairs of the whole
// file, not only the fit's J blocks.  The columns past the first 2J are streamed through a bounded
// staging buffer and counted on the device; counts of the first J blocks come from the packed stream.
int hk_count_double_missing_file(hk_ctx* c, const uint8_t* bed, int64_t bps, int64_t n_snp) {
  if (!c || !bed) return fail(c, HK_EINVAL, "null pointer");
  if (!c->owns_geno) return fail(c, HK_EINVAL, "call on the context that owns the genotype upload");
  if (bps < (c->I + 3) / 4 || n_snp < 2 * c->J)
    return fail(c, HK_EINVAL, "bytes_per_snp %lld < ceil(I/4) or n_snp %lld < 2J", (long long)bps, (long long)n_snp);
  HIPC(c, launch_count_dm(c->g, c->dm, c->s));
  const int64_t nblk = n_snp / 2;  // complete SNP pairs (an odd last SNP has no partner)
  const int64_t chunk = (int64_t)1 << 14;  // blocks per staging round: 2 * 16384 * bps bytes
  uint8_t* raw = nullptr;
  if (nblk > c->J) {
    const int64_t cb = nblk - c->J < chunk ? nblk - c->J : chunk;
    HIPC(c, hipMalloc(&raw, (size_t)(2 * cb * bps)));
  }
  hipError_t e = hipSuccess;
  for (int64_t b0 = c->J; b0 < nblk && e == hipSuccess; b0 += chunk) {
    const int64_t nb = nblk - b0 < chunk ? nblk - b0 : chunk;
    e = hipMemcpyAsync(raw, bed + 2 * b0 * bps, (size_t)(2 * nb * bps), hipMemcpyHostToDevice, c->s);
    if (e == hipSuccess) e = launch_count_dm_raw(raw, bps, c->I, nb, c->dm, c->s);
    if (e == hipSuccess) e = hipStreamSynchronize(c->s);  // the staging buffer is reused next round
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->s);
  if (raw) (void)hipFree(raw);
  HIPC(c, e);
  return HK_OK;
}

void hk_fit_opts_default(hk_fit_opts* o) {
  if (!o) return;
  o->em_iters = 5;
  o->n_iter = 1000;
  o->rtol = 1e-7;
  o->Qn = 3;
  o->fix_q = 0;
  o->fix_p = 0;
  o->q_penalty = 0;
  o->p_penalty = 1;
  o->rho = 1e7;
  o->eps = 1e-5;
  o->verbose = 0;
  o->profile = 0;
  o->use_graph = 1;
  o->em_fix_q = 0;
  o->guarantee_increase = 0;
  o->qn_mode = 0;
  o->batch_rounds = 1;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------------------
// Device-resident fit
// ---------------------------------------------------------------------------------------------------

// ---------------------------------------------------------------------------------------------------
// One QN iteration of admixture_qn! (algorithms_outer.jl:65-171) enqueued without host synchronisation:
//   F(x) -> x_next, F(F(x)) -> x_next2, ll_basic, ZAL -> x_tmp, projections, ll_qn, accept into x.
// Fixed buffers (x = xb[0], x_next = xb[1], x_next2 = xb[2], x_tmp = xb[3]) and device-resident
// iteration counter / log-likelihoods make the sequence replayable from a captured hipGraph.
// ---------------------------------------------------------------------------------------------------
// Log-likelihood of (q, p) on the host (one pass + a synchronisation).
static int eval_ll(hk_ctx* c, const double* q, const double* p, Timer* tm, double* ll) {
  CHK(run_qpass(c, MODE_LL, q, p, tm, T_LL, c->d_ll + 5));
  HIPC(c, hipMemcpyAsync(c->h_ll + 5, c->d_ll + 5, sizeof(double), hipMemcpyDeviceToHost, c->s));
  HIPC(c, hipStreamSynchronize(c->s));
  *ll = c->h_ll[5];
  return HK_OK;
}

// guarantee_increase line search of update_q! / update_p! (algorithms_inner.jl:170-199, 314-343): out
// already holds proj(base + delta); try proj(base + f delta) for f = 1, 1/2, ..., 2^-9 and accept the
// first whose log-likelihood is >= ll_tmp (Q step) or > ll_tmp (P step), raising ll_tmp to it.  When no
// factor qualifies, out keeps the last one tried (f = 2^-9), as the reference's loop leaves it.
static int halve_step(hk_ctx* c, const hk_fit_opts* o, Timer* tm, bool is_q, const double* base, const double* delta,
                      double* out, const double* q_ll, const double* p_ll, double* ll_tmp) {
  const int64_t n = is_q ? (int64_t)c->K * c->I : (int64_t)c->K * 4 * c->J;
  double f = 1.0;
  for (;;) {
    if (f < 1.0) {
      HIPC(c, launch_axpy_step(n, base, delta, f, out, c->s));
      if (is_q) HIPC(c, launch_project_q(c->K, c->I, out, o->eps, c->s));
      else HIPC(c, launch_project_p(c->K, c->J, out, o->eps, c->s));
    }
    double ll;
    CHK(eval_ll(c, q_ll, p_ll, tm, &ll));
    if (is_q ? ll >= *ll_tmp : ll > *ll_tmp) {
      *ll_tmp = ll;
      return HK_OK;
    }
    if (o->verbose)
      fprintf(stderr, "Step size halved in update_%c. comparison: %.17g, current: %.17g, current step size: %g\n",
              is_q ? 'q' : 'p', *ll_tmp, ll, f);
    f /= 2;
    if (f < 1e-3) return HK_OK;
  }
}

// ll_prev: the iteration's starting log-likelihood (needed by guarantee_increase only).  gate / ctl / trace:
// set when the iteration is part of a batch graph -- every kernel is gated on the fit's device gate and the
// device-side stopping rule (launch_fit_control) closes the iteration.
static int enqueue_iteration(hk_ctx* c, const hk_fit_opts* o, Timer* tm, double ll_prev = NAN, int* gate = nullptr,
                             FitCtl* ctl = nullptr, double* trace = nullptr) {
  const int K = c->K;
  const int64_t I = c->I, J = c->J;
  const int64_t N = (int64_t)K * (I + 4 * J), nq = (int64_t)K * I, np = (int64_t)K * 4 * J;
  double* x = c->xb[0];
  double* xn = c->xb[1];
  double* xn2 = c->xb[2];
  double* xt = c->xb[3];
  StepOpts qso{o->q_penalty, o->rho, o->eps};
  StepOpts pso{o->p_penalty, o->rho, o->eps};
  qso.gate = pso.gate = gate;
  // guarantee_increase: the raw QP steps live in x_tmp (d.q_tmp / d.p_tmp in the reference) until ZAL
  // overwrites it; every step is then line-searched on the host, so the iteration cannot be captured.
  const bool gi = o->guarantee_increase != 0;
  double* dq = gi ? qof(c, xt) : nullptr;
  double* dp = gi ? pof(c, xt) : nullptr;
  double ll_tmp = ll_prev;
  auto qstep = [&](const double* q, const double* p, double* q_out) -> int {
    CHK(run_qpass(c, MODE_CURV, q, p, tm, T_QPASS, nullptr, gate));
    if (tm) tm->begin(T_QSTEP);
    HIPC(c, launch_q_step(K, I, c->res_q, q, q_out, qso, c->fails, c->s, dq));
    if (tm) tm->end();
    if (gi) CHK(halve_step(c, o, tm, true, q, dq, q_out, q_out, p, &ll_tmp));
    return HK_OK;
  };
  auto pstep = [&](const double* q, const double* p, double* p_out) -> int {
    CHK(run_ppass(c, MODE_CURV, q, p, tm, T_PPASS, gate));
    if (tm) tm->begin(T_PSTEP);
    HIPC(c, launch_p_step(K, J, c->res_p, p, p_out, pso, c->fails, p_lists(c), c->s, dp));
    if (tm) tm->end();
    if (gi) CHK(halve_step(c, o, tm, false, p, dp, p_out, q, p_out, &ll_tmp));
    return HK_OK;
  };
  // F(x): q_next from (q, p), p_next from (q_next, p)  (update_q!, update_p!; Gauss-Seidel order)
  if (!o->fix_q) CHK(qstep(qof(c, x), pof(c, x), qof(c, xn)));
  else HIPC(c, launch_copy_f64(nq, qof(c, x), qof(c, xn), c->s, gate));
  if (!o->fix_p) CHK(pstep(qof(c, xn), pof(c, x), pof(c, xn)));
  else HIPC(c, launch_copy_f64(np, pof(c, x), pof(c, xn), c->s, gate));
  // F(F(x)): q_next2 from (q_next, p_next), p_next2 from (q_next2, p_next)
  if (!o->fix_q) CHK(qstep(qof(c, xn), pof(c, xn), qof(c, xn2)));
  else HIPC(c, launch_copy_f64(nq, qof(c, x), qof(c, xn2), c->s, gate));
  if (!o->fix_p) CHK(pstep(qof(c, xn2), pof(c, xn), pof(c, xn2)));
  else HIPC(c, launch_copy_f64(np, pof(c, x), pof(c, xn2), c->s, gate));
  // The QN point needs only x, F(x), F(F(x)) (algorithms_outer.jl:128-144), so it is formed first and both
  // log-likelihoods -- ll_basic of F(F(x)) (:119-126) and ll_qn of the projected QN point (:146-153) -- come
  // from one genotype sweep.
  if (tm) tm->begin(T_QN);
  HIPC(c, launch_zal(N, o->Qn, c->d_iter, c->U, c->V, x, xn, xn2, xt, c->zal_part, c->zal_coef, c->s, o->qn_mode,
                     gate));
  HIPC(c, launch_project_p(K, J, pof(c, xt), o->eps, c->s, gate));
  HIPC(c, launch_project_q(K, I, qof(c, xt), o->eps, c->s, gate));
  if (tm) tm->end();
  CHK(run_ll_pair(c, qof(c, xn2), pof(c, xn2), qof(c, xt), pof(c, xt), c->d_ll + 0, c->d_ll + 1, tm, T_LL, gate));
  HIPC(c, launch_accept(N, c->d_ll, xn2, xt, x, c->d_iter, c->s, gate));
  if (ctl) HIPC(c, launch_fit_control(c->d_ll, ctl, trace, gate, c->s));
  return HK_OK;
}

// Options that change the enqueued sequence (graph cache key).
static unsigned long long graph_key(const hk_ctx* c, const hk_fit_opts* o) {
  unsigned long long k = 1469598103934665603ull;
  auto mix = [&](unsigned long long v) { k = (k ^ v) * 1099511628211ull; };
  mix((unsigned)o->fix_q); mix((unsigned)o->fix_p); mix((unsigned)o->q_penalty); mix((unsigned)o->p_penalty);
  mix((unsigned)o->Qn); mix(*(const unsigned long long*)&o->rho); mix(*(const unsigned long long*)&o->eps);
  mix((unsigned)o->qn_mode);
  mix((unsigned)c->g.qchunks); mix((unsigned)c->g.pchunks);
  return k;
}

static void drop_graph(hk_ctx* c) {
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->graph_exec = nullptr;
  c->graph = nullptr;
  c->graph_key = 0;
}

// Capture one iteration into a graph (once per option set) and launch it; eager enqueue otherwise.
static int launch_iteration(hk_ctx* c, const hk_fit_opts* o, bool use_graph, double ll_prev) {
  if (!use_graph || o->guarantee_increase) return enqueue_iteration(c, o, nullptr, ll_prev);
  const unsigned long long key = graph_key(c, o);
  if (!c->graph_exec || c->graph_key != key) {
    drop_graph(c);
    HIPC(c, hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_iteration(c, o, nullptr);
    hipGraph_t gr = nullptr;
    hipError_t e = hipStreamEndCapture(c->s, &gr);
    if (rc != HK_OK) {
      if (gr) (void)hipGraphDestroy(gr);
      return rc;
    }
    HIPC(c, e);
    c->graph = gr;
    HIPC(c, hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0));
    c->graph_key = key;
  }
  HIPC(c, hipGraphLaunch(c->graph_exec, c->s));
  return HK_OK;
}

static int ensure_fit_buffers(hk_ctx* c, int Qn) {
  const int64_t N = (int64_t)c->K * (c->I + 4 * c->J);
  if (c->Qn_alloc < Qn) {
    drop_graph(c);
    if (c->U) HIPC(c, hipFree(c->U));
    if (c->V) HIPC(c, hipFree(c->V));
    c->U = c->V = nullptr;
    HIPC(c, hipMalloc(&c->U, sizeof(double) * N * Qn));
    HIPC(c, hipMalloc(&c->V, sizeof(double) * N * Qn));
    c->Qn_alloc = Qn;
  }
  return HK_OK;
}

static int read_ll(hk_ctx* c, int n) {
  HIPC(c, hipMemcpyAsync(c->h_ll, c->d_ll, sizeof(double) * n, hipMemcpyDeviceToHost, c->s));
  HIPC(c, hipStreamSynchronize(c->s));
  return HK_OK;
}

// init_em! pass (algorithms_outer.jl:4-16): EM update into x, log-likelihood -> d_ll[3] (want_ll), projections.
// init_em! uses the log-likelihood of an EM iteration only for its progress line and, for the last one, as
// d.ll_new; the others are skipped unless verbose.
static int enqueue_em_iteration(hk_ctx* c, const hk_fit_opts* o, Timer* tm, bool want_ll = true) {
  const int K = c->K;
  double* x = c->xb[0];
  double* xn = c->xb[1];
  const int64_t N = (int64_t)K * (c->I + 4 * c->J);
  if (tm) tm->begin(T_EM);
  if (o->em_fix_q) {  // em_p! (algorithms_inner.jl:40): P-only EM update, q kept
    HIPC(c, launch_copy_f64((int64_t)K * c->I, qof(c, x), qof(c, xn), c->s));
  } else {
    CHK(run_qpass(c, MODE_GRAD, qof(c, x), pof(c, x)));
    HIPC(c, launch_em_q(K, c->I, c->J, c->res_q, K, c->dm, qof(c, x), qof(c, xn), o->eps, c->s));
  }
  CHK(run_ppass(c, MODE_GRAD, qof(c, x), pof(c, x)));
  HIPC(c, launch_em_p(K, c->J, c->res_p, K, pof(c, x), pof(c, xn), o->eps, c->s));
  if (tm) tm->end();
  HIPC(c, launch_copy_f64(N, xn, x, c->s));  // d.p .= d.p_next; d.q .= d.q_next
  if (want_ll) CHK(run_qpass(c, MODE_LL, qof(c, x), pof(c, x), tm, T_LL, c->d_ll + 3));
  HIPC(c, launch_project_p(K, c->J, pof(c, x), o->eps, c->s));
  HIPC(c, launch_project_q(K, c->I, qof(c, x), o->eps, c->s));
  return HK_OK;
}

namespace {
// Per-fit host state of the QN loop.
struct FitRun {
  hk_ctx* c;
  const hk_fit_opts* o;
  double* ll_trace;
  double* t_trace;
  hk_fit_result* res;
  Timer tm;
  double ll_new = NAN;
  int iter = 0;
  bool active = true;
  std::chrono::steady_clock::time_point t0;
};

// One completion event per fit of a batch (no timing).
struct BatchEvents {
  std::vector<hipEvent_t> e;
  hipError_t status = hipSuccess;
  explicit BatchEvents(int n) : e(n, nullptr) {
    for (int b = 0; b < n && status == hipSuccess; ++b)
      status = hipEventCreateWithFlags(&e[b], hipEventDisableTiming);
  }
  ~BatchEvents() {
    for (auto x : e)
      if (x) (void)hipEventDestroy(x);
  }
};
}  // namespace

static int fit_prepare(FitRun& f, double ll_start) {
  hk_ctx* c = f.c;
  const hk_fit_opts* o = f.o;
  if (o->Qn < 1 || o->Qn > 4 || o->n_iter < 0 || o->em_iters < 0 || o->qn_mode < 0 || o->qn_mode > 1)
    return fail(c, HK_EINVAL, "invalid fit options (Qn in 1..4)");
  memset(f.res, 0, sizeof(*f.res));
  f.t0 = std::chrono::steady_clock::now();
  f.tm.on = o->profile != 0;
  f.tm.s = c->s;
  CHK(ensure_fit_buffers(c, o->Qn));
  HIPC(c, launch_set_int(c->fails, 0, c->s));
  HIPC(c, launch_set_int(c->fails + 1, 0, c->s));
  HIPC(c, launch_set_int(c->d_iter, 1, c->s));
  return HK_OK;
}

static int fit_em(FitRun& f, double ll_start) {
  hk_ctx* c = f.c;
  const hk_fit_opts* o = f.o;
  double ll = NAN;
  for (int it = 0; it < o->em_iters; ++it) {
    const bool want_ll = o->verbose || it + 1 == o->em_iters;
    CHK(enqueue_em_iteration(c, o, f.tm.on ? &f.tm : nullptr, want_ll));
    if (!want_ll) continue;
    CHK(read_ll(c, 4));
    ll = c->h_ll[3];
    if (o->verbose) fprintf(stderr, "EM iter %d, ll: %.17g\n", it + 1, ll);
  }
  if (o->em_iters == 0) ll = ll_start;
  if (isnan(ll)) {
    CHK(run_qpass(c, MODE_LL, qof(c, c->xb[0]), pof(c, c->xb[0]), f.tm.on ? &f.tm : nullptr, T_LL, c->d_ll + 3));
    CHK(read_ll(c, 4));
    ll = c->h_ll[3];
  }
  f.ll_new = ll;
  f.res->ll_init = ll;
  if (o->verbose) fprintf(stderr, "initial ll: %.17g\n", ll);
  if (o->n_iter == 0) f.active = false;
  return HK_OK;
}

// Host side of one iteration after its work completed: trace, accept, stopping rule.
static int fit_after_iteration(FitRun& f) {
  hk_ctx* c = f.c;
  const hk_fit_opts* o = f.o;
  const double ll_prev = f.ll_new;
  const double ll_basic = c->h_ll[0], ll_qn = c->h_ll[1];
  f.ll_new = c->h_ll[2];
  ++f.iter;
  if (f.ll_trace) f.ll_trace[f.iter - 1] = f.ll_new;
  if (f.t_trace)
    f.t_trace[f.iter - 1] = std::chrono::duration<double>(std::chrono::steady_clock::now() - f.t0).count();
  const double reldiff = fabs((f.ll_new - ll_prev) / ll_prev);
  if (o->verbose)
    fprintf(stderr, "Iteration %d: ll=%.17g, reldiff = %.6g, ll_basic=%.17g, ll_qn=%.17g\n", f.iter, f.ll_new, reldiff,
            ll_basic, ll_qn);
  if (!isfinite(f.ll_new)) return fail(c, HK_ENAN, "log-likelihood became non-finite at iteration %d", f.iter);
  if (reldiff < o->rtol) {
    f.res->converged = 1;
    f.active = false;
  } else if (f.iter >= o->n_iter) {
    f.active = false;
  }
  return HK_OK;
}

static int fit_finish(FitRun& f) {
  hk_ctx* c = f.c;
  int nf[2] = {0, 0};
  HIPC(c, hipMemcpyAsync(nf, c->fails, 2 * sizeof(int), hipMemcpyDeviceToHost, c->s));
  HIPC(c, hipStreamSynchronize(c->s));
  hk_fit_result* r = f.res;
  r->ll = f.ll_new;
  r->iters = f.iter;
  r->qp_failures = nf[0];
  r->qp_iter_limit = nf[1];
  if (f.tm.on) {
    double t[T_NCAT] = {0};
    int n[T_NCAT] = {0};
    f.tm.collect(t, n);
    r->t_qpass_ms = t[T_QPASS]; r->n_qpass = n[T_QPASS];
    r->t_ppass_ms = t[T_PPASS]; r->n_ppass = n[T_PPASS];
    r->t_ll_ms = t[T_LL]; r->n_ll = n[T_LL];
    r->t_qstep_ms = t[T_QSTEP]; r->n_qstep = n[T_QSTEP];
    r->t_pstep_ms = t[T_PSTEP]; r->n_pstep = n[T_PSTEP];
    r->t_em_ms = t[T_EM]; r->n_em = n[T_EM];
    r->t_qn_ms = t[T_QN]; r->n_qn = n[T_QN];
  }
  r->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - f.t0).count();
  return HK_OK;
}

extern "C" int hk_fit_ex(hk_ctx* c, const hk_fit_opts* o, double ll_start, double* ll_trace, double* t_trace,
                         hk_fit_result* res) {
  if (!c || !o || !res) return fail(c, HK_EINVAL, "null pointer");
  FitRun f;
  f.c = c;
  f.o = o;
  f.ll_trace = ll_trace;
  f.t_trace = t_trace;
  f.res = res;
  CHK(fit_prepare(f, ll_start));
  CHK(fit_em(f, ll_start));
  // per-kernel events and the step-halving line search need the eager path
  const bool use_graph = o->use_graph && !o->profile && !o->guarantee_increase;
  while (f.active) {
    if (use_graph) CHK(launch_iteration(c, o, true, f.ll_new));
    else CHK(enqueue_iteration(c, o, f.tm.on ? &f.tm : nullptr, f.ll_new));
    CHK(read_ll(c, 3));
    CHK(fit_after_iteration(f));
  }
  return fit_finish(f);
}

extern "C" int hk_fit(hk_ctx* c, const hk_fit_opts* o, double ll_start, double* ll_trace, hk_fit_result* res) {
  return hk_fit_ex(c, o, ll_start, ll_trace, nullptr, res);
}

namespace {
// Device state of a round-graph batch: per-fit gates and stopping-rule state, the fork / join stream and
// events, the captured round.  Released after the root stream synchronised.
struct RoundBatch {
  int n = 0;
  int* gates = nullptr;      // device, n: 1 while the fit runs
  FitCtl* ctls = nullptr;    // device, n
  int* h_gates = nullptr;    // pinned, 2 x n: the gates at the last two checks
  hipStream_t root = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join;
  hipEvent_t chk[2] = {nullptr, nullptr};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~RoundBatch() {
    if (root) (void)hipStreamSynchronize(root);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto e : join)
      if (e) (void)hipEventDestroy(e);
    for (auto e : {fork, chk[0], chk[1]})
      if (e) (void)hipEventDestroy(e);
    if (root) (void)hipStreamDestroy(root);
    if (gates) (void)hipFree(gates);
    if (ctls) (void)hipFree(ctls);
    if (h_gates) (void)hipHostFree(h_gates);
  }
};
}  // namespace

// QN iterations of every fit as ONE captured graph per round: a fork from a root stream to each fit's stream,
// kRoundIters whole iterations of the fit (every kernel gated on the fit's device gate, each iteration closed
// by the device-side stopping rule), a join.  The same executable graph is replayed round after round; a fit whose stopping rule fired skips every
// kernel of later rounds, so nothing is re-captured as fits converge, and each fit runs exactly the
// iterations hk_fit would run.  One round stays in flight while the host reads the gates of the previous one.
// QN iterations of each fit per round: a fit that stops within a round skips the rest of it (gated).
constexpr int kRoundIters = 1;

static int fit_batch_rounds(hk_ctx** ctxs, int n, const hk_fit_opts* opts, std::vector<FitRun>& runs,
                            double* seconds_at_finish, std::chrono::steady_clock::time_point t0) {
  hk_ctx* c0 = ctxs[0];
  RoundBatch rb;
  rb.n = n;
  rb.join.assign(n, nullptr);
  HIPC(c0, hipSetDevice(c0->dev));
  HIPC(c0, hipMalloc(&rb.gates, sizeof(int) * n));
  HIPC(c0, hipMalloc(&rb.ctls, sizeof(FitCtl) * n));
  HIPC(c0, hipHostMalloc((void**)&rb.h_gates, sizeof(int) * 2 * n, hipHostMallocDefault));
  HIPC(c0, hipStreamCreateWithFlags(&rb.root, hipStreamNonBlocking));
  HIPC(c0, hipEventCreateWithFlags(&rb.fork, hipEventDisableTiming));
  HIPC(c0, hipEventCreateWithFlags(&rb.chk[0], hipEventDisableTiming));
  HIPC(c0, hipEventCreateWithFlags(&rb.chk[1], hipEventDisableTiming));
  for (int b = 0; b < n; ++b) HIPC(c0, hipEventCreateWithFlags(&rb.join[b], hipEventDisableTiming));
  std::vector<int> g0(n);
  std::vector<FitCtl> ctl(n);
  for (int b = 0; b < n; ++b) {
    g0[b] = runs[b].active ? 1 : 0;
    ctl[b] = FitCtl{runs[b].ll_new, opts[b].rtol, opts[b].n_iter, 0, 0, 0};
    HIPC(ctxs[b], hipStreamSynchronize(ctxs[b]->s));  // the warm-up and set-up work of every fit is done
  }
  HIPC(c0, hipMemcpy(rb.gates, g0.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  HIPC(c0, hipMemcpy(rb.ctls, ctl.data(), sizeof(FitCtl) * n, hipMemcpyHostToDevice));
  // capture one round, the costliest fits (largest K x individuals x blocks) first: their branches set the
  // length of a round, so they are dispatched before the cheap ones
  std::vector<int> order(n);
  for (int b = 0; b < n; ++b) order[b] = b;
  auto cost = [&](int b) { return (double)ctxs[b]->K * ctxs[b]->K * (double)ctxs[b]->I * (double)ctxs[b]->J; };
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost(a) > cost(b); });
  HIPC(c0, hipStreamBeginCapture(rb.root, hipStreamCaptureModeThreadLocal));
  int rc = HK_OK;
  hipError_t e = hipEventRecord(rb.fork, rb.root);
  for (int t = 0; t < n && e == hipSuccess && rc == HK_OK; ++t) {
    const int b = order[t];
    hk_ctx* c = ctxs[b];
    e = hipStreamWaitEvent(c->s, rb.fork, 0);
    for (int r = 0; r < kRoundIters && e == hipSuccess && rc == HK_OK; ++r)
      rc = enqueue_iteration(c, &opts[b], nullptr, NAN, rb.gates + b, rb.ctls + b, nullptr);
    if (e == hipSuccess && rc == HK_OK) e = hipEventRecord(rb.join[b], c->s);
    if (e == hipSuccess && rc == HK_OK) e = hipStreamWaitEvent(rb.root, rb.join[b], 0);
  }
  hipGraph_t gr = nullptr;
  const hipError_t ee = hipStreamEndCapture(rb.root, &gr);
  rb.graph = gr;
  if (rc != HK_OK) return rc;
  HIPC(c0, e);
  HIPC(c0, ee);
  HIPC(c0, hipGraphInstantiate(&rb.exec, rb.graph, nullptr, nullptr, 0));
  // replay: launch round r + 1 before reading the gates after round r
  std::vector<char> stopped(n, 0);
  auto now_s = [&]() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
  for (int b = 0; b < n; ++b)
    if (!runs[b].active) {
      stopped[b] = 1;
      if (seconds_at_finish) seconds_at_finish[b] = now_s();
    }
  auto launch_round = [&](int slot) -> int {
    HIPC(c0, hipGraphLaunch(rb.exec, rb.root));
    HIPC(c0, hipMemcpyAsync(rb.h_gates + slot * n, rb.gates, sizeof(int) * n, hipMemcpyDeviceToHost, rb.root));
    HIPC(c0, hipEventRecord(rb.chk[slot], rb.root));
    return HK_OK;
  };
  CHK(launch_round(0));
  for (int round = 1;; ++round) {
    CHK(launch_round(round & 1));
    const int prev = (round - 1) & 1;
    HIPC(c0, hipEventSynchronize(rb.chk[prev]));
    int active = 0;
    for (int b = 0; b < n; ++b) {
      if (rb.h_gates[prev * n + b]) {
        ++active;
      } else if (!stopped[b]) {
        stopped[b] = 1;
        if (seconds_at_finish) seconds_at_finish[b] = now_s();
      }
    }
    if (active == 0) break;
  }
  HIPC(c0, hipStreamSynchronize(rb.root));
  HIPC(c0, hipMemcpy(ctl.data(), rb.ctls, sizeof(FitCtl) * n, hipMemcpyDeviceToHost));
  for (int b = 0; b < n; ++b) {
    FitRun& f = runs[b];
    if (ctl[b].status) return fail(ctxs[b], HK_ENAN, "log-likelihood became non-finite at iteration %d", ctl[b].iter);
    if (ctl[b].iter > 0) {
      f.iter = ctl[b].iter;
      f.ll_new = ctl[b].ll_prev;
      f.res->converged = ctl[b].converged;
    }
    f.active = false;
  }
  return HK_OK;
}

// Several fits of contexts over one genotype upload (hk_ctx_create_shared), each with its own buffers and
// stream: the EM warm-up in lockstep, then the QN iterations either as one captured graph per round
// (batch_rounds) or each fit's iteration enqueued as soon as its previous one completed.
extern "C" int hk_fit_batch(hk_ctx** ctxs, int n, const hk_fit_opts* opts, const double* ll_start,
                            hk_fit_result* res, double* seconds_at_finish) {
  if (!ctxs || n < 1 || !opts || !res) return fail(nullptr, HK_EINVAL, "hk_fit_batch: invalid arguments");
  std::vector<FitRun> runs(n);
  const auto t0 = std::chrono::steady_clock::now();
  for (int b = 0; b < n; ++b) {
    runs[b].c = ctxs[b];
    runs[b].o = &opts[b];
    runs[b].ll_trace = nullptr;
    runs[b].t_trace = nullptr;
    runs[b].res = &res[b];
    CHK(fit_prepare(runs[b], ll_start ? ll_start[b] : NAN));
  }
  // EM warm-up, lockstep over EM iterations.
  int max_em = 0;
  for (int b = 0; b < n; ++b) max_em = opts[b].em_iters > max_em ? opts[b].em_iters : max_em;
  std::vector<double> ll_em(n, NAN);
  for (int it = 0; it < max_em; ++it) {
    for (int b = 0; b < n; ++b)
      if (it < opts[b].em_iters)
        CHK(enqueue_em_iteration(ctxs[b], &opts[b], nullptr, opts[b].verbose || it + 1 == opts[b].em_iters));
    for (int b = 0; b < n; ++b)
      if (it + 1 == opts[b].em_iters) {
        CHK(read_ll(ctxs[b], 4));
        ll_em[b] = ctxs[b]->h_ll[3];
      }
  }
  for (int b = 0; b < n; ++b) {
    hk_fit_opts oz = opts[b];
    oz.em_iters = 0;
    runs[b].o = &oz;
    CHK(fit_em(runs[b], opts[b].em_iters > 0 ? ll_em[b] : (ll_start ? ll_start[b] : NAN)));
    runs[b].o = &opts[b];
  }
  bool rounds = opts[0].batch_rounds != 0;
  for (int b = 0; b < n; ++b)
    rounds = rounds && opts[b].use_graph && !opts[b].guarantee_increase && !opts[b].profile && !opts[b].verbose;
  if (rounds) {
    CHK(fit_batch_rounds(ctxs, n, opts, runs, seconds_at_finish, t0));
    for (int b = 0; b < n; ++b) CHK(fit_finish(runs[b]));
    return HK_OK;
  }
  // QN iterations, event-driven: a fit's next iteration is enqueued as soon as its previous one (and the
  // copy of its log-likelihoods) completed, so no fit waits for the slowest one of a round.  A fit's
  // results do not depend on the order (every context owns its buffers and stream).
  BatchEvents ev(n);
  HIPC(nullptr, ev.status);
  auto launch = [&](int b) -> int {
    hk_ctx* c = ctxs[b];
    CHK(launch_iteration(c, &opts[b], opts[b].use_graph != 0, runs[b].ll_new));
    HIPC(c, hipMemcpyAsync(c->h_ll, c->d_ll, sizeof(double) * 3, hipMemcpyDeviceToHost, c->s));
    HIPC(c, hipEventRecord(ev.e[b], c->s));
    return HK_OK;
  };
  int active = 0;
  for (int b = 0; b < n; ++b)
    if (runs[b].active) {
      CHK(launch(b));
      ++active;
    }
  while (active > 0) {
    for (int b = 0; b < n; ++b) {
      if (!runs[b].active) continue;
      const hipError_t e = hipEventQuery(ev.e[b]);
      if (e == hipErrorNotReady) continue;
      HIPC(ctxs[b], e);
      CHK(fit_after_iteration(runs[b]));
      if (runs[b].active) {
        CHK(launch(b));
      } else {
        --active;
        if (seconds_at_finish)
          seconds_at_finish[b] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      }
    }
  }
  for (int b = 0; b < n; ++b) CHK(fit_finish(runs[b]));
  return HK_OK;
}

namespace {
struct EventPair {
  hipEvent_t a = nullptr, b = nullptr;
  ~EventPair() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
};
}  // namespace

// Device time of a whole pass as a fit runs it -- parameter conversion, staging tables, the pass kernel,
// the reduction over its chunks and the finishing kernels: the chunk count trades the pass kernel's
// parallelism against the partial sums that the reduction reads.  kind 0 = log-likelihood, 1 = update_q
// curvature pass, 2 = update_p curvature pass.
static int time_full_pass(hk_ctx* c, int kind, int reps, double* ms) {
  EventPair ev;
  HIPC(c, hipEventCreate(&ev.a));
  HIPC(c, hipEventCreate(&ev.b));
  double* x = c->xb[0];
  double total = 0.0;
  for (int r = 0; r < reps; ++r) {
    HIPC(c, hipEventRecord(ev.a, c->s));
    if (kind == 0) CHK(run_qpass(c, MODE_LL, qof(c, x), pof(c, x), nullptr, 0, c->d_ll + 5));
    else if (kind == 1) CHK(run_qpass(c, MODE_CURV, qof(c, x), pof(c, x)));
    else CHK(run_ppass(c, MODE_CURV, qof(c, x), pof(c, x)));
    HIPC(c, hipEventRecord(ev.b, c->s));
    HIPC(c, hipEventSynchronize(ev.b));
    float t = 0.f;
    HIPC(c, hipEventElapsedTime(&t, ev.a, ev.b));
    total += t;
  }
  *ms = total / reps;
  return HK_OK;
}

// Autotuner: time the chunk (tile) shapes of the individual-parallel passes (log-likelihood + update_q)
// and of the block-parallel update_p pass on the context's data, each with its reduction, keep the fastest.
extern "C" int hk_autotune(hk_ctx* c, int reps, int* q_chunks, int* p_chunks) {
  if (!c || reps < 1) return fail(c, HK_EINVAL, "invalid arguments");
  drop_graph(c);
  const int q0 = c->g.qchunks, p0 = c->g.pchunks;
  const int qreq0 = c->qchunks_req, preq0 = c->pchunks_req;
  int qbest = q0, pbest = p0;
  double tq_best = 1e300, tp_best = 1e300;
  const int mult[] = {1, 2, 4, 8, 16};
  const int div[] = {4, 2};
  std::vector<int> qc, pc;
  for (int d : div) if (q0 / d >= 1) qc.push_back(q0 / d);
  for (int m : mult) qc.push_back(q0 * m);
  for (int d : div) if (p0 / d >= 1) pc.push_back(p0 / d);
  for (int m : mult) pc.push_back(p0 * m);
  // chunk counts x curvature-pass variants (the second variant only where one exists for K)
  const int nqs = has_alt_shape(c->K, c->prec, 0) ? 2 : 1, nps = has_alt_shape(c->K, c->prec, 1) ? 2 : 1;
  int qsbest = 0, psbest = 0;
  for (int qv : qc) {
    if (qv > c->g.J8) continue;
    CHK(hk_set_tuning(c, qv, preq0));
    double tl = 0;
    CHK(time_full_pass(c, 0, reps, &tl));
    for (int sh = 0; sh < nqs; ++sh) {
      CHK(hk_set_pass_shape(c, sh, 0));
      double tu = 0;
      CHK(time_full_pass(c, 1, reps, &tu));
      if (tl + tu < tq_best) { tq_best = tl + tu; qbest = c->g.qchunks; qsbest = sh; }
    }
  }
  for (int pv : pc) {
    if (pv > c->g.I8) continue;
    CHK(hk_set_tuning(c, qbest, pv));
    for (int sh = 0; sh < nps; ++sh) {
      CHK(hk_set_pass_shape(c, qsbest, sh));
      double tp = 0;
      CHK(time_full_pass(c, 2, reps, &tp));
      if (tp < tp_best) { tp_best = tp; pbest = c->g.pchunks; psbest = sh; }
    }
  }
  (void)qreq0;
  CHK(hk_set_tuning(c, qbest, pbest));
  CHK(hk_set_pass_shape(c, qsbest, psbest));
  if (q_chunks) *q_chunks = qbest;
  if (p_chunks) *p_chunks = pbest;
  return HK_OK;
}

extern "C" int hk_time_pass(hk_ctx* c, int kind, int reps, double* ms) {
  if (!c || !ms || reps < 1 || kind < 0 || kind > 3) return fail(c, HK_EINVAL, "invalid arguments");
  hipEvent_t a, b;
  HIPC(c, hipEventCreate(&a));
  HIPC(c, hipEventCreate(&b));
  double* x = c->xb[0];
  double total = 0.0;
  int rc = HK_OK;
  for (int r = 0; r < reps && rc == HK_OK; ++r) {
    const void *qq, *pp;
    rc = pass_params(c, qof(c, x), pof(c, x), &qq, &pp);
    if (rc != HK_OK) break;
    hipError_t e = hipSuccess;
    if (kind == 1 || kind == 2) e = launch_curv_tables(c->K, c->prec, kind == 1 ? 0 : 1, c->g, qq, pp, c->s);
    if (e != hipSuccess) {
      rc = fail(c, HK_EHIP, "timing failed: %s", hipGetErrorString(e));
      break;
    }
    (void)hipEventRecord(a, c->s);
    if (kind == 0)
      e = c->prec == PREC_FP64 ? launch_qside(c->K, c->prec, MODE_LL, c->g, qq, pp, c->part_q, c->s)
                               : launch_ll_mm(c->K, c->g, (const float*)qq, (const float*)pp, nullptr, nullptr,
                                              c->part_q, c->ll_stage, c->s);
    else if (kind == 1) e = launch_qside(c->K, c->prec, MODE_CURV, c->g, qq, pp, c->part_q, c->s);
    else if (kind == 2)
      e = launch_pside(c->K, c->prec, MODE_CURV, c->g, qq, pp, c->g.pchunks == 1 ? c->res_p : c->part_p, c->s);
    else e = launch_pside(c->K, c->prec, MODE_GRAD, c->g, qq, pp, c->g.pchunks == 1 ? c->res_p : c->part_p, c->s);
    (void)hipEventRecord(b, c->s);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, a, b);
    if (e != hipSuccess) rc = fail(c, HK_EHIP, "timing failed: %s", hipGetErrorString(e));
    total += t;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (rc == HK_OK) *ms = total / reps;
  return rc;
}

extern "C" int hk_copy_bench(hk_ctx* c, int64_t bytes, int reps, double* gbps) {
  if (!c || bytes < 16 || reps < 1 || !gbps) return fail(c, HK_EINVAL, "invalid arguments");
  const int64_t n16 = bytes / 16;
  void *a = nullptr, *b = nullptr;
  HIPC(c, hipMalloc(&a, n16 * 16));
  hipError_t e = hipMalloc(&b, n16 * 16);
  hipEvent_t t0 = nullptr, t1 = nullptr;
  if (e == hipSuccess) e = hipMemsetAsync(a, 0, n16 * 16, c->s);
  if (e == hipSuccess) e = hipEventCreate(&t0);
  if (e == hipSuccess) e = hipEventCreate(&t1);
  if (e == hipSuccess) e = launch_copy16(a, b, n16, c->s);  // warm
  if (e == hipSuccess) e = hipEventRecord(t0, c->s);
  for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch_copy16(a, b, n16, c->s);
  if (e == hipSuccess) e = hipEventRecord(t1, c->s);
  if (e == hipSuccess) e = hipEventSynchronize(t1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
  if (t0) (void)hipEventDestroy(t0);
  if (t1) (void)hipEventDestroy(t1);
  (void)hipFree(a);
  if (b) (void)hipFree(b);
  HIPC(c, e);
  *gbps = 2.0 * (double)(n16 * 16) * reps / (ms * 1e-3) / 1e9;
  return HK_OK;
}

extern "C" int hk_mlp_predict(hk_mlp* m, hk_ctx* c, int n_blocks, uint64_t seed, float* P) {
  if (!c || !m || !P || n_blocks < 1 || n_blocks > c->J)
    return fail(c, HK_EINVAL, "invalid arguments (1 <= n_blocks <= J)");
  HIPC(c, hipSetDevice(c->dev));
  HIPC(c, hipStreamSynchronize(c->s));  // genotype streams ready before the network's stream reads them
  const int rc = mlp_predict_from(m, c->g, n_blocks, seed, P);
  if (rc != HK_OK) return fail(c, rc, "hk_mlp_predict: %s", hk_mlp_last_error(m));
  return HK_OK;
}

extern "C" int hk_features(hk_ctx* c, int n_blocks, uint64_t seed, float* F) {
  if (!c || !F || n_blocks < 1 || n_blocks > c->J) return fail(c, HK_EINVAL, "invalid arguments (1 <= n_blocks <= J)");
  return project_features(c->g, n_blocks, seed, F, c->s, c->err, sizeof c->err);
}

extern "C" int hk_skfr(hk_ctx* c, int k, int sparsity, int mode, int max_iter, int tries, uint64_t seed,
                       const int32_t* init, int32_t* clusters, int32_t* aims, int* n_aims, double* wss, int* iters,
                       double* prof) {
  if (!c || !clusters || !aims || !n_aims) return fail(c, HK_EINVAL, "null pointer");
  HIPC(c, hipSetDevice(c->dev));
  const int rc = skfr_run(c->g, k, sparsity, mode, max_iter, tries, seed, init, nullptr, clusters, aims, n_aims, wss,
                          iters, prof, c->s, c->err, sizeof c->err);
  return rc;
}

extern "C" int hk_skfr_path(hk_ctx* c, int k, int n_s, const int32_t* sparsities, int mode, int max_iter, int tries,
                            uint64_t seed, int32_t* clusters, int32_t* aims, int* n_aims, double* wss, int* iters) {
  if (!c || !sparsities || !clusters || !aims || !n_aims || n_s < 1) return fail(c, HK_EINVAL, "invalid arguments");
  for (int t = 1; t < n_s; ++t)
    if (sparsities[t] > sparsities[t - 1]) return fail(c, HK_EINVAL, "sparsities should be decreasing");
  HIPC(c, hipSetDevice(c->dev));
  int64_t off = 0;
  for (int t = 0; t < n_s; ++t) {
    const int s = sparsities[t];
    const int64_t room = 2 * (c->J < (int64_t)k * s ? c->J : (int64_t)k * s);
    int na = 0;
    // the first sparsity: the best of `tries` seeded starts; each next one from the previous labels
    const int rc = skfr_run(c->g, k, s, mode, max_iter, t == 0 ? tries : 1, seed, nullptr,
                            t == 0 ? nullptr : clusters + (int64_t)(t - 1) * c->I, clusters + (int64_t)t * c->I,
                            aims + off, &na, wss ? wss + t : nullptr, iters ? iters + t : nullptr, nullptr, c->s,
                            c->err, sizeof c->err);
    if (rc != 0) return rc;
    n_aims[t] = na;
    off += room;
  }
  return HK_OK;
}
