// hk_api.hip — C ABI of libhaplo_hip.so: context, parameter transfer, the reference's GPU runner entry
// points and the device-resident QN/SQP fit loop.  Host code; the kernels live in hk_pass.hip and
// hk_step.hip.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <chrono>
#include <utility>
#include <vector>

#include "../../include/haplo_hip.h"
#include "hk_internal.h"

using namespace hk;

struct hk_ctx {
  int64_t I = 0, J = 0;
  int K = 0, prec = 0, dev = 0;
  PassGeom g{};
  hipStream_t s = nullptr;
  uint32_t* gq = nullptr;
  uint32_t* gp = nullptr;
  void* qq = nullptr;
  void* pp = nullptr;
  int64_t* dm = nullptr;
  double* xb[4] = {nullptr, nullptr, nullptr, nullptr};  // parameter buffers [q | p | pad]
  int64_t xlen = 0;                                      // K * (I + 4 Jp)
  float* x32 = nullptr;                                  // FP32 mirror of one parameter buffer
  double* res_q = nullptr;   // [I][NT]
  double* res_p = nullptr;   // [4Jp][NT]
  double* part_q = nullptr;  // [qchunks][NT][Ip]
  double* part_p = nullptr;  // [pchunks][4Jp][NT]
  double* ll_dev = nullptr;  // scalar
  double* ll_rows = nullptr; // [I]
  double* U = nullptr;
  double* V = nullptr;
  int Qn_alloc = 0;
  double* zal_part = nullptr;
  double* zal_coef = nullptr;
  int* fails = nullptr;
  double* full_a = nullptr;  // ABI expansion buffers
  double* full_b = nullptr;
  int64_t full_len = 0;
  int qchunks_req = 0, pchunks_req = 0;
  int64_t part_q_len = 0, part_p_len = 0;
  char err[512] = {0};
};

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

static int64_t round_up(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

static int ntq(int K) { return K + K * (K + 1) / 2; }

// Chunking of the passes so that one launch has >= ~2048 workgroups (8 per CU on 256 CUs).
static void plan_chunks(hk_ctx* c) {
  PassGeom& g = c->g;
  const int64_t wgq = g.Ip / 256, wgp = 4 * g.Jp / 256;
  int64_t qc = c->qchunks_req > 0 ? c->qchunks_req : (2048 + wgq - 1) / wgq;
  int64_t pc = c->pchunks_req > 0 ? c->pchunks_req : (2048 + wgp - 1) / wgp;
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
  const int NT = ntq(c->K);
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
static int pass_params(hk_ctx* c, const double* q, const double* p, const void** qq, const void** pp) {
  if (c->prec == PREC_FP64) {
    *qq = q;
    *pp = p;
    return HK_OK;
  }
  const int64_t nq = (int64_t)c->K * c->I, np = (int64_t)c->K * 4 * c->g.Jp;
  HIPC(c, launch_to_f32(q, c->x32, nq, c->s));
  HIPC(c, launch_to_f32(p, c->x32 + nq, np, c->s));
  *qq = c->x32;
  *pp = c->x32 + nq;
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

// Q-side pass (individual-parallel): MODE_LL -> ll_dev / ll_rows, else res_q.
static int run_qpass(hk_ctx* c, int mode, const double* q, const double* p, Timer* tm = nullptr, int cat = 0) {
  const void *qq, *pp;
  CHK(pass_params(c, q, p, &qq, &pp));
  if (tm) tm->begin(cat);
  HIPC(c, launch_qside(c->K, c->prec, mode, c->g, qq, pp, c->part_q, c->s));
  if (tm) tm->end();
  if (mode == MODE_LL) HIPC(c, launch_reduce_ll(c->part_q, c->g.qchunks, c->g, c->ll_rows, c->ll_dev, c->s));
  else if (mode == MODE_CURV) HIPC(c, launch_reduce_q_curv(c->part_q, c->g.qchunks, c->K, c->g, q, c->res_q, c->s));
  else HIPC(c, launch_reduce_q(c->part_q, c->g.qchunks, nacc(c->K, mode), c->g, c->res_q, c->s));
  return HK_OK;
}

// P-side pass (block-parallel) -> res_p [4Jp][NT].
static int run_ppass(hk_ctx* c, int mode, const double* q, const double* p, Timer* tm = nullptr, int cat = 0) {
  const void *qq, *pp;
  CHK(pass_params(c, q, p, &qq, &pp));
  const int NT = nacc(c->K, mode);
  if (c->g.pchunks == 1) {
    if (tm) tm->begin(cat);
    HIPC(c, launch_pside(c->K, c->prec, mode, c->g, qq, pp, c->res_p, c->s));
    if (tm) tm->end();
  } else {
    if (tm) tm->begin(cat);
    HIPC(c, launch_pside(c->K, c->prec, mode, c->g, qq, pp, c->part_p, c->s));
    if (tm) tm->end();
    HIPC(c, launch_reduce_p(c->part_p, c->g.pchunks, NT, c->g, c->res_p, c->s));
  }
  if (mode == MODE_CURV) HIPC(c, launch_finish_p(c->K, c->g, p, c->res_p, c->s));
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
  void* ptrs[] = {c->gq, c->gp, c->qq, c->pp, c->dm, c->xb[0], c->xb[1], c->xb[2], c->xb[3], c->x32, c->res_q, c->res_p,
                  c->part_q, c->part_p, c->ll_dev, c->ll_rows, c->U, c->V, c->zal_part, c->zal_coef, c->fails,
                  c->full_a, c->full_b};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (c->s) (void)hipStreamDestroy(c->s);
  delete c;
  return HK_OK;
}

int hk_ctx_create(hk_ctx** out, int64_t I, int64_t J, int K, const uint8_t* bed, int64_t bps, int64_t n_snp,
                  int precision) {
  if (!out) return fail(nullptr, HK_EINVAL, "out is NULL");
  *out = nullptr;
  if (I <= 0 || J <= 0 || K < 1 || K > HK_MAX_K || !bed || bps < (I + 3) / 4 || n_snp < 2 * J ||
      (precision != HK_PREC_FP64 && precision != HK_PREC_MIXED))
    return fail(nullptr, HK_EINVAL,
                "invalid arguments: I=%lld J=%lld K=%d bytes_per_snp=%lld n_snp=%lld precision=%d (need K in 1..%d, "
                "bytes_per_snp >= ceil(I/4), n_snp >= 2J)",
                (long long)I, (long long)J, K, (long long)bps, (long long)n_snp, precision, HK_MAX_K);
  if (I > (int64_t)1 << 31 || J > (int64_t)1 << 29)
    return fail(nullptr, HK_EINVAL, "I or J too large for this build");
  hk_ctx* c = new hk_ctx();
  c->I = I;
  c->J = J;
  c->K = K;
  c->prec = precision;
  hipError_t e = hipGetDevice(&c->dev);
  if (e != hipSuccess) {
    fail(nullptr, HK_EHIP, "hipGetDevice: %s", hipGetErrorString(e));
    delete c;
    return HK_EHIP;
  }
#define CREATE_HIP(call)                                                                   \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess) {                                                                \
      fail(nullptr, e_ == hipErrorOutOfMemory ? HK_ENOMEM : HK_EHIP, "%s failed: %s", #call, \
           hipGetErrorString(e_));                                                         \
      hk_ctx_destroy(c);                                                                   \
      return e_ == hipErrorOutOfMemory ? HK_ENOMEM : HK_EHIP;                              \
    }                                                                                      \
  } while (0)
  CREATE_HIP(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
  PassGeom& g = c->g;
  g.I = I;
  g.J = J;
  g.Ip = round_up(I, 256);
  g.Jp = round_up(J, 64);
  g.J8 = g.Jp / 8;
  g.I8 = g.Ip / 8;
  CREATE_HIP(hipMalloc(&c->gq, sizeof(uint32_t) * g.J8 * g.Ip));
  CREATE_HIP(hipMalloc(&c->gp, sizeof(uint32_t) * g.I8 * g.Jp));
  {
    const int qs = 16 * ((K * (K + 1) / 2 + 15) / 16);
    const size_t el = precision == HK_PREC_FP64 ? sizeof(double) : sizeof(float);
    CREATE_HIP(hipMalloc(&c->qq, el * g.Ip * qs));
    CREATE_HIP(hipMalloc(&c->pp, el * 4 * g.Jp * qs));
  }
  g.qq = c->qq;
  g.pp = c->pp;
  g.use_mfma = 1;
  g.gq = c->gq;
  g.gp = c->gp;
  // Upload the first 2J SNP columns (each bps bytes) and repack into the two nibble streams.
  {
    uint8_t* raw = nullptr;
    const int64_t nbytes = 2 * J * bps;
    CREATE_HIP(hipMalloc(&raw, (size_t)nbytes));
    e = hipMemcpyAsync(raw, bed, (size_t)nbytes, hipMemcpyHostToDevice, c->s);
    if (e == hipSuccess) e = launch_repack(raw, bps, g, c->s);
    if (e == hipSuccess) e = hipStreamSynchronize(c->s);
    (void)hipFree(raw);
    CREATE_HIP(e);
  }
  CREATE_HIP(hipMalloc(&c->dm, sizeof(int64_t) * I));
  CREATE_HIP(launch_count_dm(g, c->dm, c->s));
  c->xlen = (int64_t)K * (I + 4 * g.Jp);
  for (int b = 0; b < 4; ++b) {
    CREATE_HIP(hipMalloc(&c->xb[b], sizeof(double) * c->xlen));
    CREATE_HIP(launch_fill(c->xb[b], (int64_t)K * I, 1.0 / K, c->s));
    CREATE_HIP(launch_fill(c->xb[b] + (int64_t)K * I, (int64_t)K * 4 * g.Jp, 0.25, c->s));
  }
  if (precision == HK_PREC_MIXED) CREATE_HIP(hipMalloc(&c->x32, sizeof(float) * c->xlen));
  const int NT = ntq(K);
  CREATE_HIP(hipMalloc(&c->res_q, sizeof(double) * I * NT));
  CREATE_HIP(hipMalloc(&c->res_p, sizeof(double) * 4 * g.Jp * NT));
  CREATE_HIP(hipMalloc(&c->ll_dev, sizeof(double) * 4));
  CREATE_HIP(hipMalloc(&c->ll_rows, sizeof(double) * I));
  CREATE_HIP(hipMalloc(&c->fails, sizeof(int)));
  CREATE_HIP(hipMalloc(&c->zal_part, sizeof(double) * 512 * 64));
  CREATE_HIP(hipMalloc(&c->zal_coef, sizeof(double) * 8));
  plan_chunks(c);
  if (alloc_parts(c) != HK_OK) {
    snprintf(g_create_err, sizeof g_create_err, "%s", c->err);
    hk_ctx_destroy(c);
    return HK_ENOMEM;
  }
  CREATE_HIP(hipStreamSynchronize(c->s));
#undef CREATE_HIP
  *out = c;
  return HK_OK;
}

int hk_set_tuning(hk_ctx* c, int qc, int pc) {
  if (!c || qc < 0 || pc < 0) return fail(c, HK_EINVAL, "invalid tuning");
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
  HIPC(c, hipMemsetAsync(c->fails, 0, sizeof(int), c->s));
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
  HIPC(c, hipMemsetAsync(c->fails, 0, sizeof(int), c->s));
  HIPC(c, launch_p_step(c->K, c->J, c->res_p, pof(c, x), pof(c, xn), so, c->fails, c->s));
  int nf = 0;
  HIPC(c, hipMemcpyAsync(&nf, c->fails, sizeof(int), hipMemcpyDeviceToHost, c->s));
  HIPC(c, hipMemcpyAsync(p_next, pof(c, xn), sizeof(double) * c->K * 4 * c->J, hipMemcpyDeviceToHost, c->s));
  HIPC(c, hipStreamSynchronize(c->s));
  if (nf) return fail(c, HK_EQP, "%d row QPs failed", nf);
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
}

}  // extern "C"

// ---------------------------------------------------------------------------------------------------
// Device-resident fit
// ---------------------------------------------------------------------------------------------------

static int fit_loglik(hk_ctx* c, Timer& tm, double* x, double* out) {
  CHK(run_qpass(c, MODE_LL, qof(c, x), pof(c, x), &tm, T_LL));
  HIPC(c, hipMemcpyAsync(out, c->ll_dev, sizeof(double), hipMemcpyDeviceToHost, c->s));
  HIPC(c, hipStreamSynchronize(c->s));
  return HK_OK;
}

// F-map half sweeps (update_q! / update_p!, src/algorithms_inner.jl:97-355).
static int fit_qstep(hk_ctx* c, Timer& tm, const double* q, const double* p, double* q_out, const StepOpts& so) {
  CHK(run_qpass(c, MODE_CURV, q, p, &tm, T_QPASS));
  tm.begin(T_QSTEP);
  HIPC(c, launch_q_step(c->K, c->I, c->res_q, q, q_out, so, c->fails, c->s));
  tm.end();
  return HK_OK;
}

static int fit_pstep(hk_ctx* c, Timer& tm, const double* q, const double* p, double* p_out, const StepOpts& so) {
  CHK(run_ppass(c, MODE_CURV, q, p, &tm, T_PPASS));
  tm.begin(T_PSTEP);
  HIPC(c, launch_p_step(c->K, c->J, c->res_p, p, p_out, so, c->fails, c->s));
  tm.end();
  return HK_OK;
}

extern "C" int hk_fit_ex(hk_ctx* c, const hk_fit_opts* o, double ll_start, double* ll_trace, double* t_trace,
                         hk_fit_result* res) {
  if (!c || !o || !res) return fail(c, HK_EINVAL, "null pointer");
  if (o->Qn < 1 || o->Qn > 4 || o->n_iter < 0 || o->em_iters < 0)
    return fail(c, HK_EINVAL, "invalid fit options (Qn in 1..4)");
  const auto t0 = std::chrono::steady_clock::now();
  memset(res, 0, sizeof(*res));
  const int K = c->K;
  const int64_t I = c->I, J = c->J;
  const int64_t N = (int64_t)K * (I + 4 * J);
  if (c->Qn_alloc < o->Qn) {
    if (c->U) HIPC(c, hipFree(c->U));
    if (c->V) HIPC(c, hipFree(c->V));
    c->U = c->V = nullptr;
    HIPC(c, hipMalloc(&c->U, sizeof(double) * N * o->Qn));
    HIPC(c, hipMalloc(&c->V, sizeof(double) * N * o->Qn));
    c->Qn_alloc = o->Qn;
  }
  Timer tm;
  tm.on = o->profile != 0;
  tm.s = c->s;
  HIPC(c, hipMemsetAsync(c->fails, 0, sizeof(int), c->s));
  double*& x = c->xb[0];
  double*& xn = c->xb[1];
  double*& xn2 = c->xb[2];
  double*& xt = c->xb[3];
  const StepOpts qso{o->q_penalty, o->rho, o->eps};
  const StepOpts pso{o->p_penalty, o->rho, o->eps};
  const int NTg = K;  // gradient-only accumulator count
  double ll_new = NAN;
  // init_em! (src/algorithms_outer.jl:1-22); driver.jl:182 runs it without fix_q/fix_p.
  for (int it = 0; it < o->em_iters; ++it) {
    tm.begin(T_EM);
    CHK(run_qpass(c, MODE_GRAD, qof(c, x), pof(c, x)));
    HIPC(c, launch_em_q(K, I, J, c->res_q, NTg, c->dm, qof(c, x), qof(c, xn), o->eps, c->s));
    CHK(run_ppass(c, MODE_GRAD, qof(c, x), pof(c, x)));
    HIPC(c, launch_em_p(K, J, c->res_p, NTg, pof(c, x), pof(c, xn), o->eps, c->s));
    tm.end();
    std::swap(x, xn);  // d.p .= d.p_next; d.q .= d.q_next
    CHK(fit_loglik(c, tm, x, &ll_new));
    HIPC(c, launch_project_p(K, J, pof(c, x), o->eps, c->s));
    HIPC(c, launch_project_q(K, I, qof(c, x), o->eps, c->s));
    if (o->verbose) fprintf(stderr, "EM iter %d, ll: %.17g\n", it + 1, ll_new);
  }
  if (o->em_iters == 0) ll_new = ll_start;
  if (isnan(ll_new)) CHK(fit_loglik(c, tm, x, &ll_new));
  res->ll_init = ll_new;
  if (o->verbose) fprintf(stderr, "initial ll: %.17g\n", ll_new);
  const int64_t nq = (int64_t)K * I, np = (int64_t)K * 4 * J;
  int iters = 0, conv = 0;
  for (int i = 1; i <= o->n_iter; ++i) {
    iters = i;
    const double ll_prev = ll_new;
    // F(x): q_next from (q, p); p_next from (q_next, p)
    if (!o->fix_q) CHK(fit_qstep(c, tm, qof(c, x), pof(c, x), qof(c, xn), qso));
    else HIPC(c, hipMemcpyAsync(qof(c, xn), qof(c, x), sizeof(double) * nq, hipMemcpyDeviceToDevice, c->s));
    if (!o->fix_p) CHK(fit_pstep(c, tm, qof(c, xn), pof(c, x), pof(c, xn), pso));
    else HIPC(c, hipMemcpyAsync(pof(c, xn), pof(c, x), sizeof(double) * np, hipMemcpyDeviceToDevice, c->s));
    // F(F(x)): q_next2 from (q_next, p_next); p_next2 from (q_next2, p_next)
    if (!o->fix_q) CHK(fit_qstep(c, tm, qof(c, xn), pof(c, xn), qof(c, xn2), qso));
    else HIPC(c, hipMemcpyAsync(qof(c, xn2), qof(c, x), sizeof(double) * nq, hipMemcpyDeviceToDevice, c->s));
    if (!o->fix_p) CHK(fit_pstep(c, tm, qof(c, xn2), pof(c, xn), pof(c, xn2), pso));
    else HIPC(c, hipMemcpyAsync(pof(c, xn2), pof(c, x), sizeof(double) * np, hipMemcpyDeviceToDevice, c->s));
    double ll_basic, ll_qn;
    CHK(fit_loglik(c, tm, xn2, &ll_basic));
    tm.begin(T_QN);
    HIPC(c, launch_zal(N, o->Qn, i, c->U, c->V, x, xn, xn2, xt, c->zal_part, c->zal_coef, c->s));
    HIPC(c, launch_project_p(K, J, pof(c, xt), o->eps, c->s));
    HIPC(c, launch_project_q(K, I, qof(c, xt), o->eps, c->s));
    tm.end();
    CHK(fit_loglik(c, tm, xt, &ll_qn));
    if (ll_basic < ll_qn) {
      std::swap(x, xt);
      ll_new = ll_qn;
    } else {
      std::swap(x, xn2);
      ll_new = ll_basic;
    }
    if (ll_trace) ll_trace[i - 1] = ll_new;
    if (t_trace) t_trace[i - 1] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const double reldiff = fabs((ll_new - ll_prev) / ll_prev);
    if (o->verbose)
      fprintf(stderr, "Iteration %d: ll=%.17g, reldiff = %.6g, ll_basic=%.17g, ll_qn=%.17g\n", i, ll_new, reldiff,
              ll_basic, ll_qn);
    if (!isfinite(ll_new)) return fail(c, HK_ENAN, "log-likelihood became non-finite at iteration %d", i);
    if (reldiff < o->rtol) {
      conv = 1;
      break;
    }
  }
  int nf = 0;
  HIPC(c, hipMemcpyAsync(&nf, c->fails, sizeof(int), hipMemcpyDeviceToHost, c->s));
  HIPC(c, hipStreamSynchronize(c->s));
  res->ll = ll_new;
  res->iters = iters;
  res->converged = conv;
  res->qp_failures = nf;
  if (tm.on) {
    double t[T_NCAT] = {0};
    int n[T_NCAT] = {0};
    tm.collect(t, n);
    res->t_qpass_ms = t[T_QPASS]; res->n_qpass = n[T_QPASS];
    res->t_ppass_ms = t[T_PPASS]; res->n_ppass = n[T_PPASS];
    res->t_ll_ms = t[T_LL]; res->n_ll = n[T_LL];
    res->t_qstep_ms = t[T_QSTEP]; res->n_qstep = n[T_QSTEP];
    res->t_pstep_ms = t[T_PSTEP]; res->n_pstep = n[T_PSTEP];
    res->t_em_ms = t[T_EM]; res->n_em = n[T_EM];
    res->t_qn_ms = t[T_QN]; res->n_qn = n[T_QN];
  }
  res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return HK_OK;
}

extern "C" int hk_fit(hk_ctx* c, const hk_fit_opts* o, double ll_start, double* ll_trace, hk_fit_result* res) {
  return hk_fit_ex(c, o, ll_start, ll_trace, nullptr, res);
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
    (void)hipEventRecord(a, c->s);
    hipError_t e;
    if (kind == 0) e = launch_qside(c->K, c->prec, MODE_LL, c->g, qq, pp, c->part_q, c->s);
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
