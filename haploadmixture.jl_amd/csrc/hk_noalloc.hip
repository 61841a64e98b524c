// hk_noalloc.hip — hk_loglik_noalloc: the log-likelihood pass of hk_loglik for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops).  It allocates nothing: the pass's
// partial sums and staged parameters live in `scratch`, a file-scope device array of fixed size that each GPU
// holds for the life of the process.  Every call on one GPU shares that GPU's array, so calls that overlap on
// one GPU race; the pass overwrites every scratch word it reads, so nothing has to start at zero.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>

#include <mutex>

#include "../../include/haplo_hip.h"
#include "hk_internal.h"

using namespace hk;

static constexpr size_t kScratchBytes = (size_t)64 << 20;  // the stated maximum: 64 MiB per GPU

// Host code never uses this name as an address: it is a host-side stand-in.  The device address of the
// current GPU's copy comes from hipGetSymbolAddress (scratch_ptr below).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices];

int nfail(hk_ctx* c, int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(c->err, sizeof c->err, fmt, ap);
  va_end(ap);
  return code;
}

// The scratch address on device dev, resolved on the first call that uses dev (with dev made current for
// hipGetSymbolAddress, the caller's current device restored afterwards) and kept per device.
hipError_t scratch_ptr(int dev, void** out) {
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != dev && (e = hipSetDevice(dev)) != hipSuccess) return e;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess) return e;
    g_ptr[dev] = p;
  }
  *out = g_ptr[dev];
  return hipSuccess;
}

// Pass parameters in scratch: p (K x 4J) followed by 0.25 for the padding blocks (the pass reads p for every
// padded block; their double-missing nibbles make the value irrelevant), in the pass precision; q converted to
// FP32 for the MIXED engine (the FP64 engine reads the caller's q in place, qd = nullptr).
template <typename TP>
__global__ void k_noalloc_params(const double* __restrict__ q, int64_t nq, const double* __restrict__ p, int64_t np,
                                 int64_t np_pad, TP* __restrict__ qd, TP* __restrict__ pd) {
  const int64_t n = np_pad + (qd ? nq : 0);
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    if (t < np_pad) pd[t] = t < np ? (TP)p[t] : (TP)0.25;
    else qd[t - np_pad] = (TP)q[t - np_pad];
  }
}

size_t up256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

extern "C" int hk_loglik_noalloc(hk_ctx* c, const double* q, const double* p, double* ll, double* rows) {
  if (!c) return HK_EINVAL;
  if (!q || !p || !ll) return nfail(c, HK_EINVAL, "hk_loglik_noalloc: null pointer");
  const PassGeom& g0 = c->g;
  const bool f32 = c->prec != PREC_FP64;
  const size_t el = f32 ? sizeof(float) : sizeof(double);
  const int64_t nq = (int64_t)c->K * c->I, np = (int64_t)c->K * 4 * c->J, np_pad = (int64_t)c->K * 4 * g0.Jp;
  // layout: [p padded | q (MIXED) | BF16 terms of p (MIXED) | rows (unless the caller passes them) |
  //          chunk partials, workgroup sums]
  const size_t b_p = up256(el * (size_t)np_pad), b_q = f32 ? up256(el * (size_t)nq) : 0;
  const size_t b_st = f32 ? up256(ll_stage_bytes(c->K, g0.J8)) : 0;
  const size_t b_rows = rows ? 0 : up256(sizeof(double) * (size_t)c->I);
  const size_t nb = (size_t)((c->I + 255) / 256);
  const size_t fixed = b_p + b_q + b_st + b_rows + sizeof(double) * nb;
  const size_t per_chunk = sizeof(double) * (size_t)g0.Ip;
  if (fixed + per_chunk > kScratchBytes)
    return nfail(c, HK_EINVAL,
                 "hk_loglik_noalloc: I=%lld J=%lld K=%d needs at least %zu bytes of scratch; its maximum is %zu",
                 (long long)c->I, (long long)c->J, c->K, fixed + per_chunk, kScratchBytes);
  void* base = nullptr;
  hipError_t e = scratch_ptr(c->dev, &base);
  if (e != hipSuccess) return nfail(c, HK_EHIP, "hk_loglik_noalloc: scratch address: %s", hipGetErrorString(e));
  unsigned char* sp = (unsigned char*)base;
  void* pd = sp;
  void* qd = f32 ? (void*)(sp + b_p) : nullptr;
  void* stage = f32 ? (void*)(sp + b_p + b_q) : nullptr;
  double* rows_d = rows ? rows : (double*)(sp + b_p + b_q + b_st);
  double* part = (double*)(sp + b_p + b_q + b_st + b_rows);
  // The context's j chunks when they fit (the sums are then those of hk_loglik, bit for bit), fewer otherwise.
  PassGeom g = g0;
  int64_t nch = g0.qchunks;
  const int64_t room = (int64_t)((kScratchBytes - fixed) / per_chunk);
  if (nch > room) nch = room;
  g.qwpc = (int)((g.J8 + nch - 1) / nch);
  g.qchunks = (int)((g.J8 + g.qwpc - 1) / g.qwpc);
  const int64_t n = np_pad + (f32 ? nq : 0);
  const unsigned grid = (unsigned)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096);
  if (f32)
    hipLaunchKernelGGL(k_noalloc_params<float>, dim3(grid), dim3(256), 0, c->s, q, nq, p, np, np_pad, (float*)qd,
                       (float*)pd);
  else
    hipLaunchKernelGGL(k_noalloc_params<double>, dim3(grid), dim3(256), 0, c->s, q, nq, p, np, np_pad,
                       (double*)nullptr, (double*)pd);
  e = hipGetLastError();
  if (e == hipSuccess)
    e = f32 ? launch_ll_mm(c->K, g, (const float*)qd, (const float*)pd, nullptr, nullptr, part, stage, c->s)
            : launch_qside(c->K, c->prec, MODE_LL, g, (const void*)q, pd, part, c->s);
  if (e == hipSuccess) e = launch_reduce_ll(part, g.qchunks, g, rows_d, ll, c->s);
  if (e != hipSuccess) return nfail(c, HK_EHIP, "hk_loglik_noalloc: launch failed: %s", hipGetErrorString(e));
  return HK_OK;
}
