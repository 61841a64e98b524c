/* abi_driver.c — a plain C caller of the engine's ABI (include/haplo_hip.h) with no Python in the process,
 * in the role of the reference's runner layer (src/cuda/runners.jl: create the device data, evaluate the
 * passes, run the fit).  Reads a BED payload and starting parameters from DIR, writes every result back to
 * DIR as raw little-endian doubles.
 *
 *   abi_driver DIR I J K PRECISION
 *     in:  DIR/bed.bin (2J SNPs x ceil(I/4) bytes), DIR/q.bin (K x I), DIR/p.bin (K x 4J)
 *     out: DIR/ll.bin (ll, then the I rows), DIR/uq.bin (XtX_q, Xtz_q), DIR/up.bin (XtX_p, Xtz_p),
 *          DIR/em.bin (q_next, p_next), DIR/fit.bin (ll, ll_init, iters, converged, qp_failures, then the
 *          log-likelihood trace of the iterations), DIR/qp.bin (q, p after the fit)
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "haplo_hip.h"

static void* slurp(const char* dir, const char* name, size_t want) {
  char path[4096];
  snprintf(path, sizeof path, "%s/%s", dir, name);
  FILE* f = fopen(path, "rb");
  if (!f) return NULL;
  void* buf = malloc(want ? want : 1);
  const size_t got = fread(buf, 1, want, f);
  const int extra = fgetc(f) != EOF;
  fclose(f);
  if (got != want || extra) {
    fprintf(stderr, "%s: expected %zu bytes\n", path, want);
    free(buf);
    return NULL;
  }
  return buf;
}

static int dump(const char* dir, const char* name, const double* a, size_t na, const double* b, size_t nb) {
  char path[4096];
  snprintf(path, sizeof path, "%s/%s", dir, name);
  FILE* f = fopen(path, "wb");
  if (!f) return 1;
  const int bad = fwrite(a, sizeof(double), na, f) != na || (nb && fwrite(b, sizeof(double), nb, f) != nb);
  return fclose(f) != 0 || bad;
}

int main(int argc, char** argv) {
  if (argc != 6) {
    fprintf(stderr, "usage: %s DIR I J K PRECISION\n", argv[0]);
    return 2;
  }
  const char* dir = argv[1];
  const int64_t I = atoll(argv[2]), J = atoll(argv[3]);
  const int K = atoi(argv[4]), prec = atoi(argv[5]);
  if (I < 1 || J < 1 || K < 1 || K > HK_MAX_K) {
    fprintf(stderr, "bad sizes\n");
    return 2;
  }
  const int64_t bps = (I + 3) / 4, nq = (int64_t)K * I, np = (int64_t)K * 4 * J;
  uint8_t* bed = slurp(dir, "bed.bin", (size_t)(2 * J * bps));
  double* q = slurp(dir, "q.bin", sizeof(double) * (size_t)nq);
  double* p = slurp(dir, "p.bin", sizeof(double) * (size_t)np);
  if (!bed || !q || !p) return 2;

  hk_ctx* ctx = NULL;
  int rc = hk_ctx_create(&ctx, I, J, K, bed, bps, 2 * J, prec);
  if (rc != HK_OK) {
    fprintf(stderr, "hk_ctx_create: %d %s\n", rc, hk_create_error());
    return 1;
  }
#define CK(call)                                                              \
  do {                                                                        \
    const int rc_ = (call);                                                   \
    if (rc_ != HK_OK) {                                                       \
      fprintf(stderr, "%s: %d %s\n", #call, rc_, hk_last_error(ctx));         \
      hk_ctx_destroy(ctx);                                                    \
      return 1;                                                               \
    }                                                                         \
  } while (0)

  CK(hk_set_params(ctx, q, p));
  double* rows = malloc(sizeof(double) * (size_t)(I + 1));
  CK(hk_loglik(ctx, &rows[0], rows + 1));
  double* XtXq = malloc(sizeof(double) * (size_t)(nq * K));
  double* Xtzq = malloc(sizeof(double) * (size_t)nq);
  double* XtXp = malloc(sizeof(double) * (size_t)(np * K));
  double* Xtzp = malloc(sizeof(double) * (size_t)np);
  double* qn = malloc(sizeof(double) * (size_t)nq);
  double* pn = malloc(sizeof(double) * (size_t)np);
  CK(hk_update_q(ctx, XtXq, Xtzq));
  CK(hk_update_p(ctx, XtXp, Xtzp));
  CK(hk_em(ctx, 0, qn, pn));

  hk_fit_opts o;
  hk_fit_opts_default(&o);
  hk_fit_result r;
  double* fit = calloc((size_t)(5 + o.n_iter), sizeof(double));
  CK(hk_fit(ctx, &o, NAN, fit + 5, &r));
  fit[0] = r.ll;
  fit[1] = r.ll_init;
  fit[2] = r.iters;
  fit[3] = r.converged;
  fit[4] = r.qp_failures;
  CK(hk_get_params(ctx, q, p));
  CK(hk_ctx_destroy(ctx));
  ctx = NULL;

  int bad = dump(dir, "ll.bin", rows, (size_t)(I + 1), NULL, 0);
  bad |= dump(dir, "uq.bin", XtXq, (size_t)(nq * K), Xtzq, (size_t)nq);
  bad |= dump(dir, "up.bin", XtXp, (size_t)(np * K), Xtzp, (size_t)np);
  bad |= dump(dir, "em.bin", qn, (size_t)nq, pn, (size_t)np);
  bad |= dump(dir, "fit.bin", fit, (size_t)(5 + r.iters), NULL, 0);
  bad |= dump(dir, "qp.bin", q, (size_t)nq, p, (size_t)np);
  printf("ll %.17g iters %d converged %d\n", r.ll, r.iters, r.converged);
  free(bed); free(q); free(p); free(rows); free(XtXq); free(Xtzq); free(XtXp); free(Xtzp); free(qn); free(pn);
  free(fit);
  return bad ? 1 : 0;
}
