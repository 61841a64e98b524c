// hk_internal.h — declarations shared between the kernel translation units and the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef struct hk_mlp hk_mlp;

namespace hk {

enum { MODE_LL = 0, MODE_GRAD = 1, MODE_CURV = 2 };
enum { PREC_FP64 = 0, PREC_MIXED = 1 };

// Geometry of one fit's genotype passes.
struct PassGeom {
  int64_t I, J;    // individuals, SNP blocks
  int64_t Ip, Jp;  // padded: Ip multiple of 256, Jp multiple of 64
  int64_t J8, I8;  // genotype words: Jp / 8, Ip / 8
  const uint32_t* gq;  // [J8][Ip]
  const uint32_t* gp;  // [I8][Jp]
  int qchunks, qwpc;   // Q side: number of j chunks, words per chunk
  int pchunks, pwpc;   // P side: number of i chunks, words per chunk
  int flush_curv;      // MIXED: genotype words per FP32 partial of the curvature passes (then FP64)
  int flush_grad;      // MIXED: genotype words per FP32 partial of the gradient (EM) passes
  int acc32;           // reference Float32 semantics: no FP64 flushes, chunk partials summed in FP32
  void* qq;            // staging table of the P-side curvature pass (TP; rows per individual)
  void* pp;            // staging table of the Q-side curvature pass (TP; rows per block)
  const int* gate;     // device gate of the fit's iteration kernels (nullptr: always run)
  // Curvature-pass variant per side (FP32 passes, K <= 8): 0 = the default for K, 1 = the other staging form
  // where one compiles without spills (K <= 6: without the register prefetch; Q side at K = 7: with it).  Both
  // forms do the same arithmetic in the same order; the autotuner times them.
  int qshape = 0, pshape = 0;
};
// Whether the FP32 curvature pass of side (0 = Q, 1 = P) has a second variant at K.
inline bool has_alt_shape(int K, int prec, int side) {
  return prec != 0 && (K <= 6 || (K == 7 && side == 0));
}

int nacc(int K, int mode);
// Workgroups of 256 for a loop over n elements, capped at 2^22 (2^30 work-items): a dispatch packet holds the grid
// size in work-items as a 32-bit count, so element loops that can pass 2^32 stride over a capped grid instead.
inline unsigned grid_1d(int64_t n) {
  const int64_t nb = (n + 255) / 256;
  return (unsigned)(nb < 1 ? 1 : (nb < (1 << 22) ? nb : (1 << 22)));
}
// Element counts of the curvature passes' staging tables (P side: nq, Q side: np).
void stage_sizes(int K, int64_t Ip, int64_t Jp, int64_t* nq, int64_t* np);

hipError_t launch_repack(const uint8_t* raw, int64_t bps, const PassGeom& g, hipStream_t s);
hipError_t launch_count_dm(const PassGeom& g, int64_t* dm, hipStream_t s);
// Adds the double-missing pairs of nb blocks held as 2 nb raw BED columns (bps bytes each) to dm.
hipError_t launch_count_dm_raw(const uint8_t* raw, int64_t bps, int64_t I, int64_t nb, int64_t* dm, hipStream_t s);
hipError_t launch_decode(const PassGeom& g, uint8_t* codes, hipStream_t s);
// MODE_CURV passes read staging tables built from the same (q, p) by launch_curv_tables (side 0: the
// Q-side pass, side 1: the P-side pass), which must be enqueued first.
hipError_t launch_curv_tables(int K, int prec, int side, const PassGeom& g, const void* q, const void* p,
                              hipStream_t s);
hipError_t launch_qside(int K, int prec, int mode, const PassGeom& g, const void* q, const void* p, double* part,
                        hipStream_t s);
// Log-likelihood pass of the MIXED engine on the matrix cores: one parameter set (q2 = p2 = nullptr) or two in
// one genotype sweep; per-chunk row partials of set s at part + s * qchunks * Ip (launch_reduce_ll on each).
// stage: ll_stage_bytes(K, g.J8) per set (16-byte aligned), the BF16 terms of p that the pass reads.
hipError_t launch_ll_mm(int K, const PassGeom& g, const float* q1, const float* p1, const float* q2, const float* p2,
                        double* part, void* stage, hipStream_t s);
size_t ll_stage_bytes(int K, int64_t J8);
hipError_t launch_pside(int K, int prec, int mode, const PassGeom& g, const void* q, const void* p, double* out,
                        hipStream_t s);
// 9 <= K <= 16 (hk_wide.hip): curvature passes over windows of packed columns, the P-side EM pass on the
// matrix cores, their staging tables and the P-side finishing step (the launchers above route to them).
hipError_t launch_curv_tables_wide(int K, int prec, int side, const PassGeom& g, const void* x, hipStream_t s);
hipError_t launch_qside_curv_wide(int K, int prec, const PassGeom& g, const void* q, double* part, hipStream_t s);
hipError_t launch_pside_wide(int K, int prec, int mode, const PassGeom& g, const void* q, const void* p,
                             double* out, hipStream_t s);
void stage_sizes_wide(int K, int64_t Ip, int64_t Jp, int64_t* nq, int64_t* np);
hipError_t launch_reduce_q(const double* part, int nchunk, int NT, const PassGeom& g, double* res, hipStream_t s);
hipError_t launch_reduce_q_curv(const double* part, int nchunk, int K, const PassGeom& g, const double* q, double* res,
                               hipStream_t s);
// Gradient slots of the P-side curvature pass; with flags (J bytes) also the early list of the blocks the
// P row step must solve densely (dense: J + 1 ints, [count, blocks]), reset here.
hipError_t launch_finish_p(int K, const PassGeom& g, const double* p, double* res, unsigned char* flags, int* dense,
                           hipStream_t s);
hipError_t launch_reduce_p(const double* part, int nchunk, int NT, const PassGeom& g, double* res, hipStream_t s);
hipError_t launch_reduce_ll(const double* part, int nchunk, const PassGeom& g, double* rows, double* total,
                            hipStream_t s);
hipError_t launch_expand(const double* res, int K, int NT, int64_t n, double* XtX, double* Xtz, hipStream_t s);
hipError_t launch_em_expand(const double* res, int K, int NT, int64_t n, const double* par, double* out,
                            hipStream_t s);

// Row steps, projections, EM normalisation, ZAL (hk_step.hip).
struct StepOpts {
  int penalty;  // Q: H + rho 11' without equality; P: rho on the (l, l2) block diagonals
  double rho;
  double eps;   // pseudocount of the simplex projections
  const int* gate = nullptr;  // device gate (PassGeom::gate)
};
// dq / dp (optional): the raw QP steps before projection, for the step-halving line search.
hipError_t launch_q_step(int K, int64_t I, const double* res_q, const double* q, double* q_next, const StepOpts& o,
                         int* fails, hipStream_t s, double* dq = nullptr);
// Device lists of the P row step.  dense / flags: the blocks launch_finish_p marked, solved by the dense
// one-wave solver in the same launch as the four-lane QPs; late: J + 1 ints for blocks the four-lane
// kernel hands over itself.
struct PStepLists {
  const int* dense;
  const unsigned char* flags;
  int* late;
};
// Occupancy caps of the P-step kernels for K, queried once per process at context set-up (outside any capture);
// launch_p_step returns hipErrorNotReady for a K that was never primed.
hipError_t prime_p_step_caps(int K);
hipError_t launch_p_step(int K, int64_t J, const double* res_p, const double* p, double* p_next, const StepOpts& o,
                         int* fails, const PStepLists& lists, hipStream_t s, double* dp = nullptr);
// 9 <= K <= 16 (hk_step_wide.hip): the four-lane kernels of those K, one quad per slot, the first `ndb`
// workgroups on the early dense list; the late list after.  p_step_cap_wide: resident workgroups (set-up).
unsigned p_step_cap_wide(int K, bool penalty);
hipError_t launch_p_step_wide(int K, int64_t J, const double* res_p, const double* p, double* p_next, const StepOpts& o,
                              int* fails, const PStepLists& lists, int ndb, hipStream_t s, double* dp);
hipError_t launch_project_q(int K, int64_t I, double* q, double eps, hipStream_t s, const int* gate = nullptr);
hipError_t launch_project_p(int K, int64_t J, double* p, double eps, hipStream_t s, const int* gate = nullptr);
hipError_t launch_em_q(int K, int64_t I, int64_t J, const double* res_q, int NT, const int64_t* dm, const double* q,
                       double* q_next, double eps, hipStream_t s);
hipError_t launch_em_p(int K, int64_t J, const double* res_p, int NT, const double* p, double* p_next, double eps,
                       hipStream_t s);
hipError_t launch_pair_to_f32(const double* q, int64_t nq, const double* p, int64_t np, float* dst, hipStream_t s,
                              const int* gate = nullptr);
// Quasi-Newton extrapolation of admixture_qn! (mode 0: ZAL, 1: LBQN); dots_part holds 512 x 64 doubles.
hipError_t launch_zal(int64_t N, int Qn, const int* d_iter, double* U, double* V, const double* x0, const double* x1,
                      const double* x2, double* x_out, double* dots_part, double* coef, hipStream_t s, int mode,
                      const int* gate = nullptr);
hipError_t launch_accept(int64_t N, double* ll, const double* x_basic, const double* x_qn, double* x, int* d_iter,
                         hipStream_t s, const int* gate = nullptr);
hipError_t launch_copy_f64(int64_t n, const double* src, double* dst, hipStream_t s, const int* gate = nullptr);
// Device-side stopping rule of a fit run as part of a batch graph (admixture_qn!'s reldiff < rtol or
// n_iter iterations, algorithms_outer.jl:161-165), evaluated after the accept step of every iteration.
struct FitCtl {
  double ll_prev;  // ll_new of the previous iteration (before the first: the QN loop's start value)
  double rtol;
  int n_iter;
  int iter;        // QN iterations run
  int converged;   // 1 once the relative change fell below rtol
  int status;      // 0, or 1 when the log-likelihood became non-finite
};
// Appends ll[2] (ll_new) to trace, updates ctl and clears *gate when the fit stops.
hipError_t launch_fit_control(const double* ll, FitCtl* ctl, double* trace, int* gate, hipStream_t s);

hipError_t launch_set_int(int* p, int v, hipStream_t s, const int* gate = nullptr);
hipError_t launch_axpy_step(int64_t n, const double* a, const double* d, double f, double* out, hipStream_t s);
hipError_t launch_fill(double* p, int64_t n, double v, hipStream_t s);
hipError_t launch_copy16(const void* src, void* dst, int64_t n16, hipStream_t s);

// Random-projection features (hk_mlp.hip): F_host (I x 256) from the first n_blocks blocks of gq.
int project_features(const PassGeom& g, int n_blocks, uint64_t seed, float* F_host, hipStream_t s, char* err,
                     size_t errlen);
// The same into a device buffer, with reusable scratch.
struct ProjBuffers;
hipError_t project_features_into(const PassGeom& g, int n_blocks, uint64_t seed, float* F_dev, ProjBuffers& b,
                                 hipStream_t s);
// Network prediction from a context's features computed on the device (hk_mlp_predict).
int mlp_predict_from(hk_mlp* m, const PassGeom& g, int n_blocks, uint64_t seed, float* P);

// SKFR marker pre-selection (hk_skfr.hip) on the genotype streams of g; returns 0, -1 (bad argument), -2 (HIP
// error) or -3 (out of device memory) with err filled.
// init_labels (I labels, or NULL): the first try starts from these cluster labels (a path step).
int skfr_run(const PassGeom& g, int k, int sparsity, int local, int max_iter, int tries, uint64_t seed,
             const int32_t* init, const int32_t* init_labels, int32_t* clusters, int32_t* aims, int* n_aims,
             double* wss, int* iters, double* prof, hipStream_t s, char* err, size_t errlen);

}  // namespace hk

// Context of one fit (the C ABI's hk_ctx): genotype streams, parameters, pass buffers and the stream.
struct hk_ctx {
  int64_t I = 0, J = 0;
  int K = 0, prec = 0, dev = 0;
  hk::PassGeom g{};
  hipStream_t s = nullptr;
  uint32_t* gq = nullptr;
  uint32_t* gp = nullptr;
  void* qq = nullptr;
  void* pp = nullptr;
  int64_t* dm = nullptr;
  double* xb[4] = {nullptr, nullptr, nullptr, nullptr};  // parameter buffers [q | p | pad]
  int64_t xlen = 0;                                      // K * (I + 4 Jp)
  float* x32 = nullptr;                                  // FP32 mirror of one parameter buffer
  float* x32b = nullptr;                                 // second mirror (two-set log-likelihood sweep)
  void* ll_stage = nullptr;  // BF16 terms of p for the MIXED log-likelihood pass (two sets)
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
  int* redo = nullptr;       // J + 1 ints: early list of the blocks for the dense P-step solver
  int* redo_late = nullptr;  // J + 1 ints: blocks the four-lane P-step kernel hands over itself
  unsigned char* pflag = nullptr;  // J: block on the early list
  double* full_a = nullptr;  // ABI expansion buffers
  double* full_b = nullptr;
  int64_t full_len = 0;
  // Runner-level device arrays of CuAdmixData (cuda/structs.jl:1-15) beyond x / x_next: XtX_q, Xtz_q,
  // XtX_p, Xtz_p in the reference's full layouts, allocated on first use.
  double* named[4] = {nullptr, nullptr, nullptr, nullptr};
  int qchunks_req = 0, pchunks_req = 0;
  // fit state for the sync-free iteration: ll[0] ll_basic, ll[1] ll_qn, ll[2] ll_new, ll[3] scratch
  double* d_ll = nullptr;
  int* d_iter = nullptr;
  double* h_ll = nullptr;  // pinned mirror of d_ll
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  unsigned long long graph_key = 0;
  bool owns_geno = true;
  int64_t part_q_len = 0, part_p_len = 0;
  char err[512] = {0};
};
