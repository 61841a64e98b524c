// hk_step_wide.hip — update_p! row step for 9 <= K <= 16 (algorithms_inner.jl:259-312 assembly and QP,
// :344-351 step and projection).
//
// The four-lane P solver (hk_step.hip) keeps a haplotype block's K x K curvature and its Cholesky factor in one
// lane's registers, which stops at K = 8.  Here every block's QP (4K <= 64 variables) goes to the dense one-wave
// active-set solver, the same one the K <= 8 path uses for blocks with a singular haplotype curvature: lane r
// owns row r of the curvature, so the QP is solved with the same rule, ratio test and tie-breaks at any K.
// The penalty form stays in registers (accumulation registers take the overflow); the equality form's Schur
// complement over K groups spills to scratch at K >= 12.
#include "hk_device.h"
#include "hk_internal.h"
#include "hk_qp.h"

namespace hk {

// One wavefront per block, grid-stride over the J blocks.
template <int K, bool PEN>
__global__ __launch_bounds__(64) void k_p_step_all(int64_t J, const double* __restrict__ res,
                                                   const double* __restrict__ p, double* __restrict__ p_next,
                                                   double rho, double eps, int* __restrict__ fails,
                                                   double* __restrict__ dp, const int* __restrict__ gate) {
  if (gate_off(gate)) return;
  __shared__ double lds[4 * K * 4 * K];
  for (int64_t j = blockIdx.x; j < J; j += gridDim.x)
    p_step_dense<K, PEN>(j, res, p, p_next, rho, eps, fails, dp, lds);
}

template <int K>
static void p_step_all_k(int64_t J, const double* res_p, const double* p, double* p_next, const StepOpts& o,
                         int* fails, hipStream_t s, double* dp) {
  const unsigned grid = (unsigned)(J < 8192 ? J : 8192);
  if (o.penalty)
    hipLaunchKernelGGL((k_p_step_all<K, true>), dim3(grid), dim3(64), 0, s, J, res_p, p, p_next, o.rho, o.eps, fails,
                       dp, o.gate);
  else
    hipLaunchKernelGGL((k_p_step_all<K, false>), dim3(grid), dim3(64), 0, s, J, res_p, p, p_next, o.rho, o.eps,
                       fails, dp, o.gate);
}

hipError_t launch_p_step_wide(int K, int64_t J, const double* res_p, const double* p, double* p_next, const StepOpts& o,
                              int* fails, hipStream_t s, double* dp) {
  if (J < 1) return hipSuccess;
  switch (K) {
    case 9: p_step_all_k<9>(J, res_p, p, p_next, o, fails, s, dp); break;
    case 10: p_step_all_k<10>(J, res_p, p, p_next, o, fails, s, dp); break;
    case 11: p_step_all_k<11>(J, res_p, p, p_next, o, fails, s, dp); break;
    case 12: p_step_all_k<12>(J, res_p, p, p_next, o, fails, s, dp); break;
    case 13: p_step_all_k<13>(J, res_p, p, p_next, o, fails, s, dp); break;
    case 14: p_step_all_k<14>(J, res_p, p, p_next, o, fails, s, dp); break;
    case 15: p_step_all_k<15>(J, res_p, p, p_next, o, fails, s, dp); break;
    case 16: p_step_all_k<16>(J, res_p, p, p_next, o, fails, s, dp); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace hk
