// hk_step_wide.hip — update_p! row step for 9 <= K <= 16 (algorithms_inner.jl:259-312 assembly and QP,
// :344-351 step and projection).
//
// The same one-launch scheme as K <= 8 (hk_step.hip): the four-lane block-structured solver (hk_pstep4.h) on
// every block the P-side finishing kernel did not list, the dense one-wave solver on the listed blocks (a free
// haplotype block with a vanishing Cholesky pivot) in the first workgroups of the same launch, and the late
// list after it.  From K = 9 on, a lane's K x K block curvature, factor and inverse exceed the register file:
// the four-lane kernels run one wave per SIMD and spill to scratch.  They still do about a sixteenth of the
// dense solver's arithmetic per QP (four K x K factorisations against one 4K x 4K) and hold sixteen QPs per
// wavefront instead of one.  One quad per slot (the grid covers every block): the striding form that K <= 8
// uses for large J spills about twice as many registers here and measured 7-19 % slower.
#include "hk_device.h"
#include "hk_internal.h"
#include "hk_pstep4.h"

namespace hk {

template <int K>
static unsigned p_cap_k(bool pen) {  // resident workgroups (primed at set-up like K <= 8; the launch needs no cap)
  return pen ? resident_blocks(k_p_step4<K, true, false>, 256) : resident_blocks(k_p_step4<K, false, false>, 256);
}

unsigned p_step_cap_wide(int K, bool pen) {
  switch (K) {
    case 9: return p_cap_k<9>(pen);
    case 10: return p_cap_k<10>(pen);
    case 11: return p_cap_k<11>(pen);
    case 12: return p_cap_k<12>(pen);
    case 13: return p_cap_k<13>(pen);
    case 14: return p_cap_k<14>(pen);
    case 15: return p_cap_k<15>(pen);
    case 16: return p_cap_k<16>(pen);
    default: return 0;
  }
}

template <int K, bool PEN>
static void p_step_wide_k(int64_t J, const double* res_p, const double* p, double* p_next, const StepOpts& o,
                          int* fails, const PStepLists& lists, int ndb, hipStream_t s, double* dp) {
  const unsigned nq = (unsigned)((4 * J + 255) / 256), nd = (unsigned)(J < 1024 ? J : 1024);
  hipLaunchKernelGGL((k_p_step4<K, PEN, false>), dim3(ndb + nq), dim3(256), 0, s, J, res_p, p, p_next, o.rho, o.eps,
                     dp, lists.flags, lists.dense, ndb, fails, lists.late, o.gate);
  hipLaunchKernelGGL((k_p_step_list<K, PEN>), dim3(nd), dim3(64), 0, s, res_p, p, p_next, o.rho, o.eps, fails, dp,
                     lists.late, o.gate);
}

template <int K>
static void p_step_wide_pen(int64_t J, const double* res_p, const double* p, double* p_next, const StepOpts& o,
                            int* fails, const PStepLists& lists, int ndb, hipStream_t s, double* dp) {
  if (o.penalty) p_step_wide_k<K, true>(J, res_p, p, p_next, o, fails, lists, ndb, s, dp);
  else p_step_wide_k<K, false>(J, res_p, p, p_next, o, fails, lists, ndb, s, dp);
}

hipError_t launch_p_step_wide(int K, int64_t J, const double* res_p, const double* p, double* p_next, const StepOpts& o,
                              int* fails, const PStepLists& lists, int ndb, hipStream_t s, double* dp) {
  switch (K) {
    case 9: p_step_wide_pen<9>(J, res_p, p, p_next, o, fails, lists, ndb, s, dp); break;
    case 10: p_step_wide_pen<10>(J, res_p, p, p_next, o, fails, lists, ndb, s, dp); break;
    case 11: p_step_wide_pen<11>(J, res_p, p, p_next, o, fails, lists, ndb, s, dp); break;
    case 12: p_step_wide_pen<12>(J, res_p, p, p_next, o, fails, lists, ndb, s, dp); break;
    case 13: p_step_wide_pen<13>(J, res_p, p, p_next, o, fails, lists, ndb, s, dp); break;
    case 14: p_step_wide_pen<14>(J, res_p, p, p_next, o, fails, lists, ndb, s, dp); break;
    case 15: p_step_wide_pen<15>(J, res_p, p, p_next, o, fails, lists, ndb, s, dp); break;
    case 16: p_step_wide_pen<16>(J, res_p, p, p_next, o, fails, lists, ndb, s, dp); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace hk
