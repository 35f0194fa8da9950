// lbm_kernels_nt.hip — the 3D step kernels with non-temporal population stores, for lattices
// whose buffers cannot stay in the caches between steps (a translation unit of its own so that it
// compiles in parallel with the ordinary instantiations).
#include "lbm_step.hpp"

namespace lbm {

hipError_t launch_step_scalar_nt(const Config &c, const StepParams &p, hipStream_t s) {
  return with_full(c, [&](auto Qc, auto STc, auto BCc, auto COc) -> hipError_t {
    constexpr int Q = decltype(Qc)::value;
    if constexpr (Q == 9) {
      return hipErrorInvalidValue;  // D2Q9 lattices use ordinary stores
    } else {
      hipLaunchKernelGGL((k_step<Q, decltype(STc)::value, decltype(BCc)::value, decltype(COc)::value, true>),
                         dim3(p.ntiles), dim3(BLOCK_THREADS), 0, s, p);
      return hipGetLastError();
    }
  });
}

hipError_t launch_step_batch_scalar_nt(const Config &c, const BatchArgs &b, int grid, hipStream_t s) {
  return with_full(c, [&](auto Qc, auto STc, auto BCc, auto COc) -> hipError_t {
    constexpr int Q = decltype(Qc)::value, ST = decltype(STc)::value, BC = decltype(BCc)::value,
                  CO = decltype(COc)::value;
    if constexpr (Q == 9) {
      return hipErrorInvalidValue;  // D2Q9 groups use ordinary stores
    } else {
      if (b.one_tile)
        hipLaunchKernelGGL((k_step_batch<Q, ST, BC, CO, true, true>), dim3(grid), dim3(BLOCK_THREADS), 0, s, b);
      else
        hipLaunchKernelGGL((k_step_batch<Q, ST, BC, CO, false, true>), dim3(grid), dim3(BLOCK_THREADS), 0, s, b);
      return hipGetLastError();
    }
  });
}

}  // namespace lbm
