// Launcher of the discrete adjoint Rosenbrock integrator, included by its own
// generated translation unit (kpphip_<root>_adj.hip) after the mechanism's
// <root>_hip.h.
#pragma once
#include <hip/hip_runtime.h>

#include "adj_device.h"
#include "mech_ops.h"

namespace kpphip {
// built for mechanisms compiled with the adjoint's model functions whose LU
// fits a plain tile of modest size
template <class M>
constexpr bool has_adj() {
  if constexpr (M::ADJOINT) return M::LU_NONZERO <= 4096;
  return false;
}
// the fixed rows of a lane's work tile (the adjoint vectors' rows and the
// checkpoints are sized by the host)
template <class M>
int adj_rows_mech(int) {
  if constexpr (has_adj<M>()) return AdjLayout<M>::FIXED;
  return 0;
}
template <class M>
hipError_t adj_launch_mech(const RosParams& p, int nblocks, hipStream_t st) {
  if constexpr (has_adj<M>()) {
    hipLaunchKernelGGL((adj_plain_kernel<M>), dim3(nblocks), dim3(TILE), 0, st, p);
    return hipGetLastError();
  }
  return hipErrorInvalidValue;
}
}  // namespace kpphip

// Defines kpphip_mech_adj_<NAME>(): the discrete adjoint integrator of NAME.
#define KPPHIP_DEFINE_MECH_ADJ(NAME)                                                    \
  extern "C" const kpphip::SdirkOps* kpphip_mech_adj_##NAME() {                         \
    static const kpphip::SdirkOps ops = {kpphip::adj_rows_mech<kppm_##NAME>,            \
                                         kpphip::adj_launch_mech<kppm_##NAME>};         \
    return &ops;                                                                        \
  }
