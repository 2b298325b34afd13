// Launchers of the implicit integrators (SDIRK, Runge-Kutta), included by
// their own generated translation units (kpphip_<root>_sdirk.hip,
// kpphip_<root>_rk.hip) after the mechanism's <root>_hip.h, so that the
// Rosenbrock translation units do not depend on them.
#pragma once
#include <hip/hip_runtime.h>

#include "mech_ops.h"
#include "rk_device.h"
#include "sdirk_device.h"

namespace kpphip {
// SDIRK kernels, built for mechanisms whose LU fits a plain tile of modest
// size (not for the largest ones, whose model routines are split into parts)
template <class M>
constexpr bool has_sdirk() {
  return M::LU_NONZERO <= 4096;
}
template <class M>
int sdirk_rows_mech(int S) {
  if constexpr (has_sdirk<M>()) {
    switch (S) {
      case 2: return SdLayout<M, 2>::ROWS;
      case 3: return SdLayout<M, 3>::ROWS;
      case 5: return SdLayout<M, 5>::ROWS;
    }
  }
  return 0;
}
template <class M>
hipError_t sdirk_launch_mech(const RosParams& p, int nblocks, hipStream_t st) {
  if constexpr (has_sdirk<M>()) {
    switch (p.sd.S) {
      case 2:
        return launch_plain_or_queue(sdirk_plain_kernel<M, 2>, sdirk_persistent_kernel<M, 2>, p,
                                     nblocks, st);
      case 3:
        return launch_plain_or_queue(sdirk_plain_kernel<M, 3>, sdirk_persistent_kernel<M, 3>, p,
                                     nblocks, st);
      case 5:
        return launch_plain_or_queue(sdirk_plain_kernel<M, 5>, sdirk_persistent_kernel<M, 5>, p,
                                     nblocks, st);
      default: return hipErrorInvalidValue;
    }
  }
  return hipErrorInvalidValue;
}
}  // namespace kpphip

namespace kpphip {
// Runge-Kutta kernels, built where SDIRK's are
template <class M>
int rk_rows_mech(int) {
  if constexpr (has_sdirk<M>()) return RkLayout<M>::ROWS;
  return 0;
}
template <class M>
hipError_t rk_launch_mech(const RosParams& p, int nblocks, hipStream_t st) {
  if constexpr (has_sdirk<M>())
    return launch_plain_or_queue(rk_plain_kernel<M>, rk_persistent_kernel<M>, p, nblocks, st);
  return hipErrorInvalidValue;
}
}  // namespace kpphip

// Defines kpphip_mech_rk_<NAME>(): the Runge-Kutta integrator of NAME.
#define KPPHIP_DEFINE_MECH_RK(NAME)                                                     \
  extern "C" const kpphip::SdirkOps* kpphip_mech_rk_##NAME() {                          \
    static const kpphip::SdirkOps ops = {kpphip::rk_rows_mech<kppm_##NAME>,             \
                                         kpphip::rk_launch_mech<kppm_##NAME>};          \
    return &ops;                                                                        \
  }

// Defines kpphip_mech_sdirk_<NAME>(): the SDIRK integrator of NAME.
#define KPPHIP_DEFINE_MECH_SDIRK(NAME)                                                  \
  extern "C" const kpphip::SdirkOps* kpphip_mech_sdirk_##NAME() {                       \
    static const kpphip::SdirkOps ops = {kpphip::sdirk_rows_mech<kppm_##NAME>,          \
                                         kpphip::sdirk_launch_mech<kppm_##NAME>};       \
    return &ops;                                                                        \
  }

