// Work queue of the implicit integrators' persistent kernels (SDIRK,
// Runge-Kutta), included by their device headers.
#pragma once
#include "ros_device.h"

namespace kpphip {

// The work queue of the implicit integrators' persistent kernels (SDIRK,
// Runge-Kutta): a lane whose cell has finished takes the next cell from the
// queue (p.queue, zero at launch); the wavefront leaves when the queue is
// exhausted and no lane is active.  Cells are independent, so each cell's
// results are those of the one-cell-per-lane kernel.
template <class CellT, class StateT>
__device__ __forceinline__ void refill_loop(const RosParams& p, const CellT& c) {
  const int lane = threadIdx.x;
  StateT s;
  const bool worker = lane < p.lanes;
  bool active = false;
  bool exhausted = false;  // wavefront-uniform
  for (;;) {
    const uint64_t idle = __ballot(worker && !active);
    const int n_idle = __popcll(idle);
    if (!exhausted && n_idle > 0) {
      const int first = __ffsll((long long)idle) - 1;
      int base = 0;
      if (lane == first) base = atomicAdd(p.queue, n_idle);
      base = __shfl(base, first);
      if (base + n_idle >= p.ncells) exhausted = true;
      if (worker && !active) {
        const int k = base + __popcll(idle & ((1ull << lane) - 1ull));
        if (k < p.ncells) {
          c.start(p, p.order ? p.order[k] : k, s);
          active = true;
        }
      }
    }
    if (__ballot(active) == 0) break;
    if (active && c.attempt(p, s) != 0) active = false;
  }
}

}  // namespace kpphip
