// Work queue of the persistent kernels of the SDIRK and Runge-Kutta
// integrators, and their launch, included by their device headers.
#pragma once
#include <hip/hip_runtime.h>

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

// the resident grid of a persistent kernel: as many one-wavefront blocks as
// the device holds at once, at most one per 64 cells
template <class K>
int resident_blocks(K kernel, int nblocks) {
  int occ = 0, dev = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, TILE, 0) != hipSuccess ||
      hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return nblocks;
  const int nb = (occ > 0 ? occ : 1) * (cus > 0 ? cus : 1);
  return nb < nblocks ? nb : nblocks;
}
// one-cell-per-lane grid, or (p.queue set, zero on the stream) the
// persistent one
template <class K1, class K2>
hipError_t launch_plain_or_queue(K1 plain, K2 persistent, const RosParams& p, int nblocks,
                                 hipStream_t st) {
  if (p.queue)
    hipLaunchKernelGGL(persistent, dim3(resident_blocks(persistent, nblocks)), dim3(TILE), 0, st,
                       p);
  else
    hipLaunchKernelGGL(plain, dim3(nblocks), dim3(TILE), 0, st, p);
  return hipGetLastError();
}
}  // namespace kpphip
