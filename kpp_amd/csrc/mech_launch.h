// Per-mechanism kernels and launchers.  Included by one generated
// translation unit per mechanism after the mechanism's <root>_hip.h.
#pragma once
#include <hip/hip_runtime.h>

#include "mech_ops.h"
#include "ros_coop.h"
#include "ros_device.h"

namespace kpphip {

// Accessor over the caller's SoA arrays: element i of cell c at b[i*ld + c].
struct Strided {
  using value_type = double;
  double* __restrict__ b;
  size_t ld;
  __device__ __forceinline__ double& operator[](int i) const { return b[(size_t)i * ld]; }
};

// Batched model functions with the generated-model ABI of KPP
// (Fun/Jac_SP/KppDecomp/KppSolve/Update_RCONST; prototypes in
// int/rosenbrock.c:54-91), one cell per lane over SoA arrays.
template <class M>
__global__ __launch_bounds__(256) void k_fun(int n, size_t ld, const double* V, const double* F,
                                             const double* R, double* out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  Strided v{(double*)V + c, ld}, f{(double*)F + c, ld}, r{(double*)R + c, ld}, o{out + c, ld};
  M::Fun(v, f, r, o);
}
template <class M>
__global__ __launch_bounds__(256) void k_jac(int n, size_t ld, const double* V, const double* F,
                                             const double* R, double* out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  Strided v{(double*)V + c, ld}, f{(double*)F + c, ld}, r{(double*)R + c, ld}, o{out + c, ld};
  M::Jac_SP(v, f, r, o);
}
template <class M>
__global__ __launch_bounds__(256) void k_decomp(int n, size_t ld, double* JVS, int* ising) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  Strided g{JVS + c, ld};
  ising[c] = M::KppDecompChecked(g);
}
template <class M>
__global__ __launch_bounds__(256) void k_solve(int n, size_t ld, const double* JVS, double* X) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  Strided g{(double*)JVS + c, ld}, x{X + c, ld};
  M::KppSolve(g, x);
}
// Update_RCONST at (TEMP, CFACTOR, SUN, TIME) per cell.  NUMBER-type rates
// are constants inside Fun/Jac_SP and their slot is left as zero.
template <class M>
__global__ __launch_bounds__(256) void k_rates(int n, size_t ld, const double* temp,
                                               const double* cfactor, const double* sun,
                                               const double* time, const double* C, double* R) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  Strided r{R + c, ld};
  for (int i = 0; i < M::NREACT; i++) r[i] = 0.0;
  const double tm = time ? time[c] : 0.0;
  M::rates_fixed(temp[c], cfactor[c], r);
  M::rates_sun(temp[c], cfactor[c], sun[c], tm, r);
  if (M::RATES_USE_STATE && C)
    M::rates_state(temp[c], cfactor[c], sun[c], tm, Strided{const_cast<double*>(C) + c, ld},
                   Strided{const_cast<double*>(C) + (size_t)M::NVAR * ld + c, ld}, r);
}

static __global__ void k_zero_int(int* q) { *q = 0; }

// Persistent kernel with an LDS mirror of NL LU entries per lane (NL > 0:
// FP64 only, and only if the mechanism is built with a mirror that large);
// NL == 0 is the plain tile layout, FP64 or mixed.
template <class M, bool MIX, int NL>
constexpr bool has_lds_split() {
  return NL == 0 || (!MIX && NL <= M::LDS_MAX);
}
template <class M, int S, bool MIX, int NL>
hipError_t launch_persistent(const RosParams& p, int nblocks, hipStream_t st) {
  if constexpr (has_lds_split<M, MIX, NL>()) {
    hipLaunchKernelGGL((ros_persistent_kernel<M, S, MIX, NL>), dim3(nblocks), dim3(TILE),
                       (size_t)NL * TILE * sizeof(double), st, p);
    return hipSuccess;
  } else {
    return hipErrorInvalidValue;
  }
}
template <class M, int S, bool MIX, int NL>
int occupancy_persistent() {
  int n = 0;
  if constexpr (has_lds_split<M, MIX, NL>()) {
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, ros_persistent_kernel<M, S, MIX, NL>,
                                                     TILE, (size_t)NL * TILE * sizeof(double)) !=
        hipSuccess)
      n = 0;
  }
  return n;
}

// Cooperative kernel (one wavefront per cell): FP64 only, mechanisms built
// with cooperative programs.
template <class M, class C, int S, bool MIX>
constexpr bool has_coop() {
  return !MIX && C::ENABLED;
}
template <class M, class C, int S, bool MIX>
hipError_t launch_coop(const RosParams& p, int nblocks, hipStream_t st) {
  if constexpr (has_coop<M, C, S, MIX>()) {
    const size_t lds = CoopLds<M, C>::BYTES;
    auto k = ros_coop_kernel<M, C, S>;
    if (hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds) != hipSuccess)
      return hipErrorInvalidValue;
    hipLaunchKernelGGL(k, dim3(nblocks), dim3(TILE), lds, st, p);
    return hipSuccess;
  } else {
    return hipErrorInvalidValue;
  }
}
template <class M, class C, int S, bool MIX>
int occupancy_coop() {
  int n = 0;
  if constexpr (has_coop<M, C, S, MIX>()) {
    const size_t lds = CoopLds<M, C>::BYTES;
    auto k = ros_coop_kernel<M, C, S>;
    (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, TILE, lds) != hipSuccess) n = 0;
  }
  return n;
}

template <class M, class C, int S, bool MIX>
hipError_t launch_s(const RosParams& p, int variant, int nblocks, hipStream_t st) {
  if (variant == KPP_VARIANT_COOP) {
    if constexpr (has_coop<M, C, S, MIX>()) {
      hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, st, p.queue);
      if (p.rpre)
        hipLaunchKernelGGL(ros_rates_prepass<M>, dim3((p.ncells + 255) / 256), dim3(256), 0, st,
                           p);
      const hipError_t e = launch_coop<M, C, S, MIX>(p, nblocks, st);
      if (e != hipSuccess) return e;
    } else {
      return hipErrorInvalidValue;
    }
  } else if (variant == KPP_VARIANT_PERSISTENT) {
    if (p.lds_entries != 0 && p.lds_entries != KPP_LDS_SPLIT_A &&
        p.lds_entries != KPP_LDS_SPLIT_B)
      return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, st, p.queue);
    if (p.rpre)
      hipLaunchKernelGGL(ros_rates_prepass<M>, dim3((p.ncells + 255) / 256), dim3(256), 0, st, p);
    hipError_t e;
    if (p.lds_entries == KPP_LDS_SPLIT_A)
      e = launch_persistent<M, S, MIX, KPP_LDS_SPLIT_A>(p, nblocks, st);
    else if (p.lds_entries == KPP_LDS_SPLIT_B)
      e = launch_persistent<M, S, MIX, KPP_LDS_SPLIT_B>(p, nblocks, st);
    else
      e = launch_persistent<M, S, MIX, 0>(p, nblocks, st);
    if (e != hipSuccess) return e;
  } else if (variant == KPP_VARIANT_REGISTER) {
    if constexpr (fits_registers<M>() && !MIX)
      hipLaunchKernelGGL((ros_reg_kernel<M, S>), dim3(nblocks), dim3(TILE), 0, st, p);
    else
      return hipErrorInvalidValue;
  } else {
    hipLaunchKernelGGL((ros_plain_kernel<M, S, MIX>), dim3(nblocks), dim3(TILE), 0, st, p);
  }
  return hipGetLastError();
}

template <class M, class C, bool MIX>
hipError_t launch_mech(const RosParams& p, int variant, int nblocks, hipStream_t st) {
  switch (p.coef.S) {
    case 2: return launch_s<M, C, 2, MIX>(p, variant, nblocks, st);
    case 3: return launch_s<M, C, 3, MIX>(p, variant, nblocks, st);
    case 4: return launch_s<M, C, 4, MIX>(p, variant, nblocks, st);
    case 6: return launch_s<M, C, 6, MIX>(p, variant, nblocks, st);
  }
  return hipErrorInvalidValue;
}

// Resident wavefronts per CU of a kernel (its register/LDS limits).
template <class M, class C, int S, bool MIX>
int occupancy_s(int variant, int lds) {
  int n = 0;
  hipError_t e;
  if (variant == KPP_VARIANT_COOP) {
    return occupancy_coop<M, C, S, MIX>();
  } else if (variant == KPP_VARIANT_PERSISTENT) {
    if (lds == KPP_LDS_SPLIT_A) return occupancy_persistent<M, S, MIX, KPP_LDS_SPLIT_A>();
    if (lds == KPP_LDS_SPLIT_B) return occupancy_persistent<M, S, MIX, KPP_LDS_SPLIT_B>();
    if (lds != 0) return 0;
    return occupancy_persistent<M, S, MIX, 0>();
  } else if (variant == KPP_VARIANT_REGISTER) {
    if constexpr (fits_registers<M>() && !MIX)
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, ros_reg_kernel<M, S>, TILE, 0);
    else
      return 0;
  } else {
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, ros_plain_kernel<M, S, MIX>, TILE, 0);
  }
  return e == hipSuccess ? n : 0;
}
template <class M, class C, bool MIX>
int occupancy_mech(int variant, int S, int lds) {
  switch (S) {
    case 2: return occupancy_s<M, C, 2, MIX>(variant, lds);
    case 3: return occupancy_s<M, C, 3, MIX>(variant, lds);
    case 4: return occupancy_s<M, C, 4, MIX>(variant, lds);
    case 6: return occupancy_s<M, C, 6, MIX>(variant, lds);
  }
  return 0;
}

template <class M>
int coop_rows_mech(int S) {
  switch (S) {
    case 2: return CoopLayout<M, 2>::ROWS;
    case 3: return CoopLayout<M, 3>::ROWS;
    case 4: return CoopLayout<M, 4>::ROWS;
    case 6: return CoopLayout<M, 6>::ROWS;
  }
  return 0;
}

template <class M, bool MIX>
int rows_mech(int S) {
  switch (S) {
    case 2: return Layout<M, 2, MIX>::ROWS;
    case 3: return Layout<M, 3, MIX>::ROWS;
    case 4: return Layout<M, 4, MIX>::ROWS;
    case 6: return Layout<M, 6, MIX>::ROWS;
  }
  return 0;
}

template <class M>
hipError_t component_mech(int which, int n, size_t ld, const double* a, const double* b,
                          const double* c, const double* d, const double* e, double* out,
                          int* iout, hipStream_t st) {
  const int bs = 256, nb = (n + bs - 1) / bs;
  if (n <= 0) return hipSuccess;
  switch (which) {
    case KPP_COMP_FUN: hipLaunchKernelGGL(k_fun<M>, dim3(nb), dim3(bs), 0, st, n, ld, a, b, c, out); break;
    case KPP_COMP_JAC: hipLaunchKernelGGL(k_jac<M>, dim3(nb), dim3(bs), 0, st, n, ld, a, b, c, out); break;
    case KPP_COMP_DECOMP: hipLaunchKernelGGL(k_decomp<M>, dim3(nb), dim3(bs), 0, st, n, ld, out, iout); break;
    case KPP_COMP_SOLVE: hipLaunchKernelGGL(k_solve<M>, dim3(nb), dim3(bs), 0, st, n, ld, a, out); break;
    case KPP_COMP_RATES:
      hipLaunchKernelGGL(k_rates<M>, dim3(nb), dim3(bs), 0, st, n, ld, a, b, c, d, e, out);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <class M, class C>
constexpr size_t coop_lds_bytes() {
  if constexpr (C::ENABLED) return CoopLds<M, C>::BYTES;
  else return 0;
}

template <class M, class D, class C>
MechOps make_ops(const char* name) {
  MechOps o;
  o.name = name;
  o.nvar = M::NVAR;
  o.nfix = M::NFIX;
  o.nspec = M::NSPEC;
  o.nreact = M::NREACT;
  o.nonzero = M::NONZERO;
  o.lu_nonzero = M::LU_NONZERO;
  o.rates_use_time = M::RATES_USE_TIME ? 1 : 0;
  o.n_sun_rates = M::N_SUN_RATES;
  o.rates_use_state = M::RATES_USE_STATE ? 1 : 0;
  o.n_state_rates = M::N_STATE_RATES;
  o.fits_registers = fits_registers<M>() ? 1 : 0;
  o.lds_max = M::LDS_MAX >= KPP_LDS_SPLIT_B ? KPP_LDS_SPLIT_B
            : (M::LDS_MAX >= KPP_LDS_SPLIT_A ? KPP_LDS_SPLIT_A : 0);
  o.species = D::names();
  o.lu_icol = D::icol();
  o.lu_crow = D::crow();
  o.lu_diag = D::diag();
  o.init = D::init();
  o.has_coop = C::ENABLED ? 1 : 0;
  o.coop_lds = coop_lds_bytes<M, C>();
  o.rows = rows_mech<M, false>;
  o.coop_rows = coop_rows_mech<M>;
  o.occupancy = occupancy_mech<M, C, false>;
  o.launch = launch_mech<M, C, false>;
  return o;
}

}  // namespace kpphip

// Defines the registry hook kpphip_mech_<NAME>() for mechanism NAME.
#define KPPHIP_DEFINE_MECH(NAME)                                                        \
  namespace {                                                                           \
  struct Data_##NAME {                                                                  \
    static const char* const* names() { return kppm_##NAME##_data::SPC_NAMES; }         \
    static const int* icol() { return kppm_##NAME##_data::LU_ICOL; }                    \
    static const int* crow() { return kppm_##NAME##_data::LU_CROW; }                    \
    static const int* diag() { return kppm_##NAME##_data::LU_DIAG; }                    \
    static const double* init() { return kppm_##NAME##_data::INIT; }                    \
  };                                                                                    \
  }                                                                                     \
  extern "C" const kpphip::MechOps* kpphip_mech_##NAME() {                              \
    static const kpphip::MechOps ops =                                                  \
        kpphip::make_ops<kppm_##NAME, Data_##NAME, kppm_##NAME##_coop>(#NAME);          \
    return &ops;                                                                        \
  }

// Defines kpphip_mech_mixed_<NAME>(): the mixed-precision integrator of NAME.
#define KPPHIP_DEFINE_MECH_MIXED(NAME)                                                  \
  extern "C" const kpphip::MixedOps* kpphip_mech_mixed_##NAME() {                       \
    static const kpphip::MixedOps ops = {                                               \
        kpphip::rows_mech<kppm_##NAME, true>,                                           \
        kpphip::occupancy_mech<kppm_##NAME, kppm_##NAME##_coop, true>,                  \
        kpphip::launch_mech<kppm_##NAME, kppm_##NAME##_coop, true>};                    \
    return &ops;                                                                        \
  }

// Defines kpphip_mech_comp_<NAME>(): the batched model functions of NAME.
#define KPPHIP_DEFINE_MECH_COMPONENTS(NAME)                                              \
  extern "C" kpphip::ComponentFn kpphip_mech_comp_##NAME() {                             \
    return kpphip::component_mech<kppm_##NAME>;                                          \
  }
