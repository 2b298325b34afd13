// Host-visible description of one compiled mechanism.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "kpphip_params.h"

#define KPP_VARIANT_PLAIN 0
#define KPP_VARIANT_PERSISTENT 1
#define KPP_VARIANT_REGISTER 2
#define KPP_VARIANT_COOP 3

// LDS mirror sizes (LU entries per lane) the persistent kernel is built with,
// where the mechanism's LDS_MAX allows (24 x 64 lanes x 8 B = 12 KiB per
// one-wavefront workgroup).  Larger mirrors fit in LDS but were slower on
// saprc99: the mirrored values held in registers through the triangular
// solves push the compiler into a one-load-at-a-time schedule there.
#define KPP_LDS_SPLIT_A 20
#define KPP_LDS_SPLIT_B 24

#define KPP_COMP_FUN 0
#define KPP_COMP_JAC 1
#define KPP_COMP_DECOMP 2
#define KPP_COMP_SOLVE 3
#define KPP_COMP_RATES 4

namespace kpphip {

struct MechOps {
  const char* name;
  int nvar, nfix, nspec, nreact, nonzero, lu_nonzero;
  int rates_use_time, n_sun_rates, rates_use_state, n_state_rates;
  int fits_registers;  // KPP_VARIANT_REGISTER is compiled for this mechanism
  int has_coop;        // KPP_VARIANT_COOP is compiled for this mechanism
  size_t coop_lds;     // LDS bytes per cell of the cooperative kernel
  int lds_max;         // largest LDS mirror (LU entries) the persistent kernel is built with
  const char* const* species;
  const int *lu_icol, *lu_crow, *lu_diag;
  const double* init;  // Initialize(): #INITVALUES times CFACTOR, [NSPEC]
  int (*rows)(int S);  // tile rows per lane for S stages
  int (*coop_rows)(int S);  // doubles per wavefront block of the cooperative kernel
  // resident wavefronts per CU (0 if unknown) with `lds` mirrored LU entries per lane
  int (*occupancy)(int variant, int S, int lds);
  hipError_t (*launch)(const RosParams& p, int variant, int nblocks, hipStream_t st);
};

// Batched model functions of a mechanism (their own translation unit).
typedef hipError_t (*ComponentFn)(int which, int n, size_t ld, const double* a, const double* b,
                                  const double* c, const double* d, const double* e, double* out,
                                  int* iout, hipStream_t st);

// The SDIRK integrator of a mechanism (its own translation unit), and in the
// same form the Runge-Kutta integrator (rows' argument unused); rows = 0
// where it is not built.
struct SdirkOps {
  int (*rows)(int S);
  hipError_t (*launch)(const RosParams& p, int nblocks, hipStream_t st);
};

// The mixed-precision integrator of a mechanism (its own translation unit):
// FP32 LU and triangular solves with one FP64 refinement sweep per solve.
struct MixedOps {
  int (*rows)(int S);
  int (*occupancy)(int variant, int S, int lds);
  hipError_t (*launch)(const RosParams& p, int variant, int nblocks, hipStream_t st);
};

}  // namespace kpphip
