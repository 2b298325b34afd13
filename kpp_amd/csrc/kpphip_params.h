// Launch parameters shared by the host launcher and the device integrator.
#pragma once
#include <stdint.h>

#define KPP_SUN_FIXED 0 /* SUN given per cell and held fixed */
#define KPP_SUN_TOD 1   /* SUN from Update_SUN(TIME) at every rate update */
#define KPP_DELTAMIN 1.0e-6
#define KPP_SMAX 6
/* Per-cell codes with no counterpart in int/rosenbrock.c, whose reject loop
 * (:537-619) has no bound of its own: a NaN error estimate or a non-finite
 * step size (-9), or more than KPP_MAX_REJECT rejected attempts in a row
 * (-10), end the cell. */
#define KPP_IERR_NONFINITE (-9)
#define KPP_IERR_REJECTS (-10)
#define KPP_MAX_REJECT 100000

namespace kpphip {

// Rosenbrock method table in the layout of int/rosenbrock.c:779-1222
// (A and C strictly lower triangular, row-wise: A(i,j) = A[(i-1)(i-2)/2+j-1]).
struct RosCoef {
  int S;
  int NewF[KPP_SMAX];
  double A[KPP_SMAX * (KPP_SMAX - 1) / 2];
  double C[KPP_SMAX * (KPP_SMAX - 1) / 2];
  double M[KPP_SMAX], E[KPP_SMAX], Alpha[KPP_SMAX], Gamma[KPP_SMAX];
  double ELO;
};

struct RosParams {
  int ncells;
  int ldc;                 // leading dimension of the SoA arrays (>= ncells)
  double* C;               // [NSPEC][ldc] concentrations, in/out
  const double* temp;      // [ncells]
  const double* cfactor;   // [ncells]
  const double* sun;       // [ncells] SUN (fixed mode) or null
  int sun_mode;            // KPP_SUN_FIXED / KPP_SUN_TOD
  const double* tstart;    // [ncells] or null -> tstart0
  const double* tend;      // [ncells] or null -> tend0
  double tstart0, tend0;
  const double* atol;      // [NVAR] (vector tolerances) or null
  const double* rtol;
  double atol0, rtol0;     // scalar tolerances (vector_tol == 0)
  int vector_tol;
  int autonomous;
  int max_steps;
  double Roundoff, Hmin, Hmax_user, Hstart_user, FacMin, FacMax, FacRej, FacSafe;
  const double* hstart;    // [ncells] per-cell starting step or null
  int hstart_auto;         // IPAR[4] == 1: estimate the starting step where none is given
  const double* rconst;    // [NREACT][ldc] given rate constants or null
  double* rpre;            // [NREACT][ncells] rate constants computed by a pre-pass, or null
  int rates_given;
  int dfdt_zero;           // rates do not depend on T: dF/dT is exactly zero
  int stat0[8];            // IPAR[10..17] on entry
  int* ierr;               // [ncells]
  int* istat;              // [8][ldc] or null
  double* rstat;           // [2][ldc] (Texit, Hexit) or null
  double* hfirst;          // [ncells] size of each cell's first accepted step, or null
  int* fallbacks;          // [ncells] attempts that fell back to the FP64 LU (mixed), or null
  const int* order;        // [ncells] launch order (cell permutation) or null
  double* scratch;         // tiles
  int* queue;              // persistent kernel: next-cell counter (starts at 0)
  int refill_min;          // persistent kernel: refill when >= this many lanes idle
  int lanes;               // cells per wavefront (1..64)
  int lds_entries;         // persistent kernel: LU entries per lane mirrored in LDS (0: none)
  RosCoef coef;
};

}  // namespace kpphip
