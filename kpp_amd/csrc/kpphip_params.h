// Launch parameters shared by the host launcher and the device integrator.
#pragma once
#include <stdint.h>

#define KPP_SUN_FIXED 0 /* SUN given per cell and held fixed */
#define KPP_SUN_TOD 1   /* SUN from Update_SUN(TIME) at every rate update */
#define KPP_DELTAMIN 1.0e-6
/* DeltaMin of int/rosenbrock.f90's Rosenbrock and ros_Integrator (:287, :537):
 * the default starting step of the F90 integrator */
#define KPP_DELTAMIN_F90 1.0e-5
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

// SDIRK method and controls (int/sdirk.c:170-525, tables :988-1306) as the
// integrator uses them: Theta[j][i], Alpha[j][i] as rkTheta/rkAlpha.
struct SdCoef {
  int S;
  double Gamma, C[5], D[5], E[5], ELO, Theta[5][5], Alpha[5][5];
  int NewtonMaxit, StartNewton, Max_no_steps, ITOL;
  double Hmin, Hmax_user, Hstart_user, FacMin, FacMax, FacRej, FacSafe, ThetaMin, NewtonTol,
      Qmin, Qmax;
};

// The fully implicit Runge-Kutta integrator (int/runge_kutta.c): the
// method's tables and the controls of RungeKutta (:203-611).
struct RkCoef {
  int kind;  // 1 Radau-2A, 2 Radau-1A, 3 Lobatto-3C, 4 Gauss (the reference's rkMethod)
  double A[3][3], C[3], D[3], E[4], Bgam[5], Theta[3], Gamma, Alpha, Beta, ELO;
  double T[3][3], TinvAinv[3][3];
  int NewtonMaxit, StartNewton, Gustafsson, SdirkError, Max_no_steps, ITOL;
  double Hmin_user, Hmax_user, Hstart_user, FacMin, FacMax, FacRej, FacSafe, ThetaMin,
      NewtonTol, Qmin, Qmax;
};

// The discrete adjoint Rosenbrock integrator (kpp_rosenbrock_adj_dev)
struct AdjCoef {
  int adjoint;      // 1: discrete adjoint (AdjointType 2); 0: forward only (AdjointType 1)
  int nadj;         // adjoint vectors per cell
  double* lambda;   // [nadj*NVAR][ldc]: dg/dy at the end on entry, at the start on return
  double* chk;      // checkpoint tiles, nchk slots per cell
  int nchk;         // checkpoint slots per cell
  int save_lu;      // ICNTRL[7]: the factored matrices are checkpointed
  int first_step;   // 1: the backward loop includes the first accepted step
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
  // The F90 integrator of int/rosenbrock.f90 (kpp_integrate_f90_dev) instead
  // of int/rosenbrock.c: its start rule, outputs, unskipped array updates,
  // error norm over non-passive species with a 1e-10 floor, clipping.
  int f90;
  int nonneg;              // f90: ICNTRL(16) = 1, accepted states clipped at zero
  const double* np_w;      // f90: [NVAR] 0.0 for a passive species, 1.0 otherwise; null = none
  int np_count;            // f90: NonPassiveSpc_Count
  int rates_frozen;        // rates computed once when a cell starts, never updated
  int h211b;               // f90: the H211b step-size filter (int/rosenbrock_h211b_qssa.f90)
  double bb, kk;           // its parameters, RCNTRL(15) and RCNTRL(16)
  SdCoef sd;               // the SDIRK integrator (kpp_sdirk_dev)
  RkCoef rk;               // the Runge-Kutta integrator (kpp_rk_dev)
  AdjCoef adj;             // the discrete adjoint integrator (kpp_rosenbrock_adj_dev)
};

}  // namespace kpphip
