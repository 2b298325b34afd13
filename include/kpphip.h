/* kpphip.h — C ABI of the MI355X batched Rosenbrock solver for KPP mechanisms.
 *
 * Drop-in boundary for KPP's C integration path.  KPP's reference surface
 * integrates ONE cell per call over global state:
 *   void INTEGRATE(double TIN, double TOUT)                int/rosenbrock.c:94-133
 *   int  Rosenbrock(double Y[], double Tstart, double Tend,
 *                   double AbsTol[], double RelTol[], ode_Fun, ode_Jac,
 *                   double RPAR[], int IPAR[])             int/rosenbrock.c:137-446
 * and the generated model ABI Fun/Jac_SP/KppDecomp/KppSolve/Update_RCONST
 * (prototypes int/rosenbrock.c:54-91; generated by src/gen.c:642-760,
 * 1185-1319, 1960-2033, 2343-2473; KppDecomp util/sutil.c:2-23).
 * The functions below do the same for many independent cells at once on one
 * GPU.  IPAR/RPAR keep the index meanings of int/rosenbrock.c:201-254:
 *   IPAR[0] autonomous, IPAR[1] scalar tolerances (0 = vector), IPAR[2] max
 *   steps, IPAR[3] method (1 Ros2, 2 Ros3, 3 Ros4, 4 Rodas3, 5 Rodas4),
 *   IPAR[10..17] Nfun Njac Nstp Nacc Nrej Ndec Nsol Nsng on entry;
 *   IPAR[4] (unused by the reference) = 1: where no starting step is given
 *   (RPAR[2] or hstart[c] = 0), estimate one per cell (Hairer, Norsett &
 *   Wanner, Solving ODEs I, II.4; two function evaluations, counted in Nfun)
 *   instead of the reference's 1e-6 s;
 *   RPAR[0..6] Hmin Hmax Hstart FacMin FacMax FacRej FacSafe.
 * Per-cell results use Rosenbrock's codes (1 success, -1..-8 failures) and
 * two of this library's own, for cells whose reject loop would never end
 * (int/rosenbrock.c:537-619 has no bound; a GPU wavefront must finish):
 *   -9  the error estimate is NaN or the new step size is not finite (a NaN
 *       or Inf in the cell's inputs, rates or Jacobian);
 *  -10  more than 100000 rejected attempts in a row (step factors that do
 *       not shrink the step, e.g. FacMin = FacRej = 1).
 * Per-cell statistics come back in the IPAR[10..17] order, Texit/Hexit in
 * the RPAR[10..11] order.  Parameter errors (-1..-5) are returned by the
 * call itself, as Rosenbrock returns them before integrating.  API misuse
 * returns KPP_E_* (<= -100).  Nothing is printed.
 *
 * Arrays are structure-of-arrays: element s of cell c is X[s*ld + c].
 * Concentrations C hold NSPEC rows: NVAR variable species, then NFIX fixed
 * species, in the mechanism's KPP order (kpp_mech_species()).
 */
#ifndef KPPHIP_H
#define KPPHIP_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KPP_E_MECH (-101)      /* unknown mechanism name */
#define KPP_E_ARG (-102)       /* bad argument (null pointer, size) */
#define KPP_E_HIP (-103)       /* HIP runtime error */
#define KPP_E_WORKSPACE (-104) /* workspace too small */
#define KPP_E_NODEV (-105)     /* no GPU device */
#define KPP_E_WEIGHTS (-106)   /* scheduler weights file missing or not for this mechanism */

#define KPP_SUN_FIXED 0 /* SUN given per cell and held fixed (Update_SUN bypassed) */
#define KPP_SUN_TOD 1   /* SUN from Update_SUN(TIME) (util/UpdateSun.c:1-20) */

#define KPP_VARIANT_PLAIN 0      /* one lane per cell, cells in launch order */
#define KPP_VARIANT_PERSISTENT 1 /* persistent work-stealing wavefronts */
#define KPP_VARIANT_REGISTER 2   /* one lane per cell, working set in registers: small
                                    mechanisms only (<= 64 LU entries and reactions),
                                    else the call returns KPP_E_ARG */
#define KPP_VARIANT_COOP 3       /* one wavefront per cell (its 64 lanes share the model
                                    routines; the LU in LDS): large mechanisms; FP64, for
                                    mechanisms built with cooperative programs, else
                                    KPP_E_ARG.  Persistent grid: kpp_launch.nblocks
                                    wavefronts (0 = all that fit), lanes 0 or 64 */

/* ---- mechanisms compiled into the library ---- */
int kpp_mech_count(void);
const char* kpp_mech_name(int i);
/* sizes: out[0..9] = NVAR NFIX NSPEC NREACT NONZERO LU_NONZERO
 *        RATES_USE_TIME N_SUN_RATES RATES_USE_STATE N_STATE_RATES
 * (ROOT_Parameters.h, src/gen.c:2627-2720; the last four describe which rate
 * constants depend on TIME/SUN and on the concentrations, e.g. MCM's RO2 sum) */
int kpp_mech_info(const char* mech, int out[10]);
const char* kpp_mech_species(const char* mech, int k);
/* Initialize() (src/gen.c:3252-3290): C[NSPEC] = #INITVALUES times CFACTOR */
int kpp_mech_initial(const char* mech, double* C);
/* LU_ICOL[LU_NONZERO], LU_CROW[NVAR+1], LU_DIAG[NVAR+1] (ROOT_JacobianSP,
 * src/gen.c:530-588), 0-based */
int kpp_mech_lu_pattern(const char* mech, int* icol, int* crow, int* diag);

/* Rosenbrock's parameter checks (int/rosenbrock.c:286-382) on host data:
 * returns 0 if valid, else -1..-5 exactly as Rosenbrock would. */
int kpp_check_params(const int IPAR[20], const double RPAR[20], const double* atol,
                     const double* rtol, int nvar);

/* ---- device-resident batch (inputs already in GPU memory) ---- */
typedef struct {
  int ncells;            /* number of cells */
  int ld;                /* leading dimension of every SoA array (>= ncells) */
  double* C;             /* device [NSPEC][ld] concentrations, in/out */
  const double* temp;    /* device [ncells] TEMP (K) */
  const double* cfactor; /* device [ncells] CFACTOR (M/1e6 for saprc99 rate laws) */
  const double* sun;     /* device [ncells] SUN in fixed mode, else NULL */
  int sun_mode;          /* KPP_SUN_FIXED or KPP_SUN_TOD */
  const double* tstart;  /* device [ncells] or NULL (then tstart0) */
  const double* tend;    /* device [ncells] or NULL (then tend0) */
  double tstart0, tend0;
  const double* hstart;  /* device [ncells] starting step per cell (>=0) or NULL (RPAR[2]) */
  const double* rconst;  /* device [NREACT][ld] rate constants supplied by the caller or NULL */
  int* ierr;             /* device [ncells] Rosenbrock return code per cell */
  int* istat;            /* device [8][ld] IPAR[10..17] per cell, or NULL */
  double* rstat;         /* device [2][ld] Texit, Hexit per cell, or NULL
                            (kpp_integrate_f90_dev, kpp_sdirk_dev: [3][ld]) */
  const int* order;      /* device [ncells] processing order (a permutation) or NULL */
  double* hfirst;        /* device [ncells] step log: size of each cell's first accepted step
                            (0 if none was accepted), or NULL.  Not a reference output. */
  int* fallbacks;        /* device [ncells]: attempts that fell back from the FP32 to the FP64
                            factorisation (kpp_launch.lu_fp32), or NULL */
} kpp_batch;

typedef struct {
  int variant;    /* KPP_VARIANT_PLAIN / _PERSISTENT / _REGISTER */
  int nblocks;    /* persistent: wavefronts in the grid (0 = library default) */
  int refill_min; /* persistent: refill when this many lanes are idle (0 = default 1) */
  int lanes;      /* cells per wavefront, 1..64 (0 = library default: 64, fewer for
                     batches too small to fill the device) */
  int lu_fp32;    /* 0: FP64 KppDecomp/KppSolve (the reference's arithmetic).
                     1: mixed precision, no counterpart in KPP: the iteration matrix is
                     factored in FP32 and every stage solve is an FP32 solve followed by
                     one FP64 iterative-refinement sweep (FP64 residual against the FP64
                     matrix, FP32 correction); a cell whose sweep does not converge
                     (correction above 1e-5 of the solution, or not finite) finishes the
                     attempt with the FP64 factorisation.  State, stage sums, error norm
                     and step control stay FP64. */
  int lds_entries; /* persistent, FP64 only: the factored values of this many LU entries
                      per cell (those KppDecomp re-reads most) are kept in LDS instead of
                      the workspace tile; no counterpart in KPP, results bitwise the same.
                      0 (none), 20 or 24; 20 and 24 only for mechanisms built with an LDS
                      mirror (saprc99, saprcnov), anything else is KPP_E_ARG. */
} kpp_launch;

/* Device bytes of workspace needed for a launch. */
size_t kpp_workspace_bytes(const char* mech, int method, const kpp_launch* l, int ncells);

/* Batched Rosenbrock (int/rosenbrock.c:137-627) on device-resident data.
 * atol/rtol are HOST arrays of NVAR entries (IPAR[1]==0) or 1 entry.
 * workspace: device memory of kpp_workspace_bytes() bytes; stream: a
 * hipStream_t (NULL = default stream).  Asynchronous: results are ready
 * when the stream is.  Returns 0, a parameter error -1..-5, or KPP_E_*. */
int kpp_rosenbrock_dev(const char* mech, const kpp_batch* b, const int IPAR[20],
                       const double RPAR[20], const double* atol, const double* rtol,
                       const kpp_launch* l, void* workspace, size_t workspace_bytes,
                       void* stream);

/* kpp_rosenbrock_dev without a workspace argument.  No counterpart in KPP: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, hosts that forbid allocation).  It calls no hipMalloc,
 * hipFree or other allocating routine: the workspace is a library-owned
 * device array of KPP_STATIC_SCRATCH_BYTES per GPU, held until the process
 * ends.  The persistent kernel computes each cell's rate constants when the
 * cell starts (no pre-pass buffer) and runs as many wavefronts as fit; the
 * plain kernel needs a tile per 64 cells and a batch whose tiles exceed the
 * array returns KPP_E_WORKSPACE, as does any launch that does not fit.
 * Results are those of kpp_rosenbrock_dev bit for bit.  Every call on a GPU
 * shares that GPU's array, so calls that overlap on one GPU (two streams,
 * two host threads) race: serialise them. */
#define KPP_STATIC_SCRATCH_BYTES ((size_t)256 << 20)
int kpp_rosenbrock_dev_static(const char* mech, const kpp_batch* b, const int IPAR[20],
                              const double RPAR[20], const double* atol, const double* rtol,
                              const kpp_launch* l, void* stream);

/* ---- the F90 integrator (int/rosenbrock.f90) ----
 * KPP's Fortran 90 Rosenbrock is a separate implementation with its own
 * controls and outputs.  kpp_integrate_f90_dev replaces its
 *   SUBROUTINE INTEGRATE(TIN, TOUT, ICNTRL_U, RCNTRL_U, ISTATUS_U, RSTATUS_U,
 *                        IERR_U)                   int/rosenbrock.f90:43-125
 * for every cell of a device batch, with the semantics of its Rosenbrock
 * (:128-446) and ros_Integrator (:492-686):
 *  - ICNTRL_U/RCNTRL_U are merged as INTEGRATE does (ICNTRL = 0 with
 *    ICNTRL(15) = 5, then ICNTRL_U where /= 0; RCNTRL_U where > 0); either
 *    may be NULL (an absent optional argument);
 *  - ICNTRL(1) autonomous, (2) 0 = vector / 1 = scalar tolerances, (3) method
 *    0,4 Rodas3, 1 Ros2, 2 Ros3, 3 Ros4, 5 Rodas4, 6 Rang3, 7 Rodas3.1,
 *    (4) max steps (0 = 200000), (15) rate updates (util/util.f90:133-159:
 *    bit 1 Update_RCONST at every Fun/Jac, with bit 4 SUN from Update_SUN of
 *    the time, else the cell's b->sun; without bit 1 the rates stay at their
 *    values at TIN), (16) = 1 clips accepted states at zero;
 *  - RCNTRL(1:7) Hmin, Hmax, Hstart (0: MAX(Hmin, 1e-5)), FacMin, FacMax,
 *    FacRej, FacSafe;
 *  - b->tstart/tend (or tstart0/tend0) are TIN/TOUT; b->sun_mode is ignored;
 *    b->hstart per cell overrides RCNTRL(3); b->rconst supplies the rates;
 *  - b->istat [8][ld] is ISTATUS(1:8); b->rstat must hold [3][ld]:
 *    RSTATUS(1:3) = Texit, Hexit, Hnew of the last accepted step (0 if none);
 *  - passive: host [NVAR] or NULL; passive[i] != 0 leaves species i out of
 *    the error norm (NonPassiveSpc, src/gen.c:3355-3384: ATOL(i) >=
 *    PassiveSpc_ATOL_Threshold);
 *  - plain, persistent and register kernels; FP64 only (lu_fp32 and
 *    KPP_VARIANT_COOP are KPP_E_ARG).
 * Per-cell codes 1, -1..-8 as int/rosenbrock.f90:452-489; -9 also for a NaN
 * error estimate (gfortran's MAX(Err, 1e-10) would accept it).  Returns 0, a
 * parameter error -1..-5 (every cell's ierr set to it), or KPP_E_*. */
int kpp_integrate_f90_dev(const char* mech, const kpp_batch* b, const int ICNTRL_U[20],
                          const double RCNTRL_U[20], const double* atol, const double* rtol,
                          const int* passive, const kpp_launch* l, void* workspace,
                          size_t workspace_bytes, void* stream);
/* The INTEGRATE of int/rosenbrock_h211b_qssa.f90 (:56-138): the F90
 * integrator above with its step-size control replaced, when RCNTRL(15) = bb
 * and RCNTRL(16) = kk are both set, by the H211b digital filter (:882-892;
 * Soderlind, ACM TOMS 29 (2003)): Fac = (1/Err)^(1/(bb kk)) *
 * (1/Err_prev)^(1/(bb kk)) * Fac_prev^(-1/bb), without the FacMin/FacMax
 * bounds, the filter's memory starting at 1; the error norm over all
 * species.  The QSSA options (ICNTRL(18) /= 0) are not built: KPP_E_ARG. */
int kpp_integrate_h211b_dev(const char* mech, const kpp_batch* b, const int ICNTRL_U[20],
                            const double RCNTRL_U[20], const double* atol, const double* rtol,
                            const kpp_launch* l, void* workspace, size_t workspace_bytes,
                            void* stream);
/* workspace of kpp_integrate_f90_dev for method ICNTRL(3) = icntrl3 */
size_t kpp_workspace_bytes_f90(const char* mech, int icntrl3, const kpp_launch* l, int ncells);
/* the parameter checks of int/rosenbrock.f90:306-433 after INTEGRATE's merge */
int kpp_check_params_f90(const int ICNTRL_U[20], const double RCNTRL_U[20], const double* atol,
                         const double* rtol, int nvar);

/* ---- the C SDIRK integrator (int/sdirk.c) on the same device functions ----
 * kpp_sdirk_dev replaces
 *   int SDIRK(int N, KPP_REAL Tinitial, KPP_REAL Tfinal, KPP_REAL Y[],
 *             KPP_REAL RelTol[], KPP_REAL AbsTol[], KPP_REAL RCNTRL[],
 *             int ICNTRL[], KPP_REAL RSTATUS[], int ISTATUS[])   int/sdirk.c:170
 * (and its INTEGRATE, :91-166) for every cell of a device batch, one lane per
 * cell at a time (a resident grid whose lanes take cells from a work queue,
 * in b->order when given), over [b->tstart, b->tend] (or tstart0/tend0):
 *  - ICNTRL/RCNTRL as SDIRK takes them (0-based, no merge): ICNTRL[1] 0 =
 *    vector tolerances, ICNTRL[2] method 0/1 Sdirk2a, 2 Sdirk2b, 3 Sdirk3a,
 *    4 Sdirk4a, 5 Sdirk4b (any other: Sdirk2a), [3] max steps (0: 200000),
 *    [4] max Newton iterations (0: 8), [5] 0 = extrapolated Newton start;
 *    RCNTRL[0..10] Hmin, Hmax, Hstart, FacMin (0.2), FacMax (10), FacRej
 *    (0.1), FacSafe (0.9), ThetaMin (1e-3), NewtonTol (3e-2), Qmin (1),
 *    Qmax (1.2);
 *  - b->sun_mode, b->hstart (overrides RCNTRL[2]) and b->rconst as for
 *    kpp_rosenbrock_dev; b->istat [8][ld] = ISTATUS[1..8] (Nfun..Nsng);
 *    b->rstat must hold [3][ld] = RSTATUS[1..3] (Texit, Hexit, Hnew);
 *  - the reference's SDIRK_PrepareMatrix halves only its local step on a
 *    singular matrix and never reports -8; so does this one;
 *  - deviations: parameter errors -1..-5 are returned (the reference prints
 *    them and integrates with the setting uninitialised), failures -6/-7 are
 *    returned per cell (the reference returns 0), -9/-10 as for Rosenbrock.
 * Not built for the largest mechanisms (LU_NONZERO > 4096): KPP_E_ARG.
 * Asynchronous on `stream`; the workspace holds kpp_workspace_bytes_sdirk(). */
size_t kpp_workspace_bytes_sdirk(const char* mech, int icntrl2, int ncells);
int kpp_sdirk_dev(const char* mech, const kpp_batch* b, const int ICNTRL[20],
                  const double RCNTRL[20], const double* atol, const double* rtol,
                  void* workspace, size_t workspace_bytes, void* stream);

/* ---- the C fully implicit Runge-Kutta integrator (int/runge_kutta.c) ----
 * kpp_rk_dev replaces
 *   void RungeKutta(int N, KPP_REAL T, KPP_REAL Tend, KPP_REAL Y[],
 *                   KPP_REAL RelTol[], KPP_REAL AbsTol[], KPP_REAL RCNTRL[],
 *                   int ICNTRL[], KPP_REAL RSTATUS[], int ISTATUS[],
 *                   int* IERR)                         int/runge_kutta.c:203
 * (and its INTEGRATE, :114-200) for every cell of a device batch, one lane
 * per cell at a time (a work queue as kpp_sdirk_dev's), over [b->tstart,
 * b->tend] (or tstart0/tend0), with the complex
 * sparse LU of util/sutil.c:28-87 on the device:
 *  - ICNTRL/RCNTRL as RungeKutta takes them (0-based, no merge): ICNTRL[1]
 *    0 = vector tolerances, [2] method 0/1 Radau-2A, 2 Lobatto-3C, 3 Gauss,
 *    4 Radau-1A (the reference's switch), [3] max steps (0: 200000), [4] max
 *    Newton iterations (0: 8), [5] 0 = extrapolated Newton start, [9] 1 =
 *    SDIRK error estimator (INTEGRATE's choice), 0 classical, [10] 0 =
 *    Gustafsson's predictive controller; RCNTRL[0..10] Hmin, Hmax, Hstart,
 *    FacMin (0.2), FacMax (8), FacRej (0.1), FacSafe (0.9), ThetaMin (1e-3),
 *    NewtonTol (3e-2), Qmin (1), Qmax (1.2);
 *  - b->sun_mode, b->hstart (overrides RCNTRL[2]) and b->rconst as for
 *    kpp_rosenbrock_dev; b->istat [8][ld] = ISTATUS[0..7] (Nfun..Nsng, the
 *    reference's counting: RK_PrepareRHS's evaluations and the solves of the
 *    error estimators are not counted); b->rstat must hold [3][ld] =
 *    RSTATUS[0..2] (Texit, Hacc, Hnew);
 *  - every parameter check runs and the last failing one's code, -1..-8 or
 *    -13, is returned (every cell's ierr set to it), as the reference's IERR;
 *  - the -10 (step size) and -12 (five singular matrices in a row)
 *    conditions are only reported by the reference, which goes on; so does
 *    this one;
 *  - deviations: -9 (more than ICNTRL[3] steps) ends the cell (the
 *    reference prints it and goes on), a NaN error estimate or non-finite
 *    step ends it with -14, more than 100000 passes in a row in which T does
 *    not advance with -15; ICNTRL[2] = 5 (Lobatto-3A, "not yet implemented"
 *    in the reference) is -13.
 * Not built for the largest mechanisms (LU_NONZERO > 4096): KPP_E_ARG.
 * Asynchronous on `stream`; the workspace holds kpp_workspace_bytes_rk(). */
size_t kpp_workspace_bytes_rk(const char* mech, int ncells);
int kpp_rk_dev(const char* mech, const kpp_batch* b, const int ICNTRL[20],
               const double RCNTRL[20], const double* atol, const double* rtol, void* workspace,
               size_t workspace_bytes, void* stream);

/* ---- the C discrete adjoint Rosenbrock integrator (int/rosenbrock_adj.c) ----
 * kpp_rosenbrock_adj_dev replaces
 *   int RosenbrockADJ(KPP_REAL Y[], int NADJ, KPP_REAL Lambda[][NVAR],
 *                     KPP_REAL Tstart, KPP_REAL Tend, KPP_REAL AbsTol[],
 *                     KPP_REAL RelTol[], KPP_REAL AbsTol_adj[][NVAR],
 *                     KPP_REAL RelTol_adj[][NVAR], KPP_REAL RCNTRL[],
 *                     int ICNTRL[], KPP_REAL RSTATUS[], int ISTATUS[])
 *                                                  int/rosenbrock_adj.c:191
 * (and its INTEGRATE_ADJ, :116-188) for every cell of a device batch, one
 * lane per cell, forward then backward, the adjoint's model functions
 * (Hessian, HessTR_Vec, JacTR_SP_Vec, KppSolveTR) on the device:
 *  - ICNTRL/RCNTRL as RosenbrockADJ takes them (0-based, no merge):
 *    ICNTRL[0] autonomous, [1] 0 = vector tolerances, [2] method 0/4 Rodas3,
 *    1 Ros2, 2 Ros3, 3 Ros4, 5 Rodas4, 7 Rodas3_1 (the reference's switch and
 *    tables), [3] must be 0 (the reference's check of it always fails), [5]
 *    CadjMethod (checked, unused), [6] AdjointType 0/2 discrete, 1 none, 3/4
 *    continuous (not built: -13), [7] SaveLU; RCNTRL[0..6] Hmin, Hmax,
 *    Hstart (0: max(Hmin, 1e-5)), FacMin (0.2), FacMax (6), FacRej (0.1),
 *    FacSafe (0.9); AbsTol_adj/RelTol_adj only serve the continuous types;
 *  - lambda [nadj*NVAR][b->ld] (row a*NVAR + i is Lambda[a][i]) holds
 *    dg_a/dy(tend) on entry and dg_a/dy(tstart) on return; nadj may be 0;
 *  - nchk checkpoint slots per cell (the reference's bufsize is 200000): a
 *    cell with more accepted steps ends with -12 (the reference exits);
 *  - first_step 0 keeps the reference's backward loop (while stack_ptr > 0),
 *    which leaves out the first accepted step; 1 differentiates every step
 *    (the F90 integrator's stack, int/rosenbrock_adj.f90);
 *  - b->sun_mode and b->rconst as for kpp_rosenbrock_dev (b->hstart is not
 *    read); b->istat [8][ld] = ISTATUS[1..8] (Nfun..Nsng, backward pass
 *    included); b->rstat must hold [3][ld] = RSTATUS[1..3] (Texit, Hexit,
 *    Hnew);
 *  - parameter errors (-1..-5, -9, -13; the first failing check in the
 *    reference's order) set every cell's ierr and are returned;
 *  - deviations: a NaN error estimate or non-finite step ends the cell with
 *    -10, more than 100000 rejections in a row with -11.
 * Built for mechanisms compiled with the adjoint's functions (NVAR <= 300)
 * whose LU_NONZERO <= 4096, else KPP_E_ARG.  Asynchronous on `stream`; the
 * workspace holds kpp_workspace_bytes_adj() (0: not built or bad method). */
size_t kpp_workspace_bytes_adj(const char* mech, int ncells, const int ICNTRL[20], int nadj,
                               int nchk);
int kpp_rosenbrock_adj_dev(const char* mech, const kpp_batch* b, const int ICNTRL[20],
                           const double RCNTRL[20], const double* atol, const double* rtol,
                           int nadj, double* lambda, int nchk, int first_step, void* workspace,
                           size_t workspace_bytes, void* stream);

/* ---- batched generated-model functions (device SoA arrays, ld >= n) ---- */
/* Fun(V,F,RCT,Vdot)                 src/gen.c:642-760 */
int kpp_fun_dev(const char* mech, int n, int ld, const double* V, const double* F,
                const double* RCT, double* Vdot, void* stream);
/* Jac_SP(V,F,RCT,JVS)               src/gen.c:1185-1319 */
int kpp_jac_sp_dev(const char* mech, int n, int ld, const double* V, const double* F,
                   const double* RCT, double* JVS, void* stream);
/* KppDecomp(JVS) -> ising per cell  util/sutil.c:2-23 */
int kpp_decomp_dev(const char* mech, int n, int ld, double* JVS, int* ising, void* stream);
/* KppSolve(JVS,X)                   src/gen.c:1960-2033 */
int kpp_solve_dev(const char* mech, int n, int ld, const double* JVS, double* X, void* stream);
/* Update_RCONST at per-cell TEMP/CFACTOR/SUN/TIME (src/gen.c:2343-2473;
 * rate laws util/UserRateLaws.c:9-78).  NUMBER-type rates are constants
 * inside Fun/Jac_SP and are written as 0. */
int kpp_rates_dev(const char* mech, int n, int ld, const double* temp, const double* cfactor,
                  const double* sun, const double* time, double* RCONST, void* stream);
/* the same for mechanisms whose rate code reads the concentrations (the F90
 * rate code inlined into Update_RCONST, e.g. MCM's RO2 sum): C is the device
 * [NSPEC][ld] state the rates see */
int kpp_rates_c_dev(const char* mech, int n, int ld, const double* temp, const double* cfactor,
                    const double* sun, const double* time, const double* C, double* RCONST,
                    void* stream);

/* ---- host-buffer convenience API (allocates device memory itself) ---- */
typedef struct kpp_ctx kpp_ctx;
/* Context flags (or-ed with KPP_SUN_TOD).  None has a counterpart in KPP; the
 * default (0) integrates every cell with the plain one-lane-per-cell kernel
 * in KPP's FP64 arithmetic. */
#define KPP_CTX_PERSISTENT 0x10  /* persistent work-stealing kernel, default launch shape */
#define KPP_CTX_AUTOTUNE 0x20    /* persistent kernel with the launch shape kpp_autotune
                                    picks on the context's first integration (it integrates
                                    that batch once per candidate, then restores it) */
#define KPP_CTX_LEARNED 0x40     /* learned step scheduler: starting step per cell where the
                                    call gives none, and the launch order; weights from
                                    <library dir>/<mech>_step_mlp.kppw (written by the
                                    build) or kpp_ctx_load_scheduler; KPP_E_WEIGHTS if none */
#define KPP_CTX_MIXED 0x80       /* FP32 LU with FP64 refinement (kpp_launch.lu_fp32 = 1);
                                    not with KPP_CTX_AUTOTUNE */
#define KPP_CTX_HSTART_AUTO 0x100 /* INTEGRATE_BATCH: estimated starting step (IPAR[4] = 1)
                                     where STEPMIN is 0 */
/* method: IPAR[3] code; flags: KPP_SUN_TOD to drive SUN by time of day, KPP_CTX_* */
int kpp_ctx_create(kpp_ctx** ctx, const char* mech, int ncells, int method, int flags);
void kpp_ctx_destroy(kpp_ctx* ctx);
/* KPP_CTX_LEARNED: replace the scheduler's weights (a file kpp_sched_load reads) */
int kpp_ctx_load_scheduler(kpp_ctx* ctx, const char* path);
/* the launch shape the context integrates with (after autotuning, the chosen one) */
int kpp_ctx_get_launch(kpp_ctx* ctx, kpp_launch* l);
/* host [ncells]: the starting step each cell was given in the last integration
 * (the caller's, the carried STEPMIN, the scheduler's; 0 where RPAR[2]'s rule applied) */
int kpp_ctx_get_hstart(kpp_ctx* ctx, double* hstart);
/* host SoA C[NSPEC][ncells]; per-cell temp, cfactor, sun (NULL in TOD mode) */
int kpp_set_state(kpp_ctx* ctx, const double* C, const double* temp, const double* cfactor,
                  const double* sun);
/* Rosenbrock over all cells, host arrays; tstart/tend per cell; hstart per
 * cell or NULL; IPAR[3] must equal the context's method. Synchronous. */
int kpp_rosenbrock_batch(kpp_ctx* ctx, const double* tstart, const double* tend,
                         const double* atol, const double* rtol, const int IPAR[20],
                         const double RPAR[20], const double* hstart);
/* results to host: C[NSPEC][ncells], ierr[ncells], istat[8][ncells], rstat[2][ncells]
 * (any may be NULL) */
int kpp_get_state(kpp_ctx* ctx, double* C, int* ierr, int* istat, double* rstat);
/* INTEGRATE(TIN,TOUT) for every cell (int/rosenbrock.c:94-133): non-autonomous,
 * Rodas4 unless the context was created with another method, tolerances
 * set by kpp_ctx_set_tol (default ATOL 1, RTOL 1e-3 scalar), STEPMIN per
 * cell carried from the previous call's Hexit. */
int kpp_ctx_set_tol(kpp_ctx* ctx, const double* atol, const double* rtol, int vector_tol);
int INTEGRATE_BATCH(kpp_ctx* ctx, double TIN, double TOUT);
/* The F90 INTEGRATE(TIN, TOUT, ICNTRL_U, RCNTRL_U, ISTATUS_U, RSTATUS_U,
 * IERR_U) (int/rosenbrock.f90:43-125) for every cell, through
 * kpp_integrate_f90_dev: the context's state, tolerances (vector tolerances
 * are required when ICNTRL(2) = 0) and passive species; host outputs, each
 * optional: ISTATUS_U [8][ncells], RSTATUS_U [3][ncells], IERR_U [ncells].
 * Synchronous.  Not with KPP_CTX_MIXED; the learned scheduler is not used. */
int INTEGRATE_F90_BATCH(kpp_ctx* ctx, double TIN, double TOUT, const int ICNTRL_U[20],
                        const double RCNTRL_U[20], int* ISTATUS_U, double* RSTATUS_U,
                        int* IERR_U);
/* passive species of INTEGRATE_F90_BATCH: host [NVAR], nonzero = passive, or
 * NULL for none (NonPassiveSpc_Indices of the F90 Initialize) */
int kpp_ctx_set_passive(kpp_ctx* ctx, const int* passive);

/* ---- autotuner (no counterpart in KPP) ----
 * Times candidate launch shapes of the persistent kernel (cells per
 * wavefront, refill threshold, grid size) on the caller's batch and returns
 * the fastest in *best (and its time in *best_ms).  C0 is a device copy of
 * the batch's initial C ([NSPEC][ld]); the batch's C is reset from it before
 * every trial and once more at the end.  Synchronises the stream.  The
 * workspace must hold kpp_autotune_workspace_bytes() bytes. */
size_t kpp_autotune_workspace_bytes(const char* mech, int method, int ncells);
int kpp_autotune(const char* mech, const kpp_batch* b, const double* C0, const int IPAR[20],
                 const double RPAR[20], const double* atol, const double* rtol, void* workspace,
                 size_t workspace_bytes, void* stream, kpp_launch* best, float* best_ms,
                 int* ntried);

/* ---- graphs: record a sequence of the asynchronous calls above once, replay it ----
 * No counterpart in KPP.  kpp_capture_begin/kpp_capture_end bracket calls
 * made on one non-default stream (capture mode thread-local); the recorded
 * kernels keep the pointers and arguments they were called with, so a replay
 * integrates whatever those device arrays hold at launch time.
 * kpp_rosenbrock_dev records only with scalar tolerances (IPAR[1] != 0):
 * vector tolerances need a host-to-device copy and return KPP_E_ARG under
 * capture.  Nothing the library records clears memory with a memset node. */
typedef struct kpp_graph kpp_graph;
int kpp_capture_begin(void* stream);
int kpp_capture_end(void* stream, kpp_graph** graph);
int kpp_graph_launch(kpp_graph* graph, void* stream);
void kpp_graph_destroy(kpp_graph* graph);

/* ---- learned step scheduler (no counterpart in KPP) ----
 * A 4-layer MLP (in_pad -> 128 -> 128 -> 128 -> 16, GELU) on BF16 matrix
 * cores with FP32 accumulation predicts, from a cell's initial state, the
 * size of its first accepted step (output 0, standardised log10) and one of
 * 8 step-count buckets (outputs 1..8, logits).  The prediction only sets the
 * per-cell starting step (hstart, Rosenbrock's RPAR[2]) and the launch order
 * (order); Rosenbrock's error control still judges every step.
 * Parameters: FP32 master copy `w` and BF16 copy `wb`, both laid out as
 * W1[128][in_pad] b1[128] W2[128][128] b2[128] W3[128][128] b3[128]
 * W4[16][128] b4[16] (kpp_mlp_param_offsets). */
#define KPP_MLP_HIDDEN 128
#define KPP_MLP_OUT 16
#define KPP_MLP_BUCKETS 8
typedef struct {
  int nin;    /* inputs used: NSPEC + 2 (log10 concentrations, TEMP, SUN) */
  int in_pad; /* padded input width, a multiple of 32 */
  float* w;   /* device FP32 parameters, kpp_mlp_param_count(in_pad) floats */
  void* wb;   /* device BF16 parameters, same layout */
} kpp_mlp;
size_t kpp_mlp_param_count(int in_pad);
int kpp_mlp_param_offsets(int in_pad, size_t w_off[4], size_t b_off[4]);
/* refresh wb from w */
int kpp_mlp_sync_bf16(const kpp_mlp* net, void* stream);
/* features X[n][in_pad] (BF16): (log10(max(C,cfloor)) - mu)/sd per species,
 * then TEMP and SUN (0 if sun is NULL) standardised the same way, then zeros */
int kpp_mlp_features(int n, int ld, int nspec, const double* C, const double* temp,
                     const double* sun, const float* mu, const float* sd, double cfloor,
                     int in_pad, void* X, void* stream);
size_t kpp_mlp_work_bytes(int n);
/* out[n][16] = network outputs (FP32) */
int kpp_mlp_forward(const kpp_mlp* net, const void* X, int n, float* out, void* work,
                    size_t work_bytes, void* stream);
/* gradients of the mean loss  w*(o0-y)^2 + cross-entropy(o[1..8], bucket)
 * in the parameter layout (FP32); *loss (device) = summed loss */
int kpp_mlp_grad(const kpp_mlp* net, const void* X, const float* y, const int* bucket,
                 const float* wreg, int n, float* grad, float* loss, void* work,
                 size_t work_bytes, void* stream);
/* kpp_mlp_grad followed by one Adam step (beta 0.9/0.999, eps 1e-8) on w, wb */
int kpp_mlp_train_step(const kpp_mlp* net, const void* X, const float* y, const int* bucket,
                       const float* wreg, int n, float* grad, float* m1, float* m2, int step,
                       float lr, float* loss, void* work, size_t work_bytes, void* stream);
/* prediction: hstart[n] = 10^(o0*sd_y+mu_y) clamped to [hmin,hmax], bucket[n],
 * order[n] = cells sorted by expected step count, most steps first (stable):
 * sum_j softmax(o[1..8])_j * centres[j] on 64 levels between centres[0] and
 * centres[7] (host array of 8 increasing floats, the mean step count of each
 * bucket), or the expected bucket index if centres is NULL;
 * any output may be NULL */
size_t kpp_mlp_schedule_work_bytes(int n);
int kpp_mlp_schedule(const kpp_mlp* net, const void* X, int n, float mu_y, float sd_y,
                     double hmin, double hmax, const float* centres, double* hstart, int* bucket,
                     int* order, void* work, size_t work_bytes, void* stream);

/* The scheduler as one object for C hosts: network, input standardisation
 * and step scale loaded from a weights file (the flat image kpp_amd/learned.py
 * writes beside a trained .npz; the build exports <lib dir>/<mech>_step_mlp.kppw).
 * kpp_sched_run writes, for the cells of a device batch, hstart[ncells] (the
 * predicted first step times the trained scale, clamped to [hmin, hmax]) and
 * order[ncells] (most expected steps first); either may be NULL.  Allocates
 * its work buffers on first use. */
typedef struct kpp_sched kpp_sched;
int kpp_sched_load(const char* path, const char* mech, kpp_sched** out);
void kpp_sched_destroy(kpp_sched* s);
int kpp_sched_run(kpp_sched* s, const kpp_batch* b, double hmin, double hmax, double* hstart,
                  int* order, void* stream);

#ifdef __cplusplus
}
#endif
#endif
