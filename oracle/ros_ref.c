/* ros_ref.c — CPU oracle for the batched Rosenbrock solver.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker or the reported CPU baseline, never as the product path.
 *
 * A reentrant restatement of KPP's C Rosenbrock path, one cell per call:
 *   Rosenbrock             int/rosenbrock.c:137-446 (parameter defaults,
 *                          checks and error codes, incl. the IPAR[3]==0
 *                          quirk of :303-310)
 *   RosenbrockIntegrator   int/rosenbrock.c:450-627
 *   ros_ErrorNorm          int/rosenbrock.c:632-659
 *   ros_FunTimeDerivative  int/rosenbrock.c:663-682
 *   ros_PrepareMatrix      int/rosenbrock.c:686-737
 *   Ros2/Ros3/Ros4/Rodas3/Rodas3_1/Rodas4 tables  int/rosenbrock.c:779-1222
 *   Fun/Jac/Decomp/Solve templates               int/rosenbrock.c:1227-1296
 *   KppDecomp              util/sutil.c:2-23
 *   WCOPY/WAXPY/WSCAL/WLAMCH  util/blas.c:18-180
 *   Update_SUN             util/UpdateSun.c:1-20
 * The model functions (Fun, Jac_SP, KppSolve, Update_RCONST, rate laws of
 * util/UserRateLaws.c) come from the mechanism compiler's C output, included
 * through KPP_MODEL_HEADER.
 *
 * Global state of the reference (statistics counters, TIME, SUN, RCONST,
 * FIX) lives in a per-cell context, so cells run in parallel under OpenMP.
 * Two documented extensions of the reference's driver: SUN can be held
 * fixed per cell (sun_mode 0: Update_SUN is bypassed, as the fixed-sun
 * workloads specify), and the rate constants can be supplied by the caller
 * (rconst != NULL) instead of computed, to compare arithmetic bit for bit
 * with a device whose libm differs.
 *
 * Parity pins: see tests/test_oracle_*.py (golden vectors from the
 * reference's documentation, order conditions, closed forms, dense LU,
 * scipy reference integrations).
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include KPP_MODEL_HEADER
#ifndef KPP_FUN_SPLIT
#define KPP_FUN_SPLIT 0
#endif

#define MAX(a, b) (((a) >= (b)) ? (a) : (b))
#define MIN(b, c) (((b) < (c)) ? (b) : (c))
#define ABS(x) (((x) >= 0) ? (x) : (-x))
#define SQRT(d) (pow((d), 0.5))
#define ZERO 0.0
#define ONE 1.0
#define HALF 0.5
#define DeltaMin 1.0e-6
#define Smax 6
/* codes beyond the reference's -1..-8 (see the reject branch below) */
#define IERR_NONFINITE (-9)
#define IERR_REJECTS (-10)
#define MAX_REJECT 100000

typedef struct {
  int Nfun, Njac, Nstp, Nacc, Nrej, Ndec, Nsol, Nsng;
} ros_stats;

typedef struct {
  double TEMP, CFACTOR, SUN;
  int sun_mode; /* 0: SUN fixed per cell; 1: SUN from Update_SUN(TIME) */
  double FIX[KPP_NFIX > 0 ? KPP_NFIX : 1];
  double RCONST[KPP_NREACT];
  int rates_given;
  ros_stats st;
  double Hacc1; /* step size of the first accepted step of the call (step log), 0 if none */
  const double* Yacc; /* the accepted state (the integrator's Y): what rate code reading
                         concentrations sees, as KPP's global C */
  int fun_split; /* F90 flavour of a #FUNCTION SPLIT mechanism: FunTemplate calls Fun_SPLIT */
} ros_cell;

typedef struct {
  int S;
  double A[Smax * (Smax - 1) / 2], C[Smax * (Smax - 1) / 2];
  double M[Smax], E[Smax], Alpha[Smax], Gamma[Smax], ELO;
  char NewF[Smax];
} ros_method;

/* ---------- util/blas.c ---------- */
static void WCOPY(int N, const double* X, double* Y) {
  for (int i = 0; i < N; i++) Y[i] = X[i];
}
static void WAXPY(int N, double Alpha, const double* X, double* Y) {
  if (Alpha == ZERO) return;
  for (int i = 0; i < N; i++) Y[i] = Y[i] + Alpha * X[i];
}
static void WSCAL(int N, double Alpha, double* X) {
  if (Alpha == ONE) return;
  if (Alpha == -ONE) {
    for (int i = 0; i < N; i++) X[i] = -X[i];
  } else if (Alpha == ZERO) {
    for (int i = 0; i < N; i++) X[i] = ZERO;
  } else {
    for (int i = 0; i < N; i++) X[i] = Alpha * X[i];
  }
}
static double WLAMCH_ADD(double A, double B) { return A + B; }
double ros_ref_wlamch(void) {
  /* halving loop of util/blas.c:150-180 */
  volatile double Eps = pow(HALF, 16), Suma;
  int i;
  for (i = 17; i <= 80; i++) {
    Eps = Eps * HALF;
    Suma = WLAMCH_ADD(ONE, Eps);
    if (Suma <= ONE) break;
  }
  if (i == 80) return 2.2e-16;
  return Eps * 2.0;
}

/* ---------- util/UpdateSun.c ---------- */
double ros_ref_update_sun(double TIME) {
  double SunRise = 4.5, SunSet = 19.5, Thour, Tlocal, Ttmp;
  const double PI = 3.14159265358979;
  Thour = TIME / 3600.0;
  Tlocal = Thour - ((int)Thour / 24) * 24;
  if ((Tlocal >= SunRise) && (Tlocal <= SunSet)) {
    Ttmp = (2.0 * Tlocal - SunRise - SunSet) / (SunSet - SunRise);
    if (Ttmp > 0) Ttmp = Ttmp * Ttmp;
    else Ttmp = -Ttmp * Ttmp;
    return (1.0 + cos(PI * Ttmp)) / 2.0;
  }
  return 0.0;
}

/* ---------- util/sutil.c KppDecomp ---------- */
static int KppDecomp(double* JVS) {
  double W[KPP_NVAR];
  double a;
  int k, kk, j, jj;
  for (k = 0; k < KPP_NVAR; k++) {
    if (JVS[kpp_LU_DIAG[k]] == 0.0) return k + 1;
    for (kk = kpp_LU_CROW[k]; kk < kpp_LU_CROW[k + 1]; kk++) W[kpp_LU_ICOL[kk]] = JVS[kk];
    for (kk = kpp_LU_CROW[k]; kk < kpp_LU_DIAG[k]; kk++) {
      j = kpp_LU_ICOL[kk];
      a = -W[j] / JVS[kpp_LU_DIAG[j]];
      W[j] = -a;
      for (jj = kpp_LU_DIAG[j] + 1; jj < kpp_LU_CROW[j + 1]; jj++) W[kpp_LU_ICOL[jj]] += a * JVS[jj];
    }
    for (kk = kpp_LU_CROW[k]; kk < kpp_LU_CROW[k + 1]; kk++) JVS[kk] = W[kpp_LU_ICOL[kk]];
  }
  return 0;
}

/* ---------- templates (int/rosenbrock.c:1227-1296) ---------- */
static void update_rates(ros_cell* c, double T) {
  if (c->rates_given) return;
  if (c->sun_mode == 1) c->SUN = ros_ref_update_sun(T);
  kpp_Update_RCONST(c->TEMP, c->SUN, c->CFACTOR, T, c->Yacc, c->FIX, c->RCONST);
}
static void FunTemplate(ros_cell* c, double T, const double* Y, double* Ydot) {
  update_rates(c, T);
#if KPP_FUN_SPLIT
  /* KPP_FUN_OR_FUN_SPLIT of the F90 integrators (src/code.c:279-280, 361-362) */
  if (c->fun_split) {
    kpp_Fun_SPLIT(Y, c->FIX, c->RCONST, Ydot);
    c->st.Nfun++;
    return;
  }
#endif
  kpp_Fun(Y, c->FIX, c->RCONST, Ydot);
  c->st.Nfun++;
}
static void JacTemplate(ros_cell* c, double T, const double* Y, double* Jcb) {
  update_rates(c, T);
  kpp_Jac_SP(Y, c->FIX, c->RCONST, Jcb);
  c->st.Njac++;
}

/* ---------- method tables (int/rosenbrock.c:779-1222) ---------- */
static void Ros2(ros_method* m) {
  double g = 1.70710678118655;
  m->S = 2;
  m->A[0] = 1.0 / g;
  m->C[0] = (-2.0) / g;
  m->NewF[0] = 1; m->NewF[1] = 1;
  m->M[0] = (3.0) / (2.0 * g);
  m->M[1] = (1.0) / (2.0 * g);
  m->E[0] = 1.0 / (2.0 * g);
  m->E[1] = 1.0 / (2.0 * g);
  m->ELO = 2.0;
  m->Alpha[0] = 0.0; m->Alpha[1] = 1.0;
  m->Gamma[0] = g; m->Gamma[1] = -g;
}
static void Ros3(ros_method* m) {
  m->S = 3;
  m->A[0] = 1.0; m->A[1] = 1.0; m->A[2] = 0.0;
  m->C[0] = -1.0156171083877702091975600115545;
  m->C[1] = 4.0759956452537699824805835358067;
  m->C[2] = 9.2076794298330791242156818474003;
  m->NewF[0] = 1; m->NewF[1] = 1; m->NewF[2] = 0;
  m->M[0] = 1.0;
  m->M[1] = 6.1697947043828245592553615689730;
  m->M[2] = -0.4277225654321857332623837380651;
  m->E[0] = 0.5;
  m->E[1] = -2.9079558716805469821718236208017;
  m->E[2] = 0.2235406989781156962736090927619;
  m->ELO = 3.0;
  m->Alpha[0] = 0.0;
  m->Alpha[1] = 0.43586652150845899941601945119356;
  m->Alpha[2] = 0.43586652150845899941601945119356;
  m->Gamma[0] = 0.43586652150845899941601945119356;
  m->Gamma[1] = 0.24291996454816804366592249683314;
  m->Gamma[2] = 2.1851380027664058511513169485832;
}
static void Ros4(ros_method* m) {
  m->S = 4;
  m->A[0] = 0.2000000000000000e+01;
  m->A[1] = 0.1867943637803922e+01;
  m->A[2] = 0.2344449711399156;
  m->A[3] = m->A[1];
  m->A[4] = m->A[2];
  m->A[5] = 0.0;
  m->C[0] = -0.7137615036412310e+01;
  m->C[1] = 0.2580708087951457e+01;
  m->C[2] = 0.6515950076447975;
  m->C[3] = -0.2137148994382534e+01;
  m->C[4] = -0.3214669691237626;
  m->C[5] = -0.6949742501781779;
  m->NewF[0] = 1; m->NewF[1] = 1; m->NewF[2] = 1; m->NewF[3] = 0;
  m->M[0] = 0.2255570073418735e+01;
  m->M[1] = 0.2870493262186792;
  m->M[2] = 0.4353179431840180;
  m->M[3] = 0.1093502252409163e+01;
  m->E[0] = -0.2815431932141155;
  m->E[1] = -0.7276199124938920e-01;
  m->E[2] = -0.1082196201495311;
  m->E[3] = -0.1093502252409163e+01;
  m->ELO = 4.0;
  m->Alpha[0] = 0.0;
  m->Alpha[1] = 0.1145640000000000e+01;
  m->Alpha[2] = 0.6552168638155900;
  m->Alpha[3] = m->Alpha[2];
  m->Gamma[0] = 0.5728200000000000;
  m->Gamma[1] = -0.1769193891319233e+01;
  m->Gamma[2] = 0.7592633437920482;
  m->Gamma[3] = -0.1049021087100450;
}
static void Rodas3(ros_method* m) {
  m->S = 4;
  m->A[0] = 0.0; m->A[1] = 2.0; m->A[2] = 0.0; m->A[3] = 2.0; m->A[4] = 0.0; m->A[5] = 1.0;
  m->C[0] = 4.0; m->C[1] = 1.0; m->C[2] = -1.0; m->C[3] = 1.0; m->C[4] = -1.0;
  m->C[5] = -2.66666666666667;
  m->NewF[0] = 1; m->NewF[1] = 0; m->NewF[2] = 1; m->NewF[3] = 1;
  m->M[0] = 2.0; m->M[1] = 0.0; m->M[2] = 1.0; m->M[3] = 1.0;
  m->E[0] = 0.0; m->E[1] = 0.0; m->E[2] = 0.0; m->E[3] = 1.0;
  m->ELO = 3.0;
  m->Alpha[0] = 0.0; m->Alpha[1] = 0.0; m->Alpha[2] = 1.0; m->Alpha[3] = 1.0;
  m->Gamma[0] = 0.5; m->Gamma[1] = 1.5; m->Gamma[2] = 0.0; m->Gamma[3] = 0.0;
}
static void Rodas3_1(ros_method* m) {
  m->S = 4;
  m->A[0] = 0.000000000000000; m->A[1] = 0.646601929740551; m->A[2] = 0.409567801987914;
  m->A[3] = 0.646601929740551; m->A[4] = 0.409567801987914; m->A[5] = 1.000000000000000;
  m->C[0] = 4.198495621784201; m->C[1] = 3.711590161613010; m->C[2] = -1.787771994729384;
  m->C[3] = 4.458898153216104; m->C[4] = -2.024095448516552; m->C[5] = -2.626700600119396;
  m->NewF[0] = 1; m->NewF[1] = 0; m->NewF[2] = 1; m->NewF[3] = 1;
  m->M[0] = 0.646601929740551; m->M[1] = 0.409567801987914; m->M[2] = 1.0; m->M[3] = 1.0;
  m->E[0] = 0.0; m->E[1] = 0.0; m->E[2] = 0.0; m->E[3] = 1.0;
  m->ELO = 3.0;
  m->Alpha[0] = 0.0; m->Alpha[1] = 0.0; m->Alpha[2] = 1.0; m->Alpha[3] = 1.0;
  m->Gamma[0] = 0.515; m->Gamma[1] = 1.628546001287715; m->Gamma[2] = 0.0; m->Gamma[3] = 0.0;
}
static void Rodas4(ros_method* m) {
  m->S = 6;
  m->Alpha[0] = 0.000; m->Alpha[1] = 0.386; m->Alpha[2] = 0.210;
  m->Alpha[3] = 0.630; m->Alpha[4] = 1.000; m->Alpha[5] = 1.000;
  m->Gamma[0] = 0.2500000000000000; m->Gamma[1] = -0.1043000000000000;
  m->Gamma[2] = 0.1035000000000000; m->Gamma[3] = -0.3620000000000023e-01;
  m->Gamma[4] = 0.0; m->Gamma[5] = 0.0;
  m->A[0] = 0.1544000000000000e+01;
  m->A[1] = 0.9466785280815826;
  m->A[2] = 0.2557011698983284;
  m->A[3] = 0.3314825187068521e+01;
  m->A[4] = 0.2896124015972201e+01;
  m->A[5] = 0.9986419139977817;
  m->A[6] = 0.1221224509226641e+01;
  m->A[7] = 0.6019134481288629e+01;
  m->A[8] = 0.1253708332932087e+02;
  m->A[9] = -0.6878860361058950;
  m->A[10] = m->A[6]; m->A[11] = m->A[7]; m->A[12] = m->A[8]; m->A[13] = m->A[9];
  m->A[14] = 1.0;
  m->C[0] = -0.5668800000000000e+01;
  m->C[1] = -0.2430093356833875e+01;
  m->C[2] = -0.2063599157091915;
  m->C[3] = -0.1073529058151375;
  m->C[4] = -0.9594562251023355e+01;
  m->C[5] = -0.2047028614809616e+02;
  m->C[6] = 0.7496443313967647e+01;
  m->C[7] = -0.1024680431464352e+02;
  m->C[8] = -0.3399990352819905e+02;
  m->C[9] = 0.1170890893206160e+02;
  m->C[10] = 0.8083246795921522e+01;
  m->C[11] = -0.7981132988064893e+01;
  m->C[12] = -0.3152159432874371e+02;
  m->C[13] = 0.1631930543123136e+02;
  m->C[14] = -0.6058818238834054e+01;
  m->M[0] = m->A[6]; m->M[1] = m->A[7]; m->M[2] = m->A[8]; m->M[3] = m->A[9];
  m->M[4] = 1.0; m->M[5] = 1.0;
  for (int i = 0; i < 5; i++) m->E[i] = 0.0;
  m->E[5] = 1.0;
  for (int i = 0; i < 6; i++) m->NewF[i] = 1;
  m->ELO = 4.0;
}

/* Fill a method table; code as IPAR[3] (1..5, 7). Returns 0 if unknown. */
int ros_ref_method(int code, int* S, double* A, double* C, double* M, double* E, double* Alpha,
                   double* Gamma, double* ELO, int* NewF) {
  ros_method m;
  memset(&m, 0, sizeof m);
  switch (code) {
    case 1: Ros2(&m); break;
    case 2: Ros3(&m); break;
    case 3: Ros4(&m); break;
    case 4: Rodas3(&m); break;
    case 5: Rodas4(&m); break;
    case 7: Rodas3_1(&m); break;
    default: return 0;
  }
  *S = m.S;
  for (int i = 0; i < Smax * (Smax - 1) / 2; i++) { A[i] = m.A[i]; C[i] = m.C[i]; }
  for (int i = 0; i < Smax; i++) {
    M[i] = m.M[i]; E[i] = m.E[i]; Alpha[i] = m.Alpha[i]; Gamma[i] = m.Gamma[i];
    NewF[i] = m.NewF[i];
  }
  *ELO = m.ELO;
  return 1;
}

/* ---------- ros_ErrorNorm (:632-659) ---------- */
static double ros_ErrorNorm(const double* Y, const double* Ynew, const double* Yerr,
                            const double* AbsTol, const double* RelTol, int VectorTol) {
  double Err = ZERO, Scale, Ymax;
  for (int i = 0; i < KPP_NVAR; i++) {
    Ymax = MAX(ABS(Y[i]), ABS(Ynew[i]));
    if (VectorTol) Scale = AbsTol[i] + RelTol[i] * Ymax;
    else Scale = AbsTol[0] + RelTol[0] * Ymax;
    Err = Err + (Yerr[i] * Yerr[i]) / (Scale * Scale);
  }
  return SQRT(Err / (double)KPP_NVAR);
}

/* ---------- ros_FunTimeDerivative (:663-682) ---------- */
static void ros_FunTimeDerivative(ros_cell* c, double T, double Roundoff, const double* Y,
                                  const double* Fcn0, double* dFdT) {
  double Delta = SQRT(Roundoff) * MAX(DeltaMin, ABS(T));
  FunTemplate(c, T + Delta, Y, dFdT);
  WAXPY(KPP_NVAR, (-ONE), Fcn0, dFdT);
  WSCAL(KPP_NVAR, (ONE / Delta), dFdT);
}

/* ---------- ros_PrepareMatrix (:686-737) ---------- */
static int ros_PrepareMatrix(ros_cell* c, double* H, int Direction, double gam,
                             const double* Jac0, double* Ghimj) {
  int Nconsecutive = 0, ising;
  double ghinv;
  for (;;) {
    WCOPY(KPP_LU_NONZERO, Jac0, Ghimj);
    WSCAL(KPP_LU_NONZERO, (-ONE), Ghimj);
    ghinv = ONE / (Direction * (*H) * gam);
    for (int i = 0; i < KPP_NVAR; i++) Ghimj[kpp_LU_DIAG[i]] = Ghimj[kpp_LU_DIAG[i]] + ghinv;
    ising = KppDecomp(Ghimj);
    c->st.Ndec++;
    if (ising == 0) return 0;
    c->st.Nsng++;
    Nconsecutive++;
    if (Nconsecutive <= 5) *H = (*H) * HALF;
    else return 1;
  }
}

/* ---------- RosenbrockIntegrator (:450-627) ---------- */
static int RosenbrockIntegrator(ros_cell* c, double* Y, double Tstart, double Tend,
                                const double* AbsTol, const double* RelTol, const ros_method* m,
                                int Autonomous, int VectorTol, int Max_no_steps, double Roundoff,
                                double Hmin, double Hmax, double Hstart, double FacMin,
                                double FacMax, double FacRej, double FacSafe, double* Texit,
                                double* Hexit) {
  double Ynew[KPP_NVAR], Fcn0[KPP_NVAR], Fcn[KPP_NVAR], dFdT[KPP_NVAR];
  double Jac0[KPP_LU_NONZERO], Ghimj[KPP_LU_NONZERO];
  double K[KPP_NVAR * Smax];
  double H, T, Hnew, HC, HG, Fac, Tau, Err, Yerr[KPP_NVAR];
  int Direction, ioffset, j, istage;
  int RejectLastH, RejectMoreH;
  int Nrun = 0; /* rejected attempts since the last accepted step */
  const int ros_S = m->S;

  T = Tstart;
  *Hexit = 0.0;
  c->Yacc = Y;
  H = MIN(Hstart, Hmax);
  if (ABS(H) <= 10.0 * Roundoff) H = DeltaMin;
  Direction = (Tend >= Tstart) ? +1 : -1;
  RejectLastH = 0;
  RejectMoreH = 0;

  while (((Direction > 0) && ((T - Tend) + Roundoff <= ZERO)) ||
         ((Direction < 0) && ((Tend - T) + Roundoff <= ZERO))) {
    if (c->st.Nstp > Max_no_steps) { *Texit = T; return -6; }
    if (((T + 0.1 * H) == T) || (H <= Roundoff)) { *Texit = T; return -7; }
    *Hexit = H;
    H = MIN(H, ABS(Tend - T));
    FunTemplate(c, T, Y, Fcn0);
    if (!Autonomous) ros_FunTimeDerivative(c, T, Roundoff, Y, Fcn0, dFdT);
    JacTemplate(c, T, Y, Jac0);
    for (;;) {
      if (ros_PrepareMatrix(c, &H, Direction, m->Gamma[0], Jac0, Ghimj)) {
        *Texit = T;
        return -8;
      }
      for (istage = 1; istage <= ros_S; istage++) {
        ioffset = KPP_NVAR * (istage - 1);
        if (istage == 1) {
          WCOPY(KPP_NVAR, Fcn0, Fcn);
        } else if (m->NewF[istage - 1]) {
          WCOPY(KPP_NVAR, Y, Ynew);
          for (j = 1; j <= istage - 1; j++)
            WAXPY(KPP_NVAR, m->A[(istage - 1) * (istage - 2) / 2 + j - 1], &K[KPP_NVAR * (j - 1)],
                  Ynew);
          Tau = T + m->Alpha[istage - 1] * Direction * H;
          FunTemplate(c, Tau, Ynew, Fcn);
        }
        WCOPY(KPP_NVAR, Fcn, &K[ioffset]);
        for (j = 1; j <= istage - 1; j++) {
          HC = m->C[(istage - 1) * (istage - 2) / 2 + j - 1] / (Direction * H);
          WAXPY(KPP_NVAR, HC, &K[KPP_NVAR * (j - 1)], &K[ioffset]);
        }
        if ((!Autonomous) && (m->Gamma[istage - 1])) {
          HG = Direction * H * m->Gamma[istage - 1];
          WAXPY(KPP_NVAR, HG, dFdT, &K[ioffset]);
        }
        kpp_KppSolve(Ghimj, &K[ioffset]);
        c->st.Nsol++;
      }
      WCOPY(KPP_NVAR, Y, Ynew);
      for (j = 1; j <= ros_S; j++) WAXPY(KPP_NVAR, m->M[j - 1], &K[KPP_NVAR * (j - 1)], Ynew);
      WSCAL(KPP_NVAR, ZERO, Yerr);
      for (j = 1; j <= ros_S; j++) WAXPY(KPP_NVAR, m->E[j - 1], &K[KPP_NVAR * (j - 1)], Yerr);
      Err = ros_ErrorNorm(Y, Ynew, Yerr, AbsTol, RelTol, VectorTol);
      Fac = MIN(FacMax, MAX(FacMin, FacSafe / pow(Err, ONE / m->ELO)));
      Hnew = H * Fac;
      c->st.Nstp++;
      if ((Err <= ONE) || (H <= Hmin)) {
        if (c->Hacc1 == ZERO) c->Hacc1 = H;
        c->st.Nacc++;
        WCOPY(KPP_NVAR, Ynew, Y);
        T += Direction * H;
        Hnew = MAX(Hmin, MIN(Hnew, Hmax));
        if (RejectLastH) Hnew = MIN(Hnew, H);
        RejectLastH = 0;
        RejectMoreH = 0;
        H = Hnew;
        Nrun = 0;
        break;
      } else {
        if (c->st.Nacc >= 1) c->st.Nrej++;
        if (RejectMoreH) Hnew = H * FacRej;
        RejectMoreH = RejectLastH;
        RejectLastH = 1;
        H = Hnew;
        /* Deviation from int/rosenbrock.c:537-619, whose reject loop has no
         * bound: a NaN error estimate or a non-finite step size ends the
         * cell with -9, more than MAX_REJECT rejections in a row with -10
         * (the GPU integrator must end every cell). */
        if (Err != Err || !(ABS(Hnew) <= DBL_MAX)) { *Texit = T; return IERR_NONFINITE; }
        if (++Nrun > MAX_REJECT) { *Texit = T; return IERR_REJECTS; }
      }
    }
  }
  *Texit = T;
  return 1;
}

/* ---------- starting step estimate (library extension, IPAR[4] = 1) ----------
 * Not in int/rosenbrock.c (whose default start is DeltaMin): Hairer, Norsett
 * & Wanner, Solving ODEs I, II.4, with the weights of ros_ErrorNorm
 * sc_i = AbsTol_i + RelTol_i |y0_i|; two function evaluations, counted. */
static double ros_InitialStep(ros_cell* c, double T, int Direction, double* Y,
                              const double* AbsTol, const double* RelTol, int VectorTol,
                              double ELO) {
  double F0[KPP_NVAR], Y1[KPP_NVAR], F1[KPP_NVAR];
  double d0 = ZERO, d1 = ZERO, d2 = ZERO, h0, h1, dm, dh, w, a, b;
  c->Yacc = Y;
  FunTemplate(c, T, Y, F0);
  for (int i = 0; i < KPP_NVAR; i++) {
    w = VectorTol ? (AbsTol[i] + RelTol[i] * ABS(Y[i])) : (AbsTol[0] + RelTol[0] * ABS(Y[i]));
    a = Y[i] / w;
    b = F0[i] / w;
    d0 = d0 + a * a;
    d1 = d1 + b * b;
  }
  d0 = sqrt(d0 / (double)KPP_NVAR);
  d1 = sqrt(d1 / (double)KPP_NVAR);
  h0 = (d0 < 1.0e-5 || d1 < 1.0e-5) ? 1.0e-6 : 0.01 * (d0 / d1);
  dh = (double)Direction * h0;
  for (int i = 0; i < KPP_NVAR; i++) Y1[i] = Y[i] + dh * F0[i];
  FunTemplate(c, T + dh, Y1, F1);
  for (int i = 0; i < KPP_NVAR; i++) {
    w = VectorTol ? (AbsTol[i] + RelTol[i] * ABS(Y[i])) : (AbsTol[0] + RelTol[0] * ABS(Y[i]));
    a = (F1[i] - F0[i]) / w;
    d2 = d2 + a * a;
  }
  d2 = sqrt(d2 / (double)KPP_NVAR) / h0;
  dm = MAX(d1, d2);
  h1 = (dm <= 1.0e-15) ? MAX(1.0e-6, 1.0e-3 * h0) : pow(0.01 / dm, ONE / ELO);
  return MIN(100.0 * h0, h1);
}

/* ---------- Rosenbrock (:137-446) ---------- */
static int Rosenbrock(ros_cell* c, double* Y, double Tstart, double Tend, const double* AbsTol,
                      const double* RelTol, double* RPAR, int* IPAR) {
  ros_method m;
  int Max_no_steps, IERR, UplimTol, Method, Autonomous, VectorTol;
  double Roundoff, FacMin, FacMax, FacRej, FacSafe, Hmin, Hmax, Hstart, Hexit = 0, Texit = 0;

  c->st.Nfun = IPAR[10];
  c->st.Njac = IPAR[11];
  c->st.Nstp = IPAR[12];
  c->st.Nacc = IPAR[13];
  c->st.Nrej = IPAR[14];
  c->st.Ndec = IPAR[15];
  c->st.Nsol = IPAR[16];
  c->st.Nsng = IPAR[17];
  Autonomous = !(IPAR[0] == 0);
  if (IPAR[1] == 0) { VectorTol = 1; UplimTol = KPP_NVAR; }
  else { VectorTol = 0; UplimTol = 1; }
  Max_no_steps = (IPAR[2] == 0) ? 100000 : IPAR[2];
  if (Max_no_steps < 0) return -1;
  Method = (IPAR[3] == 0) ? 3 : IPAR[3];
  if ((IPAR[3] < 1) || (IPAR[3] > 5)) return -2;
  Roundoff = ros_ref_wlamch();
  Hmin = RPAR[0];
  if (RPAR[0] < ZERO) return -3;
  if (RPAR[1] == ZERO) Hmax = ABS(Tend - Tstart);
  else Hmax = MIN(ABS(RPAR[1]), ABS(Tend - Tstart));
  if (RPAR[1] < ZERO) return -3;
  if (RPAR[2] == ZERO) Hstart = MAX(Hmin, DeltaMin);
  else Hstart = MIN(ABS(RPAR[2]), ABS(Tend - Tstart));
  if (RPAR[2] < ZERO) return -3;
  FacMin = (RPAR[3] == ZERO) ? 0.2 : RPAR[3];
  if (RPAR[3] < ZERO) return -4;
  FacMax = (RPAR[4] == ZERO) ? 6.0 : RPAR[4];
  if (RPAR[4] < ZERO) return -4;
  FacRej = (RPAR[5] == ZERO) ? 0.1 : RPAR[5];
  if (RPAR[5] < ZERO) return -4;
  FacSafe = (RPAR[6] == ZERO) ? 0.9 : RPAR[6];
  if (RPAR[6] < ZERO) return -4;
  for (int i = 0; i < UplimTol; i++)
    if ((AbsTol[i] <= ZERO) || (RelTol[i] <= 10.0 * Roundoff) || (RelTol[i] >= ONE)) return -5;
  memset(&m, 0, sizeof m);
  switch (Method) {
    case 1: Ros2(&m); break;
    case 2: Ros3(&m); break;
    case 3: Ros4(&m); break;
    case 4: Rodas3(&m); break;
    case 5: Rodas4(&m); break;
    case 7: Rodas3_1(&m); break;
    default: return -2;
  }
  if (RPAR[2] == ZERO && IPAR[4] == 1) {
    double h = ros_InitialStep(c, Tstart, (Tend >= Tstart) ? +1 : -1, Y, AbsTol, RelTol,
                               VectorTol, m.ELO);
    Hstart = MAX(Hmin, MIN(h, ABS(Tend - Tstart)));
  }
  IERR = RosenbrockIntegrator(c, Y, Tstart, Tend, AbsTol, RelTol, &m, Autonomous, VectorTol,
                              Max_no_steps, Roundoff, Hmin, Hmax, Hstart, FacMin, FacMax, FacRej,
                              FacSafe, &Texit, &Hexit);
  IPAR[10] = c->st.Nfun;
  IPAR[11] = c->st.Njac;
  IPAR[12] = c->st.Nstp;
  IPAR[13] = c->st.Nacc;
  IPAR[14] = c->st.Nrej;
  IPAR[15] = c->st.Ndec;
  IPAR[16] = c->st.Nsol;
  IPAR[17] = c->st.Nsng;
  RPAR[10] = Texit;
  RPAR[11] = Hexit;
  return IERR;
}

/* ======================= F90 Rosenbrock (int/rosenbrock.f90) =======================
 * KPP's Fortran 90 integrator is a separate implementation with its own
 * method list, defaults and outputs:
 *   INTEGRATE        :43-125   ICNTRL(15) = 5 by default; ICNTRL_U overrides
 *                               where /= 0, RCNTRL_U where > 0
 *   Rosenbrock       :128-446  ICNTRL(3) map (0,4 Rodas3; 1 Ros2; 2 Ros3;
 *                               3 Ros4; 5 Rodas4; 6 Rang3; 7 Rodas3.1),
 *                               200000 steps, Hstart = MAX(Hmin, 1e-5),
 *                               -1..-5 checks
 *   ros_Integrator   :492-686  H = MIN(MAX(|Hmin|,|Hstart|),|Hmax|) signed by
 *                               the direction; array updates without the
 *                               zero-coefficient skips of the C WAXPY;
 *                               ICNTRL(16) clipping; RSTATUS(1:3) = T, H, Hnew
 *                               of the last accepted step
 *   ros_ErrorNorm    :690-748  non-passive species only, (Yerr/Scale)**2,
 *                               MAX(Err, 1e-10)
 *   Ros2 g = 1 + 1/SQRT(2); Rodas3 C(6) = -(8/3); Rang3 :1327-1394.
 * Its FunTemplate/JacTemplate (:1405-1478) call Update_SUN/Update_RCONST as
 * ICNTRL(15) selects (util/util.f90:133-159); the rate code of the supported
 * mechanisms reads concentrations from the global C, i.e. the accepted state,
 * as in the C path.  int/rosenbrock_h211b_qssa.f90 is the same integrator
 * (INTEGRATE :56-138, ros_Integrator :612-935, the norm over all species
 * :939-969) with the H211b step-size filter; its QSSA options are not
 * restated.  Deviations, both shared with the device: a NaN error
 * estimate ends the cell with -9 (gfortran's MAX would make it 1e-10 and
 * accept the step), and more than MAX_REJECT rejections in a row with -10. */
#define DeltaMinF90 1.0e-5
#define F90MAX(a, b) (((a) >= (b)) ? (a) : (b))

static void Ros2_f90(ros_method* m) {
  Ros2(m);
  double g = 1.0 + 1.0 / sqrt(2.0);
  m->A[0] = 1.0 / g;
  m->C[0] = (-2.0) / g;
  m->M[0] = (3.0) / (2.0 * g);
  m->M[1] = (1.0) / (2.0 * g);
  m->E[0] = 1.0 / (2.0 * g);
  m->E[1] = 1.0 / (2.0 * g);
  m->Gamma[0] = g;
  m->Gamma[1] = -g;
}
static void Rodas3_f90(ros_method* m) {
  Rodas3(m);
  m->C[5] = -(8.0 / 3.0);
}
static void Rang3(ros_method* m) {
  /* Rang & Angermann, BIT 45 (2005) 761-787, Tables 4.1-4.2, as tabled in
   * int/rosenbrock.f90:1348-1392 */
  m->S = 4;
  m->A[0] = 5.09052051067020; m->A[1] = 5.09052051067020; m->A[2] = 0.0;
  m->A[3] = 4.97628111010787; m->A[4] = 2.77268164715849e-02; m->A[5] = 2.29428036027904e-01;
  m->C[0] = -1.16790812312283e+01; m->C[1] = -1.64057326467367e+01;
  m->C[2] = -2.77268164715850e-01; m->C[3] = -8.38103960500476;
  m->C[4] = -8.48328409199343e-01; m->C[5] = 2.87009860433106e-01;
  m->M[0] = 5.22582761233094; m->M[1] = -5.56971148154165e-01;
  m->M[2] = 3.57979469353645e-01; m->M[3] = 1.72337398521064;
  m->E[0] = -5.16845212784040; m->E[1] = -1.26351942603842;
  m->E[2] = -1.11022302462516e-16; m->E[3] = 2.22044604925031e-16;
  m->Alpha[0] = 0.0; m->Alpha[1] = 2.21878746765329; m->Alpha[2] = 2.21878746765329;
  m->Alpha[3] = 1.55392337535788;
  m->Gamma[0] = 4.35866521508459e-01; m->Gamma[1] = -1.78292094614483;
  m->Gamma[2] = -2.46541900496934; m->Gamma[3] = -8.05529997906370e-01;
  for (int i = 0; i < 4; i++) m->NewF[i] = 1;
  m->ELO = 3.0;
}
/* the ICNTRL(3) map of int/rosenbrock.f90:321-340; 0 if unknown */
static int method_f90(int code, ros_method* m) {
  memset(m, 0, sizeof *m);
  switch (code) {
    case 0: case 4: Rodas3_f90(m); return 1;
    case 1: Ros2_f90(m); return 1;
    case 2: Ros3(m); return 1;
    case 3: Ros4(m); return 1;
    case 5: Rodas4(m); return 1;
    case 6: Rang3(m); return 1;
    case 7: Rodas3_1(m); return 1;
  }
  return 0;
}
int ros_ref_method_f90(int code, int* S, double* A, double* C, double* M, double* E,
                       double* Alpha, double* Gamma, double* ELO, int* NewF) {
  ros_method m;
  if (!method_f90(code, &m)) return 0;
  *S = m.S;
  for (int i = 0; i < Smax * (Smax - 1) / 2; i++) { A[i] = m.A[i]; C[i] = m.C[i]; }
  for (int i = 0; i < Smax; i++) {
    M[i] = m.M[i]; E[i] = m.E[i]; Alpha[i] = m.Alpha[i]; Gamma[i] = m.Gamma[i];
    NewF[i] = m.NewF[i];
  }
  *ELO = m.ELO;
  return 1;
}

typedef struct {
  const int* passive; /* [NVAR] 1 = passive (left out of the error norm), or NULL */
  int np_count;       /* NonPassiveSpc_Count */
  int nonneg;         /* ICNTRL(16) == 1 */
  int h211b;          /* int/rosenbrock_h211b_qssa.f90 with bb, kk > 0: the H211b controller */
  double bb, kk;      /* RCNTRL(15), RCNTRL(16) */
} f90_opts;

/* ros_ErrorNorm, int/rosenbrock.f90:690-748 (returns the norm before the
 * MAX with 1e-10, which the caller applies) */
static double ros_ErrorNorm_f90(const double* Y, const double* Ynew, const double* Yerr,
                                const double* AbsTol, const double* RelTol, int VectorTol,
                                const f90_opts* o) {
  double Err = ZERO, Scale, Ymax, t;
  for (int i = 0; i < KPP_NVAR; i++) {
    if (o->passive && o->passive[i]) continue;
    Ymax = F90MAX(fabs(Y[i]), fabs(Ynew[i]));
    Scale = VectorTol ? (AbsTol[i] + RelTol[i] * Ymax) : (AbsTol[0] + RelTol[0] * Ymax);
    t = Yerr[i] / Scale;
    Err = Err + t * t;
  }
  return sqrt(Err / (double)o->np_count);
}

/* ros_Integrator, int/rosenbrock.f90:492-686; rstatus[0..2] = RSTATUS(1:3) */
static int ros_Integrator_f90(ros_cell* c, double* Y, double Tstart, double Tend,
                              const double* AbsTol, const double* RelTol, const ros_method* m,
                              int Autonomous, int VectorTol, int Max_no_steps, double Roundoff,
                              double Hmin, double Hmax, double Hstart, double FacMin,
                              double FacMax, double FacRej, double FacSafe, const f90_opts* o,
                              double* rstatus) {
  double Ynew[KPP_NVAR], Fcn0[KPP_NVAR], Fcn[KPP_NVAR], dFdT[KPP_NVAR];
  double Jac0[KPP_LU_NONZERO], Ghimj[KPP_LU_NONZERO];
  double K[KPP_NVAR * Smax], Yerr[KPP_NVAR];
  double H, T, Hnew, HC, HG, Fac, Tau, Err, Delta;
  double FacOld = 1.0, ErrOld = 1.0; /* the H211b filter's memory */
  int Direction, j, istage, i;
  int RejectLastH, RejectMoreH;
  int Nrun = 0;
  const int N = KPP_NVAR, ros_S = m->S;

  T = Tstart;
  rstatus[1] = ZERO;
  c->Yacc = Y;
  H = MIN(F90MAX(fabs(Hmin), fabs(Hstart)), fabs(Hmax));
  if (fabs(H) <= 10.0 * Roundoff) H = DeltaMinF90;
  Direction = (Tend >= Tstart) ? +1 : -1;
  H = Direction * H;
  RejectLastH = 0;
  RejectMoreH = 0;

  while (((Direction > 0) && ((T - Tend) + Roundoff <= ZERO)) ||
         ((Direction < 0) && ((Tend - T) + Roundoff <= ZERO))) {
    if (c->st.Nstp > Max_no_steps) return -6;
    if (((T + 0.1 * H) == T) || (H <= Roundoff)) return -7;
    H = MIN(H, fabs(Tend - T));
    FunTemplate(c, T, Y, Fcn0);
    if (!Autonomous) {
      /* ros_FunTimeDerivative :754-778 (DeltaMin 1e-6 there) */
      Delta = sqrt(Roundoff) * F90MAX(DeltaMin, fabs(T));
      FunTemplate(c, T + Delta, Y, dFdT);
      for (i = 0; i < N; i++) dFdT[i] = dFdT[i] - Fcn0[i];
      for (i = 0; i < N; i++) dFdT[i] = dFdT[i] * (ONE / Delta);
    }
    JacTemplate(c, T, Y, Jac0);
    for (;;) {
      if (ros_PrepareMatrix(c, &H, Direction, m->Gamma[0], Jac0, Ghimj)) return -8;
      for (istage = 1; istage <= ros_S; istage++) {
        double* Ki = &K[N * (istage - 1)];
        if (istage == 1) {
          for (i = 0; i < N; i++) Fcn[i] = Fcn0[i];
        } else if (m->NewF[istage - 1]) {
          for (i = 0; i < N; i++) Ynew[i] = Y[i];
          for (j = 1; j <= istage - 1; j++) {
            const double a = m->A[(istage - 1) * (istage - 2) / 2 + j - 1];
            for (i = 0; i < N; i++) Ynew[i] = Ynew[i] + a * K[N * (j - 1) + i];
          }
          Tau = T + m->Alpha[istage - 1] * Direction * H;
          FunTemplate(c, Tau, Ynew, Fcn);
        }
        for (i = 0; i < N; i++) Ki[i] = Fcn[i];
        for (j = 1; j <= istage - 1; j++) {
          HC = m->C[(istage - 1) * (istage - 2) / 2 + j - 1] / (Direction * H);
          for (i = 0; i < N; i++) Ki[i] = Ki[i] + HC * K[N * (j - 1) + i];
        }
        if ((!Autonomous) && (m->Gamma[istage - 1] != ZERO)) {
          HG = Direction * H * m->Gamma[istage - 1];
          for (i = 0; i < N; i++) Ki[i] = Ki[i] + HG * dFdT[i];
        }
        kpp_KppSolve(Ghimj, Ki);
        c->st.Nsol++;
      }
      for (i = 0; i < N; i++) Ynew[i] = Y[i];
      for (j = 1; j <= ros_S; j++)
        for (i = 0; i < N; i++) Ynew[i] = Ynew[i] + m->M[j - 1] * K[N * (j - 1) + i];
      for (i = 0; i < N; i++) Yerr[i] = ZERO;
      for (j = 1; j <= ros_S; j++)
        for (i = 0; i < N; i++) Yerr[i] = Yerr[i] + m->E[j - 1] * K[N * (j - 1) + i];
      Err = ros_ErrorNorm_f90(Y, Ynew, Yerr, AbsTol, RelTol, VectorTol, o);
      if (Err != Err) return IERR_NONFINITE;
      Err = F90MAX(Err, 1.0e-10);
      if (o->h211b) {
        /* H211b controller (int/rosenbrock_h211b_qssa.f90:882-892, after
         * Soderlind 2003): no FacMin/FacMax bounds, memory of the previous
         * attempt's error and factor */
        Fac = pow(ONE / Err, 1.0 / (o->bb * o->kk)) * pow(ONE / ErrOld, 1.0 / (o->bb * o->kk)) *
              pow(FacOld, -1.0 / o->bb);
        FacOld = Fac;
        ErrOld = Err;
      } else {
        Fac = MIN(FacMax, F90MAX(FacMin, FacSafe / pow(Err, ONE / m->ELO)));
      }
      Hnew = H * Fac;
      c->st.Nstp++;
      if ((Err <= ONE) || (H <= Hmin)) {
        if (c->Hacc1 == ZERO) c->Hacc1 = H;
        c->st.Nacc++;
        if (o->nonneg)
          for (i = 0; i < N; i++) Y[i] = F90MAX(Ynew[i], ZERO);
        else
          for (i = 0; i < N; i++) Y[i] = Ynew[i];
        T = T + Direction * H;
        Hnew = F90MAX(Hmin, MIN(Hnew, Hmax));
        if (RejectLastH) Hnew = MIN(Hnew, H);
        rstatus[1] = H;
        rstatus[2] = Hnew;
        rstatus[0] = T;
        RejectLastH = 0;
        RejectMoreH = 0;
        H = Hnew;
        Nrun = 0;
        break;
      } else {
        if (RejectMoreH) Hnew = H * FacRej;
        RejectMoreH = RejectLastH;
        RejectLastH = 1;
        H = Hnew;
        if (c->st.Nacc >= 1) c->st.Nrej++;
        if (!(fabs(Hnew) <= DBL_MAX)) return IERR_NONFINITE;
        if (++Nrun > MAX_REJECT) return IERR_REJECTS;
      }
    }
  }
  return 1;
}

/* Rosenbrock, int/rosenbrock.f90:128-446: ICNTRL/RCNTRL as merged by
 * INTEGRATE; istatus[0..7], rstatus[0..2] */
static int Rosenbrock_f90(ros_cell* c, double* Y, double Tstart, double Tend,
                          const double* AbsTol, const double* RelTol, const double* RCNTRL,
                          const int* ICNTRL, const f90_opts* o, int* istatus, double* rstatus) {
  ros_method m;
  int Max_no_steps, UplimTol, Autonomous, VectorTol, IERR;
  double Roundoff, Hmin, Hmax, Hstart, FacMin, FacMax, FacRej, FacSafe;
  const double span = fabs(Tend - Tstart);
  memset(&c->st, 0, sizeof c->st);
  for (int i = 0; i < 3; i++) rstatus[i] = ZERO;
  Autonomous = !(ICNTRL[0] == 0);
  if (ICNTRL[1] == 0) { VectorTol = 1; UplimTol = KPP_NVAR; }
  else { VectorTol = 0; UplimTol = 1; }
  IERR = 0;
  if (!method_f90(ICNTRL[2], &m)) IERR = -2;
  if (!IERR) {
    if (ICNTRL[3] == 0) Max_no_steps = 200000;
    else if (ICNTRL[3] > 0) Max_no_steps = ICNTRL[3];
    else IERR = -1;
  }
  Roundoff = 2.220446049250313e-16; /* EPSILON(0.0_dp) = 2^-52 */
  if (!IERR) {
    if (RCNTRL[0] == ZERO) Hmin = ZERO;
    else if (RCNTRL[0] > ZERO) Hmin = RCNTRL[0];
    else IERR = -3;
  }
  if (!IERR) {
    if (RCNTRL[1] == ZERO) Hmax = span;
    else if (RCNTRL[1] > ZERO) Hmax = MIN(fabs(RCNTRL[1]), span);
    else IERR = -3;
  }
  if (!IERR) {
    if (RCNTRL[2] == ZERO) Hstart = F90MAX(Hmin, DeltaMinF90);
    else if (RCNTRL[2] > ZERO) Hstart = MIN(fabs(RCNTRL[2]), span);
    else IERR = -3;
  }
  if (!IERR) {
    const double def[4] = {0.2, 6.0, 0.1, 0.9};
    double* dst[4] = {&FacMin, &FacMax, &FacRej, &FacSafe};
    for (int k = 0; k < 4 && !IERR; k++) {
      if (RCNTRL[3 + k] == ZERO) *dst[k] = def[k];
      else if (RCNTRL[3 + k] > ZERO) *dst[k] = RCNTRL[3 + k];
      else IERR = -4;
    }
  }
  if (!IERR)
    for (int i = 0; i < UplimTol; i++)
      if ((AbsTol[i] <= ZERO) || (RelTol[i] <= 10.0 * Roundoff) || (RelTol[i] >= ONE)) {
        IERR = -5;
        break;
      }
  if (!IERR)
    IERR = ros_Integrator_f90(c, Y, Tstart, Tend, AbsTol, RelTol, &m, Autonomous, VectorTol,
                              Max_no_steps, Roundoff, Hmin, Hmax, Hstart, FacMin, FacMax, FacRej,
                              FacSafe, o, rstatus);
  istatus[0] = c->st.Nfun; istatus[1] = c->st.Njac; istatus[2] = c->st.Nstp;
  istatus[3] = c->st.Nacc; istatus[4] = c->st.Nrej; istatus[5] = c->st.Ndec;
  istatus[6] = c->st.Nsol; istatus[7] = c->st.Nsng;
  return IERR;
}

/* The merge of INTEGRATE (int/rosenbrock.f90:57-80): ICNTRL = 0 with
 * ICNTRL(15) = 5, then ICNTRL_U where /= 0; RCNTRL = 0, then RCNTRL_U where
 * > 0.  (Either user array may be NULL: not PRESENT.) */
void ros_ref_f90_merge(const int* ICNTRL_U, const double* RCNTRL_U, int* ICNTRL,
                       double* RCNTRL) {
  for (int i = 0; i < 20; i++) { ICNTRL[i] = 0; RCNTRL[i] = ZERO; }
  ICNTRL[14] = 5;
  for (int i = 0; i < 20; i++) {
    if (ICNTRL_U && ICNTRL_U[i] != 0) ICNTRL[i] = ICNTRL_U[i];
    if (RCNTRL_U && RCNTRL_U[i] > ZERO) RCNTRL[i] = RCNTRL_U[i];
  }
}

/* Batched F90 INTEGRATE over SoA concentrations, cell c over [tin[c], tout[c]].
 * ICNTRL_U/RCNTRL_U as the F90 caller passes them (merged as INTEGRATE does);
 * ICNTRL(15) selects the rate updates (util/util.f90:133-159): RCONST bit 1
 * -> rates recomputed at every Fun/Jac (with Update_SUN from TIME if bit 4,
 * else the cell's SUN), otherwise the cell's rates stay as computed at entry
 * (or as given in rconst).  passive: [NVAR] or NULL.  istatus [8][ncells],
 * rstatus [3][ncells] (Texit, Hexit, Hnew), ierr [ncells]. */
void ros_ref_f90_batch(int ncells, double* C, const double* temp, const double* cfactor,
                       const double* sun, const double* tin, const double* tout,
                       const double* atol, const double* rtol, const int* ICNTRL_U,
                       const double* RCNTRL_U, const int* passive, const double* rconst,
                       int* ierr, int* istatus, double* rstatus, int h211b, int nthreads) {
  int ICNTRL[20];
  double RCNTRL[20];
  ros_ref_f90_merge(ICNTRL_U, RCNTRL_U, ICNTRL, RCNTRL);
  const int opt = ICNTRL[14];
  const int do_rconst = (opt != -1) && (opt & 1), do_sun = (opt != -1) && (opt & 4);
  f90_opts o;
  o.passive = passive;
  o.nonneg = (ICNTRL[15] == 1);
  o.np_count = KPP_NVAR;
  /* h211b: the integrator of int/rosenbrock_h211b_qssa.f90 (its QSSA
   * options, ICNTRL(18), are not restated); H211b control when RCNTRL(15)
   * and RCNTRL(16) are both set (:490-514), else the F90 controller */
  o.bb = RCNTRL[14];
  o.kk = RCNTRL[15];
  o.h211b = h211b && o.bb > ZERO && o.kk > ZERO;
  if (passive) {
    o.np_count = 0;
    for (int i = 0; i < KPP_NVAR; i++) o.np_count += passive[i] ? 0 : 1;
  }
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
#endif
  for (int cidx = 0; cidx < ncells; cidx++) {
    ros_cell c;
    double Y[KPP_NVAR];
    int ist[8];
    double rst[3];
    memset(&c, 0, sizeof c);
    c.TEMP = temp[cidx];
    c.CFACTOR = cfactor[cidx];
    c.SUN = sun ? sun[cidx] : 0.0;
    c.sun_mode = (do_rconst && do_sun) ? 1 : 0;
    c.fun_split = KPP_FUN_SPLIT;
    for (int i = 0; i < KPP_NVAR; i++) Y[i] = C[(size_t)i * ncells + cidx];
    for (int i = 0; i < KPP_NFIX; i++) c.FIX[i] = C[(size_t)(KPP_NVAR + i) * ncells + cidx];
    if (rconst) {
      c.rates_given = 1;
      for (int i = 0; i < KPP_NREACT; i++) c.RCONST[i] = rconst[(size_t)i * ncells + cidx];
    } else if (!do_rconst) {
      /* rates the caller's Update_RCONST left before INTEGRATE: computed
       * once from the cell's state at entry, then held */
      kpp_Update_RCONST(c.TEMP, c.SUN, c.CFACTOR, tin[cidx], Y, c.FIX, c.RCONST);
      c.rates_given = 1;
    }
    int e = Rosenbrock_f90(&c, Y, tin[cidx], tout[cidx], atol, rtol, RCNTRL, ICNTRL, &o, ist,
                           rst);
    for (int i = 0; i < KPP_NVAR; i++) C[(size_t)i * ncells + cidx] = Y[i];
    ierr[cidx] = e;
    if (istatus)
      for (int i = 0; i < 8; i++) istatus[(size_t)i * ncells + cidx] = ist[i];
    if (rstatus)
      for (int i = 0; i < 3; i++) rstatus[(size_t)i * ncells + cidx] = rst[i];
  }
}

/* the C SDIRK integrator (int/sdirk.c) on the same model functions */
#include "sdirk_ref.h"
#include "rk_ref.h"
/* the C discrete adjoint Rosenbrock integrator (int/rosenbrock_adj.c) */
#if defined(KPP_ADJOINT) && KPP_ADJOINT
#include "rosadj_ref.h"
#endif

/* ======================= exported entry points ======================= */
int ros_ref_sizes(int* out) {
  out[0] = KPP_NVAR;
  out[1] = KPP_NFIX;
  out[2] = KPP_NSPEC;
  out[3] = KPP_NREACT;
  out[4] = KPP_NONZERO;
  out[5] = KPP_LU_NONZERO;
  return 6;
}
const char* ros_ref_root(void) { return KPP_ROOT; }
void ros_ref_lu_pattern(int* icol, int* crow, int* diag) {
  for (int i = 0; i < KPP_LU_NONZERO; i++) icol[i] = kpp_LU_ICOL[i];
  for (int i = 0; i <= KPP_NVAR; i++) { crow[i] = kpp_LU_CROW[i]; diag[i] = kpp_LU_DIAG[i]; }
}
const char* ros_ref_species_name(int k) { return (k >= 0 && k < KPP_NSPEC) ? kpp_SPC_NAMES[k] : 0; }
void ros_ref_initialize(double* C) { kpp_Initialize(C); }
void ros_ref_rconst(double TEMP, double SUN, double CFACTOR, double TIME, double* RCONST) {
  static const double zero[KPP_NSPEC] = {0};
  kpp_Update_RCONST(TEMP, SUN, CFACTOR, TIME, zero, zero + KPP_NVAR, RCONST);
}
/* rate constants seen by a cell whose concentrations are C[NSPEC] */
void ros_ref_rconst_c(double TEMP, double SUN, double CFACTOR, double TIME, const double* C,
                      double* RCONST) {
  kpp_Update_RCONST(TEMP, SUN, CFACTOR, TIME, C, C + KPP_NVAR, RCONST);
}
void ros_ref_fun(const double* V, const double* F, const double* RCT, double* Vdot) {
  kpp_Fun(V, F, RCT, Vdot);
}
void ros_ref_jac_sp(const double* V, const double* F, const double* RCT, double* JVS) {
  kpp_Jac_SP(V, F, RCT, JVS);
}
int ros_ref_decomp(double* JVS) { return KppDecomp(JVS); }
/* Fun_SPLIT for tests: 1 and Vdot = P - D*V for a #FUNCTION SPLIT mechanism, else 0 */
int ros_ref_fun_split(const double* V, const double* F, const double* RCT, double* Vdot) {
#if KPP_FUN_SPLIT
  kpp_Fun_SPLIT(V, F, RCT, Vdot);
  return 1;
#else
  (void)V; (void)F; (void)RCT; (void)Vdot;
  return 0;
#endif
}
void ros_ref_solve(const double* JVS, double* X) { kpp_KppSolve(JVS, X); }
/* Ghimj build + decomposition of ros_PrepareMatrix for one fixed H */
int ros_ref_prepare(const double* Jac0, double H, double gam, double* Ghimj) {
  ros_cell c;
  memset(&c, 0, sizeof c);
  double h = H;
  int r = ros_PrepareMatrix(&c, &h, +1, gam, Jac0, Ghimj);
  return r ? -1 : c.st.Nsng;
}
double ros_ref_error_norm(const double* Y, const double* Ynew, const double* Yerr,
                          const double* AbsTol, const double* RelTol, int VectorTol) {
  return ros_ErrorNorm(Y, Ynew, Yerr, AbsTol, RelTol, VectorTol);
}

/* Batched integration over SoA concentrations C[s*ncells + c] (NSPEC rows).
 * temp/cfactor/sun: per-cell arrays; sun_mode 0 = fixed SUN, 1 = Update_SUN.
 * tstart/tend: per-cell arrays.  atol/rtol: NVAR entries (IPAR[1]==0) or 1.
 * hstart: per-cell starting step (overrides RPAR[2]) or NULL.
 * rconst: per-cell rate constants (SoA NREACT x ncells) or NULL.
 * ierr[c], istat[8*ncells] (Nfun..Nsng, SoA), rstat[2*ncells] (Texit, Hexit).
 * hacc1[c]: step size of the cell's first accepted step, or NULL (a step log
 * for the learned scheduler; not part of the reference's outputs).
 */
void ros_ref_batch(int ncells, double* C, const double* temp, const double* cfactor,
                   const double* sun, int sun_mode, const double* tstart, const double* tend,
                   const double* atol, const double* rtol, const int* IPAR_in,
                   const double* RPAR_in, const double* hstart, const double* rconst, int* ierr,
                   int* istat, double* rstat, double* hacc1, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
#endif
  for (int cidx = 0; cidx < ncells; cidx++) {
    ros_cell c;
    double Y[KPP_NVAR];
    int IPAR[20];
    double RPAR[20];
    memset(&c, 0, sizeof c);
    for (int i = 0; i < 20; i++) { IPAR[i] = IPAR_in[i]; RPAR[i] = RPAR_in[i]; }
    if (hstart) RPAR[2] = hstart[cidx];
    c.TEMP = temp[cidx];
    c.CFACTOR = cfactor[cidx];
    c.SUN = sun ? sun[cidx] : 0.0;
    c.sun_mode = sun_mode;
    for (int i = 0; i < KPP_NVAR; i++) Y[i] = C[(size_t)i * ncells + cidx];
    for (int i = 0; i < KPP_NFIX; i++) c.FIX[i] = C[(size_t)(KPP_NVAR + i) * ncells + cidx];
    if (rconst) {
      c.rates_given = 1;
      for (int i = 0; i < KPP_NREACT; i++) c.RCONST[i] = rconst[(size_t)i * ncells + cidx];
    }
    int e = Rosenbrock(&c, Y, tstart[cidx], tend[cidx], atol, rtol, RPAR, IPAR);
    for (int i = 0; i < KPP_NVAR; i++) C[(size_t)i * ncells + cidx] = Y[i];
    ierr[cidx] = e;
    if (istat)
      for (int i = 0; i < 8; i++) istat[(size_t)i * ncells + cidx] = IPAR[10 + i];
    if (rstat) {
      rstat[cidx] = RPAR[10];
      rstat[(size_t)ncells + cidx] = RPAR[11];
    }
    if (hacc1) hacc1[cidx] = c.Hacc1;
  }
}

int ros_ref_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
