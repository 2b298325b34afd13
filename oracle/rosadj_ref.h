/* rosadj_ref.h -- CPU oracle of KPP's C discrete adjoint Rosenbrock
 * integrator (int/rosenbrock_adj.c).
 *
 * TEST INFRASTRUCTURE ONLY (included by ros_ref.c, whose header says who may
 * load the library).  A reentrant restatement, one cell per call, of the
 * adjoint types 1 (none) and 2 (discrete):
 *   RosenbrockADJ          int/rosenbrock_adj.c:191-597 (controls, checks)
 *   ros_DPush / ros_DPop   :779-855 (checkpoints in a per-cell buffer)
 *   ros_FwdInt             :949-1186
 *   ros_DadjInt            :1189-1324
 *   ros_ErrorNorm          :1747-1770
 *   ros_FunTimeDerivative  :1773-1788
 *   ros_JacTimeDerivative  :1791-1810
 *   ros_PrepareMatrix      :1813-1872
 *   method tables          :2054-2470 (Ros2 with g = 1 + 1/SQRT(2), Rodas3
 *                          with C(6) = -(8/3), this file's Rodas3_1)
 *   FunTemplate/JacTemplate/HessTemplate  :2474-2544 (rates at T)
 * and the model functions Hessian, HessTR_Vec, JacTR_SP_Vec, KppSolveTR of
 * the mechanism compiler's C output (src/gen.c:1327-1600, 1829-1870,
 * 2035-2106).
 *
 * Kept from the reference: ros_PrepareMatrix takes H by value (a singular
 * matrix halves only its local step; the stages use the caller's H with that
 * LU) and gives up after the sixth singular matrix (-8); ICNTRL[3] != 0 is
 * always -1 (the test reads the uninitialised Max_no_steps, 0); the maximum
 * number of steps is bufsize - 1 = 199999; RSTATUS[2] (Hexit) is set to H
 * before the step is cut to Tend and to the accepted H; the backward loop
 * runs while stack_ptr > 0, so the first accepted step is not differentiated
 * (first_step = 0) -- first_step = 1 processes it too, the complete discrete
 * adjoint (int/rosenbrock_adj.f90's stack is 1-based).  ISTATUS[1..8] and
 * RSTATUS[1..3] are indexed from 1 (Nfun = 1 ... Nsng = 8).
 * Deviations: a NaN error estimate ends the cell with -10 (the reference's
 * MAX(Err, 1e-10) turns it into 1e-10 and accepts), more than MAX_REJECT
 * rejections in a row with -11, a full checkpoint buffer with -12 (the
 * reference exits the process); adjoint types 3 and 4 (continuous) are not
 * restated (-13). */

#define ADJ_NONFINITE (-10)
#define ADJ_REJECTS (-11)
#define ADJ_CHECKPOINTS (-12)
#define ADJ_NOT_BUILT (-13)
#define ADJ_DELTAMIN 1.0e-5

/* planted mistakes for the tests that check the checks (0 = none): 1 drops
 * the Hessian term, 2 the time-derivative term H dJdT^T sum(gamma_i U_i), 3
 * transposes the solve (KppSolve in place of KppSolveTR) */
static int adj_plant = 0;
void adj_ref_plant(int k) { adj_plant = k; }

/* the method tables of int/rosenbrock_adj.c:2054-2470 */
static int adj_method(int code, ros_method* m) {
  memset(m, 0, sizeof *m);
  switch (code) {
    case 0: case 4: Rodas3_f90(m); return 1; /* C(6) = -(8/3) */
    case 1: Ros2_f90(m); return 1;            /* g = 1 + 1/SQRT(2); pow(2, 0.5) == sqrt(2) */
    case 2: Ros3(m); return 1;
    case 3: Ros4(m); return 1;
    case 5: Rodas4(m); return 1;
    case 7: { /* this file's Rodas3_1 (:2302-2370) */
      static const double A[6] = {0.0, 1.5382237953138116, -0.36440683885434433,
                                  1.538223795313811, -0.36440683885434433, 1.0};
      static const double C[6] = {-4.0919303685081028, -3.0551174378039538e-002,
                                  1.7259281281917580, 0.19561160936073679,
                                  1.9301670595355112, -2.6267006001193960};
      m->S = 4;
      for (int i = 0; i < 6; i++) { m->A[i] = A[i]; m->C[i] = C[i]; }
      m->NewF[0] = 1; m->NewF[1] = 0; m->NewF[2] = 1; m->NewF[3] = 1;
      m->M[0] = 1.5382237953138116; m->M[1] = -0.36440683885434444;
      m->M[2] = 1.0000000000000002; m->M[3] = 1.0;
      m->E[3] = 1.0;
      m->ELO = 3.0;
      m->Alpha[2] = 1.0; m->Alpha[3] = 1.0;
      m->Gamma[0] = 0.5150000000000000; m->Gamma[1] = -0.57028223198756145;
      return 1;
    }
  }
  return 0;
}

typedef struct {
  double Hmin, Hmax, Hstart, FacMin, FacMax, FacRej, FacSafe, Roundoff;
  int Max_no_steps, Autonomous, VectorTol, SaveLU, first_step;
} adj_opts;

/* the per-cell checkpoint buffer (chk_H, chk_T, chk_Y, chk_K, chk_J) */
typedef struct {
  int cap, ptr, S;
  double *H, *T, *Y, *K, *J;
} adj_stack;

static double adj_error_norm(const double* Y, const double* Ynew, const double* Yerr,
                             const double* AbsTol, const double* RelTol, int VectorTol) {
  double Err = ZERO;
  for (int i = 0; i < KPP_NVAR; i++) {
    const double Ymax = MAX(ABS(Y[i]), ABS(Ynew[i]));
    const double Scale = VectorTol ? AbsTol[i] + RelTol[i] * Ymax : AbsTol[0] + RelTol[0] * Ymax;
    Err = Err + pow((Yerr[i] / Scale), 2);
  }
  return SQRT(Err / KPP_NVAR); /* the caller takes MAX(Err, 1e-10) */
}

/* ros_PrepareMatrix with H by value: 1 if still singular after five halvings */
static int adj_prepare(ros_cell* c, double H, int Direction, double gam, const double* Jac0,
                       double* Ghimj) {
  int Nconsecutive = 0;
  for (;;) {
    for (int i = 0; i < KPP_LU_NONZERO; i++) Ghimj[i] = -Jac0[i];
    const double ghinv = ONE / (Direction * H * gam);
    for (int i = 0; i < KPP_NVAR; i++) Ghimj[kpp_LU_DIAG[i]] = Ghimj[kpp_LU_DIAG[i]] + ghinv;
    const int ising = KppDecomp(Ghimj);
    c->st.Ndec++;
    if (ising == 0) return 0;
    c->st.Nsng++;
    Nconsecutive++;
    if (Nconsecutive <= 5) H = H * HALF;
    else return 1;
  }
}

static void HessTemplate(ros_cell* c, double T, const double* Y, double* Hes) {
  update_rates(c, T);
  kpp_Hessian(Y, c->FIX, c->RCONST, Hes);
}

/* ros_FwdInt (:949-1186), discrete adjoint checkpoints when stk != NULL */
static int adj_fwd(ros_cell* c, double* Y, double Tstart, double Tend, const double* AbsTol,
                   const double* RelTol, const ros_method* m, const adj_opts* o, adj_stack* stk,
                   double* rstatus) {
  const int N = KPP_NVAR, S = m->S;
  double Ynew[KPP_NVAR], Fcn0[KPP_NVAR], Fcn[KPP_NVAR], dFdT[KPP_NVAR], Yerr[KPP_NVAR];
  double K[KPP_NVAR * Smax], Ystage[KPP_NVAR * Smax];
  double Jac0[KPP_LU_NONZERO], Ghimj[KPP_LU_NONZERO];
  double H, Hnew, Fac, Err, T = Tstart;
  int RejectLastH = 0, RejectMoreH = 0, Nrun = 0;
  const int Direction = (Tend >= Tstart) ? 1 : -1;
  rstatus[2] = ZERO;
  H = MIN(MAX(ABS(o->Hmin), ABS(o->Hstart)), ABS(o->Hmax));
  if (ABS(H) <= 10.0 * o->Roundoff) H = ADJ_DELTAMIN;
  H = Direction * H;
  c->Yacc = Y;
  memset(Ynew, 0, sizeof Ynew);
  while (((Direction > 0) && ((T - Tend) + o->Roundoff <= ZERO)) ||
         ((Direction < 0) && ((Tend - T) + o->Roundoff <= ZERO))) {
    if (c->st.Nstp > o->Max_no_steps) return -6;
    if (((T + 0.1 * H) == T) || (H <= o->Roundoff)) return -7;
    rstatus[2] = H;
    H = MIN(H, ABS(Tend - T));
    FunTemplate(c, T, Y, Fcn0);
    if (!o->Autonomous) {
      const double Delta = SQRT(o->Roundoff) * MAX(ADJ_DELTAMIN, ABS(T));
      FunTemplate(c, T + Delta, Y, dFdT);
      WAXPY(N, (-ONE), Fcn0, dFdT);
      WSCAL(N, (ONE / Delta), dFdT);
    }
    JacTemplate(c, T, Y, Jac0);
    for (;;) {
      if (adj_prepare(c, H, Direction, m->Gamma[0], Jac0, Ghimj)) return -8;
      for (int istage = 0; istage < S; istage++) {
        double* Ki = K + N * istage;
        if (istage == 0) {
          WCOPY(N, Fcn0, Fcn);
          if (stk) {
            for (int i = 0; i < N; i++) Ystage[i] = Y[i];
            WCOPY(N, Y, Ynew);
          }
        } else if (m->NewF[istage]) {
          WCOPY(N, Y, Ynew);
          for (int j = 0; j < istage; j++)
            WAXPY(N, m->A[istage * (istage - 1) / 2 + j], K + N * j, Ynew);
          FunTemplate(c, T + m->Alpha[istage] * Direction * H, Ynew, Fcn);
        }
        if (istage > 0 && stk)
          for (int i = 0; i < N; i++) Ystage[N * istage + i] = Ynew[i];
        WCOPY(N, Fcn, Ki);
        for (int j = 0; j < istage; j++) {
          const double HC = m->C[istage * (istage - 1) / 2 + j] / (Direction * H);
          WAXPY(N, HC, K + N * j, Ki);
        }
        if (!o->Autonomous && m->Gamma[istage] != ZERO)
          WAXPY(N, Direction * H * m->Gamma[istage], dFdT, Ki);
        kpp_KppSolve(Ghimj, Ki);
        c->st.Nsol++;
      }
      WCOPY(N, Y, Ynew);
      for (int j = 0; j < S; j++) WAXPY(N, m->M[j], K + N * j, Ynew);
      WSCAL(N, ZERO, Yerr);
      for (int j = 0; j < S; j++) WAXPY(N, m->E[j], K + N * j, Yerr);
      Err = adj_error_norm(Y, Ynew, Yerr, AbsTol, RelTol, o->VectorTol);
      if (Err != Err) return ADJ_NONFINITE;
      Err = MAX(Err, 1.0e-10);
      Fac = MIN(o->FacMax, MAX(o->FacMin, o->FacSafe / pow(Err, (ONE / m->ELO))));
      Hnew = H * Fac;
      c->st.Nstp++;
      if ((Err <= ONE) || (H <= o->Hmin)) {
        c->st.Nacc++;
        if (stk) { /* ros_DPush */
          if (stk->ptr + 1 >= stk->cap) return ADJ_CHECKPOINTS;
          const int k = ++stk->ptr;
          stk->H[k] = H;
          stk->T[k] = T;
          for (int i = 0; i < N * S; i++) {
            stk->Y[(size_t)k * N * S + i] = Ystage[i];
            stk->K[(size_t)k * N * S + i] = K[i];
          }
          if (o->SaveLU)
            for (int i = 0; i < KPP_LU_NONZERO; i++)
              stk->J[(size_t)k * KPP_LU_NONZERO + i] = Ghimj[i];
        }
        WCOPY(N, Ynew, Y);
        T = T + Direction * H;
        Hnew = MAX(o->Hmin, MIN(Hnew, o->Hmax));
        if (RejectLastH) Hnew = MIN(Hnew, H);
        rstatus[2] = H;
        rstatus[3] = Hnew;
        rstatus[1] = T;
        RejectLastH = 0;
        RejectMoreH = 0;
        H = Hnew;
        Nrun = 0;
        break;
      } else {
        if (RejectMoreH) Hnew = H * o->FacRej;
        RejectMoreH = RejectLastH;
        RejectLastH = 1;
        H = Hnew;
        if (c->st.Nacc >= 1) c->st.Nrej++;
        if (!(ABS(H) <= DBL_MAX)) return ADJ_NONFINITE;
        if (++Nrun > MAX_REJECT) return ADJ_REJECTS;
      }
    }
  }
  return 1;
}

/* ros_DadjInt (:1189-1324); Lambda [NADJ][NVAR] */
static int adj_dadj(ros_cell* c, int NADJ, double* Lambda, double Tstart, double Tend,
                    const ros_method* m, const adj_opts* o, adj_stack* stk) {
  const int N = KPP_NVAR, S = m->S;
  double* U = (double*)malloc(sizeof(double) * (size_t)NADJ * N * S);
  double* V = (double*)malloc(sizeof(double) * (size_t)NADJ * N * S);
  double Jac[KPP_LU_NONZERO], dJdT[KPP_LU_NONZERO], Ghimj[KPP_LU_NONZERO], Hes0[KPP_NHESS];
  double Tmp[KPP_NVAR], Tmp2[KPP_NVAR];
  const int Direction = (Tend >= Tstart) ? 1 : -1;
  const int last = o->first_step ? 0 : 1;
  while (stk->ptr >= last) {
    const int k = stk->ptr--;
    const double T = stk->T[k], H = stk->H[k];
    const double* Ystage = stk->Y + (size_t)k * N * S;
    const double* K = stk->K + (size_t)k * N * S;
    if (o->SaveLU) {
      for (int i = 0; i < KPP_LU_NONZERO; i++) Ghimj[i] = stk->J[(size_t)k * KPP_LU_NONZERO + i];
    } else {
      JacTemplate(c, T, Ystage, Ghimj);
      const double Tau = ONE / (Direction * H * m->Gamma[0]);
      WSCAL(KPP_LU_NONZERO, (-ONE), Ghimj);
      for (int i = 0; i < N; i++) Ghimj[kpp_LU_DIAG[i]] = Ghimj[kpp_LU_DIAG[i]] + Tau;
      (void)KppDecomp(Ghimj);
      c->st.Ndec++;
    }
    HessTemplate(c, T, Ystage, Hes0);
    for (int istage = S - 1; istage >= 0; istage--) {
      const int istart = N * istage;
      for (int a = 0; a < NADJ; a++) {
        double* Ui = U + (size_t)a * N * S + istart;
        WCOPY(N, Lambda + (size_t)a * N, Ui);
        WSCAL(N, m->M[istage], Ui);
      }
      for (int j = istage + 1; j < S; j++) {
        const int jstart = N * j;
        const double HA = m->A[j * (j - 1) / 2 + istage];
        const double HC = m->C[j * (j - 1) / 2 + istage] / (Direction * H);
        for (int a = 0; a < NADJ; a++) {
          WAXPY(N, HA, V + (size_t)a * N * S + jstart, U + (size_t)a * N * S + istart);
          WAXPY(N, HC, U + (size_t)a * N * S + jstart, U + (size_t)a * N * S + istart);
        }
      }
      for (int a = 0; a < NADJ; a++) {
        double* Ui = U + (size_t)a * N * S + istart;
        if (adj_plant == 3) kpp_KppSolve(Ghimj, Ui);
        else kpp_KppSolveTR(Ghimj, Ui, Ui);
        c->st.Nsol++;
      }
      JacTemplate(c, T + m->Alpha[istage] * Direction * H, Ystage + istart, Jac);
      for (int a = 0; a < NADJ; a++)
        kpp_JacTR_SP_Vec(Jac, U + (size_t)a * N * S + istart, V + (size_t)a * N * S + istart);
    }
    if (!o->Autonomous) { /* ros_JacTimeDerivative with the last Jac (stage 1) */
      const double Delta = SQRT(o->Roundoff) * MAX(ADJ_DELTAMIN, ABS(T));
      JacTemplate(c, T + Delta, Ystage, dJdT);
      WAXPY(KPP_LU_NONZERO, (-ONE), Jac, dJdT);
      WSCAL(KPP_LU_NONZERO, (ONE / Delta), dJdT);
    }
    for (int istage = 0; istage < S; istage++) {
      const int istart = N * istage;
      for (int a = 0; a < NADJ; a++) {
        double* La = Lambda + (size_t)a * N;
        WAXPY(N, ONE, V + (size_t)a * N * S + istart, La);
        kpp_HessTR_Vec(Hes0, U + (size_t)a * N * S + istart, K + istart, Tmp);
        if (adj_plant != 1) WAXPY(N, ONE, Tmp, La);
      }
    }
    if (!o->Autonomous && adj_plant != 2) {
      for (int a = 0; a < NADJ; a++) {
        for (int i = 0; i < N; i++) Tmp[i] = ZERO;
        for (int istage = 0; istage < S; istage++)
          WAXPY(N, m->Gamma[istage], U + (size_t)a * N * S + N * istage, Tmp);
        kpp_JacTR_SP_Vec(dJdT, Tmp, Tmp2);
        WAXPY(N, H, Tmp2, Lambda + (size_t)a * N);
      }
    }
  }
  free(U);
  free(V);
  return 1;
}

/* RosenbrockADJ (:191-597) for one cell; istatus[0..7] = ISTATUS[1..8],
 * rstatus[0..2] = RSTATUS[1..3] */
static int RosenbrockADJ_cell(ros_cell* c, double* Y, int NADJ, double* Lambda, double Tstart,
                              double Tend, const double* AbsTol, const double* RelTol,
                              const double* RCNTRL, const int* ICNTRL, int first_step,
                              int cap, int* istatus, double* rstatus) {
  ros_method m;
  adj_opts o;
  double rst[4] = {0, 0, 0, 0};
  int IERR, AdjointType;
  memset(&c->st, 0, sizeof c->st);
  memset(&o, 0, sizeof o);
  o.first_step = first_step;
  o.Autonomous = !(ICNTRL[0] == 0);
  o.VectorTol = (ICNTRL[1] == 0);
  const int UplimTol = o.VectorTol ? KPP_NVAR : 1;
  IERR = 1;
  if (!adj_method(ICNTRL[2], &m)) IERR = -2;
  else if (ICNTRL[3] != 0) IERR = -1; /* the reference tests its Max_no_steps, still 0 */
  else {
    o.Max_no_steps = 200000 - 1;
    if (!(ICNTRL[5] == 0 || (ICNTRL[5] >= 1 && ICNTRL[5] <= 5))) IERR = -2;
    else if (!(ICNTRL[6] == 0 || (ICNTRL[6] >= 1 && ICNTRL[6] <= 4))) IERR = -9;
  }
  AdjointType = (ICNTRL[6] == 0) ? 2 : ICNTRL[6];
  o.SaveLU = (ICNTRL[7] != 0);
  o.Roundoff = ros_ref_wlamch();
  const double span = ABS(Tend - Tstart);
  if (IERR == 1) {
    if (RCNTRL[0] == ZERO) o.Hmin = ZERO;
    else if (RCNTRL[0] > ZERO) o.Hmin = RCNTRL[0];
    else IERR = -3;
  }
  if (IERR == 1) {
    if (RCNTRL[1] == ZERO) o.Hmax = span;
    else if (RCNTRL[1] > ZERO) o.Hmax = MIN(ABS(RCNTRL[1]), span);
    else IERR = -3;
  }
  if (IERR == 1) {
    if (RCNTRL[2] == ZERO) o.Hstart = MAX(o.Hmin, ADJ_DELTAMIN);
    else if (RCNTRL[2] > ZERO) o.Hstart = MIN(ABS(RCNTRL[2]), span);
    else IERR = -3;
  }
  if (IERR == 1) {
    const double def[4] = {0.2, 6.0, 0.1, 0.9};
    double* dst[4] = {&o.FacMin, &o.FacMax, &o.FacRej, &o.FacSafe};
    for (int k = 0; k < 4 && IERR == 1; k++) {
      if (RCNTRL[3 + k] == ZERO) *dst[k] = def[k];
      else if (RCNTRL[3 + k] > ZERO) *dst[k] = RCNTRL[3 + k];
      else IERR = -4;
    }
  }
  if (IERR == 1)
    for (int i = 0; i < UplimTol; i++)
      if ((AbsTol[i] <= ZERO) || (RelTol[i] <= 10.0 * o.Roundoff) || (RelTol[i] >= 1.0)) IERR = -5;
  if (IERR == 1 && (AdjointType == 3 || AdjointType == 4)) IERR = ADJ_NOT_BUILT;
  if (IERR == 1) {
    adj_stack stk, *sp = NULL;
    memset(&stk, 0, sizeof stk);
    if (AdjointType == 2) {
      stk.cap = cap;
      stk.ptr = -1;
      stk.S = m.S;
      stk.H = (double*)malloc(sizeof(double) * cap);
      stk.T = (double*)malloc(sizeof(double) * cap);
      stk.Y = (double*)malloc(sizeof(double) * (size_t)cap * KPP_NVAR * m.S);
      stk.K = (double*)malloc(sizeof(double) * (size_t)cap * KPP_NVAR * m.S);
      stk.J = o.SaveLU ? (double*)malloc(sizeof(double) * (size_t)cap * KPP_LU_NONZERO) : NULL;
      sp = &stk;
    }
    IERR = adj_fwd(c, Y, Tstart, Tend, AbsTol, RelTol, &m, &o, sp, rst);
    if (IERR >= 0 && AdjointType == 2) IERR = adj_dadj(c, NADJ, Lambda, Tstart, Tend, &m, &o, sp);
    if (sp) {
      free(stk.H); free(stk.T); free(stk.Y); free(stk.K); free(stk.J);
    }
  }
  istatus[0] = c->st.Nfun; istatus[1] = c->st.Njac; istatus[2] = c->st.Nstp;
  istatus[3] = c->st.Nacc; istatus[4] = c->st.Nrej; istatus[5] = c->st.Ndec;
  istatus[6] = c->st.Nsol; istatus[7] = c->st.Nsng;
  rstatus[0] = rst[1]; rstatus[1] = rst[2]; rstatus[2] = rst[3];
  return IERR;
}

/* Batched RosenbrockADJ over SoA arrays: C [NSPEC][n], Lambda [NADJ*NVAR][n]
 * (row a*NVAR + i = Lambda[a][i]), ICNTRL/RCNTRL as RosenbrockADJ takes them
 * (0-based, no merge), cap checkpoints per cell; sun_mode as ros_ref_batch;
 * rconst or NULL.  istatus [8][n], rstatus [3][n]. */
void adj_ref_batch(int ncells, double* C, const double* temp, const double* cfactor,
                   const double* sun, int sun_mode, const double* tin, const double* tout,
                   const double* atol, const double* rtol, const int* ICNTRL, const double* RCNTRL,
                   const double* rconst, int NADJ, double* Lambda, int first_step, int cap,
                   int* ierr, int* istatus, double* rstatus, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4)
#endif
  for (int cidx = 0; cidx < ncells; cidx++) {
    ros_cell c;
    double Y[KPP_NVAR];
    double* L = (double*)malloc(sizeof(double) * (size_t)(NADJ > 0 ? NADJ : 1) * KPP_NVAR);
    int ist[8];
    double rst[3];
    memset(&c, 0, sizeof c);
    c.TEMP = temp[cidx];
    c.CFACTOR = cfactor[cidx];
    c.SUN = sun ? sun[cidx] : 0.0;
    c.sun_mode = sun_mode;
    for (int i = 0; i < KPP_NVAR; i++) Y[i] = C[(size_t)i * ncells + cidx];
    for (int i = 0; i < KPP_NFIX; i++) c.FIX[i] = C[(size_t)(KPP_NVAR + i) * ncells + cidx];
    for (int r = 0; r < NADJ * KPP_NVAR; r++) L[r] = Lambda[(size_t)r * ncells + cidx];
    if (rconst) {
      c.rates_given = 1;
      for (int i = 0; i < KPP_NREACT; i++) c.RCONST[i] = rconst[(size_t)i * ncells + cidx];
    }
    int e = RosenbrockADJ_cell(&c, Y, NADJ, L, tin[cidx], tout[cidx], atol, rtol, RCNTRL, ICNTRL,
                               first_step, cap, ist, rst);
    for (int i = 0; i < KPP_NVAR; i++) C[(size_t)i * ncells + cidx] = Y[i];
    for (int r = 0; r < NADJ * KPP_NVAR; r++) Lambda[(size_t)r * ncells + cidx] = L[r];
    free(L);
    ierr[cidx] = e;
    if (istatus)
      for (int i = 0; i < 8; i++) istatus[(size_t)i * ncells + cidx] = ist[i];
    if (rstatus)
      for (int i = 0; i < 3; i++) rstatus[(size_t)i * ncells + cidx] = rst[i];
  }
}

/* the adjoint's model functions for tests */
int adj_ref_nhess(void) { return KPP_NHESS; }
void adj_ref_hessian(const double* V, const double* F, const double* RCT, double* HESS, int* ii,
                     int* jj, int* kk) {
  kpp_Hessian(V, F, RCT, HESS);
  for (int m = 0; m < KPP_NHESS; m++) {
    ii[m] = kpp_IHESS_I[m]; jj[m] = kpp_IHESS_J[m]; kk[m] = kpp_IHESS_K[m];
  }
}
void adj_ref_hesstr_vec(const double* HESS, const double* U1, const double* U2, double* HTU) {
  kpp_HessTR_Vec(HESS, U1, U2, HTU);
}
void adj_ref_jactr_vec(const double* JVS, const double* U, double* V) {
  kpp_JacTR_SP_Vec(JVS, U, V);
}
void adj_ref_jac_vec(const double* JVS, const double* U, double* V) { kpp_Jac_SP_Vec(JVS, U, V); }
void adj_ref_solve_tr(const double* JVS, const double* X, double* XX) { kpp_KppSolveTR(JVS, X, XX); }
