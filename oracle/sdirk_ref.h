/* sdirk_ref.h -- CPU oracle of KPP's C SDIRK integrator (int/sdirk.c).
 *
 * TEST INFRASTRUCTURE ONLY (included by ros_ref.c, whose header says who may
 * load the library).  A reentrant restatement, one cell per call:
 *   SDIRK                  int/sdirk.c:170-525 (defaults and checks)
 *   SDIRK_Integrator       int/sdirk.c:528-812
 *   SDIRK_ErrorScale/Norm  int/sdirk.c:815-851
 *   SDIRK_PrepareMatrix    int/sdirk.c:905-960
 *   SDIRK_Solve            int/sdirk.c:963-985
 *   Sdirk4a/4b/2a/2b/3a    int/sdirk.c:988-1306
 *   FUN_CHEM/JAC_CHEM      int/sdirk.c:1309-1361 (Update_SUN and
 *                          Update_RCONST before every call)
 *
 * The reference's own behaviour is kept where it is defined:
 *  - SDIRK_PrepareMatrix takes H, SkipJac and ISING by value: its step
 *    halving on a singular matrix changes only its local copy (the LU is of
 *    the halved step, the solves scale by the caller's), and the caller's
 *    singularity flag stays 0, so -8 never reaches the caller;
 *  - the Newton predicted error is NewtonIncrement *
 *    Theta^((NewtonMaxit-(iter+1))/(1-Theta)) and the step reduction
 *    0.8*Qnewton^(-1/(1+NewtonMaxit-iter+1)), as the C file has them;
 *  - ISTATUS/RSTATUS are indexed from 1 (ISTATUS[1] = Nfun ... [8] = Nsng,
 *    RSTATUS[1..3] = Texit, Hexit, Hnew); the batch returns them as rows.
 * Deviations, where the reference's behaviour is undefined or would not end:
 *  - a parameter error returns its code (-1..-5): the reference prints it,
 *    discards SDIRK_ErrorMsg's result and integrates with the setting left
 *    uninitialised;
 *  - the failures inside the integrator return their codes -6/-7 (the
 *    reference returns its Ierr argument, 0);
 *  - a NaN error estimate or a non-finite step ends the cell with -9, more
 *    than MAX_REJECT rejected attempts in a row (error or Newton failures)
 *    with -10. */

typedef struct {
  int S;
  double Gamma, C[5], D[5], E[5], ELO, Theta[5][5], Alpha[5][5];
} sdirk_method;

/* the method tables of int/sdirk.c:988-1306 (rkA/rkB/rkBhat define the
 * method; the integrator uses Gamma, C, D, E, Theta, Alpha and ELO) */
static int sdirk_method_of(int code, sdirk_method* m) {
  memset(m, 0, sizeof *m);
  switch (code) {
    case 0: case 1: default:  /* Sdirk2a (ICNTRL[2] 0, 1 and unknown) */
      m->S = 2;
      m->Gamma = 0.2928932188134524755991556378951510;
      m->C[0] = 0.292893218813452475599155637895151; m->C[1] = 1.0;
      m->D[1] = 1.0;
      m->E[0] = 0.4714045207910316829338962414032326;
      m->E[1] = -0.1380711874576983496005629080698993;
      m->ELO = 2;
      m->Theta[0][1] = 2.414213562373095048801688724209698;
      m->Alpha[0][1] = 3.414213562373095048801688724209698;
      return code >= 0 && code <= 5 ? 1 : 2;
    case 2:  /* Sdirk2b */
      m->S = 2;
      m->Gamma = 1.707106781186547524400844362104849;
      m->C[0] = 1.707106781186547524400844362104849; m->C[1] = 1.0;
      m->D[1] = 1.0;
      m->E[0] = -0.4714045207910316829338962414032326;
      m->E[1] = 0.8047378541243650162672295747365659;
      m->ELO = 2;
      m->Theta[0][1] = -0.414213562373095048801688724209698;
      m->Alpha[0][1] = 0.5857864376269049511983112757903019;
      return 1;
    case 3:  /* Sdirk3a */
      m->S = 3;
      m->Gamma = 0.2113248654051871177454256097490213;
      m->C[0] = 0.2113248654051871177454256097490213;
      m->C[1] = 0.4226497308103742354908512194980427; m->C[2] = 1.0;
      m->D[2] = 1.0;
      m->E[0] = 0.9106836025229590978424821138352906;
      m->E[1] = -1.244016935856292431175815447168624;
      m->E[2] = 0.3333333333333333333333333333333333;
      m->ELO = 2;
      m->Theta[0][1] = 1.0;
      m->Theta[0][2] = -1.732050807568877293527446341505872;
      m->Theta[1][2] = 2.732050807568877293527446341505872;
      m->Alpha[0][1] = 2.0;
      m->Alpha[0][2] = -12.92820323027550917410978536602349;
      m->Alpha[1][2] = 8.83012701892219323381861585376468;
      return 1;
    case 4:  /* Sdirk4a */
      m->S = 5;
      m->Gamma = 0.2666666666666666666666666666666667;
      m->C[0] = 0.2666666666666666666666666666666667;
      m->C[1] = 0.7666666666666666666666666666666667;
      m->C[2] = 0.5666666666666666666666666666666667;
      m->C[3] = 0.3661315380631796996374935266701191; m->C[4] = 1.0;
      m->D[4] = 1.0;
      m->E[0] = -0.6804000050475287124787034884002302;
      m->E[1] = 1.558961944525217193393931795738823;
      m->E[2] = -13.55893003128907927748632408763868;
      m->E[3] = 15.48522576958521253098585004571302;
      m->E[4] = 1.0;
      m->ELO = 4.0;
      m->Theta[0][1] = 1.875000000000000000000000000000000;
      m->Theta[0][2] = 1.708847304091539528432732316227462;
      m->Theta[1][2] = -0.2030773231622746185852981969486824;
      m->Theta[0][3] = 0.2680325578937783958847157206823118;
      m->Theta[1][3] = -0.1828840955527181631794050728644549;
      m->Theta[2][3] = 0.2968246986151577397160821594427966;
      m->Theta[0][4] = 0.9096171815241460655379433581446771;
      m->Theta[1][4] = -3.108254967778352416114774430509465;
      m->Theta[2][4] = 12.33727431701306195581826123274001;
      m->Theta[3][4] = -11.24557012450885560524143016037523;
      m->Alpha[0][1] = 2.875000000000000000000000000000000;
      m->Alpha[0][2] = 0.8500000000000000000000000000000000;
      m->Alpha[1][2] = 0.4434782608695652173913043478260870;
      m->Alpha[0][3] = 0.7352046091658870564637910527807370;
      m->Alpha[1][3] = -0.9525565003057343527941920657462074e-01;
      m->Alpha[2][3] = 0.4290111305453813852259481840631738;
      m->Alpha[0][4] = -16.10898993405067684831655675112808;
      m->Alpha[1][4] = 6.559571569643355712998131800797873;
      m->Alpha[2][4] = -15.90772144271326504260996815012482;
      m->Alpha[3][4] = 25.34908987169226073668861694892683;
      return 1;
    case 5:  /* Sdirk4b */
      m->S = 5;
      m->Gamma = 0.25;
      m->C[0] = 0.25; m->C[1] = 0.75; m->C[2] = 0.55; m->C[3] = 0.5; m->C[4] = 1.0;
      m->D[4] = 1.0;
      m->E[0] = 0.5750; m->E[1] = 0.2125; m->E[2] = -4.6875; m->E[3] = 4.2500;
      m->E[4] = 0.1500;
      m->ELO = 4;
      m->Theta[0][1] = 2.0;
      m->Theta[0][2] = 1.680000000000000000000000000000000;
      m->Theta[1][2] = -0.1600000000000000000000000000000000;
      m->Theta[0][3] = 1.308823529411764705882352941176471;
      m->Theta[1][3] = -0.1838235294117647058823529411764706;
      m->Theta[2][3] = 0.1102941176470588235294117647058824;
      m->Theta[0][4] = -3.083333333333333333333333333333333;
      m->Theta[1][4] = -4.291666666666666666666666666666667;
      m->Theta[2][4] = 34.37500000000000000000000000000000;
      m->Theta[3][4] = -28.3333333333333333333333333333;
      m->Alpha[0][1] = 3.0;
      m->Alpha[0][2] = 0.8800000000000000000000000000000000;
      m->Alpha[1][2] = 0.4400000000000000000000000000000000;
      m->Alpha[0][3] = 0.1666666666666666666666666666666667;
      m->Alpha[1][3] = -0.8333333333333333333333333333333333e-01;
      m->Alpha[2][3] = 0.9469696969696969696969696969696970;
      m->Alpha[0][4] = -6.0;
      m->Alpha[1][4] = 9.0;
      m->Alpha[2][4] = -56.81818181818181818181818181818182;
      m->Alpha[3][4] = 54.0;
      return 1;
  }
}

/* SIGN(x,y) of int/sdirk.c:5, unparenthesised: ((x*y) >= 0) ? x : -x */
#define SD_SIGN(x, y) ((((x) * (y)) >= 0) ? (x) : (-(x)))

static void sdirk_scale(const double* AbsTol, const double* RelTol, int ITOL, const double* Y,
                        double* SCAL) {
  for (int i = 0; i < KPP_NVAR; i++)
    SCAL[i] = (ITOL == 0) ? ONE / (AbsTol[0] + RelTol[0] * ABS(Y[i]))
                          : ONE / (AbsTol[i] + RelTol[i] * ABS(Y[i]));
}
static double sdirk_norm(const double* Y, const double* SCAL) {
  double Err = ZERO;
  for (int i = 0; i < KPP_NVAR; i++) Err = Err + pow(Y[i] * SCAL[i], 2);
  return MAX(SQRT(Err / (double)KPP_NVAR), 1.0e-10);
}
static void WADD(const double* X, const double* Y, double* Z) {
  for (int i = 0; i < KPP_NVAR; i++) Z[i] = X[i] + Y[i];
}
static void sdirk_solve(ros_cell* c, const double* E, double H, double gam, double* RHS) {
  WSCAL(KPP_NVAR, ONE / (H * gam), RHS);
  kpp_KppSolve(E, RHS);
  c->st.Nsol++;
}

/* SDIRK_PrepareMatrix with its by-value H, SkipJac and ISING */
static void sdirk_prepare(ros_cell* c, double H, double T, const double* Y, double* FJAC,
                          double* E, int SkipJac, double gam) {
  int ISING = 1, cons = 0;
  while (ISING != 0) {
    const double HGammaInv = ONE / (H * gam);
    if (SkipJac == 0) JacTemplate(c, T, Y, FJAC);
    for (int i = 0; i < KPP_LU_NONZERO; i++) E[i] = -FJAC[i];
    for (int i = 0; i < KPP_NVAR; i++) E[kpp_LU_DIAG[i]] = E[kpp_LU_DIAG[i]] + HGammaInv;
    ISING = KppDecomp(E);
    c->st.Ndec++;
    if (ISING != 0) {
      c->st.Nsng++;
      cons++;
      if (cons >= 6) return;
      H = 0.5 * H;
      SkipJac = 1;
    }
  }
}

typedef struct {
  double Hmin, Hmax, Hstart, FacMin, FacMax, FacRej, FacSafe, ThetaMin, NewtonTol, Qmin, Qmax;
  int Max_no_steps, NewtonMaxit, StartNewton, ITOL;
} sdirk_opts;

static int SDIRK_Integrator(ros_cell* c, double Tinitial, double Tfinal, double* Y,
                            const double* AbsTol, const double* RelTol, const sdirk_method* m,
                            const sdirk_opts* o, double Roundoff, double* rstat) {
  double Z[5][KPP_NVAR];
  double G[KPP_NVAR], TMP[KPP_NVAR], SCAL[KPP_NVAR], RHS[KPP_NVAR], FJAC[KPP_LU_NONZERO],
      E[KPP_LU_NONZERO];
  double T, H, Theta = 0, Hratio, NewtonRate = 0, NewtonPredictedErr, Qnewton, Err = 0, Fac,
               Hnew, Tdirection, NewtonIncrement = 0, NewtonIncrementOld = 0;
  int Reject, FirstStep, SkipJac, SkipLU, NewtonDone, CycleTloop, Nrun = 0;
  const int N = KPP_NVAR;

  T = Tinitial;
  Tdirection = SD_SIGN(ONE, Tfinal - Tinitial);
  H = MAX(ABS(o->Hmin), ABS(o->Hstart));
  if (ABS(H) <= 10.0 * Roundoff) H = 1.0e-06;
  H = MIN(ABS(H), o->Hmax);
  H = SD_SIGN(H, Tdirection);
  SkipLU = 0;
  SkipJac = 0;
  Reject = 0;
  FirstStep = 1;
  CycleTloop = 0;
  c->Yacc = Y;
  sdirk_scale(AbsTol, RelTol, o->ITOL, Y, SCAL);

  while ((Tfinal - T) * Tdirection - Roundoff > ZERO) {
    if (SkipLU == 0) sdirk_prepare(c, H, T, Y, FJAC, E, SkipJac, m->Gamma);
    if (c->st.Nstp > o->Max_no_steps) return -6;
    if ((T + 0.1 * H == T) || (ABS(H) <= Roundoff)) return -7;
    for (int is = 0; is < m->S; is++) {
      for (int i = 0; i < N; i++) Z[is][i] = ZERO;
      for (int i = 0; i < N; i++) G[i] = ZERO;
      for (int j = 0; j < is; j++) {
        WAXPY(N, m->Theta[j][is], Z[j], G);
        if (o->StartNewton == 1) WAXPY(N, m->Alpha[j][is], Z[j], Z[is]);
      }
      NewtonDone = 0;
      Fac = 0.5;
      for (int it = 0; it < o->NewtonMaxit; it++) {
        WADD(Y, Z[is], TMP);
        FunTemplate(c, T + m->C[is] * H, TMP, RHS);
        WSCAL(N, H * m->Gamma, RHS);
        WAXPY(N, -ONE, Z[is], RHS);
        WAXPY(N, ONE, G, RHS);
        sdirk_solve(c, E, H, m->Gamma, RHS);
        NewtonIncrement = sdirk_norm(RHS, SCAL);
        if (it == 0) {
          Theta = ABS(o->ThetaMin);
          NewtonRate = 2.0;
        } else {
          Theta = NewtonIncrement / NewtonIncrementOld;
          if (Theta < 0.99) {
            NewtonRate = Theta / (ONE - Theta);
            NewtonPredictedErr =
                NewtonIncrement * pow(Theta, (o->NewtonMaxit - (it + 1)) / (ONE - Theta));
            if (NewtonPredictedErr >= o->NewtonTol) {
              Qnewton = MIN(10.0, NewtonPredictedErr / o->NewtonTol);
              Fac = 0.8 * pow(Qnewton, (-ONE / (1 + o->NewtonMaxit - it + 1)));
              break;
            }
          } else {
            break;
          }
        }
        NewtonIncrementOld = NewtonIncrement;
        WAXPY(N, ONE, RHS, Z[is]);
        NewtonDone = (NewtonRate * NewtonIncrement <= o->NewtonTol);
        if (NewtonDone == 1) break;
      }
      if (NewtonDone == 0) {
        H = Fac * H;
        Reject = 1;
        SkipJac = 1;
        SkipLU = 0;
        CycleTloop = 1;
      }
      if (CycleTloop == 1) break;
    }
    if (CycleTloop == 1) {
      CycleTloop = 0;
      if (!(ABS(H) <= DBL_MAX)) return IERR_NONFINITE;
      if (++Nrun > MAX_REJECT) return IERR_REJECTS;
      continue;
    }
    c->st.Nstp++;
    for (int i = 0; i < N; i++) TMP[i] = ZERO;
    for (int i = 0; i < m->S; i++)
      if (m->E[i] != ZERO) WAXPY(N, m->E[i], Z[i], TMP);
    sdirk_solve(c, E, H, m->Gamma, TMP);
    Err = sdirk_norm(TMP, SCAL);
    Fac = o->FacSafe * pow(Err, (-ONE / m->ELO));
    Fac = MAX(o->FacMin, MIN(o->FacMax, Fac));
    Hnew = H * Fac;
    if (Err < ONE) {
      FirstStep = 0;
      c->st.Nacc++;
      if (c->Hacc1 == ZERO) c->Hacc1 = H;
      T = T + H;
      for (int i = 0; i < m->S; i++)
        if (m->D[i] != ZERO) WAXPY(N, m->D[i], Z[i], Y);
      sdirk_scale(AbsTol, RelTol, o->ITOL, Y, SCAL);
      Hnew = Tdirection * MIN(ABS(Hnew), o->Hmax);
      rstat[0] = T;
      rstat[1] = H;
      rstat[2] = Hnew;
      if (Reject == 1) Hnew = Tdirection * MIN(ABS(Hnew), ABS(H));
      Reject = 0;
      if ((T + Hnew / o->Qmin - Tfinal) * Tdirection > ZERO) {
        H = Tfinal - T;
      } else {
        Hratio = Hnew / H;
        SkipLU = ((Theta <= o->ThetaMin) && (Hratio >= o->Qmin) && (Hratio <= o->Qmax));
        if (SkipLU == 0) H = Hnew;
      }
      SkipJac = 0;
      Nrun = 0;
    } else {
      if ((FirstStep == 1) || (Reject == 1)) H = o->FacRej * H;
      else H = Hnew;
      Reject = 1;
      SkipJac = 1;
      SkipLU = 0;
      if (c->st.Nacc >= 1) c->st.Nrej++;
      if (Err != Err || !(ABS(H) <= DBL_MAX)) return IERR_NONFINITE;
      if (++Nrun > MAX_REJECT) return IERR_REJECTS;
    }
  }
  return 1;
}

/* SDIRK, int/sdirk.c:170-525: controls, defaults, checks */
static int SDIRK_cell(ros_cell* c, double Tinitial, double Tfinal, double* Y,
                      const double* AbsTol, const double* RelTol, const double* RCNTRL,
                      const int* ICNTRL, int* istatus, double* rstat) {
  sdirk_method m;
  sdirk_opts o;
  const double Roundoff = ros_ref_wlamch();
  const double span = ABS(Tfinal - Tinitial);
  int IERR = 0;
  memset(&c->st, 0, sizeof c->st);
  rstat[0] = rstat[1] = rstat[2] = ZERO;
  o.ITOL = (ICNTRL[1] == 0) ? 1 : 0;
  sdirk_method_of(ICNTRL[2], &m);
  if (ICNTRL[3] == 0) o.Max_no_steps = 200000;
  else if (ICNTRL[3] > 0) o.Max_no_steps = ICNTRL[3];
  else IERR = -1;
  if (!IERR) {
    if (ICNTRL[4] == 0) o.NewtonMaxit = 8;
    else if (ICNTRL[4] > 0) o.NewtonMaxit = ICNTRL[4];
    else IERR = -2;
  }
  o.StartNewton = (ICNTRL[5] == 0) ? 1 : 0;
  if (!IERR) {
    if (RCNTRL[0] == ZERO) o.Hmin = ZERO;
    else if (RCNTRL[0] > ZERO) o.Hmin = RCNTRL[0];
    else IERR = -3;
  }
  if (!IERR) {
    if (RCNTRL[1] == ZERO) o.Hmax = span;
    else if (RCNTRL[1] > ZERO) o.Hmax = MIN(ABS(RCNTRL[1]), span);
    else IERR = -3;
  }
  if (!IERR) {
    if (RCNTRL[2] == ZERO) o.Hstart = MAX(o.Hmin, Roundoff);
    else if (RCNTRL[2] > ZERO) o.Hstart = MIN(ABS(RCNTRL[2]), span);
    else IERR = -3;
  }
  if (!IERR) {
    const double def[4] = {0.2, 10.0, 0.1, 0.9};
    double* dst[4] = {&o.FacMin, &o.FacMax, &o.FacRej, &o.FacSafe};
    for (int k = 0; k < 4 && !IERR; k++) {
      if (RCNTRL[3 + k] == ZERO) *dst[k] = def[k];
      else if (RCNTRL[3 + k] > ZERO) *dst[k] = RCNTRL[3 + k];
      else IERR = -4;
    }
  }
  o.ThetaMin = (RCNTRL[7] == ZERO) ? 1.0e-03 : RCNTRL[7];
  o.NewtonTol = (RCNTRL[8] == ZERO) ? 3.0e-02 : RCNTRL[8];
  o.Qmin = (RCNTRL[9] == ZERO) ? ONE : RCNTRL[9];
  o.Qmax = (RCNTRL[10] == ZERO) ? 1.2 : RCNTRL[10];
  if (!IERR) {
    /* int/sdirk.c:491-503 as written: the scalar test compares the bool
     * (AbsTol <= 0 || RelTol != 0) with 10*Roundoff */
    if (o.ITOL == 0) {
      if ((AbsTol[0] <= ZERO || RelTol[0]) <= (10.0 * Roundoff)) IERR = -5;
    } else {
      for (int i = 0; i < KPP_NVAR; i++)
        if ((AbsTol[i] <= ZERO) || (RelTol[i] <= 10.0 * Roundoff)) IERR = -5;
    }
  }
  if (!IERR)
    IERR = SDIRK_Integrator(c, Tinitial, Tfinal, Y, AbsTol, RelTol, &m, &o, Roundoff, rstat);
  istatus[0] = c->st.Nfun; istatus[1] = c->st.Njac; istatus[2] = c->st.Nstp;
  istatus[3] = c->st.Nacc; istatus[4] = c->st.Nrej; istatus[5] = c->st.Ndec;
  istatus[6] = c->st.Nsol; istatus[7] = c->st.Nsng;
  return IERR;
}

/* SDIRK method table for tests: S, Gamma, C[5], D[5], E[5], ELO,
 * Theta[5][5], Alpha[5][5] (row j, column i as rkTheta[j][i]); returns 1 for
 * codes 0..5, 2 for an unknown code (Sdirk2a, as the switch's default) */
int sdirk_ref_method(int code, int* S, double* Gamma, double* C, double* D, double* E,
                     double* ELO, double* Theta, double* Alpha) {
  sdirk_method m;
  int r = sdirk_method_of(code, &m);
  *S = m.S;
  *Gamma = m.Gamma;
  *ELO = m.ELO;
  for (int i = 0; i < 5; i++) {
    C[i] = m.C[i]; D[i] = m.D[i]; E[i] = m.E[i];
    for (int j = 0; j < 5; j++) {
      Theta[j * 5 + i] = m.Theta[j][i];
      Alpha[j * 5 + i] = m.Alpha[j][i];
    }
  }
  return r;
}

/* Batched SDIRK (int/sdirk.c INTEGRATE/SDIRK) over SoA concentrations:
 * cell c over [tin[c], tout[c]] with ICNTRL/RCNTRL as SDIRK takes them
 * (0-based C arrays, no merge), sun_mode as ros_ref_batch, rconst or NULL.
 * istatus [8][n] = ISTATUS[1..8], rstatus [3][n] = RSTATUS[1..3]. */
void sdirk_ref_batch(int ncells, double* C, const double* temp, const double* cfactor,
                     const double* sun, int sun_mode, const double* tin, const double* tout,
                     const double* atol, const double* rtol, const int* ICNTRL,
                     const double* RCNTRL, const double* rconst, int* ierr, int* istatus,
                     double* rstatus, int nthreads) {
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
    c.sun_mode = sun_mode;
    for (int i = 0; i < KPP_NVAR; i++) Y[i] = C[(size_t)i * ncells + cidx];
    for (int i = 0; i < KPP_NFIX; i++) c.FIX[i] = C[(size_t)(KPP_NVAR + i) * ncells + cidx];
    if (rconst) {
      c.rates_given = 1;
      for (int i = 0; i < KPP_NREACT; i++) c.RCONST[i] = rconst[(size_t)i * ncells + cidx];
    }
    int e = SDIRK_cell(&c, tin[cidx], tout[cidx], Y, atol, rtol, RCNTRL, ICNTRL, ist, rst);
    for (int i = 0; i < KPP_NVAR; i++) C[(size_t)i * ncells + cidx] = Y[i];
    ierr[cidx] = e;
    if (istatus)
      for (int i = 0; i < 8; i++) istatus[(size_t)i * ncells + cidx] = ist[i];
    if (rstatus)
      for (int i = 0; i < 3; i++) rstatus[(size_t)i * ncells + cidx] = rst[i];
  }
}
