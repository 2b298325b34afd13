/* rk_ref.h -- CPU oracle of KPP's C fully implicit 3-stage Runge-Kutta
 * integrator (int/runge_kutta.c).
 *
 * TEST INFRASTRUCTURE ONLY (included by ros_ref.c, whose header says who may
 * load the library).  A reentrant restatement, one cell per call:
 *   RungeKutta             int/runge_kutta.c:203-611 (defaults and checks)
 *   RK_Integrator          int/runge_kutta.c:616-1001
 *   RK_ErrorScale          int/runge_kutta.c:1063-1083
 *   RK_Interpolate         int/runge_kutta.c:1108-1149
 *   RK_PrepareRHS          int/runge_kutta.c:1152-1187
 *   RK_Decomp              int/runge_kutta.c:1190-1246
 *   RK_Solve               int/runge_kutta.c:1249-1295
 *   RK_ErrorEstimate       int/runge_kutta.c:1298-1351
 *   RK_ErrorNorm           int/runge_kutta.c:1354-1364
 *   Radau2A/Lobatto3C/Gauss/Radau1A coefficients  int/runge_kutta.c:1369-1863
 *   KppDecompCmplxR/KppSolveCmplxR                util/sutil.c:28-87
 *   FUN_CHEM/JAC_CHEM      int/runge_kutta.c:1997-2024 (Update_SUN and
 *                          Update_RCONST before every call)
 *
 * Kept from the reference: ICNTRL[2] 0/1 Radau-2A, 2 Lobatto-3C, 3 Gauss,
 * 4 Radau-1A (the switch, not the comment above it); every parameter check
 * runs and the last failing one's code is returned; Hmin/Hstart taken with
 * ABS (no -3); the SDIRK error estimator's b0 chosen by ICNTRL[9] before the
 * tables are filled; the solves of RK_ErrorEstimate and of the SDIRK stage
 * are not counted in Nsol; SkipLU left as it was when the step is cut to
 * Tend; the ABS and SIGN macros unparenthesised (ABS(Tend-T) is -Tend-T for
 * Tend < T).
 * The reference only prints its -10 (step size) and -12 (five singular
 * matrices in a row) messages and goes on, returning 1 if the loop ends;
 * so does the restatement.  Deviations, where the reference would not end:
 *  - -9 (more than Max_no_steps steps) ends the cell (the reference prints
 *    it and goes on);
 *  - a NaN error estimate or a non-finite step ends the cell with -14, and
 *    more than MAX_REJECT passes in a row in which T does not advance with
 *    -15 (codes beyond the reference's -1..-13);
 *  - ICNTRL[2] = 5 (Lobatto3A_Coefficients, listed as "not yet implemented"
 *    by the reference, whose 4-stage table the 3-stage integrator uses only
 *    in part) is -13, as an unknown method. */

#define RK_R2A 1
#define RK_R1A 2
#define RK_L3C 3
#define RK_GAU 4
#define RK_IERR_NONFINITE (-14)
#define RK_IERR_REJECTS (-15)

typedef struct {
  int kind;
  double A[3][3], B[3], C[3], D[3], E[4], Bhat[4], Bgam[5], Theta[3];
  double Gamma, Alpha, Beta, ELO, T[3][3], TinvAinv[3][3];
} rk_method;

static void rk_set3(double dst[3][3], const double src[9]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) dst[i][j] = src[3 * i + j];
}

/* the coefficient routines; b0 tunes the error estimators (values as the
 * reference computes them at run time, operation for operation) */
static int rk_method_of(int code, int sdirk_error, rk_method* m) {
  memset(m, 0, sizeof *m);
  double b0;
  switch (code) {
    case 0:
    case 1: { /* Radau-2A, int/runge_kutta.c:1369-1493 */
      static const double A[9] = {
          1.968154772236604258683861429918299e-01, -6.55354258501983881085227825696087e-02,
          2.377097434822015242040823210718965e-02, 3.944243147390872769974116714584975e-01,
          2.920734116652284630205027458970589e-01, -4.154875212599793019818600988496743e-02,
          3.764030627004672750500754423692808e-01, 5.124858261884216138388134465196080e-01,
          1.111111111111111111111111111111111e-01};
      static const double T[9] = {
          9.443876248897524148749007950641664e-02, -1.412552950209542084279903838077973e-01,
          -3.00291941051474244918611170890539e-02, 2.502131229653333113765090675125018e-01,
          2.041293522937999319959908102983381e-01, 3.829421127572619377954382335998733e-01,
          1.0, 1.0, 0.0};
      static const double TA[9] = {
          1.520148562492775501049204957366528e+01, 1.192055789400527921212348994770778,
          1.903956760517560343018332287285119, -9.669512977505946748632625374449567,
          -8.724028436822336183071773193986487, 3.096043239482439656981667712714881,
          -1.409513259499574544876303981551774e+01, 5.895975725255405108079130152868952,
          -1.441236197545344702389881889085515e-01};
      b0 = sdirk_error == 1 ? 0.2e-01 : 0.5e-01;
      m->kind = RK_R2A;
      rk_set3(m->A, A);
      rk_set3(m->T, T);
      rk_set3(m->TinvAinv, TA);
      m->B[0] = A[6]; m->B[1] = A[7]; m->B[2] = A[8];
      m->C[0] = 1.550510257216821901802715925294109e-01;
      m->C[1] = 6.449489742783178098197284074705891e-01;
      m->C[2] = ONE;
      m->D[2] = ONE;
      m->E[0] = ONE * b0;
      m->E[1] = -10.04880939982741556246032950764708 * b0;
      m->E[2] = 1.382142733160748895793662840980412 * b0;
      m->E[3] = -.3333333333333333333333333333333333 * b0;
      m->Bgam[0] = b0;
      m->Bgam[1] = .3764030627004672750500754423692807 - 1.558078204724922382431975370686279 * b0;
      m->Bgam[2] = .8914115380582557157653087040196118 * b0 + .5124858261884216138388134465196077;
      m->Bgam[3] = -.1637777184845662566367174924883037 - .3333333333333333333333333333333333 * b0;
      m->Bgam[4] = .2748888295956773677478286035994148;
      m->Theta[0] = -1.520677486405081647234271944611547 - 10.04880939982741556246032950764708 * b0;
      m->Theta[1] = 2.070455145596436382729929151810376 + 1.382142733160748895793662840980413 * b0;
      m->Theta[2] = -.3333333333333333333333333333333333 * b0 - .3744441479783868387391430179970741;
      m->ELO = (b0 == ZERO) ? 6.0 : 4.0;
      m->Gamma = 3.637834252744495732208418513577775;
      m->Alpha = 2.681082873627752133895790743211112;
      m->Beta = 3.050430199247410569426377624787569;
      return 1;
    }
    case 2: { /* Lobatto-3C, int/runge_kutta.c:1499-1617 */
      static const double A[9] = {
          .1666666666666666666666666666666667, -.3333333333333333333333333333333333,
          .1666666666666666666666666666666667, .1666666666666666666666666666666667,
          .4166666666666666666666666666666667, -.8333333333333333333333333333333333e-01,
          .1666666666666666666666666666666667, .6666666666666666666666666666666667,
          .1666666666666666666666666666666667};
      static const double T[9] = {
          1.0, 1.0, 0.0, .4554100411010284672111720348287483,
          -.6027050205505142336055860174143743, -.4309321229203225731070721341350346,
          2.195823345445647152832799205549709, -1.097911672722823576416399602774855,
          .7850032632435902184104551358922130};
      static const double TA[9] = {
          1.104302803159744452668648155627548, .916122120694355522658740710823143,
          .5029105849749601702795812241441172, 1.895697196840255547331351844372453,
          3.083877879305644477341259289176857, -1.502910584974960170279581224144117,
          .8362439183082935036129145574774502, -3.344975673233174014451658229909802,
          .312908409479233358005944466882642};
      b0 = sdirk_error == 1 ? 0.2 : 0.5;
      m->kind = RK_L3C;
      rk_set3(m->A, A);
      rk_set3(m->T, T);
      rk_set3(m->TinvAinv, TA);
      m->B[0] = A[6]; m->B[1] = A[7]; m->B[2] = A[8];
      m->C[0] = ZERO; m->C[1] = 0.5; m->C[2] = ONE;
      m->Bhat[0] = b0;
      m->Bhat[1] = .16666666666666666666666666666666667 - b0;
      m->Bhat[2] = .66666666666666666666666666666666667;
      m->Bhat[3] = .16666666666666666666666666666666667;
      m->D[2] = ONE;
      m->E[0] = .3808338772072650364017425226487022 * b0;
      m->E[1] = -1.142501631621795109205227567946107 * b0;
      m->E[2] = -1.523335508829060145606970090594809 * b0;
      m->E[3] = .3808338772072650364017425226487022 * b0;
      m->Bgam[0] = b0;
      m->Bgam[1] = .1666666666666666666666666666666667 - 1.0 * b0;
      m->Bgam[2] = .6666666666666666666666666666666667;
      m->Bgam[3] = -.2141672105405983697350758559820354;
      m->Bgam[4] = .3808338772072650364017425226487021;
      m->Theta[0] = -3.0 * b0 - .3808338772072650364017425226487021;
      m->Theta[1] = -4.0 * b0 + 1.523335508829060145606970090594808;
      m->Theta[2] = -.142501631621795109205227567946106 + b0;
      m->ELO = (b0 == ZERO) ? 5.0 : 4.0;
      m->Gamma = 2.625816818958466716011888933765284;
      m->Alpha = 1.687091590520766641994055533117359;
      m->Beta = 2.508731754924880510838743672432351;
      return 1;
    }
    case 3: { /* Gauss, int/runge_kutta.c:1623-1738 */
      static const double A[9] = {
          .1388888888888888888888888888888889, -.359766675249389034563954710966045e-01,
          .97894440153083260495800422294756e-02, .3002631949808645924380249472131556,
          .2222222222222222222222222222222222, -.224854172030868146602471694353778e-01,
          .2679883337624694517281977355483022, .4804211119693833479008399155410489,
          .1388888888888888888888888888888889};
      static const double T[9] = {
          .7215185205520017032081769924397664e-01, -.8224123057363067064866206597516454e-01,
          -.6012073861930850173085948921439054e-01, .1188325787412778070708888193730294,
          .5306509074206139504614411373957448e-01, .3162050511322915732224862926182701,
          1.0, 1.0, 0.0};
      static const double TA[9] = {
          27.82766708436744962047620566703329, 5.290933503982655311815946575100597,
          2.007817718512643701322151051660114, -17.66368928942422710690385180065675,
          -14.45491129892587782538830044147713, 2.992182281487356298677848948339886,
          -25.60678350282974256072419392007303, 6.762434375611708328910623303779923,
          1.043979339483109825041215970036771};
      b0 = 0.1;
      m->kind = RK_GAU;
      rk_set3(m->A, A);
      rk_set3(m->T, T);
      rk_set3(m->TinvAinv, TA);
      m->B[0] = .2777777777777777777777777777777778;
      m->B[1] = .4444444444444444444444444444444444;
      m->B[2] = .2777777777777777777777777777777778;
      m->C[0] = .1127016653792583114820734600217600;
      m->C[1] = .5000000000000000000000000000000000;
      m->C[2] = .8872983346207416885179265399782400;
      m->Bhat[0] = b0;
      m->Bhat[1] = -1.4788305577012361475298775666303999 * b0 + .27777777777777777777777777777777778;
      m->Bhat[2] = .44444444444444444444444444444444444 + .66666666666666666666666666666666667 * b0;
      m->Bhat[3] = -.18783610896543051913678910003626672 * b0 + .27777777777777777777777777777777778;
      m->D[0] = .1666666666666666666666666666666667e+01;
      m->D[1] = -.1333333333333333333333333333333333e+01;
      m->D[2] = .1666666666666666666666666666666667e+01;
      m->E[0] = .2153144231161121782447335303806954 * b0;
      m->E[1] = -2.825278112319014084275808340593191 * b0;
      m->E[2] = .2870858974881495709929780405075939 * b0;
      m->E[3] = -.4558086256248162565397206448274867e-01 * b0;
      m->Bgam[0] = ZERO;
      m->Bgam[1] = .2373339543355109188382583162660537;
      m->Bgam[2] = .5879873931885192299409334646982414;
      m->Bgam[3] = -.4063577064014232702392531134499046e-01;
      m->Bgam[4] = .2153144231161121782447335303806955;
      m->Theta[0] = -2.594040933093095272574031876464493;
      m->Theta[1] = 1.824611539036311947589425112250199;
      m->Theta[2] = .1856563166634371860478043996459493;
      m->ELO = 4.0;
      m->Gamma = 4.644370709252171185822941421408064;
      m->Alpha = 3.677814645373914407088529289295970;
      m->Beta = 3.508761919567443321903661209182446;
      return 1;
    }
    case 4: { /* Radau-1A, int/runge_kutta.c:1746-1857 */
      static const double A[9] = {
          .1111111111111111111111111111111111, -.1916383190435098943442935597058829,
          .8052720793239878323318244859477174e-01, .1111111111111111111111111111111111,
          .2920734116652284630205027458970589, -.481334970546573839513422644787591e-01,
          .1111111111111111111111111111111111, .5370223859435462728402311533676479,
          .1968154772236604258683861429918299};
      static const double T[9] = {
          .424293819848497965354371036408369, -.3235571519651980681202894497035503,
          -.522137786846287839586599927945048, .57594609499806128896291585429339e-01,
          .3148663231849760131614374283783e-02, .452429247674359778577728510381731,
          1.0, 1.0, 0.0};
      static const double TA[9] = {
          4.487354449794728738538663081025420, 5.178748573958397475446442544234494,
          1.435609490412123627047824222335563, -2.854361287939276673073807031221493,
          -1.003648660720543859000994063139137e+01, 1.789135380979465422050817815017383,
          -4.160768067752685525282947313530352, 1.124128569859216916690209918405860,
          1.700644430961823796581896350418417};
      b0 = 0.1;
      m->kind = RK_R1A;
      rk_set3(m->A, A);
      rk_set3(m->T, T);
      rk_set3(m->TinvAinv, TA);
      m->B[0] = .1111111111111111111111111111111111;
      m->B[1] = .5124858261884216138388134465196080;
      m->B[2] = .3764030627004672750500754423692808;
      m->C[0] = ZERO;
      m->C[1] = .3550510257216821901802715925294109;
      m->C[2] = .8449489742783178098197284074705891;
      m->Bhat[0] = b0;
      m->Bhat[1] = .11111111111111111111111111111111111 - b0;
      m->Bhat[2] = .51248582618842161383881344651960810;
      m->Bhat[3] = .37640306270046727505007544236928079;
      m->D[0] = .3333333333333333333333333333333333;
      m->D[1] = -.8914115380582557157653087040196127;
      m->D[2] = 1.558078204724922382431975370686279;
      m->E[0] = .2748888295956773677478286035994148 * b0;
      m->E[1] = -1.374444147978386838739143017997074 * b0;
      m->E[2] = -1.335337922441686804550326197041126 * b0;
      m->E[3] = .235782604058977333559011782643466 * b0;
      m->Bgam[0] = ZERO;
      m->Bgam[1] = .1948150124588532186183490991130616e-01;
      m->Bgam[2] = .7575249005733381398986810981093584;
      m->Bgam[3] = -.518952314149008295083446116200793e-01;
      m->Bgam[4] = .2748888295956773677478286035994148;
      m->Theta[0] = -1.224370034375505083904362087063351;
      m->Theta[1] = .9340045331532641409047527962010133;
      m->Theta[2] = .4656990124352088397561234800640929;
      m->ELO = 4.0;
      m->Gamma = 3.637834252744495732208418513577775;
      m->Alpha = 2.681082873627752133895790743211112;
      m->Beta = 3.050430199247410569426377624787569;
      return 1;
    }
    default:
      return 0;
  }
}

/* SIGN(x,y) of int/runge_kutta.c:6 */
#define RK_SIGN(x, y) (((y) >= 0) ? (ABS(x)) : (-ABS(x)))

/* util/sutil.c:28-58 */
static int KppDecompCmplxR(double* JVSR, double* JVSI) {
  double WR[KPP_NVAR], WI[KPP_NVAR];
  for (int k = 0; k < KPP_NVAR; k++) {
    const int dk = kpp_LU_DIAG[k];
    if (JVSR[dk] == 0.0 || JVSI[dk] == 0.0) return k + 1;
    for (int kk = kpp_LU_CROW[k]; kk < kpp_LU_CROW[k + 1]; kk++) {
      WR[kpp_LU_ICOL[kk]] = JVSR[kk];
      WI[kpp_LU_ICOL[kk]] = JVSI[kk];
    }
    for (int kk = kpp_LU_CROW[k]; kk < kpp_LU_DIAG[k]; kk++) {
      const int j = kpp_LU_ICOL[kk], dj = kpp_LU_DIAG[j];
      const double den = JVSR[dj] * JVSR[dj] + JVSI[dj] * JVSI[dj];
      const double ar = -(WR[j] * JVSR[dj] + WI[j] * JVSI[dj]) / den;
      const double ai = -(WI[j] * JVSR[dj] - WR[j] * JVSI[dj]) / den;
      WR[j] = -ar;
      WI[j] = -ai;
      for (int jj = dj + 1; jj < kpp_LU_CROW[j + 1]; jj++) {
        const int c = kpp_LU_ICOL[jj];
        WR[c] = WR[c] + ar * JVSR[jj] - ai * JVSI[jj];
        WI[c] = WI[c] + ar * JVSI[jj] + ai * JVSR[jj];
      }
    }
    for (int kk = kpp_LU_CROW[k]; kk < kpp_LU_CROW[k + 1]; kk++) {
      JVSR[kk] = WR[kpp_LU_ICOL[kk]];
      JVSI[kk] = WI[kpp_LU_ICOL[kk]];
    }
  }
  return 0;
}

/* util/sutil.c:63-87 */
static void KppSolveCmplxR(const double* JVSR, const double* JVSI, double* XR, double* XI) {
  for (int i = 0; i < KPP_NVAR; i++)
    for (int j = kpp_LU_CROW[i]; j < kpp_LU_DIAG[i]; j++) {
      const int c = kpp_LU_ICOL[j];
      XR[i] = XR[i] - (JVSR[j] * XR[c] - JVSI[j] * XI[c]);
      XI[i] = XI[i] - (JVSR[j] * XI[c] + JVSI[j] * XR[c]);
    }
  for (int i = KPP_NVAR - 1; i >= 0; i--) {
    double sumr = XR[i], sumi = XI[i];
    for (int j = kpp_LU_DIAG[i] + 1; j < kpp_LU_CROW[i + 1]; j++) {
      const int c = kpp_LU_ICOL[j];
      sumr = sumr - (JVSR[j] * XR[c] - JVSI[j] * XI[c]);
      sumi = sumi - (JVSR[j] * XI[c] + JVSI[j] * XR[c]);
    }
    const int d = kpp_LU_DIAG[i];
    const double den = JVSR[d] * JVSR[d] + JVSI[d] * JVSI[d];
    XR[i] = (sumr * JVSR[d] + sumi * JVSI[d]) / den;
    XI[i] = (sumi * JVSR[d] - sumr * JVSI[d]) / den;
  }
}

static void rk_scale(int ITOL, const double* AbsTol, const double* RelTol, const double* Y,
                     double* SCAL) {
  for (int i = 0; i < KPP_NVAR; i++)
    SCAL[i] = (ITOL == 0) ? ONE / (AbsTol[0] + RelTol[0] * ABS(Y[i]))
                          : ONE / (AbsTol[i] + RelTol[i] * ABS(Y[i]));
}
static double rk_norm(const double* SCAL, const double* DY) {
  double s = ZERO;
  for (int i = 0; i < KPP_NVAR; i++) s = s + (DY[i] * SCAL[i]) * (DY[i] * SCAL[i]);
  return MAX(SQRT(s / KPP_NVAR), 1.0e-10);
}

/* RK_Interpolate "make" (:1117-1134) and "eval" (:1136-1148) */
static void rk_interp_make(const rk_method* m, const double* Z1, const double* Z2,
                           const double* Z3, double CONT[][3]) {
  const double* c = m->C;
  const double den = (c[2] - c[1]) * (c[1] - c[0]) * (c[0] - c[2]);
  for (int i = 0; i < KPP_NVAR; i++) {
    CONT[i][0] = (-c[2] * c[2] * c[1] * Z1[i] + Z3[i] * c[1] * c[0] * c[0] +
                  c[1] * c[1] * c[2] * Z1[i] - c[1] * c[1] * c[0] * Z3[i] +
                  c[2] * c[2] * c[0] * Z2[i] - Z2[i] * c[2] * c[0] * c[0]) / den - Z3[i];
    CONT[i][1] = -(c[0] * c[0] * (Z3[i] - Z2[i]) + c[1] * c[1] * (Z1[i] - Z3[i]) +
                   c[2] * c[2] * (Z2[i] - Z1[i])) / den;
    CONT[i][2] = (c[0] * (Z3[i] - Z2[i]) + c[1] * (Z1[i] - Z3[i]) + c[2] * (Z2[i] - Z1[i])) / den;
  }
}
static void rk_interp_eval(const rk_method* m, double H, double Hold, double* Z1, double* Z2,
                           double* Z3, double CONT[][3]) {
  const double r = H / Hold;
  const double x1 = ONE + m->C[0] * r, x2 = ONE + m->C[1] * r, x3 = ONE + m->C[2] * r;
  for (int i = 0; i < KPP_NVAR; i++) {
    Z1[i] = CONT[i][0] + x1 * (CONT[i][1] + x1 * CONT[i][2]);
    Z2[i] = CONT[i][0] + x2 * (CONT[i][1] + x2 * CONT[i][2]);
    Z3[i] = CONT[i][0] + x3 * (CONT[i][1] + x3 * CONT[i][2]);
  }
}

/* RK_PrepareRHS (:1152-1187) without the Lobatto-3A branch: R = Z - h A F;
 * its three function evaluations are not counted in Nfun */
static void rk_prepare_rhs(ros_cell* c, const rk_method* m, double T, double H, const double* Y,
                           const double* Z1, const double* Z2, const double* Z3, double* R1,
                           double* R2, double* R3) {
  double TMP[KPP_NVAR], F[KPP_NVAR];
  const double* Z[3] = {Z1, Z2, Z3};
  WCOPY(KPP_NVAR, Z1, R1);
  WCOPY(KPP_NVAR, Z2, R2);
  WCOPY(KPP_NVAR, Z3, R3);
  for (int j = 0; j < 3; j++) {
    for (int i = 0; i < KPP_NVAR; i++) TMP[i] = Y[i] + Z[j][i];
    /* FUN_CHEM without a count: RK_PrepareRHS has no ISTATUS */
    update_rates(c, T + m->C[j] * H);
    kpp_Fun(TMP, c->FIX, c->RCONST, F);
    WAXPY(KPP_NVAR, -H * m->A[0][j], F, R1);
    WAXPY(KPP_NVAR, -H * m->A[1][j], F, R2);
    WAXPY(KPP_NVAR, -H * m->A[2][j], F, R3);
  }
}

/* RK_Decomp (:1190-1246): E1 = Gamma/H - FJAC (real), E2 = (Alpha + i Beta)/H
 * - FJAC (complex); Ndec counts a singular E1 too */
static int rk_decomp(ros_cell* c, const rk_method* m, double H, const double* FJAC, double* E1,
                     double* E2R, double* E2I) {
  const double Gamma = m->Gamma / H, Alpha = m->Alpha / H, Beta = m->Beta / H;
  for (int i = 0; i < KPP_LU_NONZERO; i++) E1[i] = -FJAC[i];
  for (int i = 0; i < KPP_NVAR; i++) E1[kpp_LU_DIAG[i]] = E1[kpp_LU_DIAG[i]] + Gamma;
  int ising = KppDecomp(E1);
  if (ising != 0) {
    c->st.Ndec++;
    return ising;
  }
  for (int i = 0; i < KPP_LU_NONZERO; i++) {
    E2R[i] = -FJAC[i];
    E2I[i] = ZERO;
  }
  for (int i = 0; i < KPP_NVAR; i++) {
    const int j = kpp_LU_DIAG[i];
    E2R[j] = E2R[j] + Alpha;
    E2I[j] = E2I[j] + Beta;
  }
  ising = KppDecompCmplxR(E2R, E2I);
  c->st.Ndec++;
  return ising;
}

/* RK_Solve (:1249-1295) */
static void rk_solve(ros_cell* c, const rk_method* m, double H, const double* E1,
                     const double* E2R, const double* E2I, double* R1, double* R2, double* R3) {
  for (int i = 0; i < KPP_NVAR; i++) {
    const double x1 = R1[i] / H, x2 = R2[i] / H, x3 = R3[i] / H;
    R1[i] = m->TinvAinv[0][0] * x1 + m->TinvAinv[0][1] * x2 + m->TinvAinv[0][2] * x3;
    R2[i] = m->TinvAinv[1][0] * x1 + m->TinvAinv[1][1] * x2 + m->TinvAinv[1][2] * x3;
    R3[i] = m->TinvAinv[2][0] * x1 + m->TinvAinv[2][1] * x2 + m->TinvAinv[2][2] * x3;
  }
  kpp_KppSolve(E1, R1);
  KppSolveCmplxR(E2R, E2I, R2, R3);
  for (int i = 0; i < KPP_NVAR; i++) {
    const double x1 = R1[i], x2 = R2[i], x3 = R3[i];
    R1[i] = m->T[0][0] * x1 + m->T[0][1] * x2 + m->T[0][2] * x3;
    R2[i] = m->T[1][0] * x1 + m->T[1][1] * x2 + m->T[1][2] * x3;
    R3[i] = m->T[2][0] * x1 + m->T[2][1] * x2 + m->T[2][2] * x3;
  }
  c->st.Nsol++;
}

/* RK_ErrorEstimate (:1298-1351) */
static double rk_error_estimate(ros_cell* c, const rk_method* m, double H, double T,
                                const double* Y, const double* FO, const double* E1,
                                const double* Z1, const double* Z2, const double* Z3,
                                const double* SCAL, int FirstStep, int Reject) {
  double F1[KPP_NVAR], F2[KPP_NVAR], TMP[KPP_NVAR];
  const double HrkE1 = m->E[1] / H, HrkE2 = m->E[2] / H, HrkE3 = m->E[3] / H;
  for (int i = 0; i < KPP_NVAR; i++) {
    F2[i] = HrkE1 * Z1[i] + HrkE2 * Z2[i] + HrkE3 * Z3[i];
    TMP[i] = m->E[0] * FO[i] + F2[i];
  }
  kpp_KppSolve(E1, TMP);
  if (m->kind == RK_R1A || m->kind == RK_GAU) kpp_KppSolve(E1, TMP);
  if (m->kind == RK_GAU) kpp_KppSolve(E1, TMP);
  double Err = rk_norm(SCAL, TMP);
  if (Err < ONE) return Err;
  if (FirstStep == 1 || Reject == 1) {
    for (int i = 0; i < KPP_NVAR; i++) TMP[i] = Y[i] + TMP[i];
    FunTemplate(c, T, TMP, F1);
    for (int i = 0; i < KPP_NVAR; i++) TMP[i] = F1[i] + F2[i];
    kpp_KppSolve(E1, TMP);
    Err = rk_norm(SCAL, TMP);
  }
  return Err;
}

typedef struct {
  double Hmin, Hmax, Hstart, FacMin, FacMax, FacRej, FacSafe, ThetaMin, NewtonTol, Qmin, Qmax;
  int Max_no_steps, NewtonMaxit, StartNewton, Gustafsson, ITOL, SdirkError;
} rk_opts;

/* RK_Integrator (:616-1001) */
static int RK_Integrator(ros_cell* c, double* Tp, double Tend, double* Y, const double* AbsTol,
                         const double* RelTol, const rk_method* m, const rk_opts* o,
                         double Roundoff, double* rstat) {
  const int N = KPP_NVAR;
  double FJAC[KPP_LU_NONZERO], E1[KPP_LU_NONZERO], E2R[KPP_LU_NONZERO], E2I[KPP_LU_NONZERO];
  double Z1[KPP_NVAR], Z2[KPP_NVAR], Z3[KPP_NVAR], Z4[KPP_NVAR], SCAL[KPP_NVAR], DZ1[KPP_NVAR],
      DZ2[KPP_NVAR], DZ3[KPP_NVAR], DZ4[KPP_NVAR], G[KPP_NVAR], TMP[KPP_NVAR], FO[KPP_NVAR];
  double CONT[KPP_NVAR][3];
  double T = *Tp, H, Hacc = ZERO, Hnew, Hold, Fac, FacGus, Theta = ZERO, Err, ErrOld = ZERO,
         NewtonRate = ZERO, NewtonIncrement, Hratio, Qnewton, NewtonPredictedErr,
         NewtonIncrementOld = ZERO, ThetaSD;
  int NewtonIter, NewtonIterMain, Reject, FirstStep, SkipJac, NewtonDone, SkipLU;
  int Nrun = 0;

  const double Tdirection = RK_SIGN(ONE, Tend - T);
  H = MIN(MAX(ABS(o->Hmin), ABS(o->Hstart)), o->Hmax);
  if (ABS(H) <= 10.0 * Roundoff) H = 1.0e-06;
  H = RK_SIGN(H, Tdirection);
  Hold = H;
  Reject = 0;
  FirstStep = 1;
  SkipJac = 0;
  SkipLU = 0;
  if ((T + H * 1.0001 - Tend) * Tdirection >= ZERO) H = Tend - T;
  c->Yacc = Y;
  rk_scale(o->ITOL, AbsTol, RelTol, Y, SCAL);

  while ((Tend - T) * Tdirection - Roundoff > ZERO) {
    FunTemplate(c, T, Y, FO);
    if (SkipLU == 0) {
      if (SkipJac == 0) JacTemplate(c, T, Y, FJAC);
      const int ising = rk_decomp(c, m, H, FJAC, E1, E2R, E2I);
      if (ising != 0) {
        c->st.Nsng++;
        /* the reference counts singular matrices in a row only to print -12
         * from the fifth on, and goes on halving */
        H = H * 0.5;
        Reject = 1;
        SkipJac = 1;
        SkipLU = 0;
        if (!(ABS(H) <= DBL_MAX) || H == ZERO) { *Tp = T; return RK_IERR_NONFINITE; }
        if (++Nrun > MAX_REJECT) { *Tp = T; return RK_IERR_REJECTS; }
        continue;
      }
    }
    c->st.Nstp++;
    if (c->st.Nstp > o->Max_no_steps) { *Tp = T; return -9; }
    /* 0.1*ABS(H) <= ABS(T)*Roundoff: the reference prints -10 and goes on */

    /* simplified Newton iterations */
    if (FirstStep == 1 || o->StartNewton == 0) {
      for (int i = 0; i < N; i++) Z1[i] = Z2[i] = Z3[i] = ZERO;
    } else {
      rk_interp_eval(m, H, Hold, Z1, Z2, Z3, CONT);
    }
    NewtonDone = 0;
    Fac = 0.5;
    for (NewtonIter = 1; NewtonIter <= o->NewtonMaxit; NewtonIter++) {
      rk_prepare_rhs(c, m, T, H, Y, Z1, Z2, Z3, DZ1, DZ2, DZ3);
      rk_solve(c, m, H, E1, E2R, E2I, DZ1, DZ2, DZ3);
      const double n1 = rk_norm(SCAL, DZ1), n2 = rk_norm(SCAL, DZ2), n3 = rk_norm(SCAL, DZ3);
      NewtonIncrement = SQRT((n1 * n1 + n2 * n2 + n3 * n3) / 3.0);
      if (NewtonIter == 1) {
        Theta = ABS(o->ThetaMin);
        NewtonRate = 2.0;
      } else {
        Theta = NewtonIncrement / NewtonIncrementOld;
        if (Theta < 0.99) {
          NewtonRate = Theta / (ONE - Theta);
        } else {
          break;
        }
        if (NewtonIter < o->NewtonMaxit) {
          NewtonPredictedErr =
              NewtonIncrement * pow(Theta, (o->NewtonMaxit - NewtonIter)) / (ONE - Theta);
          if (NewtonPredictedErr >= o->NewtonTol) {
            Qnewton = MIN(10.0, NewtonPredictedErr / o->NewtonTol);
            Fac = 0.8 * pow(Qnewton, (-ONE / (1 + o->NewtonMaxit - NewtonIter)));
            break;
          }
        }
      }
      NewtonIncrementOld = MAX(NewtonIncrement, Roundoff);
      WAXPY(N, -ONE, DZ1, Z1);
      WAXPY(N, -ONE, DZ2, Z2);
      WAXPY(N, -ONE, DZ3, Z3);
      NewtonDone = (NewtonRate * NewtonIncrement <= o->NewtonTol);
      if (NewtonDone == 1) break;
    }
    NewtonIterMain = NewtonIter;
    if (NewtonDone == 0) {
      H = Fac * H;
      Reject = 1;
      SkipJac = 1;
      SkipLU = 0;
      if (!(ABS(H) <= DBL_MAX)) { *Tp = T; return RK_IERR_NONFINITE; }
      if (++Nrun > MAX_REJECT) { *Tp = T; return RK_IERR_REJECTS; }
      continue;
    }

    /* SDIRK stage */
    if (o->SdirkError == 1) {
      for (int i = 0; i < N; i++) Z4[i] = Z3[i];
      for (int i = 0; i < N; i++) G[i] = ZERO;
      WAXPY(N, m->Bgam[0] * H, FO, G);
      WAXPY(N, m->Theta[0], Z1, G);
      WAXPY(N, m->Theta[1], Z2, G);
      WAXPY(N, m->Theta[2], Z3, G);
      NewtonDone = 0;
      Fac = 0.5;
      for (NewtonIter = 1; NewtonIter <= o->NewtonMaxit; NewtonIter++) {
        for (int i = 0; i < N; i++) TMP[i] = Y[i] + Z4[i];
        FunTemplate(c, T + H, TMP, DZ4);
        WAXPY(N, -ONE * m->Gamma / H, Z4, DZ4);
        WAXPY(N, m->Gamma / H, G, DZ4);
        kpp_KppSolve(E1, DZ4);
        NewtonIncrement = rk_norm(SCAL, DZ4);
        if (NewtonIter == 1) {
          ThetaSD = ABS(o->ThetaMin);
          NewtonRate = 2.0;
        } else {
          ThetaSD = NewtonIncrement / NewtonIncrementOld;
          if (ThetaSD < 0.99) {
            NewtonRate = ThetaSD / (ONE - ThetaSD);
            NewtonPredictedErr =
                NewtonIncrement * pow(ThetaSD, (o->NewtonMaxit - NewtonIter)) / (ONE - ThetaSD);
            if (NewtonPredictedErr >= o->NewtonTol) {
              Qnewton = MIN(10.0, NewtonPredictedErr / o->NewtonTol);
              Fac = 0.8 * pow(Qnewton, (-ONE / (1 + o->NewtonMaxit - NewtonIter)));
              break;
            }
          } else {
            break;
          }
        }
        NewtonIncrementOld = NewtonIncrement;
        WAXPY(N, ONE, DZ4, Z4);
        NewtonDone = (NewtonRate * NewtonIncrement <= o->NewtonTol);
        if (NewtonDone == 1) break;
      }
      if (NewtonDone == 0) {
        H = Fac * H;
        Reject = 1;
        SkipJac = 1;
        SkipLU = 0;
        if (!(ABS(H) <= DBL_MAX)) { *Tp = T; return RK_IERR_NONFINITE; }
        if (++Nrun > MAX_REJECT) { *Tp = T; return RK_IERR_REJECTS; }
        continue;
      }
    }

    /* error estimation */
    if (o->SdirkError == 1) {
      for (int i = 0; i < N; i++) DZ4[i] = ZERO;
      if (m->D[0] != ZERO) WAXPY(N, m->D[0], Z1, DZ4);
      if (m->D[1] != ZERO) WAXPY(N, m->D[1], Z2, DZ4);
      if (m->D[2] != ZERO) WAXPY(N, m->D[2], Z3, DZ4);
      WAXPY(N, -ONE, Z4, DZ4);
      Err = rk_norm(SCAL, DZ4);
    } else {
      Err = rk_error_estimate(c, m, H, T, Y, FO, E1, Z1, Z2, Z3, SCAL, FirstStep, Reject);
    }

    /* new step size (BUG 4 FIX of the reference: the main loop's count) */
    Fac = pow(Err, (-ONE / m->ELO)) *
          MIN(o->FacSafe, (ONE + 2 * o->NewtonMaxit) / (NewtonIterMain + 2 * o->NewtonMaxit));
    Fac = MIN(o->FacMax, MAX(o->FacMin, Fac));
    Hnew = Fac * H;

    if (Err < ONE) {
      FirstStep = 0;
      c->st.Nacc++;
      if (c->Hacc1 == ZERO) c->Hacc1 = H;
      if (o->Gustafsson == 1) {
        if (c->st.Nacc > 1) {
          FacGus = o->FacSafe * (H / Hacc) * pow(Err * Err / ErrOld, -0.25);
          FacGus = MIN(o->FacMax, MAX(o->FacMin, FacGus));
          Fac = MIN(Fac, FacGus);
          Hnew = Fac * H;
        }
        Hacc = H;
        ErrOld = MAX(1.0e-02, Err);
      }
      Hold = H;
      const double Tprev = T;
      T = T + H;
      if (m->D[0] != ZERO) WAXPY(N, m->D[0], Z1, Y);
      if (m->D[1] != ZERO) WAXPY(N, m->D[1], Z2, Y);
      if (m->D[2] != ZERO) WAXPY(N, m->D[2], Z3, Y);
      if (o->StartNewton == 1) rk_interp_make(m, Z1, Z2, Z3, CONT);
      rk_scale(o->ITOL, AbsTol, RelTol, Y, SCAL);
      rstat[0] = T;
      rstat[2] = Hnew;
      rstat[1] = H;
      Hnew = Tdirection * MIN(MAX(ABS(Hnew), o->Hmin), o->Hmax);
      if (Reject == 1) Hnew = Tdirection * MIN(ABS(Hnew), ABS(H));
      Reject = 0;
      if ((T + Hnew / o->Qmin - Tend) * Tdirection >= ZERO) {
        H = Tend - T;
      } else {
        Hratio = Hnew / H;
        SkipLU = (Theta <= o->ThetaMin) && (Hratio >= o->Qmin) && (Hratio <= o->Qmax);
        if (SkipLU == 0) H = Hnew;
      }
      SkipJac = 0;
      /* passes without progress are bounded: an accepted step of
       * T + H == T does not reset the count */
      if (T != Tprev) Nrun = 0;
      else if (++Nrun > MAX_REJECT) { *Tp = T; return RK_IERR_REJECTS; }
    } else {
      if (FirstStep == 1 || Reject == 1) H = o->FacRej * H;
      else H = Hnew;
      Reject = 1;
      SkipJac = 1;
      SkipLU = 0;
      if (c->st.Nacc >= 1) c->st.Nrej++;
      if (Err != Err || !(ABS(H) <= DBL_MAX)) { *Tp = T; return RK_IERR_NONFINITE; }
      if (++Nrun > MAX_REJECT) { *Tp = T; return RK_IERR_REJECTS; }
    }
  }
  *Tp = T;
  return 1;
}

/* RungeKutta (:203-611): defaults and checks; the last failing check's code */
static int RungeKutta_cell(ros_cell* c, double T, double Tend, double* Y, const double* RelTol,
                           const double* AbsTol, const double* RCNTRL, const int* ICNTRL,
                           int* istatus, double* rstat) {
  rk_method m;
  rk_opts o;
  int IERR = 0;
  const double Roundoff = ros_ref_wlamch();
  memset(&c->st, 0, sizeof c->st);
  rstat[0] = rstat[1] = rstat[2] = ZERO;
  o.ITOL = (ICNTRL[1] == 0) ? 1 : 0;
  o.SdirkError = (ICNTRL[9] == 0) ? 0 : 1;
  const int code = ICNTRL[2];
  if (code == 5 || !rk_method_of(code, o.SdirkError, &m)) IERR = -13;
  if (ICNTRL[3] == 0) {
    o.Max_no_steps = 200000;
  } else {
    o.Max_no_steps = ICNTRL[3];
    if (o.Max_no_steps <= 0) IERR = -1;
  }
  if (ICNTRL[4] == 0) {
    o.NewtonMaxit = 8;
  } else {
    o.NewtonMaxit = ICNTRL[4];
    if (o.NewtonMaxit <= 0) IERR = -2;
  }
  o.StartNewton = (ICNTRL[5] == 0) ? 1 : 0;
  o.Gustafsson = (ICNTRL[10] == 0) ? 1 : 0;
  o.Hmin = (RCNTRL[0] == ZERO) ? ZERO : MIN(ABS(RCNTRL[0]), ABS(Tend - T));
  o.Hmax = (RCNTRL[1] == ZERO) ? ABS(Tend - T) : MIN(ABS(RCNTRL[1]), ABS(Tend - T));
  o.Hstart = (RCNTRL[2] == ZERO) ? ZERO : MIN(ABS(RCNTRL[2]), ABS(Tend - T));
  o.FacMin = (RCNTRL[3] == ZERO) ? 0.2 : RCNTRL[3];
  o.FacMax = (RCNTRL[4] == ZERO) ? 8.0 : RCNTRL[4];
  o.FacRej = (RCNTRL[5] == ZERO) ? 0.1 : RCNTRL[5];
  o.FacSafe = (RCNTRL[6] == ZERO) ? 0.9 : RCNTRL[6];
  if (o.FacMax < ONE || o.FacMin > ONE || o.FacSafe <= 1.0e-03 || o.FacSafe >= ONE) IERR = -4;
  if (RCNTRL[7] == ZERO) {
    o.ThetaMin = 1.0e-03;
  } else {
    o.ThetaMin = RCNTRL[7];
    if (o.ThetaMin <= 0.0 || o.ThetaMin >= 1.0) IERR = -5;
  }
  if (RCNTRL[8] == ZERO) {
    o.NewtonTol = 3.0e-02;
  } else {
    o.NewtonTol = RCNTRL[8];
    if (o.NewtonTol <= Roundoff) IERR = -6;
  }
  o.Qmin = (RCNTRL[9] == ZERO) ? ONE : RCNTRL[9];
  o.Qmax = (RCNTRL[10] == ZERO) ? 1.2 : RCNTRL[10];
  if (o.Qmin > ONE || o.Qmax < ONE) IERR = -7;
  if (o.ITOL == 0) {
    if (AbsTol[0] <= ZERO || RelTol[0] <= 10.0 * Roundoff) IERR = -8;
  } else {
    for (int i = 0; i < KPP_NVAR; i++)
      if (AbsTol[i] <= ZERO || RelTol[i] <= 10.0 * Roundoff) IERR = -8;
  }
  if (IERR >= 0) {
    double Tc = T;
    IERR = RK_Integrator(c, &Tc, Tend, Y, AbsTol, RelTol, &m, &o, Roundoff, rstat);
  }
  istatus[0] = c->st.Nfun; istatus[1] = c->st.Njac; istatus[2] = c->st.Nstp;
  istatus[3] = c->st.Nacc; istatus[4] = c->st.Nrej; istatus[5] = c->st.Ndec;
  istatus[6] = c->st.Nsol; istatus[7] = c->st.Nsng;
  return IERR;
}

/* method table for tests: kind, A[9], B[3], C[3], D[3], E[4], Bhat[4],
 * Bgam[5], Theta[3], {Gamma, Alpha, Beta, ELO}, T[9], TinvAinv[9]; returns
 * 1 for a built method, 0 otherwise */
int rk_ref_method(int code, int sdirk_error, int* kind, double* A, double* B, double* C,
                  double* D, double* E, double* Bhat, double* Bgam, double* Theta, double* gabe,
                  double* T, double* TinvAinv) {
  rk_method m;
  const int r = (code == 5) ? 0 : rk_method_of(code, sdirk_error, &m);
  if (!r) return 0;
  *kind = m.kind;
  for (int i = 0; i < 3; i++) {
    B[i] = m.B[i]; C[i] = m.C[i]; D[i] = m.D[i]; Theta[i] = m.Theta[i];
    for (int j = 0; j < 3; j++) {
      A[3 * i + j] = m.A[i][j];
      T[3 * i + j] = m.T[i][j];
      TinvAinv[3 * i + j] = m.TinvAinv[i][j];
    }
  }
  for (int i = 0; i < 4; i++) { E[i] = m.E[i]; Bhat[i] = m.Bhat[i]; }
  for (int i = 0; i < 5; i++) Bgam[i] = m.Bgam[i];
  gabe[0] = m.Gamma; gabe[1] = m.Alpha; gabe[2] = m.Beta; gabe[3] = m.ELO;
  return 1;
}

/* the complex LU and solve on the mechanism's pattern, for tests */
int rk_ref_decomp_cmplx(double* JVSR, double* JVSI) { return KppDecompCmplxR(JVSR, JVSI); }
void rk_ref_solve_cmplx(const double* JVSR, const double* JVSI, double* XR, double* XI) {
  KppSolveCmplxR(JVSR, JVSI, XR, XI);
}

/* Batched RungeKutta (int/runge_kutta.c INTEGRATE/RungeKutta) over SoA
 * concentrations: cell c over [tin[c], tout[c]] with ICNTRL/RCNTRL as
 * RungeKutta takes them (0-based, no merge), sun_mode as ros_ref_batch,
 * rconst or NULL.  istatus [8][n] = ISTATUS[0..7], rstatus [3][n] =
 * RSTATUS[0..2] (Texit, Hacc, Hnew). */
void rk_ref_batch(int ncells, double* C, const double* temp, const double* cfactor,
                  const double* sun, int sun_mode, const double* tin, const double* tout,
                  const double* atol, const double* rtol, const int* ICNTRL, const double* RCNTRL,
                  const double* rconst, int* ierr, int* istatus, double* rstatus, int nthreads) {
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
    int e = RungeKutta_cell(&c, tin[cidx], tout[cidx], Y, rtol, atol, RCNTRL, ICNTRL, ist, rst);
    for (int i = 0; i < KPP_NVAR; i++) C[(size_t)i * ncells + cidx] = Y[i];
    ierr[cidx] = e;
    if (istatus)
      for (int i = 0; i < 8; i++) istatus[(size_t)i * ncells + cidx] = ist[i];
    if (rstatus)
      for (int i = 0; i < 3; i++) rstatus[(size_t)i * ncells + cidx] = rst[i];
  }
}
