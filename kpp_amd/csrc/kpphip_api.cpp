// C ABI of the batched Rosenbrock solver (declarations in include/kpphip.h).
#include "kpphip.h"

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <dlfcn.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "mech_ops.h"

#define KPPHIP_MECH(N)                                        \
  extern "C" const kpphip::MechOps* kpphip_mech_##N();        \
  extern "C" const kpphip::MixedOps* kpphip_mech_mixed_##N(); \
  extern "C" kpphip::ComponentFn kpphip_mech_comp_##N();      \
  extern "C" const kpphip::SdirkOps* kpphip_mech_sdirk_##N(); \
  extern "C" const kpphip::SdirkOps* kpphip_mech_rk_##N();   \
  extern "C" const kpphip::SdirkOps* kpphip_mech_adj_##N();
#include "kpphip_registry.inc"
#undef KPPHIP_MECH

namespace {

using kpphip::MechOps;
using kpphip::RosCoef;
using kpphip::RosParams;

const MechOps* (*const g_mechs[])() = {
#define KPPHIP_MECH(N) kpphip_mech_##N,
#include "kpphip_registry.inc"
#undef KPPHIP_MECH
};
constexpr int g_nmechs = (int)(sizeof(g_mechs) / sizeof(g_mechs[0]));
const kpphip::MixedOps* (*const g_mixed[])() = {
#define KPPHIP_MECH(N) kpphip_mech_mixed_##N,
#include "kpphip_registry.inc"
#undef KPPHIP_MECH
};

kpphip::ComponentFn (*const g_comp[])() = {
#define KPPHIP_MECH(N) kpphip_mech_comp_##N,
#include "kpphip_registry.inc"
#undef KPPHIP_MECH
};

const kpphip::SdirkOps* (*const g_sdirk[])() = {
#define KPPHIP_MECH(N) kpphip_mech_sdirk_##N,
#include "kpphip_registry.inc"
#undef KPPHIP_MECH
};

const kpphip::SdirkOps* sdirk_of(const MechOps* m) {
  for (int i = 0; i < g_nmechs; i++)
    if (g_mechs[i]() == m) return g_sdirk[i]();
  return nullptr;
}

const kpphip::SdirkOps* (*const g_rk[])() = {
#define KPPHIP_MECH(N) kpphip_mech_rk_##N,
#include "kpphip_registry.inc"
#undef KPPHIP_MECH
};

const kpphip::SdirkOps* rk_of(const MechOps* m) {
  for (int i = 0; i < g_nmechs; i++)
    if (g_mechs[i]() == m) return g_rk[i]();
  return nullptr;
}

const kpphip::SdirkOps* (*const g_adj[])() = {
#define KPPHIP_MECH(N) kpphip_mech_adj_##N,
#include "kpphip_registry.inc"
#undef KPPHIP_MECH
};

const kpphip::SdirkOps* adj_of(const MechOps* m) {
  for (int i = 0; i < g_nmechs; i++)
    if (g_mechs[i]() == m) return g_adj[i]();
  return nullptr;
}

const kpphip::MixedOps* mixed_of(const MechOps* m) {
  for (int i = 0; i < g_nmechs; i++)
    if (g_mechs[i]() == m) return g_mixed[i]();
  return nullptr;
}

kpphip::ComponentFn component_of(const MechOps* m) {
  for (int i = 0; i < g_nmechs; i++)
    if (g_mechs[i]() == m) return g_comp[i]();
  return nullptr;
}

const MechOps* find_mech(const char* name) {
  if (!name) return nullptr;
  for (int i = 0; i < g_nmechs; i++) {
    const MechOps* m = g_mechs[i]();
    if (strcmp(m->name, name) == 0) return m;
  }
  return nullptr;
}

// Unit roundoff by the halving loop of util/blas.c:150-180 (WLAMCH('E')).
double unit_roundoff() {
  volatile double eps = pow(0.5, 16), s;
  int i;
  for (i = 17; i <= 80; i++) {
    eps = eps * 0.5;
    s = 1.0 + eps;
    if (s <= 1.0) break;
  }
  if (i == 80) return 2.2e-16;
  return eps * 2.0;
}

// Method tables of int/rosenbrock.c:779-1222 (numbers as printed there).
bool fill_method(int code, RosCoef& m) {
  memset(&m, 0, sizeof m);
  switch (code) {
    case 1: {  // Ros2 :779-830
      const double g = 1.70710678118655;
      m.S = 2;
      m.A[0] = 1.0 / g;
      m.C[0] = (-2.0) / g;
      m.NewF[0] = 1; m.NewF[1] = 1;
      m.M[0] = (3.0) / (2.0 * g); m.M[1] = (1.0) / (2.0 * g);
      m.E[0] = 1.0 / (2.0 * g); m.E[1] = 1.0 / (2.0 * g);
      m.ELO = 2.0;
      m.Alpha[0] = 0.0; m.Alpha[1] = 1.0;
      m.Gamma[0] = g; m.Gamma[1] = -g;
      return true;
    }
    case 2:  // Ros3 :834-892
      m.S = 3;
      m.A[0] = 1.0; m.A[1] = 1.0; m.A[2] = 0.0;
      m.C[0] = -1.0156171083877702091975600115545;
      m.C[1] = 4.0759956452537699824805835358067;
      m.C[2] = 9.2076794298330791242156818474003;
      m.NewF[0] = 1; m.NewF[1] = 1; m.NewF[2] = 0;
      m.M[0] = 1.0; m.M[1] = 6.1697947043828245592553615689730;
      m.M[2] = -0.4277225654321857332623837380651;
      m.E[0] = 0.5; m.E[1] = -2.9079558716805469821718236208017;
      m.E[2] = 0.2235406989781156962736090927619;
      m.ELO = 3.0;
      m.Alpha[0] = 0.0; m.Alpha[1] = 0.43586652150845899941601945119356;
      m.Alpha[2] = 0.43586652150845899941601945119356;
      m.Gamma[0] = 0.43586652150845899941601945119356;
      m.Gamma[1] = 0.24291996454816804366592249683314;
      m.Gamma[2] = 2.1851380027664058511513169485832;
      return true;
    case 3:  // Ros4 :898-973
      m.S = 4;
      m.A[0] = 0.2000000000000000e+01; m.A[1] = 0.1867943637803922e+01;
      m.A[2] = 0.2344449711399156; m.A[3] = m.A[1]; m.A[4] = m.A[2]; m.A[5] = 0.0;
      m.C[0] = -0.7137615036412310e+01; m.C[1] = 0.2580708087951457e+01;
      m.C[2] = 0.6515950076447975; m.C[3] = -0.2137148994382534e+01;
      m.C[4] = -0.3214669691237626; m.C[5] = -0.6949742501781779;
      m.NewF[0] = 1; m.NewF[1] = 1; m.NewF[2] = 1; m.NewF[3] = 0;
      m.M[0] = 0.2255570073418735e+01; m.M[1] = 0.2870493262186792;
      m.M[2] = 0.4353179431840180; m.M[3] = 0.1093502252409163e+01;
      m.E[0] = -0.2815431932141155; m.E[1] = -0.7276199124938920e-01;
      m.E[2] = -0.1082196201495311; m.E[3] = -0.1093502252409163e+01;
      m.ELO = 4.0;
      m.Alpha[0] = 0.0; m.Alpha[1] = 0.1145640000000000e+01;
      m.Alpha[2] = 0.6552168638155900; m.Alpha[3] = m.Alpha[2];
      m.Gamma[0] = 0.5728200000000000; m.Gamma[1] = -0.1769193891319233e+01;
      m.Gamma[2] = 0.7592633437920482; m.Gamma[3] = -0.1049021087100450;
      return true;
    case 4:  // Rodas3 :976-1045
      m.S = 4;
      m.A[0] = 0.0; m.A[1] = 2.0; m.A[2] = 0.0; m.A[3] = 2.0; m.A[4] = 0.0; m.A[5] = 1.0;
      m.C[0] = 4.0; m.C[1] = 1.0; m.C[2] = -1.0; m.C[3] = 1.0; m.C[4] = -1.0;
      m.C[5] = -2.66666666666667;
      m.NewF[0] = 1; m.NewF[1] = 0; m.NewF[2] = 1; m.NewF[3] = 1;
      m.M[0] = 2.0; m.M[1] = 0.0; m.M[2] = 1.0; m.M[3] = 1.0;
      m.E[3] = 1.0;
      m.ELO = 3.0;
      m.Alpha[2] = 1.0; m.Alpha[3] = 1.0;
      m.Gamma[0] = 0.5; m.Gamma[1] = 1.5;
      return true;
    case 5:  // Rodas4 :1121-1222
      m.S = 6;
      m.Alpha[0] = 0.000; m.Alpha[1] = 0.386; m.Alpha[2] = 0.210;
      m.Alpha[3] = 0.630; m.Alpha[4] = 1.000; m.Alpha[5] = 1.000;
      m.Gamma[0] = 0.2500000000000000; m.Gamma[1] = -0.1043000000000000;
      m.Gamma[2] = 0.1035000000000000; m.Gamma[3] = -0.3620000000000023e-01;
      m.A[0] = 0.1544000000000000e+01; m.A[1] = 0.9466785280815826;
      m.A[2] = 0.2557011698983284; m.A[3] = 0.3314825187068521e+01;
      m.A[4] = 0.2896124015972201e+01; m.A[5] = 0.9986419139977817;
      m.A[6] = 0.1221224509226641e+01; m.A[7] = 0.6019134481288629e+01;
      m.A[8] = 0.1253708332932087e+02; m.A[9] = -0.6878860361058950;
      m.A[10] = m.A[6]; m.A[11] = m.A[7]; m.A[12] = m.A[8]; m.A[13] = m.A[9]; m.A[14] = 1.0;
      m.C[0] = -0.5668800000000000e+01; m.C[1] = -0.2430093356833875e+01;
      m.C[2] = -0.2063599157091915; m.C[3] = -0.1073529058151375;
      m.C[4] = -0.9594562251023355e+01; m.C[5] = -0.2047028614809616e+02;
      m.C[6] = 0.7496443313967647e+01; m.C[7] = -0.1024680431464352e+02;
      m.C[8] = -0.3399990352819905e+02; m.C[9] = 0.1170890893206160e+02;
      m.C[10] = 0.8083246795921522e+01; m.C[11] = -0.7981132988064893e+01;
      m.C[12] = -0.3152159432874371e+02; m.C[13] = 0.1631930543123136e+02;
      m.C[14] = -0.6058818238834054e+01;
      m.M[0] = m.A[6]; m.M[1] = m.A[7]; m.M[2] = m.A[8]; m.M[3] = m.A[9];
      m.M[4] = 1.0; m.M[5] = 1.0;
      m.E[5] = 1.0;
      for (int i = 0; i < 6; i++) m.NewF[i] = 1;
      m.ELO = 4.0;
      return true;
  }
  return false;
}

// The ICNTRL(3) map of int/rosenbrock.f90:321-340 and its tables where they
// differ from the C ones: Ros2's g = 1 + 1/SQRT(2) (:921), Rodas3's
// C(6) = -(8/3) (:1122), and Rang3 (:1327-1394; Rang & Angermann, BIT 45
// (2005) 761-787, Tables 4.1-4.2), which the C integrator does not have.
bool fill_method_f90(int code, RosCoef& m) {
  switch (code) {
    case 0: case 4:
      fill_method(4, m);
      m.C[5] = -(8.0 / 3.0);
      return true;
    case 1: {
      fill_method(1, m);
      const double g = 1.0 + 1.0 / sqrt(2.0);
      m.A[0] = 1.0 / g;
      m.C[0] = (-2.0) / g;
      m.M[0] = (3.0) / (2.0 * g); m.M[1] = (1.0) / (2.0 * g);
      m.E[0] = 1.0 / (2.0 * g); m.E[1] = 1.0 / (2.0 * g);
      m.Gamma[0] = g; m.Gamma[1] = -g;
      return true;
    }
    case 2: case 3: case 5:
      return fill_method(code, m);
    case 7:  // Rodas3.1 (:1157-1221; the C integrator cannot select it)
      memset(&m, 0, sizeof m);
      m.S = 4;
      m.A[0] = 0.000000000000000; m.A[1] = 0.646601929740551; m.A[2] = 0.409567801987914;
      m.A[3] = 0.646601929740551; m.A[4] = 0.409567801987914; m.A[5] = 1.000000000000000;
      m.C[0] = 4.198495621784201; m.C[1] = 3.711590161613010; m.C[2] = -1.787771994729384;
      m.C[3] = 4.458898153216104; m.C[4] = -2.024095448516552; m.C[5] = -2.626700600119396;
      m.NewF[0] = 1; m.NewF[1] = 0; m.NewF[2] = 1; m.NewF[3] = 1;
      m.M[0] = 0.646601929740551; m.M[1] = 0.409567801987914; m.M[2] = 1.0; m.M[3] = 1.0;
      m.E[3] = 1.0;
      m.ELO = 3.0;
      m.Alpha[2] = 1.0; m.Alpha[3] = 1.0;
      m.Gamma[0] = 0.515; m.Gamma[1] = 1.628546001287715;
      return true;
    case 6:
      memset(&m, 0, sizeof m);
      m.S = 4;
      m.A[0] = 5.09052051067020; m.A[1] = 5.09052051067020; m.A[2] = 0.0;
      m.A[3] = 4.97628111010787; m.A[4] = 2.77268164715849e-02; m.A[5] = 2.29428036027904e-01;
      m.C[0] = -1.16790812312283e+01; m.C[1] = -1.64057326467367e+01;
      m.C[2] = -2.77268164715850e-01; m.C[3] = -8.38103960500476;
      m.C[4] = -8.48328409199343e-01; m.C[5] = 2.87009860433106e-01;
      m.M[0] = 5.22582761233094; m.M[1] = -5.56971148154165e-01;
      m.M[2] = 3.57979469353645e-01; m.M[3] = 1.72337398521064;
      m.E[0] = -5.16845212784040; m.E[1] = -1.26351942603842;
      m.E[2] = -1.11022302462516e-16; m.E[3] = 2.22044604925031e-16;
      m.Alpha[1] = 2.21878746765329; m.Alpha[2] = 2.21878746765329;
      m.Alpha[3] = 1.55392337535788;
      m.Gamma[0] = 4.35866521508459e-01; m.Gamma[1] = -1.78292094614483;
      m.Gamma[2] = -2.46541900496934; m.Gamma[3] = -8.05529997906370e-01;
      for (int i = 0; i < 4; i++) m.NewF[i] = 1;
      m.ELO = 3.0;
      return true;
  }
  return false;
}

// SDIRK method tables of int/sdirk.c:988-1306 (the coefficients the
// integrator uses: Gamma, C, D, E, ELO, Theta, Alpha); ICNTRL[2] 0, 1 and
// any unknown code select Sdirk2a (the switch's default, :316-338)
void fill_sdirk(int code, kpphip::SdCoef& m) {
  memset(&m, 0, sizeof m);
  switch (code) {
    case 2:  // Sdirk2b
      m.S = 2;
      m.Gamma = 1.707106781186547524400844362104849;
      m.C[0] = 1.707106781186547524400844362104849; m.C[1] = 1.0;
      m.D[1] = 1.0;
      m.E[0] = -0.4714045207910316829338962414032326;
      m.E[1] = 0.8047378541243650162672295747365659;
      m.ELO = 2;
      m.Theta[0][1] = -0.414213562373095048801688724209698;
      m.Alpha[0][1] = 0.5857864376269049511983112757903019;
      return;
    case 3:  // Sdirk3a
      m.S = 3;
      m.Gamma = 0.2113248654051871177454256097490213;
      m.C[0] = 0.2113248654051871177454256097490213;
      m.C[1] = 0.4226497308103742354908512194980427; m.C[2] = 1.0;
      m.D[2] = 1.0;
      m.E[0] = 0.9106836025229590978424821138352906;
      m.E[1] = -1.244016935856292431175815447168624;
      m.E[2] = 0.3333333333333333333333333333333333;
      m.ELO = 2;
      m.Theta[0][1] = 1.0;
      m.Theta[0][2] = -1.732050807568877293527446341505872;
      m.Theta[1][2] = 2.732050807568877293527446341505872;
      m.Alpha[0][1] = 2.0;
      m.Alpha[0][2] = -12.92820323027550917410978536602349;
      m.Alpha[1][2] = 8.83012701892219323381861585376468;
      return;
    case 4:  // Sdirk4a
      m.S = 5;
      m.Gamma = 0.2666666666666666666666666666666667;
      m.C[0] = 0.2666666666666666666666666666666667;
      m.C[1] = 0.7666666666666666666666666666666667;
      m.C[2] = 0.5666666666666666666666666666666667;
      m.C[3] = 0.3661315380631796996374935266701191; m.C[4] = 1.0;
      m.D[4] = 1.0;
      m.E[0] = -0.6804000050475287124787034884002302;
      m.E[1] = 1.558961944525217193393931795738823;
      m.E[2] = -13.55893003128907927748632408763868;
      m.E[3] = 15.48522576958521253098585004571302;
      m.E[4] = 1.0;
      m.ELO = 4.0;
      m.Theta[0][1] = 1.875000000000000000000000000000000;
      m.Theta[0][2] = 1.708847304091539528432732316227462;
      m.Theta[1][2] = -0.2030773231622746185852981969486824;
      m.Theta[0][3] = 0.2680325578937783958847157206823118;
      m.Theta[1][3] = -0.1828840955527181631794050728644549;
      m.Theta[2][3] = 0.2968246986151577397160821594427966;
      m.Theta[0][4] = 0.9096171815241460655379433581446771;
      m.Theta[1][4] = -3.108254967778352416114774430509465;
      m.Theta[2][4] = 12.33727431701306195581826123274001;
      m.Theta[3][4] = -11.24557012450885560524143016037523;
      m.Alpha[0][1] = 2.875000000000000000000000000000000;
      m.Alpha[0][2] = 0.8500000000000000000000000000000000;
      m.Alpha[1][2] = 0.4434782608695652173913043478260870;
      m.Alpha[0][3] = 0.7352046091658870564637910527807370;
      m.Alpha[1][3] = -0.9525565003057343527941920657462074e-01;
      m.Alpha[2][3] = 0.4290111305453813852259481840631738;
      m.Alpha[0][4] = -16.10898993405067684831655675112808;
      m.Alpha[1][4] = 6.559571569643355712998131800797873;
      m.Alpha[2][4] = -15.90772144271326504260996815012482;
      m.Alpha[3][4] = 25.34908987169226073668861694892683;
      return;
    case 5:  // Sdirk4b
      m.S = 5;
      m.Gamma = 0.25;
      m.C[0] = 0.25; m.C[1] = 0.75; m.C[2] = 0.55; m.C[3] = 0.5; m.C[4] = 1.0;
      m.D[4] = 1.0;
      m.E[0] = 0.5750; m.E[1] = 0.2125; m.E[2] = -4.6875; m.E[3] = 4.2500; m.E[4] = 0.1500;
      m.ELO = 4;
      m.Theta[0][1] = 2.0;
      m.Theta[0][2] = 1.680000000000000000000000000000000;
      m.Theta[1][2] = -0.1600000000000000000000000000000000;
      m.Theta[0][3] = 1.308823529411764705882352941176471;
      m.Theta[1][3] = -0.1838235294117647058823529411764706;
      m.Theta[2][3] = 0.1102941176470588235294117647058824;
      m.Theta[0][4] = -3.083333333333333333333333333333333;
      m.Theta[1][4] = -4.291666666666666666666666666666667;
      m.Theta[2][4] = 34.37500000000000000000000000000000;
      m.Theta[3][4] = -28.3333333333333333333333333333;
      m.Alpha[0][1] = 3.0;
      m.Alpha[0][2] = 0.8800000000000000000000000000000000;
      m.Alpha[1][2] = 0.4400000000000000000000000000000000;
      m.Alpha[0][3] = 0.1666666666666666666666666666666667;
      m.Alpha[1][3] = -0.8333333333333333333333333333333333e-01;
      m.Alpha[2][3] = 0.9469696969696969696969696969696970;
      m.Alpha[0][4] = -6.0;
      m.Alpha[1][4] = 9.0;
      m.Alpha[2][4] = -56.81818181818181818181818181818182;
      m.Alpha[3][4] = 54.0;
      return;
    default:  // Sdirk2a
      m.S = 2;
      m.Gamma = 0.2928932188134524755991556378951510;
      m.C[0] = 0.292893218813452475599155637895151; m.C[1] = 1.0;
      m.D[1] = 1.0;
      m.E[0] = 0.4714045207910316829338962414032326;
      m.E[1] = -0.1380711874576983496005629080698993;
      m.ELO = 2;
      m.Theta[0][1] = 2.414213562373095048801688724209698;
      m.Alpha[0][1] = 3.414213562373095048801688724209698;
      return;
  }
}

// SDIRK's controls (int/sdirk.c:305-503): defaults, and the checks' codes,
// which the reference prints and then ignores (returned here instead)
int sdirk_controls(const int* ICNTRL, const double* RCNTRL, const double* atol,
                   const double* rtol, int nvar, kpphip::SdCoef& m) {
  fill_sdirk(ICNTRL[2], m);
  m.ITOL = (ICNTRL[1] == 0) ? 1 : 0;
  if (ICNTRL[3] < 0) return -1;
  m.Max_no_steps = (ICNTRL[3] == 0) ? 200000 : ICNTRL[3];
  if (ICNTRL[4] < 0) return -2;
  m.NewtonMaxit = (ICNTRL[4] == 0) ? 8 : ICNTRL[4];
  m.StartNewton = (ICNTRL[5] == 0) ? 1 : 0;
  const double roundoff = unit_roundoff();
  for (int i = 0; i <= 2; i++)
    if (RCNTRL[i] < 0.0) return -3;
  for (int i = 3; i <= 6; i++)
    if (RCNTRL[i] < 0.0) return -4;
  m.Hmin = RCNTRL[0];
  m.Hmax_user = RCNTRL[1];
  m.Hstart_user = RCNTRL[2];
  m.FacMin = (RCNTRL[3] == 0.0) ? 0.2 : RCNTRL[3];
  m.FacMax = (RCNTRL[4] == 0.0) ? 10.0 : RCNTRL[4];
  m.FacRej = (RCNTRL[5] == 0.0) ? 0.1 : RCNTRL[5];
  m.FacSafe = (RCNTRL[6] == 0.0) ? 0.9 : RCNTRL[6];
  m.ThetaMin = (RCNTRL[7] == 0.0) ? 1.0e-03 : RCNTRL[7];
  m.NewtonTol = (RCNTRL[8] == 0.0) ? 3.0e-02 : RCNTRL[8];
  m.Qmin = (RCNTRL[9] == 0.0) ? 1.0 : RCNTRL[9];
  m.Qmax = (RCNTRL[10] == 0.0) ? 1.2 : RCNTRL[10];
  if (m.ITOL == 0) {
    // :492-495 as written: the bool (AbsTol <= 0 || RelTol != 0) against 10*Roundoff
    if ((double)((atol[0] <= 0.0) || (rtol[0] != 0.0)) <= 10.0 * roundoff) return -5;
  } else {
    for (int i = 0; i < nvar; i++)
      if ((atol[i] <= 0.0) || (rtol[i] <= 10.0 * roundoff)) return -5;
  }
  return 0;
}

// The Runge-Kutta tables of int/runge_kutta.c:1369-1863 as the integrator
// uses them (A, C, D, E, Bgam, Theta, the eigenvalues Gamma, Alpha +- i Beta
// of A^-1 and the transformation T, T^-1 A^-1); b0 of the error estimators
// by the estimator choice, the b0 expressions evaluated as the reference
// does at run time.  False for a code the library does not build (5,
// Lobatto-3A, which the reference lists as not yet implemented).
bool fill_rk(int code, int sdirk_error, kpphip::RkCoef& m) {
  auto set3 = [](double (*d)[3], std::initializer_list<double> v) {
    int k = 0;
    for (double x : v) d[k / 3][k % 3] = x, k++;
  };
  double b0;
  for (int i = 0; i < 3; i++) m.D[i] = 0.0;
  switch (code) {
    case 0:
    case 1:  // Radau-2A
      b0 = sdirk_error == 1 ? 0.2e-01 : 0.5e-01;
      m.kind = 1;
      set3(m.A, {1.968154772236604258683861429918299e-01, -6.55354258501983881085227825696087e-02,
                 2.377097434822015242040823210718965e-02, 3.944243147390872769974116714584975e-01,
                 2.920734116652284630205027458970589e-01, -4.154875212599793019818600988496743e-02,
                 3.764030627004672750500754423692808e-01, 5.124858261884216138388134465196080e-01,
                 1.111111111111111111111111111111111e-01});
      set3(m.T, {9.443876248897524148749007950641664e-02, -1.412552950209542084279903838077973e-01,
                 -3.00291941051474244918611170890539e-02, 2.502131229653333113765090675125018e-01,
                 2.041293522937999319959908102983381e-01, 3.829421127572619377954382335998733e-01,
                 1.0, 1.0, 0.0});
      set3(m.TinvAinv,
           {1.520148562492775501049204957366528e+01, 1.192055789400527921212348994770778,
            1.903956760517560343018332287285119, -9.669512977505946748632625374449567,
            -8.724028436822336183071773193986487, 3.096043239482439656981667712714881,
            -1.409513259499574544876303981551774e+01, 5.895975725255405108079130152868952,
            -1.441236197545344702389881889085515e-01});
      m.C[0] = 1.550510257216821901802715925294109e-01;
      m.C[1] = 6.449489742783178098197284074705891e-01;
      m.C[2] = 1.0;
      m.D[2] = 1.0;
      m.E[0] = 1.0 * b0;
      m.E[1] = -10.04880939982741556246032950764708 * b0;
      m.E[2] = 1.382142733160748895793662840980412 * b0;
      m.E[3] = -.3333333333333333333333333333333333 * b0;
      m.Bgam[0] = b0;
      m.Bgam[1] = .3764030627004672750500754423692807 - 1.558078204724922382431975370686279 * b0;
      m.Bgam[2] = .8914115380582557157653087040196118 * b0 + .5124858261884216138388134465196077;
      m.Bgam[3] = -.1637777184845662566367174924883037 - .3333333333333333333333333333333333 * b0;
      m.Bgam[4] = .2748888295956773677478286035994148;
      m.Theta[0] = -1.520677486405081647234271944611547 - 10.04880939982741556246032950764708 * b0;
      m.Theta[1] = 2.070455145596436382729929151810376 + 1.382142733160748895793662840980413 * b0;
      m.Theta[2] = -.3333333333333333333333333333333333 * b0 - .3744441479783868387391430179970741;
      m.ELO = (b0 == 0.0) ? 6.0 : 4.0;
      m.Gamma = 3.637834252744495732208418513577775;
      m.Alpha = 2.681082873627752133895790743211112;
      m.Beta = 3.050430199247410569426377624787569;
      return true;
    case 2:  // Lobatto-3C
      b0 = sdirk_error == 1 ? 0.2 : 0.5;
      m.kind = 3;
      set3(m.A, {.1666666666666666666666666666666667, -.3333333333333333333333333333333333,
                 .1666666666666666666666666666666667, .1666666666666666666666666666666667,
                 .4166666666666666666666666666666667, -.8333333333333333333333333333333333e-01,
                 .1666666666666666666666666666666667, .6666666666666666666666666666666667,
                 .1666666666666666666666666666666667});
      set3(m.T, {1.0, 1.0, 0.0, .4554100411010284672111720348287483,
                 -.6027050205505142336055860174143743, -.4309321229203225731070721341350346,
                 2.195823345445647152832799205549709, -1.097911672722823576416399602774855,
                 .7850032632435902184104551358922130});
      set3(m.TinvAinv,
           {1.104302803159744452668648155627548, .916122120694355522658740710823143,
            .5029105849749601702795812241441172, 1.895697196840255547331351844372453,
            3.083877879305644477341259289176857, -1.502910584974960170279581224144117,
            .8362439183082935036129145574774502, -3.344975673233174014451658229909802,
            .312908409479233358005944466882642});
      m.C[0] = 0.0;
      m.C[1] = 0.5;
      m.C[2] = 1.0;
      m.D[2] = 1.0;
      m.E[0] = .3808338772072650364017425226487022 * b0;
      m.E[1] = -1.142501631621795109205227567946107 * b0;
      m.E[2] = -1.523335508829060145606970090594809 * b0;
      m.E[3] = .3808338772072650364017425226487022 * b0;
      m.Bgam[0] = b0;
      m.Bgam[1] = .1666666666666666666666666666666667 - 1.0 * b0;
      m.Bgam[2] = .6666666666666666666666666666666667;
      m.Bgam[3] = -.2141672105405983697350758559820354;
      m.Bgam[4] = .3808338772072650364017425226487021;
      m.Theta[0] = -3.0 * b0 - .3808338772072650364017425226487021;
      m.Theta[1] = -4.0 * b0 + 1.523335508829060145606970090594808;
      m.Theta[2] = -.142501631621795109205227567946106 + b0;
      m.ELO = (b0 == 0.0) ? 5.0 : 4.0;
      m.Gamma = 2.625816818958466716011888933765284;
      m.Alpha = 1.687091590520766641994055533117359;
      m.Beta = 2.508731754924880510838743672432351;
      return true;
    case 3:  // Gauss
      b0 = 0.1;
      m.kind = 4;
      set3(m.A, {.1388888888888888888888888888888889, -.359766675249389034563954710966045e-01,
                 .97894440153083260495800422294756e-02, .3002631949808645924380249472131556,
                 .2222222222222222222222222222222222, -.224854172030868146602471694353778e-01,
                 .2679883337624694517281977355483022, .4804211119693833479008399155410489,
                 .1388888888888888888888888888888889});
      set3(m.T, {.7215185205520017032081769924397664e-01, -.8224123057363067064866206597516454e-01,
                 -.6012073861930850173085948921439054e-01, .1188325787412778070708888193730294,
                 .5306509074206139504614411373957448e-01, .3162050511322915732224862926182701,
                 1.0, 1.0, 0.0});
      set3(m.TinvAinv,
           {27.82766708436744962047620566703329, 5.290933503982655311815946575100597,
            2.007817718512643701322151051660114, -17.66368928942422710690385180065675,
            -14.45491129892587782538830044147713, 2.992182281487356298677848948339886,
            -25.60678350282974256072419392007303, 6.762434375611708328910623303779923,
            1.043979339483109825041215970036771});
      m.C[0] = .1127016653792583114820734600217600;
      m.C[1] = .5000000000000000000000000000000000;
      m.C[2] = .8872983346207416885179265399782400;
      m.D[0] = .1666666666666666666666666666666667e+01;
      m.D[1] = -.1333333333333333333333333333333333e+01;
      m.D[2] = .1666666666666666666666666666666667e+01;
      m.E[0] = .2153144231161121782447335303806954 * b0;
      m.E[1] = -2.825278112319014084275808340593191 * b0;
      m.E[2] = .2870858974881495709929780405075939 * b0;
      m.E[3] = -.4558086256248162565397206448274867e-01 * b0;
      m.Bgam[0] = 0.0;
      m.Bgam[1] = .2373339543355109188382583162660537;
      m.Bgam[2] = .5879873931885192299409334646982414;
      m.Bgam[3] = -.4063577064014232702392531134499046e-01;
      m.Bgam[4] = .2153144231161121782447335303806955;
      m.Theta[0] = -2.594040933093095272574031876464493;
      m.Theta[1] = 1.824611539036311947589425112250199;
      m.Theta[2] = .1856563166634371860478043996459493;
      m.ELO = 4.0;
      m.Gamma = 4.644370709252171185822941421408064;
      m.Alpha = 3.677814645373914407088529289295970;
      m.Beta = 3.508761919567443321903661209182446;
      return true;
    case 4:  // Radau-1A
      b0 = 0.1;
      m.kind = 2;
      set3(m.A, {.1111111111111111111111111111111111, -.1916383190435098943442935597058829,
                 .8052720793239878323318244859477174e-01, .1111111111111111111111111111111111,
                 .2920734116652284630205027458970589, -.481334970546573839513422644787591e-01,
                 .1111111111111111111111111111111111, .5370223859435462728402311533676479,
                 .1968154772236604258683861429918299});
      set3(m.T, {.424293819848497965354371036408369, -.3235571519651980681202894497035503,
                 -.522137786846287839586599927945048, .57594609499806128896291585429339e-01,
                 .3148663231849760131614374283783e-02, .452429247674359778577728510381731,
                 1.0, 1.0, 0.0});
      set3(m.TinvAinv,
           {4.487354449794728738538663081025420, 5.178748573958397475446442544234494,
            1.435609490412123627047824222335563, -2.854361287939276673073807031221493,
            -1.003648660720543859000994063139137e+01, 1.789135380979465422050817815017383,
            -4.160768067752685525282947313530352, 1.124128569859216916690209918405860,
            1.700644430961823796581896350418417});
      m.C[0] = 0.0;
      m.C[1] = .3550510257216821901802715925294109;
      m.C[2] = .8449489742783178098197284074705891;
      m.D[0] = .3333333333333333333333333333333333;
      m.D[1] = -.8914115380582557157653087040196127;
      m.D[2] = 1.558078204724922382431975370686279;
      m.E[0] = .2748888295956773677478286035994148 * b0;
      m.E[1] = -1.374444147978386838739143017997074 * b0;
      m.E[2] = -1.335337922441686804550326197041126 * b0;
      m.E[3] = .235782604058977333559011782643466 * b0;
      m.Bgam[0] = 0.0;
      m.Bgam[1] = .1948150124588532186183490991130616e-01;
      m.Bgam[2] = .7575249005733381398986810981093584;
      m.Bgam[3] = -.518952314149008295083446116200793e-01;
      m.Bgam[4] = .2748888295956773677478286035994148;
      m.Theta[0] = -1.224370034375505083904362087063351;
      m.Theta[1] = .9340045331532641409047527962010133;
      m.Theta[2] = .4656990124352088397561234800640929;
      m.ELO = 4.0;
      m.Gamma = 3.637834252744495732208418513577775;
      m.Alpha = 2.681082873627752133895790743211112;
      m.Beta = 3.050430199247410569426377624787569;
      return true;
    default:
      return false;
  }
}

// RungeKutta's controls (int/runge_kutta.c:392-605): defaults, and every
// check in the reference's order, the last failing one's code returned (the
// reference records each in IERR and returns at the end of the checks)
int rk_controls(const int* ICNTRL, const double* RCNTRL, const double* atol, const double* rtol,
                int nvar, kpphip::RkCoef& m) {
  int ierr = 0;
  m.ITOL = (ICNTRL[1] == 0) ? 1 : 0;
  m.SdirkError = (ICNTRL[9] == 0) ? 0 : 1;
  if (!fill_rk(ICNTRL[2], m.SdirkError, m)) ierr = -13;
  m.Max_no_steps = (ICNTRL[3] == 0) ? 200000 : ICNTRL[3];
  if (m.Max_no_steps <= 0) ierr = -1;
  m.NewtonMaxit = (ICNTRL[4] == 0) ? 8 : ICNTRL[4];
  if (m.NewtonMaxit <= 0) ierr = -2;
  m.StartNewton = (ICNTRL[5] == 0) ? 1 : 0;
  m.Gustafsson = (ICNTRL[10] == 0) ? 1 : 0;
  const double roundoff = unit_roundoff();
  m.Hmin_user = RCNTRL[0];
  m.Hmax_user = RCNTRL[1];
  m.Hstart_user = RCNTRL[2];
  m.FacMin = (RCNTRL[3] == 0.0) ? 0.2 : RCNTRL[3];
  m.FacMax = (RCNTRL[4] == 0.0) ? 8.0 : RCNTRL[4];
  m.FacRej = (RCNTRL[5] == 0.0) ? 0.1 : RCNTRL[5];
  m.FacSafe = (RCNTRL[6] == 0.0) ? 0.9 : RCNTRL[6];
  if (m.FacMax < 1.0 || m.FacMin > 1.0 || m.FacSafe <= 1.0e-03 || m.FacSafe >= 1.0) ierr = -4;
  m.ThetaMin = (RCNTRL[7] == 0.0) ? 1.0e-03 : RCNTRL[7];
  if (RCNTRL[7] != 0.0 && (m.ThetaMin <= 0.0 || m.ThetaMin >= 1.0)) ierr = -5;
  m.NewtonTol = (RCNTRL[8] == 0.0) ? 3.0e-02 : RCNTRL[8];
  if (RCNTRL[8] != 0.0 && m.NewtonTol <= roundoff) ierr = -6;
  m.Qmin = (RCNTRL[9] == 0.0) ? 1.0 : RCNTRL[9];
  m.Qmax = (RCNTRL[10] == 0.0) ? 1.2 : RCNTRL[10];
  if (m.Qmin > 1.0 || m.Qmax < 1.0) ierr = -7;
  if (m.ITOL == 0) {
    if (atol[0] <= 0.0 || rtol[0] <= 10.0 * roundoff) ierr = -8;
  } else {
    for (int i = 0; i < nvar; i++)
      if (atol[i] <= 0.0 || rtol[i] <= 10.0 * roundoff) ierr = -8;
  }
  return ierr;
}

int device_cus() {
  int dev = 0, n = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  return n;
}

constexpr size_t kHead = 256;  // queue counter
constexpr int TILE_LANES = 64;  // cells per tile
inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct Geometry {
  int nblocks;
  int refill_min;
  int lanes;
  int lds_entries;
  size_t tol_bytes, tile_bytes, rpre_bytes, total;
};

bool geometry(const MechOps* m, int S, const kpp_launch* l, int ncells, Geometry& g,
              bool no_rpre = false) {
  const bool mixed = l && l->lu_fp32;
  const kpphip::MixedOps* mx = mixed ? mixed_of(m) : nullptr;
  if (mixed && !mx) return false;
  const int rows = mixed ? mx->rows(S) : m->rows(S);
  if (rows <= 0) return false;
  const int variant = l ? l->variant : KPP_VARIANT_PLAIN;
  if (variant != KPP_VARIANT_PLAIN && variant != KPP_VARIANT_PERSISTENT &&
      variant != KPP_VARIANT_REGISTER && variant != KPP_VARIANT_COOP)
    return false;
  if (variant == KPP_VARIANT_REGISTER && (mixed || !m->fits_registers)) return false;
  if (variant == KPP_VARIANT_COOP) {
    // one wavefront per cell: a block of coop_rows doubles per wavefront,
    // LU and scratch in LDS; FP64 only, no LDS mirror, 64 lanes
    if (mixed || !m->has_coop || (l && l->lds_entries != 0) || (l && l->lanes != 0 && l->lanes != 64))
      return false;
    int cus = device_cus();
    if (cus <= 0) cus = 256;
    int occ = m->occupancy(variant, S, 0);
    if (occ <= 0) occ = 1;
    int nb = (l && l->nblocks > 0) ? l->nblocks : cus * occ;
    if (nb > ncells) nb = ncells > 0 ? ncells : 1;
    g.nblocks = nb;
    g.refill_min = 0;
    g.lanes = 64;
    g.lds_entries = 0;
    g.tol_bytes = align256(3 * (size_t)m->nvar * sizeof(double));
    g.tile_bytes = align256((size_t)m->coop_rows(S) * sizeof(double));
    g.rpre_bytes = no_rpre ? 0 : align256((size_t)m->nreact * (size_t)ncells * sizeof(double));
    g.total = kHead + g.tol_bytes + g.rpre_bytes + (size_t)g.nblocks * g.tile_bytes;
    return true;
  }
  if (l && l->lu_fp32 != 0 && l->lu_fp32 != 1) return false;
  // LDS mirror of LU entries: persistent FP64 kernel, sizes the mechanism is built with
  const int lds = l ? l->lds_entries : 0;
  if (lds != 0) {
    if (variant != KPP_VARIANT_PERSISTENT || mixed || lds > m->lds_max ||
        (lds != KPP_LDS_SPLIT_A && lds != KPP_LDS_SPLIT_B))
      return false;
  }
  g.lds_entries = lds;
  int cus = device_cus();
  if (cus <= 0) cus = 256;
  int occ = mixed ? mx->occupancy(variant, S, 0) : m->occupancy(variant, S, lds);
  if (occ <= 0) occ = 4;
  const int resident = cus * occ;  // wavefronts the device holds at once
  int lanes = (l && l->lanes > 0) ? l->lanes : 0;
  if (lanes > 64) return false;
  if (lanes == 0) {
    // 64 cells per wavefront unless that leaves resident wavefront slots
    // empty; then fewer (down to 16), so more wavefronts share the latency.
    // Not for large mechanisms: with thousands of LU entries per cell, full
    // wavefronts win (MCM, 16,384 cells on MI355X: 64 lanes 1.74 s, 32 lanes
    // 1.78 s, 16 lanes 1.88 s, 8 lanes slower still).
    const int min_lanes = (m->lu_nonzero > 4096) ? 64 : 16;
    lanes = 64;
    while (lanes > min_lanes && (long)(ncells + lanes - 1) / lanes < resident) lanes /= 2;
  }
  g.lanes = lanes;
  if (variant == KPP_VARIANT_PERSISTENT) {
    int nb = (l && l->nblocks > 0) ? l->nblocks : resident;
    const int need = (ncells + lanes - 1) / lanes;
    if (nb > need) nb = need > 0 ? need : 1;
    g.nblocks = nb;
    g.refill_min = (l && l->refill_min > 0) ? l->refill_min : 1;
  } else {
    g.nblocks = (ncells + lanes - 1) / lanes;
    g.refill_min = 0;
  }
  g.tol_bytes = align256(3 * (size_t)m->nvar * sizeof(double));
  // the register kernel keeps the working set in VGPRs: no tiles
  g.tile_bytes = (variant == KPP_VARIANT_REGISTER) ? 0 : (size_t)rows * 64 * sizeof(double);
  // persistent kernel: rate constants of every cell from a pre-pass
  g.rpre_bytes = (variant == KPP_VARIANT_PERSISTENT && !no_rpre)
                     ? align256((size_t)m->nreact * (size_t)ncells * sizeof(double)) : 0;
  g.total = kHead + g.tol_bytes + g.rpre_bytes + (size_t)g.nblocks * g.tile_bytes;
  return true;
}

int check_params(const int* IPAR, const double* RPAR, const double* atol, const double* rtol,
                 int nvar) {
  // order and codes of int/rosenbrock.c:286-382
  const int max_steps = (IPAR[2] == 0) ? 100000 : IPAR[2];
  if (max_steps < 0) return -1;
  if ((IPAR[3] < 1) || (IPAR[3] > 5)) return -2;
  const double roundoff = unit_roundoff();
  if (RPAR[0] < 0.0) return -3;
  if (RPAR[1] < 0.0) return -3;
  if (RPAR[2] < 0.0) return -3;
  for (int i = 3; i <= 6; i++)
    if (RPAR[i] < 0.0) return -4;
  const int ntol = (IPAR[1] == 0) ? nvar : 1;
  for (int i = 0; i < ntol; i++)
    if ((atol[i] <= 0.0) || (rtol[i] <= 10.0 * roundoff) || (rtol[i] >= 1.0)) return -5;
  return 0;
}

// INTEGRATE's merge (int/rosenbrock.f90:57-80): ICNTRL = 0 and ICNTRL(15) =
// 5, then the caller's ICNTRL_U where /= 0; RCNTRL = 0, then RCNTRL_U where
// > 0 (a null array is an absent optional argument)
void f90_merge(const int* ICNTRL_U, const double* RCNTRL_U, int ICNTRL[20], double RCNTRL[20]) {
  for (int i = 0; i < 20; i++) {
    ICNTRL[i] = 0;
    RCNTRL[i] = 0.0;
  }
  ICNTRL[14] = 5;
  for (int i = 0; i < 20; i++) {
    if (ICNTRL_U && ICNTRL_U[i] != 0) ICNTRL[i] = ICNTRL_U[i];
    if (RCNTRL_U && RCNTRL_U[i] > 0.0) RCNTRL[i] = RCNTRL_U[i];
  }
}

// order and codes of the checks of int/rosenbrock.f90:306-433
int check_params_f90(const int* ICNTRL, const double* RCNTRL, const double* atol,
                     const double* rtol, int nvar) {
  RosCoef c;
  if (!fill_method_f90(ICNTRL[2], c)) return -2;
  if (ICNTRL[3] < 0) return -1;
  for (int i = 0; i <= 2; i++)
    if (RCNTRL[i] < 0.0) return -3;
  for (int i = 3; i <= 6; i++)
    if (RCNTRL[i] < 0.0) return -4;
  const double roundoff = 2.220446049250313e-16;  // EPSILON(0.0_dp)
  const int ntol = (ICNTRL[1] == 0) ? nvar : 1;
  for (int i = 0; i < ntol; i++)
    if ((atol[i] <= 0.0) || (rtol[i] <= 10.0 * roundoff) || (rtol[i] >= 1.0)) return -5;
  return 0;
}

// An F90 call: the merged controls, the passive species and the rate updates
// ICNTRL(15) selects (util/util.f90:133-159).
struct F90Call {
  const int* ICNTRL;
  const double* RCNTRL;
  const int* passive;  // host [NVAR] (1 = passive) or null
  bool h211b;          // int/rosenbrock_h211b_qssa.f90 instead of int/rosenbrock.f90
};

__global__ void k_fill_int(int* p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// The ICNTRL[2] map of int/rosenbrock_adj.c:350-373 and its tables: those
// of the F90 integrator (Ros2 with g = 1 + 1/SQRT(2), Rodas3 with
// C(6) = -(8/3), Ros3, Ros4, Rodas4) and this file's own Rodas3_1
// (:2302-2370), which differs from the C integrator's.
bool fill_method_adj(int code, RosCoef& m) {
  if (code == 6) return false;
  if (code != 7) return fill_method_f90(code, m);
  static const double A[6] = {0.0, 1.5382237953138116, -0.36440683885434433,
                              1.538223795313811, -0.36440683885434433, 1.0};
  static const double C[6] = {-4.0919303685081028, -3.0551174378039538e-002,
                              1.7259281281917580, 0.19561160936073679,
                              1.9301670595355112, -2.6267006001193960};
  memset(&m, 0, sizeof m);
  m.S = 4;
  for (int i = 0; i < 6; i++) {
    m.A[i] = A[i];
    m.C[i] = C[i];
  }
  m.NewF[0] = 1; m.NewF[1] = 0; m.NewF[2] = 1; m.NewF[3] = 1;
  m.M[0] = 1.5382237953138116; m.M[1] = -0.36440683885434444;
  m.M[2] = 1.0000000000000002; m.M[3] = 1.0;
  m.E[3] = 1.0;
  m.ELO = 3.0;
  m.Alpha[2] = 1.0; m.Alpha[3] = 1.0;
  m.Gamma[0] = 0.5150000000000000; m.Gamma[1] = -0.57028223198756145;
  return true;
}

// RosenbrockADJ's checks (int/rosenbrock_adj.c:325-470), in its order, into
// RosParams; 0 or the error code.  ICNTRL[3] != 0 is always -1 (the
// reference tests its own Max_no_steps, still 0 there); adjoint types 3 and 4
// (continuous) are not built (-13).
int adj_controls(const int* ICNTRL, const double* RCNTRL, const double* atol,
                 const double* rtol, int nvar, RosParams& p) {
  if (!fill_method_adj(ICNTRL[2], p.coef)) return -2;
  if (ICNTRL[3] != 0) return -1;
  p.max_steps = 200000 - 1;
  if (!(ICNTRL[5] == 0 || (ICNTRL[5] >= 1 && ICNTRL[5] <= 5))) return -2;
  if (!(ICNTRL[6] == 0 || (ICNTRL[6] >= 1 && ICNTRL[6] <= 4))) return -9;
  const int type = (ICNTRL[6] == 0) ? 2 : ICNTRL[6];
  p.autonomous = ICNTRL[0] != 0;
  p.vector_tol = ICNTRL[1] == 0;
  p.adj.save_lu = ICNTRL[7] != 0;
  p.adj.adjoint = type == 2;
  const double roundoff = unit_roundoff();
  p.Roundoff = roundoff;
  if (RCNTRL[0] < 0.0 || RCNTRL[1] < 0.0 || RCNTRL[2] < 0.0) return -3;
  p.Hmin = RCNTRL[0];
  p.Hmax_user = RCNTRL[1];
  p.Hstart_user = RCNTRL[2];
  const double def[4] = {0.2, 6.0, 0.1, 0.9};
  double* dst[4] = {&p.FacMin, &p.FacMax, &p.FacRej, &p.FacSafe};
  for (int k = 0; k < 4; k++) {
    if (RCNTRL[3 + k] < 0.0) return -4;
    *dst[k] = (RCNTRL[3 + k] == 0.0) ? def[k] : RCNTRL[3 + k];
  }
  const int up = p.vector_tol ? nvar : 1;
  for (int i = 0; i < up; i++)
    if (atol[i] <= 0.0 || rtol[i] <= 10.0 * roundoff || rtol[i] >= 1.0) return -5;
  if (type == 3 || type == 4) return -13;
  return 0;
}

// The batch part of an implicit integrator's launch (kpp_sdirk_dev,
// kpp_rk_dev): tolerances into the workspace head (vector tolerances are a
// host-to-device copy, so not under graph capture), the batch's arrays, and
// the tiles after the tolerances.
int implicit_batch(const MechOps* m, const kpp_batch* b, bool vector_tol, const double* atol,
                   const double* rtol, char* ws, hipStream_t st, RosParams& p) {
  double* dtol = (double*)(ws + kHead);
  p.vector_tol = vector_tol;
  if (p.vector_tol) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) != hipSuccess) return KPP_E_HIP;
    if (cs != hipStreamCaptureStatusNone) return KPP_E_ARG;
    if (hipMemcpyAsync(dtol, atol, sizeof(double) * m->nvar, hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        hipMemcpyAsync(dtol + m->nvar, rtol, sizeof(double) * m->nvar, hipMemcpyHostToDevice,
                       st) != hipSuccess)
      return KPP_E_HIP;
    p.atol = dtol;
    p.rtol = dtol + m->nvar;
  } else {
    p.atol0 = atol[0];
    p.rtol0 = rtol[0];
  }
  p.ncells = b->ncells;
  p.ldc = b->ld;
  p.C = b->C;
  p.temp = b->temp;
  p.cfactor = b->cfactor;
  p.sun = b->sun;
  p.sun_mode = b->sun_mode;
  p.tstart = b->tstart;
  p.tend = b->tend;
  p.tstart0 = b->tstart0;
  p.tend0 = b->tend0;
  p.Roundoff = unit_roundoff();
  p.hstart = b->hstart;
  p.rconst = b->rconst;
  p.rates_given = b->rconst != nullptr;
  p.ierr = b->ierr;
  p.istat = b->istat;
  p.rstat = b->rstat;
  p.hfirst = b->hfirst;
  p.order = b->order;
  p.lanes = TILE_LANES;
  p.scratch = (double*)(ws + kHead + align256(3 * (size_t)m->nvar * sizeof(double)));
  // the work queue of the persistent kernels (workspace head), zeroed by a
  // kernel on the stream
  p.queue = (int*)ws;
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, st, p.queue, 1, 0);
  if (hipGetLastError() != hipSuccess) return KPP_E_HIP;
  return 0;
}

// the argument checks shared by the implicit integrators
bool implicit_args_ok(const kpp_batch* b, const int* ICNTRL, const double* RCNTRL,
                      const double* atol, const double* rtol) {
  if (!b || !ICNTRL || !RCNTRL || !atol || !rtol || b->ncells < 0 || b->ld < b->ncells)
    return false;
  if (b->ncells > 0 && (!b->C || !b->temp || !b->cfactor || !b->ierr)) return false;
  if (b->sun_mode == KPP_SUN_FIXED && b->ncells > 0 && !b->sun && !b->rconst) return false;
  return b->sun_mode == KPP_SUN_FIXED || b->sun_mode == KPP_SUN_TOD;
}

size_t tiles_bytes(const MechOps* m, int rows, int ncells) {
  const size_t nb = (size_t)(ncells + TILE_LANES - 1) / TILE_LANES;
  return kHead + align256(3 * (size_t)m->nvar * sizeof(double)) +
         nb * align256((size_t)rows * TILE_LANES * sizeof(double));
}

}  // namespace

extern "C" {

int kpp_mech_count(void) { return g_nmechs; }
const char* kpp_mech_name(int i) { return (i >= 0 && i < g_nmechs) ? g_mechs[i]()->name : nullptr; }

int kpp_mech_info(const char* mech, int out[10]) {
  const MechOps* m = find_mech(mech);
  if (!m) return KPP_E_MECH;
  if (!out) return KPP_E_ARG;
  out[0] = m->nvar; out[1] = m->nfix; out[2] = m->nspec; out[3] = m->nreact;
  out[4] = m->nonzero; out[5] = m->lu_nonzero; out[6] = m->rates_use_time;
  out[7] = m->n_sun_rates;
  out[8] = m->rates_use_state;
  out[9] = m->n_state_rates;
  return 0;
}

int kpp_mech_initial(const char* mech, double* C) {
  const MechOps* m = find_mech(mech);
  if (!m) return KPP_E_MECH;
  if (!C) return KPP_E_ARG;
  memcpy(C, m->init, sizeof(double) * m->nspec);
  return 0;
}

const char* kpp_mech_species(const char* mech, int k) {
  const MechOps* m = find_mech(mech);
  if (!m || k < 0 || k >= m->nspec) return nullptr;
  return m->species[k];
}

int kpp_mech_lu_pattern(const char* mech, int* icol, int* crow, int* diag) {
  const MechOps* m = find_mech(mech);
  if (!m) return KPP_E_MECH;
  if (icol) memcpy(icol, m->lu_icol, sizeof(int) * m->lu_nonzero);
  if (crow) memcpy(crow, m->lu_crow, sizeof(int) * (m->nvar + 1));
  if (diag) memcpy(diag, m->lu_diag, sizeof(int) * (m->nvar + 1));
  return 0;
}

int kpp_check_params(const int IPAR[20], const double RPAR[20], const double* atol,
                     const double* rtol, int nvar) {
  if (!IPAR || !RPAR || !atol || !rtol || nvar <= 0) return KPP_E_ARG;
  return check_params(IPAR, RPAR, atol, rtol, nvar);
}

size_t kpp_workspace_bytes(const char* mech, int method, const kpp_launch* l, int ncells) {
  const MechOps* m = find_mech(mech);
  RosCoef c;
  Geometry g;
  if (!m || !fill_method(method, c) || ncells < 0) return 0;
  if (!geometry(m, c.S, l, ncells, g)) return 0;
  return g.total;
}

}  // extern "C"

// kpp_rosenbrock_dev, and (fixed = true) its variant on a workspace of fixed
// size: there the persistent kernel computes each cell's rate constants when
// the cell starts instead of in a pre-pass over all cells (same formulae,
// same values), and its grid is cut to the wavefronts whose tiles fit.
static int rosenbrock_impl(const char* mech, const kpp_batch* b, const int IPAR[20],
                           const double RPAR[20], const double* atol, const double* rtol,
                           const kpp_launch* l, void* workspace, size_t workspace_bytes,
                           void* stream, bool fixed, const F90Call* f90 = nullptr) {
  const MechOps* m = find_mech(mech);
  if (!m) return KPP_E_MECH;
  if (!b || !IPAR || !RPAR || !atol || !rtol || b->ncells < 0 || b->ld < b->ncells)
    return KPP_E_ARG;
  if (b->ncells > 0 && (!b->C || !b->temp || !b->cfactor || !b->ierr)) return KPP_E_ARG;
  // F90: SUN from Update_SUN only when ICNTRL(15) has both the RCONST and
  // the SUN bit; rates held at their entry values without the RCONST bit
  int sun_mode = b->sun_mode;
  bool frozen = false;
  if (f90) {
    const int opt = f90->ICNTRL[14];
    const bool do_rconst = opt != -1 && (opt & 1), do_sun = opt != -1 && (opt & 4);
    sun_mode = (do_rconst && do_sun) ? KPP_SUN_TOD : KPP_SUN_FIXED;
    frozen = !do_rconst;
    if (l && (l->variant == KPP_VARIANT_COOP || l->lu_fp32)) return KPP_E_ARG;
  }
  if (sun_mode == KPP_SUN_FIXED && b->ncells > 0 && !b->sun && !b->rconst) return KPP_E_ARG;
  if (sun_mode != KPP_SUN_FIXED && sun_mode != KPP_SUN_TOD) return KPP_E_ARG;
  hipStream_t st = (hipStream_t)stream;
  const int e = f90 ? check_params_f90(f90->ICNTRL, f90->RCNTRL, atol, rtol, m->nvar)
                    : check_params(IPAR, RPAR, atol, rtol, m->nvar);
  if (e != 0) {
    // Rosenbrock returns before integrating; every cell reports the code
    if (b->ncells > 0)
      hipLaunchKernelGGL(k_fill_int, dim3((b->ncells + 255) / 256), dim3(256), 0, st, b->ierr,
                         b->ncells, e);
    return e;
  }
  if (b->ncells == 0) return 0;
  RosParams p;
  memset(&p, 0, sizeof p);
  if (!(f90 ? fill_method_f90(f90->ICNTRL[2], p.coef) : fill_method(IPAR[3], p.coef))) return -2;
  Geometry g;
  if (!geometry(m, p.coef.S, l, b->ncells, g, fixed)) return KPP_E_ARG;
  if (fixed && g.total > workspace_bytes && g.tile_bytes > 0 &&
      ((l ? l->variant : KPP_VARIANT_PLAIN) == KPP_VARIANT_PERSISTENT ||
       (l ? l->variant : KPP_VARIANT_PLAIN) == KPP_VARIANT_COOP)) {
    const size_t head = g.total - (size_t)g.nblocks * g.tile_bytes;
    const size_t fit = workspace_bytes > head ? (workspace_bytes - head) / g.tile_bytes : 0;
    if (fit >= 1) {
      g.nblocks = (int)std::min<size_t>(fit, (size_t)g.nblocks);
      g.total = head + (size_t)g.nblocks * g.tile_bytes;
    }
  }
  if (!workspace || workspace_bytes < g.total) return KPP_E_WORKSPACE;
  char* ws = (char*)workspace;
  double* dtol = (double*)(ws + kHead);
  p.vector_tol = f90 ? (f90->ICNTRL[1] == 0) : (IPAR[1] == 0);
  if (p.vector_tol) {
    // vector tolerances go through device memory: a host-to-device copy,
    // which a stream under graph capture cannot record from pageable memory
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) != hipSuccess) return KPP_E_HIP;
    if (cs != hipStreamCaptureStatusNone) return KPP_E_ARG;
    if (hipMemcpyAsync(dtol, atol, sizeof(double) * m->nvar, hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        hipMemcpyAsync(dtol + m->nvar, rtol, sizeof(double) * m->nvar, hipMemcpyHostToDevice,
                       st) != hipSuccess)
      return KPP_E_HIP;
    p.atol = dtol;
    p.rtol = dtol + m->nvar;
  } else {
    p.atol0 = atol[0];
    p.rtol0 = rtol[0];
  }
  if (f90) {
    // ros_ErrorNorm's species: 1.0 for non-passive, 0.0 for passive
    p.f90 = 1;
    p.nonneg = (f90->ICNTRL[15] == 1);
    // H211b control when both parameters are set (:490-514)
    p.bb = f90->RCNTRL[14];
    p.kk = f90->RCNTRL[15];
    p.h211b = (f90->h211b && p.bb > 0.0 && p.kk > 0.0) ? 1 : 0;
    p.np_count = m->nvar;
    if (f90->passive) {
      std::vector<double> w(m->nvar);
      int cnt = 0;
      for (int i = 0; i < m->nvar; i++) {
        w[i] = f90->passive[i] ? 0.0 : 1.0;
        cnt += f90->passive[i] ? 0 : 1;
      }
      if (cnt == 0) return KPP_E_ARG;
      hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
      if (hipStreamIsCapturing(st, &cs) != hipSuccess) return KPP_E_HIP;
      if (cs != hipStreamCaptureStatusNone) return KPP_E_ARG;
      double* dw = dtol + 2 * m->nvar;
      // synchronous: the host vector ends with this call
      if (hipMemcpyAsync(dw, w.data(), sizeof(double) * m->nvar, hipMemcpyHostToDevice, st) !=
              hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess)
        return KPP_E_HIP;
      p.np_w = dw;
      p.np_count = cnt;
    }
  }
  p.ncells = b->ncells;
  p.ldc = b->ld;
  p.C = b->C;
  p.temp = b->temp;
  p.cfactor = b->cfactor;
  p.sun = b->sun;
  p.sun_mode = sun_mode;
  p.tstart = b->tstart;
  p.tend = b->tend;
  p.tstart0 = b->tstart0;
  p.tend0 = b->tend0;
  // the controls: IPAR/RPAR (int/rosenbrock.c:286-382), or F90's
  // ICNTRL/RCNTRL (int/rosenbrock.f90:306-433; statistics start at zero)
  const double* R7 = f90 ? f90->RCNTRL : RPAR;
  if (f90) {
    p.autonomous = !(f90->ICNTRL[0] == 0);
    p.max_steps = (f90->ICNTRL[3] == 0) ? 200000 : f90->ICNTRL[3];
    p.Roundoff = 2.220446049250313e-16;
    p.hstart_auto = 0;
  } else {
    p.autonomous = !(IPAR[0] == 0);
    p.max_steps = (IPAR[2] == 0) ? 100000 : IPAR[2];
    p.Roundoff = unit_roundoff();
    p.hstart_auto = (IPAR[4] == 1);
    for (int i = 0; i < 8; i++) p.stat0[i] = IPAR[10 + i];
  }
  p.Hmin = R7[0];
  p.Hmax_user = R7[1];
  p.Hstart_user = R7[2];
  p.FacMin = (R7[3] == 0.0) ? 0.2 : R7[3];
  p.FacMax = (R7[4] == 0.0) ? 6.0 : R7[4];
  p.FacRej = (R7[5] == 0.0) ? 0.1 : R7[5];
  p.FacSafe = (R7[6] == 0.0) ? 0.9 : R7[6];
  p.hstart = b->hstart;
  p.rconst = b->rconst;
  p.rates_given = b->rconst != nullptr;
  p.rates_frozen = frozen && !p.rates_given;
  // dF/dT is exactly zero when no rate constant depends on T
  p.dfdt_zero = p.rates_given || p.rates_frozen ||
                (sun_mode == KPP_SUN_FIXED && !m->rates_use_time);
  p.ierr = b->ierr;
  p.istat = b->istat;
  p.rstat = b->rstat;
  p.hfirst = b->hfirst;
  p.fallbacks = b->fallbacks;
  p.order = b->order;
  p.queue = (int*)ws;
  p.rpre = (g.rpre_bytes && !p.rates_given) ? (double*)(ws + kHead + g.tol_bytes) : nullptr;
  p.scratch = (double*)(ws + kHead + g.tol_bytes + g.rpre_bytes);
  p.refill_min = g.refill_min;
  p.lanes = g.lanes;
  p.lds_entries = g.lds_entries;
  const int variant = l ? l->variant : KPP_VARIANT_PLAIN;
  const hipError_t le = (l && l->lu_fp32) ? mixed_of(m)->launch(p, variant, g.nblocks, st)
                                          : m->launch(p, variant, g.nblocks, st);
  if (le != hipSuccess) return KPP_E_HIP;
  return 0;
}

extern "C" {

int kpp_rosenbrock_dev(const char* mech, const kpp_batch* b, const int IPAR[20],
                       const double RPAR[20], const double* atol, const double* rtol,
                       const kpp_launch* l, void* workspace, size_t workspace_bytes,
                       void* stream) {
  return rosenbrock_impl(mech, b, IPAR, RPAR, atol, rtol, l, workspace, workspace_bytes, stream,
                         false);
}

int kpp_integrate_f90_dev(const char* mech, const kpp_batch* b, const int ICNTRL_U[20],
                          const double RCNTRL_U[20], const double* atol, const double* rtol,
                          const int* passive, const kpp_launch* l, void* workspace,
                          size_t workspace_bytes, void* stream) {
  int ICNTRL[20];
  double RCNTRL[20];
  f90_merge(ICNTRL_U, RCNTRL_U, ICNTRL, RCNTRL);
  const F90Call f{ICNTRL, RCNTRL, passive, false};
  static const int IPAR0[20] = {0};
  static const double RPAR0[20] = {0.0};
  return rosenbrock_impl(mech, b, IPAR0, RPAR0, atol, rtol, l, workspace, workspace_bytes, stream,
                         false, &f);
}

int kpp_integrate_h211b_dev(const char* mech, const kpp_batch* b, const int ICNTRL_U[20],
                            const double RCNTRL_U[20], const double* atol, const double* rtol,
                            const kpp_launch* l, void* workspace, size_t workspace_bytes,
                            void* stream) {
  int ICNTRL[20];
  double RCNTRL[20];
  f90_merge(ICNTRL_U, RCNTRL_U, ICNTRL, RCNTRL);
  if (ICNTRL[17] != 0) return KPP_E_ARG;  // the QSSA options are not built
  const F90Call f{ICNTRL, RCNTRL, nullptr, true};
  static const int IPAR0[20] = {0};
  static const double RPAR0[20] = {0.0};
  return rosenbrock_impl(mech, b, IPAR0, RPAR0, atol, rtol, l, workspace, workspace_bytes, stream,
                         false, &f);
}

size_t kpp_workspace_bytes_f90(const char* mech, int icntrl3, const kpp_launch* l, int ncells) {
  const MechOps* m = find_mech(mech);
  RosCoef c;
  Geometry g;
  if (!m || !fill_method_f90(icntrl3, c) || ncells < 0) return 0;
  if (!geometry(m, c.S, l, ncells, g)) return 0;
  return g.total;
}

int kpp_check_params_f90(const int ICNTRL_U[20], const double RCNTRL_U[20], const double* atol,
                         const double* rtol, int nvar) {
  if (!atol || !rtol || nvar <= 0) return KPP_E_ARG;
  int ICNTRL[20];
  double RCNTRL[20];
  f90_merge(ICNTRL_U, RCNTRL_U, ICNTRL, RCNTRL);
  return check_params_f90(ICNTRL, RCNTRL, atol, rtol, nvar);
}

size_t kpp_workspace_bytes_sdirk(const char* mech, int icntrl2, int ncells) {
  const MechOps* m = find_mech(mech);
  const kpphip::SdirkOps* so = m ? sdirk_of(m) : nullptr;
  if (!so || ncells < 0) return 0;
  kpphip::SdCoef c;
  fill_sdirk(icntrl2, c);
  const int rows = so->rows(c.S);
  if (rows <= 0) return 0;
  return tiles_bytes(m, rows, ncells);
}

int kpp_sdirk_dev(const char* mech, const kpp_batch* b, const int ICNTRL[20],
                  const double RCNTRL[20], const double* atol, const double* rtol,
                  void* workspace, size_t workspace_bytes, void* stream) {
  const MechOps* m = find_mech(mech);
  if (!m) return KPP_E_MECH;
  const kpphip::SdirkOps* so = sdirk_of(m);
  if (!implicit_args_ok(b, ICNTRL, RCNTRL, atol, rtol)) return KPP_E_ARG;
  if (!so || so->rows(2) <= 0) return KPP_E_ARG;  // not built for this mechanism
  hipStream_t st = (hipStream_t)stream;
  RosParams p;
  memset(&p, 0, sizeof p);
  const int e = sdirk_controls(ICNTRL, RCNTRL, atol, rtol, m->nvar, p.sd);
  if (e != 0) {
    if (b->ncells > 0)
      hipLaunchKernelGGL(k_fill_int, dim3((b->ncells + 255) / 256), dim3(256), 0, st, b->ierr,
                         b->ncells, e);
    return e;
  }
  if (b->ncells == 0) return 0;
  const size_t need = kpp_workspace_bytes_sdirk(mech, ICNTRL[2], b->ncells);
  if (!workspace || workspace_bytes < need) return KPP_E_WORKSPACE;
  const int r = implicit_batch(m, b, ICNTRL[1] == 0, atol, rtol, (char*)workspace, st, p);
  if (r != 0) return r;
  const int nblocks = (b->ncells + TILE_LANES - 1) / TILE_LANES;
  if (so->launch(p, nblocks, st) != hipSuccess) return KPP_E_HIP;
  return 0;
}

size_t kpp_workspace_bytes_rk(const char* mech, int ncells) {
  const MechOps* m = find_mech(mech);
  const kpphip::SdirkOps* ro = m ? rk_of(m) : nullptr;
  if (!ro || ncells < 0) return 0;
  const int rows = ro->rows(0);
  if (rows <= 0) return 0;
  return tiles_bytes(m, rows, ncells);
}

int kpp_rk_dev(const char* mech, const kpp_batch* b, const int ICNTRL[20],
               const double RCNTRL[20], const double* atol, const double* rtol, void* workspace,
               size_t workspace_bytes, void* stream) {
  const MechOps* m = find_mech(mech);
  if (!m) return KPP_E_MECH;
  const kpphip::SdirkOps* ro = rk_of(m);
  if (!implicit_args_ok(b, ICNTRL, RCNTRL, atol, rtol)) return KPP_E_ARG;
  if (!ro || ro->rows(0) <= 0) return KPP_E_ARG;  // not built for this mechanism
  hipStream_t st = (hipStream_t)stream;
  RosParams p;
  memset(&p, 0, sizeof p);
  const int e = rk_controls(ICNTRL, RCNTRL, atol, rtol, m->nvar, p.rk);
  if (e != 0) {
    if (b->ncells > 0)
      hipLaunchKernelGGL(k_fill_int, dim3((b->ncells + 255) / 256), dim3(256), 0, st, b->ierr,
                         b->ncells, e);
    return e;
  }
  if (b->ncells == 0) return 0;
  const size_t need = kpp_workspace_bytes_rk(mech, b->ncells);
  if (!workspace || workspace_bytes < need) return KPP_E_WORKSPACE;
  const int r = implicit_batch(m, b, ICNTRL[1] == 0, atol, rtol, (char*)workspace, st, p);
  if (r != 0) return r;
  const int nblocks = (b->ncells + TILE_LANES - 1) / TILE_LANES;
  if (ro->launch(p, nblocks, st) != hipSuccess) return KPP_E_HIP;
  return 0;
}

// rows of a lane's work tile and of one checkpoint slot of the adjoint
static bool adj_rows(const MechOps* m, const RosCoef& c, int nadj, bool save_lu, int& rows,
                     int& slot) {
  const kpphip::SdirkOps* ao = adj_of(m);
  const int fixed = ao ? ao->rows(0) : 0;
  if (fixed <= 0) return false;
  rows = fixed + nadj * m->nvar * (1 + 2 * c.S);
  slot = 2 + 2 * c.S * m->nvar + (save_lu ? m->lu_nonzero : 0);
  return true;
}

size_t kpp_workspace_bytes_adj(const char* mech, int ncells, const int ICNTRL[20], int nadj,
                               int nchk) {
  const MechOps* m = find_mech(mech);
  RosCoef c;
  memset(&c, 0, sizeof c);
  if (!m || !ICNTRL || ncells < 0 || nadj < 0 || nchk < 0 || !fill_method_adj(ICNTRL[2], c))
    return 0;
  int rows = 0, slot = 0;
  if (!adj_rows(m, c, nadj, ICNTRL[7] != 0, rows, slot)) return 0;
  const size_t nb = (size_t)(ncells + TILE_LANES - 1) / TILE_LANES;
  return tiles_bytes(m, rows, ncells) + nb * (size_t)nchk * slot * TILE_LANES * sizeof(double);
}

int kpp_rosenbrock_adj_dev(const char* mech, const kpp_batch* b, const int ICNTRL[20],
                           const double RCNTRL[20], const double* atol, const double* rtol,
                           int nadj, double* lambda, int nchk, int first_step, void* workspace,
                           size_t workspace_bytes, void* stream) {
  const MechOps* m = find_mech(mech);
  if (!m) return KPP_E_MECH;
  if (!implicit_args_ok(b, ICNTRL, RCNTRL, atol, rtol)) return KPP_E_ARG;
  if (nadj < 0 || nchk < 0 || (nadj > 0 && b->ncells > 0 && !lambda)) return KPP_E_ARG;
  const kpphip::SdirkOps* ao = adj_of(m);
  if (!ao || ao->rows(0) <= 0) return KPP_E_ARG;  // not built for this mechanism
  hipStream_t st = (hipStream_t)stream;
  RosParams p;
  memset(&p, 0, sizeof p);
  const int e = adj_controls(ICNTRL, RCNTRL, atol, rtol, m->nvar, p);
  if (e != 0) {
    if (b->ncells > 0)
      hipLaunchKernelGGL(k_fill_int, dim3((b->ncells + 255) / 256), dim3(256), 0, st, b->ierr,
                         b->ncells, e);
    return e;
  }
  if (b->ncells == 0) return 0;
  int rows = 0, slot = 0;
  adj_rows(m, p.coef, nadj, p.adj.save_lu != 0, rows, slot);
  const size_t need = kpp_workspace_bytes_adj(mech, b->ncells, ICNTRL, nadj, nchk);
  if (!workspace || workspace_bytes < need) return KPP_E_WORKSPACE;
  const int r = implicit_batch(m, b, ICNTRL[1] == 0, atol, rtol, (char*)workspace, st, p);
  if (r != 0) return r;
  const int nblocks = (b->ncells + TILE_LANES - 1) / TILE_LANES;
  p.adj.nadj = nadj;
  p.adj.lambda = lambda;
  p.adj.nchk = nchk;
  p.adj.first_step = first_step ? 1 : 0;
  p.adj.chk = p.scratch + (size_t)nblocks * rows * TILE_LANES;
  if (ao->launch(p, nblocks, st) != hipSuccess) return KPP_E_HIP;
  return 0;
}

static int component(const char* mech, int which, int n, int ld, const double* a, const double* b,
                     const double* c, const double* d, const double* e, double* out, int* iout,
                     void* stream) {
  const MechOps* m = find_mech(mech);
  if (!m) return KPP_E_MECH;
  if (n < 0 || ld < n) return KPP_E_ARG;
  kpphip::ComponentFn fn = component_of(m);
  if (!fn || fn(which, n, (size_t)ld, a, b, c, d, e, out, iout, (hipStream_t)stream) !=
                 hipSuccess)
    return KPP_E_HIP;
  return 0;
}

int kpp_fun_dev(const char* mech, int n, int ld, const double* V, const double* F,
                const double* RCT, double* Vdot, void* stream) {
  return component(mech, KPP_COMP_FUN, n, ld, V, F, RCT, nullptr, nullptr, Vdot, nullptr, stream);
}
int kpp_jac_sp_dev(const char* mech, int n, int ld, const double* V, const double* F,
                   const double* RCT, double* JVS, void* stream) {
  return component(mech, KPP_COMP_JAC, n, ld, V, F, RCT, nullptr, nullptr, JVS, nullptr, stream);
}
int kpp_decomp_dev(const char* mech, int n, int ld, double* JVS, int* ising, void* stream) {
  return component(mech, KPP_COMP_DECOMP, n, ld, nullptr, nullptr, nullptr, nullptr, nullptr, JVS,
                   ising, stream);
}
int kpp_solve_dev(const char* mech, int n, int ld, const double* JVS, double* X, void* stream) {
  return component(mech, KPP_COMP_SOLVE, n, ld, JVS, nullptr, nullptr, nullptr, nullptr, X,
                   nullptr, stream);
}
int kpp_rates_dev(const char* mech, int n, int ld, const double* temp, const double* cfactor,
                  const double* sun, const double* time, double* RCONST, void* stream) {
  return component(mech, KPP_COMP_RATES, n, ld, temp, cfactor, sun, time, nullptr, RCONST,
                   nullptr, stream);
}
int kpp_rates_c_dev(const char* mech, int n, int ld, const double* temp, const double* cfactor,
                    const double* sun, const double* time, const double* C, double* RCONST,
                    void* stream) {
  if (!C) return KPP_E_ARG;
  return component(mech, KPP_COMP_RATES, n, ld, temp, cfactor, sun, time, C, RCONST, nullptr,
                   stream);
}

// ---------------------------------------------------------------------------
// Host-buffer context API
// ---------------------------------------------------------------------------
struct kpp_ctx {
  const MechOps* m;
  int ncells, method, sun_mode, flags;
  double *C, *temp, *cfactor, *sun, *tstart, *tend, *hstart, *rstat;
  double* hused;   // starting steps the last call passed in (0: RPAR[2]'s rule)
  double* C0;      // KPP_CTX_AUTOTUNE: the batch's state before the trials
  int *ierr, *istat;
  int* order;      // KPP_CTX_LEARNED: launch order
  void* ws;
  size_t ws_bytes;
  std::vector<double> atol, rtol;
  int vector_tol;
  std::vector<int> passive;  // INTEGRATE_F90_BATCH: passive species (empty: none)
  kpp_launch launch;  // launch shape of the integrations
  bool tuned;         // KPP_CTX_AUTOTUNE: the launch shape has been chosen
  bool carried;       // INTEGRATE_BATCH: hstart holds the previous call's Hexit
  kpp_sched* sched;   // KPP_CTX_LEARNED
};

static void ctx_free(kpp_ctx* c) {
  if (!c) return;
  void* ptrs[] = {c->C, c->temp, c->cfactor, c->sun, c->tstart, c->tend, c->hstart, c->hused,
                  c->C0, c->rstat, c->ierr, c->istat, c->order, c->ws};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  if (c->sched) kpp_sched_destroy(c->sched);
  delete c;
}

// <directory of this library>/<mech>_step_mlp.kppw, where the build puts the
// exported scheduler weights
static std::string default_weights(const char* mech) {
  Dl_info info;
  if (!dladdr((void*)&kpp_ctx_create, &info) || !info.dli_fname) return "";
  std::string lib(info.dli_fname);
  const size_t k = lib.find_last_of('/');
  const std::string dir = (k == std::string::npos) ? "." : lib.substr(0, k);
  return dir + "/" + mech + "_step_mlp.kppw";
}

int kpp_ctx_create(kpp_ctx** out, const char* mech, int ncells, int method, int flags) {
  if (!out || ncells <= 0) return KPP_E_ARG;
  *out = nullptr;
  const MechOps* m = find_mech(mech);
  if (!m) return KPP_E_MECH;
  const int known = KPP_SUN_TOD | KPP_CTX_PERSISTENT | KPP_CTX_AUTOTUNE | KPP_CTX_LEARNED |
                    KPP_CTX_MIXED | KPP_CTX_HSTART_AUTO;
  if ((flags & ~known) || ((flags & KPP_CTX_AUTOTUNE) && (flags & KPP_CTX_MIXED)))
    return KPP_E_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return KPP_E_NODEV;
  kpp_ctx* c = new kpp_ctx();  // value-initialised: all pointers null
  c->m = m;
  c->ncells = ncells;
  c->method = method;
  c->flags = flags;
  c->sun_mode = (flags & KPP_SUN_TOD) ? KPP_SUN_TOD : KPP_SUN_FIXED;
  c->atol.assign(1, 1.0);
  c->rtol.assign(1, 1.0e-3);
  c->vector_tol = 0;
  const size_t n = (size_t)ncells;
  memset(&c->launch, 0, sizeof c->launch);
  c->launch.variant = (flags & (KPP_CTX_PERSISTENT | KPP_CTX_AUTOTUNE)) ? KPP_VARIANT_PERSISTENT
                                                                         : KPP_VARIANT_PLAIN;
  c->launch.lu_fp32 = (flags & KPP_CTX_MIXED) ? 1 : 0;
  c->ws_bytes = kpp_workspace_bytes(mech, method, &c->launch, ncells);
  if (c->ws_bytes == 0) { ctx_free(c); return -2; }
  if (flags & KPP_CTX_AUTOTUNE)
    c->ws_bytes = std::max(c->ws_bytes, kpp_autotune_workspace_bytes(mech, method, ncells));
  bool ok = hipMalloc(&c->C, sizeof(double) * n * m->nspec) == hipSuccess &&
            hipMalloc(&c->temp, sizeof(double) * n) == hipSuccess &&
            hipMalloc(&c->cfactor, sizeof(double) * n) == hipSuccess &&
            hipMalloc(&c->sun, sizeof(double) * n) == hipSuccess &&
            hipMalloc(&c->tstart, sizeof(double) * n) == hipSuccess &&
            hipMalloc(&c->tend, sizeof(double) * n) == hipSuccess &&
            hipMalloc(&c->hstart, sizeof(double) * n) == hipSuccess &&
            hipMalloc(&c->hused, sizeof(double) * n) == hipSuccess &&
            hipMalloc(&c->rstat, sizeof(double) * 3 * n) == hipSuccess &&
            hipMalloc(&c->ierr, sizeof(int) * n) == hipSuccess &&
            hipMalloc(&c->istat, sizeof(int) * 8 * n) == hipSuccess &&
            hipMalloc(&c->ws, c->ws_bytes) == hipSuccess;
  if (ok && (flags & KPP_CTX_LEARNED)) ok = hipMalloc(&c->order, sizeof(int) * n) == hipSuccess;
  if (ok && (flags & KPP_CTX_AUTOTUNE))
    ok = hipMalloc(&c->C0, sizeof(double) * n * m->nspec) == hipSuccess;
  if (!ok) { ctx_free(c); return KPP_E_HIP; }
  if (hipMemset(c->hstart, 0, sizeof(double) * n) != hipSuccess ||
      hipMemset(c->hused, 0, sizeof(double) * n) != hipSuccess ||
      hipMemset(c->sun, 0, sizeof(double) * n) != hipSuccess) {
    ctx_free(c);
    return KPP_E_HIP;
  }
  if (flags & KPP_CTX_LEARNED) {
    const std::string w = default_weights(mech);
    const int r = w.empty() ? KPP_E_WEIGHTS : kpp_sched_load(w.c_str(), mech, &c->sched);
    if (r != 0) { ctx_free(c); return r; }
  }
  *out = c;
  return 0;
}

void kpp_ctx_destroy(kpp_ctx* c) { ctx_free(c); }

int kpp_ctx_load_scheduler(kpp_ctx* c, const char* path) {
  if (!c || !path || !(c->flags & KPP_CTX_LEARNED)) return KPP_E_ARG;
  kpp_sched* s = nullptr;
  const int r = kpp_sched_load(path, c->m->name, &s);
  if (r != 0) return r;
  if (c->sched) kpp_sched_destroy(c->sched);
  c->sched = s;
  return 0;
}

int kpp_ctx_get_launch(kpp_ctx* c, kpp_launch* l) {
  if (!c || !l) return KPP_E_ARG;
  *l = c->launch;
  return 0;
}

int kpp_ctx_get_hstart(kpp_ctx* c, double* h) {
  if (!c || !h) return KPP_E_ARG;
  return hipMemcpy(h, c->hused, sizeof(double) * c->ncells, hipMemcpyDeviceToHost) == hipSuccess
             ? 0
             : KPP_E_HIP;
}

int kpp_set_state(kpp_ctx* c, const double* C, const double* temp, const double* cfactor,
                  const double* sun) {
  if (!c || !C || !temp || !cfactor) return KPP_E_ARG;
  if (c->sun_mode == KPP_SUN_FIXED && !sun) return KPP_E_ARG;
  const size_t n = (size_t)c->ncells;
  if (hipMemcpy(c->C, C, sizeof(double) * n * c->m->nspec, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->temp, temp, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->cfactor, cfactor, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess)
    return KPP_E_HIP;
  if (sun && hipMemcpy(c->sun, sun, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess)
    return KPP_E_HIP;
  return 0;
}

__global__ static void k_fill_double(double* p, int n, double v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// How a call's starting steps are given.
enum HstartMode {
  HS_RPAR,     // none: RPAR[2]'s rule for every cell
  HS_GIVEN,    // c->hstart holds the caller's values
  HS_CARRIED,  // INTEGRATE_BATCH: c->hstart holds the previous call's Hexit (if c->carried)
};

static int ctx_run(kpp_ctx* c, const int* IPAR, const double* RPAR, const double* atol,
                   const double* rtol, HstartMode hs, double span) {
  kpp_batch b;
  memset(&b, 0, sizeof b);
  b.ncells = c->ncells;
  b.ld = c->ncells;
  b.C = c->C;
  b.temp = c->temp;
  b.cfactor = c->cfactor;
  b.sun = c->sun;
  b.sun_mode = c->sun_mode;
  b.tstart = c->tstart;
  b.tend = c->tend;
  b.hstart = (hs == HS_RPAR) ? nullptr : c->hstart;
  b.ierr = c->ierr;
  b.istat = c->istat;
  b.rstat = c->rstat;
  const int n = c->ncells, nb = (n + 255) / 256;
  if (c->sched) {
    // learned starting steps where the call gives none (RPAR[2] = 0, or the
    // first INTEGRATE_BATCH of the context), and the launch order
    const bool predict = (hs == HS_RPAR && RPAR[2] == 0.0) || (hs == HS_CARRIED && !c->carried);
    const double hmin = std::max(RPAR[0], (double)KPP_DELTAMIN);
    const int r = kpp_sched_run(c->sched, &b, hmin, std::max(span, hmin),
                                predict ? c->hstart : nullptr, c->order, nullptr);
    if (r != 0) return r;
    if (predict) b.hstart = c->hstart;
    b.order = c->order;
  }
  if (b.hstart) {
    if (hipMemcpyAsync(c->hused, b.hstart, sizeof(double) * n, hipMemcpyDeviceToDevice, nullptr) !=
        hipSuccess)
      return KPP_E_HIP;
  } else {
    hipLaunchKernelGGL(k_fill_double, dim3(nb), dim3(256), 0, nullptr, c->hused, n, RPAR[2]);
  }
  if ((c->flags & KPP_CTX_AUTOTUNE) && !c->tuned) {
    // the launch shape is chosen once, by timing candidates on this batch
    if (hipMemcpyAsync(c->C0, c->C, sizeof(double) * (size_t)n * c->m->nspec,
                       hipMemcpyDeviceToDevice, nullptr) != hipSuccess)
      return KPP_E_HIP;
    kpp_launch best;
    float ms = 0.f;
    int tried = 0;
    const int r = kpp_autotune(c->m->name, &b, c->C0, IPAR, RPAR, atol, rtol, c->ws, c->ws_bytes,
                               nullptr, &best, &ms, &tried);
    if (r != 0) return r;
    c->launch = best;
    c->tuned = true;
  }
  int r = kpp_rosenbrock_dev(c->m->name, &b, IPAR, RPAR, atol, rtol, &c->launch, c->ws,
                             c->ws_bytes, nullptr);
  if (hipDeviceSynchronize() != hipSuccess) return KPP_E_HIP;
  return r;
}

int kpp_rosenbrock_batch(kpp_ctx* c, const double* tstart, const double* tend,
                         const double* atol, const double* rtol, const int IPAR[20],
                         const double RPAR[20], const double* hstart) {
  if (!c || !tstart || !tend || !atol || !rtol || !IPAR || !RPAR) return KPP_E_ARG;
  if (IPAR[3] != c->method) return KPP_E_ARG;
  const size_t n = (size_t)c->ncells;
  if (hipMemcpy(c->tstart, tstart, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->tend, tend, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess)
    return KPP_E_HIP;
  if (hstart &&
      hipMemcpy(c->hstart, hstart, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess)
    return KPP_E_HIP;
  double span = 0.0;
  for (size_t i = 0; i < n; i++) span = std::max(span, fabs(tend[i] - tstart[i]));
  return ctx_run(c, IPAR, RPAR, atol, rtol, hstart ? HS_GIVEN : HS_RPAR, span);
}

int kpp_get_state(kpp_ctx* c, double* C, int* ierr, int* istat, double* rstat) {
  if (!c) return KPP_E_ARG;
  const size_t n = (size_t)c->ncells;
  if (C && hipMemcpy(C, c->C, sizeof(double) * n * c->m->nspec, hipMemcpyDeviceToHost) != hipSuccess)
    return KPP_E_HIP;
  if (ierr && hipMemcpy(ierr, c->ierr, sizeof(int) * n, hipMemcpyDeviceToHost) != hipSuccess)
    return KPP_E_HIP;
  if (istat && hipMemcpy(istat, c->istat, sizeof(int) * 8 * n, hipMemcpyDeviceToHost) != hipSuccess)
    return KPP_E_HIP;
  if (rstat && hipMemcpy(rstat, c->rstat, sizeof(double) * 2 * n, hipMemcpyDeviceToHost) != hipSuccess)
    return KPP_E_HIP;
  return 0;
}

int kpp_ctx_set_tol(kpp_ctx* c, const double* atol, const double* rtol, int vector_tol) {
  if (!c || !atol || !rtol) return KPP_E_ARG;
  const int n = vector_tol ? c->m->nvar : 1;
  c->atol.assign(atol, atol + n);
  c->rtol.assign(rtol, rtol + n);
  c->vector_tol = vector_tol ? 1 : 0;
  return 0;
}

int kpp_ctx_set_passive(kpp_ctx* c, const int* passive) {
  if (!c) return KPP_E_ARG;
  if (!passive) {
    c->passive.clear();
    return 0;
  }
  int cnt = 0;
  for (int i = 0; i < c->m->nvar; i++) cnt += passive[i] ? 0 : 1;
  if (cnt == 0) return KPP_E_ARG;
  c->passive.assign(passive, passive + c->m->nvar);
  return 0;
}

__global__ static void k_fill_time(double* ts, double* te, int n, double t0, double t1) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) { ts[i] = t0; te[i] = t1; }
}
__global__ static void k_copy_hexit(double* hs, const double* rstat, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) hs[i] = rstat[n + i];
}

// INTEGRATE (int/rosenbrock.c:94-133): non-autonomous, RPAR[2] = STEPMIN
// (per cell: the previous call's Hexit), method fixed at context creation
// (the reference hard-codes 5 = Rodas4), tolerances of the context.  With
// KPP_CTX_LEARNED the first call starts each cell from the scheduler's step
// (STEPMIN is still 0 there); with KPP_CTX_HSTART_AUTO from the estimate.
int INTEGRATE_BATCH(kpp_ctx* c, double TIN, double TOUT) {
  if (!c) return KPP_E_ARG;
  int IPAR[20];
  double RPAR[20];
  memset(IPAR, 0, sizeof IPAR);
  memset(RPAR, 0, sizeof RPAR);
  IPAR[0] = 0;
  IPAR[1] = c->vector_tol ? 0 : 1;
  IPAR[3] = c->method;
  IPAR[4] = (c->flags & KPP_CTX_HSTART_AUTO) ? 1 : 0;
  const int nb = (c->ncells + 255) / 256;
  hipLaunchKernelGGL(k_fill_time, dim3(nb), dim3(256), 0, 0, c->tstart, c->tend, c->ncells, TIN,
                     TOUT);
  int r = ctx_run(c, IPAR, RPAR, c->atol.data(), c->rtol.data(), HS_CARRIED, fabs(TOUT - TIN));
  if (r != 0) return r;
  hipLaunchKernelGGL(k_copy_hexit, dim3(nb), dim3(256), 0, 0, c->hstart, c->rstat, c->ncells);
  if (hipDeviceSynchronize() != hipSuccess) return KPP_E_HIP;
  c->carried = true;
  return 0;
}

// INTEGRATE of int/rosenbrock.f90:43-125 for every cell of the context:
// ICNTRL_U/RCNTRL_U as the F90 caller passes them (null: absent), the
// context's tolerances (ICNTRL(2) = 0 needs vector tolerances) and passive
// species, SUN and rate updates as ICNTRL(15) selects.  The F90 INTEGRATE
// carries no step between calls (it only stores STEPMIN = RSTATUS(2)).
// Outputs to host arrays, each optional: ISTATUS_U [8][n] (ISTATUS(1:8)),
// RSTATUS_U [3][n] (Texit, Hexit, Hnew), IERR_U [n].  The plain or
// persistent kernel of the context, never the learned scheduler.
int INTEGRATE_F90_BATCH(kpp_ctx* c, double TIN, double TOUT, const int ICNTRL_U[20],
                        const double RCNTRL_U[20], int* ISTATUS_U, double* RSTATUS_U,
                        int* IERR_U) {
  if (!c || (c->flags & KPP_CTX_MIXED)) return KPP_E_ARG;
  int ICNTRL[20];
  double RCNTRL[20];
  f90_merge(ICNTRL_U, RCNTRL_U, ICNTRL, RCNTRL);
  if (ICNTRL[1] == 0 && !c->vector_tol) return KPP_E_ARG;
  const int n = c->ncells, nb = (n + 255) / 256;
  hipLaunchKernelGGL(k_fill_time, dim3(nb), dim3(256), 0, 0, c->tstart, c->tend, n, TIN, TOUT);
  kpp_batch b;
  memset(&b, 0, sizeof b);
  b.ncells = n;
  b.ld = n;
  b.C = c->C;
  b.temp = c->temp;
  b.cfactor = c->cfactor;
  b.sun = c->sun;
  b.sun_mode = KPP_SUN_FIXED;
  b.tstart = c->tstart;
  b.tend = c->tend;
  b.ierr = c->ierr;
  b.istat = c->istat;
  b.rstat = c->rstat;
  kpp_launch l = c->launch;
  l.lds_entries = 0;
  const size_t need = kpp_workspace_bytes_f90(c->m->name, ICNTRL[2], &l, n);
  if (need > c->ws_bytes) {
    void* ws = nullptr;
    if (hipMalloc(&ws, need) != hipSuccess) return KPP_E_HIP;
    (void)hipFree(c->ws);
    c->ws = ws;
    c->ws_bytes = need;
  }
  int r = kpp_integrate_f90_dev(c->m->name, &b, ICNTRL_U, RCNTRL_U, c->atol.data(),
                                c->rtol.data(), c->passive.empty() ? nullptr : c->passive.data(),
                                &l, c->ws, c->ws_bytes, nullptr);
  if (hipDeviceSynchronize() != hipSuccess) return KPP_E_HIP;
  if (r <= -100) return r;
  if ((IERR_U && hipMemcpy(IERR_U, c->ierr, sizeof(int) * n, hipMemcpyDeviceToHost) != hipSuccess) ||
      (ISTATUS_U && hipMemcpy(ISTATUS_U, c->istat, sizeof(int) * 8 * (size_t)n,
                              hipMemcpyDeviceToHost) != hipSuccess) ||
      (RSTATUS_U && hipMemcpy(RSTATUS_U, c->rstat, sizeof(double) * 3 * (size_t)n,
                              hipMemcpyDeviceToHost) != hipSuccess))
    return KPP_E_HIP;
  return r;
}

}  // extern "C"

// ---- graph capture ----
struct kpp_graph {
  hipGraph_t graph;
  hipGraphExec_t exec;
};

extern "C" int kpp_capture_begin(void* stream) {
  if (!stream) return KPP_E_ARG;  // the legacy default stream cannot be captured
  return hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal) == hipSuccess
             ? 0
             : KPP_E_HIP;
}

extern "C" int kpp_capture_end(void* stream, kpp_graph** out) {
  if (!stream || !out) return KPP_E_ARG;
  *out = nullptr;
  hipGraph_t g = nullptr;
  if (hipStreamEndCapture((hipStream_t)stream, &g) != hipSuccess || !g) return KPP_E_HIP;
  hipGraphExec_t e = nullptr;
  if (hipGraphInstantiate(&e, g, nullptr, nullptr, 0) != hipSuccess) {
    hipGraphDestroy(g);
    return KPP_E_HIP;
  }
  *out = new kpp_graph{g, e};
  return 0;
}

extern "C" int kpp_graph_launch(kpp_graph* g, void* stream) {
  if (!g) return KPP_E_ARG;
  return hipGraphLaunch(g->exec, (hipStream_t)stream) == hipSuccess ? 0 : KPP_E_HIP;
}

extern "C" void kpp_graph_destroy(kpp_graph* g) {
  if (!g) return;
  hipGraphExecDestroy(g->exec);
  hipGraphDestroy(g->graph);
  delete g;
}

// ---- autotuner ----
// Candidate launch shapes of the persistent kernel: cells per wavefront
// (lanes), refill threshold, grid size relative to the resident capacity,
// and the LU entries per lane mirrored in LDS (candidates with a mirror the
// mechanism is not built with are skipped).
namespace {
struct Cand {
  int lanes, refill_min, grid_pct, lds;
};
const Cand kCands[] = {{64, 1, 100, 0},  {64, 8, 100, 0},  {64, 1, 75, 0},
                       {64, 1, 125, 0},  {32, 1, 100, 0},  {32, 4, 100, 0},
                       {48, 1, 100, 0},  {64, 16, 100, 0}, {64, 1, 100, KPP_LDS_SPLIT_A},
                       {64, 1, 100, KPP_LDS_SPLIT_B}, {64, 1, 125, KPP_LDS_SPLIT_B}};
constexpr int kNCand = sizeof(kCands) / sizeof(kCands[0]);

bool cand_launch(const MechOps* m, int S, int ncells, const Cand& c, kpp_launch& l) {
  l.variant = KPP_VARIANT_PERSISTENT;
  l.lu_fp32 = 0;
  l.lanes = c.lanes;
  l.refill_min = c.refill_min;
  l.nblocks = 0;
  l.lds_entries = c.lds;
  if (c.lds > m->lds_max) return false;
  Geometry g;
  if (!geometry(m, S, &l, ncells, g)) return false;
  l.nblocks = std::max(1, (int)((long)g.nblocks * c.grid_pct / 100));
  return true;
}
}  // namespace

extern "C" size_t kpp_autotune_workspace_bytes(const char* mech, int method, int ncells) {
  const MechOps* m = find_mech(mech);
  RosCoef c;
  if (!m || !fill_method(method, c) || ncells <= 0) return 0;
  size_t best = 0;
  for (int i = 0; i < kNCand; i++) {
    kpp_launch l;
    if (!cand_launch(m, c.S, ncells, kCands[i], l)) continue;
    best = std::max(best, kpp_workspace_bytes(mech, method, &l, ncells));
  }
  return best;
}

extern "C" int kpp_autotune(const char* mech, const kpp_batch* b, const double* C0,
                            const int IPAR[20], const double RPAR[20], const double* atol,
                            const double* rtol, void* workspace, size_t workspace_bytes,
                            void* stream, kpp_launch* best, float* best_ms, int* ntried) {
  const MechOps* m = find_mech(mech);
  if (!m) return KPP_E_MECH;
  if (!b || !C0 || !best || !IPAR || b->ncells <= 0) return KPP_E_ARG;
  RosCoef coef;
  if (!fill_method(IPAR[3], coef)) return -2;
  hipStream_t st = (hipStream_t)stream;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return KPP_E_HIP;
  if (hipEventCreate(&e1) != hipSuccess) {
    hipEventDestroy(e0);
    return KPP_E_HIP;
  }
  const size_t cbytes = sizeof(double) * (size_t)m->nspec * b->ld;
  float best_t = 0.f;
  int tried = 0, rc = 0;
  for (int i = 0; i < kNCand && rc == 0; i++) {
    kpp_launch l;
    if (!cand_launch(m, coef.S, b->ncells, kCands[i], l)) continue;
    if (kpp_workspace_bytes(mech, IPAR[3], &l, b->ncells) > workspace_bytes) continue;
    if (hipMemcpyAsync(b->C, C0, cbytes, hipMemcpyDeviceToDevice, st) != hipSuccess ||
        hipEventRecord(e0, st) != hipSuccess) {
      rc = KPP_E_HIP;
      break;
    }
    const int r = kpp_rosenbrock_dev(mech, b, IPAR, RPAR, atol, rtol, &l, workspace,
                                     workspace_bytes, stream);
    if (r != 0) {
      rc = r;
      break;
    }
    float ms = 0.f;
    if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
      rc = KPP_E_HIP;
      break;
    }
    tried++;
    if (tried == 1 || ms < best_t) {
      best_t = ms;
      *best = l;
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc == 0 && hipMemcpyAsync(b->C, C0, cbytes, hipMemcpyDeviceToDevice, st) != hipSuccess)
    rc = KPP_E_HIP;
  if (best_ms) *best_ms = best_t;
  if (ntried) *ntried = tried;
  if (rc == 0 && tried == 0) rc = KPP_E_WORKSPACE;
  return rc;
}


// ---- workspace-free entry point ----
// `scratch` is the device memory kpp_rosenbrock_dev_static uses instead of a
// caller's workspace: one array per GPU, held until the process ends.  Host
// code never passes the symbol itself to a kernel (on the host it only
// names the array): each device's address comes from hipGetSymbolAddress,
// taken once per device with that device current.
static __device__ __align__(256) char scratch[KPP_STATIC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices];

int scratch_of_current_device(void** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return KPP_E_HIP;
  if (dev < 0 || dev >= kMaxDevices) return KPP_E_NODEV;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return KPP_E_HIP;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}
}  // namespace

extern "C" int kpp_rosenbrock_dev_static(const char* mech, const kpp_batch* b, const int IPAR[20],
                                         const double RPAR[20], const double* atol,
                                         const double* rtol, const kpp_launch* l, void* stream) {
  if (!find_mech(mech)) return KPP_E_MECH;
  void* ws = nullptr;
  const int e = scratch_of_current_device(&ws);
  if (e != 0) return e;
  return rosenbrock_impl(mech, b, IPAR, RPAR, atol, rtol, l, ws, KPP_STATIC_SCRATCH_BYTES, stream,
                         true);
}
