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
  extern "C" kpphip::ComponentFn kpphip_mech_comp_##N();
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

int device_cus() {
  int dev = 0, n = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  return n;
}

constexpr size_t kHead = 256;  // queue counter
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
    g.tol_bytes = align256(2 * (size_t)m->nvar * sizeof(double));
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
  g.tol_bytes = align256(2 * (size_t)m->nvar * sizeof(double));
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

__global__ void k_fill_int(int* p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
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
                           void* stream, bool fixed) {
  const MechOps* m = find_mech(mech);
  if (!m) return KPP_E_MECH;
  if (!b || !IPAR || !RPAR || !atol || !rtol || b->ncells < 0 || b->ld < b->ncells)
    return KPP_E_ARG;
  if (b->ncells > 0 && (!b->C || !b->temp || !b->cfactor || !b->ierr)) return KPP_E_ARG;
  if (b->sun_mode == KPP_SUN_FIXED && b->ncells > 0 && !b->sun && !b->rconst) return KPP_E_ARG;
  if (b->sun_mode != KPP_SUN_FIXED && b->sun_mode != KPP_SUN_TOD) return KPP_E_ARG;
  hipStream_t st = (hipStream_t)stream;
  const int e = check_params(IPAR, RPAR, atol, rtol, m->nvar);
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
  if (!fill_method(IPAR[3], p.coef)) return -2;
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
  p.vector_tol = (IPAR[1] == 0);
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
  p.autonomous = !(IPAR[0] == 0);
  p.max_steps = (IPAR[2] == 0) ? 100000 : IPAR[2];
  p.Roundoff = unit_roundoff();
  p.Hmin = RPAR[0];
  p.Hmax_user = RPAR[1];
  p.Hstart_user = RPAR[2];
  p.FacMin = (RPAR[3] == 0.0) ? 0.2 : RPAR[3];
  p.FacMax = (RPAR[4] == 0.0) ? 6.0 : RPAR[4];
  p.FacRej = (RPAR[5] == 0.0) ? 0.1 : RPAR[5];
  p.FacSafe = (RPAR[6] == 0.0) ? 0.9 : RPAR[6];
  p.hstart = b->hstart;
  p.hstart_auto = (IPAR[4] == 1);
  p.rconst = b->rconst;
  p.rates_given = b->rconst != nullptr;
  // dF/dT is exactly zero when no rate constant depends on T
  p.dfdt_zero = p.rates_given || (b->sun_mode == KPP_SUN_FIXED && !m->rates_use_time);
  for (int i = 0; i < 8; i++) p.stat0[i] = IPAR[10 + i];
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
            hipMalloc(&c->rstat, sizeof(double) * 2 * n) == hipSuccess &&
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
