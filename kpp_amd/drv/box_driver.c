/* Box model over many cells: the batched counterpart of KPP's C driver
 * (drv/general.c:46-90), written against the C ABI of include/kpphip.h.
 *
 * general.c initialises one cell (Initialize), then calls INTEGRATE(TIME,
 * TIME+DT) for every output interval, with STEPMIN carried from one call to
 * the next.  Here every cell starts from Initialize()'s concentrations, each
 * variable species scaled by a per-cell factor 10^u, u uniform in [-1, 1)
 * (a fixed splitmix64 stream), and INTEGRATE_BATCH advances all cells over
 * each interval at once.
 *
 * usage: box_driver MECH NCELLS METHOD TSTART TEND DT [OUT [FLAGS [TEMP CFACTOR]]]
 *   METHOD is Rosenbrock's IPAR[3] code; SUN follows the time of day
 *   (Update_SUN).  Prints one summary line; writes the final C[NSPEC][NCELLS],
 *   IERR[NCELLS] and the starting steps of the first interval HSTART[NCELLS]
 *   to OUT (raw doubles, int32, doubles) when given ("-": no file).
 *   FLAGS: "plain" (default) or a comma list of persistent, autotune,
 *   learned, mixed, hstart_auto (the KPP_CTX_* context flags; "tuned" =
 *   autotune,learned).  TEMP (default 270 K) and CFACTOR (default 1) are
 *   every cell's.  Exit status 0 when every cell returned IERR = 1.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "kpphip.h"

static int parse_flags(const char* s) {
  int f = 0;
  char buf[256];
  strncpy(buf, s, sizeof buf - 1);
  buf[sizeof buf - 1] = 0;
  for (char* t = strtok(buf, ","); t; t = strtok(NULL, ",")) {
    if (!strcmp(t, "plain")) continue;
    else if (!strcmp(t, "persistent")) f |= KPP_CTX_PERSISTENT;
    else if (!strcmp(t, "autotune")) f |= KPP_CTX_AUTOTUNE;
    else if (!strcmp(t, "learned")) f |= KPP_CTX_LEARNED;
    else if (!strcmp(t, "mixed")) f |= KPP_CTX_MIXED;
    else if (!strcmp(t, "hstart_auto")) f |= KPP_CTX_HSTART_AUTO;
    else if (!strcmp(t, "tuned")) f |= KPP_CTX_AUTOTUNE | KPP_CTX_LEARNED;
    else return -1;
  }
  return f;
}

static uint64_t splitmix64(uint64_t* s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

int main(int argc, char** argv) {
  if (argc < 7) {
    fprintf(stderr, "usage: %s MECH NCELLS METHOD TSTART TEND DT [OUT [FLAGS [TEMP CFACTOR]]]\n",
            argv[0]);
    return 2;
  }
  const int flags = argc > 8 ? parse_flags(argv[8]) : 0;
  const double T0 = argc > 10 ? atof(argv[9]) : 270.0, CF0 = argc > 10 ? atof(argv[10]) : 1.0;
  if (flags < 0) {
    fprintf(stderr, "bad FLAGS\n");
    return 2;
  }
  const char* mech = argv[1];
  const int ncells = atoi(argv[2]), method = atoi(argv[3]);
  const double tstart = atof(argv[4]), tend = atof(argv[5]), dt = atof(argv[6]);
  int info[10];
  if (kpp_mech_info(mech, info) != 0 || ncells <= 0 || !(dt > 0)) {
    fprintf(stderr, "bad mechanism or arguments\n");
    return 2;
  }
  const int nvar = info[0], nspec = info[2];
  double* C = malloc(sizeof(double) * (size_t)nspec * ncells);
  double* init = malloc(sizeof(double) * nspec);
  double* temp = malloc(sizeof(double) * ncells);
  double* cfac = malloc(sizeof(double) * ncells);
  int* ierr = malloc(sizeof(int) * ncells);
  double* h0 = malloc(sizeof(double) * ncells);
  if (!C || !init || !temp || !cfac || !ierr || !h0) return 2;
  kpp_mech_initial(mech, init);
  uint64_t st = 12345;
  for (int c = 0; c < ncells; c++) {
    for (int s = 0; s < nspec; s++) {
      double f = 1.0;
      if (s < nvar) {
        const double u = (double)(splitmix64(&st) >> 11) * (1.0 / 9007199254740992.0);
        f = pow(10.0, 2.0 * u - 1.0);
      }
      C[(size_t)s * ncells + c] = init[s] * f;
    }
    temp[c] = T0;
    cfac[c] = CF0;
  }
  kpp_ctx* ctx = NULL;
  int r = kpp_ctx_create(&ctx, mech, ncells, method, KPP_SUN_TOD | flags);
  if (r == 0) r = kpp_set_state(ctx, C, temp, cfac, NULL);
  int calls = 0;
  for (double t = tstart; r == 0 && t < tend; t += dt, calls++) {
    r = INTEGRATE_BATCH(ctx, t, t + dt);
    if (r == 0 && calls == 0) r = kpp_ctx_get_hstart(ctx, h0);
  }
  if (r == 0) r = kpp_get_state(ctx, C, ierr, NULL, NULL);
  kpp_ctx_destroy(ctx);
  if (r != 0) {
    fprintf(stderr, "kpphip call failed with %d\n", r);
    return 1;
  }
  int ok = 0;
  double sum = 0.0;
  for (int c = 0; c < ncells; c++) ok += (ierr[c] == 1);
  for (size_t i = 0; i < (size_t)nvar * ncells; i++) sum += C[i];
  printf("box_driver %s cells=%d intervals=%d ierr_ok=%d sum_var=%.17g\n", mech, ncells, calls, ok,
         sum);
  if (argc > 7 && strcmp(argv[7], "-") != 0) {
    FILE* f = fopen(argv[7], "wb");
    if (!f || fwrite(C, sizeof(double), (size_t)nspec * ncells, f) != (size_t)nspec * ncells ||
        fwrite(ierr, sizeof(int), ncells, f) != (size_t)ncells ||
        fwrite(h0, sizeof(double), ncells, f) != (size_t)ncells || fclose(f) != 0)
      return 2;
  }
  free(C); free(init); free(temp); free(cfac); free(ierr); free(h0);
  return ok == ncells ? 0 : 1;
}
