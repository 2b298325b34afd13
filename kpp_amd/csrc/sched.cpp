// Learned step scheduler loaded from a weights file, for C hosts (no
// counterpart in KPP; declarations in include/kpphip.h).
//
// The weights file is the flat little-endian image kpp_amd/learned.py
// (StepModel.save_flat) writes next to the trained .npz:
//   char magic[8] = "KPPWGT1\0"; char mech[32];
//   int32 nin, in_pad, nparam, has_centres;
//   float64 cfloor, mu_y, sd_y, h_scale;
//   float32 centres[8]; float32 mu[in_pad], sd[in_pad], w[nparam].
// kpp_sched_run computes the network's features of a device batch, runs
// it, and writes each cell's starting step (the predicted first accepted
// step times h_scale, clamped) and the launch order (most expected steps
// first), exactly what the Python host does with kpp_mlp_features and
// kpp_mlp_schedule.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "kpphip.h"

struct kpp_sched {
  int nin, in_pad, nspec;
  size_t nparam;
  double cfloor, mu_y, sd_y, h_scale;
  int has_centres;
  float centres[KPP_MLP_BUCKETS];
  float *w = nullptr, *mu = nullptr, *sd = nullptr;
  void* wb = nullptr;
  void* X = nullptr;        // BF16 features [cap][in_pad]
  void* work = nullptr;     // kpp_mlp_schedule work
  size_t x_cap = 0, work_cap = 0;
};

static void sched_free(kpp_sched* s) {
  if (!s) return;
  void* ptrs[] = {s->w, s->mu, s->sd, s->wb, s->X, s->work};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete s;
}

extern "C" {

int kpp_sched_load(const char* path, const char* mech, kpp_sched** out) {
  if (!path || !mech || !out) return KPP_E_ARG;
  *out = nullptr;
  int info[10];
  if (kpp_mech_info(mech, info) != 0) return KPP_E_MECH;
  FILE* f = fopen(path, "rb");
  if (!f) return KPP_E_WEIGHTS;
  char magic[8], name[32];
  int32_t hdr[4];
  double dv[4];
  float cen[KPP_MLP_BUCKETS];
  bool ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, "KPPWGT1", 8) == 0 &&
            fread(name, 1, 32, f) == 32 && fread(hdr, sizeof(int32_t), 4, f) == 4 &&
            fread(dv, sizeof(double), 4, f) == 4 &&
            fread(cen, sizeof(float), KPP_MLP_BUCKETS, f) == (size_t)KPP_MLP_BUCKETS;
  name[31] = 0;
  const int nspec = info[2];
  ok = ok && strcmp(name, mech) == 0 && hdr[0] == nspec + 2 && hdr[1] >= hdr[0] &&
       hdr[1] % 32 == 0 && (size_t)hdr[2] == kpp_mlp_param_count(hdr[1]);
  std::vector<float> mu, sd, w;
  if (ok) {
    mu.resize(hdr[1]);
    sd.resize(hdr[1]);
    w.resize(hdr[2]);
    ok = fread(mu.data(), sizeof(float), mu.size(), f) == mu.size() &&
         fread(sd.data(), sizeof(float), sd.size(), f) == sd.size() &&
         fread(w.data(), sizeof(float), w.size(), f) == w.size();
  }
  fclose(f);
  if (!ok) return KPP_E_WEIGHTS;
  kpp_sched* s = new kpp_sched();
  s->nin = hdr[0];
  s->in_pad = hdr[1];
  s->nparam = (size_t)hdr[2];
  s->has_centres = hdr[3];
  s->nspec = nspec;
  s->cfloor = dv[0];
  s->mu_y = dv[1];
  s->sd_y = dv[2];
  s->h_scale = dv[3];
  memcpy(s->centres, cen, sizeof cen);
  const size_t pb = sizeof(float) * s->nparam, ib = sizeof(float) * s->in_pad;
  ok = hipMalloc((void**)&s->w, pb) == hipSuccess && hipMalloc(&s->wb, pb / 2) == hipSuccess &&
       hipMalloc((void**)&s->mu, ib) == hipSuccess && hipMalloc((void**)&s->sd, ib) == hipSuccess &&
       hipMemcpy(s->w, w.data(), pb, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(s->mu, mu.data(), ib, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(s->sd, sd.data(), ib, hipMemcpyHostToDevice) == hipSuccess;
  if (ok) {
    kpp_mlp net{s->nin, s->in_pad, s->w, s->wb};
    ok = kpp_mlp_sync_bf16(&net, nullptr) == 0 && hipDeviceSynchronize() == hipSuccess;
  }
  if (!ok) {
    sched_free(s);
    return KPP_E_HIP;
  }
  *out = s;
  return 0;
}

void kpp_sched_destroy(kpp_sched* s) { sched_free(s); }

int kpp_sched_run(kpp_sched* s, const kpp_batch* b, double hmin, double hmax, double* hstart,
                  int* order, void* stream) {
  if (!s || !b || b->ncells <= 0 || !b->C || !b->temp || !(hmin > 0) || !(hmax >= hmin))
    return KPP_E_ARG;
  const int n = b->ncells;
  const size_t xb = (size_t)n * s->in_pad * 2, wk = kpp_mlp_schedule_work_bytes(n);
  if (xb > s->x_cap) {
    if (s->X) (void)hipFree(s->X);
    s->X = nullptr;
    s->x_cap = 0;
    if (hipMalloc(&s->X, xb) != hipSuccess) return KPP_E_HIP;
    s->x_cap = xb;
  }
  if (wk > s->work_cap) {
    if (s->work) (void)hipFree(s->work);
    s->work = nullptr;
    s->work_cap = 0;
    if (hipMalloc(&s->work, wk) != hipSuccess) return KPP_E_HIP;
    s->work_cap = wk;
  }
  int r = kpp_mlp_features(n, b->ld, s->nspec, b->C, b->temp,
                           b->sun_mode == KPP_SUN_FIXED ? b->sun : nullptr, s->mu, s->sd,
                           s->cfloor, s->in_pad, s->X, stream);
  if (r != 0) return r;
  kpp_mlp net{s->nin, s->in_pad, s->w, s->wb};
  // starting step = h_scale x 10^(o0*sd_y + mu_y)
  const float mu = (float)(s->mu_y + log10(s->h_scale));
  return kpp_mlp_schedule(&net, s->X, n, mu, (float)s->sd_y, hmin, hmax,
                          s->has_centres ? s->centres : nullptr, hstart, nullptr, order, s->work,
                          s->work_cap, stream);
}

}  // extern "C"
