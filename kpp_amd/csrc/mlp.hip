// Learned step scheduler: a 4-layer MLP (inputs -> 128 -> 128 -> 128 -> 16,
// GELU) that predicts each cell's first step size and step-count bucket from
// its initial state.  Weights and activations are BF16, every product
// accumulates in FP32 on the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16);
// the optimiser keeps FP32 master weights.  Training and inference both run
// through the kernels below.  The network has no counterpart in KPP: it
// only chooses the integrator's starting step (Rosenbrock's RPAR[2] per cell)
// and the order in which cells are launched; the Rosenbrock error control
// still accepts or rejects every step.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "kpphip.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int HID = KPP_MLP_HIDDEN;
constexpr int NOUT = KPP_MLP_OUT;
constexpr int NBUCKET = KPP_MLP_BUCKETS;

enum { EPI_STORE = 0, EPI_BIAS = 1, EPI_BIAS_GELU = 2, EPI_GELU_BACK = 3 };

// GELU, tanh form: 0.5 z (1 + tanh(c (z + 0.044715 z^3))), c = sqrt(2/pi)
__device__ __forceinline__ float gelu_f(float z) {
  const float c = 0.7978845608028654f;
  const float t = tanhf(c * (z + 0.044715f * z * z * z));
  return 0.5f * z * (1.0f + t);
}
__device__ __forceinline__ float gelu_grad(float z) {
  const float c = 0.7978845608028654f;
  const float t = tanhf(c * (z + 0.044715f * z * z * z));
  return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * c * (1.0f + 3.0f * 0.044715f * z * z);
}

// D[M][N] (+)= op(A)[M][K] * op(B)[K][N] on the matrix cores.
//  A: AT=false -> element (m,k) at A[m*lda + k]; AT=true -> A[k*lda + m]
//  B: BT=false -> element (k,n) at B[n*ldb + k] (a weight matrix W[N][K]);
//     BT=true  -> B[k*ldb + n]
// One wavefront computes 16 rows x N columns (N <= 128, a multiple of 16);
// four wavefronts per block.  blockIdx.y splits K into chunks of kchunk
// (a multiple of 32); with more than one chunk the partial sums are added
// atomically into Cf, which the caller has cleared.
// Fragment layout of v_mfma_f32_16x16x32_bf16 (lane l, i = 0..7 / 0..3):
//  A operand: row l%16, k = 8*(l/16) + i;  B operand: column l%16,
//  k = 8*(l/16) + i;  accumulator: row 4*(l/16) + i, column l%16.
template <bool AT, bool BT, int EPI>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const __bf16* __restrict__ A,
                                               int lda, const __bf16* __restrict__ B, int ldb,
                                               const float* __restrict__ bias, float* Cf, int ldc,
                                               __bf16* __restrict__ Hb, int ldh,
                                               const float* __restrict__ Zin, int kchunk) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int m0 = (blockIdx.x * 4 + wave) * 16;
  if (m0 >= M) return;
  const int kbeg = blockIdx.y * kchunk;
  const int kend = min(K, kbeg + kchunk);
  const int ntiles = N / 16;
  const int r = lane & 15;
  const int kq = 8 * (lane >> 4);
  floatx4 acc[HID / 16];
#pragma unroll
  for (int t = 0; t < HID / 16; t++) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
  const int am = m0 + r;
  for (int k0 = kbeg; k0 < kend; k0 += 32) {
    const int kk = k0 + kq;
    bf16x8 a;
    if (!AT && am < M && kk + 8 <= kend) {
      a = *reinterpret_cast<const bf16x8*>(A + (size_t)am * lda + kk);
    } else {
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const bool ok = am < M && kk + i < kend;
        a[i] = ok ? (AT ? A[(size_t)(kk + i) * lda + am] : A[(size_t)am * lda + kk + i])
                  : (__bf16)0.0f;
      }
    }
#pragma unroll
    for (int t = 0; t < HID / 16; t++) {
      if (t < ntiles) {
        const int bn = t * 16 + r;
        bf16x8 b;
        if (!BT && kk + 8 <= kend) {
          b = *reinterpret_cast<const bf16x8*>(B + (size_t)bn * ldb + kk);
        } else {
#pragma unroll
          for (int i = 0; i < 8; i++) {
            const bool ok = kk + i < kend;
            b[i] = ok ? (BT ? B[(size_t)(kk + i) * ldb + bn] : B[(size_t)bn * ldb + kk + i])
                      : (__bf16)0.0f;
          }
        }
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
      }
    }
  }
  const int rq = 4 * (lane >> 4);
#pragma unroll
  for (int t = 0; t < HID / 16; t++) {
    if (t >= ntiles) continue;
    const int n = t * 16 + r;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int m = m0 + rq + i;
      if (m >= M) continue;
      const float v = acc[t][i];
      const size_t o = (size_t)m * ldc + n;
      if (EPI == EPI_STORE) {
        if (gridDim.y > 1) atomicAdd(Cf + o, v);
        else Cf[o] = v;
      } else if (EPI == EPI_BIAS) {
        Cf[o] = v + bias[n];
      } else if (EPI == EPI_BIAS_GELU) {
        const float z = v + bias[n];
        if (Cf) Cf[o] = z;
        Hb[(size_t)m * ldh + n] = (__bf16)gelu_f(z);
      } else {  // EPI_GELU_BACK: v = dL/dH, Zin = pre-activation
        const float dz = v * gelu_grad(Zin[o]);
        if (Cf) Cf[o] = dz;
        Hb[(size_t)m * ldh + n] = (__bf16)dz;
      }
    }
  }
}

__global__ void k_fill(float* p, size_t n, float v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// db[n] += sum over rows m of D[m][n]  (bias gradients)
__global__ __launch_bounds__(256) void k_colsum(int M, int N, const float* __restrict__ D,
                                                 float* db) {
  const int n = threadIdx.x % N;
  const int rows_per_pass = blockDim.x / N;
  const int rsub = threadIdx.x / N;
  if (rsub >= rows_per_pass) return;
  float s = 0.f;
  for (int m = blockIdx.x * rows_per_pass + rsub; m < M; m += gridDim.x * rows_per_pass)
    s += D[(size_t)m * N + n];
  atomicAdd(db + n, s);
}

// Loss per row: w*(o0 - y)^2 + cross-entropy of the bucket logits o[1..8];
// writes dL/dO (mean over rows) in FP32 and BF16 and adds the summed loss.
__global__ __launch_bounds__(256) void k_loss(int M, const float* __restrict__ O,
                                               const float* __restrict__ y,
                                               const int* __restrict__ bucket,
                                               const float* __restrict__ wreg, float* dO,
                                               __bf16* dOb, float* loss_sum) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  float L = 0.f;
  if (m < M) {
    const float* o = O + (size_t)m * NOUT;
    const float inv = 1.0f / (float)M;
    const float w = wreg[m];
    const float e = o[0] - y[m];
    float mx = o[1];
    for (int j = 2; j <= NBUCKET; j++) mx = fmaxf(mx, o[j]);
    float se = 0.f;
    for (int j = 1; j <= NBUCKET; j++) se += expf(o[j] - mx);
    const int c = bucket[m];
    L = w * e * e + (logf(se) + mx - o[1 + c]);
    float g[NOUT];
    g[0] = 2.0f * w * e * inv;
    for (int j = 1; j <= NBUCKET; j++)
      g[j] = (expf(o[j] - mx) / se - ((j - 1) == c ? 1.0f : 0.0f)) * inv;
    for (int j = NBUCKET + 1; j < NOUT; j++) g[j] = 0.f;
    for (int j = 0; j < NOUT; j++) {
      dO[(size_t)m * NOUT + j] = g[j];
      dOb[(size_t)m * NOUT + j] = (__bf16)g[j];
    }
  }
  // block reduction of the loss
  __shared__ float red[256];
  red[threadIdx.x] = L;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss_sum, red[0]);
}

// Adam (bias-corrected) on the FP32 master weights; refreshes the BF16 copy.
__global__ void k_adam(size_t n, float* w, __bf16* wb, const float* __restrict__ g, float* m1,
                       float* m2, float lr, float b1, float b2, float eps, float c1, float c2) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.0f - b1) * gi;
  const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  const float wi = w[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
  w[i] = wi;
  wb[i] = (__bf16)wi;
}

__global__ void k_to_bf16(size_t n, const float* w, __bf16* wb) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) wb[i] = (__bf16)w[i];
}

// Features of one cell: standardised log10 of every species concentration
// (floored), TEMP and SUN; zero padding up to in_pad.
__global__ __launch_bounds__(256) void k_features(int n, int ld, int nspec,
                                                   const double* __restrict__ C,
                                                   const double* __restrict__ temp,
                                                   const double* __restrict__ sun,
                                                   const float* __restrict__ mu,
                                                   const float* __restrict__ sd, double cfloor,
                                                   int in_pad, __bf16* X) {
  // one block = 256 consecutive cells = one contiguous stretch of X; each
  // thread writes its row into LDS, then the block copies it out in 16-byte
  // pieces (a thread's own row would be a 2*in_pad-byte strided write)
  extern __shared__ __align__(16) unsigned char lds[];
  __bf16* tile = reinterpret_cast<__bf16*>(lds);
  const int c0 = blockIdx.x * 256;
  const int cell = c0 + threadIdx.x;
  if (cell < n) {
    __bf16* x = tile + (size_t)threadIdx.x * in_pad;
    for (int j = 0; j < in_pad; j++) {
      float v = 0.f;
      if (j < nspec) {
        const double c = C[(size_t)j * ld + cell];
        v = (float)log10(c > cfloor ? c : cfloor);
      } else if (j == nspec) {
        v = (float)temp[cell];
      } else if (j == nspec + 1) {
        v = sun ? (float)sun[cell] : 0.f;
      }
      x[j] = (j < nspec + 2) ? (__bf16)((v - mu[j]) / sd[j]) : (__bf16)0.0f;
    }
  }
  __syncthreads();
  const int rows = min(256, n - c0);
  const int chunks = rows * in_pad / 8;  // in_pad is a multiple of 32
  uint4* dst = reinterpret_cast<uint4*>(X + (size_t)c0 * in_pad);
  const uint4* src = reinterpret_cast<const uint4*>(tile);
  for (int i = threadIdx.x; i < chunks; i += 256) dst[i] = src[i];
}

// Prediction: first step h = 10^(o0*sd_y + mu_y) clamped to [hmin, hmax];
// bucket = argmax of the bucket logits.  The launch-order key is the
// expected step count under the logits' softmax -- sum_j p_j * centre_j, with
// centre_j the mean step count of bucket j in the training data -- mapped
// linearly from [centre_0, centre_last] onto NKEY levels and reversed, so
// cells predicted to need the most steps are launched first
// (longest-processing-time order for the persistent kernel's work queue).
// Without centres the expected bucket index is used.  Expected counts rank
// cells better than expected indices because the buckets are quantiles, not
// equal widths in step count.
constexpr int NKEY = 64;
struct Centres {
  float c[NBUCKET];
  int use;
};
__global__ void k_decode(int n, const float* __restrict__ O, float mu_y, float sd_y, double hmin,
                         double hmax, Centres cen, double* hstart, int* bucket, uint8_t* key) {
  const int cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= n) return;
  const float* o = O + (size_t)cell * NOUT;
  if (hstart) {
    double h = pow(10.0, (double)(o[0] * sd_y + mu_y));
    h = h < hmin ? hmin : (h > hmax ? hmax : h);
    hstart[cell] = h;
  }
  int best = 0;
  float mx = o[1];
  for (int j = 1; j < NBUCKET; j++)
    if (o[1 + j] > o[1 + best]) best = j;
  mx = o[1 + best];
  float se = 0.f, ex = 0.f;
  for (int j = 0; j < NBUCKET; j++) {
    const float e = expf(o[1 + j] - mx);
    se += e;
    ex += e * (cen.use ? cen.c[j] : (float)j);
  }
  if (bucket) bucket[cell] = best;
  if (key) {
    // expected count (or bucket index), position in [lo, hi] -> [0, NKEY-1]
    const float lo = cen.use ? cen.c[0] : 0.f;
    const float hi = cen.use ? cen.c[NBUCKET - 1] : (float)(NBUCKET - 1);
    const float eb = ex / se;
    int q = (int)((eb - lo) * (float)(NKEY - 1) / (hi - lo) + 0.5f);
    q = q < 0 ? 0 : (q > NKEY - 1 ? NKEY - 1 : q);
    key[cell] = (uint8_t)(NKEY - 1 - q);
  }
}

// Stable counting sort of cell indices by an NKEY-valued key, in three kernels
// with no memory clears (every count is written, none accumulated), so the
// sequence can be recorded in a graph.  Block b owns cells
// [b*SORT_CHUNK, (b+1)*SORT_CHUNK).
constexpr int SORT_THREADS = 256;
constexpr int SORT_ITEMS = 8;
constexpr int SORT_CHUNK = SORT_THREADS * SORT_ITEMS;

// rank of this thread's key among equal keys of lower threads in the block,
// and the per-key totals of the block (LDS)
__device__ __forceinline__ int block_rank(int key, bool valid, int* wave_cnt /*[4][NKEY]*/,
                                          int* tot /*[NKEY]*/) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  int rank = 0;
#pragma unroll
  for (int k = 0; k < NKEY; k++) {
    const uint64_t m = __ballot(valid && key == k);
    if (valid && key == k) rank = __popcll(m & below);
    if (lane == 0) wave_cnt[wave * NKEY + k] = __popcll(m);
  }
  __syncthreads();
  if (threadIdx.x < NKEY) {
    int t = 0;
    for (int w = 0; w < SORT_THREADS / 64; w++) t += wave_cnt[w * NKEY + threadIdx.x];
    tot[threadIdx.x] = t;
  }
  int before = 0;
  if (valid)
    for (int w = 0; w < wave; w++) before += wave_cnt[w * NKEY + key];
  __syncthreads();
  return rank + before;
}

__global__ __launch_bounds__(SORT_THREADS) void k_sort_count(int n, const uint8_t* __restrict__ key,
                                                             int nblk, int* counts) {
  __shared__ int wave_cnt[4 * NKEY], tot[NKEY];
  int acc = 0;  // thread k < NKEY accumulates key k
  const int base = blockIdx.x * SORT_CHUNK;
  for (int it = 0; it < SORT_ITEMS; it++) {
    const int c = base + it * SORT_THREADS + threadIdx.x;
    const bool valid = c < n;
    block_rank(valid ? key[c] : 0, valid, wave_cnt, tot);
    if (threadIdx.x < NKEY) acc += tot[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x < NKEY) counts[threadIdx.x * nblk + blockIdx.x] = acc;
}

// exclusive scan of counts[key][block] in key-major order (one block)
__global__ __launch_bounds__(1024) void k_sort_scan(int total, int* counts) {
  __shared__ int part[1024];
  const int per = (total + blockDim.x - 1) / blockDim.x;
  const int b = threadIdx.x * per;
  int s = 0;
  for (int i = b; i < min(total, b + per); i++) s += counts[i];
  part[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int t = 0; t < (int)blockDim.x; t++) {
      const int v = part[t];
      part[t] = run;
      run += v;
    }
  }
  __syncthreads();
  int run = part[threadIdx.x];
  for (int i = b; i < min(total, b + per); i++) {
    const int v = counts[i];
    counts[i] = run;
    run += v;
  }
}

__global__ __launch_bounds__(SORT_THREADS) void k_sort_scatter(int n, const uint8_t* __restrict__ key,
                                                               int nblk, const int* __restrict__ offs,
                                                               int* order) {
  __shared__ int wave_cnt[4 * NKEY], tot[NKEY], run[NKEY];
  if (threadIdx.x < NKEY) run[threadIdx.x] = offs[threadIdx.x * nblk + blockIdx.x];
  __syncthreads();
  const int base = blockIdx.x * SORT_CHUNK;
  for (int it = 0; it < SORT_ITEMS; it++) {
    const int c = base + it * SORT_THREADS + threadIdx.x;
    const bool valid = c < n;
    const int k = valid ? key[c] : 0;
    const int r = block_rank(k, valid, wave_cnt, tot);
    if (valid) order[run[k] + r] = c;
    __syncthreads();
    if (threadIdx.x < NKEY) run[threadIdx.x] += tot[threadIdx.x];
    __syncthreads();
  }
}

hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

struct Layout {
  size_t w[4], b[4], total;
};
Layout layout(int in_pad) {
  Layout L;
  const int kin[4] = {in_pad, HID, HID, HID};
  const int nout[4] = {HID, HID, HID, NOUT};
  size_t o = 0;
  for (int l = 0; l < 4; l++) {
    L.w[l] = o;
    o += (size_t)nout[l] * kin[l];
    L.b[l] = o;
    o += nout[l];
  }
  L.total = o;
  return L;
}

bool net_ok(const kpp_mlp* net) {
  return net && net->w && net->wb && net->in_pad > 0 && net->in_pad % 32 == 0 &&
         net->nin > 0 && net->nin <= net->in_pad;
}

template <bool AT, bool BT, int EPI>
void gemm(hipStream_t st, int M, int N, int K, const __bf16* A, int lda, const __bf16* B, int ldb,
          const float* bias, float* Cf, int ldc, __bf16* Hb, int ldh, const float* Zin,
          int splitk) {
  const int waves = (M + 15) / 16;
  int kchunk = K;
  if (splitk > 1) kchunk = (((K + splitk - 1) / splitk) + 31) / 32 * 32;
  const int ny = (K + kchunk - 1) / kchunk;
  hipLaunchKernelGGL((k_gemm<AT, BT, EPI>), dim3((waves + 3) / 4, ny), dim3(256), 0, st, M, N, K,
                     A, lda, B, ldb, bias, Cf, ldc, Hb, ldh, Zin, kchunk);
}

void fill(hipStream_t st, float* p, size_t n, float v) {
  if (n) hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
}

// Work buffer carve-up for a batch of n rows.
struct Work {
  float *Z[3], *O, *dO, *dZ, *loss;
  __bf16 *H[3], *dOb, *dZb[2];
  size_t bytes;
};
Work carve(void* base, int n) {
  Work w;
  size_t o = 0;
  auto take = [&](size_t bytes) {
    size_t at = o;
    o += (bytes + 255) / 256 * 256;
    return base ? (void*)((char*)base + at) : nullptr;
  };
  for (int l = 0; l < 3; l++) w.Z[l] = (float*)take(sizeof(float) * (size_t)n * HID);
  w.O = (float*)take(sizeof(float) * (size_t)n * NOUT);
  w.dO = (float*)take(sizeof(float) * (size_t)n * NOUT);
  w.dZ = (float*)take(sizeof(float) * (size_t)n * HID);
  w.loss = (float*)take(sizeof(float) * 64);
  for (int l = 0; l < 3; l++) w.H[l] = (__bf16*)take(2 * (size_t)n * HID);
  w.dOb = (__bf16*)take(2 * (size_t)n * NOUT);
  for (int l = 0; l < 2; l++) w.dZb[l] = (__bf16*)take(2 * (size_t)n * HID);
  w.bytes = o;
  return w;
}

// Forward pass; keeps pre-activations (Z) when `keep` is set.
void forward(hipStream_t st, const kpp_mlp* net, const __bf16* X, int n, const Work& w,
             bool keep) {
  const Layout L = layout(net->in_pad);
  const float* P = net->w;
  const __bf16* Pb = (const __bf16*)net->wb;
  gemm<false, false, EPI_BIAS_GELU>(st, n, HID, net->in_pad, X, net->in_pad, Pb + L.w[0],
                                    net->in_pad, P + L.b[0], keep ? w.Z[0] : nullptr, HID,
                                    w.H[0], HID, nullptr, 1);
  for (int l = 1; l < 3; l++)
    gemm<false, false, EPI_BIAS_GELU>(st, n, HID, HID, w.H[l - 1], HID, Pb + L.w[l], HID,
                                      P + L.b[l], keep ? w.Z[l] : nullptr, HID, w.H[l], HID,
                                      nullptr, 1);
  gemm<false, false, EPI_BIAS>(st, n, NOUT, HID, w.H[2], HID, Pb + L.w[3], HID, P + L.b[3], w.O,
                               NOUT, nullptr, 0, nullptr, 1);
}

int splitk_for(int n) { return n >= 8192 ? 64 : (n >= 1024 ? 16 : 1); }

// Forward, loss and backward; gradients (layout of the parameters) in G.
void grads(hipStream_t st, const kpp_mlp* net, const __bf16* X, const float* y, const int* bucket,
           const float* wreg, int n, float* G, const Work& w) {
  const Layout L = layout(net->in_pad);
  const __bf16* Pb = (const __bf16*)net->wb;
  forward(st, net, X, n, w, true);
  fill(st, w.loss, 1, 0.f);
  hipLaunchKernelGGL(k_loss, dim3((n + 255) / 256), dim3(256), 0, st, n, w.O, y, bucket, wreg,
                     w.dO, w.dOb, w.loss);
  fill(st, G, L.total, 0.f);
  const int sk = splitk_for(n);
  // output layer
  gemm<true, true, EPI_STORE>(st, NOUT, HID, n, w.dOb, NOUT, w.H[2], HID, nullptr, G + L.w[3],
                              HID, nullptr, 0, nullptr, sk);
  hipLaunchKernelGGL(k_colsum, dim3(64), dim3(256), 0, st, n, NOUT, w.dO, G + L.b[3]);
  // dZ3 = (dO . W4) * gelu'(Z3)
  gemm<false, true, EPI_GELU_BACK>(st, n, HID, NOUT, w.dOb, NOUT, Pb + L.w[3], HID, nullptr,
                                   w.dZ, HID, w.dZb[0], HID, w.Z[2], 1);
  int cur = 0;
  for (int l = 2; l >= 0; l--) {
    const __bf16* Hin = (l == 0) ? X : w.H[l - 1];
    const int kin = (l == 0) ? net->in_pad : HID;
    gemm<true, true, EPI_STORE>(st, HID, kin, n, w.dZb[cur], HID, Hin, kin, nullptr, G + L.w[l],
                                kin, nullptr, 0, nullptr, sk);
    hipLaunchKernelGGL(k_colsum, dim3(64), dim3(256), 0, st, n, HID, w.dZ, G + L.b[l]);
    if (l > 0) {
      gemm<false, true, EPI_GELU_BACK>(st, n, HID, HID, w.dZb[cur], HID, Pb + L.w[l], HID,
                                       nullptr, w.dZ, HID, w.dZb[1 - cur], HID, w.Z[l - 1], 1);
      cur = 1 - cur;
    }
  }
}

}  // namespace

extern "C" {

size_t kpp_mlp_param_count(int in_pad) {
  if (in_pad <= 0 || in_pad % 32) return 0;
  return layout(in_pad).total;
}

int kpp_mlp_param_offsets(int in_pad, size_t* w_off, size_t* b_off) {
  if (in_pad <= 0 || in_pad % 32 || !w_off || !b_off) return KPP_E_ARG;
  const Layout L = layout(in_pad);
  for (int l = 0; l < 4; l++) {
    w_off[l] = L.w[l];
    b_off[l] = L.b[l];
  }
  return 0;
}

size_t kpp_mlp_work_bytes(int n) { return n > 0 ? carve(nullptr, n).bytes : 0; }

int kpp_mlp_sync_bf16(const kpp_mlp* net, void* stream) {
  if (!net_ok(net)) return KPP_E_ARG;
  const size_t n = layout(net->in_pad).total;
  hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     as_stream(stream), n, net->w, (__bf16*)net->wb);
  return hipGetLastError() == hipSuccess ? 0 : KPP_E_HIP;
}

int kpp_mlp_features(int n, int ld, int nspec, const double* C, const double* temp,
                     const double* sun, const float* mu, const float* sd, double cfloor,
                     int in_pad, void* X, void* stream) {
  if (n <= 0 || ld < n || !C || !temp || !mu || !sd || !X || in_pad < nspec + 2 || in_pad % 32)
    return KPP_E_ARG;
  if ((reinterpret_cast<uintptr_t>(X) & 15) != 0) return KPP_E_ARG;
  hipLaunchKernelGGL(k_features, dim3((n + 255) / 256), dim3(256),
                     (unsigned)(256 * in_pad * sizeof(__bf16)), as_stream(stream), n, ld, nspec, C,
                     temp, sun, mu, sd, cfloor, in_pad, (__bf16*)X);
  return hipGetLastError() == hipSuccess ? 0 : KPP_E_HIP;
}

int kpp_mlp_forward(const kpp_mlp* net, const void* X, int n, float* out, void* work,
                    size_t work_bytes, void* stream) {
  if (!net_ok(net) || !X || n <= 0 || !work) return KPP_E_ARG;
  Work w = carve(work, n);
  if (work_bytes < w.bytes) return KPP_E_WORKSPACE;
  hipStream_t st = as_stream(stream);
  forward(st, net, (const __bf16*)X, n, w, false);
  if (out && hipMemcpyAsync(out, w.O, sizeof(float) * (size_t)n * NOUT, hipMemcpyDeviceToDevice,
                            st) != hipSuccess)
    return KPP_E_HIP;
  return hipGetLastError() == hipSuccess ? 0 : KPP_E_HIP;
}

int kpp_mlp_grad(const kpp_mlp* net, const void* X, const float* y, const int* bucket,
                 const float* wreg, int n, float* grad, float* loss, void* work,
                 size_t work_bytes, void* stream) {
  if (!net_ok(net) || !X || !y || !bucket || !wreg || n <= 0 || !grad || !work)
    return KPP_E_ARG;
  Work w = carve(work, n);
  if (work_bytes < w.bytes) return KPP_E_WORKSPACE;
  hipStream_t st = as_stream(stream);
  grads(st, net, (const __bf16*)X, y, bucket, wreg, n, grad, w);
  if (loss && hipMemcpyAsync(loss, w.loss, sizeof(float), hipMemcpyDeviceToDevice, st) !=
                  hipSuccess)
    return KPP_E_HIP;
  return hipGetLastError() == hipSuccess ? 0 : KPP_E_HIP;
}

int kpp_mlp_train_step(const kpp_mlp* net, const void* X, const float* y, const int* bucket,
                       const float* wreg, int n, float* grad, float* m1, float* m2, int step,
                       float lr, float* loss, void* work, size_t work_bytes, void* stream) {
  if (!m1 || !m2 || step < 1) return KPP_E_ARG;
  int r = kpp_mlp_grad(net, X, y, bucket, wreg, n, grad, loss, work, work_bytes, stream);
  if (r) return r;
  const size_t np = layout(net->in_pad).total;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, as_stream(stream),
                     np, net->w, (__bf16*)net->wb, grad, m1, m2, lr, b1, b2, 1e-8f, c1, c2);
  return hipGetLastError() == hipSuccess ? 0 : KPP_E_HIP;
}

size_t kpp_mlp_schedule_work_bytes(int n) {
  if (n <= 0) return 0;
  const int nblk = (n + SORT_CHUNK - 1) / SORT_CHUNK;
  return kpp_mlp_work_bytes(n) + ((size_t)n + 255) / 256 * 256 +
         ((size_t)NKEY * nblk * 4 + 255) / 256 * 256;
}

int kpp_mlp_schedule(const kpp_mlp* net, const void* X, int n, float mu_y, float sd_y,
                     double hmin, double hmax, const float* centres, double* hstart, int* bucket,
                     int* order, void* work, size_t work_bytes, void* stream) {
  if (!net_ok(net) || !X || n <= 0 || !work || !(hmin > 0) || !(hmax >= hmin)) return KPP_E_ARG;
  Centres cen{};
  if (centres) {
    for (int j = 0; j < NBUCKET; j++) {
      cen.c[j] = centres[j];
      if (!(centres[j] == centres[j])) return KPP_E_ARG;  // NaN
      if (j > 0 && !(centres[j] > centres[j - 1])) return KPP_E_ARG;  // increasing
    }
    cen.use = 1;
  }
  if (work_bytes < kpp_mlp_schedule_work_bytes(n)) return KPP_E_WORKSPACE;
  hipStream_t st = as_stream(stream);
  Work w = carve(work, n);
  forward(st, net, (const __bf16*)X, n, w, false);
  char* p = (char*)work + w.bytes;
  uint8_t* key = (uint8_t*)p;
  int* counts = (int*)(p + ((size_t)n + 255) / 256 * 256);
  hipLaunchKernelGGL(k_decode, dim3((n + 255) / 256), dim3(256), 0, st, n, w.O, mu_y, sd_y, hmin,
                     hmax, cen, hstart, bucket, order ? key : nullptr);
  if (order) {
    const int nblk = (n + SORT_CHUNK - 1) / SORT_CHUNK;
    hipLaunchKernelGGL(k_sort_count, dim3(nblk), dim3(SORT_THREADS), 0, st, n, key, nblk, counts);
    hipLaunchKernelGGL(k_sort_scan, dim3(1), dim3(1024), 0, st, NKEY * nblk, counts);
    hipLaunchKernelGGL(k_sort_scatter, dim3(nblk), dim3(SORT_THREADS), 0, st, n, key, nblk, counts,
                       order);
  }
  return hipGetLastError() == hipSuccess ? 0 : KPP_E_HIP;
}

}  // extern "C"
