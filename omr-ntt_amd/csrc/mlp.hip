// mlp.hip -- learned schedule cost model: a 4 -> 64 -> 64 -> 36 fp32 MLP
// trained and run by the kernels below (VALU fp32; far too small for MFMA).
//
// Inputs (omr_cost_model_features): the field as its element width
// ((bytes - 4) / 4: 0 for the 31/32-bit fields, 1 for Goldilocks -- what
// changes the kernels), log2 N / 32, log2 out_len / 32, and the batch as the
// log2 of its total elements, log2(N batch) / 40 (what decides cache
// residency), so that a field or batch size absent from the training shapes
// maps onto features the model has seen.  Outputs: one score per schedule class
// (tuning.hip), trained on autotuner timings (normalised log time); the
// chosen schedule is the argmin over the classes valid for the shape.
// Gradients are computed per sample into rows of a buffer and reduced in a
// fixed order, so training is deterministic.
#include <hip/hip_runtime.h>
#include <new>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <algorithm>
#include <vector>

#include "internal.h"

namespace omr {

constexpr int kIn = 4, kH = 64, kOut = kNumClasses;
constexpr int oW1 = 0, oB1 = oW1 + kH * kIn, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH,
              oB3 = oW3 + kOut * kH, kParams = oB3 + kOut;
static_assert(kOut <= kH, "one thread per output in the forward kernel");

// One workgroup (64 threads) per sample.
__global__ void __launch_bounds__(kH) k_mlp_forward(const float *__restrict__ P, const float *__restrict__ X,
                                                    float *__restrict__ Y, float *__restrict__ H1,
                                                    float *__restrict__ H2) {
  __shared__ float x[kIn], h1[kH], h2[kH];
  const int b = blockIdx.x, j = threadIdx.x;
  if (j < kIn) x[j] = X[b * kIn + j];
  __syncthreads();
  float a = P[oB1 + j];
  for (int i = 0; i < kIn; i++) a += P[oW1 + j * kIn + i] * x[i];
  a = a > 0.f ? a : 0.f;
  h1[j] = a;
  H1[b * kH + j] = a;
  __syncthreads();
  float c = P[oB2 + j];
  for (int i = 0; i < kH; i++) c += P[oW2 + j * kH + i] * h1[i];
  c = c > 0.f ? c : 0.f;
  h2[j] = c;
  H2[b * kH + j] = c;
  __syncthreads();
  if (j < kOut) {
    float y = P[oB3 + j];
    for (int i = 0; i < kH; i++) y += P[oW3 + j * kH + i] * h2[i];
    Y[b * kOut + j] = y;
  }
}

// Masked mean-squared error; per-sample gradient rows G[b][kParams] and
// per-sample loss terms.  inv_count = 1 / sum(mask).
__global__ void __launch_bounds__(kH) k_mlp_backward(const float *__restrict__ P, const float *__restrict__ X,
                                                     const float *__restrict__ H1, const float *__restrict__ H2,
                                                     const float *__restrict__ Y, const float *__restrict__ T,
                                                     const float *__restrict__ M, float inv_count,
                                                     float *__restrict__ G, float *__restrict__ L) {
  __shared__ float dy[kOut], dh2[kH], dh1[kH], sl[kH];
  const int b = blockIdx.x, j = threadIdx.x;
  float *g = G + (size_t)b * kParams;
  const float *h1 = H1 + b * kH, *h2 = H2 + b * kH, *x = X + b * kIn;
  sl[j] = 0.f;
  if (j < kOut) {
    const float m = M[b * kOut + j];
    const float d = Y[b * kOut + j] - T[b * kOut + j];
    dy[j] = m * 2.f * d * inv_count;
    sl[j] = m * d * d * inv_count;
  }
  __syncthreads();
  if (j < kOut) {
    for (int i = 0; i < kH; i++) g[oW3 + j * kH + i] = dy[j] * h2[i];
    g[oB3 + j] = dy[j];
  }
  {
    float s = 0.f;
    for (int k = 0; k < kOut; k++) s += P[oW3 + k * kH + j] * dy[k];
    dh2[j] = h2[j] > 0.f ? s : 0.f;
  }
  __syncthreads();
  for (int i = 0; i < kH; i++) g[oW2 + j * kH + i] = dh2[j] * h1[i];
  g[oB2 + j] = dh2[j];
  {
    float s = 0.f;
    for (int k = 0; k < kH; k++) s += P[oW2 + k * kH + j] * dh2[k];
    dh1[j] = h1[j] > 0.f ? s : 0.f;
  }
  __syncthreads();
  for (int i = 0; i < kIn; i++) g[oW1 + j * kIn + i] = dh1[j] * x[i];
  g[oB1 + j] = dh1[j];
  if (j == 0) {
    float s = 0.f;
    for (int k = 0; k < kOut; k++) s += sl[k];
    L[b] = s;
  }
}

// out[i] = sum_b rows[b][i] in order b = 0..B-1 (deterministic).
__global__ void __launch_bounds__(256) k_sum_rows_f32(const float *__restrict__ rows, uint32_t B,
                                                      uint32_t width, float *__restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= width) return;
  float s = 0.f;
  for (uint32_t b = 0; b < B; b++) s += rows[(size_t)b * width + i];
  out[i] = s;
}

__global__ void __launch_bounds__(256) k_adam(float *__restrict__ P, const float *__restrict__ g,
                                              float *__restrict__ m, float *__restrict__ v, float lr,
                                              float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kParams) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  P[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace omr

using namespace omr;

extern "C" omr_status omr_tune_classes(omr_ctx *ctx, int field, uint64_t n, uint64_t out_len,
                                       uint64_t batch, int reps, float *times_us);
extern "C" omr_status omr_tune_set_class(int field, uint64_t n, uint64_t out_len, int cls);
extern "C" omr_status omr_tune_set_class_batch(int field, uint64_t n, uint64_t out_len, uint64_t batch, int cls);

struct omr_cost_model {
  omr_ctx *ctx = nullptr;
  float *P = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;
  uint32_t cap = 0;  // samples the work buffers hold
  float *X = nullptr, *T = nullptr, *M = nullptr, *Y = nullptr, *H1 = nullptr, *H2 = nullptr, *G = nullptr,
        *L = nullptr, *Lsum = nullptr;
  bool trained = false;
  uint64_t adam_t = 0;
};

static void mlp_free_work(omr_cost_model *cm) {
  float **bufs[] = {&cm->X, &cm->T, &cm->M, &cm->Y, &cm->H1, &cm->H2, &cm->G, &cm->L, &cm->Lsum};
  for (float **b : bufs) {
    if (*b) (void)hipFree(*b);
    *b = nullptr;
  }
  cm->cap = 0;
}

static omr_status mlp_reserve(omr_cost_model *cm, uint32_t B) {
  if (B <= cm->cap) return OMR_OK;
  mlp_free_work(cm);
  struct {
    float **p;
    size_t n;
  } want[] = {{&cm->X, (size_t)B * kIn},  {&cm->T, (size_t)B * kOut}, {&cm->M, (size_t)B * kOut},
              {&cm->Y, (size_t)B * kOut}, {&cm->H1, (size_t)B * kH},  {&cm->H2, (size_t)B * kH},
              {&cm->G, (size_t)B * kParams}, {&cm->L, (size_t)B},      {&cm->Lsum, 1}};
  for (auto &w : want)
    if (hipMalloc(w.p, w.n * sizeof(float)) != hipSuccess) {
      (void)hipGetLastError();
      mlp_free_work(cm);
      return OMR_E_ALLOC;
    }
  cm->cap = B;
  return OMR_OK;
}

extern "C" int omr_cost_model_num_params(void) { return kParams; }
extern "C" int omr_cost_model_num_inputs(void) { return kIn; }
extern "C" int omr_cost_model_num_outputs(void) { return kOut; }

extern "C" omr_status omr_cost_model_destroy(omr_cost_model *cm) try {
  if (!cm) return OMR_E_NULL;
  (void)hipDeviceSynchronize();
  mlp_free_work(cm);
  for (float *p : {cm->P, cm->m, cm->v, cm->g})
    if (p) (void)hipFree(p);
  delete cm;
  return OMR_OK;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

extern "C" omr_status omr_cost_model_set_params(omr_cost_model *cm, const float *params) try {
  if (!cm || !params) return OMR_E_NULL;
  if (hipMemcpy(cm->P, params, kParams * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) return OMR_E_DEVICE;
  (void)hipMemset(cm->m, 0, kParams * sizeof(float));
  (void)hipMemset(cm->v, 0, kParams * sizeof(float));
  cm->adam_t = 0;
  return OMR_OK;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

extern "C" omr_status omr_cost_model_get_params(omr_cost_model *cm, float *params) try {
  if (!cm || !params) return OMR_E_NULL;
  return hipMemcpy(params, cm->P, kParams * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess ? OMR_OK
                                                                                                : OMR_E_DEVICE;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

// He-uniform initialisation from a seeded host generator.
extern "C" omr_status omr_cost_model_create(omr_ctx *ctx, uint32_t seed, omr_cost_model **out) try {
  if (!ctx || !out) return OMR_E_NULL;
  if (hipSetDevice(ctx->device) != hipSuccess) return OMR_E_DEVICE;
  omr_cost_model *cm = new (std::nothrow) omr_cost_model();
  if (!cm) return OMR_E_ALLOC;
  cm->ctx = ctx;
  for (float **p : {&cm->P, &cm->m, &cm->v, &cm->g})
    if (hipMalloc(p, kParams * sizeof(float)) != hipSuccess) {
      (void)hipGetLastError();
      omr_cost_model_destroy(cm);
      return OMR_E_ALLOC;
    }
  std::vector<float> h(kParams, 0.f);
  std::mt19937 rng(seed);
  auto fill = [&](int off, int fan_in, int count) {
    const float lim = std::sqrt(6.f / (float)fan_in);
    std::uniform_real_distribution<float> d(-lim, lim);
    for (int i = 0; i < count; i++) h[off + i] = d(rng);
  };
  fill(oW1, kIn, kH * kIn);
  fill(oW2, kH, kH * kH);
  fill(oW3, kH, kOut * kH);
  omr_status st = omr_cost_model_set_params(cm, h.data());
  if (st != OMR_OK) {
    omr_cost_model_destroy(cm);
    return st;
  }
  *out = cm;
  return OMR_OK;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

extern "C" omr_status omr_cost_model_features(int field, uint64_t n, uint64_t out_len, uint64_t batch,
                                              float *x) try {
  if (!x) return OMR_E_NULL;
  if (!field_ok(field)) return OMR_E_FIELD;
  if (n == 0 || out_len == 0 || batch == 0) return OMR_E_SIZE;
  x[0] = (float)(field_bytes(field) - 4) / 4.f;
  x[1] = (float)std::log2((double)n) / 32.f;
  x[2] = (float)std::log2((double)out_len) / 32.f;
  x[3] = (float)(std::log2((double)n) + std::log2((double)batch)) / 40.f;
  return OMR_OK;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

static omr_status mlp_upload_forward(omr_cost_model *cm, const float *x, uint32_t B) {
  omr_status st = mlp_reserve(cm, B);
  if (st != OMR_OK) return st;
  if (hipMemcpy(cm->X, x, (size_t)B * kIn * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) return OMR_E_DEVICE;
  hipLaunchKernelGGL(k_mlp_forward, dim3(B), dim3(kH), 0, 0, cm->P, cm->X, cm->Y, cm->H1, cm->H2);
  return hipGetLastError() == hipSuccess ? OMR_OK : OMR_E_DEVICE;
}

extern "C" omr_status omr_cost_model_forward(omr_cost_model *cm, const float *x, uint32_t B, float *y) try {
  if (!cm || !x || !y) return OMR_E_NULL;
  if (B == 0) return OMR_E_ARG;
  if (hipSetDevice(cm->ctx->device) != hipSuccess) return OMR_E_DEVICE;
  omr_status st = mlp_upload_forward(cm, x, B);
  if (st != OMR_OK) return st;
  return hipMemcpy(y, cm->Y, (size_t)B * kOut * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess
             ? OMR_OK
             : OMR_E_DEVICE;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

// One forward + backward over the (already uploaded) batch; leaves the summed
// gradient in cm->g and the loss in cm->Lsum.
static omr_status mlp_grad_step(omr_cost_model *cm, uint32_t B, float inv_count) {
  hipLaunchKernelGGL(k_mlp_forward, dim3(B), dim3(kH), 0, 0, cm->P, cm->X, cm->Y, cm->H1, cm->H2);
  hipLaunchKernelGGL(k_mlp_backward, dim3(B), dim3(kH), 0, 0, cm->P, cm->X, cm->H1, cm->H2, cm->Y, cm->T,
                     cm->M, inv_count, cm->G, cm->L);
  hipLaunchKernelGGL(k_sum_rows_f32, dim3((kParams + 255) / 256), dim3(256), 0, 0, cm->G, B, (uint32_t)kParams,
                     cm->g);
  hipLaunchKernelGGL(k_sum_rows_f32, dim3(1), dim3(256), 0, 0, cm->L, B, 1u, cm->Lsum);
  return hipGetLastError() == hipSuccess ? OMR_OK : OMR_E_DEVICE;
}

static omr_status mlp_upload_train(omr_cost_model *cm, const float *x, const float *t, const float *mask,
                                   uint32_t B, float &inv_count) {
  omr_status st = mlp_reserve(cm, B);
  if (st != OMR_OK) return st;
  double cnt = 0;
  for (size_t i = 0; i < (size_t)B * kOut; i++) cnt += mask[i];
  if (cnt <= 0) return OMR_E_ARG;
  inv_count = (float)(1.0 / cnt);
  if (hipMemcpy(cm->X, x, (size_t)B * kIn * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(cm->T, t, (size_t)B * kOut * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(cm->M, mask, (size_t)B * kOut * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    return OMR_E_DEVICE;
  return OMR_OK;
}

extern "C" omr_status omr_cost_model_gradients(omr_cost_model *cm, const float *x, const float *t,
                                               const float *mask, uint32_t B, float *grad, float *loss) try {
  if (!cm || !x || !t || !mask || !grad) return OMR_E_NULL;
  if (B == 0) return OMR_E_ARG;
  if (hipSetDevice(cm->ctx->device) != hipSuccess) return OMR_E_DEVICE;
  float inv;
  omr_status st = mlp_upload_train(cm, x, t, mask, B, inv);
  if (st != OMR_OK) return st;
  if ((st = mlp_grad_step(cm, B, inv)) != OMR_OK) return st;
  if (hipMemcpy(grad, cm->g, kParams * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) return OMR_E_DEVICE;
  if (loss && hipMemcpy(loss, cm->Lsum, sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) return OMR_E_DEVICE;
  return OMR_OK;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

// Full-batch Adam for `steps` steps; *loss_out = loss before the last step.
extern "C" omr_status omr_cost_model_train(omr_cost_model *cm, const float *x, const float *t,
                                           const float *mask, uint32_t B, uint32_t steps, float lr,
                                           float *loss_out) try {
  if (!cm || !x || !t || !mask) return OMR_E_NULL;
  if (B == 0 || !(lr > 0.f)) return OMR_E_ARG;
  if (hipSetDevice(cm->ctx->device) != hipSuccess) return OMR_E_DEVICE;
  float inv;
  omr_status st = mlp_upload_train(cm, x, t, mask, B, inv);
  if (st != OMR_OK) return st;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (uint32_t s = 0; s < steps; s++) {
    if ((st = mlp_grad_step(cm, B, inv)) != OMR_OK) return st;
    cm->adam_t++;
    const float c1 = 1.f - std::pow(b1, (float)cm->adam_t), c2 = 1.f - std::pow(b2, (float)cm->adam_t);
    hipLaunchKernelGGL(k_adam, dim3((kParams + 255) / 256), dim3(256), 0, 0, cm->P, cm->g, cm->m, cm->v, lr, b1,
                       b2, eps, c1, c2);
    if (hipGetLastError() != hipSuccess) return OMR_E_DEVICE;
  }
  if (loss_out && hipMemcpy(loss_out, cm->Lsum, sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    return OMR_E_DEVICE;
  if (hipDeviceSynchronize() != hipSuccess) return OMR_E_DEVICE;
  cm->trained = true;
  return OMR_OK;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

// Schedule class the model scores lowest among the classes valid for the shape.
extern "C" omr_status omr_cost_model_choose(omr_cost_model *cm, int field, uint64_t n, uint64_t out_len,
                                            uint64_t batch, int *cls) try {
  if (!cm || !cls) return OMR_E_NULL;
  float x[kIn], y[kOut];
  omr_status st = omr_cost_model_features(field, n, out_len, batch, x);
  if (st != OMR_OK) return st;
  if ((st = omr_cost_model_forward(cm, x, 1, y)) != OMR_OK) return st;
  const uint32_t logN = 63 - __builtin_clzll(n), logOut = 63 - __builtin_clzll(out_len);
  int best = -1;
  for (int c = 0; c < kOut; c++) {
    Schedule s;
    int m;
    uint32_t l1;
    const bool ok = out_len == n ? class_schedule(field, c, logN, s) : class_partial(field, c, logN, logOut, m, l1);
    if (ok && (best < 0 || y[c] < y[best])) best = c;
  }
  if (best < 0) return OMR_E_SCHEDULE;
  *cls = best;
  return OMR_OK;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

// Picks and installs a schedule for a shape.  With a trained model:
// use_model == 1 takes the model's best-predicted class (no timing runs);
// use_model = k >= 2 times only the model's k best-predicted valid classes
// and keeps the fastest (a model-pruned search).  use_model == 0, or an
// untrained model: exhaustive timing of every valid class.  *timed = number
// of classes timed.
extern "C" omr_status omr_autotune_shape(omr_ctx *ctx, omr_cost_model *cm, int use_model, int field,
                                         uint64_t n, uint64_t out_len, uint64_t batch, int reps, int *cls,
                                         int *timed) try {
  if (!ctx || !cls) return OMR_E_NULL;
  omr_status st;
  int chosen = -1, ntimed = 0;
  if (use_model == 1 && cm && cm->trained) {
    if ((st = omr_cost_model_choose(cm, field, n, out_len, batch, &chosen)) != OMR_OK) return st;
  } else if (use_model >= 2 && cm && cm->trained) {
    float x[kIn], y[kOut];
    if ((st = omr_cost_model_features(field, n, out_len, batch, x)) != OMR_OK) return st;
    if ((st = omr_cost_model_forward(cm, x, 1, y)) != OMR_OK) return st;
    const uint32_t logN = 63 - __builtin_clzll(n), logOut = 63 - __builtin_clzll(out_len);
    std::vector<int> valid;
    for (int c = 0; c < kOut; c++) {
      Schedule s;
      int m;
      uint32_t l1;
      if (out_len == n ? class_schedule(field, c, logN, s) : class_partial(field, c, logN, logOut, m, l1))
        valid.push_back(c);
    }
    if (valid.empty()) return OMR_E_SCHEDULE;
    std::stable_sort(valid.begin(), valid.end(), [&](int p, int q) { return y[p] < y[q]; });
    bool only[kOut] = {};
    for (size_t i = 0; i < valid.size() && i < (size_t)use_model; i++) only[valid[i]] = true;
    float t[kOut];
    if ((st = tune_classes_subset(ctx, field, n, out_len, batch, reps, only, t)) != OMR_OK) return st;
    for (int c = 0; c < kOut; c++)
      if (t[c] >= 0.f) {
        ntimed++;
        if (chosen < 0 || t[c] < t[chosen]) chosen = c;
      }
    if (chosen < 0) return OMR_E_SCHEDULE;
  } else {
    float t[kOut];
    if ((st = omr_tune_classes(ctx, field, n, out_len, batch, reps, t)) != OMR_OK) return st;
    for (int c = 0; c < kOut; c++)
      if (t[c] >= 0.f) {
        ntimed++;
        if (chosen < 0 || t[c] < t[chosen]) chosen = c;
      }
    if (chosen < 0) return OMR_E_SCHEDULE;
  }
  if ((st = omr_tune_set_class_batch(field, n, out_len, batch, chosen)) != OMR_OK) return st;
  *cls = chosen;
  if (timed) *timed = ntimed;
  return OMR_OK;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

extern "C" omr_status omr_cost_model_save(omr_cost_model *cm, const char *path) try {
  if (!cm || !path) return OMR_E_NULL;
  std::vector<float> h(kParams);
  omr_status st = omr_cost_model_get_params(cm, h.data());
  if (st != OMR_OK) return st;
  FILE *fp = fopen(path, "wb");
  if (!fp) return OMR_E_ARG;
  const uint32_t hdr[3] = {0x4F4D524Du, (uint32_t)kParams, cm->trained ? 1u : 0u};
  bool ok = fwrite(hdr, sizeof(hdr), 1, fp) == 1 && fwrite(h.data(), sizeof(float), kParams, fp) == (size_t)kParams;
  fclose(fp);
  return ok ? OMR_OK : OMR_E_ARG;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

extern "C" omr_status omr_cost_model_load(omr_cost_model *cm, const char *path) try {
  if (!cm || !path) return OMR_E_NULL;
  FILE *fp = fopen(path, "rb");
  if (!fp) return OMR_E_ARG;
  uint32_t hdr[3];
  std::vector<float> h(kParams);
  bool ok = fread(hdr, sizeof(hdr), 1, fp) == 1 && hdr[0] == 0x4F4D524Du && hdr[1] == (uint32_t)kParams &&
            fread(h.data(), sizeof(float), kParams, fp) == (size_t)kParams;
  fclose(fp);
  if (!ok) return OMR_E_ARG;
  omr_status st = omr_cost_model_set_params(cm, h.data());
  if (st == OMR_OK) cm->trained = hdr[2] != 0;
  return st;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}
