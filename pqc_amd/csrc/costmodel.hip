// Learned cost model for the autotuner: a small MLP (in -> 32 -> 32 -> 1, ReLU)
// that predicts log(step time) from a launch configuration's features.  It is
// trained and evaluated on the GPU by the kernels below; nothing here touches a
// KEM value -- the model only ranks launch configurations, which never change
// a result.
//
// Parameters, fp32, flattened in this order:
//   W1[H][D], b1[H], W2[H][H], b2[H], w3[H], b3
// Loss: mean over the batch of (y - t)^2.
// Training step (full batch): k_mlp_grad (one thread per sample: forward,
// backward, per-workgroup gradient sums in LDS), k_mlp_reduce (one thread per
// parameter: sum over workgroups), k_mlp_adam (one thread per parameter).
#include <cmath>
#include <new>
#include <vector>

#include "internal.h"

namespace pqc {
namespace costmodel {

constexpr int H = 32;
constexpr int DMAX = 16;
constexpr int GB = 128;  // samples per workgroup in k_mlp_grad

__host__ __device__ constexpr int n_params(int d) { return H * d + H + H * H + H + H + 1; }

struct Offs {
  int w1, b1, w2, b2, w3, b3;
};
__host__ __device__ inline Offs offs(int d) {
  Offs o;
  o.w1 = 0;
  o.b1 = H * d;
  o.w2 = o.b1 + H;
  o.b2 = o.w2 + H * H;
  o.w3 = o.b2 + H;
  o.b3 = o.w3 + H;
  return o;
}

// forward for one sample; keeps the activations for the backward pass
__device__ inline float forward(const float* __restrict__ p, const float* x, int d, float (&h1)[H], float (&h2)[H]) {
  const Offs o = offs(d);
#pragma unroll 4
  for (int j = 0; j < H; ++j) {
    float a = p[o.b1 + j];
    for (int i = 0; i < d; ++i) a += p[o.w1 + j * d + i] * x[i];
    h1[j] = a > 0.f ? a : 0.f;
  }
#pragma unroll 4
  for (int j = 0; j < H; ++j) {
    float a = p[o.b2 + j];
#pragma unroll
    for (int i = 0; i < H; ++i) a += p[o.w2 + j * H + i] * h1[i];
    h2[j] = a > 0.f ? a : 0.f;
  }
  float y = p[o.b3];
#pragma unroll
  for (int j = 0; j < H; ++j) y += p[o.w3 + j] * h2[j];
  return y;
}

__global__ void k_mlp_forward(int n, int d, const float* __restrict__ x, const float* __restrict__ p,
                              float* __restrict__ y) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float xs[DMAX], h1[H], h2[H];
  for (int i = 0; i < d; ++i) xs[i] = x[(size_t)s * d + i];
  y[s] = forward(p, xs, d, h1, h2);
}

// dL/dparams summed over this workgroup's samples -> part[blockIdx][P]; the
// loss sum -> lpart[blockIdx].  dL/dy = 2 (y - t) / n.
__global__ __launch_bounds__(GB) void k_mlp_grad(int n, int d, const float* __restrict__ x,
                                                 const float* __restrict__ t, const float* __restrict__ p,
                                                 float* __restrict__ part, float* __restrict__ lpart) {
  extern __shared__ float g[];  // P gradient sums + 1 loss sum
  const int P = n_params(d);
  for (int i = threadIdx.x; i <= P; i += blockDim.x) g[i] = 0.f;
  __syncthreads();
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) {
    const Offs o = offs(d);
    float xs[DMAX], h1[H], h2[H];
    for (int i = 0; i < d; ++i) xs[i] = x[(size_t)s * d + i];
    const float y = forward(p, xs, d, h1, h2);
    const float e = y - t[s];
    atomicAdd(&g[P], e * e);
    const float dy = 2.f * e / (float)n;
    // output layer
    atomicAdd(&g[o.b3], dy);
    float d2[H];
#pragma unroll
    for (int j = 0; j < H; ++j) {
      atomicAdd(&g[o.w3 + j], dy * h2[j]);
      d2[j] = h2[j] > 0.f ? dy * p[o.w3 + j] : 0.f;
    }
    // second hidden layer
    float d1[H];
#pragma unroll
    for (int i = 0; i < H; ++i) d1[i] = 0.f;
#pragma unroll 4
    for (int j = 0; j < H; ++j) {
      atomicAdd(&g[o.b2 + j], d2[j]);
#pragma unroll
      for (int i = 0; i < H; ++i) {
        atomicAdd(&g[o.w2 + j * H + i], d2[j] * h1[i]);
        d1[i] += d2[j] * p[o.w2 + j * H + i];
      }
    }
    // first hidden layer
#pragma unroll 4
    for (int j = 0; j < H; ++j) {
      const float dj = h1[j] > 0.f ? d1[j] : 0.f;
      atomicAdd(&g[o.b1 + j], dj);
      for (int i = 0; i < d; ++i) atomicAdd(&g[o.w1 + j * d + i], dj * xs[i]);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < P; i += blockDim.x) part[(size_t)blockIdx.x * P + i] = g[i];
  if (threadIdx.x == 0) lpart[blockIdx.x] = g[P];
}

__global__ void k_mlp_reduce(int nb, int P, const float* __restrict__ part, const float* __restrict__ lpart,
                             float* __restrict__ grad, float* __restrict__ loss, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < P) {
    float a = 0.f;
    for (int b = 0; b < nb; ++b) a += part[(size_t)b * P + i];
    grad[i] = a;
  } else if (i == P) {
    float a = 0.f;
    for (int b = 0; b < nb; ++b) a += lpart[b];
    *loss = a / (float)n;
  }
}

// Adam (Kingma & Ba 2015), bias-corrected, step `t` >= 1
__global__ void k_mlp_adam(int P, float* __restrict__ p, const float* __restrict__ grad, float* __restrict__ m,
                           float* __restrict__ v, float lr, float b1, float b2, float eps, int t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  const float gi = grad[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / (1.f - powf(b1, (float)t));
  const float vh = vi / (1.f - powf(b2, (float)t));
  p[i] -= lr * mh / (sqrtf(vh) + eps);
}

}  // namespace costmodel
}  // namespace pqc

using namespace pqc::costmodel;

struct pqc_costmodel {
  int device = 0;
  int d = 0, P = 0, cap = 0;  // input width, parameter count, sample capacity
  int step = 0;               // Adam steps taken
  float *p = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr, *loss = nullptr;
  float *x = nullptr, *t = nullptr, *y = nullptr, *part = nullptr, *lpart = nullptr;
  hipStream_t s = nullptr;
};

static void cm_free(pqc_costmodel* c) {
  for (float* q : {c->p, c->m, c->v, c->grad, c->loss, c->x, c->t, c->y, c->part, c->lpart})
    if (q) (void)hipFree(q);
  if (c->s) (void)hipStreamDestroy(c->s);
  delete c;
}

// splitmix64, for a reproducible weight initialisation
static uint64_t sm64(uint64_t& z) {
  uint64_t r = (z += 0x9E3779B97F4A7C15ull);
  r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ull;
  r = (r ^ (r >> 27)) * 0x94D049BB133111EBull;
  return r ^ (r >> 31);
}

extern "C" {

int pqc_costmodel_create(pqc_costmodel** out, int in_dim, int max_samples, uint64_t seed) {
  if (!out || in_dim < 1 || in_dim > DMAX || max_samples < 1) return PQC_ERR_INVALID_ARG;
  *out = nullptr;
  auto* c = new (std::nothrow) pqc_costmodel;
  if (!c) return PQC_ERR_NO_MEMORY;
  if (hipGetDevice(&c->device) != hipSuccess) {
    delete c;
    return PQC_ERR_NO_DEVICE;
  }
  c->d = in_dim;
  c->P = n_params(in_dim);
  c->cap = max_samples;
  const int nb = (max_samples + GB - 1) / GB;
  bool ok = hipStreamCreate(&c->s) == hipSuccess;
  auto alloc = [&](float** q, size_t nfl) { ok = ok && hipMalloc(q, nfl * sizeof(float)) == hipSuccess; };
  alloc(&c->p, c->P);
  alloc(&c->m, c->P);
  alloc(&c->v, c->P);
  alloc(&c->grad, c->P);
  alloc(&c->loss, 1);
  alloc(&c->x, (size_t)max_samples * in_dim);
  alloc(&c->t, max_samples);
  alloc(&c->y, max_samples);
  alloc(&c->part, (size_t)nb * c->P);
  alloc(&c->lpart, nb);
  if (!ok) {
    cm_free(c);
    return PQC_ERR_NO_MEMORY;
  }
  // He-uniform weights, zero biases
  std::vector<float> h(c->P, 0.f);
  const Offs o = offs(in_dim);
  uint64_t z = seed;
  auto uni = [&](float lim) { return (float)((double)(sm64(z) >> 11) / 9007199254740992.0 * 2.0 - 1.0) * lim; };
  for (int i = 0; i < H * in_dim; ++i) h[o.w1 + i] = uni(sqrtf(6.f / in_dim));
  for (int i = 0; i < H * H; ++i) h[o.w2 + i] = uni(sqrtf(6.f / H));
  for (int i = 0; i < H; ++i) h[o.w3 + i] = uni(sqrtf(6.f / H));
  std::vector<float> zero(c->P, 0.f);
  ok = hipMemcpy(c->p, h.data(), c->P * sizeof(float), hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(c->m, zero.data(), c->P * sizeof(float), hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(c->v, zero.data(), c->P * sizeof(float), hipMemcpyHostToDevice) == hipSuccess;
  if (!ok) {
    cm_free(c);
    return PQC_ERR_DEVICE;
  }
  *out = c;
  return PQC_OK;
}

void pqc_costmodel_destroy(pqc_costmodel* c) {
  if (c) cm_free(c);
}

int pqc_costmodel_num_params(const pqc_costmodel* c) { return c ? c->P : 0; }

int pqc_costmodel_trained_steps(const pqc_costmodel* c) { return c ? c->step : 0; }

int pqc_costmodel_get_params(const pqc_costmodel* c, float* out, size_t count) {
  if (!c || !out || count < (size_t)c->P) return PQC_ERR_INVALID_ARG;
  return hipMemcpy(out, c->p, c->P * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess ? PQC_OK : PQC_ERR_DEVICE;
}

int pqc_costmodel_set_params(pqc_costmodel* c, const float* in, size_t count) {
  if (!c || !in || count < (size_t)c->P) return PQC_ERR_INVALID_ARG;
  return hipMemcpy(c->p, in, c->P * sizeof(float), hipMemcpyHostToDevice) == hipSuccess ? PQC_OK : PQC_ERR_DEVICE;
}

static int cm_upload(pqc_costmodel* c, int n, const float* x, const float* t) {
  if (n < 1 || n > c->cap || !x) return PQC_ERR_INVALID_ARG;
  if (hipMemcpyAsync(c->x, x, (size_t)n * c->d * sizeof(float), hipMemcpyHostToDevice, c->s) != hipSuccess)
    return PQC_ERR_DEVICE;
  if (t && hipMemcpyAsync(c->t, t, (size_t)n * sizeof(float), hipMemcpyHostToDevice, c->s) != hipSuccess)
    return PQC_ERR_DEVICE;
  return PQC_OK;
}

static void cm_grad(pqc_costmodel* c, int n) {
  const int nb = (n + GB - 1) / GB;
  hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(GB), (c->P + 1) * sizeof(float), c->s, n, c->d,
                     (const float*)c->x, (const float*)c->t, (const float*)c->p, c->part, c->lpart);
  hipLaunchKernelGGL(k_mlp_reduce, dim3((c->P + 1 + 255) / 256), dim3(256), 0, c->s, nb, c->P,
                     (const float*)c->part, (const float*)c->lpart, c->grad, c->loss, n);
}

int pqc_costmodel_predict(pqc_costmodel* c, int n, const float* x, float* y) {
  if (!c || !y) return PQC_ERR_INVALID_ARG;
  int rc = cm_upload(c, n, x, nullptr);
  if (rc) return rc;
  hipLaunchKernelGGL(k_mlp_forward, dim3((n + 127) / 128), dim3(128), 0, c->s, n, c->d, (const float*)c->x,
                     (const float*)c->p, c->y);
  if (hipMemcpyAsync(y, c->y, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, c->s) != hipSuccess)
    return PQC_ERR_DEVICE;
  return hipStreamSynchronize(c->s) == hipSuccess ? PQC_OK : PQC_ERR_DEVICE;
}

int pqc_costmodel_grad(pqc_costmodel* c, int n, const float* x, const float* t, float* loss, float* grad,
                       size_t count) {
  if (!c || !t || !loss || !grad || count < (size_t)c->P) return PQC_ERR_INVALID_ARG;
  int rc = cm_upload(c, n, x, t);
  if (rc) return rc;
  cm_grad(c, n);
  if (hipMemcpyAsync(grad, c->grad, c->P * sizeof(float), hipMemcpyDeviceToHost, c->s) != hipSuccess ||
      hipMemcpyAsync(loss, c->loss, sizeof(float), hipMemcpyDeviceToHost, c->s) != hipSuccess)
    return PQC_ERR_DEVICE;
  return hipStreamSynchronize(c->s) == hipSuccess ? PQC_OK : PQC_ERR_DEVICE;
}

int pqc_costmodel_fit(pqc_costmodel* c, int n, const float* x, const float* t, int steps, float lr,
                      float* final_loss) {
  if (!c || !t || steps < 1 || !(lr > 0.f)) return PQC_ERR_INVALID_ARG;
  int rc = cm_upload(c, n, x, t);
  if (rc) return rc;
  for (int i = 0; i < steps; ++i) {
    cm_grad(c, n);
    c->step += 1;
    hipLaunchKernelGGL(k_mlp_adam, dim3((c->P + 255) / 256), dim3(256), 0, c->s, c->P, c->p, (const float*)c->grad,
                       c->m, c->v, lr, 0.9f, 0.999f, 1e-8f, c->step);
  }
  if (final_loss) {
    cm_grad(c, n);  // loss after the last update
    if (hipMemcpyAsync(final_loss, c->loss, sizeof(float), hipMemcpyDeviceToHost, c->s) != hipSuccess)
      return PQC_ERR_DEVICE;
  }
  if (hipGetLastError() != hipSuccess) return PQC_ERR_DEVICE;
  return hipStreamSynchronize(c->s) == hipSuccess ? PQC_OK : PQC_ERR_DEVICE;
}

}  // extern "C"
