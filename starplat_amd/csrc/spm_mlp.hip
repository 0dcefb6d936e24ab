// spm_mlp.hip -- the learned schedule selector's network: a 16 -> 64 -> 64 -> 8
// perceptron (ReLU hidden layers, softmax cross-entropy), trained and run on
// the GPU by the kernels below. No counterpart in the reference; it only
// chooses schedules, never the arithmetic of the graph kernels.
//
// Numerics: FP32 master parameters; forward and backward read BF16 copies of
// the weight matrices (round-to-nearest-even) and accumulate in FP32; biases
// stay FP32. The Adam update writes the FP32 masters and refreshes the BF16
// copies.
//
// Mapping: one workgroup per mini-batch chunk of up to kMlpChunk samples;
// thread j owns neuron j of the current layer and loops over the samples, so
// activations live in LDS and the weight gradient of row j is a private sum
// of that thread (no atomics inside a workgroup). Each workgroup then adds
// its partial gradient to the global one with one atomic per parameter.
#include <cmath>
#include <cstring>
#include <vector>

#include "spm_internal.hpp"

namespace spm {

constexpr int kIn = 16, kH = 64, kOut = 8;
constexpr int kMlpChunk = 64;
constexpr int kW1 = 0, kB1 = kW1 + kH * kIn, kW2 = kB1 + kH, kB2 = kW2 + kH * kH, kW3 = kB2 + kH,
              kB3 = kW3 + kOut * kH, kNParams = kB3 + kOut;

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return uint16_t(u >> 16);  // inf / nan kept
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}

// Weight value as used by the network: BF16 copy for matrices, FP32 bias.
__device__ __forceinline__ float wv(const uint16_t* wb, int k) { return bf16_to_f32(wb[k]); }

// Forward of one chunk into LDS; logits in s_o. Threads: kH (64).
__device__ void mlp_forward_chunk(const float* __restrict__ x, int n, const uint16_t* __restrict__ wb,
                                  const float* __restrict__ p, float (*s_x)[kIn], float (*s_h1)[kH],
                                  float (*s_h2)[kH], float (*s_o)[kOut]) {
  const int j = threadIdx.x;
  for (int s = j; s < n * kIn; s += kH) s_x[s / kIn][s % kIn] = x[s];
  __syncthreads();
  for (int s = 0; s < n; ++s) {
    float a = p[kB1 + j];
#pragma unroll
    for (int i = 0; i < kIn; ++i) a += wv(wb, kW1 + j * kIn + i) * s_x[s][i];
    s_h1[s][j] = a > 0.f ? a : 0.f;
  }
  __syncthreads();
  for (int s = 0; s < n; ++s) {
    float a = p[kB2 + j];
#pragma unroll 16
    for (int i = 0; i < kH; ++i) a += wv(wb, kW2 + j * kH + i) * s_h1[s][i];
    s_h2[s][j] = a > 0.f ? a : 0.f;
  }
  __syncthreads();
  if (j < kOut) {
    for (int s = 0; s < n; ++s) {
      float a = p[kB3 + j];
#pragma unroll 16
      for (int i = 0; i < kH; ++i) a += wv(wb, kW3 + j * kH + i) * s_h2[s][i];
      s_o[s][j] = a;
    }
  }
  __syncthreads();
}

__global__ __launch_bounds__(kH) void mlp_forward(const float* __restrict__ x, int n,
                                                  const uint16_t* __restrict__ wb, const float* __restrict__ p,
                                                  float* __restrict__ logits) {
  __shared__ float s_x[kMlpChunk][kIn];
  __shared__ float s_h1[kMlpChunk][kH];
  __shared__ float s_h2[kMlpChunk][kH];
  __shared__ float s_o[kMlpChunk][kOut];
  const int base = blockIdx.x * kMlpChunk;
  const int m = min(kMlpChunk, n - base);
  if (m <= 0) return;
  mlp_forward_chunk(x + size_t(base) * kIn, m, wb, p, s_x, s_h1, s_h2, s_o);
  for (int t = threadIdx.x; t < m * kOut; t += kH) logits[size_t(base) * kOut + t] = s_o[t / kOut][t % kOut];
}

// Gradient of mean cross-entropy over all n samples (the divisor is n_total);
// grad must be zeroed by the caller; loss_sum receives the summed loss.
__global__ __launch_bounds__(kH) void mlp_backward(const float* __restrict__ x, const int32_t* __restrict__ y,
                                                   int n, int n_total, const uint16_t* __restrict__ wb,
                                                   const float* __restrict__ p, float* __restrict__ grad,
                                                   float* __restrict__ loss_sum) {
  __shared__ float s_x[kMlpChunk][kIn];
  __shared__ float s_h1[kMlpChunk][kH];
  __shared__ float s_h2[kMlpChunk][kH];
  __shared__ float s_o[kMlpChunk][kOut];
  __shared__ float s_d2[kMlpChunk][kH];
  __shared__ float s_loss[kMlpChunk];
  const int base = blockIdx.x * kMlpChunk;
  const int m = min(kMlpChunk, n - base);
  if (m <= 0) return;
  const int j = threadIdx.x;
  mlp_forward_chunk(x + size_t(base) * kIn, m, wb, p, s_x, s_h1, s_h2, s_o);
  const float inv = 1.f / float(n_total);
  // softmax cross-entropy: s_o becomes dL/dlogits
  if (j < m) {
    const int s = j;
    float mx = s_o[s][0];
    for (int k = 1; k < kOut; ++k) mx = fmaxf(mx, s_o[s][k]);
    float z = 0.f;
    for (int k = 0; k < kOut; ++k) z += expf(s_o[s][k] - mx);
    const int lab = y[base + s];
    s_loss[s] = logf(z) + mx - s_o[s][lab];
    for (int k = 0; k < kOut; ++k) s_o[s][k] = (expf(s_o[s][k] - mx) / z - (k == lab ? 1.f : 0.f)) * inv;
  }
  __syncthreads();
  if (j == 0) {
    float t = 0.f;
    for (int s = 0; s < m; ++s) t += s_loss[s];
    atomicAdd(loss_sum, t);
  }
  // layer 3: thread k < 8 owns row k of W3
  if (j < kOut) {
    float gb = 0.f;
    for (int s = 0; s < m; ++s) gb += s_o[s][j];
    atomicAdd(&grad[kB3 + j], gb);
    for (int i = 0; i < kH; ++i) {
      float g = 0.f;
      for (int s = 0; s < m; ++s) g += s_o[s][j] * s_h2[s][i];
      atomicAdd(&grad[kW3 + j * kH + i], g);
    }
  }
  // dh2 (thread j = hidden unit j of layer 2)
  for (int s = 0; s < m; ++s) {
    float a = 0.f;
#pragma unroll
    for (int k = 0; k < kOut; ++k) a += wv(wb, kW3 + k * kH + j) * s_o[s][k];
    s_d2[s][j] = s_h2[s][j] > 0.f ? a : 0.f;
  }
  __syncthreads();
  {
    float gb = 0.f;
    for (int s = 0; s < m; ++s) gb += s_d2[s][j];
    atomicAdd(&grad[kB2 + j], gb);
    for (int i = 0; i < kH; ++i) {
      float g = 0.f;
      for (int s = 0; s < m; ++s) g += s_d2[s][j] * s_h1[s][i];
      atomicAdd(&grad[kW2 + j * kH + i], g);
    }
  }
  // dh1, written over s_h2 (no longer needed)
  for (int s = 0; s < m; ++s) {
    float a = 0.f;
#pragma unroll 16
    for (int k = 0; k < kH; ++k) a += wv(wb, kW2 + k * kH + j) * s_d2[s][k];
    s_h2[s][j] = s_h1[s][j] > 0.f ? a : 0.f;
  }
  __syncthreads();
  {
    float gb = 0.f;
    for (int s = 0; s < m; ++s) gb += s_h2[s][j];
    atomicAdd(&grad[kB1 + j], gb);
    for (int i = 0; i < kIn; ++i) {
      float g = 0.f;
      for (int s = 0; s < m; ++s) g += s_h2[s][j] * s_x[s][i];
      atomicAdd(&grad[kW1 + j * kIn + i], g);
    }
  }
}

// Adam (Kingma & Ba) on the FP32 masters; refreshes the BF16 copies.
__global__ void mlp_adam(float* __restrict__ p, uint16_t* __restrict__ wb, const float* __restrict__ grad,
                         float* __restrict__ m1, float* __restrict__ m2, float lr, float b1, float b2,
                         float eps, float c1, float c2) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= kNParams) return;
  const float g = grad[k];
  const float a = b1 * m1[k] + (1.f - b1) * g;
  const float b = b2 * m2[k] + (1.f - b2) * g * g;
  m1[k] = a;
  m2[k] = b;
  p[k] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
  wb[k] = f32_to_bf16(p[k]);
}

__global__ void mlp_refresh_bf16(const float* __restrict__ p, uint16_t* __restrict__ wb) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < kNParams) wb[k] = f32_to_bf16(p[k]);
}

__global__ void mlp_zero(float* __restrict__ a, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) a[k] = 0.f;
}

}  // namespace spm

using namespace spm;

struct spm_mlp {
  hipStream_t stream = nullptr;
  float* p = nullptr;       // FP32 masters [kNParams]
  uint16_t* wb = nullptr;   // BF16 copy [kNParams]
  float* grad = nullptr;    // [kNParams + 1]: gradient, then the loss sum
  float* m1 = nullptr;
  float* m2 = nullptr;
  float* x = nullptr;       // staged samples
  int32_t* y = nullptr;
  float* out = nullptr;
  int cap = 0;
  int64_t step = 0;
};

namespace {

int mlp_reserve(spm_mlp* m, int n) {
  if (n <= m->cap) return SPM_OK;
  (void)hipFree(m->x);
  (void)hipFree(m->y);
  (void)hipFree(m->out);
  m->x = nullptr;
  m->y = nullptr;
  m->out = nullptr;
  m->cap = 0;
  SPM_CHECK_HIP(hipMalloc(&m->x, sizeof(float) * size_t(n) * kIn));
  SPM_CHECK_HIP(hipMalloc(&m->y, sizeof(int32_t) * size_t(n)));
  SPM_CHECK_HIP(hipMalloc(&m->out, sizeof(float) * size_t(n) * kOut));
  m->cap = n;
  return SPM_OK;
}

int mlp_grad_dev(spm_mlp* m, int n, int off, int cnt) {
  hipLaunchKernelGGL(mlp_zero, dim3((kNParams + 1 + 255) / 256), dim3(256), 0, m->stream, m->grad, kNParams + 1);
  hipLaunchKernelGGL(mlp_backward, dim3((cnt + kMlpChunk - 1) / kMlpChunk), dim3(kH), 0, m->stream,
                     (const float*)(m->x + size_t(off) * kIn), (const int32_t*)(m->y + off), cnt, cnt,
                     (const uint16_t*)m->wb, (const float*)m->p, m->grad, m->grad + kNParams);
  (void)n;
  SPM_CHECK_HIP(hipGetLastError());
  return SPM_OK;
}

}  // namespace

extern "C" {

int spm_mlp_create(int32_t n_in, int32_t n_hidden, int32_t n_out, uint64_t seed, spm_mlp** out) {
  if (!out) return set_error(SPM_EINVAL, "spm_mlp_create: out is NULL");
  *out = nullptr;
  if (n_in != kIn || n_hidden != kH || n_out != kOut)
    return set_error(SPM_EINVAL, "spm_mlp_create: this build's network is 16 -> 64 -> 64 -> 8");
  spm_mlp* m = new spm_mlp();
  hipError_t e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&m->p, sizeof(float) * kNParams);
  if (e == hipSuccess) e = hipMalloc(&m->wb, sizeof(uint16_t) * kNParams);
  if (e == hipSuccess) e = hipMalloc(&m->grad, sizeof(float) * (kNParams + 1));
  if (e == hipSuccess) e = hipMalloc(&m->m1, sizeof(float) * kNParams);
  if (e == hipSuccess) e = hipMalloc(&m->m2, sizeof(float) * kNParams);
  if (e != hipSuccess) {
    spm_mlp_destroy(m);
    return set_error(SPM_ENOMEM, std::string("spm_mlp_create: ") + hipGetErrorString(e));
  }
  // He-uniform initialisation from a counter-based hash of (seed, index).
  std::vector<float> h(kNParams, 0.f);
  auto uni = [&](uint64_t k) {
    uint64_t z = seed * 0x9E3779B97F4A7C15ull + k + 0x632BE59BD9B4E019ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return double(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
  };
  for (int k = 0; k < kH * kIn; ++k) h[kW1 + k] = float(uni(k) * std::sqrt(6.0 / kIn));
  for (int k = 0; k < kH * kH; ++k) h[kW2 + k] = float(uni(100000 + k) * std::sqrt(6.0 / kH));
  for (int k = 0; k < kOut * kH; ++k) h[kW3 + k] = float(uni(200000 + k) * std::sqrt(6.0 / kH));
  int rc = spm_mlp_set_params(m, h.data());
  if (rc != SPM_OK) {
    spm_mlp_destroy(m);
    return rc;
  }
  *out = m;
  return SPM_OK;
}

int spm_mlp_destroy(spm_mlp* m) {
  if (!m) return SPM_OK;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  (void)hipFree(m->p);
  (void)hipFree(m->wb);
  (void)hipFree(m->grad);
  (void)hipFree(m->m1);
  (void)hipFree(m->m2);
  (void)hipFree(m->x);
  (void)hipFree(m->y);
  (void)hipFree(m->out);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return SPM_OK;
}

int spm_mlp_num_params(void) { return kNParams; }

int spm_mlp_set_params(spm_mlp* m, const float* params) {
  if (!m || !params) return set_error(SPM_EINVAL, "spm_mlp_set_params: NULL argument");
  SPM_CHECK_HIP(hipMemcpyAsync(m->p, params, sizeof(float) * kNParams, hipMemcpyHostToDevice, m->stream));
  hipLaunchKernelGGL(mlp_refresh_bf16, dim3((kNParams + 255) / 256), dim3(256), 0, m->stream, (const float*)m->p, m->wb);
  hipLaunchKernelGGL(mlp_zero, dim3((kNParams + 255) / 256), dim3(256), 0, m->stream, m->m1, kNParams);
  hipLaunchKernelGGL(mlp_zero, dim3((kNParams + 255) / 256), dim3(256), 0, m->stream, m->m2, kNParams);
  SPM_CHECK_HIP(hipGetLastError());
  SPM_CHECK_HIP(hipStreamSynchronize(m->stream));
  m->step = 0;
  return SPM_OK;
}

int spm_mlp_get_params(spm_mlp* m, float* params) {
  if (!m || !params) return set_error(SPM_EINVAL, "spm_mlp_get_params: NULL argument");
  SPM_CHECK_HIP(hipMemcpyAsync(params, m->p, sizeof(float) * kNParams, hipMemcpyDeviceToHost, m->stream));
  SPM_CHECK_HIP(hipStreamSynchronize(m->stream));
  return SPM_OK;
}

int spm_mlp_forward(spm_mlp* m, const float* x, int32_t n, float* logits) {
  if (!m || (n > 0 && (!x || !logits)) || n < 0) return set_error(SPM_EINVAL, "spm_mlp_forward: bad argument");
  if (n == 0) return SPM_OK;
  SPM_TRY(mlp_reserve(m, n));
  SPM_CHECK_HIP(hipMemcpyAsync(m->x, x, sizeof(float) * size_t(n) * kIn, hipMemcpyHostToDevice, m->stream));
  hipLaunchKernelGGL(mlp_forward, dim3((n + kMlpChunk - 1) / kMlpChunk), dim3(kH), 0, m->stream,
                     (const float*)m->x, n, (const uint16_t*)m->wb, (const float*)m->p, m->out);
  SPM_CHECK_HIP(hipGetLastError());
  SPM_CHECK_HIP(hipMemcpyAsync(logits, m->out, sizeof(float) * size_t(n) * kOut, hipMemcpyDeviceToHost, m->stream));
  SPM_CHECK_HIP(hipStreamSynchronize(m->stream));
  return SPM_OK;
}

int spm_mlp_predict(spm_mlp* m, const float* x, int32_t n, int32_t* cls) {
  if (!cls && n > 0) return set_error(SPM_EINVAL, "spm_mlp_predict: cls is NULL");
  std::vector<float> lg(size_t(n > 0 ? n : 1) * kOut);
  SPM_TRY(spm_mlp_forward(m, x, n, lg.data()));
  for (int32_t s = 0; s < n; ++s) {
    int best = 0;
    for (int k = 1; k < kOut; ++k)
      if (lg[size_t(s) * kOut + k] > lg[size_t(s) * kOut + best]) best = k;
    cls[s] = best;
  }
  return SPM_OK;
}

int spm_mlp_grad(spm_mlp* m, const float* x, const int32_t* y, int32_t n, float* loss, float* grads) {
  if (!m || !x || !y || n <= 0 || !grads) return set_error(SPM_EINVAL, "spm_mlp_grad: bad argument");
  for (int32_t s = 0; s < n; ++s)
    if (y[s] < 0 || y[s] >= kOut) return set_error(SPM_EINVAL, "spm_mlp_grad: label out of range");
  SPM_TRY(mlp_reserve(m, n));
  SPM_CHECK_HIP(hipMemcpyAsync(m->x, x, sizeof(float) * size_t(n) * kIn, hipMemcpyHostToDevice, m->stream));
  SPM_CHECK_HIP(hipMemcpyAsync(m->y, y, sizeof(int32_t) * size_t(n), hipMemcpyHostToDevice, m->stream));
  SPM_TRY(mlp_grad_dev(m, n, 0, n));
  std::vector<float> g(kNParams + 1);
  SPM_CHECK_HIP(hipMemcpyAsync(g.data(), m->grad, sizeof(float) * (kNParams + 1), hipMemcpyDeviceToHost, m->stream));
  SPM_CHECK_HIP(hipStreamSynchronize(m->stream));
  std::memcpy(grads, g.data(), sizeof(float) * kNParams);
  if (loss) *loss = g[kNParams] / float(n);
  return SPM_OK;
}

int spm_mlp_train(spm_mlp* m, const float* x, const int32_t* y, int32_t n, int32_t epochs, int32_t batch,
                  float lr, uint64_t seed, float* final_loss) {
  if (!m || !x || !y || n <= 0 || epochs < 0 || batch <= 0) return set_error(SPM_EINVAL, "spm_mlp_train: bad argument");
  for (int32_t s = 0; s < n; ++s)
    if (y[s] < 0 || y[s] >= kOut) return set_error(SPM_EINVAL, "spm_mlp_train: label out of range");
  SPM_TRY(mlp_reserve(m, n));
  // Each epoch visits the samples in a seeded permutation, staged once per epoch.
  std::vector<int32_t> perm(n);
  std::vector<float> xs(size_t(n) * kIn);
  std::vector<int32_t> ys(n);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  uint64_t st = seed ^ 0xA0761D6478BD642Full;
  for (int32_t ep = 0; ep < epochs; ++ep) {
    for (int32_t i = 0; i < n; ++i) perm[i] = i;
    for (int32_t i = n - 1; i > 0; --i) {
      st = st * 6364136223846793005ull + 1442695040888963407ull;
      const int32_t k = int32_t((st >> 33) % uint64_t(i + 1));
      std::swap(perm[i], perm[k]);
    }
    for (int32_t i = 0; i < n; ++i) {
      std::memcpy(&xs[size_t(i) * kIn], &x[size_t(perm[i]) * kIn], sizeof(float) * kIn);
      ys[i] = y[perm[i]];
    }
    SPM_CHECK_HIP(hipMemcpyAsync(m->x, xs.data(), sizeof(float) * xs.size(), hipMemcpyHostToDevice, m->stream));
    SPM_CHECK_HIP(hipMemcpyAsync(m->y, ys.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, m->stream));
    for (int32_t off = 0; off < n; off += batch) {
      const int cnt = std::min(batch, n - off);
      SPM_TRY(mlp_grad_dev(m, n, off, cnt));
      ++m->step;
      const float c1 = 1.f - std::pow(b1, float(m->step)), c2 = 1.f - std::pow(b2, float(m->step));
      hipLaunchKernelGGL(mlp_adam, dim3((kNParams + 255) / 256), dim3(256), 0, m->stream, m->p, m->wb,
                         (const float*)m->grad, m->m1, m->m2, lr, b1, b2, eps, c1, c2);
      SPM_CHECK_HIP(hipGetLastError());
    }
    SPM_CHECK_HIP(hipStreamSynchronize(m->stream));
  }
  if (final_loss) {
    std::vector<float> g(kNParams);
    // loss of the trained network on the full set
    SPM_CHECK_HIP(hipMemcpyAsync(m->x, x, sizeof(float) * size_t(n) * kIn, hipMemcpyHostToDevice, m->stream));
    SPM_CHECK_HIP(hipMemcpyAsync(m->y, y, sizeof(int32_t) * n, hipMemcpyHostToDevice, m->stream));
    SPM_TRY(mlp_grad_dev(m, n, 0, n));
    float ls = 0.f;
    SPM_CHECK_HIP(hipMemcpyAsync(&ls, m->grad + kNParams, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    SPM_CHECK_HIP(hipStreamSynchronize(m->stream));
    *final_loss = ls / float(n);
  }
  return SPM_OK;
}

}  // extern "C"
