// The scheduling network: a multilayer perceptron 8 -> 64 -> 64 -> 2 (ReLU) that predicts a DP
// call's band extent and cell count from features known before the call runs. Weights are
// stored in FP16; every product is accumulated in FP32. Forward, backward and the Adam update
// are this file's own kernels; the host side (C ABI in include/mpoa.h) drives training and
// inference. The predictions only schedule work; no result depends on them.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <cmath>
#include <string>
#include <vector>

#include "mpoa.h"

namespace {

constexpr int IN = 8, H = 64, OUT = 2;
// Parameter layout (row-major, input index first): W1[IN][H] b1[H] W2[H][H] b2[H] W3[H][OUT] b3[OUT]
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + IN * H, OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + H * H,
              OFF_W3 = OFF_B2 + H, OFF_B3 = OFF_W3 + H * OUT, N_PARAM = OFF_B3 + OUT;
constexpr int TILE = 64;  // samples per workgroup in the gradient kernel

__device__ __forceinline__ void load_weights(const __half* w, float* Wf) {
  for (int p = threadIdx.x; p < N_PARAM; p += blockDim.x) Wf[p] = __half2float(w[p]);
  __syncthreads();
}

// One sample per thread.
__global__ void __launch_bounds__(256) k_mlp_forward(const float* __restrict__ X, int64_t n,
                                                     const __half* __restrict__ w,
                                                     float* __restrict__ Y) {
  __shared__ float Wf[N_PARAM];
  load_weights(w, Wf);
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float x[IN];
#pragma unroll
  for (int k = 0; k < IN; k++) x[k] = X[s * IN + k];
  float h1[H];
#pragma unroll
  for (int i = 0; i < H; i++) {
    float a = Wf[OFF_B1 + i];
#pragma unroll
    for (int k = 0; k < IN; k++) a = fmaf(x[k], Wf[OFF_W1 + k * H + i], a);
    h1[i] = fmaxf(a, 0.f);
  }
  float y0 = Wf[OFF_B3], y1 = Wf[OFF_B3 + 1];
  for (int j = 0; j < H; j++) {
    float a = Wf[OFF_B2 + j];
#pragma unroll
    for (int i = 0; i < H; i++) a = fmaf(h1[i], Wf[OFF_W2 + i * H + j], a);
    a = fmaxf(a, 0.f);
    y0 = fmaf(a, Wf[OFF_W3 + j * OUT], y0);
    y1 = fmaf(a, Wf[OFF_W3 + j * OUT + 1], y1);
  }
  Y[s * OUT] = y0;
  Y[s * OUT + 1] = y1;
}

// Forward and backward of TILE samples per workgroup for the loss
// L = 1/(2B) * sum over samples and outputs of (y - t)^2 (B = batch). Writes this tile's
// gradient contribution to partial[blockIdx.x * N_PARAM ...] and its squared error to
// sq[blockIdx.x]; a second kernel sums the tiles in a fixed order.
__global__ void __launch_bounds__(256) k_mlp_grad(const float* __restrict__ X,
                                                  const float* __restrict__ T, int64_t n,
                                                  float inv_batch, const __half* __restrict__ w,
                                                  float* __restrict__ partial,
                                                  float* __restrict__ sq) {
  __shared__ float Wf[N_PARAM];
  __shared__ float Xs[TILE][IN];
  __shared__ float H1[TILE][H], H2[TILE][H], D1[TILE][H], D2[TILE][H];
  __shared__ float dY[TILE][OUT];
  __shared__ float red[256];
  load_weights(w, Wf);
  const int tid = threadIdx.x;
  const int64_t s0 = (int64_t)blockIdx.x * TILE;
  for (int e = tid; e < TILE * IN; e += 256) {
    const int s = e / IN, k = e % IN;
    Xs[s][k] = s0 + s < n ? X[(s0 + s) * IN + k] : 0.f;
  }
  __syncthreads();
  for (int e = tid; e < TILE * H; e += 256) {
    const int s = e / H, i = e % H;
    float a = Wf[OFF_B1 + i];
    for (int k = 0; k < IN; k++) a = fmaf(Xs[s][k], Wf[OFF_W1 + k * H + i], a);
    H1[s][i] = fmaxf(a, 0.f);
  }
  __syncthreads();
  for (int e = tid; e < TILE * H; e += 256) {
    const int s = e / H, j = e % H;
    float a = Wf[OFF_B2 + j];
    for (int i = 0; i < H; i++) a = fmaf(H1[s][i], Wf[OFF_W2 + i * H + j], a);
    H2[s][j] = fmaxf(a, 0.f);
  }
  __syncthreads();
  float err = 0.f;
  if (tid < TILE * OUT) {
    const int s = tid / OUT, o = tid % OUT;
    float y = Wf[OFF_B3 + o];
    for (int j = 0; j < H; j++) y = fmaf(H2[s][j], Wf[OFF_W3 + j * OUT + o], y);
    const bool valid = s0 + s < n;
    const float d = valid ? y - T[(s0 + s) * OUT + o] : 0.f;
    dY[s][o] = d * inv_batch;
    err = d * d;
  }
  red[tid] = err;
  __syncthreads();
  for (int e = tid; e < TILE * H; e += 256) {
    const int s = e / H, j = e % H;
    const float g = dY[s][0] * Wf[OFF_W3 + j * OUT] + dY[s][1] * Wf[OFF_W3 + j * OUT + 1];
    D2[s][j] = H2[s][j] > 0.f ? g : 0.f;
  }
  __syncthreads();
  for (int e = tid; e < TILE * H; e += 256) {
    const int s = e / H, i = e % H;
    float g = 0.f;
    for (int j = 0; j < H; j++) g = fmaf(D2[s][j], Wf[OFF_W2 + i * H + j], g);
    D1[s][i] = H1[s][i] > 0.f ? g : 0.f;
  }
  __syncthreads();
  float* out = partial + (int64_t)blockIdx.x * N_PARAM;
  for (int p = tid; p < N_PARAM; p += 256) {
    float g = 0.f;
    if (p < OFF_B1) {  // W1[k][i]
      const int k = p / H, i = p % H;
      for (int s = 0; s < TILE; s++) g = fmaf(Xs[s][k], D1[s][i], g);
    } else if (p < OFF_W2) {
      const int i = p - OFF_B1;
      for (int s = 0; s < TILE; s++) g += D1[s][i];
    } else if (p < OFF_B2) {  // W2[i][j]
      const int q = p - OFF_W2, i = q / H, j = q % H;
      for (int s = 0; s < TILE; s++) g = fmaf(H1[s][i], D2[s][j], g);
    } else if (p < OFF_W3) {
      const int j = p - OFF_B2;
      for (int s = 0; s < TILE; s++) g += D2[s][j];
    } else if (p < OFF_B3) {  // W3[j][o]
      const int q = p - OFF_W3, j = q / OUT, o = q % OUT;
      for (int s = 0; s < TILE; s++) g = fmaf(H2[s][j], dY[s][o], g);
    } else {
      const int o = p - OFF_B3;
      for (int s = 0; s < TILE; s++) g += dY[s][o];
    }
    out[p] = g;
  }
  for (int st = 128; st > 0; st >>= 1) {
    __syncthreads();
    if (tid < st) red[tid] += red[tid + st];
  }
  if (tid == 0) sq[blockIdx.x] = red[0];
}

// Sums the per-tile gradients in tile order (deterministic).
__global__ void k_mlp_reduce(const float* __restrict__ partial, int tiles,
                             float* __restrict__ grad, const float* __restrict__ sq,
                             float* __restrict__ loss, float inv_2batch) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < N_PARAM) {
    float g = 0.f;
    for (int t = 0; t < tiles; t++) g += partial[(int64_t)t * N_PARAM + p];
    grad[p] = g;
  }
  if (p == 0 && loss) {
    float l = 0.f;
    for (int t = 0; t < tiles; t++) l += sq[t];
    *loss = l * inv_2batch;
  }
}

// Adam on FP32 master weights; the FP16 working copy is refreshed from them.
__global__ void k_mlp_adam(float* __restrict__ master, float* __restrict__ m1,
                           float* __restrict__ m2, const float* __restrict__ grad,
                           __half* __restrict__ w, float lr, float b1, float b2, float eps,
                           float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N_PARAM) return;
  const float g = grad[p];
  const float a = b1 * m1[p] + (1.f - b1) * g;
  const float v = b2 * m2[p] + (1.f - b2) * g * g;
  m1[p] = a;
  m2[p] = v;
  const float x = master[p] - lr * (a / c1) / (sqrtf(v / c2) + eps);
  master[p] = x;
  w[p] = __float2half(x);
}

thread_local std::string g_mlp_err;

#define MLPCHK(expr)                                                          \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess) {                                                   \
      g_mlp_err = std::string(#expr) + ": " + hipGetErrorString(e_);          \
      return e_ == hipErrorOutOfMemory ? MPOA_E_OOM : MPOA_E_DEVICE;          \
    }                                                                         \
  } while (0)

template <class T>
struct Dev {
  T* p = nullptr;
  size_t n = 0;
  ~Dev() {
    if (p) (void)hipFree(p);
  }
  hipError_t ensure(size_t count) {
    if (count <= n) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
    if (e == hipSuccess) n = count;
    return e;
  }
};

uint64_t mix64(uint64_t& s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

}  // namespace

struct mpoa_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  Dev<__half> w;
  Dev<float> master, m1, m2, grad, partial, sq, loss;
  Dev<float> X, T, Y;
  int64_t step = 0;
};

extern "C" {

int mpoa_mlp_create(int device, uint64_t seed, mpoa_mlp** out) {
  if (!out) return MPOA_E_ARG;
  *out = nullptr;
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || device < 0 || device >= nd) return MPOA_E_DEVICE;
  MLPCHK(hipSetDevice(device));
  mpoa_mlp* m = new mpoa_mlp();
  m->device = device;
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess ||
      m->w.ensure(N_PARAM) != hipSuccess || m->master.ensure(N_PARAM) != hipSuccess ||
      m->m1.ensure(N_PARAM) != hipSuccess || m->m2.ensure(N_PARAM) != hipSuccess ||
      m->grad.ensure(N_PARAM) != hipSuccess || m->loss.ensure(1) != hipSuccess) {
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return MPOA_E_OOM;
  }
  // He-uniform initialisation from a seeded generator; biases zero.
  std::vector<float> init(N_PARAM, 0.f);
  uint64_t st = seed;
  auto fill = [&](int off, int fan_in, int count) {
    const float a = std::sqrt(6.f / (float)fan_in);
    for (int i = 0; i < count; i++) {
      const double u = (double)(mix64(st) >> 11) * (1.0 / 9007199254740992.0);
      init[off + i] = (float)((2.0 * u - 1.0) * a);
    }
  };
  fill(OFF_W1, IN, IN * H);
  fill(OFF_W2, H, H * H);
  fill(OFF_W3, H, H * OUT);
  std::vector<__half> h(N_PARAM);
  for (int p = 0; p < N_PARAM; p++) {
    h[p] = __float2half(init[p]);
    init[p] = __half2float(h[p]);
  }
  std::vector<float> zero(N_PARAM, 0.f);
  if (hipMemcpy(m->w.p, h.data(), N_PARAM * sizeof(__half), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->master.p, init.data(), N_PARAM * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->m1.p, zero.data(), N_PARAM * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->m2.p, zero.data(), N_PARAM * 4, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipStreamDestroy(m->stream);
    delete m;
    return MPOA_E_DEVICE;
  }
  *out = m;
  return MPOA_OK;
}

void mpoa_mlp_destroy(mpoa_mlp* m) {
  if (!m) return;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

int32_t mpoa_mlp_n_params(void) { return N_PARAM; }

int mpoa_mlp_get_weights(mpoa_mlp* m, uint16_t* out, int32_t cap) {
  if (!m || !out || cap < N_PARAM) return MPOA_E_ARG;
  MLPCHK(hipSetDevice(m->device));
  MLPCHK(hipMemcpy(out, m->w.p, N_PARAM * sizeof(__half), hipMemcpyDeviceToHost));
  return MPOA_OK;
}

int mpoa_mlp_set_weights(mpoa_mlp* m, const uint16_t* in, int32_t n) {
  if (!m || !in || n != N_PARAM) return MPOA_E_ARG;
  MLPCHK(hipSetDevice(m->device));
  std::vector<float> f(N_PARAM), zero(N_PARAM, 0.f);
  for (int p = 0; p < N_PARAM; p++) {
    __half_raw r;
    r.x = in[p];
    f[p] = __half2float(__half(r));
  }
  MLPCHK(hipMemcpy(m->w.p, in, N_PARAM * sizeof(__half), hipMemcpyHostToDevice));
  MLPCHK(hipMemcpy(m->master.p, f.data(), N_PARAM * 4, hipMemcpyHostToDevice));
  MLPCHK(hipMemcpy(m->m1.p, zero.data(), N_PARAM * 4, hipMemcpyHostToDevice));
  MLPCHK(hipMemcpy(m->m2.p, zero.data(), N_PARAM * 4, hipMemcpyHostToDevice));
  m->step = 0;
  return MPOA_OK;
}

int mpoa_mlp_forward(mpoa_mlp* m, const float* X, int64_t n, float* Y) {
  if (!m || n < 0 || (n > 0 && (!X || !Y))) return MPOA_E_ARG;
  if (n == 0) return MPOA_OK;
  MLPCHK(hipSetDevice(m->device));
  MLPCHK(m->X.ensure((size_t)n * IN));
  MLPCHK(m->Y.ensure((size_t)n * OUT));
  MLPCHK(hipMemcpyAsync(m->X.p, X, (size_t)n * IN * 4, hipMemcpyHostToDevice, m->stream));
  const int blocks = (int)((n + 255) / 256);
  k_mlp_forward<<<blocks, 256, 0, m->stream>>>(m->X.p, n, m->w.p, m->Y.p);
  MLPCHK(hipGetLastError());
  MLPCHK(hipMemcpyAsync(Y, m->Y.p, (size_t)n * OUT * 4, hipMemcpyDeviceToHost, m->stream));
  MLPCHK(hipStreamSynchronize(m->stream));
  return MPOA_OK;
}

namespace {
// Gradient of the loss over samples [0, n) of the device arrays X, T (n <= batch).
int grad_launch(mpoa_mlp* m, const float* dX, const float* dT, int64_t n) {
  const int tiles = (int)((n + TILE - 1) / TILE);
  MLPCHK(m->partial.ensure((size_t)tiles * N_PARAM));
  MLPCHK(m->sq.ensure((size_t)tiles));
  k_mlp_grad<<<tiles, 256, 0, m->stream>>>(dX, dT, n, 1.f / (float)n, m->w.p, m->partial.p,
                                           m->sq.p);
  MLPCHK(hipGetLastError());
  k_mlp_reduce<<<(N_PARAM + 255) / 256, 256, 0, m->stream>>>(m->partial.p, tiles, m->grad.p,
                                                              m->sq.p, m->loss.p,
                                                              0.5f / (float)n);
  MLPCHK(hipGetLastError());
  return MPOA_OK;
}
}  // namespace

int mpoa_mlp_gradients(mpoa_mlp* m, const float* X, const float* T, int64_t n, float* grad,
                       float* loss) {
  if (!m || n <= 0 || !X || !T || !grad) return MPOA_E_ARG;
  MLPCHK(hipSetDevice(m->device));
  MLPCHK(m->X.ensure((size_t)n * IN));
  MLPCHK(m->T.ensure((size_t)n * OUT));
  MLPCHK(hipMemcpyAsync(m->X.p, X, (size_t)n * IN * 4, hipMemcpyHostToDevice, m->stream));
  MLPCHK(hipMemcpyAsync(m->T.p, T, (size_t)n * OUT * 4, hipMemcpyHostToDevice, m->stream));
  int rc = grad_launch(m, m->X.p, m->T.p, n);
  if (rc != MPOA_OK) return rc;
  MLPCHK(hipMemcpyAsync(grad, m->grad.p, N_PARAM * 4, hipMemcpyDeviceToHost, m->stream));
  float l = 0.f;
  MLPCHK(hipMemcpyAsync(&l, m->loss.p, 4, hipMemcpyDeviceToHost, m->stream));
  MLPCHK(hipStreamSynchronize(m->stream));
  if (loss) *loss = l;
  return MPOA_OK;
}

int mpoa_mlp_train(mpoa_mlp* m, const float* X, const float* T, int64_t n, int32_t epochs,
                   int32_t batch, float lr, uint64_t seed, float* epoch_loss) {
  if (!m || n <= 0 || !X || !T || epochs <= 0 || batch <= 0) return MPOA_E_ARG;
  MLPCHK(hipSetDevice(m->device));
  MLPCHK(m->X.ensure((size_t)n * IN));
  MLPCHK(m->T.ensure((size_t)n * OUT));
  std::vector<float> xs((size_t)n * IN), ts((size_t)n * OUT);
  std::vector<int64_t> perm(n);
  for (int64_t i = 0; i < n; i++) perm[i] = i;
  uint64_t st = seed;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int ep = 0; ep < epochs; ep++) {
    for (int64_t i = n - 1; i > 0; i--) {  // Fisher-Yates
      const int64_t j = (int64_t)(mix64(st) % (uint64_t)(i + 1));
      const int64_t t = perm[i];
      perm[i] = perm[j];
      perm[j] = t;
    }
    for (int64_t i = 0; i < n; i++) {
      for (int k = 0; k < IN; k++) xs[i * IN + k] = X[perm[i] * IN + k];
      for (int o = 0; o < OUT; o++) ts[i * OUT + o] = T[perm[i] * OUT + o];
    }
    MLPCHK(hipMemcpyAsync(m->X.p, xs.data(), xs.size() * 4, hipMemcpyHostToDevice, m->stream));
    MLPCHK(hipMemcpyAsync(m->T.p, ts.data(), ts.size() * 4, hipMemcpyHostToDevice, m->stream));
    double ep_loss = 0;
    int64_t nb = 0;
    for (int64_t s = 0; s + batch <= n || (s < n && s == 0); s += batch) {
      const int64_t cnt = std::min<int64_t>(batch, n - s);
      int rc = grad_launch(m, m->X.p + s * IN, m->T.p + s * OUT, cnt);
      if (rc != MPOA_OK) return rc;
      m->step++;
      const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
      k_mlp_adam<<<(N_PARAM + 255) / 256, 256, 0, m->stream>>>(m->master.p, m->m1.p, m->m2.p,
                                                                 m->grad.p, m->w.p, lr, b1, b2,
                                                                 eps, c1, c2);
      MLPCHK(hipGetLastError());
      if (epoch_loss) {
        float l = 0.f;
        MLPCHK(hipMemcpyAsync(&l, m->loss.p, 4, hipMemcpyDeviceToHost, m->stream));
        MLPCHK(hipStreamSynchronize(m->stream));
        ep_loss += l;
        nb++;
      }
    }
    if (epoch_loss) epoch_loss[ep] = nb ? (float)(ep_loss / nb) : 0.f;
  }
  MLPCHK(hipStreamSynchronize(m->stream));
  return MPOA_OK;
}

const char* mpoa_mlp_error(void) { return g_mlp_err.c_str(); }

}  // extern "C"
