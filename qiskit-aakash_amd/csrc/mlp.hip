// mlp.hip -- forward, backward and Adam kernels of the fusion cost model: a
// 3-layer perceptron x -> relu(x W1 + b1) -> relu(h1 W2 + b2) -> h2 W3 + b3,
// FP32, trained on the GPU with mean squared error.  Parameter vector layout:
//   W1 [F x H] row-major, b1 [H], W2 [H x H], b2 [H], W3 [H], b3 [1].

#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "dmsim.h"

namespace dmsim_internal {
int set_error(int code, const char *msg);  // dmsim.hip: the message dm_last_error returns
}

namespace {

constexpr int kT = 256;

int mfail(int code, const char *fmt, ...) {
  char msg[256];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(msg, sizeof(msg), fmt, ap);
  va_end(ap);
  return dmsim_internal::set_error(code, msg);
}

#define MLP_TRY(expr)                                                               \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess) return mfail(DM_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

struct Layout {
  int F, H;
  size_t w1, b1, w2, b2, w3, b3, total;
};

Layout layout(int F, int H) {
  Layout L;
  L.F = F;
  L.H = H;
  L.w1 = 0;
  L.b1 = L.w1 + (size_t)F * H;
  L.w2 = L.b1 + H;
  L.b2 = L.w2 + (size_t)H * H;
  L.w3 = L.b2 + H;
  L.b3 = L.w3 + H;
  L.total = L.b3 + 1;
  return L;
}

// Y[b, o] = act(sum_i X[b, i] W[i, o] + bias[o])
__global__ void k_dense(const float *__restrict__ X, const float *__restrict__ W, const float *__restrict__ bias,
                        float *__restrict__ Y, int B, int I, int O, int relu) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= B * O) return;
  const int b = idx / O, o = idx % O;
  float s = bias[o];
  for (int i = 0; i < I; ++i) s += X[(size_t)b * I + i] * W[(size_t)i * O + o];
  Y[idx] = relu ? fmaxf(s, 0.f) : s;
}

// dZ = dY * relu'(Y) (in place when relu), Y the activated output
__global__ void k_relu_grad(const float *__restrict__ Y, float *__restrict__ dY, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && Y[i] <= 0.f) dY[i] = 0.f;
}

// dW[i, o] = sum_b X[b, i] dZ[b, o]; db[o] = sum_b dZ[b, o]
__global__ void k_dense_wgrad(const float *__restrict__ X, const float *__restrict__ dZ, float *__restrict__ dW,
                              float *__restrict__ db, int B, int I, int O) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < I * O) {
    const int i = idx / O, o = idx % O;
    float s = 0.f;
    for (int b = 0; b < B; ++b) s += X[(size_t)b * I + i] * dZ[(size_t)b * O + o];
    dW[idx] = s;
  }
  if (idx < O) {
    float s = 0.f;
    for (int b = 0; b < B; ++b) s += dZ[(size_t)b * O + idx];
    db[idx] = s;
  }
}

// dX[b, i] = sum_o dZ[b, o] W[i, o]
__global__ void k_dense_xgrad(const float *__restrict__ dZ, const float *__restrict__ W, float *__restrict__ dX,
                              int B, int I, int O) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= B * I) return;
  const int b = idx / I, i = idx % I;
  float s = 0.f;
  for (int o = 0; o < O; ++o) s += dZ[(size_t)b * O + o] * W[(size_t)i * O + o];
  dX[idx] = s;
}

// d(loss)/d(pred) for loss = mean (pred - y)^2, and the loss itself (one block)
__global__ void k_mse(const float *__restrict__ pred, const float *__restrict__ y, float *__restrict__ dpred,
                      float *__restrict__ loss, int B) {
  __shared__ float part[kT];
  float s = 0.f;
  for (int b = threadIdx.x; b < B; b += kT) {
    const float d = pred[b] - y[b];
    dpred[b] = 2.f * d / B;
    s += d * d;
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int off = kT / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) part[threadIdx.x] += part[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[0] = part[0] / B;
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void k_gather_rows(const float *__restrict__ src, const int *__restrict__ idx, float *__restrict__ dst,
                              int rows, int cols) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows * cols) return;
  const int r = t / cols, c = t % cols;
  dst[t] = src[(size_t)idx[r] * cols + c];
}

inline int blocks(size_t n) { return (int)((n + kT - 1) / kT); }

// Device workspace for one batch size.
struct Net {
  Layout L;
  int B;
  float *p, *g, *h1, *h2, *out, *d1, *d2, *dout, *loss;
};

int net_alloc(Net *net, int F, int H, int B) {
  net->L = layout(F, H);
  net->B = B;
  const size_t P = net->L.total;
  MLP_TRY(hipMalloc(&net->p, P * sizeof(float)));
  MLP_TRY(hipMalloc(&net->g, P * sizeof(float)));
  MLP_TRY(hipMalloc(&net->h1, (size_t)B * H * sizeof(float)));
  MLP_TRY(hipMalloc(&net->h2, (size_t)B * H * sizeof(float)));
  MLP_TRY(hipMalloc(&net->out, (size_t)B * sizeof(float)));
  MLP_TRY(hipMalloc(&net->d1, (size_t)B * H * sizeof(float)));
  MLP_TRY(hipMalloc(&net->d2, (size_t)B * H * sizeof(float)));
  MLP_TRY(hipMalloc(&net->dout, (size_t)B * sizeof(float)));
  MLP_TRY(hipMalloc(&net->loss, sizeof(float)));
  return DM_OK;
}

void net_free(Net *net) {
  for (float *x : {net->p, net->g, net->h1, net->h2, net->out, net->d1, net->d2, net->dout, net->loss})
    if (x) hipFree(x);
}

void forward(const Net &n, const float *X, int B) {
  const Layout &L = n.L;
  hipLaunchKernelGGL(k_dense, dim3(blocks((size_t)B * L.H)), dim3(kT), 0, 0, X, n.p + L.w1, n.p + L.b1, n.h1, B, L.F,
                     L.H, 1);
  hipLaunchKernelGGL(k_dense, dim3(blocks((size_t)B * L.H)), dim3(kT), 0, 0, n.h1, n.p + L.w2, n.p + L.b2, n.h2, B,
                     L.H, L.H, 1);
  hipLaunchKernelGGL(k_dense, dim3(blocks(B)), dim3(kT), 0, 0, n.h2, n.p + L.w3, n.p + L.b3, n.out, B, L.H, 1, 0);
}

// gradient of the mean squared error over the batch into n.g; loss into n.loss
void backward(const Net &n, const float *X, const float *y, int B) {
  const Layout &L = n.L;
  hipLaunchKernelGGL(k_mse, dim3(1), dim3(kT), 0, 0, n.out, y, n.dout, n.loss, B);
  // layer 3
  hipLaunchKernelGGL(k_dense_wgrad, dim3(blocks(std::max(L.H, 1))), dim3(kT), 0, 0, n.h2, n.dout, n.g + L.w3,
                     n.g + L.b3, B, L.H, 1);
  hipLaunchKernelGGL(k_dense_xgrad, dim3(blocks((size_t)B * L.H)), dim3(kT), 0, 0, n.dout, n.p + L.w3, n.d2, B, L.H, 1);
  hipLaunchKernelGGL(k_relu_grad, dim3(blocks((size_t)B * L.H)), dim3(kT), 0, 0, n.h2, n.d2, B * L.H);
  // layer 2
  hipLaunchKernelGGL(k_dense_wgrad, dim3(blocks((size_t)L.H * L.H)), dim3(kT), 0, 0, n.h1, n.d2, n.g + L.w2,
                     n.g + L.b2, B, L.H, L.H);
  hipLaunchKernelGGL(k_dense_xgrad, dim3(blocks((size_t)B * L.H)), dim3(kT), 0, 0, n.d2, n.p + L.w2, n.d1, B, L.H,
                     L.H);
  hipLaunchKernelGGL(k_relu_grad, dim3(blocks((size_t)B * L.H)), dim3(kT), 0, 0, n.h1, n.d1, B * L.H);
  // layer 1
  hipLaunchKernelGGL(k_dense_wgrad, dim3(blocks((size_t)L.F * L.H)), dim3(kT), 0, 0, X, n.d1, n.g + L.w1, n.g + L.b1,
                     B, L.F, L.H);
}

uint64_t splitmix(uint64_t *s) {
  uint64_t z = (*s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

double uniform01(uint64_t *s) { return ((splitmix(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

}  // namespace

extern "C" {

int dm_mlp_param_count(int n_features, int hidden) {
  if (n_features < 1 || hidden < 1) return mfail(DM_ERR_ARG, "bad network size");
  return (int)layout(n_features, hidden).total;
}

int dm_mlp_init(int n_features, int hidden, uint64_t seed, float *params) {
  if (!params || n_features < 1 || hidden < 1) return mfail(DM_ERR_ARG, "bad arguments");
  const Layout L = layout(n_features, hidden);
  uint64_t s = seed;
  auto normal = [&]() {
    const double u1 = uniform01(&s), u2 = uniform01(&s);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
  };
  for (size_t i = 0; i < L.total; ++i) params[i] = 0.f;
  for (size_t i = 0; i < (size_t)n_features * hidden; ++i) params[L.w1 + i] = (float)(normal() * sqrt(2.0 / n_features));
  for (size_t i = 0; i < (size_t)hidden * hidden; ++i) params[L.w2 + i] = (float)(normal() * sqrt(2.0 / hidden));
  for (int i = 0; i < hidden; ++i) params[L.w3 + i] = (float)(normal() * sqrt(1.0 / hidden));
  return DM_OK;
}

int dm_mlp_predict(const float *params, int n_features, int hidden, const float *X, int n, float *out) {
  if (!params || !X || !out || n < 0) return mfail(DM_ERR_ARG, "bad arguments");
  if (n == 0) return DM_OK;
  Net net;
  memset(&net, 0, sizeof(net));
  int rc = net_alloc(&net, n_features, hidden, n);
  float *dX = nullptr;
  if (rc == DM_OK) {
    hipError_t e = hipMalloc(&dX, (size_t)n * n_features * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(net.p, params, net.L.total * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dX, X, (size_t)n * n_features * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      forward(net, dX, n);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, net.out, (size_t)n * sizeof(float), hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = mfail(DM_ERR_HIP, "mlp predict: %s", hipGetErrorString(e));
  }
  if (dX) hipFree(dX);
  net_free(&net);
  return rc;
}

int dm_mlp_loss_grad(const float *params, int n_features, int hidden, const float *X, const float *y, int n,
                     float *loss, float *grad) {
  if (!params || !X || !y || n < 1) return mfail(DM_ERR_ARG, "bad arguments");
  Net net;
  memset(&net, 0, sizeof(net));
  int rc = net_alloc(&net, n_features, hidden, n);
  float *dX = nullptr, *dy = nullptr;
  if (rc == DM_OK) {
    hipError_t e = hipMalloc(&dX, (size_t)n * n_features * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&dy, (size_t)n * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(net.p, params, net.L.total * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dX, X, (size_t)n * n_features * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dy, y, (size_t)n * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      forward(net, dX, n);
      backward(net, dX, dy, n);
      e = hipGetLastError();
    }
    if (e == hipSuccess && loss) e = hipMemcpy(loss, net.loss, sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess && grad) e = hipMemcpy(grad, net.g, net.L.total * sizeof(float), hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = mfail(DM_ERR_HIP, "mlp grad: %s", hipGetErrorString(e));
  }
  if (dX) hipFree(dX);
  if (dy) hipFree(dy);
  net_free(&net);
  return rc;
}

int dm_mlp_train(float *params, int n_features, int hidden, const float *X, const float *y, int n, int epochs,
                 int batch, float lr, uint64_t seed, float *loss_per_epoch) {
  if (!params || !X || !y || n < 1 || epochs < 0 || batch < 1) return mfail(DM_ERR_ARG, "bad arguments");
  if (batch > n) batch = n;
  Net net;
  memset(&net, 0, sizeof(net));
  int rc = net_alloc(&net, n_features, hidden, batch);
  float *dXall = nullptr, *dyall = nullptr, *dXb = nullptr, *dyb = nullptr, *m = nullptr, *v = nullptr;
  int *didx = nullptr;
  const size_t P = net.L.total;
  hipError_t e = hipSuccess;
  if (rc == DM_OK) {
    e = hipMalloc(&dXall, (size_t)n * n_features * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&dyall, (size_t)n * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&dXb, (size_t)batch * n_features * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&dyb, (size_t)batch * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&m, P * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&v, P * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&didx, (size_t)batch * sizeof(int));
    if (e == hipSuccess) e = hipMemset(m, 0, P * sizeof(float));
    if (e == hipSuccess) e = hipMemset(v, 0, P * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(net.p, params, P * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dXall, X, (size_t)n * n_features * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dyall, y, (size_t)n * sizeof(float), hipMemcpyHostToDevice);
  }
  if (rc == DM_OK && e == hipSuccess) {
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    uint64_t s = seed;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    int t = 0;
    for (int ep = 0; ep < epochs && e == hipSuccess; ++ep) {
      for (int i = n - 1; i > 0; --i) {  // Fisher-Yates
        const int j = (int)(splitmix(&s) % (uint64_t)(i + 1));
        std::swap(order[i], order[j]);
      }
      double ep_loss = 0.0;
      int nb = 0;
      for (int start = 0; start + batch <= n && e == hipSuccess; start += batch) {
        e = hipMemcpy(didx, order.data() + start, (size_t)batch * sizeof(int), hipMemcpyHostToDevice);
        if (e != hipSuccess) break;
        hipLaunchKernelGGL(k_gather_rows, dim3(blocks((size_t)batch * n_features)), dim3(kT), 0, 0, dXall, didx, dXb,
                           batch, n_features);
        hipLaunchKernelGGL(k_gather_rows, dim3(blocks(batch)), dim3(kT), 0, 0, dyall, didx, dyb, batch, 1);
        forward(net, dXb, batch);
        backward(net, dXb, dyb, batch);
        ++t;
        const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
        hipLaunchKernelGGL(k_adam, dim3(blocks(P)), dim3(kT), 0, 0, net.p, net.g, m, v, (int)P, lr, b1, b2, eps, c1, c2);
        float l = 0.f;
        e = hipMemcpy(&l, net.loss, sizeof(float), hipMemcpyDeviceToHost);
        ep_loss += l;
        ++nb;
      }
      if (loss_per_epoch) loss_per_epoch[ep] = nb ? (float)(ep_loss / nb) : 0.f;
    }
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(params, net.p, P * sizeof(float), hipMemcpyDeviceToHost);
  }
  if (e != hipSuccess && rc == DM_OK) rc = mfail(DM_ERR_HIP, "mlp train: %s", hipGetErrorString(e));
  for (void *x : {(void *)dXall, (void *)dyall, (void *)dXb, (void *)dyb, (void *)m, (void *)v, (void *)didx})
    if (x) hipFree(x);
  net_free(&net);
  return rc;
}

}  // extern "C"
