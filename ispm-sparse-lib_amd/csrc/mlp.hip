/* mlp.hip -- the learned planner's network: 32 features -> 64 -> 64 -> (C plan scores, 1
 * log2-droptol), ReLU hidden layers, FP32, trained and evaluated on the GPU by the kernels
 * below.  Loss per sample: soft cross-entropy of the plan scores against a target
 * distribution over candidate formats (built from measured SpMV times) plus
 * drop_weight * (droptol output - target)^2; samples without a label in a head are masked.
 * Parameters are one flat array: W1[64x32] b1[64] W2[64x64] b2[64] W3[Ox64] b3[O], O = C+1.
 * Tiny dense layers (at most 64 wide): one thread per output element, FP32 accumulation.
 * No counterpart in the reference. */
#include <hip/hip_runtime.h>
#include <cmath>
#include <vector>
#include "internal.h"

namespace ispm {
namespace {

constexpr int kIn = 32, kH = 64;

struct mlp_dims {
  int O;                  /* outputs: C plan scores + 1 droptol */
  size_t w1, b1, w2, b2, w3, b3, total;
};

mlp_dims dims(int O)
{
  mlp_dims d;
  d.O = O;
  d.w1 = 0;
  d.b1 = d.w1 + (size_t)kH * kIn;
  d.w2 = d.b1 + kH;
  d.b2 = d.w2 + (size_t)kH * kH;
  d.w3 = d.b2 + kH;
  d.b3 = d.w3 + (size_t)O * kH;
  d.total = d.b3 + O;
  return d;
}

/* Y[b,o] = act(sum_i X[b,i] W[o,i] + bias[o]) */
__global__ void k_fwd(const float *__restrict__ X, const float *__restrict__ W, const float *__restrict__ bias,
                      float *__restrict__ Y, int B, int in, int out, int relu)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * out) return;
  const int b = t / out, o = t - b * out;
  float s = bias[o];
  for (int i = 0; i < in; i++) s += X[b * in + i] * W[o * in + i];
  Y[t] = relu ? fmaxf(s, 0.f) : s;
}

/* per-sample loss and gradient of the output layer */
__global__ void k_loss(const float *__restrict__ Z, const float *__restrict__ Tp, const float *__restrict__ Mp,
                       const float *__restrict__ Td, const float *__restrict__ Md, float *__restrict__ dZ,
                       float *__restrict__ lossv, int B, int C, float wdrop)
{
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const int O = C + 1;
  const float *z = Z + b * O;
  float *dz = dZ + b * O;
  float loss = 0.f;
  if (Mp[b] > 0.f) {
    float m = -INFINITY;
    for (int c = 0; c < C; c++) m = fmaxf(m, z[c]);
    float se = 0.f;
    for (int c = 0; c < C; c++) se += expf(z[c] - m);
    const float lse = m + logf(se);
    for (int c = 0; c < C; c++) {
      const float q = Tp[b * C + c];
      const float p = expf(z[c] - lse);
      dz[c] = (p - q) / B;
      loss -= q * (z[c] - lse);
    }
  } else {
    for (int c = 0; c < C; c++) dz[c] = 0.f;
  }
  if (Md[b] > 0.f) {
    const float e = z[C] - Td[b];
    dz[C] = 2.f * wdrop * e / B;
    loss += wdrop * e * e;
  } else {
    dz[C] = 0.f;
  }
  lossv[b] = loss / B;
}

/* dW[o,i] = sum_b dY[b,o] X[b,i];  db[o] = sum_b dY[b,o] */
__global__ void k_bwd_w(const float *__restrict__ dY, const float *__restrict__ X, float *__restrict__ dW,
                        float *__restrict__ db, int B, int in, int out)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < out * in) {
    const int o = t / in, i = t - o * in;
    float s = 0.f;
    for (int b = 0; b < B; b++) s += dY[b * out + o] * X[b * in + i];
    dW[t] = s;
  } else if (t < out * in + out) {
    const int o = t - out * in;
    float s = 0.f;
    for (int b = 0; b < B; b++) s += dY[b * out + o];
    db[o] = s;
  }
}

/* dX[b,i] = (sum_o dY[b,o] W[o,i]) * [Xact[b,i] > 0]  (ReLU of the layer below) */
__global__ void k_bwd_x(const float *__restrict__ dY, const float *__restrict__ W, const float *__restrict__ Xact,
                        float *__restrict__ dX, int B, int in, int out)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * in) return;
  const int b = t / in, i = t - b * in;
  float s = 0.f;
  for (int o = 0; o < out; o++) s += dY[b * out + o] * W[o * in + i];
  dX[t] = Xact[t] > 0.f ? s : 0.f;
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                       int n, float lr, float b1, float b2, float c1, float c2, float eps)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float gi = g[t];
  const float mi = b1 * m[t] + (1.f - b1) * gi;
  const float vi = b2 * v[t] + (1.f - b2) * gi * gi;
  m[t] = mi;
  v[t] = vi;
  p[t] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void k_sum(const float *__restrict__ a, int n, float *__restrict__ out)
{
  if (blockIdx.x || threadIdx.x) return;
  float s = 0.f;
  for (int i = 0; i < n; i++) s += a[i];
  *out = s;
}

/* workspace of activations and gradients, grown on demand */
struct workspace {
  float *h1 = nullptr, *h2 = nullptr, *z = nullptr, *dz = nullptr, *dh2 = nullptr, *dh1 = nullptr,
        *grad = nullptr, *lossv = nullptr;
  int B = 0, O = 0;
};
workspace g_ws;

int ws_ensure(int B, int O)
{
  if (g_ws.B >= B && g_ws.O == O) return ISPM_OK;
  float **ps[] = {&g_ws.h1, &g_ws.h2, &g_ws.z, &g_ws.dz, &g_ws.dh2, &g_ws.dh1, &g_ws.grad, &g_ws.lossv};
  for (float **p : ps) { ispm_free(*p); *p = nullptr; }
  const mlp_dims d = dims(O);
  int st;
  if ((st = ispm_malloc((void **)&g_ws.h1, sizeof(float) * B * kH)) ||
      (st = ispm_malloc((void **)&g_ws.h2, sizeof(float) * B * kH)) ||
      (st = ispm_malloc((void **)&g_ws.z, sizeof(float) * B * O)) ||
      (st = ispm_malloc((void **)&g_ws.dz, sizeof(float) * B * O)) ||
      (st = ispm_malloc((void **)&g_ws.dh2, sizeof(float) * B * kH)) ||
      (st = ispm_malloc((void **)&g_ws.dh1, sizeof(float) * B * kH)) ||
      (st = ispm_malloc((void **)&g_ws.grad, sizeof(float) * d.total)) ||
      (st = ispm_malloc((void **)&g_ws.lossv, sizeof(float) * B)))
    return st;
  g_ws.B = B;
  g_ws.O = O;
  return ISPM_OK;
}

inline dim3 blocks(long long n) { return dim3((unsigned)((n + 255) / 256)); }

int forward(const float *P, const float *X, int B, const mlp_dims &d, float *out)
{
  hipStream_t s = stream();
  hipLaunchKernelGGL(k_fwd, blocks((long long)B * kH), dim3(256), 0, s, X, P + d.w1, P + d.b1, g_ws.h1, B, kIn, kH, 1);
  hipLaunchKernelGGL(k_fwd, blocks((long long)B * kH), dim3(256), 0, s, (const float *)g_ws.h1, P + d.w2, P + d.b2, g_ws.h2, B, kH, kH, 1);
  hipLaunchKernelGGL(k_fwd, blocks((long long)B * d.O), dim3(256), 0, s, (const float *)g_ws.h2, P + d.w3, P + d.b3, out, B, kH, d.O, 0);
  ISPM_HIP(hipGetLastError());
  return ISPM_OK;
}

}  // namespace
}  // namespace ispm

using namespace ispm;

extern "C" {

int ispm_mlp_param_count(int n_plan_classes)
{
  return (int)dims(n_plan_classes + 1).total;
}

int ispm_mlp_forward(int n_plan_classes, const float *params, const float *X, int B, float *out)
{
  if (!params || !X || !out || B <= 0 || n_plan_classes <= 0)
    return fail(ISPM_ERR_INVALID, "mlp_forward: bad argument");
  int st = ensure_init();
  if (st) return st;
  const mlp_dims d = dims(n_plan_classes + 1);
  if ((st = ws_ensure(B, d.O))) return st;
  return forward(params, X, B, d, out);
}

int ispm_mlp_gradient(int n_plan_classes, const float *params, const float *X, const float *Tp,
                      const float *Mp, const float *Td, const float *Md, int B, float drop_weight,
                      float *grad, float *loss)
{
  if (!params || !X || !Tp || !Mp || !Td || !Md || !grad || B <= 0 || n_plan_classes <= 0)
    return fail(ISPM_ERR_INVALID, "mlp_gradient: bad argument");
  int st = ensure_init();
  if (st) return st;
  const int C = n_plan_classes;
  const mlp_dims d = dims(C + 1);
  if ((st = ws_ensure(B, d.O))) return st;
  if ((st = forward(params, X, B, d, g_ws.z))) return st;
  hipStream_t s = stream();
  hipLaunchKernelGGL(k_loss, blocks(B), dim3(256), 0, s, (const float *)g_ws.z, Tp, Mp, Td, Md, g_ws.dz, g_ws.lossv, B, C, drop_weight);
  /* output layer */
  hipLaunchKernelGGL(k_bwd_w, blocks((long long)d.O * kH + d.O), dim3(256), 0, s, (const float *)g_ws.dz, (const float *)g_ws.h2, grad + d.w3, grad + d.b3, B, kH, d.O);
  hipLaunchKernelGGL(k_bwd_x, blocks((long long)B * kH), dim3(256), 0, s, (const float *)g_ws.dz, params + d.w3, (const float *)g_ws.h2, g_ws.dh2, B, kH, d.O);
  /* hidden layer 2 */
  hipLaunchKernelGGL(k_bwd_w, blocks((long long)kH * kH + kH), dim3(256), 0, s, (const float *)g_ws.dh2, (const float *)g_ws.h1, grad + d.w2, grad + d.b2, B, kH, kH);
  hipLaunchKernelGGL(k_bwd_x, blocks((long long)B * kH), dim3(256), 0, s, (const float *)g_ws.dh2, params + d.w2, (const float *)g_ws.h1, g_ws.dh1, B, kH, kH);
  /* hidden layer 1 */
  hipLaunchKernelGGL(k_bwd_w, blocks((long long)kH * kIn + kH), dim3(256), 0, s, (const float *)g_ws.dh1, X, grad + d.w1, grad + d.b1, B, kIn, kH);
  if (loss) hipLaunchKernelGGL(k_sum, dim3(1), dim3(64), 0, s, (const float *)g_ws.lossv, B, loss);
  ISPM_HIP(hipGetLastError());
  return ISPM_OK;
}

int ispm_mlp_train_step(int n_plan_classes, float *params, float *m, float *v, int step,
                        const float *X, const float *Tp, const float *Mp, const float *Td,
                        const float *Md, int B, float lr, float drop_weight, float *loss)
{
  if (!m || !v || step < 1) return fail(ISPM_ERR_INVALID, "mlp_train_step: bad argument");
  const mlp_dims d = dims(n_plan_classes + 1);
  int st = ensure_init();
  if (st) return st;
  if ((st = ws_ensure(B, d.O))) return st;
  if ((st = ispm_mlp_gradient(n_plan_classes, params, X, Tp, Mp, Td, Md, B, drop_weight, g_ws.grad, loss))) return st;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
  hipLaunchKernelGGL(k_adam, blocks(d.total), dim3(256), 0, stream(), params, (const float *)g_ws.grad, m, v,
                     (int)d.total, lr, b1, b2, c1, c2, 1e-8f);
  ISPM_HIP(hipGetLastError());
  return ISPM_OK;
}

}  // extern "C"
