/* batched.hip -- many independent small SPD systems solved in ONE persistent launch.
 *
 * Layout: systems are stored back to back; system i has n_i rows padded to a multiple of
 * 64 (padding rows are identity rows with a zero right-hand side) and is held as 64-row
 * ELL slices (column-major: slot (r,k) of slice s at slice_off[s] + r + 64k, width
 * slice_w[s]) with system-local column indices, FP32 values and its inverse diagonal.
 * Kernel: a fixed grid (a few workgroups per CU) pulls systems from an atomic work queue
 * (systems sorted by decreasing size: longest-processing-time-first) and solves each one
 * completely inside the workgroup with Jacobi-preconditioned CG: the SpMV gives one 64-row
 * slice to each wave (one row per lane), dot products are wave64 swizzle reductions
 * combined in LDS, and phases are separated by workgroup barriers -- no grid-wide
 * synchronisation, so the launch is capturable in a hipGraph.  Every workgroup leaves when
 * the queue is empty and every system stops at its iteration cap.  Scalars and reductions
 * in FP64, vectors in FP32.  No counterpart in the reference. */
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>
#include "device.h"
#include "internal.h"

namespace ispm {
namespace {

struct sys_desc {
  long long row0;       /* first (padded) row in the concatenated vectors */
  int n;                /* rows */
  int slice0;           /* first slice */
  int n_slices;
  int pad;
};

constexpr int kNT = 512;           /* threads per workgroup: 8 waves */
constexpr int kWaves = kNT / 64;

__device__ __forceinline__ double wg_sum(double v, double *lds)
{
  v = dev::wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0;
#pragma unroll
  for (int i = 0; i < kWaves; i++) s += lds[i];
  return s;
}

__global__ void __launch_bounds__(kNT)
batched_pcg_kernel(const sys_desc *__restrict__ sys, const int *__restrict__ order, int n_sys,
                   const long long *__restrict__ slice_off, const int *__restrict__ slice_w,
                   const unsigned *__restrict__ cols, const float *__restrict__ vals,
                   const float *__restrict__ dinv, const float *__restrict__ b, float *__restrict__ x,
                   float *__restrict__ r, float *__restrict__ p, float *__restrict__ q, double tol,
                   int maxit, unsigned *queue, int *iters_out, double *relres_out)
{
  __shared__ double lds[kWaves];
  __shared__ int s_job;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (;;) {
    if (threadIdx.x == 0) s_job = (int)atomicAdd(queue, 1u);
    __syncthreads();
    const int job = s_job;
    __syncthreads();
    if (job >= n_sys) return;
    const int id = order[job];
    const sys_desc d = sys[id];
    const int npad = d.n_slices * 64;
    const float *bi = b + d.row0, *di = dinv + d.row0;
    float *xi = x + d.row0, *ri = r + d.row0, *pi = p + d.row0, *qi = q + d.row0;
    /* x = 0, r = b, p = z = D^-1 r */
    double rz = 0, bb = 0;
    for (int i = threadIdx.x; i < npad; i += kNT) {
      const float bv = bi[i];
      const float z = bv * di[i];
      xi[i] = 0.f;
      ri[i] = bv;
      pi[i] = z;
      rz += (double)bv * z;
      bb += (double)bv * bv;
    }
    rz = wg_sum(rz, lds);
    bb = wg_sum(bb, lds);
    const double thr = tol * tol * bb;
    int it = 0;
    double rr = bb;
    while (rr > thr && it < maxit) {
      __syncthreads();   /* p complete before the product gathers from it */
      double pq = 0;
      for (int s = wave; s < d.n_slices; s += kWaves) {
        const int gs = d.slice0 + s;
        const long long sb = slice_off[gs] + lane;
        const int w = slice_w[gs];
        float acc = 0.f;
        for (int k = 0; k < w; k++) {
          const long long slot = sb + 64LL * k;
          acc += vals[slot] * pi[cols[slot]];
        }
        const int row = s * 64 + lane;
        qi[row] = acc;
        pq += (double)pi[row] * acc;
      }
      pq = wg_sum(pq, lds);
      const float a = (float)(rz / pq);
      double rz2 = 0, rr2 = 0;
      for (int i = threadIdx.x; i < npad; i += kNT) {
        xi[i] += a * pi[i];
        const float rv = ri[i] - a * qi[i];
        ri[i] = rv;
        rz2 += (double)rv * (rv * di[i]);
        rr2 += (double)rv * rv;
      }
      rz2 = wg_sum(rz2, lds);
      rr2 = wg_sum(rr2, lds);
      const float beta = (float)(rz2 / rz);
      rz = rz2;
      rr = rr2;
      for (int i = threadIdx.x; i < npad; i += kNT) pi[i] = ri[i] * di[i] + beta * pi[i];
      it++;
    }
    if (threadIdx.x == 0) {
      iters_out[id] = it;
      relres_out[id] = bb > 0 ? sqrt(rr / bb) : 0.0;
    }
    __syncthreads();
  }
}

__global__ void reset_queue(unsigned *q) { *q = 0u; }

}  // namespace
}  // namespace ispm

using namespace ispm;

struct ispm_batch {
  int n_sys = 0;
  long long rows_padded = 0;
  std::vector<long long> row_off;
  sys_desc *d_sys = nullptr;
  int *d_order = nullptr;
  long long *d_soff = nullptr;
  int *d_sw = nullptr;
  unsigned *d_cols = nullptr;
  float *d_vals = nullptr, *d_dinv = nullptr, *d_r = nullptr, *d_p = nullptr, *d_q = nullptr;
  unsigned *d_queue = nullptr;
  int *d_iters = nullptr;
  double *d_rel = nullptr;
  int grid = 1;
  size_t slots = 0;
};

extern "C" {

int ispm_batch_create(ispm_batch **out, int n_sys, const int *sys_rows, const int *row_ptr,
                      const int *cols, const float *vals)
{
  if (!out || n_sys <= 0 || !sys_rows || !row_ptr || !cols || !vals)
    return fail(ISPM_ERR_INVALID, "batch_create: bad argument");
  *out = nullptr;
  int st = ensure_init();
  if (st) return st;
  ispm_batch *B = new ispm_batch;
  B->n_sys = n_sys;
  std::vector<sys_desc> sd(n_sys);
  std::vector<long long> soff;
  std::vector<int> sw;
  std::vector<unsigned> hc;
  std::vector<float> hv, hd;
  long long grow = 0, prow = 0;   /* global CSR row, padded row */
  for (int i = 0; i < n_sys; i++) {
    const int n = sys_rows[i];
    if (n <= 0) { delete B; return fail(ISPM_ERR_INVALID, "batch_create: empty system %d", i); }
    const int ns = (n + 63) / 64;
    sd[i] = sys_desc{prow, n, (int)soff.size(), ns, 0};
    B->row_off.push_back(prow);
    hd.resize(prow + (long long)ns * 64, 1.0f);
    for (int s = 0; s < ns; s++) {
      int w = 0;
      for (int rl = s * 64; rl < std::min(n, s * 64 + 64); rl++)
        w = std::max(w, row_ptr[grow + rl + 1] - row_ptr[grow + rl]);
      const long long base = (long long)hv.size();
      soff.push_back(base);
      sw.push_back(w);
      hv.resize(base + 64LL * w, 0.0f);
      hc.resize(base + 64LL * w, 0u);
      for (int l = 0; l < 64; l++) {
        const int rl = s * 64 + l;
        if (rl >= n) {                    /* padding row: identity, column = itself */
          for (int k = 0; k < w; k++) hc[base + l + 64LL * k] = (unsigned)rl;
          if (w > 0) hv[base + l] = 1.0f;
          continue;
        }
        const long long gr = grow + rl;
        int k = 0;
        float diag = 0.f;
        for (int e = row_ptr[gr]; e < row_ptr[gr + 1]; e++, k++) {
          const long long c = cols[e] - grow;
          if (c < 0 || c >= n) { delete B; return fail(ISPM_ERR_INVALID, "batch_create: system %d couples to another system", i); }
          hc[base + l + 64LL * k] = (unsigned)c;
          hv[base + l + 64LL * k] = vals[e];
          if (c == rl) diag += vals[e];
        }
        for (; k < w; k++) hc[base + l + 64LL * k] = (unsigned)rl;
        hd[prow + rl] = diag != 0.f ? 1.0f / diag : 1.0f;
      }
    }
    grow += n;
    prow += (long long)ns * 64;
  }
  B->rows_padded = prow;
  B->slots = hv.size();
  /* longest systems first */
  std::vector<int> order(n_sys);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return sd[a].n > sd[b].n; });
  auto up = [&](void **dst, const void *src, size_t bytes) -> int {
    int e = ispm_malloc(dst, bytes ? bytes : 8);
    if (e) return e;
    return bytes ? ispm_memcpy(*dst, src, bytes, ISPM_H2D) : ISPM_OK;
  };
  if ((st = up((void **)&B->d_sys, sd.data(), sizeof(sys_desc) * n_sys)) ||
      (st = up((void **)&B->d_order, order.data(), sizeof(int) * n_sys)) ||
      (st = up((void **)&B->d_soff, soff.data(), sizeof(long long) * soff.size())) ||
      (st = up((void **)&B->d_sw, sw.data(), sizeof(int) * sw.size())) ||
      (st = up((void **)&B->d_cols, hc.data(), sizeof(unsigned) * hc.size())) ||
      (st = up((void **)&B->d_vals, hv.data(), sizeof(float) * hv.size())) ||
      (st = up((void **)&B->d_dinv, hd.data(), sizeof(float) * hd.size())) ||
      (st = ispm_malloc((void **)&B->d_r, sizeof(float) * prow)) ||
      (st = ispm_malloc((void **)&B->d_p, sizeof(float) * prow)) ||
      (st = ispm_malloc((void **)&B->d_q, sizeof(float) * prow)) ||
      (st = ispm_malloc((void **)&B->d_queue, sizeof(unsigned))) ||
      (st = ispm_malloc((void **)&B->d_iters, sizeof(int) * n_sys)) ||
      (st = ispm_malloc((void **)&B->d_rel, sizeof(double) * n_sys))) {
    ispm_batch_destroy(B);
    return st;
  }
  B->grid = std::min(n_sys, num_cus() * 2);
  *out = B;
  return ISPM_OK;
}

int ispm_batch_layout(const ispm_batch *B, long long *rows_padded, long long *row_offsets)
{
  if (!B) return fail(ISPM_ERR_INVALID, "batch_layout: null batch");
  if (rows_padded) *rows_padded = B->rows_padded;
  if (row_offsets) std::copy(B->row_off.begin(), B->row_off.end(), row_offsets);
  return ISPM_OK;
}

int ispm_batch_solve(ispm_batch *B, const float *b, float *x, double tol, int maxit,
                     int *iters, double *relres)
{
  if (!B || !b || !x) return fail(ISPM_ERR_INVALID, "batch_solve: null argument");
  hipStream_t s = stream();
  hipLaunchKernelGGL(reset_queue, dim3(1), dim3(1), 0, s, B->d_queue);
  hipLaunchKernelGGL(batched_pcg_kernel, dim3(B->grid), dim3(kNT), 0, s, (const sys_desc *)B->d_sys,
                     (const int *)B->d_order, B->n_sys, (const long long *)B->d_soff, (const int *)B->d_sw,
                     (const unsigned *)B->d_cols, (const float *)B->d_vals, (const float *)B->d_dinv, b, x,
                     B->d_r, B->d_p, B->d_q, tol, maxit > 0 ? maxit : 10000, B->d_queue, B->d_iters, B->d_rel);
  ISPM_HIP(hipGetLastError());
  if (iters) ISPM_HIP(hipMemcpyAsync(iters, B->d_iters, sizeof(int) * B->n_sys, hipMemcpyDeviceToHost, s));
  if (relres) ISPM_HIP(hipMemcpyAsync(relres, B->d_rel, sizeof(double) * B->n_sys, hipMemcpyDeviceToHost, s));
  if (iters || relres) ISPM_HIP(hipStreamSynchronize(s));
  return ISPM_OK;
}

void ispm_batch_destroy(ispm_batch *B)
{
  if (!B) return;
  void *ps[] = {B->d_sys, B->d_order, B->d_soff, B->d_sw, B->d_cols, B->d_vals, B->d_dinv,
                B->d_r, B->d_p, B->d_q, B->d_queue, B->d_iters, B->d_rel};
  for (void *p : ps) ispm_free(p);
  delete B;
}

}  // extern "C"
