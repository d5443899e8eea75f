/* sblas_kernels.h -- device templates shared by sblas.hip and solver.hip: 16-byte vector
 * packs, the single-pass (last block finishes) reduction and the fused Ghysels-Vanroose
 * recurrence update (util/cuda/sblas.cu:177-237 of the reference). */
#ifndef ISPM_SBLAS_KERNELS_H
#define ISPM_SBLAS_KERNELS_H

#include <hip/hip_runtime.h>
#include "device.h"
#include "internal.h"

namespace ispm {

template<typename T> struct vec16;
template<> struct vec16<double> {
  typedef double2 type;
  static constexpr int n = 2;
  static __device__ __forceinline__ double dot(double2 a, double2 b) { return a.x * b.x + a.y * b.y; }
};
template<> struct vec16<float> {
  typedef float4 type;
  static constexpr int n = 4;
  static __device__ __forceinline__ double dot(float4 a, float4 b)
  {
    return (double)a.x * b.x + (double)a.y * b.y + (double)a.z * b.z + (double)a.w * b.w;
  }
};

template<typename T> struct alignas(16) pack16 { T v[16 / sizeof(T)]; };

template<typename T> __device__ __forceinline__ pack16<T> ld16(const T *p, int i)
{
  return reinterpret_cast<const pack16<T> *>(p)[i];
}
template<typename T> __device__ __forceinline__ void st16(T *p, int i, const pack16<T> &v)
{
  reinterpret_cast<pack16<T> *>(p)[i] = v;
}

/* Block (256 threads) reduces NS sums; block partials go to parts[s*G + block]; the last
 * block to arrive sums them in block order.  Returns true in the last block, with the
 * totals in out (valid in every thread of that block).  The counter is reset for the next
 * reduction by the finishing block. */
template<int NS>
__device__ __forceinline__ bool finish_reduction(double (&v)[NS], double *parts, unsigned *counter,
                                                 double *lds, double (&out)[NS])
{
  __shared__ unsigned s_last;
  dev::block_sum<256, NS>(v, lds);
  const unsigned G = gridDim.x;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int s = 0; s < NS; s++) parts[s * G + blockIdx.x] = v[s];
    unsigned prev = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    s_last = (prev == G - 1) ? 1u : 0u;
  }
  __syncthreads();
  if (!s_last) return false;
  __threadfence();
  double a[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) a[s] = 0;
  for (unsigned i = threadIdx.x; i < G; i += 256)
#pragma unroll
    for (int s = 0; s < NS; s++) a[s] += __builtin_nontemporal_load(&parts[s * G + i]);
  dev::block_sum<256, NS>(a, lds);
#pragma unroll
  for (int s = 0; s < NS; s++) out[s] = a[s];
  if (threadIdx.x == 0) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

/* Fused recurrence update, one pass over 10 input and 8 output vectors:
 *   z = n + b z;  q = m + b q;  s = w + b s;  p = u + b p;
 *   u -= a q;  w -= a z;  r -= a s;  x += a p;
 * with sums sum|r|, (r,u), (w,u), (r,r) accumulated in FP64.  Vector arithmetic in T as in
 * the reference (sblas.cu:192-200). */
template<typename T, bool VEC>
__global__ void __launch_bounds__(256)
ppcg_update_kernel(T *resnorm, T *gamma, T *delta, double *r2out, const T *alpha, const T *beta,
                   const T *__restrict__ n, const T *__restrict__ m, T *__restrict__ p,
                   T *__restrict__ s, T *__restrict__ q, T *__restrict__ z, T *__restrict__ x,
                   T *__restrict__ r, T *__restrict__ u, T *__restrict__ w, int N,
                   double *parts, unsigned *counter)
{
  __shared__ double lds[4 * 4];
  const T a = *alpha, b = *beta;
  double acc[4] = {0, 0, 0, 0};
  const int gt = blockIdx.x * 256 + threadIdx.x, gs = gridDim.x * 256;
  auto one = [&](T &zi, T &qi, T &si, T &pi, T &ui, T &wi, T &ri, T &xi, T ni, T mi) {
    zi = ni + b * zi;
    qi = mi + b * qi;
    si = wi + b * si;
    pi = ui + b * pi;
    ui = ui - a * qi;
    wi = wi - a * zi;
    ri = ri - a * si;
    xi = xi + a * pi;
    acc[0] += fabs((double)ri);
    acc[1] += (double)ri * (double)ui;
    acc[2] += (double)wi * (double)ui;
    acc[3] += (double)ri * (double)ri;
  };
  int tail0 = 0;
  if (VEC) {
    constexpr int L = 16 / sizeof(T);
    const int nv = N / L;
    for (int i = gt; i < nv; i += gs) {
      pack16<T> zn = ld16(n, i), mm = ld16(m, i), zz = ld16(z, i), qq = ld16(q, i),
                ss = ld16(s, i), pp = ld16(p, i), uu = ld16(u, i), ww = ld16(w, i),
                rr = ld16(r, i), xx = ld16(x, i);
#pragma unroll
      for (int l = 0; l < L; l++)
        one(zz.v[l], qq.v[l], ss.v[l], pp.v[l], uu.v[l], ww.v[l], rr.v[l], xx.v[l], zn.v[l], mm.v[l]);
      st16(z, i, zz); st16(q, i, qq); st16(s, i, ss); st16(p, i, pp);
      st16(u, i, uu); st16(w, i, ww); st16(r, i, rr); st16(x, i, xx);
    }
    tail0 = nv * L;
  }
  for (int i = tail0 + gt; i < N; i += gs) {
    T zi = z[i], qi = q[i], si = s[i], pi = p[i], ui = u[i], wi = w[i], ri = r[i], xi = x[i];
    one(zi, qi, si, pi, ui, wi, ri, xi, n[i], m[i]);
    z[i] = zi; q[i] = qi; s[i] = si; p[i] = pi; u[i] = ui; w[i] = wi; r[i] = ri; x[i] = xi;
  }
  double out[4];
  if (finish_reduction<4>(acc, parts, counter, lds, out) && threadIdx.x == 0) {
    *resnorm = (T)out[0];
    *gamma = (T)out[1];
    *delta = (T)out[2];
    if (r2out) *r2out = out[3];
  }
}

}  // namespace ispm

#endif
