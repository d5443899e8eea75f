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

/* Two-phase deterministic reductions without atomics.  Phase 1 (every reduction kernel,
 * 256-thread blocks): the block reduces its NS per-thread sums and thread 0 stores the block
 * totals to parts[s*G + block], G = gridDim.x.  Phase 2 (a one-block finishing kernel of
 * kFinThreads threads, launched next on the same stream): sum_partials adds the G partials
 * of each sum in a fixed order.  A device-scope atomic per block (the "last block finishes"
 * pattern) costs a cache write-back and invalidate per block on this multi-XCD part, which
 * measured at ~27 ns per block serialised -- 225 us for an 8192-block dot -- while the
 * extra launch costs a few microseconds. */
template<int NS>
__device__ __forceinline__ void store_partials(double (&v)[NS], double *parts, double *lds)
{
  dev::block_sum<256, NS>(v, lds);
  if (threadIdx.x == 0) {
    const unsigned G = gridDim.x;
#pragma unroll
    for (int s = 0; s < NS; s++) parts[s * G + blockIdx.x] = v[s];
  }
}

constexpr int kFinThreads = 1024;
static_assert(kMaxBlocks % kFinThreads == 0, "partials per finishing thread");

/* Sum G <= GMAX partials of each of NS sums (block of kFinThreads threads; lds holds
 * NS * kFinThreads/64 doubles).  All loads are issued before the adds; the order of the adds
 * is fixed, so the result is reproducible run to run.  out[] is valid in every thread. */
template<int NS, unsigned GMAX = kMaxBlocks>
struct partials {
  static_assert(GMAX % kFinThreads == 0, "partials per finishing thread");
  static constexpr int K = GMAX / kFinThreads;
  double t[NS][K];
  /* requests this thread's share of the G partials of each sum */
  __device__ __forceinline__ void load(const double *parts, unsigned G)
  {
#pragma unroll
    for (int s = 0; s < NS; s++)
#pragma unroll
      for (int k = 0; k < K; k++) {
        const unsigned i = threadIdx.x + k * kFinThreads;
        t[s][k] = i < G ? parts[s * G + i] : 0.0;
      }
  }
  /* the sums, valid in every thread (all threads of the block take part) */
  __device__ __forceinline__ void sum(double *lds, double (&out)[NS])
  {
#pragma unroll
    for (int s = 0; s < NS; s++) {
      double a = 0;
#pragma unroll
      for (int k = 0; k < K; k++) a += t[s][k];
      out[s] = a;
    }
    dev::block_sum<kFinThreads, NS>(out, lds);
  }
};

template<int NS, unsigned GMAX = kMaxBlocks>
__device__ __forceinline__ void sum_partials(const double *parts, unsigned G, double *lds, double (&out)[NS])
{
  partials<NS, GMAX> p;
  p.load(parts, G);
  p.sum(lds, out);
}

/* Fused recurrence update, one pass over 10 input and 8 output vectors:
 *   z = n + b z;  q = m + b q;  s = w + b s;  p = u + b p;
 *   u -= a q;  w -= a z;  r -= a s;  x += a p;
 * with sums sum|r|, (r,u), (w,u), (r,r) accumulated in FP64.  Vector arithmetic in T as in
 * the reference (sblas.cu:192-200). */
template<typename T, bool VEC>
__global__ void __launch_bounds__(256)
ppcg_update_kernel(const T *alpha, const T *beta,
                   const T *__restrict__ n, const T *__restrict__ m, T *__restrict__ p,
                   T *__restrict__ s, T *__restrict__ q, T *__restrict__ z, T *__restrict__ x,
                   T *__restrict__ r, T *__restrict__ u, T *__restrict__ w, int N,
                   double *parts)
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
  store_partials<4>(acc, parts, lds);
}

}  // namespace ispm

#endif
