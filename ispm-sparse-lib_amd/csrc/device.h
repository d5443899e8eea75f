/* device.h -- device-side helpers shared by the HIP translation units: element loads,
 * wave64 reductions on ds_swizzle / readlane, and the last-block-finishes reduction. */
#ifndef ISPM_DEVICE_H
#define ISPM_DEVICE_H

#include <hip/hip_runtime.h>
#include <cstdint>

namespace ispm {
namespace dev {

typedef uint16_t bf16_t;

/* ---- element conversion: stored matrix element E -> accumulation type T ---- */
template<typename T> __device__ __forceinline__ T cvt(float v) { return (T)v; }
template<typename T> __device__ __forceinline__ T cvt_d(double v) { return (T)v; }

/* Matrix arrays are streamed once per product: they are read with nontemporal loads so
 * they do not evict the x windows that neighbouring slices reuse from L2. */
typedef float nt_f2 __attribute__((ext_vector_type(2)));
typedef double nt_d2 __attribute__((ext_vector_type(2)));
typedef unsigned nt_u2 __attribute__((ext_vector_type(2)));

template<typename T, typename E> struct elt;
template<typename T> struct elt<T, float> {
  static __device__ __forceinline__ T get(const float *p, unsigned i) { return (T)__builtin_nontemporal_load(p + i); }
  static __device__ __forceinline__ void get2(const float *p, unsigned i, T &a, T &b)
  {
    nt_f2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_f2 *>(p + i));
    a = (T)v.x; b = (T)v.y;
  }
};
template<typename T> struct elt<T, double> {
  static __device__ __forceinline__ T get(const double *p, unsigned i) { return (T)__builtin_nontemporal_load(p + i); }
  static __device__ __forceinline__ void get2(const double *p, unsigned i, T &a, T &b)
  {
    nt_d2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_d2 *>(p + i));
    a = (T)v.x; b = (T)v.y;
  }
};
__device__ __forceinline__ float bf16_to_f32(unsigned short u) { return __uint_as_float((unsigned)u << 16); }
template<typename T> struct elt<T, bf16_t> {
  static __device__ __forceinline__ T get(const bf16_t *p, unsigned i) { return (T)bf16_to_f32(__builtin_nontemporal_load(p + i)); }
  static __device__ __forceinline__ void get2(const bf16_t *p, unsigned i, T &a, T &b)
  {
    unsigned v = __builtin_nontemporal_load(reinterpret_cast<const unsigned *>(p + i));
    a = (T)__uint_as_float(v << 16);
    b = (T)__uint_as_float(v & 0xffff0000u);
  }
};
/* VEC consecutive elements (16 bytes, 8 for BF16) from a VEC-aligned index, converted to T */
typedef float nt_f4 __attribute__((ext_vector_type(4)));
typedef unsigned nt_u4 __attribute__((ext_vector_type(4)));
template<typename E> struct evec { static constexpr unsigned n = 4; };
template<> struct evec<double> { static constexpr unsigned n = 2; };
template<typename T> __device__ __forceinline__ void elt_vec(const double *p, unsigned i, T (&v)[2])
{
  nt_d2 w = __builtin_nontemporal_load(reinterpret_cast<const nt_d2 *>(p + i));
  v[0] = (T)w.x; v[1] = (T)w.y;
}
template<typename T> __device__ __forceinline__ void elt_vec(const float *p, unsigned i, T (&v)[4])
{
  nt_f4 w = __builtin_nontemporal_load(reinterpret_cast<const nt_f4 *>(p + i));
  v[0] = (T)w.x; v[1] = (T)w.y; v[2] = (T)w.z; v[3] = (T)w.w;
}
template<typename T> __device__ __forceinline__ void elt_vec(const bf16_t *p, unsigned i, T (&v)[4])
{
  nt_u2 w = __builtin_nontemporal_load(reinterpret_cast<const nt_u2 *>(p + i));
  v[0] = (T)__uint_as_float(w.x << 16); v[1] = (T)__uint_as_float(w.x & 0xffff0000u);
  v[2] = (T)__uint_as_float(w.y << 16); v[3] = (T)__uint_as_float(w.y & 0xffff0000u);
}
__device__ __forceinline__ void col_vec(const unsigned *p, unsigned i, unsigned (&c)[2])
{
  nt_u2 w = __builtin_nontemporal_load(reinterpret_cast<const nt_u2 *>(p + i));
  c[0] = w.x; c[1] = w.y;
}
__device__ __forceinline__ void col_vec(const unsigned *p, unsigned i, unsigned (&c)[4])
{
  nt_u4 w = __builtin_nontemporal_load(reinterpret_cast<const nt_u4 *>(p + i));
  c[0] = w.x; c[1] = w.y; c[2] = w.z; c[3] = w.w;
}

__device__ __forceinline__ unsigned col_load(const unsigned *p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void col_load2(const unsigned *p, unsigned &a, unsigned &b)
{
  nt_u2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_u2 *>(p));
  a = v.x; b = v.y;
}

/* ---- wave64 reductions ---- */
/* ds_swizzle in bit mode: lane l reads lane (l ^ m) within its 32-lane half. */
template<int M> __device__ __forceinline__ float swz_xor(float v)
{
  return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x1f | (M << 10)));
}
template<int M> __device__ __forceinline__ double swz_xor(double v)
{
  long long b = __double_as_longlong(v);
  int lo = __builtin_amdgcn_ds_swizzle((int)(b & 0xffffffff), 0x1f | (M << 10));
  int hi = __builtin_amdgcn_ds_swizzle((int)(b >> 32), 0x1f | (M << 10));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double readlane(double v, int l)
{
  long long b = __double_as_longlong(v);
  int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), l);
  int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float readlane(float v, int l)
{
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

/* Sum over the 64 lanes of a wave; the result is wave-uniform.  Butterfly in each 32-lane
 * half with ds_swizzle, then the two halves combined through readlane.  The summation
 * order is fixed, so the result is deterministic. */
template<typename T> __device__ __forceinline__ T wave_sum(T v)
{
  v += swz_xor<1>(v);
  v += swz_xor<2>(v);
  v += swz_xor<4>(v);
  v += swz_xor<8>(v);
  v += swz_xor<16>(v);
  return readlane(v, 0) + readlane(v, 32);
}

/* Block-wide sum of NS values per thread (block of NT threads, NT a multiple of 64).
 * Result valid in every thread.  lds must hold NS * NT/64 values of T. */
template<int NT, int NS, typename T> __device__ __forceinline__ void block_sum(T (&v)[NS], T *lds)
{
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int s = 0; s < NS; s++) v[s] = wave_sum(v[s]);
  if (NT > 64) {
    if (lane == 0) {
#pragma unroll
      for (int s = 0; s < NS; s++) lds[s * (NT / 64) + w] = v[s];
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < NS; s++) {
      T a = 0;
#pragma unroll
      for (int i = 0; i < NT / 64; i++) a += lds[s * (NT / 64) + i];
      v[s] = a;
    }
    __syncthreads();
  }
}

}  // namespace dev
}  // namespace ispm

#endif
