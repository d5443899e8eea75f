/* refresh.hip -- on-device coefficient refresh: new CSR values for the plan's sparsity
 * pattern are scattered into the SlELL element slots through the CSR-entry -> slot map the
 * conversion recorded.  Replaces change_host_matrix + update_device_matrix_coefs
 * (plan/plan.h:39-56 of the reference: host re-conversion and a full host-to-device copy)
 * with one kernel that reads nnz values and nnz slot indices. */
#include <hip/hip_runtime.h>
#include <vector>
#include "device.h"
#include "internal.h"
#include "plan_internal.h"

namespace ispm {
namespace {

template<typename V> __device__ __forceinline__ float to_f32(V v) { return (float)v; }

__device__ __forceinline__ unsigned short f32_bits_to_bf16(float f)
{
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (unsigned short)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

template<typename V, typename E>
__global__ void __launch_bounds__(256) scatter_kernel(E *__restrict__ elms, const V *__restrict__ vals,
                                                      const unsigned *__restrict__ map, long long nnz)
{
  const long long stride = (long long)gridDim.x * 256;
  for (long long k = (long long)blockIdx.x * 256 + threadIdx.x; k < nnz; k += stride) {
    const V v = vals[k];
    if constexpr (sizeof(E) == 2) elms[map[k]] = f32_bits_to_bf16((float)v);
    else elms[map[k]] = (E)v;
  }
}

/* symmetric plans: slot ~0u (a lower entry, the mirror of a stored one) is skipped; a slot
 * in a plane's first four rows is also written to the previous slice's repeated rows */
template<typename V, typename E>
__global__ void __launch_bounds__(256) sym_scatter_kernel(E *__restrict__ elms, const V *__restrict__ vals,
                                                          const unsigned *__restrict__ map, long long nnz,
                                                          unsigned ks)
{
  const long long stride = (long long)gridDim.x * 256;
  for (long long k = (long long)blockIdx.x * 256 + threadIdx.x; k < nnz; k += stride) {
    const unsigned s = map[k];
    if (s == ~0u) continue;
    E e;
    if constexpr (sizeof(E) == 2) e = f32_bits_to_bf16((float)vals[k]);
    else e = (E)vals[k];
    elms[s] = e;
    if (s % kSymPlane < 4 && s >= ks) elms[s - ks + 256] = e;
  }
}

template<typename V>
int scatter(ispm_plan *p, const V *vals, long long nnz)
{
  const unsigned *map = static_cast<const unsigned *>(p->dev_map);
  long long blocks = (nnz + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  dim3 g((unsigned)blocks), b(256);
  void *e = const_cast<void *>(p->dev.elms);
  if (p->sym) {
    const unsigned ks = p->symd.K * kSymPlane;
    switch (p->E) {
      case ISPM_F64: hipLaunchKernelGGL((sym_scatter_kernel<V, double>), g, b, 0, stream(), (double *)e, vals, map, nnz, ks); break;
      case ISPM_F32: hipLaunchKernelGGL((sym_scatter_kernel<V, float>), g, b, 0, stream(), (float *)e, vals, map, nnz, ks); break;
      default: hipLaunchKernelGGL((sym_scatter_kernel<V, unsigned short>), g, b, 0, stream(), (unsigned short *)e, vals, map, nnz, ks); break;
    }
    ISPM_HIP(hipGetLastError());
    return ISPM_OK;
  }
  switch (p->E) {
    case ISPM_F64: hipLaunchKernelGGL((scatter_kernel<V, double>), g, b, 0, stream(), (double *)e, vals, map, nnz); break;
    case ISPM_F32: hipLaunchKernelGGL((scatter_kernel<V, float>), g, b, 0, stream(), (float *)e, vals, map, nnz); break;
    default: hipLaunchKernelGGL((scatter_kernel<V, unsigned short>), g, b, 0, stream(), (unsigned short *)e, vals, map, nnz); break;
  }
  ISPM_HIP(hipGetLastError());
  return ISPM_OK;
}

/* lduMatrix coefficients into CSR value order: source s (diag, then upper, then lower)
 * goes to vals[pos[s]]; positions outside [0, nnz) are skipped */
__global__ void __launch_bounds__(256) ldu_scatter_kernel(double *__restrict__ vals, const int *__restrict__ pos,
                                                          const double *__restrict__ diag, const double *__restrict__ upper,
                                                          const double *__restrict__ lower, int n, int nf)
{
  const long long nnz = (long long)n + 2LL * nf, stride = (long long)gridDim.x * 256;
  for (long long s = (long long)blockIdx.x * 256 + threadIdx.x; s < nnz; s += stride) {
    const double v = s < n ? diag[s] : s < (long long)n + nf ? upper[s - n] : lower[s - n - nf];
    const long long k = pos[s];
    if (k >= 0 && k < nnz) vals[k] = v;
  }
}

}  // namespace
}  // namespace ispm

using namespace ispm;

extern "C" int ispm_ldu_values_device(int n, int n_faces, const int *pos, const double *diag, const double *upper,
                                      const double *lower, double *vals)
{
  if (n < 0 || n_faces < 0) return fail(ISPM_ERR_INVALID, "ldu_values_device: negative size");
  if (!vals || !pos || (n && !diag) || (n_faces && !upper)) return fail(ISPM_ERR_INVALID, "ldu_values_device: null argument");
  int st = ensure_init();
  if (st) return st;
  const long long nnz = (long long)n + 2LL * n_faces;
  if (nnz == 0) return ISPM_OK;
  long long blocks = (nnz + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(ldu_scatter_kernel, dim3((unsigned)blocks), dim3(256), 0, stream(), vals, pos, diag, upper,
                     lower ? lower : upper, n, n_faces);
  ISPM_HIP(hipGetLastError());
  return ISPM_OK;
}

extern "C" int ispm_plan_update_coefs_device(ispm_plan *p, const void *vals, int val_type)
{
  if (!p || !vals) return fail(ISPM_ERR_INVALID, "plan_update_coefs_device: null argument");
  if (!p->has_dev) return fail(ISPM_ERR_INVALID, "plan_update_coefs_device: plan has no device matrix");
  if (p->split) return fail(ISPM_ERR_UNSUPPORTED, "plan_update_coefs_device: split-diagonal plans are rebuilt");
  int st = ensure_init();
  if (st) return st;
  const long long nnz = (long long)p->L.map.size();
  if (!p->dev_map) {
    /* slot indices fit 32 bits: conversion refuses layouts beyond 2^32 slots; a symmetric
     * plan's lower entries (negative: mirrors) are skipped on the device (the caller keeps
     * the matrix symmetric, as at creation) */
    std::vector<unsigned> m(p->L.map.size());
    for (size_t k = 0; k < m.size(); k++) m[k] = p->L.map[k] < 0 ? ~0u : (unsigned)p->L.map[k];
    void *d = nullptr;
    if ((st = ispm_malloc(&d, sizeof(unsigned) * (nnz ? nnz : 1)))) return st;
    if (nnz && (st = ispm_memcpy(d, m.data(), sizeof(unsigned) * nnz, ISPM_H2D))) { ispm_free(d); return st; }
    p->dev_map = d;
  }
  if (nnz == 0) return ISPM_OK;
  if (val_type == ISPM_F64) return scatter(p, static_cast<const double *>(vals), nnz);
  if (val_type == ISPM_F32) return scatter(p, static_cast<const float *>(vals), nnz);
  return fail(ISPM_ERR_UNSUPPORTED, "plan_update_coefs_device: values must be F32 or F64");
}
