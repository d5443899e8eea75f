/* sblas.hip -- pipelined-PCG vector kernels (replaces util/cuda/sblas.cu:18-237).
 *
 * Differences in mechanism from the reference, not in results:
 *  - reductions are two-phase without atomics: each block reduces with wave64 butterflies
 *    (device.h) and stores its partial; a one-block finishing kernel sums the partials in
 *    block order (sblas_kernels.h), so the result is deterministic run to run;
 *  - dot products and the three recurrence sums accumulate in FP64 for both FP32 and
 *    FP64 vectors;
 *  - the fused update reads 10 and writes 8 vectors in one pass with 16-byte accesses.
 */
#include <hip/hip_runtime.h>
#include "device.h"
#include "internal.h"
#include "sblas_kernels.h"

namespace ispm {
namespace {

template<typename T>
__global__ void __launch_bounds__(256) fill_zero_kernel(unsigned char *p, size_t bytes)
{
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  size_t stride = (size_t)gridDim.x * 256;
  size_t nv = bytes / 16;
  uint4 *q = reinterpret_cast<uint4 *>(p);
  for (size_t k = i; k < nv; k += stride) q[k] = make_uint4(0, 0, 0, 0);
  for (size_t k = nv * 16 + i; k < bytes; k += stride) p[k] = 0;
}

template<typename T>
__global__ void __launch_bounds__(256) copy_indexed_kernel(T *dst, const T *src, const int *idx, int n)
{
  int k = blockIdx.x * 256 + threadIdx.x;
  if (k < n) dst[k] = src[idx[k]];
}

template<typename T>
__global__ void __launch_bounds__(256) add_indexed_kernel(T *dst, const T *src, const int *idx, int n)
{
  int k = blockIdx.x * 256 + threadIdx.x;
  if (k < n) dst[idx[k]] += src[k];   /* non-atomic, like sblas.cu:42: indices must be distinct */
}

template<typename T>
__global__ void ppcg_scalars_kernel(T *alpha, T *beta, const T *gamma, const T *gammaold,
                                    const T *delta)
{
  double g = (double)*gamma;
  double b = g / (double)*gammaold;
  *beta = (T)b;
  *alpha = (T)(g / ((double)*delta - b * g / (double)*alpha));
}

template<typename T, bool VEC>
__global__ void __launch_bounds__(256) dot_kernel(const T *x, const T *y, int n, double *parts)
{
  __shared__ double lds[4];
  double v[1] = {0.0};
  const int gt = blockIdx.x * 256 + threadIdx.x, gs = gridDim.x * 256;
  if (VEC) {
    typedef typename vec16<T>::type V16;
    constexpr int L = vec16<T>::n;
    const int nv = n / L;
    const V16 *xv = reinterpret_cast<const V16 *>(x), *yv = reinterpret_cast<const V16 *>(y);
    for (int i = gt; i < nv; i += gs) {
      V16 a = xv[i], b = yv[i];
      v[0] += vec16<T>::dot(a, b);
    }
    for (int i = nv * L + gt; i < n; i += gs) v[0] += (double)x[i] * (double)y[i];
  } else {
    for (int i = gt; i < n; i += gs) v[0] += (double)x[i] * (double)y[i];
  }
  store_partials<1>(v, parts, lds);
}

template<typename T>
__global__ void __launch_bounds__(kFinThreads) dot_finish(const double *parts, int G, T *s)
{
  __shared__ double lds[kFinThreads / 64];
  double out[1];
  sum_partials<1>(parts, G, lds, out);
  if (threadIdx.x == 0) *s = (T)out[0];
}

template<typename T>
__global__ void __launch_bounds__(kFinThreads)
ppcg_update_finish(const double *parts, int G, T *resnorm, T *gamma, T *delta)
{
  __shared__ double lds[4 * kFinThreads / 64];
  double out[4];
  sum_partials<4>(parts, G, lds, out);
  if (threadIdx.x == 0) {
    *resnorm = (T)out[0];
    *gamma = (T)out[1];
    *delta = (T)out[2];
  }
}

template<typename T>
void launch_ppcg_update(bool vec, int grid, double *parts, T *resnorm, T *gamma, T *delta,
                        const T *alpha, const T *beta, const T *n, const T *m, T *p, T *s, T *q,
                        T *z, T *x, T *r, T *u, T *w, int n_elts)
{
  if (vec) hipLaunchKernelGGL((ppcg_update_kernel<T, true>), dim3(grid), dim3(256), 0, stream(), alpha, beta, n, m, p, s, q, z, x, r, u, w, n_elts, parts);
  else     hipLaunchKernelGGL((ppcg_update_kernel<T, false>), dim3(grid), dim3(256), 0, stream(), alpha, beta, n, m, p, s, q, z, x, r, u, w, n_elts, parts);
  hipLaunchKernelGGL(ppcg_update_finish<T>, dim3(1), dim3(kFinThreads), 0, stream(), (const double *)parts, grid, resnorm, gamma, delta);
}

}  // namespace

}  // namespace ispm

/* Scratch of ispm_ppcg_update_vectors_noalloc: the per-workgroup partials of its four sums
 * (parts[s*G + block], G <= kMaxBlocks workgroups) in a file-scope device array instead of
 * the library's allocated scratch, so that the call allocates nothing.  Every GPU holds its
 * own copy until the process ends.  In host code the array's name is not a device address:
 * the address is taken once per device with hipGetSymbolAddress at initialisation and kept in
 * g_noalloc_parts. */
static constexpr int kNoAllocSums = 4;
static __device__ __align__(256) double noalloc_partials[ispm::kMaxBlocks * kNoAllocSums];
static double *g_noalloc_parts[ispm::kMaxDevices];

namespace ispm {
int noalloc_scratch_init(int dev)
{
  if (dev < 0 || dev >= kMaxDevices) return fail(ISPM_ERR_DEVICE, "noalloc scratch: device %d", dev);
  void *p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_partials));
  if (e != hipSuccess) return hip_status(e, "hipGetSymbolAddress(noalloc_partials)");
  g_noalloc_parts[dev] = static_cast<double *>(p);
  return ISPM_OK;
}
}  // namespace ispm

using namespace ispm;

static int grid_for(int n)
{
  /* about one 16-byte access per thread (two elements for FP64, four for FP32), capped by
   * the reduction-partials capacity */
  int need = (n + 511) / 512;
  if (need < 1) need = 1;
  return need < kMaxBlocks ? need : kMaxBlocks;
}

static bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

extern "C" int ispm_zero(void *p, size_t bytes)
{
  int st = ensure_init();
  if (st) return st;
  if (!bytes) return ISPM_OK;
  if (!p) return fail(ISPM_ERR_INVALID, "ispm_zero: null pointer");
  if (!aligned16(p)) {
    ISPM_HIP(hipMemsetAsync(p, 0, bytes, stream()));
    return ISPM_OK;
  }
  size_t blocks = (bytes / 16 + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(fill_zero_kernel<char>, dim3((unsigned)blocks), dim3(256), 0, stream(),
                     (unsigned char *)p, bytes);
  ISPM_HIP(hipGetLastError());
  return ISPM_OK;
}

extern "C" int ispm_copy_indexed(int type, void *dst, const void *src, const int *idx, int n)
{
  int st = ensure_init();
  if (st) return st;
  if (n <= 0) return ISPM_OK;
  if (!dst || !src || !idx) return fail(ISPM_ERR_INVALID, "copy_indexed: null pointer");
  dim3 g((n + 255) / 256), b(256);
  if (type == ISPM_F64)
    hipLaunchKernelGGL(copy_indexed_kernel<double>, g, b, 0, stream(), (double *)dst, (const double *)src, idx, n);
  else if (type == ISPM_F32)
    hipLaunchKernelGGL(copy_indexed_kernel<float>, g, b, 0, stream(), (float *)dst, (const float *)src, idx, n);
  else return fail(ISPM_ERR_UNSUPPORTED, "copy_indexed: type %d", type);
  ISPM_HIP(hipGetLastError());
  return ISPM_OK;
}

extern "C" int ispm_add_indexed(int type, void *dst, const void *src, const int *idx, int n)
{
  int st = ensure_init();
  if (st) return st;
  if (n <= 0) return ISPM_OK;
  if (!dst || !src || !idx) return fail(ISPM_ERR_INVALID, "add_indexed: null pointer");
  dim3 g((n + 255) / 256), b(256);
  if (type == ISPM_F64)
    hipLaunchKernelGGL(add_indexed_kernel<double>, g, b, 0, stream(), (double *)dst, (const double *)src, idx, n);
  else if (type == ISPM_F32)
    hipLaunchKernelGGL(add_indexed_kernel<float>, g, b, 0, stream(), (float *)dst, (const float *)src, idx, n);
  else return fail(ISPM_ERR_UNSUPPORTED, "add_indexed: type %d", type);
  ISPM_HIP(hipGetLastError());
  return ISPM_OK;
}

extern "C" int ispm_dot(int type, const void *x, const void *y, void *s, int n)
{
  int st = ensure_init();
  if (st) return st;
  if (!x || !y || !s) return fail(ISPM_ERR_INVALID, "dot: null pointer");
  if (n < 0) return fail(ISPM_ERR_INVALID, "dot: n < 0");
  const int grid = grid_for(n);
  double *parts = scratch().partials;
  const bool vec = aligned16(x) && aligned16(y);
  if (type == ISPM_F64) {
    if (vec) hipLaunchKernelGGL((dot_kernel<double, true>), dim3(grid), dim3(256), 0, stream(), (const double *)x, (const double *)y, n, parts);
    else     hipLaunchKernelGGL((dot_kernel<double, false>), dim3(grid), dim3(256), 0, stream(), (const double *)x, (const double *)y, n, parts);
    hipLaunchKernelGGL(dot_finish<double>, dim3(1), dim3(kFinThreads), 0, stream(), (const double *)parts, grid, (double *)s);
  } else if (type == ISPM_F32) {
    if (vec) hipLaunchKernelGGL((dot_kernel<float, true>), dim3(grid), dim3(256), 0, stream(), (const float *)x, (const float *)y, n, parts);
    else     hipLaunchKernelGGL((dot_kernel<float, false>), dim3(grid), dim3(256), 0, stream(), (const float *)x, (const float *)y, n, parts);
    hipLaunchKernelGGL(dot_finish<float>, dim3(1), dim3(kFinThreads), 0, stream(), (const double *)parts, grid, (float *)s);
  } else return fail(ISPM_ERR_UNSUPPORTED, "dot: type %d", type);
  ISPM_HIP(hipGetLastError());
  return ISPM_OK;
}

extern "C" int ispm_ppcg_update_scalars(int type, void *alpha, void *beta, const void *gamma,
                                        const void *gammaold, const void *delta)
{
  int st = ensure_init();
  if (st) return st;
  if (!alpha || !beta || !gamma || !gammaold || !delta) return fail(ISPM_ERR_INVALID, "ppcg_update_scalars: null pointer");
  if (type == ISPM_F64)
    hipLaunchKernelGGL(ppcg_scalars_kernel<double>, dim3(1), dim3(1), 0, stream(), (double *)alpha, (double *)beta, (const double *)gamma, (const double *)gammaold, (const double *)delta);
  else if (type == ISPM_F32)
    hipLaunchKernelGGL(ppcg_scalars_kernel<float>, dim3(1), dim3(1), 0, stream(), (float *)alpha, (float *)beta, (const float *)gamma, (const float *)gammaold, (const float *)delta);
  else return fail(ISPM_ERR_UNSUPPORTED, "ppcg_update_scalars: type %d", type);
  ISPM_HIP(hipGetLastError());
  return ISPM_OK;
}

extern "C" int ispm_ppcg_update_vectors(int type, void *resnorm, void *gamma, void *delta,
                                        const void *alpha, const void *beta, const void *n,
                                        const void *m, void *p, void *s, void *q, void *z,
                                        void *x, void *r, void *u, void *w, int n_elts)
{
  int st = ensure_init();
  if (st) return st;
  if (!resnorm || !gamma || !delta || !alpha || !beta || !n || !m || !p || !s || !q || !z || !x || !r || !u || !w)
    return fail(ISPM_ERR_INVALID, "ppcg_update_vectors: null pointer");
  if (n_elts < 0) return fail(ISPM_ERR_INVALID, "ppcg_update_vectors: n < 0");
  if (type != ISPM_F64 && type != ISPM_F32) return fail(ISPM_ERR_UNSUPPORTED, "ppcg_update_vectors: type %d", type);
  const bool vec = aligned16(n) && aligned16(m) && aligned16(p) && aligned16(s) && aligned16(q) &&
                   aligned16(z) && aligned16(x) && aligned16(r) && aligned16(u) && aligned16(w);
  const int grid = grid_for(n_elts);
  double *parts = scratch().partials;
  if (type == ISPM_F64) {
    typedef double T;
    launch_ppcg_update<T>(vec, grid, parts, (T *)resnorm, (T *)gamma, (T *)delta, (const T *)alpha, (const T *)beta,
                          (const T *)n, (const T *)m, (T *)p, (T *)s, (T *)q, (T *)z, (T *)x, (T *)r,
                          (T *)u, (T *)w, n_elts);
  } else {
    typedef float T;
    launch_ppcg_update<T>(vec, grid, parts, (T *)resnorm, (T *)gamma, (T *)delta, (const T *)alpha, (const T *)beta,
                          (const T *)n, (const T *)m, (T *)p, (T *)s, (T *)q, (T *)z, (T *)x, (T *)r,
                          (T *)u, (T *)w, n_elts);
  }
  ISPM_HIP(hipGetLastError());
  return ISPM_OK;
}

extern "C" int ispm_ppcg_update_vectors_noalloc(int type, void *resnorm, void *gamma, void *delta,
                                                const void *alpha, const void *beta,
                                                const void *n, const void *m, void *p, void *s,
                                                void *q, void *z, void *x, void *r, void *u,
                                                void *w, int n_elts)
{
  /* no ensure_init(): initialising a device allocates; the caller has done it */
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return fail(ISPM_ERR_DEVICE, "ppcg_update_vectors_noalloc: no usable current device");
  if (!initialized())
    return fail(ISPM_ERR_INVALID, "ppcg_update_vectors_noalloc: the library is not initialised on "
                                  "device %d; call ispm_init() there first (this entry point "
                                  "allocates nothing)", dev);
  double *parts = g_noalloc_parts[dev];
  if (!parts) return fail(ISPM_ERR_RUNTIME, "ppcg_update_vectors_noalloc: no scratch address for device %d", dev);
  if (!resnorm || !gamma || !delta || !alpha || !beta || !n || !m || !p || !s || !q || !z || !x || !r || !u || !w)
    return fail(ISPM_ERR_INVALID, "ppcg_update_vectors_noalloc: null pointer");
  if (n_elts < 0) return fail(ISPM_ERR_INVALID, "ppcg_update_vectors_noalloc: n < 0");
  if (type != ISPM_F64 && type != ISPM_F32)
    return fail(ISPM_ERR_UNSUPPORTED, "ppcg_update_vectors_noalloc: type %d", type);
  const bool vec = aligned16(n) && aligned16(m) && aligned16(p) && aligned16(s) && aligned16(q) &&
                   aligned16(z) && aligned16(x) && aligned16(r) && aligned16(u) && aligned16(w);
  /* the same grid as ispm_ppcg_update_vectors, so both sum the partials in the same order;
   * grid_for caps it at kMaxBlocks (grid-stride loops take any n_elts), which is what the
   * array holds */
  const int grid = grid_for(n_elts);
  if (grid > kMaxBlocks)
    return fail(ISPM_ERR_INVALID, "ppcg_update_vectors_noalloc: %d workgroups exceed the %d partials "
                                  "of the static scratch", grid, kMaxBlocks);
  if (type == ISPM_F64) {
    typedef double T;
    launch_ppcg_update<T>(vec, grid, parts, (T *)resnorm, (T *)gamma, (T *)delta, (const T *)alpha,
                          (const T *)beta, (const T *)n, (const T *)m, (T *)p, (T *)s, (T *)q,
                          (T *)z, (T *)x, (T *)r, (T *)u, (T *)w, n_elts);
  } else {
    typedef float T;
    launch_ppcg_update<T>(vec, grid, parts, (T *)resnorm, (T *)gamma, (T *)delta, (const T *)alpha,
                          (const T *)beta, (const T *)n, (const T *)m, (T *)p, (T *)s, (T *)q,
                          (T *)z, (T *)x, (T *)r, (T *)u, (T *)w, n_elts);
  }
  ISPM_HIP(hipGetLastError());
  return ISPM_OK;
}
