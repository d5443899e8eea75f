/* sblas.h -- pipelined-PCG vector primitives (util/cuda/sblas.h:1-56 of the reference), same
 * names and arguments; the definitions live in libispm0 and run on the library stream. */
#ifndef ISPM_SBLAS_H
#define ISPM_SBLAS_H

#include "devmem.h"

namespace sblas
{

/* dst[k] = src[i[k]] */
template<typename T>
void copy_indexed(devmem<T> *dst, const devmem<T> *src, const devmem<int> *i, int n_elts);

/* dst[i[k]] += src[k]; indices must be distinct (no atomics, like the reference) */
template<typename T>
void add_indexed(devmem<T> *dst, const devmem<T> *src, const devmem<int> *i, int n_elts);

/* *s = sum x_k y_k (result in device memory) */
template<typename T>
void dot(const devmem<T> *x, const devmem<T> *y, devmem<T> *s, int n_elts);

/* beta = gamma/gammaold; alpha = gamma / (delta - beta gamma / alpha) */
template<typename T>
void ppcg_update_scalars(devmem<T> *alpha, devmem<T> *beta, const devmem<T> *gamma,
                         const devmem<T> *gammaold, const devmem<T> *delta);

/* z = n + b z; q = m + b q; s = w + b s; p = u + b p; u -= a q; w -= a z; r -= a s;
 * x += a p; resnorm = sum |r|, gamma = (r,u), delta = (w,u) */
template<typename T>
void ppcg_update_vectors(devmem<T> *resnorm, devmem<T> *gamma, devmem<T> *delta,
                         const devmem<T> *alpha, const devmem<T> *beta,
                         const devmem<T> *n, const devmem<T> *m,
                         devmem<T> *p, devmem<T> *s, devmem<T> *q, devmem<T> *z, devmem<T> *x,
                         devmem<T> *r, devmem<T> *u, devmem<T> *w, int n_elts);

#define ISPM_SBLAS_EXTERN(T)                                                                     \
  extern template void copy_indexed<T>(devmem<T> *, const devmem<T> *, const devmem<int> *, int); \
  extern template void add_indexed<T>(devmem<T> *, const devmem<T> *, const devmem<int> *, int);  \
  extern template void dot<T>(const devmem<T> *, const devmem<T> *, devmem<T> *, int);            \
  extern template void ppcg_update_scalars<T>(devmem<T> *, devmem<T> *, const devmem<T> *,       \
                                              const devmem<T> *, const devmem<T> *);             \
  extern template void ppcg_update_vectors<T>(devmem<T> *, devmem<T> *, devmem<T> *,             \
                                              const devmem<T> *, const devmem<T> *,              \
                                              const devmem<T> *, const devmem<T> *, devmem<T> *, \
                                              devmem<T> *, devmem<T> *, devmem<T> *, devmem<T> *, \
                                              devmem<T> *, devmem<T> *, devmem<T> *, int);
ISPM_SBLAS_EXTERN(float)
ISPM_SBLAS_EXTERN(double)
#undef ISPM_SBLAS_EXTERN

}  // namespace sblas

#endif
