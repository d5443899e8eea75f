/* cudamemcpy.h (reference include path) -- direction-typed copies: host <-> device is chosen
 * from the pointer types as in util/cuda/cudamemcpy.h:8-43 of the reference, and performed
 * by the library (ispm_memcpy / ispm_memcpy_async on the library stream). */
#ifndef ISPM_MEMCPY_H
#define ISPM_MEMCPY_H

#include <cstddef>
#include "check.h"
#include "devmem.h"

struct ispm_copy_direction
{
  template<class T> int operator()(devmem<T> *, const devmem<T> *) { return ISPM_D2D; }
  template<class T> int operator()(devmem<T> *, const T *) { return ISPM_H2D; }
  template<class T> int operator()(T *, const devmem<T> *) { return ISPM_D2H; }
};

template<typename T, typename S>
static inline void copy(T *dst, const S *src, size_t n_elts)
{
  ISPM_CHECK(ispm_memcpy((void *)dst, (const void *)src, n_elts * sizeof(T), ispm_copy_direction()(dst, src)));
}

template<typename T, typename S>
static inline void copy_async(T *dst, const S *src, size_t n_elts)
{
  ISPM_CHECK(ispm_memcpy_async((void *)dst, (const void *)src, n_elts * sizeof(T), ispm_copy_direction()(dst, src)));
}

#endif
