/* pinner.h -- device aliases of mapped pinned host memory (util/cuda/pinner.h:7-48 of the
 * reference).  Memory must come from ispm_pinned_allocator / ispm_host_alloc_mapped:
 * registering memory allocated elsewhere (the reference's cuda_pinned_region) is not
 * provided. */
#ifndef ISPM_PINNER_H
#define ISPM_PINNER_H

#include "check.h"
#include "devmem.h"

template<typename T>
static inline devmem<T> *host2devmem(T *p)
{
  void *d = nullptr;
  ISPM_CHECK(ispm_host_device_pointer(&d, (void *)p));
  return static_cast<devmem<T> *>(d);
}

template<typename T>
static inline const devmem<T> *host2devmem(const T *p)
{
  void *d = nullptr;
  ISPM_CHECK(ispm_host_device_pointer(&d, (void *)p));
  return static_cast<const devmem<T> *>(d);
}

template<typename T>
class pinmemptr
{
  T *ptr_;
  devmem<T> *dev_;
public:
  explicit pinmemptr(T *p) : ptr_(p), dev_(host2devmem(p)) {}
  T &operator*() { return *ptr_; }
  devmem<T> *operator()() { return dev_; }
};

#endif
