/* pinner.h -- device aliases of mapped pinned host memory (util/cuda/pinner.h:7-48 of the
 * reference).  Memory must come from cuda_pinned_allocator (= ispm_pinned_allocator) or
 * ispm_host_alloc_mapped.  Registering memory allocated elsewhere (the reference's
 * cuda_pinned_region, pinner.h:50-71) is not available on this target: the class exists so
 * that a caller naming it in a declaration compiles, but constructing one from a pointer
 * fails to compile with the message below. */
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

template<typename T>
class cuda_pinned_region
{
  template<typename> struct never { static const bool value = false; };
public:
  cuda_pinned_region() {}
  cuda_pinned_region(const cuda_pinned_region &) = delete;
  cuda_pinned_region &operator=(const cuda_pinned_region &) = delete;
  template<typename P>
  cuda_pinned_region(P *, size_t)
  {
    static_assert(never<P>::value, "cuda_pinned_region: registering existing host memory is not "
                                   "supported on this target; allocate it with cuda_pinned_allocator "
                                   "and take its device alias with host2devmem");
  }
  devmem<T> *operator()() { return nullptr; }
};

#endif
