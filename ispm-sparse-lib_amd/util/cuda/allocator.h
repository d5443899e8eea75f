/* allocator.h -- std::allocator-compatible device (zero-filled) and mapped-pinned host
 * allocators (util/cuda/allocator.h:7-78 of the reference) over the C-ABI. */
#ifndef ISPM_ALLOCATOR_H
#define ISPM_ALLOCATOR_H

#include <cstddef>
#include <memory>
#include "check.h"

template<typename T>
class ispm_device_allocator
{
public:
  typedef T value_type;
  typedef T *pointer;
  typedef size_t size_type;
  ispm_device_allocator() {}
  template<typename O> ispm_device_allocator(const ispm_device_allocator<O> &) {}
  T *allocate(size_t n)
  {
    void *p = nullptr;
    ISPM_CHECK(ispm_malloc(&p, n * sizeof(T)));
    return static_cast<T *>(p);
  }
  void deallocate(T *p, size_t) { ISPM_CHECK(ispm_free(p)); }
  /* device memory is never touched from the host: elements are not constructed */
  template<typename U, typename... A> void construct(U *, A &&...) {}
  template<typename U> void destroy(U *) {}
  template<typename O> struct rebind { typedef ispm_device_allocator<O> other; };
};
template<typename A, typename B>
inline bool operator==(const ispm_device_allocator<A> &, const ispm_device_allocator<B> &) { return true; }
template<typename A, typename B>
inline bool operator!=(const ispm_device_allocator<A> &, const ispm_device_allocator<B> &) { return false; }

template<typename T>
class ispm_pinned_allocator
{
public:
  typedef T value_type;
  typedef T *pointer;
  typedef size_t size_type;
  ispm_pinned_allocator() {}
  template<typename O> ispm_pinned_allocator(const ispm_pinned_allocator<O> &) {}
  T *allocate(size_t n)
  {
    void *p = nullptr;
    ISPM_CHECK(ispm_host_alloc_mapped(&p, n * sizeof(T)));
    return static_cast<T *>(p);
  }
  void deallocate(T *p, size_t) { ISPM_CHECK(ispm_host_free(p)); }
  template<typename O> struct rebind { typedef ispm_pinned_allocator<O> other; };
};
template<typename A, typename B>
inline bool operator==(const ispm_pinned_allocator<A> &, const ispm_pinned_allocator<B> &) { return true; }
template<typename A, typename B>
inline bool operator!=(const ispm_pinned_allocator<A> &, const ispm_pinned_allocator<B> &) { return false; }

/* the reference's names for the two allocators (util/cuda/allocator.h:8,45) */
template<typename T> using cuda_device_allocator = ispm_device_allocator<T>;
template<typename T> using cuda_pinned_allocator = ispm_pinned_allocator<T>;

#endif
