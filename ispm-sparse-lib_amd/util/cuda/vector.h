/* vector.h -- host and device vectors tagged by memory space (util/cuda/vector.h:10-89 of
 * the reference): vector<T, host_memory_space_tag> is a std::vector; vector<T,
 * device_memory_space_tag> owns zero-filled device memory, data() gives devmem<T>* and
 * udata() the raw device pointer.  Constructing one space from the other copies. */
#ifndef ISPM_VECTOR_H
#define ISPM_VECTOR_H

#include <cstddef>
#include <utility>
#include <vector>
#include "allocator.h"
#include "cudamemcpy.h"
#include "devmem.h"

struct host_memory_space_tag {};
struct device_memory_space_tag {};

template<typename T, typename memory_space = host_memory_space_tag>
class vector;

template<typename T>
class vector<T, device_memory_space_tag>;

template<typename T>
class vector<T, host_memory_space_tag> : public std::vector<T>
{
  typedef std::vector<T> parent;
public:
  vector() {}
  explicit vector(size_t n) : parent(n) {}
  vector(size_t n, T v) : parent(n, v) {}
  vector(const vector<T, device_memory_space_tag> &v);
  vector &operator=(const vector<T, device_memory_space_tag> &v);
};

template<typename T>
class vector<T, device_memory_space_tag>
{
  devmem<T> *p_ = nullptr;
  size_t n_ = 0;
  void alloc(size_t n)
  {
    void *p = nullptr;
    if (n) ISPM_CHECK(ispm_malloc(&p, n * sizeof(T)));
    p_ = static_cast<devmem<T> *>(p);
    n_ = n;
  }
public:
  vector() {}
  explicit vector(size_t n) { alloc(n); }
  vector(const vector<T, host_memory_space_tag> &v)
  {
    alloc(v.size());
    if (n_) copy(p_, v.data(), n_);
  }
  vector(const vector &v)
  {
    alloc(v.n_);
    if (n_) copy(p_, (const devmem<T> *)v.p_, n_);
  }
  vector(vector &&v) noexcept : p_(v.p_), n_(v.n_) { v.p_ = nullptr; v.n_ = 0; }
  template<typename TI> vector(const TI first, const TI last) { assign(first, last); }
  ~vector()
  {
    if (p_) ispm_free(p_);
  }
  vector &operator=(const vector &v)
  {
    if (this != &v) {
      if (n_ < v.n_) resize(v.n_);
      if (v.n_) copy(p_, (const devmem<T> *)v.p_, v.n_);
    }
    return *this;
  }
  vector &operator=(const vector<T, host_memory_space_tag> &v)
  {
    if (n_ < v.size()) resize(v.size());
    if (v.size()) copy(p_, v.data(), v.size());
    return *this;
  }
  template<typename TI> void assign(const TI first, const TI last)
  {
    if (size_t(last - first) > n_) resize(last - first);
    if (last > first) copy(p_, &*first, last - first);
  }
  template<typename TI> void assign_async(const TI first, const TI last)
  {
    if (size_t(last - first) > n_) resize(last - first);
    if (last > first) copy_async(p_, &*first, last - first);
  }
  /* resize keeps the first min(old, new) elements; new elements are zero */
  void resize(size_t n)
  {
    if (n == n_) return;
    devmem<T> *old = p_;
    size_t on = n_;
    alloc(n);
    if (old) {
      size_t k = on < n ? on : n;
      if (k) copy(p_, (const devmem<T> *)old, k);
      ispm_free(old);
    }
  }
  size_t size() const { return n_; }
  devmem<T> *data() { return p_; }
  const devmem<T> *data() const { return p_; }
  T *udata() { return uncast_devmem(p_); }
  const T *udata() const { return uncast_devmem((const devmem<T> *)p_); }
};

template<typename T>
vector<T, host_memory_space_tag>::vector(const vector<T, device_memory_space_tag> &v) : parent(v.size())
{
  if (v.size()) copy(this->data(), v.data(), v.size());
}

template<typename T>
vector<T, host_memory_space_tag> &
vector<T, host_memory_space_tag>::operator=(const vector<T, device_memory_space_tag> &v)
{
  if (this->size() < v.size()) this->resize(v.size());
  if (v.size()) copy(this->data(), v.data(), v.size());
  return *this;
}

#endif
