/* devmem.h -- opaque element of device memory: a devmem<T>* is a device address that host
 * code cannot dereference by mistake (util/cuda/devmem.h:4-22 of the reference). */
#ifndef ISPM_DEVMEM_H
#define ISPM_DEVMEM_H

template<class T>
class devmem
{
  unsigned char opaque_[sizeof(T)];
};

template<typename T>
static inline T *uncast_devmem(devmem<T> *p) { return reinterpret_cast<T *>(p); }

template<typename T>
static inline const T *uncast_devmem(const devmem<T> *p) { return reinterpret_cast<const T *>(p); }

#endif
