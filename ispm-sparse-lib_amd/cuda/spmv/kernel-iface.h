/* kernel-iface.h (reference include path cuda/spmv/) -- SpMV launch parameters and the
 * launcher (cuda/spmv/kernel-iface.h:4-74 of the reference).  N = threads per workgroup
 * (64..1024), S = slice height, H = rows per thread (1, 2); a launch is valid iff
 * N >= S/H and S >= H, otherwise it does nothing (dispatch.h:202-203). */
#ifndef ISPM_KERNEL_IFACE_H
#define ISPM_KERNEL_IFACE_H

#include "formats/formats.h"

#define SPMV_PARAMS_VALID_P(N, S, H) ((N) >= (S) / (H) && (S) >= (H))

struct spmv_launch_params
{
  unsigned N;
  unsigned S;
  unsigned H;
  bool use_tex_sliceptr;   /* accepted, no effect: x is read through the cache hierarchy */
  bool use_tex_x;
  int mode;                /* 0: workgroup per slice (reference mapping); 1: packed rows for
                              slice heights below 64; 2: streamed elements for plain
                              slices up to 256 rows (see ispm_launch) */

  bool valid() const { return SPMV_PARAMS_VALID_P(N, S, H); }
};

template<class T, class E>
void launch_spmv_kernel(spmv_launch_params &params, slell_matrix<E, device_memory_space_tag> &m,
                        vector<T, device_memory_space_tag> &x, vector<T, device_memory_space_tag> &y);

template<class T, class E>
void launch_spmv_kernel(spmv_launch_params &params, slell_matrix<E, device_memory_space_tag> &m,
                        T *xptr, T *yptr);

#define ISPM_LAUNCH_EXTERN(T, E)                                                                      \
  extern template void launch_spmv_kernel<T, E>(spmv_launch_params &, slell_matrix<E, device_memory_space_tag> &, \
                                                vector<T, device_memory_space_tag> &,                \
                                                vector<T, device_memory_space_tag> &);               \
  extern template void launch_spmv_kernel<T, E>(spmv_launch_params &, slell_matrix<E, device_memory_space_tag> &, \
                                                T *, T *);
ISPM_LAUNCH_EXTERN(float, float)
ISPM_LAUNCH_EXTERN(double, float)
ISPM_LAUNCH_EXTERN(double, double)
ISPM_LAUNCH_EXTERN(float, ispm_bf16)
#undef ISPM_LAUNCH_EXTERN

#endif
