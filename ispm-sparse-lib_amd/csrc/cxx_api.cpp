/* cxx_api.cpp -- the reference's link-level C++ symbols, defined over the C-ABI:
 * launch_spmv_kernel<T,E> (cuda/spmv/kernel-iface.h:20-74), sblas::* (util/cuda/sblas.h),
 * fastainv_sym<double, float|double> (fastainv/fastainv.h:8-30), ispm_initialize() and
 * get_device_string() (util/cuda/initialize.h). */
#include <cstring>
#include <vector>
#include "fastainv/fastainv.h"
#include "fastainv_impl.h"
#include "plan/plan.h"
#include "util/cuda/initialize.h"
#include "util/cuda/sblas.h"

template<class T, class E>
void launch_spmv_kernel(spmv_launch_params &params, slell_matrix<E, device_memory_space_tag> &m,
                        T *xptr, T *yptr)
{
  ispm_launch p{params.N, params.S, params.H, params.use_tex_sliceptr, params.use_tex_x, params.mode};
  ispm_slell_desc d = m.ispm_desc();
  int st = ispm_spmv(&p, &d, ispm_type_code<T>::value, xptr, yptr);
  /* invalid or unsupported launch parameters leave y unwritten, as in the reference */
  if (st && st != ISPM_ERR_UNSUPPORTED) ispm_report_error(st);
}

template<class T, class E>
void launch_spmv_kernel(spmv_launch_params &params, slell_matrix<E, device_memory_space_tag> &m,
                        vector<T, device_memory_space_tag> &x, vector<T, device_memory_space_tag> &y)
{
  launch_spmv_kernel<T, E>(params, m, x.udata(), y.udata());
}

#define ISPM_LAUNCH_INST(T, E)                                                                       \
  template void launch_spmv_kernel<T, E>(spmv_launch_params &, slell_matrix<E, device_memory_space_tag> &, \
                                         vector<T, device_memory_space_tag> &,                       \
                                         vector<T, device_memory_space_tag> &);                      \
  template void launch_spmv_kernel<T, E>(spmv_launch_params &, slell_matrix<E, device_memory_space_tag> &, T *, T *);
ISPM_LAUNCH_INST(float, float)
ISPM_LAUNCH_INST(double, float)
ISPM_LAUNCH_INST(double, double)
ISPM_LAUNCH_INST(float, ispm_bf16)

namespace sblas
{
template<typename T> static constexpr int code() { return ispm_type_code<T>::value; }

template<typename T>
void copy_indexed(devmem<T> *dst, const devmem<T> *src, const devmem<int> *i, int n)
{
  ISPM_CHECK(ispm_copy_indexed(code<T>(), dst, src, uncast_devmem(i), n));
}
template<typename T>
void add_indexed(devmem<T> *dst, const devmem<T> *src, const devmem<int> *i, int n)
{
  ISPM_CHECK(ispm_add_indexed(code<T>(), dst, src, uncast_devmem(i), n));
}
template<typename T>
void dot(const devmem<T> *x, const devmem<T> *y, devmem<T> *s, int n)
{
  ISPM_CHECK(ispm_dot(code<T>(), x, y, s, n));
}
template<typename T>
void ppcg_update_scalars(devmem<T> *alpha, devmem<T> *beta, const devmem<T> *gamma,
                         const devmem<T> *gammaold, const devmem<T> *delta)
{
  ISPM_CHECK(ispm_ppcg_update_scalars(code<T>(), alpha, beta, gamma, gammaold, delta));
}
template<typename T>
void ppcg_update_vectors(devmem<T> *resnorm, devmem<T> *gamma, devmem<T> *delta,
                         const devmem<T> *alpha, const devmem<T> *beta, const devmem<T> *n,
                         const devmem<T> *m, devmem<T> *p, devmem<T> *s, devmem<T> *q,
                         devmem<T> *z, devmem<T> *x, devmem<T> *r, devmem<T> *u, devmem<T> *w,
                         int n_elts)
{
  ISPM_CHECK(ispm_ppcg_update_vectors(code<T>(), resnorm, gamma, delta, alpha, beta, n, m, p, s, q, z,
                                      x, r, u, w, n_elts));
}

#define ISPM_SBLAS_INST(T)                                                                     \
  template void copy_indexed<T>(devmem<T> *, const devmem<T> *, const devmem<int> *, int);      \
  template void add_indexed<T>(devmem<T> *, const devmem<T> *, const devmem<int> *, int);       \
  template void dot<T>(const devmem<T> *, const devmem<T> *, devmem<T> *, int);                 \
  template void ppcg_update_scalars<T>(devmem<T> *, devmem<T> *, const devmem<T> *,            \
                                      const devmem<T> *, const devmem<T> *);                   \
  template void ppcg_update_vectors<T>(devmem<T> *, devmem<T> *, devmem<T> *, const devmem<T> *, \
                                      const devmem<T> *, const devmem<T> *, const devmem<T> *, \
                                      devmem<T> *, devmem<T> *, devmem<T> *, devmem<T> *,      \
                                      devmem<T> *, devmem<T> *, devmem<T> *, devmem<T> *, int);
ISPM_SBLAS_INST(float)
ISPM_SBLAS_INST(double)
}  // namespace sblas

template<class in_real, class out_real>
static void to_csr(csr_matrix<out_real> &M, int n, std::vector<int> &r, std::vector<int> &c,
                   std::vector<out_real> &v)
{
  M.n_rows = M.n_cols = n;
  M.n_nz = (int)c.size();
  M.row_ptr.assign(r.begin(), r.end());
  M.cols.assign(c.begin(), c.end());
  M.elms.assign(v.begin(), v.end());
}

static bool factor(int n, const int *rp, const int *ci, const double *a, long long max_elts,
                   double droptol, progress_f progress, void *v, std::vector<int> &r1,
                   std::vector<int> &c1, std::vector<float> &v1, std::vector<int> &r2,
                   std::vector<int> &c2, std::vector<float> &v2)
{
  return ispm::fastainv_f64_f32(n, rp, ci, a, max_elts, droptol, progress, v, r1, c1, v1, r2, c2, v2);
}

static bool factor(int n, const int *rp, const int *ci, const double *a, long long max_elts,
                   double droptol, progress_f progress, void *v, std::vector<int> &r1,
                   std::vector<int> &c1, std::vector<double> &v1, std::vector<int> &r2,
                   std::vector<int> &c2, std::vector<double> &v2)
{
  return ispm::fastainv_f64_f64(n, rp, ci, a, max_elts, droptol, progress, v, r1, c1, v1, r2, c2, v2);
}

template<class in_real, class out_real>
bool fastainv_sym(csr_matrix<out_real> &Ainv1, csr_matrix<out_real> &Ainv2,
                  const csr_matrix<in_real> &A, progress_f progress, void *v,
                  long long max_elts, double droptol)
{
  std::vector<int> r1, c1, r2, c2;
  std::vector<out_real> v1, v2;
  if (!factor(A.n_rows, A.row_ptr.data(), A.cols.data(), A.elms.data(), max_elts, droptol,
              progress, v, r1, c1, v1, r2, c2, v2))
    return false;
  to_csr<in_real, out_real>(Ainv1, A.n_rows, r1, c1, v1);
  to_csr<in_real, out_real>(Ainv2, A.n_rows, r2, c2, v2);
  return true;
}

template bool fastainv_sym<double, float>(csr_matrix<float> &, csr_matrix<float> &,
                                          const csr_matrix<double> &, progress_f, void *,
                                          long long, double);
template bool fastainv_sym<double, double>(csr_matrix<double> &, csr_matrix<double> &,
                                           const csr_matrix<double> &, progress_f, void *,
                                           long long, double);

void ispm_initialize()
{
  int st = ispm_init();
  if (st) ispm_report_error(st);
}

DeviceIDString get_device_string()
{
  DeviceIDString s;
  std::memset(s.str, 0, sizeof s.str);
  ISPM_CHECK(ispm_device_string(s.str, sizeof s.str));
  return s;
}
