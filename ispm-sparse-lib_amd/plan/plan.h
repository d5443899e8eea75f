/* plan.h -- SpMV plan (plan/plan.h:1-161 of the reference), same members and methods.
 * PLAN_DEFAULT and PLAN_ESTIMATE (one case, as in plan.h:74-75 of the reference) take the
 * guessed format with a CDNA4 launch shape; PLAN_EXHAUSTIVE times every format and launch
 * shape on the device (ispm_plan_tune).  The fewest-stored-bytes planner is reached through
 * the C-ABI only (ISPM_PLAN_MIN_BYTES, ispm_plan_estimate). */
#ifndef ISPM_PLAN_H
#define ISPM_PLAN_H

#include <cstdio>
#include <cstdlib>
#include "formats/formats.h"
#include "formats/conversion.h"
#include "cuda/spmv/kernel-iface.h"
#include "util/cuda/timer.h"

enum spmv_plan_kind {
  PLAN_DEFAULT,
  PLAN_ESTIMATE,
  PLAN_EXHAUSTIVE,
};

enum spmv_plan_flags {
  SPMV_NO_TEXTURE_MEMORY = 001,
  SPMV_PLAN_NO_DEVICE = 002
};

namespace ispm_detail {
/* the minimum slice height plans convert with: a tuned fixed height below 4 (the library
 * times heights from 1) is taken as it is instead of being refused */
inline int plan_min_height(unsigned S, unsigned V) { return !V && S < 4 ? (int)S : 4; }
}  // namespace ispm_detail

template<class T, class E = T>
struct spmv_plan
{
  slell_matrix<E, host_memory_space_tag> *host_matrix;
  slell_matrix<E, device_memory_space_tag> *device_matrix;
  slell_params mparams;
  spmv_launch_params params;

  spmv_plan() : host_matrix(nullptr), device_matrix(nullptr) {}
  spmv_plan(const csr_matrix<E> &m, spmv_plan_kind kind = PLAN_DEFAULT, unsigned flags = 0);
  /* a plan owns its matrices: copying would free them twice */
  spmv_plan(const spmv_plan &) = delete;
  spmv_plan &operator=(const spmv_plan &) = delete;
  ~spmv_plan()
  {
    delete host_matrix;
    delete device_matrix;
  }

  void update_device_matrix_coefs() { device_matrix->elms = host_matrix->elms; }

  void upload()
  {
    delete device_matrix;
    device_matrix = new slell_matrix<E, device_memory_space_tag>(*host_matrix);
  }

  void change_host_matrix(const csr_matrix<E> &m)
  {
    delete host_matrix;
    host_matrix = new slell_matrix<E, host_memory_space_tag>(
        csr_to_slell(m, mparams.S, mparams.H, mparams.V, mparams.D, ispm_detail::plan_min_height(mparams.S, mparams.V)));
  }

  void execute_spmv(vector<T, device_memory_space_tag> &x, vector<T, device_memory_space_tag> &y)
  {
    launch_spmv_kernel<T, E>(params, *device_matrix, x, y);
  }

  void execute_spmv(T *xptr, T *yptr) { launch_spmv_kernel<T, E>(params, *device_matrix, xptr, yptr); }
};

template<class T, class E>
spmv_plan<T, E>::spmv_plan(const csr_matrix<E> &m, spmv_plan_kind kind, unsigned flags)
  : host_matrix(nullptr), device_matrix(nullptr)
{
  const bool usetex = !(flags & SPMV_NO_TEXTURE_MEMORY);
  const bool up = !(flags & SPMV_PLAN_NO_DEVICE);
  ispm_launch lp;
  if (kind == PLAN_EXHAUSTIVE) {
    ispm_csr c = ispm_detail::csr_view(m);
    unsigned S, W;
    int V, D;
    double secs;
    ISPM_CHECK(ispm_plan_tune(ispm_type_code<T>::value, ispm_type_code<E>::value, &c, &S, &W, &V, &D, &lp, &secs));
    mparams = slell_params{S, W, (unsigned)V, (unsigned)D};
    host_matrix = new slell_matrix<E, host_memory_space_tag>(csr_to_slell(m, S, W, V != 0, D != 0, ispm_detail::plan_min_height(S, V)));
    upload();
  } else {
    /* PLAN_DEFAULT, PLAN_ESTIMATE */
    host_matrix = new slell_matrix<E, host_memory_space_tag>(csr_to_slell_guess_format(m));
    if (up) upload();
    mparams = slell_params{host_matrix->slice_height, host_matrix->hblock,
                           (unsigned)host_matrix->var_height, (unsigned)host_matrix->diags};
    ISPM_CHECK(ispm_plan_default_launch(host_matrix->slice_height, host_matrix->var_height, &lp));
  }
  params = spmv_launch_params{lp.N, lp.S, lp.H, usetex, usetex, lp.mode};
}

#endif
