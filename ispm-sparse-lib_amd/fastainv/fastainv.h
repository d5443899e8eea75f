/* fastainv.h -- host approximate inverse M = Ainv2 * Ainv1 ~= A^-1 (fastainv/fastainv.h:6-30
 * of the reference), same signature and defaults; built into libispm0. */
#ifndef ISPM_FASTAINV_H
#define ISPM_FASTAINV_H

#include "formats/formats.h"

typedef void (*progress_f)(void *);

template<class in_real, class out_real>
bool fastainv_sym(csr_matrix<out_real> &Ainv1, csr_matrix<out_real> &Ainv2,
                  const csr_matrix<in_real> &A, progress_f progress, void *v,
                  long long max_elts = 8192000, double droptol = 0.125);

extern template bool fastainv_sym<double, float>(csr_matrix<float> &, csr_matrix<float> &,
                                                 const csr_matrix<double> &, progress_f, void *,
                                                 long long, double);
extern template bool fastainv_sym<double, double>(csr_matrix<double> &, csr_matrix<double> &,
                                                  const csr_matrix<double> &, progress_f, void *,
                                                  long long, double);

#endif
