/* fastainv_impl.h -- host approximate-inverse factorisation (fastainv.cpp), C++ entry points
 * used by the C-ABI and by the fastainv_sym<> template instantiations. */
#ifndef ISPM_FASTAINV_IMPL_H
#define ISPM_FASTAINV_IMPL_H

#include <vector>
#include "ispm.h"

namespace ispm {
bool fastainv_f64_f32(int n, const int *rp, const int *ci, const double *a, long long max_elts,
                      double droptol, void (*progress)(void *), void *pv, std::vector<int> &r1,
                      std::vector<int> &c1, std::vector<float> &v1, std::vector<int> &r2,
                      std::vector<int> &c2, std::vector<float> &v2);
bool fastainv_f64_f64(int n, const int *rp, const int *ci, const double *a, long long max_elts,
                      double droptol, void (*progress)(void *), void *pv, std::vector<int> &r1,
                      std::vector<int> &c1, std::vector<double> &v1, std::vector<int> &r2,
                      std::vector<int> &c2, std::vector<double> &v2);
}

#endif
