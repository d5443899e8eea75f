/* check.h -- error check for C-ABI calls: print the library's message and abort, as the
 * reference's CUDA_CHECK does for runtime errors (util/cuda/check.h:8-15). */
#ifndef ISPM_CHECK_H
#define ISPM_CHECK_H

#include <cstdio>
#include <cstdlib>
#include "ispm.h"

inline void ispm_report_error(int code)
{
  std::fprintf(stderr, "ispm error %d: %s\n", code, ispm_last_error());
  std::abort();
}

#define ISPM_CHECK(x)                   \
  do {                                  \
    int ispm_e_ = (x);                  \
    if (ispm_e_) ispm_report_error(ispm_e_); \
  } while (0)

#endif
