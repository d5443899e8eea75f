/* conversion.h -- CSR -> sliced-ELLPACK on the host (formats/conversion.h:338-492 of the
 * reference): csr_to_slell and csr_to_slell_guess_format, same arguments.  As in the
 * reference the caller's CSR is padded in place to a multiple of 512 rows; the conversion
 * itself runs in libispm0 (ispm_csr_to_slell). */
#ifndef ISPM_FORMAT_CONVERSION_H
#define ISPM_FORMAT_CONVERSION_H

#include <algorithm>
#include "formats/formats.h"

enum { MIN_SLICE_HEIGHT = 32, MAX_SLICE_HEIGHT = 512 };

namespace ispm_detail {
template<class T>
inline ispm_csr csr_view(const csr_matrix<T> &A)
{
  ispm_csr c;
  c.n_rows = A.n_rows; c.n_cols = A.n_cols; c.n_nz = A.n_nz;
  c.row_ptr = A.row_ptr.data(); c.cols = A.cols.data(); c.elms = A.elms.data();
  c.elm_type = ispm_type_code<T>::value;
  return c;
}

template<class E>
inline slell_matrix<E, host_memory_space_tag> take(ispm_slell_host &h)
{
  slell_matrix<E, host_memory_space_tag> B;
  B.n_rows = h.n_rows; B.n_cols = h.n_cols; B.n_nz = h.n_nz;
  B.slice_height = h.slice_height; B.n_slices = h.n_slices; B.hblock = h.hblock;
  B.var_height = h.var_height != 0; B.diags = h.diags != 0;
  B.slice_ptr.assign(h.slice_ptr, h.slice_ptr + h.slice_ptr_len);
  B.cols.assign(h.cols, h.cols + h.cols_len);
  const E *e = static_cast<const E *>(h.elms);
  B.elms.assign(e, e + h.elms_len);
  ispm_slell_host_free(&h);
  return B;
}
}  // namespace ispm_detail

template<class T>
slell_matrix<T, host_memory_space_tag>
csr_to_slell(const csr_matrix<T> &A, int req_slice_height, int hblock, bool var_height,
             bool use_diags, int min_height = 4)
{
  const_cast<csr_matrix<T> &>(A).pad(MAX_SLICE_HEIGHT);
  ispm_csr c = ispm_detail::csr_view(A);
  ispm_slell_host h;
  ISPM_CHECK(ispm_csr_to_slell(&c, req_slice_height, hblock, var_height, use_diags, min_height,
                               ispm_type_code<T>::value, &h));
  return ispm_detail::take<T>(h);
}

template<class T>
slell_matrix<T, host_memory_space_tag>
csr_to_slell_guess_format(const csr_matrix<T> &A, int examine_ratio = 1, int min_height = 4)
{
  ispm_csr c = ispm_detail::csr_view(A);
  ispm_slell_host h;
  ISPM_CHECK(ispm_csr_to_slell_guess(&c, examine_ratio, min_height, ispm_type_code<T>::value, &h));
  const_cast<csr_matrix<T> &>(A).pad(MAX_SLICE_HEIGHT);
  return ispm_detail::take<T>(h);
}

#endif
