/* formats.h -- host CSR and sliced-ELLPACK matrix types (formats/formats.h:6-102 of the
 * reference), same members.  slell_matrix<E, device_memory_space_tag> constructed from a
 * host one uploads it; ispm_desc() exposes the device view the C-ABI takes. */
#ifndef ISPM_FORMATS_H
#define ISPM_FORMATS_H

#include <cstdint>
#include "util/cuda/vector.h"

/* bfloat16 storage type for matrix elements (FP32 with the low 16 bits dropped) */
struct ispm_bf16 {
  uint16_t bits;
};

template<typename E> struct ispm_type_code;
template<> struct ispm_type_code<float> { enum { value = ISPM_F32 }; };
template<> struct ispm_type_code<double> { enum { value = ISPM_F64 }; };
template<> struct ispm_type_code<ispm_bf16> { enum { value = ISPM_BF16 }; };

struct matrix_shape
{
  int n_rows, n_cols, n_nz;
};

template<class T>
struct coo_matrix : matrix_shape
{
  vector<int> rows, cols;
  vector<T> elms;
};

template<class T>
struct csr_matrix : matrix_shape
{
  vector<int, host_memory_space_tag> row_ptr, cols;
  vector<T, host_memory_space_tag> elms;

  /* append empty rows up to a multiple of k (a power of two); n_cols is unchanged */
  void pad(int k)
  {
    k--;
    const int target = (n_rows + k) & ~k;
    const int last = row_ptr.back();
    for (; n_rows < target; n_rows++) row_ptr.push_back(last);
  }

  /* 4(n+1) + 4 nnz + sizeof(T) (nnz + 2 n) */
  unsigned long long spmv_bytes() const
  {
    return sizeof(unsigned) * (unsigned long long)(row_ptr.size() + cols.size()) +
           sizeof(T) * (unsigned long long)(elms.size() + 2 * (size_t)n_rows);
  }
};

struct slell_params
{
  unsigned S, H, V, D;   /* slice height, hblock width, variable height, diagonal slices */
};

template<typename E, typename memory_space>
struct slell_matrix : matrix_shape
{
  vector<unsigned, memory_space> slice_ptr;
  vector<unsigned, memory_space> cols;
  vector<E, memory_space> elms;
  unsigned slice_height;
  unsigned n_slices;
  unsigned hblock;
  bool var_height;
  bool diags;

  slell_matrix(const slell_matrix<E, host_memory_space_tag> &m)
   : matrix_shape(m), slice_ptr(m.slice_ptr), cols(m.cols), elms(m.elms),
     slice_height(m.slice_height), n_slices(m.n_slices), hblock(m.hblock),
     var_height(m.var_height), diags(m.diags)
  {}
  slell_matrix(const matrix_shape &m)
   : matrix_shape(m), slice_height(0), n_slices(0), hblock(1), var_height(false), diags(false) {}
  slell_matrix() : matrix_shape{0, 0, 0}, slice_height(0), n_slices(0), hblock(1), var_height(false), diags(false) {}

  /* stored bytes of one SpMV: slice records, columns, elements, and x plus y once with
   * vec_elt_size bytes per vector entry (the reference counts 4 bytes there,
   * formats.h:77); effective = count nnz instead of stored slots */
  unsigned long long spmv_bytes(bool effective = false, int vec_elt_size = sizeof(E))
  {
    return sizeof(unsigned) * (unsigned long long)slice_ptr.size() +
           sizeof(unsigned) * (unsigned long long)(effective ? (size_t)n_nz : cols.size()) +
           sizeof(E) * (unsigned long long)(effective ? (size_t)n_nz : elms.size()) +
           (unsigned long long)vec_elt_size * 2 * (unsigned long long)n_rows;
  }

  /* address of the first element of the slice starting at `row` (NULL if no slice starts
   * there); for variable-height slices the search starts at slice_hint and updates it */
  E *slice_elt_ptr(unsigned row, unsigned &slice_hint)
  {
    const unsigned P = slice_ptr.size() / (n_slices + 1);
    unsigned slice;
    if (!var_height) {
      slice = row / slice_height;
      if (row % slice_height) return nullptr;
    } else {
      for (slice = slice_hint; slice <= n_slices && slice_ptr[slice * P + 1] != row; slice++) {}
      if (slice > n_slices) return nullptr;
      slice_hint = slice;
    }
    return &elms[slice_ptr[slice * P]];
  }

  /* device view for the C-ABI (device-space matrices only) */
  ispm_slell_desc ispm_desc()
  {
    ispm_slell_desc d;
    d.n_rows = n_rows; d.n_cols = n_cols; d.n_nz = n_nz;
    d.n_slices = n_slices; d.slice_height = slice_height; d.hblock = hblock;
    d.var_height = var_height; d.diags = diags;
    d.slice_ptr = slice_ptr.udata(); d.slice_ptr_len = slice_ptr.size();
    d.cols = cols.udata(); d.cols_len = cols.size();
    d.elms = elms.udata(); d.elms_len = elms.size();
    d.elm_type = ispm_type_code<E>::value;
    return d;
  }
};

#endif
