/* convert.h -- host CSR -> SlELL conversion (see convert.cpp). */
#ifndef ISPM_CONVERT_H
#define ISPM_CONVERT_H

#include <cstddef>
#include <vector>

namespace ispm {
namespace conv {

struct layout {
  int n_rows = 0, n_cols = 0, n_nz = 0;   /* n_rows padded to a multiple of 512 */
  unsigned n_slices = 0, S = 0, W = 1;
  bool V = false, D = false;
  std::vector<unsigned> sp, cols;
  size_t n_elms = 0;
  std::vector<long long> map;             /* CSR entry -> element slot */
};

struct guess_info {
  int rows_examined, elts_examined;
  long long waste;
  int n_pairs, n_diags, S, hblock;
  bool var_height, diags;
};

int pad_row_ptr(const int *rp, int n_rows, std::vector<int> &out);

int convert(const int *rp, int n_rows, const int *ci, int n_cols, int n_nz, int elm_size,
            int req_height, int hblock, bool var_height, bool diags, int min_height,
            layout &out);

/* the guess heuristic's sample statistics and decision, without converting */
int guess_stats(const int *rp, int n_rows, const int *ci, int elm_size, int examine_ratio,
                guess_info &out);

/* variable slice heights (estimate_slice_height, raised to min_height) over a padded CSR */
int var_slice_heights(const int *rp_pad, int n_pad, int cap, long long nnz, int min_height,
                      std::vector<int> &h);

int guess(const int *rp, int n_rows, const int *ci, int n_cols, int n_nz, int elm_size,
          int examine_ratio, int min_height, layout &out, guess_info *gi);

}  // namespace conv
}  // namespace ispm

#endif
