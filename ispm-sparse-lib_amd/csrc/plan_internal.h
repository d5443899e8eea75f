/* plan_internal.h -- the SpMV plan object behind the opaque ispm_plan handle. */
#ifndef ISPM_PLAN_INTERNAL_H
#define ISPM_PLAN_INTERNAL_H

#include <vector>
#include "convert.h"
#include "diag_pattern.h"
#include "ispm.h"
#include "internal.h"

struct ispm_plan {
  int T = ISPM_F64, E = ISPM_F64;         /* vector and element types */
  ispm::conv::layout L;                   /* layout + CSR-entry -> slot map */
  std::vector<unsigned char> elms;        /* host copy of the element values */
  ispm_slell_desc dev{};
  bool has_dev = false;
  ispm_launch launch{};
  double tuned = 0;
  size_t stored = 0, csr = 0;
  void *dev_map = nullptr;                /* device copy of the slot map (on-device refresh) */
  bool has_pat = false;                   /* diagonal pattern shared by most slices */
  ispm::diag_pattern pat{};
  unsigned *dev_uni = nullptr;            /* device bit mask of the slices that have it */
  unsigned pat_slices = 0;
  /* split-diagonal plan (ISPM_SPMV_PLAN_SPLIT_DIAG): L / dev / launch describe the remainder */
  bool split = false;
  void *dev_diag = nullptr;               /* split_n entries of E */
  unsigned *dev_rmap = nullptr;           /* remainder row -> matrix row (~0u: padding) */
  unsigned *dev_rmask = nullptr;          /* bit per matrix row: a remainder row */
  unsigned split_n = 0, split_m = 0;      /* padded rows of the matrix; remainder rows */
  ispm::layout_hints hints{};             /* launch facts of the layout (streamed mapping) */
  unsigned *dev_xwin = nullptr;           /* hblock x windows (hints.xwin) */
  /* symmetric diagonal plan (ISPM_SPMV_PLAN_SYMMETRIC): dev.elms holds the upper planes */
  bool sym = false;
  ispm::sym_desc symd{};
};

/* rows (padded) of the matrix a plan multiplies */
inline unsigned plan_rows(const ispm_plan *p) { return p->split ? p->split_n : (unsigned)p->L.n_rows; }

namespace ispm {
size_t tsize(int t);
unsigned short f32_to_bf16(float f);
int fill_values(const conv::layout &L, const ispm_csr *A, int et, void *elms);
int upload_layout(const conv::layout &L, const void *elms, int et, ispm_slell_desc *d);
ispm_launch default_launch(const conv::layout &L);
bool build_pattern(const conv::layout &L, diag_pattern &pat, std::vector<unsigned> &mask);
void compute_hints(const conv::layout &L, layout_hints &h);
int upload_xwin(const conv::layout &L, const ispm_csr *A, int vt, unsigned *&dev, layout_hints &h);
}

#endif
