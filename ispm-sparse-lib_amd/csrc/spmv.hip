/* spmv.hip -- C-ABI entry point of the single-vector SlELL SpMV (kernel in spmv_kernel.h). */
#include "spmv_kernel.h"

namespace ispm {
/* defined in spmm.hip */
int spmm_dispatch(const ispm_launch *p, const ispm_slell_desc *A, int vec_type, int k,
                  const void *X, size_t ldx, void *Y, size_t ldy, const layout_hints *hints);
}

namespace ispm {
/* SpMV for the solver and plans: same checks as ispm_spmv, optional device skip flag, and
 * the plan's shared diagonal pattern when it has one. */
int spmv_internal(const ispm_launch *p, const ispm_slell_desc *A, int vec_type, const void *x,
                  void *y, const unsigned *skip, const diag_pattern *pat, const unsigned *uni,
                  const layout_hints *hints)
{
  unsigned H;
  int st = spmvk::check_args(p, A, vec_type, x, y, H);
  if (st) return st;
  if (A->n_slices == 0) return ISPM_OK;
  if (spmvk::use_diagpat(p, A, vec_type, H, pat, uni)) return spmvk::launch_diagpat_types(A, vec_type, *pat, uni, x, y, skip);
  if (spmvk::use_stream(p, A))
    return spmvk::launch_stream_types(A, vec_type, x, y, skip, nullptr, nullptr, nullptr, hints);
  if (spmvk::use_hwin(p, A, vec_type, H, hints)) return spmvk::launch_hwin_types(A, vec_type, x, y, skip, hints);
  if (spmvk::use_packed(p, A)) return spmvk::launch_packed_types(p, A, vec_type, x, y, skip);
  return spmvk::launch_types<1>(p, A, vec_type, x, 0, y, 0, H, skip);
}

int spmv_dot_internal(const ispm_launch *p, const ispm_slell_desc *A, int vec_type, const void *x,
                      void *y, const unsigned *skip, const diag_pattern *pat, const unsigned *uni,
                      const void *dotv, double *dparts, unsigned *dgrid,
                      const layout_hints *hints)
{
  *dgrid = 0;
  unsigned H;
  int st = spmvk::check_args(p, A, vec_type, x, y, H);
  if (st) return st;
  if (A->n_slices == 0) return ISPM_OK;
  if (spmvk::use_diagpat(p, A, vec_type, H, pat, uni))
    return spmvk::launch_diagpat_types(A, vec_type, *pat, uni, x, y, skip, dotv, dparts, dgrid);
  if (spmvk::use_stream(p, A))
    return spmvk::launch_stream_types(A, vec_type, x, y, skip, dotv, dparts, dgrid, hints);
  return spmv_internal(p, A, vec_type, x, y, skip, pat, uni, hints);
}

int spmv_split(const split_diag &sd, const ispm_launch *p, const ispm_slell_desc *R, int vec_type,
               const void *x, void *y, const unsigned *skip, const void *dotv, double *dparts,
               unsigned *dgrid, const layout_hints *hints, const split_update *up)
{
  if (dgrid) *dgrid = 0;
  if (!x || !y || !sd.diag) return fail(ISPM_ERR_INVALID, "spmv: null argument");
  const int et = sd.elm_type;
  const unsigned V = et == ISPM_F64 ? 2u : 4u;
  if (sd.n % V) return fail(ISPM_ERR_INVALID, "spmv: split-diagonal rows not a multiple of %u", V);
  const bool rem = sd.m > 0 && R && R->n_slices > 0;
  if (rem && (!spmvk::use_stream(p, R) || R->elm_type != et || !sd.rmap || !sd.rmask))
    return fail(ISPM_ERR_INVALID, "spmv: split-diagonal remainder needs the streamed mapping");
  if (vec_type == ISPM_F64 && et == ISPM_F64)
    return spmvk::launch_split<double, double>(R, sd.diag, sd.rmap, sd.rmask, sd.n, rem, x, y, skip, dotv, dparts, dgrid, hints, up);
  if (vec_type == ISPM_F64 && et == ISPM_F32)
    return spmvk::launch_split<double, float>(R, sd.diag, sd.rmap, sd.rmask, sd.n, rem, x, y, skip, dotv, dparts, dgrid, hints, up);
  if (vec_type == ISPM_F32 && et == ISPM_F32)
    return spmvk::launch_split<float, float>(R, sd.diag, sd.rmap, sd.rmask, sd.n, rem, x, y, skip, dotv, dparts, dgrid, hints, up);
  if (vec_type == ISPM_F32 && et == ISPM_BF16)
    return spmvk::launch_split<float, spmvk::bf16_t>(R, sd.diag, sd.rmap, sd.rmask, sd.n, rem, x, y, skip, dotv, dparts, dgrid, hints, up);
  return fail(ISPM_ERR_UNSUPPORTED, "spmv: unsupported (vector, element) types (%d, %d)", vec_type, et);
}

int spmv_sym(const sym_desc &sd, int vec_type, const void *x, void *y, const unsigned *skip,
             const void *dotv, double *dparts, unsigned *dgrid, unsigned x_guard)
{
  if (dgrid) *dgrid = 0;
  if (!x || !y || !sd.elms) return fail(ISPM_ERR_INVALID, "spmv: null argument");
  if (sd.K < 1 || sd.K > spmvk::kSymPlanes) return fail(ISPM_ERR_INVALID, "spmv: %u symmetric planes", sd.K);
  if (sd.n_slices == 0) return ISPM_OK;
  spmvk::sym_args a{};
  a.elms = sd.elms;
  a.n_slices = sd.n_slices;
  a.n_cols = sd.n_cols;
  a.K = sd.K;
  for (unsigned k = 0; k < spmvk::kSymPlanes; k++) a.d[k] = sd.d[k];
  const int et = sd.elm_type;
  if (vec_type == ISPM_F64 && et == ISPM_F64) return spmvk::launch_symdia<double, double>(a, sd.pat, x, y, skip, dotv, dparts, dgrid, x_guard);
  if (vec_type == ISPM_F64 && et == ISPM_F32) return spmvk::launch_symdia<double, float>(a, sd.pat, x, y, skip, dotv, dparts, dgrid, x_guard);
  if (vec_type == ISPM_F32 && et == ISPM_F32) return spmvk::launch_symdia<float, float>(a, sd.pat, x, y, skip, dotv, dparts, dgrid, x_guard);
  if (vec_type == ISPM_F32 && et == ISPM_BF16) return spmvk::launch_symdia<float, spmvk::bf16_t>(a, sd.pat, x, y, skip, dotv, dparts, dgrid, x_guard);
  return fail(ISPM_ERR_UNSUPPORTED, "spmv: unsupported (vector, element) types (%d, %d)", vec_type, et);
}

int spmm_sym(const sym_desc &sd, int vec_type, int k, const void *X, size_t ldx, void *Y, size_t ldy)
{
  if (!X || !Y || !sd.elms) return fail(ISPM_ERR_INVALID, "spmm: null argument");
  if (k < 1 || k > 8) return fail(ISPM_ERR_UNSUPPORTED, "spmm: k=%d outside 1..8", k);
  if (k > 1 && (ldx < sd.n_cols || ldy < (size_t)sd.n_slices * 256))
    return fail(ISPM_ERR_INVALID, "spmm: leading dimensions %zu, %zu below the plan's columns and rows", ldx, ldy);
  if (k == 1) return spmv_sym(sd, vec_type, X, Y, nullptr);
  if (sd.K < 1 || sd.K > spmvk::kSymPlanes) return fail(ISPM_ERR_INVALID, "spmm: %u symmetric planes", sd.K);
  if (sd.n_slices == 0) return ISPM_OK;
  spmvk::sym_args a{};
  a.elms = sd.elms;
  a.n_slices = sd.n_slices;
  a.n_cols = sd.n_cols;
  a.K = sd.K;
  for (unsigned q = 0; q < spmvk::kSymPlanes; q++) a.d[q] = sd.d[q];
  const int et = sd.elm_type;
  if (vec_type == ISPM_F64 && et == ISPM_F64) return spmvk::launch_symdia_spmm_k<double, double>(a, sd.pat, k, X, ldx, Y, ldy);
  if (vec_type == ISPM_F64 && et == ISPM_F32) return spmvk::launch_symdia_spmm_k<double, float>(a, sd.pat, k, X, ldx, Y, ldy);
  if (vec_type == ISPM_F32 && et == ISPM_F32) return spmvk::launch_symdia_spmm_k<float, float>(a, sd.pat, k, X, ldx, Y, ldy);
  return fail(ISPM_ERR_UNSUPPORTED, "spmm: symmetric plans of (vector, element) types (%d, %d)", vec_type, et);
}

/* SpMM for plans and the block solver: the checks of ispm_spmm, with the plan's layout facts
 * and its shared diagonal pattern (both may be null) */
int spmm_internal(const ispm_launch *p, const ispm_slell_desc *A, int vec_type, int k,
                  const void *X, size_t ldx, void *Y, size_t ldy, const layout_hints *hints,
                  const diag_pattern *pat, const unsigned *uni)
{
  int st = ensure_init();
  if (st) return st;
  unsigned H;
  st = spmvk::check_args(p, A, vec_type, X, Y, H);
  if (st) return st;
  if (k < 1 || k > 8) return fail(ISPM_ERR_UNSUPPORTED, "spmm: k=%d outside 1..8", k);
  if (A->n_slices == 0) return ISPM_OK;
  if (k == 1) return spmv_internal(p, A, vec_type, X, Y, nullptr, pat, uni, hints);
  if (spmvk::use_diagpat_spmm(p, A, vec_type, pat, uni)) {
    switch (k) {
      case 2: return spmvk::launch_diagpat_spmm_types<2>(A, vec_type, *pat, uni, X, ldx, Y, ldy);
      case 3: return spmvk::launch_diagpat_spmm_types<3>(A, vec_type, *pat, uni, X, ldx, Y, ldy);
      case 4: return spmvk::launch_diagpat_spmm_types<4>(A, vec_type, *pat, uni, X, ldx, Y, ldy);
      case 5: return spmvk::launch_diagpat_spmm_types<5>(A, vec_type, *pat, uni, X, ldx, Y, ldy);
      case 6: return spmvk::launch_diagpat_spmm_types<6>(A, vec_type, *pat, uni, X, ldx, Y, ldy);
      case 7: return spmvk::launch_diagpat_spmm_types<7>(A, vec_type, *pat, uni, X, ldx, Y, ldy);
      default: return spmvk::launch_diagpat_spmm_types<8>(A, vec_type, *pat, uni, X, ldx, Y, ldy);
    }
  }
  return spmm_dispatch(p, A, vec_type, k, X, ldx, Y, ldy, hints);
}
}  // namespace ispm

using namespace ispm;
using namespace ispm::spmvk;

extern "C" int ispm_spmv(const ispm_launch *p, const ispm_slell_desc *A, int vec_type,
                         const void *x, void *y)
{
  int st = ensure_init();
  if (st) return st;
  unsigned H;
  st = check_args(p, A, vec_type, x, y, H);
  if (st) return st;
  if (A->n_slices == 0) return ISPM_OK;
  if (use_stream(p, A)) return launch_stream_types(A, vec_type, x, y, nullptr);
  if (use_packed(p, A)) return launch_packed_types(p, A, vec_type, x, y, nullptr);
  return launch_types<1>(p, A, vec_type, x, 0, y, 0, H);
}

extern "C" int ispm_spmm(const ispm_launch *p, const ispm_slell_desc *A, int vec_type, int k,
                         const void *X, size_t ldx, void *Y, size_t ldy)
{
  int st = ensure_init();
  if (st) return st;
  unsigned H;
  st = check_args(p, A, vec_type, X, Y, H);
  if (st) return st;
  if (k < 1 || k > 8) return fail(ISPM_ERR_UNSUPPORTED, "spmm: k=%d outside 1..8", k);
  if (A->n_slices == 0) return ISPM_OK;
  if (k == 1) return spmv_internal(p, A, vec_type, X, Y, nullptr);
  return spmm_dispatch(p, A, vec_type, k, X, ldx, Y, ldy, nullptr);
}

unsigned ispm::sym_x_guard(const sym_desc &sd)
{
  return spmvk::sym_r4_guard(sd.pat, sd.n_slices, sd.n_cols);
}
