/* spmv.hip -- C-ABI entry point of the single-vector SlELL SpMV (kernel in spmv_kernel.h). */
#include "spmv_kernel.h"

namespace ispm {
/* defined in spmm.hip */
int spmm_dispatch(const ispm_launch *p, const ispm_slell_desc *A, int vec_type, int k,
                  const void *X, size_t ldx, void *Y, size_t ldy);
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
    return spmvk::launch_stream_types(A, vec_type, x, y, skip, nullptr, nullptr, nullptr, nullptr, kMaxDotParts, hints);
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
    return spmvk::launch_stream_types(A, vec_type, x, y, skip, dotv, dparts, dgrid, nullptr, kMaxDotParts, hints);
  return spmv_internal(p, A, vec_type, x, y, skip, pat, uni, hints);
}

int spmv_split(const split_diag &sd, const ispm_launch *p, const ispm_slell_desc *R, int vec_type,
               const void *x, void *y, const unsigned *skip, const void *dotv, double *dparts,
               unsigned *dgrid, const layout_hints *hints)
{
  if (dgrid) *dgrid = 0;
  if (!x || !y || !sd.diag) return fail(ISPM_ERR_INVALID, "spmv: null argument");
  const int et = sd.elm_type;
  const unsigned V = et == ISPM_F64 ? 2u : 4u;
  if (sd.n % V) return fail(ISPM_ERR_INVALID, "spmv: split-diagonal rows not a multiple of %u", V);
  /* one vector of rows per thread and step, at most 2048 workgroups (grid-stride) */
  const unsigned grid = std::max(1u, std::min(2048u, (sd.n / V + 255) / 256));
  const bool fuse = dotv && dparts && dgrid;
  double *dp = fuse ? dparts : nullptr;
  int st;
  if (vec_type == ISPM_F64 && et == ISPM_F64) st = spmvk::launch_split_diag<double, double>(sd.diag, x, y, sd.n, skip, dotv, dp, grid);
  else if (vec_type == ISPM_F64 && et == ISPM_F32) st = spmvk::launch_split_diag<double, float>(sd.diag, x, y, sd.n, skip, dotv, dp, grid);
  else if (vec_type == ISPM_F32 && et == ISPM_F32) st = spmvk::launch_split_diag<float, float>(sd.diag, x, y, sd.n, skip, dotv, dp, grid);
  else if (vec_type == ISPM_F32 && et == ISPM_BF16) st = spmvk::launch_split_diag<float, spmvk::bf16_t>(sd.diag, x, y, sd.n, skip, dotv, dp, grid);
  else return fail(ISPM_ERR_UNSUPPORTED, "spmv: unsupported (vector, element) types (%d, %d)", vec_type, et);
  if (st) return st;
  if (sd.m == 0 || !R || R->n_slices == 0) {
    if (fuse) *dgrid = grid;
    return ISPM_OK;
  }
  if (!spmvk::use_stream(p, R) || R->elm_type != et)
    return fail(ISPM_ERR_INVALID, "spmv: split-diagonal remainder needs the streamed mapping");
  unsigned G2 = 0;
  const bool fuse2 = fuse && grid < kMaxDotParts;
  st = spmvk::launch_stream_types(R, vec_type, x, y, skip, fuse2 ? dotv : nullptr,
                                  fuse2 ? dparts + grid : nullptr, fuse2 ? &G2 : nullptr, sd.rmap,
                                  kMaxDotParts - grid, hints);
  if (st) return st;
  /* both parts' partials, or none (the caller then takes the dot product itself) */
  if (fuse && G2) *dgrid = grid + G2;
  return ISPM_OK;
}

/* SpMM for the block solver: same checks as ispm_spmm */
int spmm_internal(const ispm_launch *p, const ispm_slell_desc *A, int vec_type, int k,
                  const void *X, size_t ldx, void *Y, size_t ldy)
{
  return ispm_spmm(p, A, vec_type, k, X, ldx, Y, ldy);
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
  if (k == 1) return launch_types<1>(p, A, vec_type, X, 0, Y, 0, H);
  return spmm_dispatch(p, A, vec_type, k, X, ldx, Y, ldy);
}
