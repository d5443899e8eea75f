/* internal.h -- shared declarations of the library's translation units (not installed). */
#ifndef ISPM_INTERNAL_H
#define ISPM_INTERNAL_H

#include <hip/hip_runtime_api.h>
#include <cstddef>
#include <cstdint>
#include <string>
#include "ispm.h"
#include "diag_pattern.h"

namespace ispm {

/* Record a message for ispm_last_error() and return code. */
int fail(int code, const char *fmt, ...);

/* Map a HIP status to ISPM_ERR_RUNTIME (with message) or ISPM_OK. */
int hip_status(hipError_t e, const char *what);

#define ISPM_HIP(call)                                              \
  do {                                                              \
    hipError_t e_ = (call);                                         \
    if (e_ != hipSuccess) return ::ispm::hip_status(e_, #call);     \
  } while (0)

/* Library stream (created by ispm_init / lazily). */
hipStream_t stream();

/* Make sure the device is initialised (gfx950 check); returns ISPM_OK or an error. */
int ensure_init();

/* Whether the current device has been initialised successfully (allocates nothing). */
bool initialized();

/* Devices the library keeps state for. */
constexpr int kMaxDevices = 16;

/* Take the device address of the no-allocation entry point's static scratch on device dev
 * (current); called once per device during initialisation. */
int noalloc_scratch_init(int dev);

/* Per-device reduction scratch: block partial sums of the two-phase reductions of the
 * sblas entry points.  Allocated once per device at initialisation. */
struct reduce_scratch {
  double *partials;         /* kMaxBlocks * kMaxSums doubles */
};
constexpr int kMaxBlocks = 8192;
constexpr int kMaxSums = 8;
reduce_scratch &scratch();

/* Number of compute units (cached). */
int num_cus();

size_t type_size(int t);

/* Facts about a layout that plans compute once on the host and launches may use. */
struct layout_hints {
  /* streamed mapping: the largest element range (from its first element rounded down to a
   * multiple of 8) of any group of 256*RPT consecutive rows, RPT = 1, 2, 4; 0 = unknown */
  unsigned stream_span[3];
  /* the same for groups of 64 and 128 consecutive rows */
  unsigned stream_span_small[2];
  /* hblock slices (fixed height 64..512): per slice (first column, length) of the x window
   * its entries read, staged in LDS (length 0: the slice gathers from global memory);
   * device array of 2 n_slices words, and the longest window */
  const unsigned *xwin;
  unsigned xwin_max;
  unsigned xwin_slices;
};

/* SpMV with the checks of ispm_spmv, an optional device skip flag (nothing is done while
 * *skip != 0), a plan's shared diagonal pattern with its slice mask (both may be null) and
 * its layout facts (may be null). */
int spmv_internal(const ispm_launch *p, const ispm_slell_desc *A, int vec_type, const void *x,
                  void *y, const unsigned *skip, const diag_pattern *pat = nullptr,
                  const unsigned *uni = nullptr, const layout_hints *hints = nullptr);

/* SpMM (k = 1..8 vectors, column j at X + j ldx) with the checks of ispm_spmm, a plan's
 * layout facts and its shared diagonal pattern with its slice mask (all may be null). */
int spmm_internal(const ispm_launch *p, const ispm_slell_desc *A, int vec_type, int k,
                  const void *X, size_t ldx, void *Y, size_t ldy, const layout_hints *hints = nullptr,
                  const diag_pattern *pat = nullptr, const unsigned *uni = nullptr);

/* Split-diagonal form of a square matrix: the rows with off-diagonal entries form a SlELL
 * layout R of m rows (streamed mapping) whose row i is matrix row rmap[i]; y = diag .* x on
 * every other row, y = diag .* x + (R x) on those (one launch, see spmv_kernel.h). */
struct split_diag {
  const void *diag;        /* n entries of elm_type (padded rows 0) */
  const unsigned *rmap;    /* R row -> matrix row, ~0u for R's padding rows */
  const unsigned *rmask;   /* n/32 words: bit r set when matrix row r is a row of R */
  unsigned n;              /* rows (a multiple of 512) */
  unsigned m;              /* rows of R with entries */
  int elm_type;
};

/* CG's residual update fused into a split-diagonal product: the product multiplies
 * x - (*alpha) q (one fused multiply-add per entry), stores that vector to xout and leaves
 * the partials of its squared norm in rrparts (as many as the dot product's) */
struct split_update {
  const void *q;
  const double *alpha;
  void *xout;
  double *rrparts;
};

/* SpMV of a split-diagonal matrix (R given by its launch and descriptor; may be empty), with
 * the fused dot product as spmv_dot_internal (dparts/dgrid may be null); with up, the
 * residual update in front (needs the fused dot). */
int spmv_split(const split_diag &sd, const ispm_launch *p, const ispm_slell_desc *R, int vec_type,
               const void *x, void *y, const unsigned *skip, const void *dotv = nullptr,
               double *dparts = nullptr, unsigned *dgrid = nullptr,
               const layout_hints *hints = nullptr, const split_update *up = nullptr);

/* Symmetric diagonal storage of a symmetric matrix whose entries lie on at most 16 upper
 * diagonals (ISPM_SPMV_PLAN_SYMMETRIC): plane k of 256-row slice s holds A[row, row + d[k]]
 * at elms[(s K + k) kSymPlane + r], d[0] = 0 < d[1] < ..., with rows 256 .. 259 of a plane
 * repeating the next slice's rows 0 .. 3; the lower entries are read from the mirror rows.  pat lists the 2K-1 column offsets -d[K-1] .. -d[1], d[0] .. d[K-1] in
 * that order for the x windows. */
constexpr unsigned kSymPlane = 260;   /* stored length of a 256-row plane */
struct sym_desc {
  const void *elms;
  int elm_type;
  unsigned n_slices, n_cols, K;
  int d[16];
  diag_pattern pat;
};

/* y = A x for symmetric diagonal storage, with the fused dot product as spmv_dot_internal
 * (dotv, dparts, dgrid may be null).  x_guard: elements before x[0] and after x[n_cols - 1]
 * that are readable and zero (vectors the solvers allocate with guard zones); with enough of
 * them (sym_x_guard) FP32 products take the four-rows-per-lane kernel. */
int spmv_sym(const sym_desc &sd, int vec_type, const void *x, void *y, const unsigned *skip,
             const void *dotv = nullptr, double *dparts = nullptr, unsigned *dgrid = nullptr,
             unsigned x_guard = 0);
/* Y = A X for k = 1..8 vectors (column j at X + j ldx, Y + j ldy) in symmetric diagonal
 * storage, every column as spmv_sym computes it (FP64 and FP32 vectors; BF16 elements
 * multiply one vector). */
int spmm_sym(const sym_desc &sd, int vec_type, int k, const void *X, size_t ldx, void *Y, size_t ldy);
/* the guard (elements on each side) the four-rows-per-lane symmetric kernel needs */
unsigned sym_x_guard(const sym_desc &sd);

/* Partial sums a fused SpMV + dot may store (one per workgroup). */
constexpr unsigned kMaxDotParts = 32768;

/* spmv_internal that also asks for the dot product (dotv, y): when the kernel chosen can fuse
 * it, its workgroups store their partial sums at dparts[0 .. *dgrid) and *dgrid is set;
 * otherwise *dgrid is 0 and the caller computes the dot product itself. */
int spmv_dot_internal(const ispm_launch *p, const ispm_slell_desc *A, int vec_type, const void *x,
                      void *y, const unsigned *skip, const diag_pattern *pat, const unsigned *uni,
                      const void *dotv, double *dparts, unsigned *dgrid,
                      const layout_hints *hints = nullptr);

}  // namespace ispm

#endif
