/* ispm.h -- C-ABI of the MI355X (gfx950) sliced-ELLPACK SpMV / pipelined-PCG library.
 *
 * Every entry point takes plain pointers and sizes; no HIP or C++ type crosses this
 * boundary.  The C++ drop-in headers (formats/formats.h, plan/plan.h, fastainv/fastainv.h,
 * util/cuda/sblas.h, util/cuda/initialize.h under ispm-sparse-lib_amd/) are thin inline
 * wrappers over these functions, so a caller built against the reference's headers
 * recompiles against ours, provided it calls no CUDA runtime function directly and does not
 * register existing host memory (cuda_pinned_region, unsupported on this target).  Each
 * declaration names the reference interface it replaces (paths relative to the reference
 * tree).
 *
 * Conventions
 *  - Return value: ISPM_OK (0) or an ISPM_ERR_* code; ispm_last_error() gives a message.
 *  - Device pointers come from ispm_malloc (or any hipMalloc'd memory on the current device).
 *  - All device work is enqueued on ONE library stream (ispm_get_stream / ispm_set_stream);
 *    results are visible after ispm_synchronize() or an event on that stream.
 *  - Scalar type codes: ISPM_F32, ISPM_F64, ISPM_BF16 (bf16 only as a matrix element type).
 */
#ifndef ISPM_H
#define ISPM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum ispm_status {
  ISPM_OK = 0,
  ISPM_ERR_INVALID = 1,     /* bad argument */
  ISPM_ERR_UNSUPPORTED = 2, /* valid but unsupported combination (reference: silent no-op) */
  ISPM_ERR_RUNTIME = 3,     /* HIP runtime failure */
  ISPM_ERR_NOMEM = 4,
  ISPM_ERR_DEVICE = 5,      /* wrong / missing device (reference: arch-mismatch abort) */
  ISPM_ERR_NOT_CONVERGED = 6
};

enum ispm_type { ISPM_F32 = 0, ISPM_F64 = 1, ISPM_BF16 = 2 };

enum ispm_copy_kind { ISPM_H2D = 0, ISPM_D2H = 1, ISPM_D2D = 2, ISPM_DEFAULT = 3 };

const char *ispm_last_error(void);

/* ------------------------------------------------------------------ runtime ------------ */

/* Replaces ispm_initialize() (util/cuda/initialize.h:4, initialize.cu:17-30): checks that the
 * current device is gfx950 (the only ISA this library is built for) and creates the library
 * stream.  Returns ISPM_ERR_DEVICE instead of aborting. */
int ispm_init(void);

/* Replaces get_device_string() (util/cuda/initialize.h:6-10, initialize.cu:32-42):
 * "dddd:bb:dd (name)" into buf (at most len bytes, NUL-terminated). */
int ispm_device_string(char *buf, int len);

/* Replaces cuda_device_allocator::allocate/deallocate (util/cuda/allocator.h:14-25):
 * device allocation, zero-filled like the reference. */
int ispm_malloc(void **p, size_t bytes);
int ispm_free(void *p);

/* Replaces copy()/copy_async() (util/cuda/cudamemcpy.h:27-43).  kind: ispm_copy_kind. */
int ispm_memcpy(void *dst, const void *src, size_t bytes, int kind);
int ispm_memcpy_async(void *dst, const void *src, size_t bytes, int kind);
/* Fill device memory with zero bytes, on the library stream (a kernel, not a memset node). */
int ispm_zero(void *p, size_t bytes);

/* Replaces cuda_pinned_allocator (util/cuda/allocator.h:44-56) and host2devmem
 * (util/cuda/pinner.h:7-22): mapped pinned host memory and its device alias. */
int ispm_host_alloc_mapped(void **hp, size_t bytes);
int ispm_host_free(void *hp);
int ispm_host_device_pointer(void **dp, void *hp);

int ispm_synchronize(void);
void *ispm_get_stream(void);            /* hipStream_t of the library stream */
int ispm_set_stream(void *hip_stream);  /* NULL restores the library's own stream */

/* Replaces cudaEventScoped / cudatimer (util/cuda/timer.h:6-52). Events record on the
 * library stream. */
int ispm_event_create(void **ev);
int ispm_event_record(void *ev);
int ispm_event_synchronize(void *ev);
int ispm_event_elapsed_ms(void *start, void *stop, float *ms);
int ispm_event_destroy(void *ev);

/* ------------------------------------------------------------ SlELL SpMV --------------- */

/* Device view of slell_matrix<E, device_memory_space_tag> (formats/formats.h:44-102). */
typedef struct ispm_slell_desc {
  int n_rows, n_cols, n_nz;
  unsigned n_slices, slice_height, hblock;
  int var_height, diags;
  const unsigned *slice_ptr; /* (n_slices+1) records of P = 1+[hblock>1]+[V]+2[D] words,
                                then the diagonal positions when diags */
  size_t slice_ptr_len;
  const unsigned *cols;
  size_t cols_len;
  const void *elms;          /* elm_type */
  size_t elms_len;
  int elm_type;
} ispm_slell_desc;

/* spmv_launch_params (cuda/spmv/kernel-iface.h:6-18).  N = threads per workgroup
 * (64..1024, power of two), S = slice height the launch is specialised for, H = rows per
 * thread (1 or 2).  Valid iff N >= S/H and S >= H.  use_tex_* are accepted and ignored:
 * x windows are read through the cache hierarchy, never through texture units.
 * mode (new): 0 = one workgroup per slice, threads split a slice's planes when N*H > S (the
 * reference's mapping); 1 = packed, for fixed slice heights below 64: each 64-lane wave
 * takes 64/S consecutive slices with one row per lane (N then only sets the workgroup
 * size); 2 = streamed, for fixed slice heights up to 256 without hblock pairs or diagonal
 * planes: a 256-thread workgroup reads the contiguous elements of 256 rows with consecutive
 * lanes on consecutive elements, stages the products in LDS and sums each row in plane
 * order (other layouts given mode 2 run packed when S < 64, else mode 0).  Any other mode
 * is rejected. */
typedef struct ispm_launch {
  unsigned N, S, H;
  int use_tex_sliceptr, use_tex_x;
  int mode;
} ispm_launch;

/* Replaces launch_spmv_kernel<T,E> (cuda/spmv/kernel-iface.h:20-74, dispatch.h:184-205):
 * y[0:n_rows) = A x.  (vec_type, A.elm_type) in {(F32,F32), (F64,F32), (F64,F64),
 * (F32,BF16)}.  x holds n_cols entries (entries outside [0,n_cols) read as 0), y holds the
 * padded n_rows.  Invalid launch parameters return ISPM_ERR_UNSUPPORTED and leave y
 * unwritten (the reference returns silently, dispatch.h:202-203). */
int ispm_spmv(const ispm_launch *p, const ispm_slell_desc *A, int vec_type,
              const void *x, void *y);

/* Multi-RHS: Y[:,j] = A X[:,j] for j < k (1..8), column j of X at X + j*ldx elements.
 * One read of A serves all k vectors.  No reference counterpart (new in this library). */
int ispm_spmm(const ispm_launch *p, const ispm_slell_desc *A, int vec_type, int k,
              const void *X, size_t ldx, void *Y, size_t ldy);

/* --------------------------------------------------------------- sblas ---------------- */
/* Replace sblas::{copy_indexed, add_indexed, dot, ppcg_update_scalars, ppcg_update_vectors}
 * (util/cuda/sblas.h:1-25, sblas.cu:18-237), type ISPM_F32 or ISPM_F64.  All pointers are
 * device pointers (scalars may be mapped host memory).  Dot products accumulate in FP64
 * for both types. */
int ispm_copy_indexed(int type, void *dst, const void *src, const int *idx, int n);
int ispm_add_indexed(int type, void *dst, const void *src, const int *idx, int n);
int ispm_dot(int type, const void *x, const void *y, void *s, int n);
int ispm_ppcg_update_scalars(int type, void *alpha, void *beta, const void *gamma,
                             const void *gammaold, const void *delta);
int ispm_ppcg_update_vectors(int type, void *resnorm, void *gamma, void *delta,
                             const void *alpha, const void *beta,
                             const void *n, const void *m,
                             void *p, void *s, void *q, void *z, void *x, void *r,
                             void *u, void *w, int n_elts);
/* No counterpart in the reference: ispm_ppcg_update_vectors for callers that cannot allocate
 * device memory (code under hipGraph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, kernels and results (bit-identical); the partial sums of its
 * reductions live in a static device array of 8192 x 4 doubles (256 KiB per GPU, held until
 * the process ends) instead of the library's allocated scratch.  Every n_elts in
 * [0, 2^31-1] fits (at most 8192 workgroups with grid-stride loops); n_elts < 0 is rejected.
 * Requires ispm_init() on the current device beforehand (ISPM_ERR_INVALID otherwise), and
 * calls no allocating routine.  All calls on one GPU share that array: calls that overlap
 * (on different streams, or from several threads) race. */
int ispm_ppcg_update_vectors_noalloc(int type, void *resnorm, void *gamma, void *delta,
                                     const void *alpha, const void *beta,
                                     const void *n, const void *m,
                                     void *p, void *s, void *q, void *z, void *x, void *r,
                                     void *u, void *w, int n_elts);

/* ------------------------------------------------------ host formats (CPU) ------------ */

/* csr_matrix<T> (formats/formats.h:18-37) as plain arrays. */
typedef struct ispm_csr {
  int n_rows, n_cols, n_nz;
  const int *row_ptr;   /* n_rows+1 */
  const int *cols;      /* n_nz */
  const void *elms;     /* n_nz of elm_type */
  int elm_type;
} ispm_csr;

/* CSR arrays owned by the library (outputs of ispm_fastainv_sym and ispm_mm_read); release
 * with ispm_csr_out_free. */
typedef struct ispm_csr_out {
  int n_rows, n_cols, n_nz;
  int *row_ptr;
  int *cols;
  void *elms;
  int elm_type;
} ispm_csr_out;

/* slell_matrix<E, host_memory_space_tag> (formats/formats.h:44-102) owned by the library;
 * release with ispm_slell_host_free. */
typedef struct ispm_slell_host {
  int n_rows, n_cols, n_nz;
  unsigned n_slices, slice_height, hblock;
  int var_height, diags;
  unsigned *slice_ptr;
  size_t slice_ptr_len;
  unsigned *cols;
  size_t cols_len;
  void *elms;
  size_t elms_len;
  int elm_type;
} ispm_slell_host;

/* Replaces csr_to_slell (formats/conversion.h:338-414).  The reference pads the caller's
 * CSR to a multiple of 512 rows in place (conversion.h:343); here A is read-only and the
 * result's n_rows is the padded count.  out->elm_type = A->elm_type, or ISPM_BF16/ISPM_F32
 * narrowing when out_elm_type differs (values rounded to nearest even). */
int ispm_csr_to_slell(const ispm_csr *A, int slice_height, int hblock, int var_height,
                      int diags, int min_height, int out_elm_type, ispm_slell_host *out);

/* Replaces csr_to_slell_guess_format (formats/conversion.h:416-492). */
int ispm_csr_to_slell_guess(const ispm_csr *A, int examine_ratio, int min_height,
                            int out_elm_type, ispm_slell_host *out);
void ispm_slell_host_free(ispm_slell_host *m);

/* Copy a host SlELL to the device (slell_matrix<E,device>(host), formats.h:56-66).
 * Release with ispm_slell_device_free. */
int ispm_slell_upload(const ispm_slell_host *h, ispm_slell_desc *d);
void ispm_slell_device_free(ispm_slell_desc *d);

/* ------------------------------------------------------------- SpMV plan --------------- */

/* ISPM_PLAN_DEFAULT and ISPM_PLAN_ESTIMATE (the same kind, as in plan/plan.h:74-75): the
 * guessed format (conversion.h:416-492); ISPM_PLAN_EXHAUSTIVE: plan/plan.h:97-144;
 * ISPM_PLAN_MIN_BYTES (no reference counterpart): the format with the fewest stored bytes,
 * as ispm_plan_estimate chooses it. */
enum ispm_plan_kind { ISPM_PLAN_DEFAULT = 0, ISPM_PLAN_ESTIMATE = 1, ISPM_PLAN_EXHAUSTIVE = 2,
                      ISPM_PLAN_MIN_BYTES = 3 };
/* ISPM_SPMV_PLAN_SPLIT_DIAG (no reference counterpart): store a square matrix as its main
 * diagonal (a dense vector) plus the rows that have off-diagonal entries (a compressed
 * SlELL layout of those rows and a row map, streamed mapping); y = diag .* x, then those
 * rows add their sums.  For approximate-inverse factors and their products, whose rows are
 * mostly the diagonal alone.  Execution and the solver only (no SpMM, no coefficient
 * refresh); plan info reports the remainder's layout with the full padded row count. */
/* ISPM_SPMV_PLAN_SYMMETRIC (no reference counterpart): for an exactly symmetric square
 * matrix whose entries lie on at most 16 upper diagonals (stencils): store only the upper
 * diagonals as dense planes of 256-row slices; each lower entry is read from its mirror, which
 * an earlier slice streamed moments before, so HBM carries about half the matrix.  Plan
 * creation fails with ISPM_ERR_INVALID when the matrix is not symmetric or not on few
 * diagonals.  Execution, SpMM (ispm_plan_execute_multi and the block solver; FP64 and FP32
 * vectors, each column bit-identical to the single-vector product), the solvers and
 * coefficient refresh. */
enum ispm_plan_flags { ISPM_SPMV_NO_TEXTURE_MEMORY = 1, ISPM_SPMV_PLAN_NO_DEVICE = 2,
                       ISPM_SPMV_PLAN_SPLIT_DIAG = 4, ISPM_SPMV_PLAN_SYMMETRIC = 8 };

typedef struct ispm_plan ispm_plan;

typedef struct ispm_plan_info {
  unsigned S, hblock, V, D;    /* slell_params mparams (formats/formats.h:39-42) */
  ispm_launch launch;          /* spmv_launch_params params */
  double tuned_seconds;        /* best timed launch (exhaustive only) */
  size_t stored_bytes;         /* 4|slice_ptr| + 4|cols| + sizeof(E)|elms| + sizeof(T)(n_cols+n_rows) */
  size_t csr_bytes;            /* 4(n+1) + (4+sizeof(E))nnz + sizeof(T)*2n */
  ispm_slell_desc device;      /* device matrix (valid unless NO_DEVICE) */
  unsigned pattern_planes;     /* diagonal planes of the pattern most slices share (0: none) */
  unsigned pattern_slices;     /* slices that have exactly that pattern */
  unsigned window_slices;      /* hblock slices whose x window is staged in LDS */
} ispm_plan_info;

/* Replaces spmv_plan<T,E>::spmv_plan(csr, kind, flags) (plan/plan.h:64-146):
 * vec_type = T, A->elm_type (or elm_type override, e.g. ISPM_BF16) = E. */
int ispm_plan_create(ispm_plan **out, int vec_type, int elm_type, const ispm_csr *A,
                     int kind, unsigned flags);
/* Plan from explicit parameters (no guessing / tuning). */
int ispm_plan_create_explicit(ispm_plan **out, int vec_type, int elm_type, const ispm_csr *A,
                              unsigned S, unsigned hblock, int V, int D,
                              const ispm_launch *launch);
int ispm_plan_info_get(const ispm_plan *p, ispm_plan_info *info);
/* Launch shape the default plan uses for a format (CDNA4 heuristic: one row per thread,
 * workgroups of max(64, min(1024, S)) threads; one 64-lane wave per slice for
 * variable-height slices). */
int ispm_plan_default_launch(unsigned slice_height, int var_height, ispm_launch *out);
/* No reference counterpart (ISPM_PLAN_MIN_BYTES; the reference's PLAN_ESTIMATE is an alias of
 * PLAN_DEFAULT, plan/plan.h:74-96): the format with the fewest stored bytes (exact for plain ELL at every slice height, fixed and
 * variable; the diagonal/hblock variant suggested by the sampled statistics converted once
 * and compared), with the default launch shape. */
int ispm_plan_estimate(int vec_type, int elm_type, const ispm_csr *A, unsigned *S,
                       unsigned *hblock, int *V, int *D, ispm_launch *launch);
/* No reference counterpart (the learned planner's training labels): SpMV seconds (median
 * of 5 launches, default launch shape of each format) of `count` formats (S[c], hblock[c],
 * V[c], D[c]); +inf where a format cannot be built for A. */
int ispm_plan_time_formats(int vec_type, int elm_type, const ispm_csr *A, int count,
                           const unsigned *S, const unsigned *hblock, const int *V, const int *D,
                           double *seconds);
/* No reference counterpart (labelling processes that share one GPU): ispm_plan_time_formats
 * with the layouts converted on `threads` host threads (0: up to 8) and the device held
 * through an advisory lock, flock(lock_fd, LOCK_EX), while each batch of converted layouts is
 * uploaded and timed (lock_fd < 0: no lock).  Host conversion runs outside the lock. */
int ispm_plan_time_formats_shared(int vec_type, int elm_type, const ispm_csr *A, int count,
                                  const unsigned *S, const unsigned *hblock, const int *V,
                                  const int *D, double *seconds, int threads, int lock_fd);
/* The PLAN_EXHAUSTIVE search (plan/plan.h:97-144): times every (S, V, D, hblock) format and
 * launch shape on the device and returns the fastest. */
int ispm_plan_tune(int vec_type, int elm_type, const ispm_csr *A, unsigned *S, unsigned *hblock,
                   int *V, int *D, ispm_launch *launch, double *seconds);
/* spmv_plan::execute_spmv (plan/plan.h:148-161). */
int ispm_plan_execute(ispm_plan *p, const void *x, void *y);
/* No reference counterpart: ispm_plan_execute for a caller whose x is readable and zero on
 * x_guard elements before x[0] and after x[n_cols - 1] (a vector allocated with guard zones).
 * Symmetric plans with FP32 vectors then take a kernel that reads its x windows without
 * bounds checks (at least ispm_plan_x_guard(p) elements are needed, else the checked kernel
 * runs); other plans ignore the guard.  Results equal ispm_plan_execute's bit for bit. */
int ispm_plan_execute_guarded(ispm_plan *p, const void *x, void *y, unsigned x_guard);
unsigned ispm_plan_x_guard(const ispm_plan *p);
int ispm_plan_execute_multi(ispm_plan *p, int k, const void *X, size_t ldx, void *Y, size_t ldy);
/* change_host_matrix + update_device_matrix_coefs (plan/plan.h:39-56): new values, same
 * sparsity pattern; conversion reuses the plan's format parameters.  Symmetric plans take
 * the upper entries and check every lower one against its mirror (ISPM_ERR_INVALID when
 * they differ); split-diagonal plans are rebuilt instead (ISPM_ERR_UNSUPPORTED). */
int ispm_plan_update_coefs(ispm_plan *p, const ispm_csr *A);
/* On-device refresh: vals (device pointer, nnz values of val_type F32 or F64, in the CSR
 * order the plan was built from) are scattered into the element slots through the
 * conversion's entry -> slot map; no host re-conversion and no full re-upload.  Symmetric
 * plans read the upper entries only (the caller keeps the matrix symmetric). */
int ispm_plan_update_coefs_device(ispm_plan *p, const void *vals, int val_type);
void ispm_plan_destroy(ispm_plan *p);

/* ------------------------------------------------------------- PCG solver -------------- */
/* Device-resident FSAI-preconditioned CG.  No counterpart in the reference library: this is
 * the PCG loop of its caller (foam-extend ispmCudaSolvers, README.md:3-9 of the reference),
 * built from the reference's primitives (execute_spmv, sblas::ppcg_update_scalars,
 * sblas::ppcg_update_vectors) plus fused kernels.
 *   method ISPM_SOLVER_PPCG: Ghysels-Vanroose pipelined PCG with the reference's recurrences
 *     (util/cuda/sblas.cu:156-217); M = Ainv2*Ainv1 applied as two SpMVs.
 *   method ISPM_SOLVER_PCG : Hestenes-Stiefel PCG.
 *   method ISPM_SOLVER_PCG1: single-reduction PCG (Chronopoulos-Gear): z = M r, w = A z,
 *     both dot products (r,z), (z,w) and the convergence sum in one finishing step, then one
 *     fused update of p, s = A p, x and r; the reference's scalar formulas.
 *   precision ISPM_PREC_FP64 : every vector and scalar in FP64.
 *   precision ISPM_PREC_MIXED: FP32 vectors, FP32/BF16 matrices, FP64 dot products and
 *     scalars; x accumulated in FP64 and the residual replaced by the FP64 true residual
 *     b - A x whenever the FP32 recursive residual has dropped by reliable_delta (PCG and
 *     PCG1).
 * Convergence is always decided on the FP64 true residual ||b - A x||_2 / ||b||_2 <= tol.
 * Iterations run in hipGraph-captured batches of check_every iterations; a device flag
 * turns the rest of a batch into no-ops once the recursive residual meets the target. */
enum ispm_solver_method { ISPM_SOLVER_PPCG = 0, ISPM_SOLVER_PCG = 1, ISPM_SOLVER_PCG1 = 2 };
enum ispm_solver_precision { ISPM_PREC_FP64 = 0, ISPM_PREC_MIXED = 1 };

typedef struct ispm_solver_opts {
  int method, precision;
  double tol;
  int maxit;
  int check_every;        /* iterations per graph replay / host check (default 16) */
  int use_graph;          /* 1: capture batches in a hipGraph */
  double reliable_delta;  /* mixed: FP32 residual reduction between FP64 replacements (0.01) */
} ispm_solver_opts;

typedef struct ispm_solve_stats {
  int iterations, replacements, converged;
  double relres;          /* FP64 true relative residual at exit */
  double l1_resnorm;      /* sum |r_i| of the recursive residual (sblas.cu:202 semantics) */
} ispm_solve_stats;

typedef struct ispm_solver ispm_solver;

/* A: plan of the system matrix (E = F64 for FP64, E = F32 for mixed; the mixed solver uses
 * it both with FP32 and FP64 vectors, so its values must be those of A).  Ainv1 / Ainv2:
 * plans of the factors (T = F64 for FP64, T = F32 for mixed) applied as z = Ainv2 (Ainv1 r);
 * Ainv1 alone is a one-matrix preconditioner z = Ainv1 r (for instance the product
 * Ainv2 * Ainv1 formed by ispm_csr_multiply); both NULL: no preconditioner.  Plans must
 * outlive the solver. */
int ispm_solver_create(ispm_solver **out, ispm_plan *A, ispm_plan *Ainv1, ispm_plan *Ainv2,
                       const ispm_solver_opts *opts);
/* Mixed precision evaluates the FP64 true residual with A's plan read into FP64
 * accumulators, which is exact only when A's values are FP32-representable.  Otherwise give
 * an FP64-element plan of the same matrix here (used only for the residual replacements). */
int ispm_solver_set_residual_plan(ispm_solver *s, ispm_plan *A64);
/* b, x: device FP64 vectors of the padded length; x is the initial guess and the result. */
int ispm_solver_solve(ispm_solver *s, const double *b, double *x, ispm_solve_stats *stats);

/* The solve with the convergence test of the reference's caller, a foam-extend (OpenFOAM)
 * linear solver (no counterpart in the reference library, README.md:3-9): OpenFOAM's
 * normalised residual sum |b - A x| / normFactor, normFactor = sum |A x0 - sumA xbar| +
 * |b - sumA xbar| + 1e-20 with sumA the row sums of A and xbar the mean of the initial x
 * (lduMatrix::solver::normFactor), over the matrix's rows.  Converged when the residual is
 * below tolerance, or below rel_tol times the initial residual when rel_tol > 0; an initial
 * residual already below tolerance returns at once with 0 iterations (x unchanged).  The
 * iteration limit is opts.maxit.  Returns ISPM_ERR_NOT_CONVERGED like ispm_solver_solve. */
typedef struct ispm_norm_stats {
  int iterations, replacements, converged;
  double initial_residual;   /* normalised residual of the initial x */
  double final_residual;     /* normalised residual at exit */
  double norm_factor;
  double l1_resnorm;         /* sum |r_i| of the recursive residual */
} ispm_norm_stats;
int ispm_solver_solve_normalised(ispm_solver *s, const double *b, double *x, double tolerance,
                                 double rel_tol, ispm_norm_stats *stats);
void ispm_solver_destroy(ispm_solver *s);

/* ------------------------------------------------------------- lduMatrix input -------- */
/* No counterpart in the reference library: its caller converts OpenFOAM's lduMatrix to the
 * csr_matrix it passes in.  n cells, n_faces internal faces; face f joins cells
 * lower_addr[f] < upper_addr[f] with coefficients upper[f] = A[lower][upper] and lower[f] =
 * A[upper][lower] (lower = NULL: symmetric, lower = upper); diag[n].  Output: FP64 CSR with
 * columns ascending in every row (library-owned, ispm_csr_out_free).  pos (optional, n +
 * 2 n_faces ints): the CSR index of diag[i] at pos[i], of upper[f] at pos[n + f], of
 * lower[f] at pos[n + n_faces + f].  Fails with ISPM_ERR_INVALID for addresses out of range
 * or not lower < upper, and for two faces joining the same cells. */
int ispm_ldu_to_csr(int n, int n_faces, const int *lower_addr, const int *upper_addr,
                    const double *diag, const double *upper, const double *lower,
                    ispm_csr_out *out, int *pos);
/* New coefficients of the same addressing in CSR value order (vals[pos[s]] = source s),
 * on the host, or on the device (pos, diag, upper, lower, vals device pointers; positions
 * outside the CSR are skipped), e.g. before ispm_plan_update_coefs_device. */
int ispm_ldu_values(int n, int n_faces, const int *pos, const double *diag, const double *upper,
                    const double *lower, double *vals);
int ispm_ldu_values_device(int n, int n_faces, const int *pos, const double *diag,
                           const double *upper, const double *lower, double *vals);

/* DroptolTuner (fastainv/droptol-tuning.h:4-103) for callers without the C++ header: a
 * golden-section search over log2(droptol) in [log2_lo, log2_hi] driven by the times the
 * caller reports for each probed drop tolerance. */
typedef struct ispm_droptol_tuner ispm_droptol_tuner;
int ispm_droptol_tuner_create(double log2_lo, double log2_hi, ispm_droptol_tuner **out);
double ispm_droptol_tuner_droptol(const ispm_droptol_tuner *t);   /* getDroptol */
double ispm_droptol_tuner_best(const ispm_droptol_tuner *t);      /* getBestDroptol */
void ispm_droptol_tuner_note_time(ispm_droptol_tuner *t, double seconds);   /* noteTime */
void ispm_droptol_tuner_restart(ispm_droptol_tuner *t);           /* restart */
int ispm_droptol_tuner_iterations(const ispm_droptol_tuner *t);
void ispm_droptol_tuner_destroy(ispm_droptol_tuner *t);

/* ------------------------------------------------------------- on-disk formats -------- */
/* No reference counterpart (the reference has no persistence).  MatrixMarket coordinate
 * files (real / integer / pattern; general or symmetric, expanded) to CSR (duplicates add),
 * and back; output is library-owned (ispm_csr_out_free). */
int ispm_mm_read(const char *path, ispm_csr_out *out);
int ispm_mm_write(const char *path, const ispm_csr *A);
/* 64-bit FNV-1a fingerprint of (n_rows, n_cols, row_ptr, cols). */
int ispm_csr_fingerprint(const ispm_csr *A, unsigned long long *fp);
/* Plan cache: text file of tuned plans keyed by (fingerprint, vector type, element type).
 * ispm_plan_create(PLAN_EXHAUSTIVE) consults and extends the file named by the environment
 * variable ISPM_PLAN_CACHE.  lookup returns ISPM_ERR_INVALID when the key is absent. */
int ispm_plan_cache_lookup(const char *path, unsigned long long fp, int vec_type, int elm_type,
                           unsigned *S, unsigned *hblock, int *V, int *D, ispm_launch *launch,
                           double *seconds);
int ispm_plan_cache_store(const char *path, unsigned long long fp, int vec_type, int elm_type,
                          unsigned S, unsigned hblock, int V, int D, const ispm_launch *launch,
                          double seconds);

/* ------------------------------------------------- multi-RHS and batched solves ------- */
/* k = 1..8 right-hand sides sharing one matrix: independent FP64 PCG recurrences per column
 * whose matrix products are SpMMs (one pass over A, Ainv1, Ainv2 per iteration for all
 * columns).  B, X: device, column j at B + j*ldb / X + j*ldx (padded length); X holds the
 * initial guesses.  stats: array of k.  opts->use_graph: batches of check_every iterations
 * are captured once (per X) and replayed as a hipGraph.  No reference counterpart. */
typedef struct ispm_block_solver ispm_block_solver;
int ispm_block_solver_create(ispm_block_solver **out, ispm_plan *A, ispm_plan *Ainv1,
                             ispm_plan *Ainv2, int k, const ispm_solver_opts *opts);
int ispm_block_solver_solve(ispm_block_solver *s, const double *B, size_t ldb, double *X,
                            size_t ldx, ispm_solve_stats *stats);
void ispm_block_solver_destroy(ispm_block_solver *s);

/* Many independent small SPD systems (FP32) solved together with Jacobi-preconditioned
 * single-reduction CG: every launch covers the segments (1024-row pieces) of all systems still
 * iterating, iterations are replayed from a captured hipGraph, and each system stops on its
 * own residual.  The systems are given as one block-diagonal CSR (global row/column
 * numbering, system i owns sys_rows[i] consecutive rows, may not couple to another system and
 * has rows of at most 255 entries); the library renumbers each system internally (reverse
 * Cuthill-McKee).  b and x use the padded per-system layout reported by ispm_batch_layout
 * (each system padded to a multiple of 64 rows; padding entries of b must be 0; x is written
 * on return and need not be initialised).  iters / relres (host arrays of n_sys, may be NULL)
 * receive each system's iteration count and FP32 recursive relative residual.  No reference
 * counterpart. */
typedef struct ispm_batch ispm_batch;
int ispm_batch_create(ispm_batch **out, int n_sys, const int *sys_rows, const int *row_ptr,
                      const int *cols, const float *vals);
int ispm_batch_layout(const ispm_batch *b, long long *rows_padded, long long *row_offsets);
/* bytes of the stored matrix (values, columns, slice offsets and widths): what one product
 * over every system reads besides its vectors */
int ispm_batch_matrix_bytes(const ispm_batch *b, long long *bytes);
int ispm_batch_solve(ispm_batch *b, const float *rhs, float *x, double tol, int maxit,
                     int *iters, double *relres);
void ispm_batch_destroy(ispm_batch *b);

/* ------------------------------------------------- approximate inverse (CPU) ---------- */

/* Replaces fastainv_sym<in_real,out_real> (fastainv/fastainv.h:8-30, fastainv.cpp:395-566):
 * AINV-type factorisation with M = Ainv2 * Ainv1 ~= A^-1.  in type ISPM_F64, out type
 * ISPM_F32 or ISPM_F64.  progress may be NULL.  Outputs are library-owned CSR arrays;
 * release with ispm_csr_out_free.  Returns ISPM_ERR_INVALID for n <= 0 (reference: false). */

typedef void (*ispm_progress_fn)(void *);

int ispm_fastainv_sym(int out_type, const ispm_csr *A, ispm_progress_fn progress, void *v,
                      long long max_elts, double droptol,
                      ispm_csr_out *Ainv1, ispm_csr_out *Ainv2);

/* No reference counterpart (its solver applies the two factors in turn).  C = A * B on the
 * host with FP64 accumulation, output type ISPM_F64 or ISPM_F32, sorted columns; forms the
 * one-matrix preconditioner Ainv2 * Ainv1 when that has fewer nonzeros than the two factors
 * together.  Output is library-owned (ispm_csr_out_free). */
int ispm_csr_multiply(const ispm_csr *A, const ispm_csr *B, int out_type, ispm_csr_out *C);

/* No reference counterpart: the learned planner's network (32 features -> 64 -> 64 ->
 * C plan scores + 1 log2 drop tolerance, ReLU, FP32), evaluated and trained by HIP kernels.
 * All pointers are device pointers; params holds ispm_mlp_param_count(C) floats laid out
 * W1[64x32] b1[64] W2[64x64] b2[64] W3[(C+1)x64] b3[C+1].  X is B x 32 (row-major); Tp B x C
 * target distributions over the formats, Mp / Md B masks (> 0: the sample has a label in
 * that head), Td B target log2 drop tolerances.  Loss = mean over samples of the masked soft
 * cross-entropy of the plan scores plus drop_weight * masked squared error of the drop
 * tolerance head.  One set of activation buffers per process: calls must not overlap. */
int ispm_mlp_param_count(int n_plan_classes);
int ispm_mlp_forward(int n_plan_classes, const float *params, const float *X, int B, float *out);
int ispm_mlp_gradient(int n_plan_classes, const float *params, const float *X, const float *Tp,
                      const float *Mp, const float *Td, const float *Md, int B, float drop_weight,
                      float *grad, float *loss);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8, bias-corrected for step >= 1) on params with
 * moment buffers m, v (zero before step 1). */
int ispm_mlp_train_step(int n_plan_classes, float *params, float *m, float *v, int step,
                        const float *X, const float *Tp, const float *Mp, const float *Td,
                        const float *Md, int B, float lr, float drop_weight, float *loss);
void ispm_csr_out_free(ispm_csr_out *m);

#ifdef __cplusplus
}
#endif

#endif /* ISPM_H */
