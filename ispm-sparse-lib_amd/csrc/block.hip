/* block.hip -- k right-hand sides solved as one block of independent FSAI-PCG recurrences
 * that share every matrix read: A*P, Ainv1*R and Ainv2*T are SpMMs over the k columns
 * (one pass over the matrix for all columns), and the per-column dot products, scalar
 * updates and vector updates run as 2-D grids (blockIdx.y = column).  Each column keeps its
 * own alpha, beta, rho and convergence flag; a converged column stops updating while the
 * others continue.  FP64; convergence of every column is decided on its FP64 true residual.
 * No counterpart in the reference (its caller solves one right-hand side at a time). */
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>
#include "device.h"
#include "internal.h"
#include "plan_internal.h"
#include "sblas_kernels.h"

namespace ispm {

namespace {

struct colsc {
  double rho, pq, alpha, beta, rr, tol2, r64;
  unsigned active, iters;
  /* X_j += alpha P_j of the last iteration not yet applied: the next k_mp applies it (it reads
   * P_j anyway), or k_mflush once the column stops */
  unsigned pend;
};

constexpr int kMaxK = 8;

enum { OP_RZ, OP_PQ, OP_R64, OP_XR };

/* per-column dot (X_j, Y_j) with the scalar epilogue of PCG */
template<int OP>
__global__ void __launch_bounds__(256)
k_mdot(const double *__restrict__ X, const double *__restrict__ Y, size_t ld, int n, const colsc *cs,
       double *parts)
{
  __shared__ double lds[4];
  const int j = blockIdx.y;
  const colsc *c = cs + j;
  if (OP != OP_R64 && !c->active) return;
  const double *x = X + j * ld, *y = Y + j * ld;
  double v[1] = {0.0};
  const int gt = blockIdx.x * 256 + threadIdx.x, gs = gridDim.x * 256, nv = n / 2;
  for (int i = gt; i < nv; i += gs) {
    pack16<double> a = ld16(x, i), b = ld16(y, i);
    v[0] += a.v[0] * b.v[0] + a.v[1] * b.v[1];
  }
  for (int i = nv * 2 + gt; i < n; i += gs) v[0] += x[i] * y[i];
  store_partials<1>(v, parts + (size_t)j * kMaxBlocks, lds);
}

/* Second phase of the per-column reductions (block j sums column j's partials) with the
 * scalar epilogue of PCG; columns already converged are left as they are. */
template<int OP>
__global__ void __launch_bounds__(kFinThreads) k_mfin(const double *parts, int G, colsc *cs)
{
  __shared__ double lds[kFinThreads / 64];
  const int j = blockIdx.x;
  colsc *c = cs + j;
  if (OP != OP_R64 && !c->active) return;
  double out[1];
  sum_partials<1>(parts + (size_t)j * kMaxBlocks, G, lds, out);
  if (threadIdx.x) return;
  const double r = out[0];
  if (OP == OP_RZ) { c->beta = r / c->rho; c->rho = r; }
  if (OP == OP_PQ) { c->pq = r; c->alpha = c->rho / r; }
  if (OP == OP_R64) c->r64 = r;
  if (OP == OP_XR) {
    c->rr = r;
    c->pend = 1;
    c->iters++;
    if (r <= c->tol2) c->active = 0;
  }
}

/* X_j += alpha_j P_j (the previous iteration's update, if pending: alpha_j is still that
 * iteration's and P_j not yet overwritten); P_j = Z_j + beta_j P_j */
__global__ void __launch_bounds__(256) k_mp(double *__restrict__ X, size_t ldx, double *__restrict__ P,
                                            const double *__restrict__ Z, size_t ld, int n, const colsc *cs)
{
  const int j = blockIdx.y;
  if (!cs[j].active) return;
  const double b = cs[j].beta, a = cs[j].pend ? cs[j].alpha : 0.0;
  double *p = P + j * ld, *x = X + j * ldx;
  const double *z = Z + j * ld;
  const int gt = blockIdx.x * 256 + threadIdx.x, gs = gridDim.x * 256;
  for (int i = gt; i < n; i += gs) {
    const double pi = p[i];
    x[i] += a * pi;
    p[i] = z[i] + b * pi;
  }
}

/* the pending X_j += alpha_j P_j of columns that stopped */
__global__ void __launch_bounds__(256) k_mflush(double *__restrict__ X, size_t ldx, const double *__restrict__ P,
                                                size_t ld, int n, const colsc *cs)
{
  const int j = blockIdx.y;
  if (!cs[j].pend) return;
  const double a = cs[j].alpha;
  double *x = X + j * ldx;
  const double *p = P + j * ld;
  const int gt = blockIdx.x * 256 + threadIdx.x, gs = gridDim.x * 256;
  for (int i = gt; i < n; i += gs) x[i] += a * p[i];
}

/* R_j -= alpha_j Q_j; (R_j, R_j); column done below tol2 (its X update follows in k_mp or
 * k_mflush) */
__global__ void __launch_bounds__(256)
k_mxr(double *__restrict__ R, const double *__restrict__ Q, size_t ld, int n, const colsc *cs, double *parts)
{
  __shared__ double lds[4];
  const int j = blockIdx.y;
  const colsc *c = cs + j;
  if (!c->active) return;
  const double a = c->alpha;
  double *r = R + j * ld;
  const double *q = Q + j * ld;
  double v[1] = {0.0};
  const int gt = blockIdx.x * 256 + threadIdx.x, gs = gridDim.x * 256;
  for (int i = gt; i < n; i += gs) {
    const double ri = r[i] - a * q[i];
    r[i] = ri;
    v[0] += ri * ri;
  }
  store_partials<1>(v, parts + (size_t)j * kMaxBlocks, lds);
}

/* R_j = B_j - AX_j */
__global__ void __launch_bounds__(256)
k_mres(const double *__restrict__ B, size_t ldb, const double *__restrict__ AX, double *__restrict__ R, size_t ld, int n)
{
  const int j = blockIdx.y;
  const double *b = B + j * ldb, *ax = AX + j * ld;
  double *r = R + j * ld;
  const int gt = blockIdx.x * 256 + threadIdx.x, gs = gridDim.x * 256;
  for (int i = gt; i < n; i += gs) r[i] = b[i] - ax[i];
}

/* dst_j = src_j (matrix of k columns), used for the preconditioner-free case */
__global__ void __launch_bounds__(256) k_mcopy(double *__restrict__ D, const double *__restrict__ S, size_t ld, int n)
{
  const int j = blockIdx.y;
  const int gt = blockIdx.x * 256 + threadIdx.x, gs = gridDim.x * 256;
  for (int i = gt; i < n; i += gs) D[j * ld + i] = S[j * ld + i];
}

}  // namespace
}  // namespace ispm

using namespace ispm;

struct ispm_block_solver {
  ispm_plan *A = nullptr, *M1 = nullptr, *M2 = nullptr;
  ispm_solver_opts o{};
  int k = 1, n = 0, grid = 1;
  size_t ld = 0;
  double *R = nullptr, *Z = nullptr, *P = nullptr, *Q = nullptr, *T = nullptr;
  colsc *cs = nullptr, *hcs = nullptr;
  double *parts = nullptr;
  /* a batch of check_every iterations captured as a hipGraph, for the X it was captured with */
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double *graph_X = nullptr;
  size_t graph_ldx = 0;
  int graph_iters = 0;
};

namespace {

int mspmm(ispm_plan *P, int k, const double *X, size_t ldx, double *Y, size_t ldy)
{
  if (P->sym) return spmm_sym(P->symd, ISPM_F64, k, X, ldx, Y, ldy);
  return spmm_internal(&P->launch, &P->dev, ISPM_F64, k, X, ldx, Y, ldy, &P->hints,
                       P->has_pat ? &P->pat : nullptr, P->dev_uni);
}

template<int OP> void mfinish(ispm_block_solver *S)
{
  hipLaunchKernelGGL(k_mfin<OP>, dim3(S->k), dim3(kFinThreads), 0, stream(), (const double *)S->parts, S->grid, S->cs);
}

int pull(ispm_block_solver *S)
{
  ISPM_HIP(hipMemcpyAsync(S->hcs, S->cs, sizeof(colsc) * S->k, hipMemcpyDeviceToHost, stream()));
  ISPM_HIP(hipStreamSynchronize(stream()));
  return ISPM_OK;
}
int push(ispm_block_solver *S)
{
  ISPM_HIP(hipMemcpyAsync(S->cs, S->hcs, sizeof(colsc) * S->k, hipMemcpyHostToDevice, stream()));
  ISPM_HIP(hipStreamSynchronize(stream()));
  return ISPM_OK;
}

int iterate(ispm_block_solver *S, double *X, size_t ldx)
{
  int st;
  const int k = S->k, n = S->n;
  const size_t ld = S->ld;
  dim3 g(S->grid, k), b(256);
  hipStream_t sm = stream();
  if (S->M1) {
    if (S->M2) {
      if ((st = mspmm(S->M1, k, S->R, ld, S->T, ld))) return st;
      if ((st = mspmm(S->M2, k, S->T, ld, S->Z, ld))) return st;
    } else if ((st = mspmm(S->M1, k, S->R, ld, S->Z, ld))) {
      return st;
    }
  } else {
    hipLaunchKernelGGL(k_mcopy, g, b, 0, sm, S->Z, (const double *)S->R, ld, n);
  }
  hipLaunchKernelGGL((k_mdot<OP_RZ>), g, b, 0, sm, (const double *)S->R, (const double *)S->Z, ld, n, S->cs, S->parts);
  mfinish<OP_RZ>(S);
  hipLaunchKernelGGL(k_mp, g, b, 0, sm, X, ldx, S->P, (const double *)S->Z, ld, n, (const colsc *)S->cs);
  if ((st = mspmm(S->A, k, S->P, ld, S->Q, ld))) return st;
  hipLaunchKernelGGL((k_mdot<OP_PQ>), g, b, 0, sm, (const double *)S->P, (const double *)S->Q, ld, n, S->cs, S->parts);
  mfinish<OP_PQ>(S);
  hipLaunchKernelGGL(k_mxr, g, b, 0, sm, S->R, (const double *)S->Q, ld, n, S->cs, S->parts);
  mfinish<OP_XR>(S);
  ISPM_HIP(hipGetLastError());
  return ISPM_OK;
}

/* it iterations, replayed from a captured hipGraph when opts.use_graph is set (the
 * kernels read the per-column flags and scalars on the device, so a batch needs no host
 * step); eager launches otherwise or under ISPM_NO_GRAPH=1 */
int run_batch(ispm_block_solver *S, double *X, size_t ldx, int it)
{
  const char *ng = std::getenv("ISPM_NO_GRAPH");
  if (!S->o.use_graph || (ng && ng[0] == '1')) {
    for (int i = 0; i < it; i++) {
      const int st = iterate(S, X, ldx);
      if (st) return st;
    }
    return ISPM_OK;
  }
  if (!S->exec || S->graph_iters != it || S->graph_X != X || S->graph_ldx != ldx) {
    if (S->exec) { (void)hipGraphExecDestroy(S->exec); S->exec = nullptr; }
    if (S->graph) { (void)hipGraphDestroy(S->graph); S->graph = nullptr; }
    hipStream_t sm = stream();
    ISPM_HIP(hipStreamBeginCapture(sm, hipStreamCaptureModeThreadLocal));
    int st = ISPM_OK;
    for (int i = 0; i < it && !st; i++) st = iterate(S, X, ldx);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(sm, &g);
    if (st) { if (g) (void)hipGraphDestroy(g); return st; }
    if (e != hipSuccess) return hip_status(e, "hipStreamEndCapture");
    S->graph = g;
    ISPM_HIP(hipGraphInstantiate(&S->exec, g, nullptr, nullptr, 0));
    S->graph_iters = it;
    S->graph_X = X;
    S->graph_ldx = ldx;
  }
  ISPM_HIP(hipGraphLaunch(S->exec, stream()));
  return ISPM_OK;
}

/* R = B - A X, per-column ||R_j||^2 into cs[j].r64 */
int residuals(ispm_block_solver *S, const double *B, size_t ldb, const double *X, size_t ldx)
{
  int st;
  dim3 g(S->grid, S->k), b(256);
  if ((st = mspmm(S->A, S->k, X, ldx, S->Q, S->ld))) return st;
  hipLaunchKernelGGL(k_mres, g, b, 0, stream(), B, ldb, (const double *)S->Q, S->R, S->ld, S->n);
  hipLaunchKernelGGL((k_mdot<OP_R64>), g, b, 0, stream(), (const double *)S->R, (const double *)S->R, S->ld, S->n, S->cs, S->parts);
  mfinish<OP_R64>(S);
  ISPM_HIP(hipGetLastError());
  return ISPM_OK;
}

}  // namespace

extern "C" {

int ispm_block_solver_create(ispm_block_solver **out, ispm_plan *A, ispm_plan *M1, ispm_plan *M2,
                             int k, const ispm_solver_opts *o)
{
  if (!out || !A || !o) return fail(ISPM_ERR_INVALID, "block_solver_create: null argument");
  *out = nullptr;
  int st = ensure_init();
  if (st) return st;
  if (k < 1 || k > kMaxK) return fail(ISPM_ERR_UNSUPPORTED, "block_solver_create: k=%d outside 1..8", k);
  if (!M1 && M2) return fail(ISPM_ERR_INVALID, "block_solver_create: second factor without the first");
  if (!A->has_dev || A->T != ISPM_F64) return fail(ISPM_ERR_INVALID, "block_solver_create: A needs an FP64-vector device plan");
  for (const ispm_plan *M : {A, M1, M2})
    if (M && M->split) return fail(ISPM_ERR_UNSUPPORTED, "block_solver_create: split-diagonal plans multiply one vector");
  for (const ispm_plan *M : {M1, M2})
    if (M && (!M->has_dev || M->T != ISPM_F64 || M->L.n_rows != A->L.n_rows))
      return fail(ISPM_ERR_INVALID, "block_solver_create: preconditioner plans must be FP64 device plans of A's size");
  ispm_block_solver *S = new ispm_block_solver;
  S->A = A; S->M1 = M1; S->M2 = M2; S->o = *o; S->k = k;
  if (S->o.check_every <= 0) S->o.check_every = 16;
  if (S->o.maxit <= 0) S->o.maxit = 100000;
  S->n = A->L.n_rows;
  S->ld = ((size_t)S->n + 63) / 64 * 64;
  S->grid = std::max(1, std::min((S->n + 511) / 512, kMaxBlocks));
  const size_t vb = sizeof(double) * S->ld * k;
  double **vs[] = {&S->R, &S->Z, &S->P, &S->Q, &S->T};
  for (double **v : vs)
    if ((st = ispm_malloc((void **)v, vb))) { ispm_block_solver_destroy(S); return st; }
  if ((st = ispm_malloc((void **)&S->cs, sizeof(colsc) * kMaxK)) ||
      (st = ispm_malloc((void **)&S->parts, sizeof(double) * kMaxBlocks * kMaxK))) {
    ispm_block_solver_destroy(S);
    return st;
  }
  hipError_t e = hipHostMalloc((void **)&S->hcs, sizeof(colsc) * kMaxK, hipHostMallocDefault);
  if (e != hipSuccess) { ispm_block_solver_destroy(S); return hip_status(e, "hipHostMalloc"); }
  *out = S;
  return ISPM_OK;
}

int ispm_block_solver_solve(ispm_block_solver *S, const double *B, size_t ldb, double *X,
                            size_t ldx, ispm_solve_stats *stats)
{
  if (!S || !B || !X) return fail(ISPM_ERR_INVALID, "block_solver_solve: null argument");
  if (ldb < (size_t)S->n || ldx < (size_t)S->n) return fail(ISPM_ERR_INVALID, "block_solver_solve: leading dimension below n");
  int st;
  const int k = S->k;
  const double tol = S->o.tol > 0 ? S->o.tol : 1e-10;
  const double margin = 0.9;
  /* ||B_j|| */
  std::vector<double> bn(k), rec(k);
  std::memset(S->hcs, 0, sizeof(colsc) * k);
  if ((st = push(S))) return st;
  {
    dim3 g(S->grid, k), b(256);
    hipLaunchKernelGGL((k_mdot<OP_R64>), g, b, 0, stream(), B, B, ldb, S->n, S->cs, S->parts);
    mfinish<OP_R64>(S);
    ISPM_HIP(hipGetLastError());
    if ((st = pull(S))) return st;
    for (int j = 0; j < k; j++) bn[j] = S->hcs[j].r64 > 0 ? std::sqrt(S->hcs[j].r64) : 1.0;
  }
  if ((st = residuals(S, B, ldb, X, ldx)) || (st = pull(S))) return st;
  for (int j = 0; j < k; j++) {
    colsc &c = S->hcs[j];
    const double rel = std::sqrt(c.r64) / bn[j];
    rec[j] = margin * tol * bn[j];
    c.rho = INFINITY;
    c.tol2 = rec[j] * rec[j];
    c.active = rel > tol;
    c.iters = 0;
  }
  if ((st = push(S)) || (st = ispm_zero(S->P, sizeof(double) * S->ld * k))) return st;
  std::vector<int> converged(k, 0), repl(k, 0);
  int total_iters = 0;
  for (int guard = 0; guard < 100000; guard++) {
    if ((st = run_batch(S, X, ldx, S->o.check_every))) return st;
    if ((st = pull(S))) return st;
    int active = 0, maxit = 0;
    for (int j = 0; j < k; j++) { active += S->hcs[j].active ? 1 : 0; maxit = std::max(maxit, (int)S->hcs[j].iters); }
    total_iters = maxit;
    if (active && maxit < S->o.maxit) continue;
    /* every column stopped (or out of iterations): apply the pending X updates, then check
     * the true residuals */
    {
      dim3 g(S->grid, k), b(256);
      hipLaunchKernelGGL(k_mflush, g, b, 0, stream(), X, ldx, (const double *)S->P, S->ld, S->n, (const colsc *)S->cs);
      ISPM_HIP(hipGetLastError());
    }
    std::vector<colsc> before(S->hcs, S->hcs + k);
    if ((st = residuals(S, B, ldb, X, ldx)) || (st = pull(S))) return st;
    int again = 0;
    for (int j = 0; j < k; j++) {
      colsc &c = S->hcs[j];
      const double rel = std::sqrt(c.r64) / bn[j];
      converged[j] = rel <= tol;
      c.iters = before[j].iters;
      c.pend = 0;   /* flushed above */
      if (!converged[j] && maxit < S->o.maxit) {
        /* drifted: restart this column from the true residual with a tighter target */
        rec[j] *= 0.5;
        c.rho = INFINITY;
        c.tol2 = rec[j] * rec[j];
        c.active = 1;
        repl[j]++;
        again++;
      } else {
        c.active = 0;
      }
    }
    if (!again) break;
    if ((st = push(S))) return st;
    /* restarted columns need p = 0; others are inactive and untouched */
    for (int j = 0; j < k; j++)
      if (S->hcs[j].active) ISPM_HIP(hipMemsetAsync(S->P + j * S->ld, 0, sizeof(double) * S->ld, stream()));
  }
  if ((st = pull(S))) return st;
  int all = 1;
  for (int j = 0; j < k; j++) {
    if (stats) {
      stats[j].iterations = (int)S->hcs[j].iters;
      stats[j].replacements = repl[j];
      stats[j].converged = converged[j];
      stats[j].relres = std::sqrt(S->hcs[j].r64) / bn[j];
      stats[j].l1_resnorm = 0;
    }
    all &= converged[j];
  }
  (void)total_iters;
  return all ? ISPM_OK : fail(ISPM_ERR_NOT_CONVERGED, "block solver: not every column converged");
}

void ispm_block_solver_destroy(ispm_block_solver *S)
{
  if (!S) return;
  for (double *v : {S->R, S->Z, S->P, S->Q, S->T}) ispm_free(v);
  ispm_free(S->cs);
  ispm_free(S->parts);
  if (S->hcs) (void)hipHostFree(S->hcs);
  if (S->exec) (void)hipGraphExecDestroy(S->exec);
  if (S->graph) (void)hipGraphDestroy(S->graph);
  delete S;
}

}  // extern "C"
