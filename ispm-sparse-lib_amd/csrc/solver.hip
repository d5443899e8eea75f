/* solver.hip -- device-resident FSAI-preconditioned CG on the library stream.
 *
 * Four variants share one driver:
 *  - FP64 pipelined PCG (Ghysels-Vanroose): per iteration t = Ainv1 w, m = Ainv2 t, n = A m
 *    and ONE fused kernel that evaluates the reference's scalar update
 *    (beta = g/g_old, alpha = g/(delta - beta g/alpha_old), util/cuda/sblas.cu:156-165) in every
 *    block and then the eight recurrences with their three sums (sblas.cu:177-217), plus
 *    (r,r) for the convergence test.
 *  - FP64 PCG (Hestenes-Stiefel): z = M r, rho = (r,z), p = z + beta p, q = A p, alpha,
 *    x += alpha p, r -= alpha q; x += alpha p runs in the next iteration's p update, which
 *    reads p anyway, and once more when the iterations stop.  With a split-diagonal
 *    preconditioner, r -= alpha q also moves into the next iteration: its product z = M r
 *    reads r and q, multiplies r - alpha q and writes that vector to a second residual
 *    buffer (the two swap every iteration, so no product reads an entry another workgroup
 *    already updated), with (r, r) -- four launches per iteration instead of five.
 *  - single-reduction PCG (Chronopoulos-Gear): z = M r with (r,z), w = A z with (z,w), one
 *    finishing kernel (alpha, beta by the reference's formulas; the convergence test on the
 *    (r,r) of the previous update) and one fused update p = z + beta p, s = w + beta s,
 *    x += alpha p, r -= alpha s: four launches per iteration instead of seven.
 *  - mixed precision (PCG, single-reduction PCG or pipelined PCG): FP32 vectors, FP32 A and
 *    FP32/BF16 factors; x is accumulated in FP32 between residual replacements and added to
 *    the FP64 x there, and the residual is replaced by the FP64 true residual (b - A x with
 *    the FP32-stored values read into FP64 accumulators) whenever the FP32 recursive
 *    residual has dropped by reliable_delta ("reliable updates").  PCG and PCG1 keep their
 *    search direction across a replacement (rescaled); the pipelined method restarts its
 *    recurrences from the replaced residual, as the FP64 solve does when it tightens.
 * All scalars live in device memory in FP64; every reduction is two-phase (block partials,
 * then a one-block finishing kernel) and deterministic.  Iterations are launched in
 * hipGraph-captured batches; once the recursive residual reaches its target a device flag
 * turns the rest of the batch into no-ops, and the host then decides on the FP64 true
 * residual.
 */
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

struct scalars {
  double rho, pq, alpha, beta;
  double gamma, gamma_old, delta, alpha_old;
  double rr, l1, tol2, r64;
  double rscale;
  unsigned done, iters;
  unsigned rr_valid;   /* PCG1, and PCG with a preconditioner: the (r,r) partials hold a residual to test */
  double xsum, normf, l1true;   /* normalised-residual solves: sum x, norm factor, sum |b - A x| */
};

enum { OP_PQ, OP_RZ, OP_GAMMA, OP_DELTA, OP_R64, OP_XR, OP_PPCG, OP_SUM, OP_NORMF, OP_L1 };

template<typename T> struct vtraits {
  static constexpr int L = 16 / sizeof(T);
};

template<typename T, int OP>
__global__ void __launch_bounds__(256)
k_dot(const T *__restrict__ x, const T *__restrict__ y, int n, const scalars *s, double *parts,
      int skip_if_done)
{
  __shared__ double lds[4];
  if (skip_if_done && s->done) return;
  constexpr int L = vtraits<T>::L;
  double v[1] = {0.0};
  const int gt = blockIdx.x * 256 + threadIdx.x, gs = gridDim.x * 256, nv = n / L;
  for (int i = gt; i < nv; i += gs) {
    pack16<T> a = ld16(x, i), b = ld16(y, i);
#pragma unroll
    for (int l = 0; l < L; l++) v[0] += (double)a.v[l] * (double)b.v[l];
  }
  for (int i = nv * L + gt; i < n; i += gs) v[0] += (double)x[i] * (double)y[i];
  store_partials<1>(v, parts, lds);
}

/* Second phase of every solver reduction: one block sums the partials of the kernel before
 * it and applies the scalar part of the recurrence.  A finishing kernel skipped when the
 * solve is done leaves the scalars as they are, like its reduction kernel. */
template<int OP, unsigned GMAX>
__global__ void __launch_bounds__(kFinThreads) k_fin(const double *parts, int G, scalars *s, int skip_if_done)
{
  constexpr int NS = OP == OP_PPCG ? 4 : OP == OP_NORMF ? 2 : 1;
  __shared__ double lds[NS * kFinThreads / 64];
  /* the partials are requested together with the done flag, not after it (one memory round
   * trip less per finishing kernel); the flag is the same for the whole block */
  const unsigned done = skip_if_done ? s->done : 0u;
  partials<NS, GMAX> pt;
  pt.load(parts, G);
  if (done) return;
  double out[NS];
  pt.sum(lds, out);
  if (threadIdx.x) return;
  const double r = out[0];
  if (OP == OP_PQ) { s->pq = r; s->alpha = s->rho / r; }
  if (OP == OP_RZ) { s->beta = r / s->rho; s->rho = r; }
  if (OP == OP_GAMMA) s->gamma = r;
  if (OP == OP_DELTA) s->delta = r;
  if (OP == OP_R64) s->r64 = r;
  if (OP == OP_SUM) s->xsum = r;
  if (OP == OP_NORMF) { s->normf = out[0]; s->l1true = out[1]; }
  if (OP == OP_L1) s->l1true = r;
  if (OP == OP_XR) {
    s->rr = r;
    s->iters++;
    if (r <= s->tol2) s->done = 1;
  }
  if (OP == OP_PPCG) {
    /* the same scalar update every block of k_ppcg made from the same inputs */
    const double g = s->gamma, go = s->gamma_old, dl = s->delta, ao = s->alpha_old;
    const double bd = g / go;
    const double ad = g / (dl - bd * g / ao);
    s->gamma_old = g;
    s->alpha_old = ad;
    s->alpha = ad;
    s->beta = bd;
    s->l1 = out[0];
    s->gamma = out[1];
    s->delta = out[2];
    s->rr = out[3];
    s->iters++;
    if (out[3] <= s->tol2) s->done = 1;
  }
}

/* r -= alpha q; (r,r) partials.  The matching x += alpha p is deferred to the next k_px,
 * which reads p anyway (one vector fewer per iteration), or to k_xflush when the loop stops;
 * x gets the same values in the same order. */
template<typename T>
__global__ void __launch_bounds__(256)
k_r(T *__restrict__ r, const T *__restrict__ q, int n, const scalars *s, double *parts)
{
  __shared__ double lds[4];
  if (s->done) return;
  constexpr int L = vtraits<T>::L;
  const T a = (T)s->alpha;
  double v[1] = {0.0};
  const int gt = blockIdx.x * 256 + threadIdx.x, gs = gridDim.x * 256, nv = n / L;
  for (int i = gt; i < nv; i += gs) {
    pack16<T> rr = ld16(r, i), qq = ld16(q, i);
#pragma unroll
    for (int l = 0; l < L; l++) {
      rr.v[l] = rr.v[l] - a * qq.v[l];
      v[0] += (double)rr.v[l] * (double)rr.v[l];
    }
    st16(r, i, rr);
  }
  for (int i = nv * L + gt; i < n; i += gs) {
    T ri = r[i] - a * q[i];
    r[i] = ri;
    v[0] += (double)ri * (double)ri;
  }
  store_partials<1>(v, parts, lds);
}

/* Single-reduction PCG (Chronopoulos-Gear), the one finishing kernel of an iteration: sums
 * gamma = (r, z) (partials of the preconditioner product), delta = (z, w) (partials of
 * w = A z) and the (r, r) partials k_cg_update left one iteration earlier.  When that
 * residual meets the target the solve is done and nothing more of the batch runs; otherwise
 * beta = gamma / gamma_old and alpha = gamma / (delta - beta gamma / alpha_old), the scalar
 * update of util/cuda/sblas.cu:156-165 of the reference. */
__global__ void __launch_bounds__(kFinThreads)
k_fin_cg(const double *pg, unsigned Gg, const double *pd, unsigned Gd, const double *pr, unsigned Gr,
         scalars *s)
{
  __shared__ double lds[3 * kFinThreads / 64];
  if (s->done) return;
  double out[3] = {0.0, 0.0, 0.0};
  const unsigned t = threadIdx.x;
  /* eight loads of each sum in flight per step; a fixed order keeps the sums reproducible */
  constexpr unsigned C = 8 * kFinThreads;
  const unsigned Gm = Gg > Gd ? (Gg > Gr ? Gg : Gr) : (Gd > Gr ? Gd : Gr);
  for (unsigned i0 = 0; i0 < Gm; i0 += C) {
    double a[8], b[8], c[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const unsigned i = i0 + t + k * kFinThreads;
      a[k] = i < Gg ? pg[i] : 0.0;
      b[k] = i < Gd ? pd[i] : 0.0;
      c[k] = i < Gr ? pr[i] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < 8; k++) { out[0] += a[k]; out[1] += b[k]; out[2] += c[k]; }
  }
  dev::block_sum<kFinThreads, 3>(out, lds);
  if (t) return;
  if (s->rr_valid) {
    s->rr = out[2];
    if (out[2] <= s->tol2) { s->done = 1; return; }
  }
  const double g = out[0], dl = out[1];
  const double bd = g / s->gamma_old;
  const double ad = g / (dl - bd * g / s->alpha_old);
  s->gamma = g;
  s->delta = dl;
  s->gamma_old = g;
  s->alpha_old = ad;
  s->alpha = ad;
  s->beta = bd;
  s->rr_valid = 1;
  s->iters++;
}

/* Hestenes-Stiefel PCG with a preconditioner: one finishing kernel for the preconditioner
 * product's (r, z) partials and the (r, r) partials k_r left one iteration earlier.  When
 * that residual meets the target the solve is done (the iteration that produced it is
 * counted, as a finish right after k_r would have) and nothing more of the batch runs;
 * otherwise beta = (r, z) / rho and rho = (r, z). */
__global__ void __launch_bounds__(kFinThreads)
k_fin_rz_rr(const double *pg, unsigned Gg, const double *pr, unsigned Gr, scalars *s)
{
  __shared__ double lds[2 * kFinThreads / 64];
  /* the done flag is read with the first partials, not before them */
  const unsigned done = s->done;
  double out[2] = {0.0, 0.0};
  const unsigned t = threadIdx.x;
  /* eight loads of each sum in flight per step; a fixed order keeps the sums reproducible */
  constexpr unsigned C = 8 * kFinThreads;
  const unsigned Gm = Gg > Gr ? Gg : Gr;
  for (unsigned i0 = 0; i0 < Gm; i0 += C) {
    double a[8], c[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const unsigned i = i0 + t + k * kFinThreads;
      a[k] = i < Gg ? pg[i] : 0.0;
      c[k] = i < Gr ? pr[i] : 0.0;
    }
    if (done) return;   /* the same for the whole block */
#pragma unroll
    for (int k = 0; k < 8; k++) { out[0] += a[k]; out[1] += c[k]; }
  }
  if (done) return;
  dev::block_sum<kFinThreads, 2>(out, lds);
  if (t) return;
  if (s->rr_valid) {
    s->rr = out[1];
    s->iters++;
    if (out[1] <= s->tol2) { s->done = 1; return; }
  }
  s->beta = out[0] / s->rho;
  s->rho = out[0];
  s->rr_valid = 1;
}

/* p = z + beta p; s = w + beta s; x += alpha p; r -= alpha s; (r, r) partials */
template<typename T>
__global__ void __launch_bounds__(256)
k_cg_update(const T *__restrict__ z, const T *__restrict__ w, T *__restrict__ p, T *__restrict__ sv,
            T *__restrict__ x, T *__restrict__ r, int n, const scalars *s, double *parts)
{
  __shared__ double lds[4];
  if (s->done) return;
  constexpr int L = vtraits<T>::L;
  const T a = (T)s->alpha, b = (T)s->beta;
  double v[1] = {0.0};
  auto one = [&](T zi, T wi, T &pi, T &si, T &xi, T &ri) {
    pi = zi + b * pi;
    si = wi + b * si;
    xi = xi + a * pi;
    ri = ri - a * si;
    v[0] += (double)ri * (double)ri;
  };
  const int gt = blockIdx.x * 256 + threadIdx.x, gs = gridDim.x * 256, nv = n / L;
  for (int i = gt; i < nv; i += gs) {
    pack16<T> zz = ld16(z, i), ww = ld16(w, i), pp = ld16(p, i), ss = ld16(sv, i), xx = ld16(x, i),
              rr = ld16(r, i);
#pragma unroll
    for (int l = 0; l < L; l++) one(zz.v[l], ww.v[l], pp.v[l], ss.v[l], xx.v[l], rr.v[l]);
    st16(p, i, pp); st16(sv, i, ss); st16(x, i, xx); st16(r, i, rr);
  }
  for (int i = nv * L + gt; i < n; i += gs) {
    T pi = p[i], si = sv[i], xi = x[i], ri = r[i];
    one(z[i], w[i], pi, si, xi, ri);
    p[i] = pi; sv[i] = si; x[i] = xi; r[i] = ri;
  }
  store_partials<1>(v, parts, lds);
}

/* x += alpha p (the previous iteration's update: alpha is still that iteration's, p not yet
 * overwritten); p = z + beta p */
template<typename T>
__global__ void __launch_bounds__(256)
k_px(T *__restrict__ x, T *__restrict__ p, const T *__restrict__ z, int n, const scalars *s)
{
  if (s->done) return;
  constexpr int L = vtraits<T>::L;
  const T a = (T)s->alpha, b = (T)s->beta;
  const int gt = blockIdx.x * 256 + threadIdx.x, gs = gridDim.x * 256, nv = n / L;
  for (int i = gt; i < nv; i += gs) {
    pack16<T> xx = ld16(x, i), pp = ld16(p, i), zz = ld16(z, i);
#pragma unroll
    for (int l = 0; l < L; l++) {
      xx.v[l] = xx.v[l] + a * pp.v[l];
      pp.v[l] = zz.v[l] + b * pp.v[l];
    }
    st16(x, i, xx);
    st16(p, i, pp);
  }
  for (int i = nv * L + gt; i < n; i += gs) {
    x[i] = x[i] + a * p[i];
    p[i] = z[i] + b * p[i];
  }
}

/* x += alpha p: the update k_px would have made, when the iterations stop */
template<typename T>
__global__ void __launch_bounds__(256) k_xflush(T *__restrict__ x, const T *__restrict__ p, int n, const scalars *s)
{
  const T a = (T)s->alpha;
  const int gt = blockIdx.x * 256 + threadIdx.x, gs = gridDim.x * 256;
  for (int i = gt; i < n; i += gs) x[i] = x[i] + a * p[i];
}

template<typename T>
__global__ void __launch_bounds__(256) k_copy(T *__restrict__ d, const T *__restrict__ a, int n, const scalars *s)
{
  if (s && s->done) return;
  const int gt = blockIdx.x * 256 + threadIdx.x, gs = gridDim.x * 256;
  for (int i = gt; i < n; i += gs) d[i] = a[i];
}

template<typename T>
__global__ void __launch_bounds__(256) k_scale(T *__restrict__ d, int n, double c)
{
  const int gt = blockIdx.x * 256 + threadIdx.x, gs = gridDim.x * 256;
  for (int i = gt; i < n; i += gs) d[i] = (T)((double)d[i] * c);
}

/* Fused pipelined-PCG step: scalar update then the recurrences of sblas.cu:192-200. */
template<typename T>
__global__ void __launch_bounds__(256)
k_ppcg(const T *__restrict__ nn, const T *__restrict__ m, T *__restrict__ p, T *__restrict__ s_,
       T *__restrict__ q, T *__restrict__ z, T *__restrict__ x, T *__restrict__ r,
       T *__restrict__ u, T *__restrict__ w, int n, const scalars *s, double *parts)
{
  __shared__ double lds[16];
  if (s->done) return;
  const double g = s->gamma, go = s->gamma_old, dl = s->delta, ao = s->alpha_old;
  const double bd = g / go;
  const double ad = g / (dl - bd * g / ao);
  const T a = (T)ad, b = (T)bd;
  double acc[4] = {0, 0, 0, 0};
  constexpr int L = vtraits<T>::L;
  auto one = [&](T &zi, T &qi, T &si, T &pi, T &ui, T &wi, T &ri, T &xi, T ni, T mi) {
    zi = ni + b * zi;
    qi = mi + b * qi;
    si = wi + b * si;
    pi = ui + b * pi;
    ui = ui - a * qi;
    wi = wi - a * zi;
    ri = ri - a * si;
    xi = xi + a * pi;
    acc[0] += fabs((double)ri);
    acc[1] += (double)ri * (double)ui;
    acc[2] += (double)wi * (double)ui;
    acc[3] += (double)ri * (double)ri;
  };
  const int gt = blockIdx.x * 256 + threadIdx.x, gs = gridDim.x * 256, nv = n / L;
  for (int i = gt; i < nv; i += gs) {
    pack16<T> zn = ld16(nn, i), mm = ld16(m, i), zz = ld16(z, i), qq = ld16(q, i), ss = ld16(s_, i),
              pp = ld16(p, i), uu = ld16(u, i), ww = ld16(w, i), rr = ld16(r, i), xx = ld16(x, i);
#pragma unroll
    for (int l = 0; l < L; l++)
      one(zz.v[l], qq.v[l], ss.v[l], pp.v[l], uu.v[l], ww.v[l], rr.v[l], xx.v[l], zn.v[l], mm.v[l]);
    st16(z, i, zz); st16(q, i, qq); st16(s_, i, ss); st16(p, i, pp);
    st16(u, i, uu); st16(w, i, ww); st16(r, i, rr); st16(x, i, xx);
  }
  for (int i = nv * L + gt; i < n; i += gs) {
    T zi = z[i], qi = q[i], si = s_[i], pi = p[i], ui = u[i], wi = w[i], ri = r[i], xi = x[i];
    one(zi, qi, si, pi, ui, wi, ri, xi, nn[i], m[i]);
    z[i] = zi; q[i] = qi; s_[i] = si; p[i] = pi; u[i] = ui; w[i] = wi; r[i] = ri; x[i] = xi;
  }
  store_partials<4>(acc, parts, lds);
}

/* ---- OpenFOAM's normalised residual (lduMatrix::solver::normFactor): with
 * xbar = average(x) and sumA = A 1 (row sums),
 *   normFactor = sum_i |(A x)_i - sumA_i xbar| + |b_i - sumA_i xbar| + 1e-20,
 *   residual   = sum_i |b_i - (A x)_i| / normFactor,
 * over the nr rows of the matrix (not the padding). ---- */
__global__ void __launch_bounds__(256) k_ones(double *__restrict__ v, int nr, int n)
{
  const int gs = gridDim.x * 256;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gs) v[i] = i < nr ? 1.0 : 0.0;
}

/* sum x_i (or sum |x_i| when ABS) over i < nr */
template<bool ABS>
__global__ void __launch_bounds__(256) k_sum1(const double *__restrict__ x, int nr, double *parts)
{
  __shared__ double lds[4];
  double v[1] = {0.0};
  const int gs = gridDim.x * 256;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < nr; i += gs) v[0] += ABS ? std::fabs(x[i]) : x[i];
  store_partials<1>(v, parts, lds);
}

__global__ void __launch_bounds__(256)
k_normf(const double *__restrict__ b, const double *__restrict__ Ax, const double *__restrict__ sumA, int nr,
        const scalars *s, double *parts)
{
  __shared__ double lds[8];
  const double xbar = s->xsum / (double)nr;
  double v[2] = {0.0, 0.0};
  const int gs = gridDim.x * 256;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < nr; i += gs) {
    const double pa = sumA[i] * xbar;
    v[0] += std::fabs(Ax[i] - pa) + std::fabs(b[i] - pa);
    v[1] += std::fabs(b[i] - Ax[i]);
  }
  store_partials<2>(v, parts, lds);
}

/* r = b - Ax (FP64), (r,r) -> s->r64 */
__global__ void __launch_bounds__(256)
k_resid(const double *__restrict__ b, const double *__restrict__ Ax, double *__restrict__ r, int n,
        double *parts)
{
  __shared__ double lds[4];
  double v[1] = {0.0};
  const int gt = blockIdx.x * 256 + threadIdx.x, gs = gridDim.x * 256;
  for (int i = gt; i < n; i += gs) {
    const double ri = b[i] - Ax[i];
    r[i] = ri;
    v[0] += ri * ri;
  }
  store_partials<1>(v, parts, lds);
}

/* dst = src / ||src|| with ||src||^2 = s->r64; s->rscale = ||src|| */
template<typename T>
__global__ void __launch_bounds__(256) k_cast_scaled(T *__restrict__ d, const double *__restrict__ a, int n, scalars *s)
{
  const double nrm = sqrt(s->r64);
  const double c = nrm > 0 ? 1.0 / nrm : 0.0;
  const int gt = blockIdx.x * 256 + threadIdx.x, gs = gridDim.x * 256;
  for (int i = gt; i < n; i += gs) d[i] = (T)(a[i] * c);
  if (gt == 0) s->rscale = nrm;
}

/* x64 += rscale * xs; xs = 0 */
__global__ void __launch_bounds__(256) k_x64(double *__restrict__ x, float *__restrict__ xs, int n, const scalars *s)
{
  const double c = s->rscale;
  const int gt = blockIdx.x * 256 + threadIdx.x, gs = gridDim.x * 256;
  for (int i = gt; i < n; i += gs) {
    x[i] += c * (double)xs[i];
    xs[i] = 0.0f;
  }
}

}  // namespace
}  // namespace ispm

using namespace ispm;

struct ispm_solver {
  ispm_plan *A = nullptr, *M1 = nullptr, *M2 = nullptr;
  ispm_plan *Ares = nullptr;  /* plan used for the FP64 true residual (default: A) */
  ispm_solver_opts o{};
  int n = 0;
  int T = ISPM_F64;
  size_t ts = 8;
  /* the working vectors carry zero guard zones of xguard elements (gbytes bytes) on both
   * sides, so that products of a symmetric A may read their x windows unchecked */
  unsigned xguard = 0;
  size_t gbytes = 0;
  bool mixed = false, pipelined = false, single = false;
  /* PCG with one split-diagonal preconditioner plan: r -= alpha q fused into z = M r, which
   * writes the updated residual to r2; r and r2 swap every iteration (batches are even) */
  bool fused_r = false;
  void *r2 = nullptr;
  void *r = nullptr, *u = nullptr, *w = nullptr, *m = nullptr, *nv = nullptr, *z = nullptr,
       *q = nullptr, *sv = nullptr, *p = nullptr, *t = nullptr, *xs = nullptr;
  double *r64 = nullptr, *tmp64 = nullptr;
  double *sumA = nullptr;  /* row sums of A (normalised-residual solves; allocated on first use) */
  scalars *s = nullptr;
  scalars *hs = nullptr;   /* pinned host mirror */
  double *parts = nullptr; /* reduction partials, private to this solver */
  int grid = 1;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  int graph_iters = 0;
  const void *graph_x = nullptr;   /* x the captured graph was recorded with (FP64 variants) */
};

namespace {

int alloc(void **p, size_t bytes) { return ispm_malloc(p, bytes); }

/* a working vector with the solver's guard zones (zero-filled like every allocation) */
int alloc_guarded(ispm_solver *S, void **p, size_t bytes)
{
  void *base = nullptr;
  int st = ispm_malloc(&base, bytes + 2 * S->gbytes);
  *p = base ? static_cast<char *>(base) + S->gbytes : nullptr;
  return st;
}

void free_guarded(ispm_solver *S, void *p)
{
  if (p) ispm_free(static_cast<char *>(p) - S->gbytes);
}

int spmv(ispm_plan *P, int vt, const void *x, void *y, const unsigned *skip, unsigned x_guard = 0)
{
  if (P->split) {
    const split_diag sd{P->dev_diag, P->dev_rmap, P->dev_rmask, P->split_n, P->split_m, P->E};
    return spmv_split(sd, &P->launch, &P->dev, vt, x, y, skip, nullptr, nullptr, nullptr, &P->hints);
  }
  if (P->sym) return spmv_sym(P->symd, vt, x, y, skip, nullptr, nullptr, nullptr, x_guard);
  return spmv_internal(&P->launch, &P->dev, vt, x, y, skip, P->has_pat ? &P->pat : nullptr, P->dev_uni,
                       &P->hints);
}

/* G partials of the reduction before (the solver's grid when 0) */
template<int OP> void finish(ispm_solver *S, int skip, unsigned G = 0)
{
  if (G > (unsigned)kMaxBlocks)
    hipLaunchKernelGGL((k_fin<OP, kMaxDotParts>), dim3(1), dim3(kFinThreads), 0, stream(),
                       (const double *)S->parts, (int)G, S->s, skip);
  else
    hipLaunchKernelGGL((k_fin<OP, kMaxBlocks>), dim3(1), dim3(kFinThreads), 0, stream(),
                       (const double *)S->parts, G ? (int)G : S->grid, S->s, skip);
}

template<typename T> int launch_dot(ispm_solver *S, int op, const void *x, const void *y, int skip)
{
  const T *a = (const T *)x, *b = (const T *)y;
  hipStream_t st = stream();
  dim3 g(S->grid), bl(256);
  switch (op) {
    case OP_PQ: hipLaunchKernelGGL((k_dot<T, OP_PQ>), g, bl, 0, st, a, b, S->n, S->s, S->parts, skip); finish<OP_PQ>(S, skip); break;
    case OP_RZ: hipLaunchKernelGGL((k_dot<T, OP_RZ>), g, bl, 0, st, a, b, S->n, S->s, S->parts, skip); finish<OP_RZ>(S, skip); break;
    case OP_GAMMA: hipLaunchKernelGGL((k_dot<T, OP_GAMMA>), g, bl, 0, st, a, b, S->n, S->s, S->parts, skip); finish<OP_GAMMA>(S, skip); break;
    case OP_DELTA: hipLaunchKernelGGL((k_dot<T, OP_DELTA>), g, bl, 0, st, a, b, S->n, S->s, S->parts, skip); finish<OP_DELTA>(S, skip); break;
    default: hipLaunchKernelGGL((k_dot<T, OP_R64>), g, bl, 0, st, a, b, S->n, S->s, S->parts, skip); finish<OP_R64>(S, skip); break;
  }
  ISPM_HIP(hipGetLastError());
  return ISPM_OK;
}

int dot(ispm_solver *S, int op, const void *x, const void *y, int skip, int vt)
{
  return vt == ISPM_F64 ? launch_dot<double>(S, op, x, y, skip) : launch_dot<float>(S, op, x, y, skip);
}

/* z = M r (two SpMVs, or a copy without preconditioner) */
int apply_M(ispm_solver *S, const void *rin, void *zout, const unsigned *skip)
{
  int st;
  if (!S->M1) {
    if (S->T == ISPM_F64) hipLaunchKernelGGL(k_copy<double>, dim3(S->grid), dim3(256), 0, stream(), (double *)zout, (const double *)rin, S->n, skip ? S->s : nullptr);
    else hipLaunchKernelGGL(k_copy<float>, dim3(S->grid), dim3(256), 0, stream(), (float *)zout, (const float *)rin, S->n, skip ? S->s : nullptr);
    ISPM_HIP(hipGetLastError());
    return ISPM_OK;
  }
  if (!S->M2) return spmv(S->M1, S->T, rin, zout, skip);
  if ((st = spmv(S->M1, S->T, rin, S->t, skip))) return st;
  return spmv(S->M2, S->T, S->t, zout, skip);
}

/* y = P x with the partials of (d, y) in S->parts: fused into the SpMV when its kernel can,
 * else a separate dot kernel.  Returns the number of partials in G. */
template<typename T>
int spmv_dot(ispm_solver *S, ispm_plan *P, const void *x, void *y, const void *d,
             const unsigned *skip, unsigned &G, double *parts = nullptr)
{
  G = 0;
  if (!parts) parts = S->parts;
  int st;
  if (P->split) {
    const split_diag sd{P->dev_diag, P->dev_rmap, P->dev_rmask, P->split_n, P->split_m, P->E};
    st = spmv_split(sd, &P->launch, &P->dev, S->T, x, y, skip, d, parts, &G, &P->hints);
  } else if (P->sym) {
    st = spmv_sym(P->symd, S->T, x, y, skip, d, parts, &G, S->xguard);   /* x: a working vector */
  } else {
    st = spmv_dot_internal(&P->launch, &P->dev, S->T, x, y, skip, P->has_pat ? &P->pat : nullptr,
                           P->dev_uni, d, parts, &G, &P->hints);
  }
  if (st || G) return st;
  hipLaunchKernelGGL((k_dot<T, OP_PQ>), dim3(S->grid), dim3(256), 0, stream(), (const T *)d, (const T *)y,
                     S->n, S->s, parts, skip ? 1 : 0);
  ISPM_HIP(hipGetLastError());
  G = (unsigned)S->grid;
  return ISPM_OK;
}

const unsigned *done_flag(ispm_solver *S) { return &S->s->done; }

/* one iteration (enqueued; may be captured) */
int iterate(ispm_solver *S)
{
  int st;
  const unsigned *sk = done_flag(S);
  hipStream_t sm = stream();
  dim3 g(S->grid), b(256);
  if (S->pipelined) {
    if ((st = apply_M(S, S->w, S->m, sk))) return st;
    if ((st = spmv(S->A, S->T, S->m, S->nv, sk, S->xguard))) return st;
    if (S->T == ISPM_F64)
      hipLaunchKernelGGL(k_ppcg<double>, g, b, 0, sm, (const double *)S->nv, (const double *)S->m,
                         (double *)S->p, (double *)S->sv, (double *)S->q, (double *)S->z,
                         (double *)S->xs, (double *)S->r, (double *)S->u, (double *)S->w, S->n, S->s,
                         S->parts);
    else
      hipLaunchKernelGGL(k_ppcg<float>, g, b, 0, sm, (const float *)S->nv, (const float *)S->m,
                         (float *)S->p, (float *)S->sv, (float *)S->q, (float *)S->z,
                         (float *)S->xs, (float *)S->r, (float *)S->u, (float *)S->w, S->n, S->s,
                         S->parts);
    finish<OP_PPCG>(S, 1);
    ISPM_HIP(hipGetLastError());
    return ISPM_OK;
  }
  if (S->single) {
    /* z = M r with gamma = (r, z); w = A z with delta = (z, w); one finishing kernel; the
     * fused update with (r, r) for the next iteration's test */
    double *pg = S->parts, *pd = S->parts + kMaxDotParts, *pr = S->parts + 2 * kMaxDotParts;
    unsigned Gg = 0, Gd = 0;
    if (S->M1) {
      ispm_plan *last = S->M2 ? S->M2 : S->M1;
      const void *in = S->r;
      if (S->M2) {
        if ((st = spmv(S->M1, S->T, S->r, S->t, sk))) return st;
        in = S->t;
      }
      st = S->T == ISPM_F64 ? spmv_dot<double>(S, last, in, S->z, S->r, sk, Gg, pg)
                            : spmv_dot<float>(S, last, in, S->z, S->r, sk, Gg, pg);
    } else {
      /* no preconditioner: z = r, gamma = (r, r) */
      if ((st = apply_M(S, S->r, S->z, sk))) return st;
      if (S->T == ISPM_F64)
        hipLaunchKernelGGL((k_dot<double, OP_PQ>), g, b, 0, sm, (const double *)S->r, (const double *)S->z, S->n, S->s, pg, 1);
      else
        hipLaunchKernelGGL((k_dot<float, OP_PQ>), g, b, 0, sm, (const float *)S->r, (const float *)S->z, S->n, S->s, pg, 1);
      ISPM_HIP(hipGetLastError());
      Gg = (unsigned)S->grid;
    }
    if (st) return st;
    st = S->T == ISPM_F64 ? spmv_dot<double>(S, S->A, S->z, S->q, S->z, sk, Gd, pd)
                          : spmv_dot<float>(S, S->A, S->z, S->q, S->z, sk, Gd, pd);
    if (st) return st;
    hipLaunchKernelGGL(k_fin_cg, dim3(1), dim3(kFinThreads), 0, sm, (const double *)pg, Gg,
                       (const double *)pd, Gd, (const double *)pr, (unsigned)S->grid, S->s);
    if (S->T == ISPM_F64)
      hipLaunchKernelGGL(k_cg_update<double>, g, b, 0, sm, (const double *)S->z, (const double *)S->q,
                         (double *)S->p, (double *)S->sv, (double *)S->xs, (double *)S->r, S->n, S->s, pr);
    else
      hipLaunchKernelGGL(k_cg_update<float>, g, b, 0, sm, (const float *)S->z, (const float *)S->q,
                         (float *)S->p, (float *)S->sv, (float *)S->xs, (float *)S->r, S->n, S->s, pr);
    ISPM_HIP(hipGetLastError());
    return ISPM_OK;
  }
  unsigned G = 0;
  /* with a preconditioner, the (r, r) test of the previous iteration's update shares the
   * finishing kernel of (r, z) */
  double *pr = S->parts + 2 * kMaxDotParts;
  if (S->fused_r) {
    /* r2 = r - alpha q and z = M r2 in one launch, with (r2, z) and (r2, r2); alpha is the
     * previous iteration's (0 after a restart, and q is finite) */
    ispm_plan *P = S->M1;
    const split_diag sd{P->dev_diag, P->dev_rmap, P->dev_rmask, P->split_n, P->split_m, P->E};
    const split_update up{S->q, &S->s->alpha, S->r2, pr};
    if ((st = spmv_split(sd, &P->launch, &P->dev, S->T, S->r, S->z, sk, S->r, S->parts, &G, &P->hints, &up)))
      return st;
    if (G == 0) return fail(ISPM_ERR_RUNTIME, "solver: split product without its dot partials");
    hipLaunchKernelGGL(k_fin_rz_rr, dim3(1), dim3(kFinThreads), 0, sm, (const double *)S->parts, G,
                       (const double *)pr, G, S->s);
    std::swap(S->r, S->r2);
  } else if (S->M1) {
    /* z = M2 (M1 r), or M1 r for a one-matrix preconditioner, with (r, z) fused into the
     * last product where possible */
    ispm_plan *last = S->M2 ? S->M2 : S->M1;
    const void *in = S->r;
    if (S->M2) {
      if ((st = spmv(S->M1, S->T, S->r, S->t, sk))) return st;
      in = S->t;
    }
    st = S->T == ISPM_F64 ? spmv_dot<double>(S, last, in, S->z, S->r, sk, G)
                          : spmv_dot<float>(S, last, in, S->z, S->r, sk, G);
    if (st) return st;
    hipLaunchKernelGGL(k_fin_rz_rr, dim3(1), dim3(kFinThreads), 0, sm, (const double *)S->parts, G,
                       (const double *)pr, (unsigned)S->grid, S->s);
  } else {
    if ((st = apply_M(S, S->r, S->z, sk))) return st;
    if ((st = dot(S, OP_RZ, S->r, S->z, 1, S->T))) return st;
  }
  if (S->T == ISPM_F64)
    hipLaunchKernelGGL(k_px<double>, g, b, 0, sm, (double *)S->xs, (double *)S->p, (const double *)S->z, S->n, S->s);
  else
    hipLaunchKernelGGL(k_px<float>, g, b, 0, sm, (float *)S->xs, (float *)S->p, (const float *)S->z, S->n, S->s);
  ISPM_HIP(hipGetLastError());
  /* q = A p with (p, q) fused where possible */
  st = S->T == ISPM_F64 ? spmv_dot<double>(S, S->A, S->p, S->q, S->p, sk, G)
                        : spmv_dot<float>(S, S->A, S->p, S->q, S->p, sk, G);
  if (st) return st;
  finish<OP_PQ>(S, 1, G);
  if (S->fused_r) {
    ISPM_HIP(hipGetLastError());
    return ISPM_OK;
  }
  if (S->T == ISPM_F64)
    hipLaunchKernelGGL(k_r<double>, g, b, 0, sm, (double *)S->r, (const double *)S->q, S->n, S->s,
                       S->M1 ? pr : S->parts);
  else
    hipLaunchKernelGGL(k_r<float>, g, b, 0, sm, (float *)S->r, (const float *)S->q, S->n, S->s,
                       S->M1 ? pr : S->parts);
  if (!S->M1) finish<OP_XR>(S, 1);
  ISPM_HIP(hipGetLastError());
  return ISPM_OK;
}

int run_batch(ispm_solver *S, int k)
{
  if (!S->o.use_graph) {
    for (int i = 0; i < k; i++) {
      int st = iterate(S);
      if (st) return st;
    }
    return ISPM_OK;
  }
  if (!S->exec || S->graph_iters != k) {
    if (S->exec) { (void)hipGraphExecDestroy(S->exec); S->exec = nullptr; }
    if (S->graph) { (void)hipGraphDestroy(S->graph); S->graph = nullptr; }
    hipStream_t sm = stream();
    ISPM_HIP(hipStreamBeginCapture(sm, hipStreamCaptureModeThreadLocal));
    int st = ISPM_OK;
    for (int i = 0; i < k && !st; i++) st = iterate(S);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(sm, &g);
    if (st) { if (g) (void)hipGraphDestroy(g); return st; }
    if (e != hipSuccess) return hip_status(e, "hipStreamEndCapture");
    S->graph = g;
    ISPM_HIP(hipGraphInstantiate(&S->exec, g, nullptr, nullptr, 0));
    S->graph_iters = k;
  }
  ISPM_HIP(hipGraphLaunch(S->exec, stream()));
  return ISPM_OK;
}

int pull(ispm_solver *S)
{
  ISPM_HIP(hipMemcpyAsync(S->hs, S->s, sizeof(scalars), hipMemcpyDeviceToHost, stream()));
  ISPM_HIP(hipStreamSynchronize(stream()));
  return ISPM_OK;
}
int push(ispm_solver *S)
{
  ISPM_HIP(hipMemcpyAsync(S->s, S->hs, sizeof(scalars), hipMemcpyHostToDevice, stream()));
  ISPM_HIP(hipStreamSynchronize(stream()));
  return ISPM_OK;
}

int zero(void *p, size_t bytes) { return ispm_zero(p, bytes); }

/* r64 = b - A x (FP64 true residual), s->r64 = ||r64||^2 */
int true_residual(ispm_solver *S, const double *b, const double *x)
{
  int st = spmv(S->Ares ? S->Ares : S->A, ISPM_F64, x, S->tmp64, nullptr);
  if (st) return st;
  hipLaunchKernelGGL(k_resid, dim3(S->grid), dim3(256), 0, stream(), b, (const double *)S->tmp64, S->r64,
                     S->n, S->parts);
  finish<OP_R64>(S, 0);
  ISPM_HIP(hipGetLastError());
  return ISPM_OK;
}

}  // namespace

extern "C" {

int ispm_solver_create(ispm_solver **out, ispm_plan *A, ispm_plan *M1, ispm_plan *M2,
                       const ispm_solver_opts *o)
{
  if (!out || !A || !o) return fail(ISPM_ERR_INVALID, "solver_create: null argument");
  *out = nullptr;
  int st = ensure_init();
  if (st) return st;
  if (!M1 && M2) return fail(ISPM_ERR_INVALID, "solver_create: second factor without the first");
  if (!A->has_dev) return fail(ISPM_ERR_INVALID, "solver_create: A plan has no device matrix");
  const bool mixed = o->precision == ISPM_PREC_MIXED;
  if (o->method != ISPM_SOLVER_PPCG && o->method != ISPM_SOLVER_PCG && o->method != ISPM_SOLVER_PCG1)
    return fail(ISPM_ERR_INVALID, "solver_create: unknown method %d", o->method);
  if (!mixed && A->E != ISPM_F64) return fail(ISPM_ERR_INVALID, "solver_create: FP64 solve needs an FP64 A plan");
  if (mixed && A->E != ISPM_F32) return fail(ISPM_ERR_INVALID, "solver_create: mixed solve needs an FP32 A plan");
  const int T = mixed ? ISPM_F32 : ISPM_F64;
  for (const ispm_plan *M : {M1, M2}) {
    if (!M) continue;
    if (!M->has_dev || M->T != T)
      return fail(ISPM_ERR_INVALID, "solver_create: preconditioner plans must be on the device with vector type %s", mixed ? "F32" : "F64");
    if (plan_rows(M) != plan_rows(A)) return fail(ISPM_ERR_INVALID, "solver_create: preconditioner size differs from A");
  }
  ispm_solver *S = new ispm_solver;
  S->A = A; S->M1 = M1; S->M2 = M2;
  S->o = *o;
  if (S->o.check_every <= 0) S->o.check_every = 16;
  if (S->o.maxit <= 0) S->o.maxit = 100000;
  if (S->o.reliable_delta <= 0 || S->o.reliable_delta >= 1) S->o.reliable_delta = 0.01;
  /* ISPM_NO_GRAPH=1 launches iterations eagerly (profilers that cannot follow graph capture) */
  if (const char *ng = std::getenv("ISPM_NO_GRAPH"))
    if (ng[0] == '1') S->o.use_graph = 0;
  S->mixed = mixed;
  S->pipelined = o->method == ISPM_SOLVER_PPCG;
  S->single = o->method == ISPM_SOLVER_PCG1;
  S->T = T;
  S->ts = mixed ? 4 : 8;
  S->n = (int)plan_rows(A);
  const size_t vb = S->ts * (size_t)S->n + 64;
  /* one 16-byte vector access per thread where possible (the vector kernels are latency
   * bound with fewer, longer grid-stride loops) */
  {
    const long long per_block = 256LL * (16 / (long long)S->ts);
    S->grid = (int)std::max(1LL, std::min((long long)kMaxBlocks, ((long long)S->n + per_block - 1) / per_block));
  }
  if (A->sym) {
    S->xguard = sym_x_guard(A->symd);
    S->gbytes = (size_t)S->xguard * S->ts;
  }
  void **vecs[] = {&S->r, &S->z, &S->q, &S->p, &S->t, &S->xs};
  for (void **v : vecs) if ((st = alloc_guarded(S, v, vb))) { ispm_solver_destroy(S); return st; }
  if (S->pipelined) {
    void **pv[] = {&S->u, &S->w, &S->m, &S->nv, &S->sv};
    for (void **v : pv) if ((st = alloc_guarded(S, v, vb))) { ispm_solver_destroy(S); return st; }
  }
  if (S->single && (st = alloc_guarded(S, &S->sv, vb))) { ispm_solver_destroy(S); return st; }
  S->fused_r = !S->pipelined && !S->single && M1 && !M2 && M1->split;
  if (S->fused_r) {
    if ((st = alloc_guarded(S, &S->r2, vb))) { ispm_solver_destroy(S); return st; }
    /* iterations in pairs per batch, so that r names the same buffer after every batch */
    S->o.check_every += S->o.check_every & 1;
  }
  if ((st = alloc((void **)&S->r64, 8 * (size_t)S->n + 64)) || (st = alloc((void **)&S->tmp64, 8 * (size_t)S->n + 64)) ||
      (st = alloc((void **)&S->s, sizeof(scalars))) || (st = alloc((void **)&S->parts, sizeof(double) * (2 * (size_t)kMaxDotParts + 4 * (size_t)kMaxBlocks)))) {
    ispm_solver_destroy(S);
    return st;
  }
  hipError_t e = hipHostMalloc((void **)&S->hs, sizeof(scalars), hipHostMallocDefault);
  if (e != hipSuccess) { ispm_solver_destroy(S); return hip_status(e, "hipHostMalloc"); }
  std::memset(S->hs, 0, sizeof(scalars));
  *out = S;
  return ISPM_OK;
}

}  // extern "C"

namespace {

/* Convergence measure of a solve.  L2 (ispm_solver_solve): ||b - A x||_2 / ||b||_2 <= tol.
 * Normalised (ispm_solver_solve_normalised): OpenFOAM's sum |b - A x| / normFactor, with the
 * norm factor of the initial x, below max(tol, rel_tol * initial) (the checkConvergence rule:
 * residual < tolerance, or residual < relTol * initial residual when relTol > 0). */
struct measure {
  bool normalised = false;
  double tol = 1e-10, rel_tol = 0.0;
  /* outputs */
  double initial = 0.0, norm_factor = 0.0;
};

/* the normalised measure of the residual true_residual left in r64 (sum |r| over the rows) */
int l1_residual(ispm_solver *S, int nr, double &l1)
{
  hipLaunchKernelGGL((k_sum1<true>), dim3(S->grid), dim3(256), 0, stream(), (const double *)S->r64, nr, S->parts);
  finish<OP_L1>(S, 0);
  ISPM_HIP(hipGetLastError());
  int st = pull(S);
  l1 = S->hs->l1true;
  return st;
}

int solve_core(ispm_solver *S, const double *b, double *x, measure &mc, ispm_solve_stats *stats)
{
  if (!S || !b || !x) return fail(ISPM_ERR_INVALID, "solver_solve: null argument");
  if ((reinterpret_cast<uintptr_t>(b) | reinterpret_cast<uintptr_t>(x)) & 15)
    return fail(ISPM_ERR_INVALID, "solver_solve: b and x must be 16-byte aligned");
  int st;
  const int n = S->n;
  const size_t vb = S->ts * (size_t)n;
  hipStream_t sm = stream();
  dim3 g(S->grid), bl(256);
  ispm_solve_stats res{};
  /* ||b|| */
  std::memset(S->hs, 0, sizeof(scalars));
  if ((st = push(S))) return st;
  if ((st = launch_dot<double>(S, OP_R64, b, b, 0))) return st;
  if ((st = pull(S))) return st;
  const double bnorm = std::sqrt(S->hs->r64) > 0 ? std::sqrt(S->hs->r64) : 1.0;
  ispm_plan *Ares = S->Ares ? S->Ares : S->A;
  /* rows of the matrix (the vectors are padded beyond them) */
  const int nr = std::min(S->n, Ares->dev.n_cols > 0 ? Ares->dev.n_cols : S->n);
  if (mc.normalised) {
    /* sumA = A 1 (r64 holds the ones vector for it), then the sum of x */
    if (!S->sumA && (st = alloc((void **)&S->sumA, 8 * (size_t)n + 64))) return st;
    hipLaunchKernelGGL(k_ones, g, bl, 0, sm, S->r64, nr, n);
    ISPM_HIP(hipGetLastError());
    if ((st = spmv(Ares, ISPM_F64, S->r64, S->sumA, nullptr))) return st;
    hipLaunchKernelGGL((k_sum1<false>), g, bl, 0, sm, (const double *)x, nr, S->parts);
    finish<OP_SUM>(S, 0);
    ISPM_HIP(hipGetLastError());
  }
  /* r = b - A x */
  if ((st = true_residual(S, b, x))) return st;
  if ((st = pull(S))) return st;
  double rel, target, abs2;   /* measure, its target, and the L2 norm of a residual that meets it */
  if (mc.normalised) {
    hipLaunchKernelGGL(k_normf, g, bl, 0, sm, b, (const double *)S->tmp64, (const double *)S->sumA, nr,
                       (const scalars *)S->s, S->parts);
    finish<OP_NORMF>(S, 0);
    ISPM_HIP(hipGetLastError());
    if ((st = pull(S))) return st;
    mc.norm_factor = S->hs->normf + 1e-20;
    rel = mc.initial = S->hs->l1true / mc.norm_factor;
    target = std::max(mc.tol, mc.rel_tol > 0 ? mc.rel_tol * mc.initial : 0.0);
    /* the recursion tests ||r||_2: scale the target by the initial residual's L2/L1 ratio */
    const double l2 = std::sqrt(S->hs->r64), l1 = S->hs->l1true;
    abs2 = target * mc.norm_factor * (l1 > 0 ? l2 / l1 : 1.0 / std::sqrt((double)nr));
  } else {
    rel = std::sqrt(S->hs->r64) / bnorm;
    target = mc.tol;
    abs2 = mc.tol * bnorm;
  }
  /* an exact initial x has nothing left to reduce, whatever the targets */
  auto met = [&](double v) { return v == 0.0 || (mc.normalised ? v < target : v <= target); };
  const double margin = 0.9;
  for (void *v : {S->z, S->q, S->p, S->sv}) if (v && (st = zero(v, vb))) return st;

  auto setup_recurrence = [&](double tol2) -> int {
    int e;
    scalars &h = *S->hs;
    if (S->mixed && S->pipelined) {
      /* r = r64 / ||r64||; u = M r; w = A u; gamma = (r, u); delta = (w, u); the
       * recurrences restart (z = q = p = s = 0) */
      hipLaunchKernelGGL(k_cast_scaled<float>, g, bl, 0, sm, (float *)S->r, (const double *)S->r64, n, S->s);
      ISPM_HIP(hipGetLastError());
      if ((e = apply_M(S, S->r, S->u, nullptr))) return e;
      if ((e = spmv(S->A, ISPM_F32, S->u, S->w, nullptr, S->xguard))) return e;
      if ((e = launch_dot<float>(S, OP_GAMMA, S->r, S->u, 0))) return e;
      if ((e = launch_dot<float>(S, OP_DELTA, S->w, S->u, 0))) return e;
      for (void *v : {S->z, S->q, S->p, S->sv, S->xs}) if ((e = zero(v, vb))) return e;
      const double r64 = h.r64;
      if ((e = pull(S))) return e;
      h.rscale = std::sqrt(r64);
      h.gamma_old = INFINITY;
      h.alpha_old = 1.0;
      h.done = 0;
      h.tol2 = tol2;
      return push(S);
    }
    if (S->mixed) {
      /* r = r64 / ||r64||; rho restarts the Krylov sequence (p = 0) */
      hipLaunchKernelGGL(k_cast_scaled<float>, g, bl, 0, sm, (float *)S->r, (const double *)S->r64, n, S->s);
      ISPM_HIP(hipGetLastError());
      /* the kernel's rscale, known here without a read-back (the r64 just pulled; sqrt is
       * correctly rounded on both sides) */
      h.rscale = std::sqrt(h.r64);
      h.rho = INFINITY;
      h.gamma_old = INFINITY;
      h.alpha_old = 1.0;
      h.alpha = 0.0;   /* no pending x update (p = 0) */
      h.rr_valid = 0;
      h.done = 0;
      h.tol2 = tol2;
      if ((e = zero(S->xs, vb)) || (e = zero(S->p, vb))) return e;
      if (S->single && (e = zero(S->sv, vb))) return e;
      return push(S);
    }
    /* FP64: r = r64 */
    ISPM_HIP(hipMemcpyAsync(S->r, S->r64, 8 * (size_t)n, hipMemcpyDeviceToDevice, sm));
    h.done = 0;
    h.tol2 = tol2;
    if (S->pipelined) {
      /* u = M r; w = A u; gamma = (r,u); delta = (w,u); recurrences restart */
      if ((e = apply_M(S, S->r, S->u, nullptr))) return e;
      if ((e = spmv(S->A, ISPM_F64, S->u, S->w, nullptr))) return e;
      if ((e = launch_dot<double>(S, OP_GAMMA, S->r, S->u, 0))) return e;
      if ((e = launch_dot<double>(S, OP_DELTA, S->w, S->u, 0))) return e;
      for (void *v : {S->z, S->q, S->p, S->sv}) if ((e = zero(v, vb))) return e;
      if ((e = pull(S))) return e;
      h.gamma_old = INFINITY;
      h.alpha_old = 1.0;
      h.done = 0;
      h.tol2 = tol2;
      return push(S);
    }
    h.rho = INFINITY;
    h.gamma_old = INFINITY;
    h.alpha_old = 1.0;
    h.alpha = 0.0;
    h.rr_valid = 0;
    if ((e = zero(S->p, vb))) return e;
    if (S->single && (e = zero(S->sv, vb))) return e;
    return push(S);
  };

  if (met(rel)) {
    res.converged = 1;
    res.relres = rel;
    if (stats) *stats = res;
    return ISPM_OK;
  }
  /* FP64 variants iterate on x directly (xs aliases x); mixed accumulates in xs */
  void *xs_saved = S->xs;
  if (!S->mixed) S->xs = x;
  if (S->exec && !S->mixed && S->graph_x != x) {
    /* graph was captured with another x: rebuild */
    (void)hipGraphExecDestroy(S->exec); S->exec = nullptr;
    (void)hipGraphDestroy(S->graph); S->graph = nullptr;
  }
  S->graph_x = x;
  double rec_target;   /* target of the recursive residual, relative to its starting norm */
  if (S->mixed) {
    double rmax = std::sqrt(S->hs->r64);
    rec_target = std::max(S->o.reliable_delta, margin * abs2 / rmax);
    st = setup_recurrence(rec_target * rec_target);
  } else {
    double t = margin * abs2;
    rec_target = t;
    st = setup_recurrence(t * t);
  }
  if (st) { S->xs = xs_saved; return st; }
  int guard = 0;
  while (true) {
    if ((st = run_batch(S, S->o.check_every))) break;
    if ((st = pull(S))) break;
    scalars &h = *S->hs;
    res.iterations = (int)h.iters;
    res.l1_resnorm = h.l1;
    const bool out_of_iters = (int)h.iters >= S->o.maxit;
    if (!h.done && !out_of_iters) continue;
    if (!S->pipelined && !S->single) {
      /* the x update k_px defers to the next iteration; stopped by maxit, the update it
       * applies here belongs to a residual not yet tested, and counts as an iteration (x
       * then is the iterate after res.iterations updates, as when the solve converges) */
      if (!h.done && h.alpha != 0.0) res.iterations = (int)h.iters + 1;
      if (S->T == ISPM_F64)
        hipLaunchKernelGGL(k_xflush<double>, g, bl, 0, sm, (double *)S->xs, (const double *)S->p, n, S->s);
      else
        hipLaunchKernelGGL(k_xflush<float>, g, bl, 0, sm, (float *)S->xs, (const float *)S->p, n, S->s);
      ISPM_HIP(hipGetLastError());
    }
    if (S->mixed) {
      hipLaunchKernelGGL(k_x64, g, bl, 0, sm, x, (float *)S->xs, n, S->s);
      ISPM_HIP(hipGetLastError());
    }
    if ((st = true_residual(S, b, x)) || (st = pull(S))) break;
    if (mc.normalised) {
      double l1 = 0.0;
      if ((st = l1_residual(S, nr, l1))) break;
      rel = l1 / mc.norm_factor;
    } else {
      rel = std::sqrt(S->hs->r64) / bnorm;
    }
    res.relres = rel;
    if (met(rel)) { res.converged = 1; break; }
    if (out_of_iters || ++guard > 10000) break;
    res.replacements++;
    if (S->mixed && S->pipelined) {
      /* restart the recurrences from the replaced residual, at its own scale */
      rec_target = std::max(S->o.reliable_delta, margin * abs2 / std::sqrt(S->hs->r64));
      if ((st = setup_recurrence(rec_target * rec_target))) break;
    } else if (S->mixed) {
      const double rmax_old = S->hs->rscale, rmax = std::sqrt(S->hs->r64);
      /* keep the search direction: rescale p and rho to the new residual scale */
      const double c = rmax_old / rmax;
      hipLaunchKernelGGL(k_cast_scaled<float>, g, bl, 0, sm, (float *)S->r, (const double *)S->r64, n, S->s);
      hipLaunchKernelGGL(k_scale<float>, g, bl, 0, sm, (float *)S->p, n, c);
      /* single-reduction PCG: s = A p is rescaled with p */
      if (S->single) hipLaunchKernelGGL(k_scale<float>, g, bl, 0, sm, (float *)S->sv, n, c);
      ISPM_HIP(hipGetLastError());
      h.rscale = rmax;   /* what k_cast_scaled stored (sqrt of the pulled r64) */
      rec_target = std::max(S->o.reliable_delta, margin * abs2 / rmax);
      h.rho = h.rho * c * c;
      h.gamma_old = h.gamma_old * c * c;
      h.alpha = 0.0;   /* the pending x update was applied above: the next k_px adds nothing */
      h.rr_valid = 0;
      h.done = 0;
      h.tol2 = rec_target * rec_target;
      if ((st = push(S))) break;
    } else {
      /* the recursive residual drifted from the true one: replace it and tighten */
      rec_target *= 0.5;
      st = setup_recurrence(rec_target * rec_target);
      if (st) break;
    }
  }
  S->xs = xs_saved;
  if (stats) *stats = res;
  if (st) return st;
  return res.converged ? ISPM_OK : fail(ISPM_ERR_NOT_CONVERGED, "solver: relres %.3e after %d iterations", res.relres, res.iterations);
}

}  // namespace

extern "C" {

int ispm_solver_solve(ispm_solver *S, const double *b, double *x, ispm_solve_stats *stats)
{
  measure mc;
  if (S) mc.tol = S->o.tol > 0 ? S->o.tol : 1e-10;
  return solve_core(S, b, x, mc, stats);
}

int ispm_solver_solve_normalised(ispm_solver *S, const double *b, double *x, double tolerance, double rel_tol,
                                 ispm_norm_stats *stats)
{
  if (!(tolerance >= 0) || !(rel_tol >= 0) || (tolerance == 0 && rel_tol == 0))
    return fail(ISPM_ERR_INVALID, "solver_solve_normalised: need tolerance >= 0, rel_tol >= 0, not both 0");
  measure mc;
  mc.normalised = true;
  mc.tol = tolerance;
  mc.rel_tol = rel_tol;
  ispm_solve_stats res{};
  const int st = solve_core(S, b, x, mc, &res);
  if (stats) {
    stats->iterations = res.iterations;
    stats->replacements = res.replacements;
    stats->converged = res.converged;
    stats->initial_residual = mc.initial;
    stats->final_residual = res.iterations || res.converged ? res.relres : mc.initial;
    stats->norm_factor = mc.norm_factor;
    stats->l1_resnorm = res.l1_resnorm;
  }
  return st;
}

int ispm_solver_set_residual_plan(ispm_solver *S, ispm_plan *A64)
{
  if (!S) return fail(ISPM_ERR_INVALID, "solver_set_residual_plan: null solver");
  if (A64 && (!A64->has_dev || A64->E != ISPM_F64 || (int)plan_rows(A64) != S->n))
    return fail(ISPM_ERR_INVALID, "solver_set_residual_plan: need an FP64 device plan of the same size");
  S->Ares = A64;
  return ISPM_OK;
}

void ispm_solver_destroy(ispm_solver *S)
{
  if (!S) return;
  if (S->exec) (void)hipGraphExecDestroy(S->exec);
  if (S->graph) (void)hipGraphDestroy(S->graph);
  for (void *v : {S->r, S->r2, S->u, S->w, S->m, S->nv, S->z, S->q, S->sv, S->p, S->t, S->xs}) free_guarded(S, v);
  ispm_free(S->r64);
  ispm_free(S->tmp64);
  ispm_free(S->sumA);
  ispm_free(S->s);
  ispm_free(S->parts);
  if (S->hs) (void)hipHostFree(S->hs);
  delete S;
}

}  // extern "C"
