#ifndef ISPM_SPMV_KERNEL_H
#define ISPM_SPMV_KERNEL_H
/* spmv_kernel.h -- sliced-ELLPACK ("SlELL") SpMV / SpMM for gfx950.
 *
 * Layout (formats/formats.h:39-102, conversion.h:338-414 of the reference): slice s covers
 * rows [y0, y0+sz) (sz a power of two), its elements stored column-major, element (r,k) at
 * e0 + r + k*sz.  Slice records in slice_ptr have P = 1+[W>1]+[V]+2[D] words:
 * [elm_off, (row_start if V), (col_off if W>1 or D), (diag_off if D)].
 *   plain ELL : nell planes of (col, val).
 *   diagonal  : nell ELL planes, then ndiag dense planes whose column is
 *               diag_pos[d] + r - sz; columns outside [0, n_cols) read x as 0.
 *   hblock W=2: nell planes of singles, then block planes holding a pair index c/2 and two
 *               value planes (even, odd column).
 *
 * Mapping (one workgroup per slice, like cuda/spmv/kernel.h:1-135, re-designed for wave64):
 *   thread t owns rows row..row+H-1 of the slice with row = (t*H) mod R and walks planes
 *   k = t*H/R, +kstep, ..., where R = min(N*H, sz) and kstep = N*H/R.  Every plane load of a
 *   wave is one contiguous run of 64*H elements (fully coalesced), diagonal x runs are
 *   contiguous too.  When several threads share a row (kstep > 1) their partial sums are
 *   combined by xor-shuffles inside the wave and through LDS across waves (replacing the
 *   shared-memory halving of cuda/reduce.h:18-127).  When kstep == 1 the plane index is
 *   wave-uniform, so diagonal offsets come through scalar loads.
 *   Workgroups are remapped so each XCD (8 per MI355X, round-robin workgroup dispatch)
 *   processes one contiguous range of slices: neighbouring slices share x windows, which
 *   then stay in that XCD's L2.
 *
 * The summation order inside a row differs from CSR order; FP64 results agree with an
 * FP64 CSR product to rounding (|dy| <= 1e-13 sum_k |a_rk x_k| in the tests).
 */
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdlib>
#include <type_traits>
#include "device.h"
#include "internal.h"
#include "diag_pattern.h"

namespace ispm {
namespace spmvk {

using dev::bf16_t;
using dev::elt;

struct spmv_args {
  const unsigned *sp;
  const unsigned *cols;
  const void *elms;
  const void *x;
  void *y;
  size_t ldx, ldy;       /* SpMM: column strides in elements */
  unsigned n_slices;
  unsigned log2S;        /* fixed-height slices */
  unsigned n_cols;
  unsigned per_xcd;      /* workgroups per XCD range */
  unsigned spb;          /* consecutive slices each workgroup processes */
  const unsigned *skip;  /* if non-null and *skip != 0 the launch does nothing (solver done flag) */
  const void *dotv = nullptr;   /* fused dot: workgroup w stores sum_rows dotv[row]*y[row] */
  double *dparts = nullptr;     /* at dparts[w] (kernels that support it) */
  /* streamed mapping only (split-diagonal plans, whose layout holds the rows with
   * off-diagonal entries): layout row i is matrix row r = rmap[i] (~0u: none), and
   * y[r] = rdiag[r] * x[r] + the row's sum */
  const unsigned *rmap = nullptr;
  const void *rdiag = nullptr;
  /* split-diagonal products with the CG residual update in front (AX kernels): the product
   * multiplies x - alpha q instead of x, q = axq, alpha = *axa */
  const void *axq = nullptr;
  const double *axa = nullptr;
};

/* x - alpha q, one rounding (the CG residual update r - alpha q fused into the next product:
 * every role that needs an entry computes it the same way) */
__device__ __forceinline__ float ax_update(float x, float q, float a) { return __builtin_fmaf(-a, q, x); }
__device__ __forceinline__ double ax_update(double x, double q, double a) { return __builtin_fma(-a, q, x); }



template<typename T> __device__ __forceinline__ T xload(const T *x, unsigned i, unsigned n)
{
  T v = x[i < n ? i : n - 1];
  return i < n ? v : (T)0;
}

/* entry i of x (bounds-checked), or of x - alpha q when AX */
template<typename T, bool AX> struct xsrc {
  const T *x, *q;
  T a;
  unsigned n;
  __device__ __forceinline__ T at(unsigned i) const
  {
    const T v = xload(x, i, n);
    return AX ? ax_update(v, xload(q, i, n), a) : v;
  }
};

template<typename T, int H> struct vstore;
template<> struct vstore<double, 2> {
  static __device__ __forceinline__ void st(double *y, const double (&v)[2])
  { *reinterpret_cast<double2 *>(y) = make_double2(v[0], v[1]); }
};
template<> struct vstore<float, 2> {
  static __device__ __forceinline__ void st(float *y, const float (&v)[2])
  { *reinterpret_cast<float2 *>(y) = make_float2(v[0], v[1]); }
};
template<> struct vstore<float, 4> {
  static __device__ __forceinline__ void st(float *y, const float (&v)[4])
  { *reinterpret_cast<float4 *>(y) = make_float4(v[0], v[1], v[2], v[3]); }
};
template<> struct vstore<double, 4> {
  static __device__ __forceinline__ void st(double *y, const double (&v)[4])
  {
    reinterpret_cast<double2 *>(y)[0] = make_double2(v[0], v[1]);
    reinterpret_cast<double2 *>(y)[1] = make_double2(v[2], v[3]);
  }
};
template<typename T> struct vstore<T, 1> {
  static __device__ __forceinline__ void st(T *y, const T (&v)[1]) { *y = v[0]; }
};

template<typename T, typename E, int H, int W, bool V, bool D, int K>
__global__ void __launch_bounds__(1024) slell_kernel(spmv_args a)
{
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const unsigned N = blockDim.x;
  const unsigned t = threadIdx.x;
  const unsigned b = blockIdx.x;
  const unsigned group = (b & 7u) * a.per_xcd + (b >> 3);
  if (a.skip && *a.skip) return;
  for (unsigned it = 0; it < a.spb; it++) {
  const unsigned slice = group * a.spb + it;
  if (slice >= a.n_slices) return;
  constexpr unsigned P = 1 + (W > 1) + (V ? 1 : 0) + (D ? 2 : 0);
  const unsigned *rec = a.sp + P * slice;
  const unsigned e0 = rec[0], e1 = rec[P];
  unsigned y0, log2sz;
  if (V) {
    y0 = rec[1];
    log2sz = 31 - __clz(rec[1 + P] - y0);
  } else {
    y0 = slice << a.log2S;
    log2sz = a.log2S;
  }
  const unsigned sz = 1u << log2sz;
  unsigned c0, nell, nsec = 0, d0 = 0;
  if (D) {
    c0 = rec[1 + V];
    d0 = rec[2 + V];
    const unsigned c1 = rec[1 + V + P];
    nell = (c1 - c0) >> log2sz;
    nsec = ((e1 - e0) >> log2sz) - nell;
  } else if (W > 1) {
    c0 = rec[1 + V];
    const unsigned c1 = rec[1 + V + P];
    const unsigned nse = (W * (c1 - c0) - (e1 - e0)) / (W - 1);
    nell = nse >> log2sz;
    nsec = ((c1 - c0) >> log2sz) - nell;
  } else {
    c0 = e0;
    nell = (e1 - e0) >> log2sz;
  }

  const E *__restrict__ el = static_cast<const E *>(a.elms);
  const unsigned *__restrict__ cl = a.cols;
  const T *__restrict__ xv = static_cast<const T *>(a.x);
  T *__restrict__ yv = static_cast<T *>(a.y);
  const unsigned ncols = a.n_cols;

  const unsigned NH = N * H;
  const unsigned log2NH = 31 - __clz(NH);
  const unsigned log2R = log2NH < log2sz ? log2NH : log2sz;
  const unsigned R = 1u << log2R;
  const unsigned kstep = NH >> log2R;
  const unsigned tH = t * H;
  const unsigned rloc = tH & (R - 1);
  const unsigned kstart = tH >> log2R;
  const unsigned g = R / H;               /* threads per plane column */
  const unsigned passes = sz >> log2R;
  const unsigned ktot = nell + nsec;

  for (unsigned pass = 0; pass < passes; pass++) {
    const unsigned row = pass * R + rloc;
    T acc[K][H];
#pragma unroll
    for (int j = 0; j < K; j++)
#pragma unroll
      for (int h = 0; h < H; h++) acc[j][h] = 0;

    auto ell_plane = [&](unsigned k) {
      const unsigned off = (k << log2sz) + row;
      T e[H];
      unsigned c[H];
      if (H == 2) {
        elt<T, E>::get2(el, e0 + off, e[0], e[H - 1]);
        dev::col_load2(cl + c0 + off, c[0], c[H - 1]);
      } else {
        e[0] = elt<T, E>::get(el, e0 + off);
        c[0] = dev::col_load(cl + c0 + off);
      }
#pragma unroll
      for (int j = 0; j < K; j++)
#pragma unroll
        for (int h = 0; h < H; h++) acc[j][h] += e[h] * xload(xv + j * a.ldx, c[h], ncols);
    };
    auto sec_plane = [&](unsigned d) {
      if (D) {
        const unsigned off = ((nell + d) << log2sz) + row;
        T e[H];
        if (H == 2) elt<T, E>::get2(el, e0 + off, e[0], e[H - 1]);
        else e[0] = elt<T, E>::get(el, e0 + off);
        const unsigned xi = a.sp[d0 + d] + row - sz;
#pragma unroll
        for (int j = 0; j < K; j++)
#pragma unroll
          for (int h = 0; h < H; h++) acc[j][h] += e[h] * xload(xv + j * a.ldx, xi + h, ncols);
      } else if (W > 1) {
        const unsigned boff = (nell << log2sz);
        const unsigned coff = ((nell + d) << log2sz) + row;
        unsigned c[H];
        if (H == 2) dev::col_load2(cl + c0 + coff, c[0], c[H - 1]);
        else c[0] = dev::col_load(cl + c0 + coff);
        T ev[H], od[H];
        const unsigned eo = e0 + boff + ((2 * d) << log2sz) + row;
        if (H == 2) {
          elt<T, E>::get2(el, eo, ev[0], ev[H - 1]);
          elt<T, E>::get2(el, eo + sz, od[0], od[H - 1]);
        } else {
          ev[0] = elt<T, E>::get(el, eo);
          od[0] = elt<T, E>::get(el, eo + sz);
        }
#pragma unroll
        for (int j = 0; j < K; j++)
#pragma unroll
          for (int h = 0; h < H; h++) {
            const unsigned xc = 2 * c[h];
            acc[j][h] += ev[h] * xload(xv + j * a.ldx, xc, ncols)
                       + od[h] * xload(xv + j * a.ldx, xc + 1, ncols);
          }
      }
    };

    if (kstep == 1) {
      /* every thread walks all planes: the plane index is wave-uniform */
#pragma unroll 4
      for (unsigned k = 0; k < nell; k++) ell_plane(k);
#pragma unroll 8
      for (unsigned d = 0; d < nsec; d++) sec_plane(d);
    } else {
      unsigned k = kstart;
#pragma unroll 2
      for (; k < nell; k += kstep) ell_plane(k);
#pragma unroll 4
      for (; k < ktot; k += kstep) sec_plane(k - nell);

      /* combine the kstep partial sums of each row */
      for (unsigned off = g; off < 64 && off < N; off <<= 1)
#pragma unroll
        for (int j = 0; j < K; j++)
#pragma unroll
          for (int h = 0; h < H; h++) acc[j][h] += __shfl_xor(acc[j][h], (int)off, 64);
      const unsigned G = g > 64 ? g : 64;
      if (N > G) {
        T *lds = reinterpret_cast<T *>(lds_raw);
#pragma unroll
        for (int j = 0; j < K; j++)
#pragma unroll
          for (int h = 0; h < H; h++) lds[(j * H + h) * N + t] = acc[j][h];
        __syncthreads();
        if (t < g) {
          const unsigned parts = N / G;
#pragma unroll
          for (int j = 0; j < K; j++)
#pragma unroll
            for (int h = 0; h < H; h++) {
              T s = 0;
              for (unsigned m = 0; m < parts; m++) s += lds[(j * H + h) * N + t + m * G];
              acc[j][h] = s;
            }
        }
        __syncthreads();   /* the next slice reuses the LDS buffer */
      }
    }
    if (t < g) {
#pragma unroll
      for (int j = 0; j < K; j++) vstore<T, H>::st(yv + j * a.ldy + y0 + pass * R + t * H, acc[j]);
    }
  }
  }
}


/* Diagonal slices with H rows per thread (fixed S, one vector) of a plan whose slices mostly
 * share one diagonal pattern (diag_pattern.h).  A slice marked in `uni` has no ELL planes
 * and exactly the pattern's diagonals: the loads of its x windows (16-byte vectors of their
 * aligned covers, addresses known from the slice index alone) are issued first, then the
 * loads of all its element planes; the windows go to LDS as soon as they arrive, and every
 * plane reads its x from LDS.  Other slices take the generic path, reading positions and x
 * directly.  Both paths add the ELL
 * planes, then the diagonal planes in storage order, as slell_kernel does, so the results
 * are identical to it. */
template<typename T> using vecx = T __attribute__((ext_vector_type(16 / sizeof(T))));

/* x window runs a pattern kernel instance keeps in registers (bound by its plane count) */
constexpr unsigned pattern_runs_in_regs(unsigned NPM) { return NPM <= 8 ? 8u : NPM <= 27 ? 9u : 16u; }
/* x-window vectors per thread the pattern kernel stages in registers */
constexpr unsigned pattern_window_vectors(unsigned NPM) { return NPM <= 27 ? 5u : 8u; }

/* The pattern restated for x loads of VX elements: each run's window [rrel, rrel + rlen)
 * widened to its VX-aligned cover, covers packed at VX-aligned LDS offsets, poff moved with
 * them; the kernel's run bound NR and its NW vectors per thread (of NT threads) decide
 * vec_runs.  Slice first rows are multiples of the slice height, so the covers are aligned in
 * every slice. */
inline void cover_pattern(const diag_pattern &p, unsigned VX, unsigned NT, unsigned NR, unsigned NW, diag_pattern &c)
{
  c = p;
  unsigned used = 0;
  long long lo = 0, hi = 0;
  for (unsigned r = 0; r < p.nrun; r++) {
    const long long rel = p.rrel[r];
    const long long a = (rel >= 0 ? rel / VX : -((-rel + VX - 1) / VX)) * VX;
    const unsigned len = (unsigned)(((rel - a) + p.rlen[r] + VX - 1) / VX * VX);
    c.rrel[r] = (int)a;
    c.rlen[r] = len;
    c.rbase[r] = used;
    used += len;
    lo = r == 0 || a < lo ? a : lo;
    hi = r == 0 || a + len > hi ? a + len : hi;
  }
  for (unsigned j = 0; j < p.n; j++)
    for (unsigned r = 0; r < p.nrun; r++)
      if (p.poff[j] >= p.rbase[r] && p.poff[j] < p.rbase[r] + p.rlen[r]) {
        c.poff[j] = c.rbase[r] + (p.poff[j] - p.rbase[r]) + (unsigned)(p.rrel[r] - c.rrel[r]);
        break;
      }
  c.lds_len = used;
  c.cover_lo = lo;
  c.cover_hi = hi;
  /* the covers as one vector list, spread over the NT threads at up to NW vectors each */
  c.wtotal = used / VX;
  for (unsigned r = 0; r < (unsigned)kPatPlanes; r++) {
    const unsigned rr = r < p.nrun ? r : 0u;
    c.wvbase[r] = r < p.nrun ? c.rbase[r] / VX : ~0u;
    c.wdelta[r] = c.rrel[rr] - (int)c.rbase[rr];
  }
  c.vec_runs = p.nrun <= NR && c.wtotal > 0 && c.wtotal <= NW * NT ? 1u : 0u;
}

/* H consecutive elements (aligned to H) of one plane, converted to T */
template<typename T, typename E, int H> struct rows_load;
template<typename T, typename E> struct rows_load<T, E, 1> {
  static __device__ __forceinline__ void get(const E *p, unsigned i, T (&v)[1]) { v[0] = elt<T, E>::get(p, i); }
};
template<typename T, typename E> struct rows_load<T, E, 2> {
  static __device__ __forceinline__ void get(const E *p, unsigned i, T (&v)[2]) { elt<T, E>::get2(p, i, v[0], v[1]); }
};
template<typename T, typename E> struct rows_load<T, E, 4> {
  static __device__ __forceinline__ void get(const E *p, unsigned i, T (&v)[4])
  {
    if constexpr (sizeof(E) == 8) {
      elt<T, E>::get2(p, i, v[0], v[1]);
      elt<T, E>::get2(p, i + 2, v[2], v[3]);
    } else {
      dev::elt_vec<T>(p, i, v);
    }
  }
};

template<typename T, typename E, unsigned LOG2S, int H, unsigned NPM>
__global__ void __launch_bounds__((1u << LOG2S) / H)
slell_diagpat_kernel(spmv_args a, diag_pattern pat, const unsigned *__restrict__ uni)
{
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *win = reinterpret_cast<T *>(lds_raw);
  constexpr unsigned S = 1u << LOG2S, NT = S / H, VX = 16 / sizeof(T), NR = pattern_runs_in_regs(NPM);
  const unsigned t = threadIdx.x, r0 = t * H;   /* this thread's first row in the slice */
  /* one slice per workgroup */
  const unsigned slice = (blockIdx.x & 7u) * a.per_xcd + (blockIdx.x >> 3);
  if (a.skip && *a.skip) return;
  if (slice >= a.n_slices) {
    if (a.dparts && threadIdx.x == 0) a.dparts[blockIdx.x] = 0.0;
    return;
  }
  const E *__restrict__ el = static_cast<const E *>(a.elms);
  const unsigned *__restrict__ sp = a.sp;
  const T *__restrict__ xv = static_cast<const T *>(a.x);
  T *__restrict__ yv = static_cast<T *>(a.y);
  const unsigned ncols = a.n_cols;
  const bool xvec = (reinterpret_cast<uintptr_t>(xv) % 16) == 0;
  const unsigned row0 = slice << LOG2S;
  const unsigned e0 = sp[3 * slice];
  T acc[H];
#pragma unroll
  for (int h = 0; h < H; h++) acc[h] = 0;
  const bool on_pattern = (uni[slice >> 5] >> (slice & 31u)) & 1u;
  if (on_pattern) {
    const unsigned np = pat.n, nrun = pat.nrun;
    /* x windows, staged with the rest of their 16-byte aligned covers (rrel, rlen, rbase and
     * poff describe the covers).  When every cover lies inside x and needs at most one vector
     * per thread, the window loads are issued first, before the element loads, so staging
     * them waits for them alone; otherwise a bounds-checked loop stages them after. */
    const bool fast = xvec && pat.vec_runs && (long long)row0 + pat.cover_lo >= 0 &&
                      (long long)row0 + pat.cover_hi <= (long long)ncols;
    constexpr unsigned NW = pattern_window_vectors(NPM);
    vecx<T> w[NW];
    if (fast) {
      /* the window vectors spread evenly over the threads, a fixed NW loads each (lanes past
       * the list's end reload its last vector: one request per wave), all issued before the
       * element loads so that staging them waits for them alone */
#pragma unroll
      for (unsigned q = 0; q < NW; q++) {
        const unsigned v0 = t + q * NT, v = v0 < pat.wtotal ? v0 : pat.wtotal - 1;
        int dl = pat.wdelta[0];
#pragma unroll
        for (unsigned r = 1; r < NR; r++)
          if (r < nrun && v >= pat.wvbase[r]) dl = pat.wdelta[r];
        w[q] = *reinterpret_cast<const vecx<T> *>(xv + (unsigned)((int)row0 + dl + (int)(v * VX)));
      }
    }
    /* every plane requested at once; planes past the pattern's count read one address that
     * is the same in all lanes (a single request per wave), keeping the number of loads in
     * flight fixed so each wait above and below is for exactly the loads it needs */
    T ev[NPM][H];
#pragma unroll
    for (unsigned j = 0; j < NPM; j++)
      rows_load<T, E, H>::get(el + e0 + (j < np ? r0 : 0u), j < np ? (j << LOG2S) : 0u, ev[j]);
    if (fast) {
#pragma unroll
      for (unsigned q = 0; q < NW; q++)
        if (t + q * NT < pat.wtotal) *reinterpret_cast<vecx<T> *>(win + (t + q * NT) * VX) = w[q];
    } else {
      for (unsigned r = 0; r < nrun; r++) {
        const unsigned x0 = row0 + (unsigned)pat.rrel[r], L = pat.rlen[r], b = pat.rbase[r];
        for (unsigned i = t; i < L; i += NT) win[b + i] = xload(xv, x0 + i, ncols);
      }
    }
    __syncthreads();
#pragma unroll
    for (unsigned j = 0; j < NPM; j++)
      if (j < np)
#pragma unroll
        for (int h = 0; h < H; h++) acc[h] += ev[j][h] * win[pat.poff[j] + r0 + h];
  } else {
    const unsigned *__restrict__ cl = a.cols;
    const unsigned c0 = sp[3 * slice + 1], d0 = sp[3 * slice + 2];
    const unsigned e1 = sp[3 * slice + 3], c1 = sp[3 * slice + 4];
    const unsigned nell = (c1 - c0) >> LOG2S;
    const unsigned nsec = ((e1 - e0) >> LOG2S) - nell;
#pragma unroll 4
    for (unsigned k = 0; k < nell; k++)
#pragma unroll
      for (int h = 0; h < H; h++) {
        const unsigned off = (k << LOG2S) + r0 + h;
        acc[h] += elt<T, E>::get(el, e0 + off) * xload(xv, dev::col_load(cl + c0 + off), ncols);
      }
#pragma unroll 8
    for (unsigned d = 0; d < nsec; d++) {
      const unsigned pos = sp[d0 + d];
#pragma unroll
      for (int h = 0; h < H; h++) {
        const unsigned off = ((nell + d) << LOG2S) + r0 + h;
        acc[h] += elt<T, E>::get(el, e0 + off) * xload(xv, pos + r0 + h - S, ncols);
      }
    }
  }
  vstore<T, H>::st(yv + row0 + r0, acc);
  if (a.dparts) {
    __shared__ double red[NT / 64];
    const T *__restrict__ dv = static_cast<const T *>(a.dotv);
    /* CG's (p, A p) has x itself as the dot vector: on a pattern slice its rows are in the
     * main diagonal's window */
    const bool dx = dv == xv && on_pattern && pat.diag_plane1;
    const unsigned o = dx ? pat.poff[pat.diag_plane1 - 1] + r0 : 0u;
    double d[1] = {0.0};
#pragma unroll
    for (int h = 0; h < H; h++) d[0] += (double)(dx ? win[o + h] : dv[row0 + r0 + h]) * (double)acc[h];
    dev::block_sum<NT, 1>(d, red);
    if (t == 0) a.dparts[blockIdx.x] = d[0];
  }
}

/* the (slice height, rows per thread, plane bound) instance of the pattern kernel */
template<typename T, typename E, unsigned NPM>
int diagpat_launch(const ispm_slell_desc *A, unsigned H, dim3 g, dim3 b, size_t lds, const spmv_args &a,
                   const diag_pattern &pat, const unsigned *uni)
{
  hipStream_t st = stream();
  switch (A->slice_height * 4 + H) {
    case 64 * 4 + 1: hipLaunchKernelGGL((slell_diagpat_kernel<T, E, 6, 1, NPM>), g, b, lds, st, a, pat, uni); break;
    case 128 * 4 + 1: hipLaunchKernelGGL((slell_diagpat_kernel<T, E, 7, 1, NPM>), g, b, lds, st, a, pat, uni); break;
    case 256 * 4 + 1: hipLaunchKernelGGL((slell_diagpat_kernel<T, E, 8, 1, NPM>), g, b, lds, st, a, pat, uni); break;
    case 128 * 4 + 2: hipLaunchKernelGGL((slell_diagpat_kernel<T, E, 7, 2, NPM>), g, b, lds, st, a, pat, uni); break;
    case 256 * 4 + 2: hipLaunchKernelGGL((slell_diagpat_kernel<T, E, 8, 2, NPM>), g, b, lds, st, a, pat, uni); break;
    case 256 * 4 + 4: hipLaunchKernelGGL((slell_diagpat_kernel<T, E, 8, 4, NPM>), g, b, lds, st, a, pat, uni); break;
    case 512 * 4 + 1: hipLaunchKernelGGL((slell_diagpat_kernel<T, E, 9, 1, NPM>), g, b, lds, st, a, pat, uni); break;
    case 512 * 4 + 2: hipLaunchKernelGGL((slell_diagpat_kernel<T, E, 9, 2, NPM>), g, b, lds, st, a, pat, uni); break;
    default: return fail(ISPM_ERR_UNSUPPORTED, "spmv: pattern kernel for slice height %u", A->slice_height);
  }
  ISPM_HIP(hipGetLastError());
  return ISPM_OK;
}

template<typename T, typename E>
int launch_diagpat(const ispm_slell_desc *A, const diag_pattern &pat, const unsigned *uni,
                   const void *x, void *y, const unsigned *skip, const void *dotv = nullptr,
                   double *dparts = nullptr, unsigned *dgrid = nullptr)
{
  /* Rows per thread is the kernel's choice (each row's sum is the same either way): one,
   * measured fastest on a 27-point 192^3 stencil with 256-row slices (FP32: 137-140 us with
   * one row per thread, 146 us with two; FP64: 281-288 us with one, 348 us with two, whose
   * windows then exceed five vectors per thread and take the bounds-checked path). */
  unsigned H = 1;
  if (const char *e = std::getenv("ISPM_DIAGPAT_H")) H = (unsigned)std::atoi(e);
  if ((H != 1 && H != 2 && H != 4) || A->slice_height / H < 64 ||
      (reinterpret_cast<uintptr_t>(y) % (H * sizeof(T))) != 0 ||
      (reinterpret_cast<uintptr_t>(A->elms) % 16) != 0)
    H = 1;
  spmv_args a;
  a.skip = skip;
  a.sp = A->slice_ptr;
  a.cols = A->cols;
  a.elms = A->elms;
  a.x = x;
  a.y = y;
  a.ldx = a.ldy = 0;
  a.n_slices = A->n_slices;
  a.log2S = (unsigned)(31 - __builtin_clz(A->slice_height));
  a.n_cols = (unsigned)A->n_cols;
  a.spb = 1;
  a.per_xcd = (A->n_slices + 7) / 8;
  if (dparts && dgrid && a.per_xcd * 8 <= kMaxDotParts) {
    a.dotv = dotv;
    a.dparts = dparts;
    *dgrid = a.per_xcd * 8;
  }
  const unsigned NPM = pat.n <= 8 ? 8u : pat.n <= 27 ? 27u : (unsigned)kPatPlanes;
  diag_pattern c;
  cover_pattern(pat, 16 / sizeof(T), A->slice_height / H, pattern_runs_in_regs(NPM), pattern_window_vectors(NPM), c);
  const size_t lds = (size_t)c.lds_len * sizeof(T);
  if (lds > 65536) return fail(ISPM_ERR_UNSUPPORTED, "spmv: pattern windows need %zu bytes of LDS", lds);
  const dim3 g(a.per_xcd * 8), b(A->slice_height / H);
  if (NPM == 8) return diagpat_launch<T, E, 8>(A, H, g, b, lds, a, c, uni);
  if (NPM == 27) return diagpat_launch<T, E, 27>(A, H, g, b, lds, a, c, uni);
  return diagpat_launch<T, E, kPatPlanes>(A, H, g, b, lds, a, c, uni);
}

/* Multi-vector form of the pattern kernel (the block solver's products, k = 2..8 vectors,
 * column j at x + j ldx): one 256-row slice per workgroup, one row per thread.  A slice on
 * the pattern loads its planes once; then, KG vectors at a time, the workgroup stages those
 * vectors' x windows in LDS (16-byte loads, the windows of all KG vectors spread over the
 * threads) and each thread adds its row's products for them, plane by plane in the pattern's
 * order -- the order of the single-vector kernel, so column j equals that kernel's product of
 * x_j bit for bit.  Gathering x per plane instead (as the generic kernels do) asks L2 for
 * every vector's lines once per plane.  Other slices gather per row, ELL planes then
 * diagonal sections, as the single-vector kernel does. */
template<typename T, typename E, unsigned K, unsigned NPM>
__global__ void __launch_bounds__(256)
slell_diagpat_spmm_kernel(spmv_args a, diag_pattern pat, const unsigned *__restrict__ uni, unsigned kg)
{
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *win = reinterpret_cast<T *>(lds_raw);
  constexpr unsigned LOG2S = 8, S = 1u << LOG2S, VX = 16 / sizeof(T);
  const unsigned t = threadIdx.x;
  const unsigned slice = (blockIdx.x & 7u) * a.per_xcd + (blockIdx.x >> 3);
  if (slice >= a.n_slices) return;   /* uniform over the workgroup */
  const E *__restrict__ el = static_cast<const E *>(a.elms);
  const unsigned *__restrict__ sp = a.sp;
  const T *__restrict__ X = static_cast<const T *>(a.x);
  T *__restrict__ Y = static_cast<T *>(a.y);
  const size_t ldx = a.ldx, ldy = a.ldy;
  const unsigned ncols = a.n_cols;
  const unsigned row0 = slice << LOG2S;
  const unsigned e0 = sp[3 * slice];
  const bool on_pattern = (uni[slice >> 5] >> (slice & 31u)) & 1u;
  if (on_pattern) {
    const unsigned np = pat.n, nrun = pat.nrun, L = pat.lds_len, wt = pat.wtotal;
    /* 16-byte window loads when every column is aligned and the covers lie inside x */
    /* this thread's window vectors are t + q S, q < QM, with their offsets in x (the same for
     * every vector) computed once; longer window lists, unaligned columns and covers
     * reaching past x stage through the bounds-checked loop */
    constexpr unsigned QM = 4;
    const bool fast = ((reinterpret_cast<uintptr_t>(X) | (ldx * sizeof(T))) % 16) == 0 && wt <= QM * S &&
                      (long long)row0 + pat.cover_lo >= 0 && (long long)row0 + pat.cover_hi <= (long long)ncols;
    T ev[NPM];
#pragma unroll
    for (unsigned j = 0; j < NPM; j++) ev[j] = j < np ? elt<T, E>::get(el, e0 + (j << LOG2S) + t) : (T)0;
    int dq[QM];
#pragma unroll
    for (unsigned q = 0; q < QM; q++) {
      const unsigned vv = t + q * S;
      int dl = pat.wdelta[0];
      for (unsigned r = 1; r < nrun; r++)
        if (vv >= pat.wvbase[r]) dl = pat.wdelta[r];
      dq[q] = (int)row0 + dl + (int)(vv * VX);
    }
    for (unsigned j0 = 0; j0 < K; j0 += kg) {
      const unsigned nk = min(kg, K - j0);
      if (j0) __syncthreads();   /* the previous vectors' windows are read */
      if (fast) {
        for (unsigned jj = 0; jj < nk; jj++) {
          const T *xj = X + (size_t)(j0 + jj) * ldx;
#pragma unroll
          for (unsigned q = 0; q < QM; q++)
            if (t + q * S < wt)
              *reinterpret_cast<vecx<T> *>(win + jj * L + (t + q * S) * VX) =
                  *reinterpret_cast<const vecx<T> *>(xj + (unsigned)dq[q]);
        }
      } else {
        for (unsigned jj = 0; jj < nk; jj++) {
          const T *xj = X + (size_t)(j0 + jj) * ldx;
          for (unsigned r = 0; r < nrun; r++) {
            const long long x0 = (long long)row0 + pat.rrel[r];
            const unsigned Lr = pat.rlen[r], b = pat.rbase[r];
            for (unsigned i = t; i < Lr; i += S) {
              const long long c = x0 + i;
              win[jj * L + b + i] = (c >= 0 && c < (long long)ncols) ? xj[c] : (T)0;
            }
          }
        }
      }
      __syncthreads();
      for (unsigned jj = 0; jj < nk; jj++) {
        T acc = 0;
#pragma unroll
        for (unsigned j = 0; j < NPM; j++)
          if (j < np) acc += ev[j] * win[jj * L + pat.poff[j] + t];
        Y[(size_t)(j0 + jj) * ldy + row0 + t] = acc;
      }
    }
  } else {
    const unsigned *__restrict__ cl = a.cols;
    const unsigned c0 = sp[3 * slice + 1], d0 = sp[3 * slice + 2];
    const unsigned e1 = sp[3 * slice + 3], c1 = sp[3 * slice + 4];
    const unsigned nell = (c1 - c0) >> LOG2S;
    const unsigned nsec = ((e1 - e0) >> LOG2S) - nell;
    T acc[K];
#pragma unroll
    for (unsigned j = 0; j < K; j++) acc[j] = 0;
    for (unsigned q = 0; q < nell; q++) {
      const unsigned off = (q << LOG2S) + t;
      const T e = elt<T, E>::get(el, e0 + off);
      const unsigned c = dev::col_load(cl + c0 + off);
#pragma unroll
      for (unsigned j = 0; j < K; j++) acc[j] += e * xload(X + (size_t)j * ldx, c, ncols);
    }
    for (unsigned d = 0; d < nsec; d++) {
      const unsigned off = ((nell + d) << LOG2S) + t;
      const T e = elt<T, E>::get(el, e0 + off);
      const unsigned c = sp[d0 + d] + t - S;
#pragma unroll
      for (unsigned j = 0; j < K; j++) acc[j] += e * xload(X + (size_t)j * ldx, c, ncols);
    }
#pragma unroll
    for (unsigned j = 0; j < K; j++) Y[(size_t)j * ldy + row0 + t] = acc[j];
  }
}

/* LDS the multi-vector pattern kernel stages per workgroup: the windows of kg vectors, kg as
 * many as fit 20 KiB (at least one; four 256-thread workgroups or more per CU) */
template<typename T>
inline unsigned diagpat_spmm_group(const diag_pattern &c, unsigned k)
{
  const size_t per = (size_t)c.lds_len * sizeof(T);
  return (unsigned)std::max<size_t>(1, std::min<size_t>(k, 20480 / per));
}

template<typename T, typename E, unsigned K>
int launch_diagpat_spmm(const ispm_slell_desc *A, const diag_pattern &pat, const unsigned *uni,
                        const void *X, size_t ldx, void *Y, size_t ldy)
{
  spmv_args a;
  a.skip = nullptr;
  a.sp = A->slice_ptr;
  a.cols = A->cols;
  a.elms = A->elms;
  a.x = X;
  a.y = Y;
  a.ldx = ldx;
  a.ldy = ldy;
  a.n_slices = A->n_slices;
  a.log2S = 8;
  a.n_cols = (unsigned)A->n_cols;
  a.spb = 1;
  a.per_xcd = (A->n_slices + 7) / 8;
  diag_pattern c;
  cover_pattern(pat, 16 / sizeof(T), 256, kPatPlanes, 1u << 20, c);
  const unsigned kg = diagpat_spmm_group<T>(c, K);
  const size_t lds = (size_t)kg * c.lds_len * sizeof(T);
  if (lds > 65536) return fail(ISPM_ERR_UNSUPPORTED, "spmm: pattern windows need %zu bytes of LDS", lds);
  const dim3 g(a.per_xcd * 8), b(256);
  if (pat.n <= 8) hipLaunchKernelGGL((slell_diagpat_spmm_kernel<T, E, K, 8>), g, b, lds, stream(), a, c, uni, kg);
  else hipLaunchKernelGGL((slell_diagpat_spmm_kernel<T, E, K, 27>), g, b, lds, stream(), a, c, uni, kg);
  ISPM_HIP(hipGetLastError());
  return ISPM_OK;
}

/* ---- symmetric diagonal storage (ISPM_SPMV_PLAN_SYMMETRIC) ----
 * A symmetric matrix whose entries lie on a few diagonals (stencils) stores only its upper
 * diagonals d_0 = 0 < d_1 < ... < d_{K-1}: plane k of 256-row slice s holds A[row, row + d_k]
 * at elms[(s K + k) kSymPlane + r] (zero where the row has no such entry; rows 256 .. 259 of
 * a plane repeat the next slice's first rows).  The lower entries are
 * the mirrors: A[row, row - d_k] = A[row - d_k, row] = plane k at row - d_k, which an earlier
 * slice stored and read moments ago, so it comes from the cache while only the K upper planes
 * stream from HBM (half a 27-point stencil's bytes).  One workgroup per slice, two
 * consecutive rows per thread for FP32 vectors and elements, else one; x is read from LDS
 * windows as in the pattern kernel (the pattern lists all 2K-1 column offsets). */
constexpr unsigned kSymPlanes = 16;

struct sym_args {
  const void *elms;
  const void *x;
  void *y;
  unsigned n_slices, n_cols, K, per_xcd;
  /* level order (band != 0): slices s = z P + q; XCD x takes positions q in
   * [x P / 8, (x + 1) P / 8) and sweeps the Z full levels z, then the slices past them */
  unsigned band, P, Z;
  int d[kSymPlanes];             /* upper diagonal offsets, d[0] = 0 */
  unsigned up_off[kSymPlanes];   /* LDS index of x[row0 + d[k]] (windows as launched) */
  unsigned lo_off[kSymPlanes];   /* LDS index of x[row0 - d[k]] (k >= 1) */
  const unsigned *skip;
  const void *dotv;
  double *dparts;
};

/* Element reads of the symmetric kernel go through the caches (the mirror reads find them
 * there) and stay raw, one element per register, until the products: a conversion, or two
 * BF16 halves packed into one register, right after a load would make the wave wait for it
 * before issuing the next.  H consecutive elements of one plane: sym_rows::get from an
 * H-aligned index, get_any from any index (a 4-byte aligned FP32 pair is one 8-byte load). */
template<typename E> struct sym_raw { typedef E type; };
template<> struct sym_raw<dev::bf16_t> { typedef unsigned type; };
template<typename T, typename E> __device__ __forceinline__ T sym_cvt(typename sym_raw<E>::type v) { return (T)v; }
template<> __device__ __forceinline__ float sym_cvt<float, dev::bf16_t>(unsigned v) { return __uint_as_float(v << 16); }

template<typename E, unsigned H> struct sym_rows {
  typedef typename sym_raw<E>::type R;
  static __device__ __forceinline__ void get(const E *p, unsigned i, R (&v)[H])
  {
#pragma unroll
    for (unsigned e = 0; e < H; e++) v[e] = (R)p[i + e];
  }
  static __device__ __forceinline__ void get_any(const E *p, unsigned i, R (&v)[H]) { get(p, i, v); }
};
template<> struct sym_rows<float, 2> {
  typedef float f2a __attribute__((ext_vector_type(2)));
  typedef float f2u __attribute__((ext_vector_type(2))) __attribute__((aligned(4)));
  static __device__ __forceinline__ void get(const float *p, unsigned i, float (&v)[2])
  {
    const f2a q = *reinterpret_cast<const f2a *>(p + i);
    v[0] = q.x; v[1] = q.y;
  }
  static __device__ __forceinline__ void get_any(const float *p, unsigned i, float (&v)[2])
  {
    const f2u q = *reinterpret_cast<const f2u *>(p + i);
    v[0] = q.x; v[1] = q.y;
  }
};

/* rows per thread: two for FP32 vectors and elements (8-byte element pairs halve the load
 * requests, which bound that kernel), else one */
template<typename T, typename E> constexpr unsigned sym_rows_per_thread() { return sizeof(T) == 4 && sizeof(E) == 4 ? 2u : 1u; }

template<typename T, typename E, unsigned NK>
__global__ void __launch_bounds__((256 / sym_rows_per_thread<T, E>())) symdia_kernel(sym_args a, diag_pattern pat)
{
  extern __shared__ __align__(16) unsigned char sym_lds[];
  T *win = reinterpret_cast<T *>(sym_lds);
  constexpr unsigned S = 256, H = sym_rows_per_thread<T, E>(), NT = S / H, VX = 16 / sizeof(T), NR = 9;
  const unsigned t = threadIdx.x;
  const unsigned xcd = blockIdx.x & 7u, i = blockIdx.x >> 3;
  unsigned slice = xcd * a.per_xcd + i;
  if (a.band) {
    const unsigned q0 = xcd * a.P / 8, w = (xcd + 1) * a.P / 8 - q0;
    slice = i < w * a.Z ? (i / w) * a.P + q0 + i % w : a.Z * a.P + (i - w * a.Z) * 8 + xcd;
  }
  if (a.skip && *a.skip) return;
  if (slice >= a.n_slices) {
    if (a.dparts && t == 0) a.dparts[blockIdx.x] = 0.0;
    return;
  }
  const E *__restrict__ el = static_cast<const E *>(a.elms);
  const T *__restrict__ xv = static_cast<const T *>(a.x);
  T *__restrict__ yv = static_cast<T *>(a.y);
  const unsigned ncols = a.n_cols, K = a.K;
  const unsigned row0 = slice * S, row = row0 + H * t;
  const bool xvec = (reinterpret_cast<uintptr_t>(xv) % 16) == 0;
  const bool fast = xvec && pat.vec_runs && pat.nrun <= NR && (long long)row0 + pat.cover_lo >= 0 &&
                    (long long)row0 + pat.cover_hi <= (long long)ncols;
  /* the window vectors spread evenly over the threads (at most NW each, the launch checks) */
  constexpr unsigned NW = 5;
  vecx<T> w[NW];
  if (fast) {
#pragma unroll
    for (unsigned q = 0; q < NW; q++) {
      const unsigned v0 = t + q * NT, v = v0 < pat.wtotal ? v0 : pat.wtotal - 1;
      int dl = pat.wdelta[0];
#pragma unroll
      for (unsigned r = 1; r < NR; r++)
        if (r < pat.nrun && v >= pat.wvbase[r]) dl = pat.wdelta[r];
      w[q] = *reinterpret_cast<const vecx<T> *>(xv + (unsigned)((int)row0 + dl + (int)(v * VX)));
    }
  }
  /* own upper planes (rows row .. row+H-1) and the mirrored lower ones, all requested before
   * any wait; element indices fit 32 bits (the plan checks it).  A mirror pair that crosses a
   * slice boundary reads one element past the slice's plane (still inside the array: the
   * mirror rows lie below the last row) and takes its second value from the next slice. */
  typedef typename sym_raw<E>::type R;
  R ev[NK][H], mv[NK][H];
  constexpr unsigned PL = kSymPlane;
  const unsigned KS = K * PL, own = slice * KS + H * t;
#pragma unroll
  for (unsigned k = 0; k < NK; k++) {
#pragma unroll
    for (unsigned e = 0; e < H; e++) ev[k][e] = mv[k][e] = R(0);
    if (k < K) sym_rows<E, H>::get(el, own + k * PL, ev[k]);
    if (k >= 1 && k < K) {
      const int j = (int)row - a.d[k];
      if (j >= 0) {
        /* a pair starting at a plane's last row reads the repeated next row */
        sym_rows<E, H>::get_any(el, ((unsigned)j >> 8) * KS + k * PL + ((unsigned)j & 255u), mv[k]);
      } else if (H == 2 && j == -1) {
        mv[k][H - 1] = (R)el[k * PL];
      }
    }
  }
  if (fast) {
#pragma unroll
    for (unsigned q = 0; q < NW; q++)
      if (t + q * NT < pat.wtotal) *reinterpret_cast<vecx<T> *>(win + (t + q * NT) * VX) = w[q];
  } else {
    for (unsigned r = 0; r < pat.nrun; r++) {
      const long long x0 = (long long)row0 + pat.rrel[r];
      const unsigned L = pat.rlen[r], b = pat.rbase[r];
      for (unsigned q = t; q < L; q += NT) {
        const long long c = x0 + q;
        win[b + q] = (c >= 0 && c < (long long)ncols) ? xv[c] : (T)0;
      }
    }
  }
  __syncthreads();
  T acc[H];
#pragma unroll
  for (unsigned e = 0; e < H; e++) acc[e] = 0;
#pragma unroll
  for (unsigned k = 0; k < NK; k++)
    if (k < K) {
#pragma unroll
      for (unsigned e = 0; e < H; e++) {
        acc[e] += sym_cvt<T, E>(ev[k][e]) * win[a.up_off[k] + H * t + e];
        if (k >= 1) acc[e] += sym_cvt<T, E>(mv[k][e]) * win[a.lo_off[k] + H * t + e];
      }
    }
#pragma unroll
  for (unsigned e = 0; e < H; e++) yv[row + e] = acc[e];
  if (a.dparts) {
    __shared__ double red[4];
    const T *__restrict__ dv = static_cast<const T *>(a.dotv);
    /* CG's (p, A p) has x itself as the dot vector: its rows are in the window of offset 0 */
    const bool dx = dv == xv;
    double d[1] = {0.0};
#pragma unroll
    for (unsigned e = 0; e < H; e++)
      d[0] += (double)(dx ? win[a.up_off[0] + H * t + e] : dv[row + e]) * (double)acc[e];
    dev::block_sum<NT, 1>(d, red);
    if (t == 0) a.dparts[blockIdx.x] = d[0];
  }
}

/* FP32 vectors and elements, four rows per lane.  The kernel above is bound by the rate at
 * which the vector memory pipeline accepts load instructions (address unit busy ~85% of the
 * time), not by bytes: here every load moves 16 bytes per lane, so a 256-row slice is one
 * wave that issues 14 own-plane loads, 12 or 13 mirror loads (4 consecutive rows at 4-byte
 * alignment, inside the plane thanks to its repeated rows) and about 9 window loads, against
 * about 80 load instructions over two waves above.  A workgroup is one wave with its own LDS
 * windows: no barrier, and LDS and registers are handed on as each wave ends (four waves per
 * group ran 6% slower, 103.6 against 97 us at 27 points 192^3).  The windows go from memory
 * to LDS directly, so every load is issued before the one wait, and the products are
 * kept in plane order (a compiler barrier per plane): left free, the compiler read every
 * plane's x values from LDS before the first product and needed 221 registers (two waves
 * per SIMD); this way 120 (four).  x must be readable, and zero outside [0, n_cols), on every
 * slice's window covers (the solvers keep zero guard zones around their vectors for this);
 * mirror rows below 0 read as zero. */
constexpr unsigned kR4Waves = 1, kR4Win = 10;

/* 16 bytes per lane from global memory straight to LDS: lane l writes lds + 16 l (lds is
 * uniform over the wave); completion is counted with the wave's vector memory loads */
__device__ __forceinline__ void lds_dma16(const void *g, void *lds)
{
  __builtin_amdgcn_global_load_lds(g, (__attribute__((address_space(3))) void *)lds, 16, 0, 0);
}

/* lane l - 1's v in lane l (lane 0: 0) */
__device__ __forceinline__ float lane_up(float v)
{
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138 /* wave_shr:1 */, 0xf, 0xf, false));
}

/* the four rows of one lane (see symdia_r4_kernel); x must be readable (and 0) on the
 * window covers of every slice, and mirror rows below 0 are read at row 0 and zeroed.
 * D1: d[1] == 1, whose mirror A[row, row - 1] is plane 1 one row up: the lane's own rows
 * but the first, which is the previous lane's last (lane 0: the previous slice's row 255),
 * so it takes no load. */
template<unsigned NK, bool D1>
__device__ __forceinline__ float __attribute__((ext_vector_type(4)))
symdia_r4_rows(const sym_args &a, const diag_pattern &pat, const float *__restrict__ el,
               const float *__restrict__ xv, float *win, unsigned slice, int row0, unsigned lane)
{
  typedef float f4a __attribute__((ext_vector_type(4)));
  typedef float f4u __attribute__((ext_vector_type(4))) __attribute__((aligned(4)));
  constexpr unsigned PL = kSymPlane, K = NK, KS = K * PL;
  const int row = row0 + 4 * (int)lane;
  /* windows go straight from memory to LDS (16 bytes per lane), so every plane load can be
   * issued before the first wait; run by run, 64 vectors u .. u + 63 of a run per load, so
   * the run's offset is uniform and a lane's address is one add */
#pragma unroll
  for (unsigned r = 0; r < 9; r++)   /* fixed bound: the run table is read once, as scalars */
    if (r < pat.nrun) {
      const unsigned vb = pat.wvbase[r], ve = r + 1 < pat.nrun ? pat.wvbase[r + 1] : pat.wtotal;
      const float *xr = xv + (row0 + pat.wdelta[r]);
      for (unsigned u = vb; u < ve; u += 64)
        if (u + lane < ve) lds_dma16(xr + 4 * (u + lane), win + 4 * u);
    }
  /* 32-bit byte offsets from the uniform base (the stored planes are below 4 GiB): one
   * address register per load */
  const char *eb = reinterpret_cast<const char *>(el);
  const unsigned ownb = (slice * KS + 4 * lane) * 4u;
  f4a ev[NK];
  f4u mv[NK];
#pragma unroll
  for (unsigned k = 0; k < NK; k++) {
    ev[k] = *reinterpret_cast<const f4a *>(eb + (ownb + k * PL * 4u));
    mv[k] = f4u{0.f, 0.f, 0.f, 0.f};
    if (k >= 1 && !(D1 && k == 1)) {
      const unsigned j = (unsigned)max(row - a.d[k], 0);
      f4u m = *reinterpret_cast<const f4u *>(eb + ((j >> 8) * KS + k * PL + (j & 255u)) * 4u);
      if (row0 < a.d[k]) {   /* only the first slices: uniform over the wave */
        const int jn = row - a.d[k];
        if (jn < 0) {
          /* read rows 0..3 (the clamped start); row jn + e is element jn + e of them */
          const float m0 = m.x, m1 = m.y, m2 = m.z;
          m.x = 0.f;
          m.y = jn == -1 ? m0 : 0.f;
          m.z = jn == -2 ? m0 : jn == -1 ? m1 : 0.f;
          m.w = jn == -3 ? m0 : jn == -2 ? m1 : jn == -1 ? m2 : 0.f;
        }
      }
      mv[k] = m;
    }
  }
  const float up1 = D1 && slice > 0 ? el[(slice - 1) * KS + PL + 255] : 0.f;
  /* every load, the windows' included, is complete before LDS is read */
  __builtin_amdgcn_s_waitcnt(0x0f70);
  __builtin_amdgcn_wave_barrier();
  if (D1) {
    const float u = lane_up(ev[1].w);
    mv[1] = f4u{lane == 0 ? up1 : u, ev[1].x, ev[1].y, ev[1].z};
  }
  f4a acc = f4a{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (unsigned k = 0; k < NK; k++) {
    const f4a e = ev[k];
    const float *xu = win + a.up_off[k] + 4 * lane;
    acc.x += e.x * xu[0];
    acc.y += e.y * xu[1];
    acc.z += e.z * xu[2];
    acc.w += e.w * xu[3];
    if (k >= 1) {
      const f4u m = mv[k];
      const float *xl = win + a.lo_off[k] + 4 * lane;
      acc.x += m.x * xl[0];
      acc.y += m.y * xl[1];
      acc.z += m.z * xl[2];
      acc.w += m.w * xl[3];
    }
    /* keep each plane's products here: later planes' LDS reads are not hoisted above */
    asm volatile("" : "+v"(acc.x), "+v"(acc.y), "+v"(acc.z), "+v"(acc.w) : : "memory");
  }
  return acc;
}

/* NK: the number of planes (compile time, so that no plane needs a branch); D1: d[1] == 1 */
template<unsigned NWAVES, unsigned NK, bool D1>
__global__ void __launch_bounds__(64 * NWAVES) symdia_r4_kernel(sym_args a, diag_pattern pat)
{
  extern __shared__ __align__(16) unsigned char sym_lds[];
  typedef float f4a __attribute__((ext_vector_type(4)));
  constexpr unsigned S = 256;
  /* the wave index is uniform: read it from lane 0, so slice arithmetic stays scalar */
  const unsigned lane = threadIdx.x & 63, wv = NWAVES == 1 ? 0u : __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  float *win = reinterpret_cast<float *>(sym_lds) + wv * pat.lds_len;
  const unsigned xcd = blockIdx.x & 7u, i = (blockIdx.x >> 3) * kR4Waves + wv;
  unsigned slice = xcd * a.per_xcd + i;
  if (a.band) {
    const unsigned q0 = xcd * a.P / 8, w = (xcd + 1) * a.P / 8 - q0;
    slice = i < w * a.Z ? (i / w) * a.P + q0 + i % w : a.Z * a.P + (i - w * a.Z) * 8 + xcd;
  }
  if (a.skip && *a.skip) return;   /* uniform over the launch */
  const bool active = slice < a.n_slices && i < a.per_xcd;
  const float *__restrict__ el = static_cast<const float *>(a.elms);
  const float *__restrict__ xv = static_cast<const float *>(a.x);
  float *__restrict__ yv = static_cast<float *>(a.y);
  const int row0 = (int)(slice * S), row = row0 + 4 * (int)lane;
  f4a acc = f4a{0.f, 0.f, 0.f, 0.f};
  if (active) acc = symdia_r4_rows<NK, D1>(a, pat, el, xv, win, slice, row0, lane);
  if (active) *reinterpret_cast<f4a *>(yv + row) = acc;
  if (a.dparts) {
    __shared__ double red[kR4Waves];
    const float *__restrict__ dv = static_cast<const float *>(a.dotv);
    double d[1] = {0.0};
    if (active) {
      /* CG's (p, A p) has x as the dot vector: its rows are in the window of offset 0 */
      const bool dx = dv == xv;
      const float *src = dx ? win + a.up_off[0] + 4 * lane : dv + row;
      d[0] = (double)src[0] * (double)acc.x + (double)src[1] * (double)acc.y +
             (double)src[2] * (double)acc.z + (double)src[3] * (double)acc.w;
    }
    dev::block_sum<64 * kR4Waves, 1>(d, red);
    if (threadIdx.x == 0) a.dparts[blockIdx.x] = d[0];
  }
}

/* zero elements x needs on each side so that every slice's FP32 window covers lie inside
 * (the four-rows-per-lane kernel reads them without bounds checks) */
inline unsigned sym_r4_guard(const diag_pattern &pat0, unsigned n_slices, unsigned n_cols)
{
  diag_pattern c4;
  cover_pattern(pat0, 4, 64, 9, kR4Win, c4);
  const long long lo = -c4.cover_lo, hi = (long long)(n_slices - 1) * 256 + c4.cover_hi - (long long)n_cols;
  const long long g = std::max(0ll, std::max(lo, hi));
  return (unsigned)((g + 63) / 64 * 64);
}

template<typename T, typename E>
int launch_symdia(const sym_args &a0, const diag_pattern &pat0, const void *x, void *y,
                  const unsigned *skip, const void *dotv, double *dparts, unsigned *dgrid,
                  unsigned x_guard)
{
  sym_args a = a0;
  a.x = x;
  a.y = y;
  a.skip = skip;
  a.per_xcd = (a.n_slices + 7) / 8;
  /* Level order when the widest offset spans P slices that split (nearly) evenly over the
   * XCDs: the mirrors a slice reads P slices back were then read by the same XCD one band
   * of P / 8 slices earlier instead of P, and stay in its L2. */
  const unsigned P = (unsigned)(a.d[a.K - 1] / 256);
  a.band = 0;
  if (P >= 8 && (P % 8 == 0 || P >= 128) && a.n_slices / P >= 2) {
    a.band = 1;
    a.P = P;
    a.Z = a.n_slices / P;
    a.per_xcd = (P + 7) / 8 * a.Z + (a.n_slices - a.Z * P + 7) / 8;
  }
  a.dotv = nullptr;
  a.dparts = nullptr;
  if (dgrid) *dgrid = 0;
  if (dotv && dparts && dgrid && a.per_xcd * 8 <= kMaxDotParts) {
    a.dotv = dotv;
    a.dparts = dparts;
    *dgrid = a.per_xcd * 8;
  }
  if (std::is_same<T, float>::value && std::is_same<E, float>::value && x_guard > 0 &&
      (reinterpret_cast<uintptr_t>(x) % 16) == 0 && x_guard >= sym_r4_guard(pat0, a.n_slices, a.n_cols)) {
    diag_pattern c4;
    cover_pattern(pat0, 4, 64, 9, kR4Win, c4);
    const size_t lds4 = (size_t)c4.lds_len * sizeof(float) * kR4Waves;
    /* the 7-, 19- and 27-point stencils (4, 10, 14 upper planes) */
    if (c4.vec_runs && lds4 <= 65536 && (a.K == 4 || a.K == 10 || a.K == 14)) {
      for (unsigned k = 0; k < a.K; k++) {
        a.up_off[k] = c4.poff[(a.K - 1) + k];
        a.lo_off[k] = k ? c4.poff[(a.K - 1) - k] : 0u;
      }
      const unsigned groups = (a.per_xcd + kR4Waves - 1) / kR4Waves;
      if (dgrid && a.dparts) {
        if (groups * 8 > kMaxDotParts) { a.dotv = nullptr; a.dparts = nullptr; *dgrid = 0; }
        else *dgrid = groups * 8;
      }
      const dim3 g4(groups * 8), b4(64 * kR4Waves);
      if (a.d[1] == 1) {
        if (a.K == 4) hipLaunchKernelGGL((symdia_r4_kernel<kR4Waves, 4, true>), g4, b4, lds4, stream(), a, c4);
        else if (a.K == 10) hipLaunchKernelGGL((symdia_r4_kernel<kR4Waves, 10, true>), g4, b4, lds4, stream(), a, c4);
        else hipLaunchKernelGGL((symdia_r4_kernel<kR4Waves, 14, true>), g4, b4, lds4, stream(), a, c4);
      } else {
        if (a.K == 4) hipLaunchKernelGGL((symdia_r4_kernel<kR4Waves, 4, false>), g4, b4, lds4, stream(), a, c4);
        else if (a.K == 10) hipLaunchKernelGGL((symdia_r4_kernel<kR4Waves, 10, false>), g4, b4, lds4, stream(), a, c4);
        else hipLaunchKernelGGL((symdia_r4_kernel<kR4Waves, 14, false>), g4, b4, lds4, stream(), a, c4);
      }
      ISPM_HIP(hipGetLastError());
      return ISPM_OK;
    }
  }
  constexpr unsigned NT = 256 / sym_rows_per_thread<T, E>(), VX = 16 / sizeof(T);
  diag_pattern c;
  cover_pattern(pat0, VX, NT, 9, 5, c);
  const size_t lds = (size_t)c.lds_len * sizeof(T);
  if (lds > 65536) return fail(ISPM_ERR_UNSUPPORTED, "spmv: symmetric windows need %zu bytes of LDS", lds);
  /* pattern positions of +d[k] and -d[k] (the plan lists -d[K-1] .. -d[1], then d[0] ..
   * d[K-1]) turned into LDS indices of the windows as launched */
  for (unsigned k = 0; k < a.K; k++) {
    a.up_off[k] = c.poff[(a.K - 1) + k];
    a.lo_off[k] = k ? c.poff[(a.K - 1) - k] : 0u;
  }
  const dim3 g(a.per_xcd * 8), b(NT);
  hipStream_t st = stream();
  if (a.K <= 4) hipLaunchKernelGGL((symdia_kernel<T, E, 4>), g, b, lds, st, a, c);
  else if (a.K <= 10) hipLaunchKernelGGL((symdia_kernel<T, E, 10>), g, b, lds, st, a, c);
  else hipLaunchKernelGGL((symdia_kernel<T, E, kSymPlanes>), g, b, lds, st, a, c);
  ISPM_HIP(hipGetLastError());
  return ISPM_OK;
}

/* Multi-vector product of symmetric diagonal storage (k = 2..8 vectors, column j at x + j ldx):
 * one 256-row slice per workgroup, one row per thread, slices in the single-vector kernel's
 * (level) order.  The row's K own entries and K - 1 mirrors are loaded once; then, KG vectors
 * at a time, the vectors' x windows are staged in LDS and each row adds, plane by plane, its
 * upper product and then its mirrored one -- the order of symdia_kernel, so every column is
 * bit-identical to the single-vector symmetric product. */
template<typename T, typename E, unsigned KV, unsigned NK>
__global__ void __launch_bounds__(256)
symdia_spmm_kernel(sym_args a, diag_pattern pat, size_t ldx, size_t ldy, unsigned kg)
{
  extern __shared__ __align__(16) unsigned char sym_lds[];
  T *win = reinterpret_cast<T *>(sym_lds);
  constexpr unsigned S = 256, VX = 16 / sizeof(T), QM = 4, PL = kSymPlane;
  const unsigned t = threadIdx.x;
  const unsigned xcd = blockIdx.x & 7u, i = blockIdx.x >> 3;
  unsigned slice = xcd * a.per_xcd + i;
  if (a.band) {
    const unsigned q0 = xcd * a.P / 8, w = (xcd + 1) * a.P / 8 - q0;
    slice = i < w * a.Z ? (i / w) * a.P + q0 + i % w : a.Z * a.P + (i - w * a.Z) * 8 + xcd;
  }
  if (slice >= a.n_slices) return;   /* uniform over the workgroup */
  const E *__restrict__ el = static_cast<const E *>(a.elms);
  const T *__restrict__ X = static_cast<const T *>(a.x);
  T *__restrict__ Y = static_cast<T *>(a.y);
  const unsigned ncols = a.n_cols, K = a.K, L = pat.lds_len, wt = pat.wtotal, nrun = pat.nrun;
  const unsigned row0 = slice * S, row = row0 + t, KS = K * PL;
  typedef typename sym_raw<E>::type R;
  R ev[NK], mv[NK];
#pragma unroll
  for (unsigned k = 0; k < NK; k++) {
    ev[k] = mv[k] = R(0);
    if (k < K) ev[k] = (R)el[slice * KS + k * PL + t];
    if (k >= 1 && k < K) {
      const int j = (int)row - a.d[k];
      if (j >= 0) mv[k] = (R)el[((unsigned)j >> 8) * KS + k * PL + ((unsigned)j & 255u)];
    }
  }
  /* window vectors t + q S, q < QM, offsets computed once (longer lists, unaligned columns
   * and covers reaching past x stage through the bounds-checked loop) */
  const bool fast = ((reinterpret_cast<uintptr_t>(X) | (ldx * sizeof(T))) % 16) == 0 && wt <= QM * S &&
                    (long long)row0 + pat.cover_lo >= 0 && (long long)row0 + pat.cover_hi <= (long long)ncols;
  int dq[QM];
#pragma unroll
  for (unsigned q = 0; q < QM; q++) {
    const unsigned vv = t + q * S;
    int dl = pat.wdelta[0];
    for (unsigned r = 1; r < nrun; r++)
      if (vv >= pat.wvbase[r]) dl = pat.wdelta[r];
    dq[q] = (int)row0 + dl + (int)(vv * VX);
  }
  for (unsigned j0 = 0; j0 < KV; j0 += kg) {
    const unsigned nk = min(kg, KV - j0);
    if (j0) __syncthreads();   /* the previous vectors' windows are read */
    for (unsigned jj = 0; jj < nk; jj++) {
      const T *xj = X + (size_t)(j0 + jj) * ldx;
      if (fast) {
#pragma unroll
        for (unsigned q = 0; q < QM; q++)
          if (t + q * S < wt)
            *reinterpret_cast<vecx<T> *>(win + jj * L + (t + q * S) * VX) = *reinterpret_cast<const vecx<T> *>(xj + dq[q]);
      } else {
        for (unsigned r = 0; r < nrun; r++) {
          const long long x0 = (long long)row0 + pat.rrel[r];
          const unsigned Lr = pat.rlen[r], b = pat.rbase[r];
          for (unsigned u = t; u < Lr; u += S) {
            const long long c = x0 + u;
            win[jj * L + b + u] = (c >= 0 && c < (long long)ncols) ? xj[c] : (T)0;
          }
        }
      }
    }
    __syncthreads();
    for (unsigned jj = 0; jj < nk; jj++) {
      T acc = 0;
#pragma unroll
      for (unsigned k = 0; k < NK; k++)
        if (k < K) {
          acc += sym_cvt<T, E>(ev[k]) * win[jj * L + a.up_off[k] + t];
          if (k >= 1) acc += sym_cvt<T, E>(mv[k]) * win[jj * L + a.lo_off[k] + t];
        }
      Y[(size_t)(j0 + jj) * ldy + row] = acc;
    }
  }
}

template<typename T, typename E, unsigned KV>
int launch_symdia_spmm(const sym_args &a0, const diag_pattern &pat0, const void *X, size_t ldx,
                       void *Y, size_t ldy)
{
  sym_args a = a0;
  a.x = X;
  a.y = Y;
  a.skip = nullptr;
  a.dotv = nullptr;
  a.dparts = nullptr;
  a.per_xcd = (a.n_slices + 7) / 8;
  /* the single-vector kernel's level order (mirrors stay in the XCD's L2) */
  const unsigned P = (unsigned)(a.d[a.K - 1] / 256);
  a.band = 0;
  if (P >= 8 && (P % 8 == 0 || P >= 128) && a.n_slices / P >= 2) {
    a.band = 1;
    a.P = P;
    a.Z = a.n_slices / P;
    a.per_xcd = (P + 7) / 8 * a.Z + (a.n_slices - a.Z * P + 7) / 8;
  }
  constexpr unsigned VX = 16 / sizeof(T);
  diag_pattern c;
  cover_pattern(pat0, VX, 256, kPatPlanes, 1u << 20, c);
  for (unsigned k = 0; k < a.K; k++) {
    a.up_off[k] = c.poff[(a.K - 1) + k];
    a.lo_off[k] = k ? c.poff[(a.K - 1) - k] : 0u;
  }
  const unsigned kg = (unsigned)std::max<size_t>(1, std::min<size_t>(KV, 20480 / ((size_t)c.lds_len * sizeof(T))));
  const size_t lds = (size_t)kg * c.lds_len * sizeof(T);
  if (lds > 65536) return fail(ISPM_ERR_UNSUPPORTED, "spmm: symmetric windows need %zu bytes of LDS", lds);
  const dim3 g(a.per_xcd * 8), b(256);
  if (a.K <= 10) hipLaunchKernelGGL((symdia_spmm_kernel<T, E, KV, 10>), g, b, lds, stream(), a, c, ldx, ldy, kg);
  else hipLaunchKernelGGL((symdia_spmm_kernel<T, E, KV, kSymPlanes>), g, b, lds, stream(), a, c, ldx, ldy, kg);
  ISPM_HIP(hipGetLastError());
  return ISPM_OK;
}

template<typename T, typename E>
int launch_symdia_spmm_k(const sym_args &a, const diag_pattern &pat, int k, const void *X, size_t ldx,
                         void *Y, size_t ldy)
{
  switch (k) {
    case 2: return launch_symdia_spmm<T, E, 2>(a, pat, X, ldx, Y, ldy);
    case 3: return launch_symdia_spmm<T, E, 3>(a, pat, X, ldx, Y, ldy);
    case 4: return launch_symdia_spmm<T, E, 4>(a, pat, X, ldx, Y, ldy);
    case 5: return launch_symdia_spmm<T, E, 5>(a, pat, X, ldx, Y, ldy);
    case 6: return launch_symdia_spmm<T, E, 6>(a, pat, X, ldx, Y, ldy);
    case 7: return launch_symdia_spmm<T, E, 7>(a, pat, X, ldx, Y, ldy);
    default: return launch_symdia_spmm<T, E, 8>(a, pat, X, ldx, Y, ldy);
  }
}

/* pattern path: a plan's diagonal fixed-height slices, one row per thread, single vector */
inline bool use_diagpat(const ispm_launch *p, const ispm_slell_desc *A, int vt, unsigned H,
                        const diag_pattern *pat, const unsigned *uni)
{
  return pat && uni && pat->n && p->mode == 0 && !A->var_height && A->diags && A->hblock == 1 &&
         (H == 1 || H == 2) && p->N * H == A->slice_height && A->slice_height >= 64 &&
         A->slice_height <= 512 &&
         ((size_t)pat->lds_len + 2 * pat->nrun * (16 / type_size(vt))) * type_size(vt) <= 65536;
}

/* the multi-vector pattern kernel: 256-row slices, FP64 or FP32 (vectors and elements alike),
 * at most 27 pattern planes */
inline bool use_diagpat_spmm(const ispm_launch *p, const ispm_slell_desc *A, int vt,
                             const diag_pattern *pat, const unsigned *uni)
{
  if (!use_diagpat(p, A, vt, 1, pat, uni) || A->slice_height != 256 || pat->n > 27 ||
      !((vt == ISPM_F64 && A->elm_type == ISPM_F64) || (vt == ISPM_F32 && A->elm_type == ISPM_F32)))
    return false;
  /* at least one vector's windows fit in LDS */
  diag_pattern c;
  cover_pattern(*pat, 16 / type_size(vt), 256, kPatPlanes, 1u << 20, c);
  return (size_t)c.lds_len * type_size(vt) <= 65536;
}

template<unsigned K>
int launch_diagpat_spmm_types(const ispm_slell_desc *A, int vt, const diag_pattern &pat, const unsigned *uni,
                              const void *X, size_t ldx, void *Y, size_t ldy)
{
  if (vt == ISPM_F64) return launch_diagpat_spmm<double, double, K>(A, pat, uni, X, ldx, Y, ldy);
  return launch_diagpat_spmm<float, float, K>(A, pat, uni, X, ldx, Y, ldy);
}

inline int launch_diagpat_types(const ispm_slell_desc *A, int vt, const diag_pattern &pat,
                                const unsigned *uni, const void *x, void *y, const unsigned *skip,
                                const void *dotv = nullptr, double *dparts = nullptr,
                                unsigned *dgrid = nullptr)
{
  const int et = A->elm_type;
  if (vt == ISPM_F64 && et == ISPM_F64) return launch_diagpat<double, double>(A, pat, uni, x, y, skip, dotv, dparts, dgrid);
  if (vt == ISPM_F64 && et == ISPM_F32) return launch_diagpat<double, float>(A, pat, uni, x, y, skip, dotv, dparts, dgrid);
  if (vt == ISPM_F32 && et == ISPM_F32) return launch_diagpat<float, float>(A, pat, uni, x, y, skip, dotv, dparts, dgrid);
  if (vt == ISPM_F32 && et == ISPM_BF16) return launch_diagpat<float, bf16_t>(A, pat, uni, x, y, skip, dotv, dparts, dgrid);
  return fail(ISPM_ERR_UNSUPPORTED, "spmv: unsupported (vector, element) types (%d, %d)", vt, et);
}

/* hblock slices of fixed height S (64..512), one workgroup of S threads per slice, one row
 * per thread (the texture-pair fetches of cuda/spmv/kernel.h:93-119 become LDS reads).  A
 * slice with an x window (xwin[2 s] = first column, xwin[2 s + 1] = length, staged by the
 * plan when the slice's columns are close together) loads that stretch of x into LDS with
 * 16-byte vectors and reads every single and pair entry from there; other slices gather
 * from global memory.  Each row adds its single planes, then its pair planes, in storage
 * order (as slell_kernel does); a padding slot's column may lie outside the window, its
 * index is clamped (the value there is 0). */
template<typename T, typename E>
__global__ void __launch_bounds__(512) slell_hwin_kernel(spmv_args a, const unsigned *__restrict__ xwin)
{
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *win = reinterpret_cast<T *>(lds_raw);
  if (a.skip && *a.skip) return;
  const unsigned slice = (blockIdx.x & 7u) * a.per_xcd + (blockIdx.x >> 3);
  if (slice >= a.n_slices) return;   /* uniform over the workgroup */
  const unsigned log2S = a.log2S, S = 1u << log2S, t = threadIdx.x;
  /* records of 2 words (element offset, column offset) for hblock without variable heights
   * or diagonals */
  const unsigned *rec = a.sp + 2 * slice;
  const unsigned e0 = rec[0], c0 = rec[1], e1 = rec[2], c1 = rec[3];
  const unsigned nse = 2 * (c1 - c0) - (e1 - e0);
  const unsigned nell = nse >> log2S, nsec = ((c1 - c0) >> log2S) - nell;
  const E *__restrict__ el = static_cast<const E *>(a.elms);
  const unsigned *__restrict__ cl = a.cols;
  const T *__restrict__ xv = static_cast<const T *>(a.x);
  T *__restrict__ yv = static_cast<T *>(a.y);
  const unsigned ncols = a.n_cols;
  const unsigned wb = xwin[2 * slice], wl = xwin[2 * slice + 1];
  if (wl) {
    /* the window as aligned 16-byte vectors where x allows it, element by element at its
     * ends (and throughout when x itself is not 16-byte aligned) */
    constexpr unsigned VX = 16 / sizeof(T);
    const bool xal = (reinterpret_cast<uintptr_t>(xv) & 15u) == 0;
    const unsigned a0 = xal ? min((wb + VX - 1) & ~(VX - 1), wb + wl) : wb + wl;
    const unsigned a1 = max(a0, xal ? (wb + wl) & ~(VX - 1) : wb + wl);
    for (unsigned q = t; q < a0 - wb; q += S) win[q] = xv[wb + q];
    for (unsigned q = a1 - wb + t; q < wl; q += S) win[q] = xv[wb + q];
    const unsigned nv = (a1 - a0) / VX;
    for (unsigned v = t; v < nv; v += S) {
      const vecx<T> xx = *reinterpret_cast<const vecx<T> *>(xv + a0 + v * VX);
#pragma unroll
      for (unsigned e = 0; e < VX; e++) win[a0 - wb + v * VX + e] = xx[e];
    }
    __syncthreads();
  }
  auto X = [&](unsigned c) -> T {
    if (wl) {
      const int i = (int)c - (int)wb;
      return win[i < 0 ? 0 : (i >= (int)wl ? (int)wl - 1 : i)];
    }
    return xload(xv, c, ncols);
  };
  T acc = 0;
  for (unsigned k = 0; k < nell; k++) {
    const unsigned off = (k << log2S) + t;
    acc += elt<T, E>::get(el, e0 + off) * X(dev::col_load(cl + c0 + off));
  }
  const unsigned boff = nell << log2S;
  for (unsigned d = 0; d < nsec; d++) {
    const unsigned xc = 2 * dev::col_load(cl + c0 + ((nell + d) << log2S) + t);
    const unsigned eo = e0 + boff + ((2 * d) << log2S) + t;
    acc += elt<T, E>::get(el, eo) * X(xc) + elt<T, E>::get(el, eo + S) * X(xc + 1);
  }
  yv[(slice << log2S) + t] = acc;
}

template<typename T, typename E>
int launch_hwin(const ispm_slell_desc *A, const void *x, void *y, const unsigned *skip,
                const layout_hints *h)
{
  spmv_args a;
  a.skip = skip;
  a.sp = A->slice_ptr;
  a.cols = A->cols;
  a.elms = A->elms;
  a.x = x;
  a.y = y;
  a.ldx = a.ldy = 0;
  a.spb = 1;
  a.n_slices = A->n_slices;
  a.log2S = (unsigned)(31 - __builtin_clz(A->slice_height));
  a.n_cols = (unsigned)A->n_cols;
  a.per_xcd = (A->n_slices + 7) / 8;
  const size_t lds = (size_t)h->xwin_max * sizeof(T);
  hipLaunchKernelGGL((slell_hwin_kernel<T, E>), dim3(a.per_xcd * 8), dim3(A->slice_height), lds, stream(),
                     a, h->xwin);
  ISPM_HIP(hipGetLastError());
  return ISPM_OK;
}

/* windowed hblock path: a plan's hblock layout with fixed slices of 64..512 rows, one row per
 * thread (the default launch), and x windows on record */
inline bool use_hwin(const ispm_launch *p, const ispm_slell_desc *A, int vt, unsigned H,
                     const layout_hints *h)
{
  return h && h->xwin && h->xwin_max && p->mode == 0 && A->hblock == 2 && !A->var_height && !A->diags &&
         H == 1 && A->slice_height >= 64 && A->slice_height <= 512 && p->N == A->slice_height &&
         (size_t)h->xwin_max * type_size(vt) <= 32768;
}

inline int launch_hwin_types(const ispm_slell_desc *A, int vt, const void *x, void *y,
                             const unsigned *skip, const layout_hints *h)
{
  const int et = A->elm_type;
  if (vt == ISPM_F64 && et == ISPM_F64) return launch_hwin<double, double>(A, x, y, skip, h);
  if (vt == ISPM_F64 && et == ISPM_F32) return launch_hwin<double, float>(A, x, y, skip, h);
  if (vt == ISPM_F32 && et == ISPM_F32) return launch_hwin<float, float>(A, x, y, skip, h);
  if (vt == ISPM_F32 && et == ISPM_BF16) return launch_hwin<float, bf16_t>(A, x, y, skip, h);
  return fail(ISPM_ERR_UNSUPPORTED, "spmv: unsupported (vector, element) types (%d, %d)", vt, et);
}

/* Packed mapping for fixed slice heights S < 64: each wave covers 64/S consecutive slices,
 * one row per lane, so every lane works even when a slice has only one or two planes (the
 * slice-per-workgroup mapping would give each row 64/S lanes and leave most of them idle).
 * Lanes of one slice read contiguous runs; consecutive slices follow each other in memory. */
template<typename T, typename E, int W, bool D>
__global__ void __launch_bounds__(256) slell_packed_kernel(spmv_args a)
{
  constexpr int U = 4;                     /* independent rows per lane */
  const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned wpb = blockDim.x >> 6;
  const unsigned log2S = a.log2S, S = 1u << log2S;
  const unsigned spw = 64u >> log2S;
  const unsigned group = (blockIdx.x & 7u) * a.per_xcd + (blockIdx.x >> 3);
  if (a.skip && *a.skip) return;
  constexpr unsigned P = 1 + (W > 1) + (D ? 2 : 0);
  const E *__restrict__ el = static_cast<const E *>(a.elms);
  const unsigned *__restrict__ cl = a.cols;
  const T *__restrict__ xv = static_cast<const T *>(a.x);
  T *__restrict__ yv = static_cast<T *>(a.y);
  const unsigned ncols = a.n_cols;
  const unsigned r = lane & (S - 1);
  for (unsigned it = 0; it < a.spb; it++) {
    /* this wave handles U consecutive groups of 64 rows */
    const unsigned wg0 = ((group * a.spb + it) * wpb + wave) * U;
    if (wg0 * spw >= a.n_slices) return;   /* wave-uniform; no workgroup barriers below */
    unsigned slice[U], e0[U], c0[U], nell[U], nsec[U], d0[U];
    bool on[U];
    unsigned kmax = 0;
#pragma unroll
    for (int u = 0; u < U; u++) {
      slice[u] = (wg0 + u) * spw + (lane >> log2S);
      on[u] = slice[u] < a.n_slices;
      nell[u] = nsec[u] = 0;
      e0[u] = c0[u] = d0[u] = 0;
      if (on[u]) {
        const unsigned *rec = a.sp + P * slice[u];
        e0[u] = rec[0];
        const unsigned e1 = rec[P];
        if (D) {
          c0[u] = rec[1]; d0[u] = rec[2];
          nell[u] = (rec[1 + P] - c0[u]) >> log2S;
          nsec[u] = ((e1 - e0[u]) >> log2S) - nell[u];
        } else if (W > 1) {
          c0[u] = rec[1];
          const unsigned c1 = rec[1 + P];
          const unsigned nse = (W * (c1 - c0[u]) - (e1 - e0[u])) / (W - 1);
          nell[u] = nse >> log2S;
          nsec[u] = ((c1 - c0[u]) >> log2S) - nell[u];
        } else {
          c0[u] = e0[u];
          nell[u] = (e1 - e0[u]) >> log2S;
        }
      }
      const unsigned tot = nell[u] + nsec[u];
      kmax = tot > kmax ? tot : kmax;
    }
    T acc[U];
#pragma unroll
    for (int u = 0; u < U; u++) acc[u] = 0;
    for (unsigned k = 0; k < kmax; k++) {
#pragma unroll
      for (int u = 0; u < U; u++) {
        if (k < nell[u]) {
          const unsigned off = (k << log2S) + r;
          acc[u] += elt<T, E>::get(el, e0[u] + off) * xload(xv, dev::col_load(cl + c0[u] + off), ncols);
        } else if (k < nell[u] + nsec[u]) {
          const unsigned d = k - nell[u];
          if (D) {
            const unsigned off = (k << log2S) + r;
            acc[u] += elt<T, E>::get(el, e0[u] + off) * xload(xv, a.sp[d0[u] + d] + r - S, ncols);
          } else if (W > 1) {
            const unsigned c = 2 * dev::col_load(cl + c0[u] + (k << log2S) + r);
            const unsigned eo = e0[u] + (nell[u] << log2S) + ((2 * d) << log2S) + r;
            acc[u] += elt<T, E>::get(el, eo) * xload(xv, c, ncols) + elt<T, E>::get(el, eo + S) * xload(xv, c + 1, ncols);
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++)
      if (on[u]) yv[(slice[u] << log2S) + r] = acc[u];
  }
}

/* Streamed mapping (mode 2) for fixed-height slices without hblock pairs or diagonal
 * planes.  A 256-thread workgroup takes 256*RPT consecutive rows, whose elements form one
 * contiguous range of the element and column arrays.  Phase 1 reads that range with
 * consecutive lanes on consecutive elements (every fetched line used in full by one
 * instruction), up to eight elements per thread in flight, and leaves the products
 * e*x[col] in LDS in storage order; phase 2 gives each thread RPT rows and adds each row's
 * products plane by plane, in storage order.  The LDS product buffer holds cap elements:
 * plans size it from their layout's largest group, so every group takes this path; a group
 * whose range exceeds it (launches without layout facts) reads its rows directly, one
 * dependent gather per element.  Suited to irregular, short rows (approximate-inverse
 * factors): the work of a workgroup is three dependent memory round trips (offsets,
 * elements and columns, x). */
template<typename T, typename E, int RPT, bool AX = false>
__device__ __forceinline__ void stream_body(const spmv_args &a, unsigned cap, unsigned bid)
{
  constexpr unsigned ROWS = 256u * RPT, U = 4, VEC = dev::evec<E>::n;
  /* products (cap elements, sized by the launch), then the group's slice offsets */
  extern __shared__ __align__(16) unsigned char stream_lds[];
  T *prod = reinterpret_cast<T *>(stream_lds);
  unsigned *soff = reinterpret_cast<unsigned *>(prod + cap);
  if (a.skip && *a.skip) return;
  const unsigned log2S = a.log2S, S = 1u << log2S;
  const unsigned spg = ROWS >> log2S;
  const unsigned t = threadIdx.x;
  const E *__restrict__ el = static_cast<const E *>(a.elms);
  const unsigned *__restrict__ cl = a.cols;
  const T *__restrict__ xv = static_cast<const T *>(a.x);
  T *__restrict__ yv = static_cast<T *>(a.y);
  const T *__restrict__ dv = static_cast<const T *>(a.dotv);
  const unsigned ncols = a.n_cols;
  const xsrc<T, AX> xs{xv, static_cast<const T *>(a.axq), AX ? (T)*a.axa : (T)0, ncols};
  const unsigned group = (bid & 7u) * a.per_xcd + (bid >> 3);
  const unsigned s0 = group * spg;
  if (s0 >= a.n_slices) {   /* uniform over the workgroup */
    if (a.dparts && t == 0) a.dparts[bid] = 0.0;
    return;
  }
  const unsigned ns = min(spg, a.n_slices - s0);
  /* Destination rows (~0u: none), and what the epilogue reads there (the row map's entry,
   * the diagonal term of a split-diagonal row, the dot product's vector): requested first,
   * as they do not depend on the layout's offsets, so their latency overlaps the element
   * and x loads. */
  unsigned dst[RPT];
  T dterm[RPT], dvv[RPT];
#pragma unroll
  for (int k = 0; k < RPT; k++) {
    const unsigned row = t + k * 256u;
    dst[k] = (row >> log2S) < ns ? (s0 << log2S) + row : ~0u;
    if (a.rmap && dst[k] != ~0u) dst[k] = a.rmap[dst[k]];
    dterm[k] = 0;
    dvv[k] = 0;
    if (dst[k] != ~0u) {
      /* the preconditioner's dot (r, M r) has x itself as the dot vector: one read of x */
      const T xd = a.rmap || dv == xv ? xs.at(dst[k]) : (T)0;
      if (a.rmap) dterm[k] = elt<T, E>::get(static_cast<const E *>(a.rdiag), dst[k]) * xd;
      if (a.dparts) dvv[k] = dv == xv ? xd : dv[dst[k]];
    }
  }
  /* the element range comes from two (scalar) loads, so the element loads need not wait for
   * the offsets table, which only phase 2 reads */
  const unsigned E0 = a.sp[s0], E1 = a.sp[s0 + ns], n = E1 - E0;
  for (unsigned q = t; q <= ns; q += 256) soff[q] = a.sp[s0 + q];   /* ns + 1 <= spg + 1 */
  /* products are stored at prod[e - base], base = E0 rounded down to VEC elements, so the
   * VEC-aligned body [A0, A1) of the range is loaded and stored as whole vectors; the head
   * [E0, min(A0, E1)) and tail [A1, E1) (fewer than VEC elements each) one by one */
  const unsigned base = E0 & ~(VEC - 1);
  const bool fits = n + (E0 - base) <= cap;
  T acc[RPT];
#pragma unroll
  for (int k = 0; k < RPT; k++) acc[k] = 0;
  if (fits) {
    const unsigned A0 = (E0 + VEC - 1) & ~(VEC - 1);
    const unsigned A1 = max(A0, E1 & ~(VEC - 1));
    const unsigned hend = min(A0, E1);
    if (E0 + t < hend) {
      const unsigned e = E0 + t;
      prod[e - base] = elt<T, E>::get(el, e) * xs.at(dev::col_load(cl + e));
    }
    if (A1 + t < E1 && A1 >= hend) {
      const unsigned e = A1 + t;
      prod[e - base] = elt<T, E>::get(el, e) * xs.at(dev::col_load(cl + e));
    }
    const unsigned nv = (A1 - A0) / VEC;
    /* rounds of 256 vectors this workgroup needs (uniform), up to U per step: the element
     * and column loads of all rounds of a step, then all their x gathers, are issued before
     * the first wait; a thread past the end of a round repeats its first vector (in range)
     * and stores nothing for it */
    const unsigned rounds = (nv + 255) / 256;
    for (unsigned r0 = 0; r0 < rounds; r0 += U) {
      const unsigned v0 = t + r0 * 256;
      if (v0 >= nv) break;
      T ev[U][VEC], xg[U][VEC];
      unsigned cv[U][VEC];
#pragma unroll
      for (unsigned u = 0; u < U; u++)
        if (r0 + u < rounds) {
          const unsigned vi = v0 + u * 256 < nv ? v0 + u * 256 : v0;
          dev::elt_vec<T>(el, A0 + vi * VEC, ev[u]);
          dev::col_vec(cl, A0 + vi * VEC, cv[u]);
        }
#pragma unroll
      for (unsigned u = 0; u < U; u++)
        if (r0 + u < rounds)
#pragma unroll
          for (unsigned q = 0; q < VEC; q++) xg[u][q] = xs.at(cv[u][q]);
#pragma unroll
      for (unsigned u = 0; u < U; u++)
        if (r0 + u < rounds && v0 + u * 256 < nv) {
          const unsigned o = A0 - base + (v0 + u * 256) * VEC;
#pragma unroll
          for (unsigned q = 0; q < VEC; q++) prod[o + q] = ev[u][q] * xg[u][q];
        }
    }
  }
  __syncthreads();   /* offsets table and products */
  if (fits) {
#pragma unroll
    for (int k = 0; k < RPT; k++) {
      const unsigned row = t + k * 256u, i = row >> log2S, r = row & (S - 1);
      if (i < ns) {
        const unsigned o = soff[i] - base, nk = (soff[i + 1] - soff[i]) >> log2S;
        for (unsigned q = 0; q < nk; q++) acc[k] += prod[o + (q << log2S) + r];
      }
    }
  } else {
#pragma unroll
    for (int k = 0; k < RPT; k++) {
      const unsigned row = t + k * 256u, i = row >> log2S, r = row & (S - 1);
      if (i < ns) {
        const unsigned o = soff[i], nk = (soff[i + 1] - o) >> log2S;
        for (unsigned q = 0; q < nk; q++) {
          /* product rounded, then added, as through LDS: the result does not depend on
           * whether the group fits */
#pragma clang fp contract(off)
          const unsigned e = o + (q << log2S) + r;
          const T pr = elt<T, E>::get(el, e) * xs.at(dev::col_load(cl + e));
          acc[k] += pr;
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < RPT; k++)
    if (a.rmap) acc[k] = dterm[k] + acc[k];
#pragma unroll
  for (int k = 0; k < RPT; k++)
    if (dst[k] != ~0u) yv[dst[k]] = acc[k];
  if (a.dparts) {
    __shared__ double red[4];
    double d[1] = {0.0};
#pragma unroll
    for (int k = 0; k < RPT; k++)
      if (dst[k] != ~0u) d[0] += (double)dvv[k] * (double)acc[k];
    dev::block_sum<256, 1>(d, red);
    if (t == 0) a.dparts[bid] = d[0];
  }
}

template<typename T, typename E, int RPT>
__global__ void __launch_bounds__(256) slell_stream_kernel(spmv_args a, unsigned cap)
{
  stream_body<T, E, RPT>(a, cap, blockIdx.x);
}

/* LDS bytes of a streamed-mapping workgroup with a product buffer of cap elements */
template<typename T> constexpr size_t stream_lds_bytes(unsigned cap, unsigned rpt, unsigned log2S)
{
  return ((size_t)cap * sizeof(T) + 4 * (size_t)(((256u * rpt) >> log2S) + 1) + 15) / 16 * 16;
}

/* Rows per thread (one for FP64 vectors, two for FP32, or rpt0 when given; ISPM_STREAM_RPT
 * overrides) and the product capacity: the layout's largest group span when known (64 KiB
 * of LDS at most), else 8 (FP64) or 4 (FP32) products per row. */
template<typename T>
void stream_shape(const layout_hints *h, unsigned log2S, unsigned &rpt, unsigned &cap, unsigned rpt0 = 0)
{
  rpt = rpt0 ? rpt0 : sizeof(T) == 8 ? 1u : 2u;
  if (const char *e = std::getenv("ISPM_STREAM_RPT")) {
    const unsigned r = (unsigned)std::atoi(e);
    if ((r == 1 || r == 2 || r == 4) && (256u * r) >= (1u << log2S)) rpt = r;
  }
  const unsigned k = rpt == 1 ? 0u : rpt == 2 ? 1u : 2u;
  cap = h && h->stream_span[k] ? (h->stream_span[k] + 3) & ~3u : (sizeof(T) == 4 ? 1024u : 2048u) * rpt;
  const size_t soff = 4 * (size_t)(((256u * rpt) >> log2S) + 1) + 16;
  if ((size_t)cap * sizeof(T) + soff > 65536) cap = (unsigned)((65536 - soff) / sizeof(T)) & ~3u;
}

template<typename T, typename E>
int launch_stream(const ispm_slell_desc *A, const void *x, void *y, const unsigned *skip,
                  const void *dotv = nullptr, double *dparts = nullptr, unsigned *dgrid = nullptr,
                  const layout_hints *hints = nullptr)
{
  spmv_args a;
  a.skip = skip;
  a.sp = A->slice_ptr;
  a.cols = A->cols;
  a.elms = A->elms;
  a.x = x;
  a.y = y;
  a.ldx = a.ldy = 0;
  a.n_slices = A->n_slices;
  a.log2S = (unsigned)(31 - __builtin_clz(A->slice_height));
  a.n_cols = (unsigned)A->n_cols;
  a.spb = 1;
  unsigned rpt, cap;
  stream_shape<T>(hints, a.log2S, rpt, cap);
  const size_t lds = stream_lds_bytes<T>(cap, rpt, a.log2S);
  const unsigned spg = (256u * rpt) >> a.log2S;
  const unsigned groups = (A->n_slices + spg - 1) / spg;
  a.per_xcd = (groups + 7) / 8;
  if (dparts && dgrid && a.per_xcd * 8 <= kMaxDotParts) {
    a.dotv = dotv;
    a.dparts = dparts;
    *dgrid = a.per_xcd * 8;
  }
  const dim3 g(a.per_xcd * 8), b(256);
  if (rpt == 1) hipLaunchKernelGGL((slell_stream_kernel<T, E, 1>), g, b, lds, stream(), a, cap);
  else if (rpt == 2) hipLaunchKernelGGL((slell_stream_kernel<T, E, 2>), g, b, lds, stream(), a, cap);
  else hipLaunchKernelGGL((slell_stream_kernel<T, E, 4>), g, b, lds, stream(), a, cap);
  ISPM_HIP(hipGetLastError());
  return ISPM_OK;
}

/* Streamed mapping for k = 2..8 right-hand sides (SpMM of the block solver): a 256-thread
 * workgroup takes 256 consecutive rows; phase 1 copies their contiguous element range (values
 * converted to T, and column indices) into LDS with coalesced vector loads, phase 2 gives each
 * thread one row and adds, plane by plane in storage order, value times x[col] for every
 * column of X: one pass over the matrix serves all k vectors.  Groups whose range exceeds
 * cap elements read their rows directly. */
template<typename T, typename E, int K>
__global__ void __launch_bounds__(256) slell_stream_spmm_kernel(spmv_args a, unsigned cap)
{
  constexpr unsigned VEC = dev::evec<E>::n;
  extern __shared__ __align__(16) unsigned char spmm_lds[];
  T *lval = reinterpret_cast<T *>(spmm_lds);                   /* cap values */
  unsigned *lcol = reinterpret_cast<unsigned *>(lval + cap);   /* cap columns */
  unsigned *soff = lcol + cap;                                  /* spg + 1 offsets */
  const unsigned log2S = a.log2S, S = 1u << log2S, spg = 256u >> log2S, t = threadIdx.x;
  const E *__restrict__ el = static_cast<const E *>(a.elms);
  const unsigned *__restrict__ cl = a.cols;
  const T *__restrict__ xv = static_cast<const T *>(a.x);
  T *__restrict__ yv = static_cast<T *>(a.y);
  const unsigned ncols = a.n_cols;
  const unsigned group = (blockIdx.x & 7u) * a.per_xcd + (blockIdx.x >> 3);
  const unsigned s0 = group * spg;
  if (s0 >= a.n_slices) return;   /* uniform over the workgroup */
  const unsigned ns = min(spg, a.n_slices - s0);
  const unsigned E0 = a.sp[s0], E1 = a.sp[s0 + ns], n = E1 - E0;
  for (unsigned q = t; q <= ns; q += 256) soff[q] = a.sp[s0 + q];
  const unsigned base = E0 & ~(VEC - 1);
  const bool fits = n + (E0 - base) <= cap;
  if (fits) {
    const unsigned A0 = (E0 + VEC - 1) & ~(VEC - 1);
    const unsigned A1 = max(A0, E1 & ~(VEC - 1));
    const unsigned hend = min(A0, E1);
    if (E0 + t < hend) {
      lval[E0 + t - base] = elt<T, E>::get(el, E0 + t);
      lcol[E0 + t - base] = dev::col_load(cl + E0 + t);
    }
    if (A1 + t < E1 && A1 >= hend) {
      lval[A1 + t - base] = elt<T, E>::get(el, A1 + t);
      lcol[A1 + t - base] = dev::col_load(cl + A1 + t);
    }
    for (unsigned v = t; v < (A1 - A0) / VEC; v += 256) {
      T ev[VEC];
      unsigned cv[VEC];
      dev::elt_vec<T>(el, A0 + v * VEC, ev);
      dev::col_vec(cl, A0 + v * VEC, cv);
      const unsigned o = A0 - base + v * VEC;
#pragma unroll
      for (unsigned q = 0; q < VEC; q++) { lval[o + q] = ev[q]; lcol[o + q] = cv[q]; }
    }
  }
  __syncthreads();
  const unsigned i = t >> log2S, r = t & (S - 1);
  if (i >= ns) return;
  T acc[K];
#pragma unroll
  for (int j = 0; j < K; j++) acc[j] = 0;
  const unsigned nk = (soff[i + 1] - soff[i]) >> log2S;
  for (unsigned q = 0; q < nk; q++) {
    /* product rounded, then added: the streamed SpMV's arithmetic, so column j of the result
     * is bit-identical to the single-vector product */
#pragma clang fp contract(off)
    const unsigned e = soff[i] + (q << log2S) + r;
    const T v = fits ? lval[e - base] : elt<T, E>::get(el, e);
    const unsigned c = fits ? lcol[e - base] : dev::col_load(cl + e);
#pragma unroll
    for (int j = 0; j < K; j++) {
      const T pr = v * xload(xv + j * a.ldx, c, ncols);
      acc[j] += pr;
    }
  }
  const size_t row = ((size_t)s0 << log2S) + t;
#pragma unroll
  for (int j = 0; j < K; j++) yv[row + j * a.ldy] = acc[j];
}

template<typename T, typename E, int K>
int launch_stream_spmm(const ispm_slell_desc *A, const void *X, size_t ldx, void *Y, size_t ldy,
                       const layout_hints *hints)
{
  spmv_args a;
  a.sp = A->slice_ptr;
  a.cols = A->cols;
  a.elms = A->elms;
  a.x = X;
  a.y = Y;
  a.ldx = ldx;
  a.ldy = ldy;
  a.n_slices = A->n_slices;
  a.log2S = (unsigned)(31 - __builtin_clz(A->slice_height));
  a.n_cols = (unsigned)A->n_cols;
  a.spb = 1;
  a.skip = nullptr;
  unsigned cap = hints && hints->stream_span[0] ? (hints->stream_span[0] + 3) & ~3u : 2048u;
  const size_t soffb = 4 * (size_t)((256u >> a.log2S) + 1) + 16;
  if ((size_t)cap * (sizeof(T) + 4) + soffb > 65536) cap = (unsigned)((65536 - soffb) / (sizeof(T) + 4)) & ~3u;
  const size_t lds = ((size_t)cap * (sizeof(T) + 4) + soffb + 15) / 16 * 16;
  const unsigned spg = 256u >> a.log2S;
  const unsigned groups = (A->n_slices + spg - 1) / spg;
  a.per_xcd = (groups + 7) / 8;
  hipLaunchKernelGGL((slell_stream_spmm_kernel<T, E, K>), dim3(a.per_xcd * 8), dim3(256), lds, stream(), a, cap);
  ISPM_HIP(hipGetLastError());
  return ISPM_OK;
}

template<int K>
int launch_stream_spmm_types(const ispm_slell_desc *A, int vt, const void *X, size_t ldx, void *Y,
                             size_t ldy, const layout_hints *hints)
{
  const int et = A->elm_type;
  if (vt == ISPM_F64 && et == ISPM_F64) return launch_stream_spmm<double, double, K>(A, X, ldx, Y, ldy, hints);
  if (vt == ISPM_F64 && et == ISPM_F32) return launch_stream_spmm<double, float, K>(A, X, ldx, Y, ldy, hints);
  if (vt == ISPM_F32 && et == ISPM_F32) return launch_stream_spmm<float, float, K>(A, X, ldx, Y, ldy, hints);
  if (vt == ISPM_F32 && et == ISPM_BF16) return launch_stream_spmm<float, bf16_t, K>(A, X, ldx, Y, ldy, hints);
  return fail(ISPM_ERR_UNSUPPORTED, "spmm: unsupported (vector, element) types (%d, %d)", vt, et);
}

inline int launch_stream_types(const ispm_slell_desc *A, int vt, const void *x, void *y,
                               const unsigned *skip, const void *dotv = nullptr,
                               double *dparts = nullptr, unsigned *dgrid = nullptr,
                               const layout_hints *h = nullptr)
{
  const int et = A->elm_type;
  if (vt == ISPM_F64 && et == ISPM_F64) return launch_stream<double, double>(A, x, y, skip, dotv, dparts, dgrid, h);
  if (vt == ISPM_F64 && et == ISPM_F32) return launch_stream<double, float>(A, x, y, skip, dotv, dparts, dgrid, h);
  if (vt == ISPM_F32 && et == ISPM_F32) return launch_stream<float, float>(A, x, y, skip, dotv, dparts, dgrid, h);
  if (vt == ISPM_F32 && et == ISPM_BF16) return launch_stream<float, bf16_t>(A, x, y, skip, dotv, dparts, dgrid, h);
  return fail(ISPM_ERR_UNSUPPORTED, "spmv: unsupported (vector, element) types (%d, %d)", vt, et);
}

/* Split-diagonal product in one launch: workgroups [0, G2) are the streamed mapping over the
 * layout of the rows with off-diagonal entries (each such row's value is its diagonal term
 * plus its sum, written once), workgroups [G2, G2 + G1) the diagonal part y = d .* x of
 * every other row (a bit per row in rmask marks the rows the first group writes).  The two
 * roles touch disjoint rows, so they run side by side: the latency-bound remainder overlaps
 * the bandwidth-bound diagonal stream.  Each diagonal workgroup thread takes VEC rows per
 * step: a vector load of d (nontemporal, read once), vector loads of x (and the dot
 * product's vector) and a vector store of y when VL (x, y, dotv aligned); fused dot
 * partials go to dparts[block] for both roles.
 *
 * AX: the product of x - alpha q (CG's residual update r - alpha q, then z = M r, in one
 * pass).  The diagonal role computes the updated vector for every row, stores it to xout
 * and sums its squares into rrparts[block]; the remainder role computes the entries it
 * gathers the same way (one fused multiply-add each) and stores 0 there.  x is only read,
 * so no workgroup sees another's update. */
struct split_args {
  const void *d;
  const unsigned *rmask;
  unsigned n;        /* rows, a multiple of VEC */
  unsigned G2;       /* remainder workgroups (a multiple of 8, or 0) */
  double *dparts;    /* fused dot: G2 + G1 partials, or null */
  void *xout;        /* AX: x - alpha q, n rows */
  double *rrparts;   /* AX: G2 + G1 partials of (x - alpha q, x - alpha q) */
};

template<typename T, typename E, bool VL, bool AX>
__device__ __forceinline__ void diag_body(const spmv_args &a, const split_args &sa, unsigned bid, unsigned nb)
{
  __shared__ double red[8];
  if (a.skip && *a.skip) return;
  constexpr unsigned V = dev::evec<E>::n;
  constexpr int NS = AX ? 2 : 1;
  typedef T tv __attribute__((ext_vector_type(V)));
  const E *__restrict__ d = static_cast<const E *>(sa.d);
  const T *__restrict__ x = static_cast<const T *>(a.x);
  const T *__restrict__ dv = static_cast<const T *>(a.dotv);
  const T *__restrict__ qv = static_cast<const T *>(a.axq);
  T *__restrict__ xo = static_cast<T *>(sa.xout);
  T *__restrict__ y = static_cast<T *>(a.y);
  static_assert(VL || !AX, "the residual update runs on whole vectors");
  const T al = AX ? (T)*a.axa : (T)0;
  double acc[NS];
#pragma unroll
  for (int k = 0; k < NS; k++) acc[k] = 0.0;
  const unsigned nv = sa.n / V;
  for (unsigned i = bid * 256u + threadIdx.x; i < nv; i += nb * 256u) {
    T dd[V], xx[V], yy[V];
    dev::elt_vec<T>(d, i * V, dd);
    const unsigned bits = sa.rmask ? (sa.rmask[(i * V) >> 5] >> ((i * V) & 31u)) & ((1u << V) - 1u) : 0u;
    if (VL) {
      const tv xv = *reinterpret_cast<const tv *>(x + (size_t)i * V);
#pragma unroll
      for (unsigned q = 0; q < V; q++) xx[q] = xv[q];
      if (AX) {
        const tv qq = *reinterpret_cast<const tv *>(qv + (size_t)i * V);
        tv xn;
#pragma unroll
        for (unsigned q = 0; q < V; q++) {
          xx[q] = ax_update(xx[q], qq[q], al);
          xn[q] = xx[q];
          acc[NS - 1] += (double)xx[q] * (double)xx[q];
        }
        *reinterpret_cast<tv *>(xo + (size_t)i * V) = xn;
      }
    } else {
#pragma unroll
      for (unsigned q = 0; q < V; q++) xx[q] = x[(size_t)i * V + q];
    }
    tv yv;
#pragma unroll
    for (unsigned q = 0; q < V; q++) { yy[q] = dd[q] * xx[q]; yv[q] = yy[q]; }
    if (VL && !bits) *reinterpret_cast<tv *>(y + (size_t)i * V) = yv;
    else {
#pragma unroll
      for (unsigned q = 0; q < V; q++)
        if (!((bits >> q) & 1u)) y[(size_t)i * V + q] = yy[q];
    }
    if (sa.dparts) {
      /* the preconditioner's dot (r, M r) has x itself as the dot vector: no second read */
      const bool dx = dv == x;
#pragma unroll
      for (unsigned q = 0; q < V; q++)
        if (!((bits >> q) & 1u)) acc[0] += (double)(dx ? xx[q] : dv[(size_t)i * V + q]) * (double)yy[q];
    }
  }
  if (sa.dparts) {
    dev::block_sum<256, NS>(acc, red);
    if (threadIdx.x == 0) {
      sa.dparts[sa.G2 + bid] = acc[0];
      if (AX) sa.rrparts[sa.G2 + bid] = acc[NS - 1];
    }
  }
}

template<typename T, typename E, int RPT, bool VL, bool AX>
__global__ void __launch_bounds__(256) split_kernel(spmv_args a, unsigned cap, split_args sa)
{
  if (blockIdx.x < sa.G2) {
    stream_body<T, E, RPT, AX>(a, cap, blockIdx.x);
    if (AX && threadIdx.x == 0) sa.rrparts[blockIdx.x] = 0.0;
  } else {
    diag_body<T, E, VL, AX>(a, sa, blockIdx.x - sa.G2, gridDim.x - sa.G2);
  }
}

template<typename T, typename E>
int launch_split(const ispm_slell_desc *R, const void *diag, const unsigned *rmap,
                 const unsigned *rmask, unsigned n, bool has_rem, const void *x, void *y,
                 const unsigned *skip, const void *dotv, double *dparts, unsigned *dgrid,
                 const layout_hints *hints, const split_update *up = nullptr)
{
  constexpr unsigned V = dev::evec<E>::n;
  spmv_args a;
  a.skip = skip;
  a.x = x;
  a.y = y;
  a.ldx = a.ldy = 0;
  a.spb = 1;
  a.rmap = rmap;
  a.rdiag = diag;
  unsigned rpt = 1, cap = 0, G2 = 0;
  size_t lds = 0;
  if (has_rem) {
    a.sp = R->slice_ptr;
    a.cols = R->cols;
    a.elms = R->elms;
    a.n_slices = R->n_slices;
    a.log2S = (unsigned)(31 - __builtin_clz(R->slice_height));
    a.n_cols = (unsigned)R->n_cols;
    /* 256-row remainder groups: shorter chains per workgroup than 512 rows, which the
     * streamed kernel takes for FP32 alone (192^3 FSAI product with the residual update:
     * 41.8 against 43.5 us per launch; 128 rows 41.1, 64 rows 45.6) */
    stream_shape<T>(hints, a.log2S, rpt, cap, 1);
    lds = stream_lds_bytes<T>(cap, rpt, a.log2S);
    const unsigned spg = (256u * rpt) >> a.log2S;
    const unsigned groups = (R->n_slices + spg - 1) / spg;
    a.per_xcd = (groups + 7) / 8;
    G2 = a.per_xcd * 8;
  } else {
    a.n_slices = 0;
    a.per_xcd = 0;
  }
  /* one vector of rows per thread and step, at most 8192 diagonal workgroups (grid-stride;
   * 2048 measured 1-2 us slower on the 192^3 FSAI product).  The remainder workgroups come
   * first: they are the longer chain (gathers), and measured slower after the diagonal ones
   * or interleaved with them. */
  const unsigned G1 = std::max(1u, std::min(8192u, (n / V + 255) / 256));
  split_args sa{diag, rmask, n, G2, nullptr, nullptr, nullptr};
  if (dotv && dparts && dgrid && G1 + G2 <= kMaxDotParts) {
    a.dotv = dotv;
    a.dparts = dparts;
    sa.dparts = dparts;
    *dgrid = G1 + G2;
  }
  const uintptr_t al = reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(y) |
                       (sa.dparts ? reinterpret_cast<uintptr_t>(dotv) : 0) |
                       (up ? reinterpret_cast<uintptr_t>(up->q) | reinterpret_cast<uintptr_t>(up->xout) : 0);
  const bool vl = al % (V * sizeof(T)) == 0;
  if (up) {
    /* the update needs the fused dot (its partials count is the caller's) and whole vectors */
    if (!sa.dparts || !vl || !up->q || !up->alpha || !up->xout || !up->rrparts)
      return fail(ISPM_ERR_INVALID, "spmv: split product with residual update needs aligned vectors and the fused dot");
    a.axq = up->q;
    a.axa = up->alpha;
    sa.xout = up->xout;
    sa.rrparts = up->rrparts;
  }
  const dim3 g(G1 + G2), b(256);
  hipStream_t st = stream();
#define ISPM_SPLIT_LAUNCH(R_, VL_, AX_) hipLaunchKernelGGL((split_kernel<T, E, R_, VL_, AX_>), g, b, lds, st, a, cap, sa)
  if (up) {
    if (rpt == 1) ISPM_SPLIT_LAUNCH(1, true, true);
    else if (rpt == 2) ISPM_SPLIT_LAUNCH(2, true, true);
    else ISPM_SPLIT_LAUNCH(4, true, true);
  } else if (rpt == 1) { if (vl) ISPM_SPLIT_LAUNCH(1, true, false); else ISPM_SPLIT_LAUNCH(1, false, false); }
  else if (rpt == 2) { if (vl) ISPM_SPLIT_LAUNCH(2, true, false); else ISPM_SPLIT_LAUNCH(2, false, false); }
  else { if (vl) ISPM_SPLIT_LAUNCH(4, true, false); else ISPM_SPLIT_LAUNCH(4, false, false); }
#undef ISPM_SPLIT_LAUNCH
  ISPM_HIP(hipGetLastError());
  return ISPM_OK;
}

template<typename T, typename E>
int launch_packed(const ispm_launch *p, const ispm_slell_desc *A, const void *x, void *y,
                  const unsigned *skip)
{
  spmv_args a;
  a.skip = skip;
  a.sp = A->slice_ptr;
  a.cols = A->cols;
  a.elms = A->elms;
  a.x = x;
  a.y = y;
  a.ldx = a.ldy = 0;
  a.n_slices = A->n_slices;
  a.log2S = (unsigned)(31 - __builtin_clz(A->slice_height));
  a.n_cols = (unsigned)A->n_cols;
  const unsigned wpb = 4;                          /* 256-thread workgroups */
  const unsigned spw = 64u >> a.log2S;
  const unsigned waves = ((A->n_slices + spw - 1) / spw + 3) / 4;   /* 4 row groups per wave */
  unsigned spb = 1;
  if (const char *e = std::getenv("ISPM_SPMV_SPB")) spb = (unsigned)std::max(1, std::atoi(e));
  a.spb = spb;
  const unsigned groups = (waves + wpb * spb - 1) / (wpb * spb);
  a.per_xcd = (groups + 7) / 8;
  const unsigned grid = a.per_xcd * 8;
  hipStream_t s = stream();
  if (A->hblock == 2) hipLaunchKernelGGL((slell_packed_kernel<T, E, 2, false>), dim3(grid), dim3(64 * wpb), 0, s, a);
  else if (A->diags) hipLaunchKernelGGL((slell_packed_kernel<T, E, 1, true>), dim3(grid), dim3(64 * wpb), 0, s, a);
  else hipLaunchKernelGGL((slell_packed_kernel<T, E, 1, false>), dim3(grid), dim3(64 * wpb), 0, s, a);
  ISPM_HIP(hipGetLastError());
  return ISPM_OK;
}

inline int launch_packed_types(const ispm_launch *p, const ispm_slell_desc *A, int vt, const void *x,
                               void *y, const unsigned *skip)
{
  const int et = A->elm_type;
  if (vt == ISPM_F64 && et == ISPM_F64) return launch_packed<double, double>(p, A, x, y, skip);
  if (vt == ISPM_F64 && et == ISPM_F32) return launch_packed<double, float>(p, A, x, y, skip);
  if (vt == ISPM_F32 && et == ISPM_F32) return launch_packed<float, float>(p, A, x, y, skip);
  if (vt == ISPM_F32 && et == ISPM_BF16) return launch_packed<float, bf16_t>(p, A, x, y, skip);
  return fail(ISPM_ERR_UNSUPPORTED, "spmv: unsupported (vector, element) types (%d, %d)", vt, et);
}

template<typename T, typename E, int H, int K>
int launch_tv(const ispm_launch *p, const ispm_slell_desc *A, const void *x, size_t ldx,
              void *y, size_t ldy, const unsigned *skip = nullptr)
{
  spmv_args a;
  a.skip = skip;
  a.sp = A->slice_ptr;
  a.cols = A->cols;
  a.elms = A->elms;
  a.x = x;
  a.y = y;
  a.ldx = ldx;
  a.ldy = ldy;
  a.n_slices = A->n_slices;
  a.log2S = A->var_height ? 0u : (unsigned)(31 - __builtin_clz(A->slice_height));
  a.n_cols = (unsigned)A->n_cols;
  /* slices per workgroup: keep about 64 workgroups of work per CU in flight and let each
   * workgroup walk consecutive slices (amortises wave launch for small slices) */
  unsigned spb = 1;
  if (const char *e = std::getenv("ISPM_SPMV_SPB")) spb = (unsigned)std::max(1, std::atoi(e));
  else if (p->N == 64 && A->var_height) spb = A->n_slices > (1u << 16) ? 8u : 1u;
  a.spb = spb;
  const unsigned groups = (A->n_slices + spb - 1) / spb;
  a.per_xcd = (groups + 7) / 8;
  const unsigned N = p->N;
  const unsigned grid = a.per_xcd * 8;
  /* LDS is needed only when partial sums of one row live in different waves. */
  size_t lds = 0;
  bool need = N > 64;
  if (!A->var_height) {
    const unsigned S = A->slice_height, R = N * H < S ? N * H : S, g = R / H;
    need = N * H > S && N > (g > 64 ? g : 64);
  }
  if (need) lds = sizeof(T) * N * H * K;
  if (lds > 65536) return fail(ISPM_ERR_UNSUPPORTED, "spmv: %zu bytes of LDS exceed 64 KiB", lds);
  hipStream_t s = stream();
  const bool V = A->var_height != 0, Dg = A->diags != 0;
  const unsigned W = A->hblock;
  if (W == 2) {
    if (V) hipLaunchKernelGGL((slell_kernel<T, E, H, 2, true, false, K>), dim3(grid), dim3(N), lds, s, a);
    else   hipLaunchKernelGGL((slell_kernel<T, E, H, 2, false, false, K>), dim3(grid), dim3(N), lds, s, a);
  } else if (Dg) {
    if (V) hipLaunchKernelGGL((slell_kernel<T, E, H, 1, true, true, K>), dim3(grid), dim3(N), lds, s, a);
    else   hipLaunchKernelGGL((slell_kernel<T, E, H, 1, false, true, K>), dim3(grid), dim3(N), lds, s, a);
  } else {
    if (V) hipLaunchKernelGGL((slell_kernel<T, E, H, 1, true, false, K>), dim3(grid), dim3(N), lds, s, a);
    else   hipLaunchKernelGGL((slell_kernel<T, E, H, 1, false, false, K>), dim3(grid), dim3(N), lds, s, a);
  }
  ISPM_HIP(hipGetLastError());
  return ISPM_OK;
}

template<typename T, typename E, int K>
int launch_h(const ispm_launch *p, const ispm_slell_desc *A, const void *x, size_t ldx,
             void *y, size_t ldy, unsigned H, const unsigned *skip = nullptr)
{
  if (H == 2) return launch_tv<T, E, 2, K>(p, A, x, ldx, y, ldy, skip);
  return launch_tv<T, E, 1, K>(p, A, x, ldx, y, ldy, skip);
}

template<int K, int HFIX>
int launch_types_h(const ispm_launch *p, const ispm_slell_desc *A, int vt, const void *x,
                   size_t ldx, void *y, size_t ldy)
{
  const int et = A->elm_type;
  if (vt == ISPM_F64 && et == ISPM_F64) return launch_tv<double, double, HFIX, K>(p, A, x, ldx, y, ldy);
  if (vt == ISPM_F64 && et == ISPM_F32) return launch_tv<double, float, HFIX, K>(p, A, x, ldx, y, ldy);
  if (vt == ISPM_F32 && et == ISPM_F32) return launch_tv<float, float, HFIX, K>(p, A, x, ldx, y, ldy);
  if (vt == ISPM_F32 && et == ISPM_BF16) return launch_tv<float, bf16_t, HFIX, K>(p, A, x, ldx, y, ldy);
  return fail(ISPM_ERR_UNSUPPORTED, "spmm: unsupported (vector, element) types (%d, %d)", vt, et);
}

template<int K>
int launch_types(const ispm_launch *p, const ispm_slell_desc *A, int vt, const void *x,
                 size_t ldx, void *y, size_t ldy, unsigned H, const unsigned *skip = nullptr)
{
  const int et = A->elm_type;
  if (vt == ISPM_F64 && et == ISPM_F64) return launch_h<double, double, K>(p, A, x, ldx, y, ldy, H, skip);
  if (vt == ISPM_F64 && et == ISPM_F32) return launch_h<double, float, K>(p, A, x, ldx, y, ldy, H, skip);
  if (vt == ISPM_F32 && et == ISPM_F32) return launch_h<float, float, K>(p, A, x, ldx, y, ldy, H, skip);
  if (vt == ISPM_F32 && et == ISPM_BF16) return launch_h<float, bf16_t, K>(p, A, x, ldx, y, ldy, H, skip);
  return fail(ISPM_ERR_UNSUPPORTED, "spmv: unsupported (vector, element) types (%d, %d)", vt, et);
}

inline bool pow2(unsigned v) { return v && !(v & (v - 1)); }

/* streamed mapping: requested by the launch (mode 2), fixed heights up to 256 rows without
 * hblock pairs or diagonal planes */
inline bool use_stream(const ispm_launch *p, const ispm_slell_desc *A)
{
  /* the kernel reads elements and columns as 16-byte vectors */
  const bool aligned = (reinterpret_cast<uintptr_t>(A->elms) % 16) == 0 &&
                       (reinterpret_cast<uintptr_t>(A->cols) % 16) == 0;
  return p->mode == 2 && !A->var_height && A->slice_height <= 256 && A->slice_height >= 1 &&
         A->hblock == 1 && !A->diags && aligned;
}

/* packed mapping: requested by the launch (mode 1; also mode 2 on layouts the streamed
 * mapping does not take), fixed-height slices below 64 rows */
inline bool use_packed(const ispm_launch *p, const ispm_slell_desc *A)
{
  return (p->mode == 1 || p->mode == 2) && !A->var_height && A->slice_height < 64 &&
         A->slice_height >= 1;
}

inline int check_args(const ispm_launch *p, const ispm_slell_desc *A, int vt, const void *x, void *y,
               unsigned &H)
{
  if (!p || !A || !x || !y) return fail(ISPM_ERR_INVALID, "spmv: null argument");
  /* reference validity rule (cuda/spmv/kernel-iface.h:4) and supported sets
   * (cuda/spmv/dispatch.h:127-181): otherwise nothing is launched */
  if (!(p->N >= p->S / (p->H ? p->H : 1) && p->S >= p->H))
    return fail(ISPM_ERR_UNSUPPORTED, "spmv: invalid launch parameters N=%u S=%u H=%u", p->N, p->S, p->H);
  if (!pow2(p->N) || p->N < 64 || p->N > 1024 || (p->H != 1 && p->H != 2))
    return fail(ISPM_ERR_UNSUPPORTED, "spmv: unsupported launch N=%u H=%u", p->N, p->H);
  if (p->mode < 0 || p->mode > 2)
    return fail(ISPM_ERR_UNSUPPORTED, "spmv: launch mode %d unsupported", p->mode);
  if (A->hblock != 1 && A->hblock != 2)
    return fail(ISPM_ERR_UNSUPPORTED, "spmv: hblock %u unsupported", A->hblock);
  if (A->hblock == 2 && A->diags)
    return fail(ISPM_ERR_INVALID, "spmv: hblock and diagonal slices are exclusive");
  if (!A->var_height && !pow2(A->slice_height))
    return fail(ISPM_ERR_UNSUPPORTED, "spmv: slice height %u is not a power of two", A->slice_height);
  if (A->n_cols <= 0 && A->n_slices) return fail(ISPM_ERR_INVALID, "spmv: n_cols must be positive");
  (void)vt;
  H = p->H;
  /* two-row vector stores need 2-element alignment of y (and x is never vector-loaded) */
  const size_t ts = type_size(vt);
  if (H == 2 && (reinterpret_cast<uintptr_t>(y) % (2 * ts))) H = 1;
  if (H == 2 && !A->var_height && A->slice_height < 2) H = 1;
  return ISPM_OK;
}

}  // namespace spmvk
}  // namespace ispm

#endif
