/* batched.hip -- many independent small SPD systems (FP32) solved together with Jacobi-
 * preconditioned CG.  No counterpart in the reference.
 *
 * Layout: systems back to back, each renumbered in reverse Cuthill-McKee order (so that the
 * 64 rows of a slice gather from a few cache lines of the vector; b is read and x written
 * in the caller's numbering) and padded to a multiple of 64 rows (padding rows are identity
 * rows whose right-hand side is 0), held as 64-row ELL slices (slot (r, k) of slice
 * s at slice_off[s] + r + 64 k, width slice_w[s]) with FP32 values and system-LOCAL column
 * indices, 16-bit when every system has fewer than 65536 rows (6 bytes per stored entry
 * instead of 8; the matrix is most of the bytes an iteration moves).  A system's slices are
 * grouped into segments of up to 16 slices (1024 rows): the unit of work of a workgroup and
 * of one partial sum.
 *
 * Iteration: single-reduction CG (Chronopoulos-Gear, the library's PCG1 recurrences):
 *   upd : p = z + b p, s = w + b s, x += a p, r -= a s, z = D^-1 r, partials (r, z), (r, r)
 *   spmv: w = A z, partial (z, w)
 *   fin : per system: sums of its segment partials in segment order, convergence test on
 *         the recursive residual, b = g'/g, a = g'/(d - b g'/a)
 *   list: the segments of the systems still iterating, as one compacted list
 * Every launch is over the segment list, so converged systems cost nothing, and each XCD
 * takes a contiguous range of the list so that a system's vectors stay in that XCD's L2 while
 * its slices gather from them.  Iterations run as hipGraph replays of 8 iterations; the host
 * reads the number of systems still iterating after each replay.  Vectors in FP32, sums and
 * scalars in FP64.  A system's arithmetic never depends on the others: solving it alone or in
 * a batch gives the same bits. */
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <atomic>
#include <numeric>
#include <thread>
#include <vector>
#include "device.h"
#include "internal.h"

namespace ispm {
namespace {

constexpr int kSeg = 16;            /* slices per segment */
constexpr int kNT = 256;            /* threads per workgroup: 4 waves */
constexpr int kIterPerGraph = 8;
constexpr unsigned kMaxWindow = 16384;   /* z entries a segment may stage in LDS (64 KiB) */

struct sys_desc {
  long long row0;   /* first (padded) row */
  int slice0, n_slices;
  int seg0, n_segs;
  int n;            /* rows */
  int pad;
};

struct sys_state {
  double g, a, b;   /* (r, z), alpha, beta */
  double bb, rr;
  int it, done;
};

/* the list the row kernels walk: segments of the active systems, in system order */
struct seg_list {
  int n_act;        /* active systems */
  int total;        /* active segments */
};

__device__ __forceinline__ void locate(const int *__restrict__ act_sys, const int *__restrict__ act_off,
                                       int n_act, int t, int &sys, int &local)
{
  int lo = 0, hi = n_act - 1;
  while (lo < hi) {               /* last j with act_off[j] <= t */
    const int mid = (lo + hi + 1) >> 1;
    if (act_off[mid] <= t) lo = mid; else hi = mid - 1;
  }
  sys = act_sys[lo];
  local = t - act_off[lo];
}

/* this workgroup's share of the list: each XCD a contiguous range (dispatch deals
 * workgroups round-robin over the 8 XCDs) */
__device__ __forceinline__ void xcd_range(int total, int &t0, int &t1, int &step)
{
  const int xcd = blockIdx.x & 7, per = (int)(gridDim.x >> 3);
  const long long lo = (long long)total * xcd / 8, hi = (long long)total * (xcd + 1) / 8;
  t0 = (int)lo + (int)(blockIdx.x >> 3);
  t1 = (int)hi;
  step = per;
}

template<int NS> __device__ __forceinline__ void seg_store(double (&v)[NS], double *parts, int seg, int n_segs_total)
{
  __shared__ double lds[NS * (kNT / 64)];
  dev::block_sum<kNT, NS>(v, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NS; q++) parts[(size_t)q * n_segs_total + seg] = v[q];
  }
}

/* x = 0, r = b (taken in the solver's row order: row g holds b[perm[g]]), z = D^-1 b,
 * p = s = 0; partials (r, z), (b, b) */
__global__ void __launch_bounds__(kNT)
k_init(const sys_desc *__restrict__ sys, const seg_list *__restrict__ L, const int *__restrict__ act_sys,
       const int *__restrict__ act_off, const float *__restrict__ b, const int *__restrict__ perm,
       const float *__restrict__ dinv, float *__restrict__ x, float *__restrict__ r, float *__restrict__ z,
       float *__restrict__ p, float *__restrict__ s, double *parts, int n_segs_total)
{
  const int total = L->total, n_act = L->n_act;
  int t0, t1, step;
  xcd_range(total, t0, t1, step);
  for (int t = t0; t < t1; t += step) {
    int id, ls;
    locate(act_sys, act_off, n_act, t, id, ls);
    const sys_desc d = sys[id];
    const long long base = d.row0 + (long long)ls * kSeg * 64;
    const int rows = min(kSeg, d.n_slices - ls * kSeg) * 64;
    double v[2] = {0.0, 0.0};
    for (int i = threadIdx.x; i < rows; i += kNT) {
      const long long g = base + i;
      const float bv = b[perm[g]], zv = bv * dinv[g];
      x[g] = 0.f;
      r[g] = bv;
      z[g] = zv;
      p[g] = 0.f;
      s[g] = 0.f;
      v[0] += (double)bv * zv;
      v[1] += (double)bv * bv;
    }
    seg_store<2>(v, parts, d.seg0 + ls, n_segs_total);
  }
}

/* p = z + b p, s = w + b s, x += a p, r -= a s, z = D^-1 r; partials (r, z), (r, r) */
__global__ void __launch_bounds__(kNT)
k_upd(const sys_desc *__restrict__ sys, const sys_state *__restrict__ st, const seg_list *__restrict__ L,
      const int *__restrict__ act_sys, const int *__restrict__ act_off, const float *__restrict__ dinv,
      const float *__restrict__ w, float *__restrict__ x, float *__restrict__ r, float *__restrict__ z,
      float *__restrict__ p, float *__restrict__ s, double *parts, int n_segs_total)
{
  const int total = L->total, n_act = L->n_act;
  int t0, t1, step;
  xcd_range(total, t0, t1, step);
  for (int t = t0; t < t1; t += step) {
    int id, ls;
    locate(act_sys, act_off, n_act, t, id, ls);
    const sys_desc d = sys[id];
    const float a = (float)st[id].a, be = (float)st[id].b;
    const long long base = d.row0 + (long long)ls * kSeg * 64;
    const int rows = min(kSeg, d.n_slices - ls * kSeg) * 64;
    double v[2] = {0.0, 0.0};
    for (int i = threadIdx.x * 4; i < rows; i += kNT * 4) {
      const long long g = base + i;
      const float4 zv = *reinterpret_cast<const float4 *>(z + g);
      const float4 wv = *reinterpret_cast<const float4 *>(w + g);
      float4 pv = *reinterpret_cast<const float4 *>(p + g);
      float4 sv = *reinterpret_cast<const float4 *>(s + g);
      float4 xv = *reinterpret_cast<const float4 *>(x + g);
      float4 rv = *reinterpret_cast<const float4 *>(r + g);
      const float4 dv = *reinterpret_cast<const float4 *>(dinv + g);
      float4 zn;
#define ISPM_UPD(c)                                   \
      pv.c = zv.c + be * pv.c;                        \
      sv.c = wv.c + be * sv.c;                        \
      xv.c = xv.c + a * pv.c;                         \
      rv.c = rv.c - a * sv.c;                         \
      zn.c = rv.c * dv.c;                             \
      v[0] += (double)rv.c * zn.c;                    \
      v[1] += (double)rv.c * rv.c;
      ISPM_UPD(x) ISPM_UPD(y) ISPM_UPD(z) ISPM_UPD(w)
#undef ISPM_UPD
      *reinterpret_cast<float4 *>(p + g) = pv;
      *reinterpret_cast<float4 *>(s + g) = sv;
      *reinterpret_cast<float4 *>(x + g) = xv;
      *reinterpret_cast<float4 *>(r + g) = rv;
      *reinterpret_cast<float4 *>(z + g) = zn;
    }
    seg_store<2>(v, parts, d.seg0 + ls, n_segs_total);
  }
}

/* w = A z for one segment's slices (one slice per wave at a time, one row per lane), and
 * the partial (z, w).  The z entries the segment's columns reach (its window, a few
 * thousand for a reverse Cuthill-McKee numbering) are first copied to LDS with 16-byte
 * loads; the gathers then read LDS, whose banks serve a wave's scattered reads in a few
 * cycles where the vector memory pipeline takes one cycle per distinct cache line.  A
 * segment whose window exceeds the LDS size (wlen 0) gathers from global memory. */
template<typename C, bool LDS>
__device__ __forceinline__ float slice_rows(const float *__restrict__ vals, const C *__restrict__ cols,
                                            unsigned o, int width, const float *__restrict__ zsrc, int base)
{
  /* planes in chunks of 16, all element and column loads of a chunk issued before its
   * gathers; past the slice's width a lane re-reads the last plane (a cache hit) and adds 0 */
  constexpr int CH = 16;
  float acc = 0.f;
  for (int k = 0; k < width; k += CH) {
    float ev[CH];
    unsigned cv[CH];
#pragma unroll
    for (int q = 0; q < CH; q++) {
      const unsigned kk = (unsigned)min(k + q, width - 1);
      ev[q] = __builtin_nontemporal_load(vals + o + 64u * kk);
      cv[q] = __builtin_nontemporal_load(cols + o + 64u * kk);
    }
    float xv[CH];
#pragma unroll
    for (int q = 0; q < CH; q++) xv[q] = zsrc[(int)cv[q] - base];
#pragma unroll
    for (int q = 0; q < CH; q++) acc += (k + q < width ? ev[q] : 0.f) * xv[q];
  }
  return acc;
}

template<typename C>
__global__ void __launch_bounds__(kNT)
k_spmv(const sys_desc *__restrict__ sys, const seg_list *__restrict__ L, const int *__restrict__ act_sys,
       const int *__restrict__ act_off, const unsigned *__restrict__ soff, const unsigned char *__restrict__ sw,
       const C *__restrict__ cols, const float *__restrict__ vals, const int2 *__restrict__ win,
       const float *__restrict__ z, float *__restrict__ w, double *parts, int n_segs_total, int parts_q)
{
  extern __shared__ __align__(16) float zl[];
  const int total = L->total, n_act = L->n_act;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int t0, t1, step;
  xcd_range(total, t0, t1, step);
  for (int t = t0; t < t1; t += step) {
    int id, ls;
    locate(act_sys, act_off, n_act, t, id, ls);
    const sys_desc d = sys[id];
    const float *__restrict__ zs = z + d.row0;
    const int2 wn = win[d.seg0 + ls];          /* first local column (a multiple of 4), length */
    if (wn.y) {
      __syncthreads();                         /* the previous segment's reads are done */
      const float4 *src = reinterpret_cast<const float4 *>(zs + wn.x);
      for (int q = threadIdx.x; q < (wn.y >> 2); q += kNT) reinterpret_cast<float4 *>(zl)[q] = src[q];
      __syncthreads();
    }
    const int sl0 = ls * kSeg, sl1 = min(d.n_slices, sl0 + kSeg);
    double v[1] = {0.0};
    for (int sl = sl0 + wave; sl < sl1; sl += kNT / 64) {
      const int gs = d.slice0 + sl;
      const unsigned o = soff[gs] + lane;
      const int width = sw[gs];
      const int row = sl * 64 + lane;
      float acc, zr;
      if (wn.y) {
        acc = slice_rows<C, true>(vals, cols, o, width, zl, wn.x);
        zr = zl[row - wn.x];
      } else {
        acc = slice_rows<C, false>(vals, cols, o, width, zs, 0);
        zr = zs[row];
      }
      w[d.row0 + row] = acc;
      v[0] += (double)zr * acc;
    }
    seg_store<1>(v, parts + (size_t)parts_q * n_segs_total, d.seg0 + ls, n_segs_total);
  }
}

/* one wave per active system: its sums over its segments (in segment order, lanes then a
 * fixed butterfly), convergence, next scalars.  init != 0: the first step (g = (r, z),
 * bb = (b, b) from k_init; d = (z, w) from the first product). */
__global__ void __launch_bounds__(kNT)
k_fin(const sys_desc *__restrict__ sys, sys_state *__restrict__ st, const seg_list *__restrict__ L,
      const int *__restrict__ act_sys, const double *__restrict__ parts, int n_segs_total, double tol2,
      int maxit, int init)
{
  const int lane = threadIdx.x & 63;
  const int j = (int)(blockIdx.x * (kNT / 64) + (threadIdx.x >> 6));
  if (j >= L->n_act) return;
  const int id = act_sys[j];
  const sys_desc d = sys[id];
  double g = 0.0, q1 = 0.0, dz = 0.0;
  for (int k = lane; k < d.n_segs; k += 64) {
    g += parts[d.seg0 + k];
    q1 += parts[(size_t)n_segs_total + d.seg0 + k];
    dz += parts[2 * (size_t)n_segs_total + d.seg0 + k];
  }
  g = dev::wave_sum(g);
  q1 = dev::wave_sum(q1);
  dz = dev::wave_sum(dz);
  if (lane) return;
  sys_state s = st[id];
  if (init) {
    s.bb = q1;
    s.rr = q1;
    s.it = 0;
    s.b = 0.0;
    s.g = g;
    s.a = dz != 0.0 ? g / dz : 0.0;
    s.done = !(q1 > 0.0) || q1 <= tol2 * q1 || maxit <= 0;
  } else {
    s.it++;
    s.rr = q1;
    const double b = g / s.g;
    const double den = dz - b * g / s.a;
    s.b = b;
    s.a = den != 0.0 ? g / den : 0.0;
    s.g = g;
    s.done = q1 <= tol2 * s.bb || s.it >= maxit || !(den != 0.0) || !(q1 == q1);
  }
  st[id] = s;
}

/* one workgroup: the segment list of the systems not done (all: of every system), in
 * system order */
__global__ void __launch_bounds__(1024)
k_list(const sys_desc *__restrict__ sys, const sys_state *__restrict__ st, int n_sys, seg_list *L,
       int *__restrict__ act_sys, int *__restrict__ act_off, int all)
{
  __shared__ unsigned long long wtot[16];
  __shared__ int carry_sys, carry_seg;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (threadIdx.x == 0) { carry_sys = 0; carry_seg = 0; }
  __syncthreads();
  for (int base = 0; base < n_sys; base += 1024) {
    const int j = base + (int)threadIdx.x;
    const int act = (j < n_sys && (all || !st[j].done)) ? 1 : 0;
    const int segs = act ? sys[j].n_segs : 0;
    /* inclusive scans of (act, segs) packed in one 64-bit value within the wave */
    unsigned long long v = ((unsigned long long)act << 40) | (unsigned long long)segs;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const unsigned long long u = (unsigned long long)__shfl_up((long long)v, o);
      if (lane >= o) v += u;
    }
    if (lane == 63) wtot[wv] = v;
    __syncthreads();
    unsigned long long pre = 0;
    for (int q = 0; q < wv; q++) pre += wtot[q];
    const unsigned long long incl = pre + v;
    const int a_incl = (int)(incl >> 40), s_incl = (int)(incl & 0xffffffffffull);
    if (act) {
      act_sys[carry_sys + a_incl - 1] = j;
      act_off[carry_sys + a_incl - 1] = carry_seg + s_incl - segs;
    }
    __syncthreads();
    if (threadIdx.x == 1023) { carry_sys += a_incl; carry_seg += s_incl; }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    L->n_act = carry_sys;
    L->total = carry_seg;
  }
}

/* the solution back in the caller's row order */
__global__ void k_unperm(const float *__restrict__ xi, const int *__restrict__ perm, float *__restrict__ x,
                         long long n)
{
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n) x[perm[g]] = xi[g];
}

__global__ void k_results(const sys_state *__restrict__ st, int n_sys, int *iters, double *rel)
{
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_sys) return;
  iters[j] = st[j].it;
  rel[j] = st[j].bb > 0 ? sqrt(st[j].rr / st[j].bb) : 0.0;
}

/* ---- host: reverse Cuthill-McKee order of one system ---- */

/* breadth-first levels from `start` over the unvisited-by-mark vertices; returns the last
 * level's vertices in `last` and the number of levels */
int bfs_levels(const std::vector<int> &rp, const std::vector<int> &ci, int start, std::vector<int> &mark,
               int stamp, std::vector<int> &last)
{
  std::vector<int> cur{start}, next;
  mark[start] = stamp;
  int levels = 0;
  while (!cur.empty()) {
    levels++;
    last = cur;
    next.clear();
    for (int v : cur)
      for (int e = rp[v]; e < rp[v + 1]; e++)
        if (mark[ci[e]] != stamp) { mark[ci[e]] = stamp; next.push_back(ci[e]); }
    cur.swap(next);
  }
  return levels;
}

/* order[new] = old: Cuthill-McKee from a pseudo-peripheral vertex of each connected
 * component (neighbours taken by increasing degree), reversed */
void rcm_order(const std::vector<int> &rp, const std::vector<int> &ci, int n, std::vector<int> &order)
{
  std::vector<int> deg(n), mark(n, -1), done(n, 0), last;
  for (int v = 0; v < n; v++) deg[v] = rp[v + 1] - rp[v];
  std::vector<int> byDeg(n);
  std::iota(byDeg.begin(), byDeg.end(), 0);
  std::stable_sort(byDeg.begin(), byDeg.end(), [&](int a, int b) { return deg[a] < deg[b]; });
  order.clear();
  order.reserve(n);
  int stamp = 0;
  std::vector<int> nb;
  for (int seed : byDeg) {
    if (done[seed]) continue;
    /* pseudo-peripheral start: move to a least-degree vertex of the last level while the
     * number of levels grows (a few rounds) */
    int start = seed, levels = bfs_levels(rp, ci, start, mark, stamp++, last);
    for (int round = 0; round < 4; round++) {
      int cand = last[0];
      for (int v : last) if (deg[v] < deg[cand]) cand = v;
      const int l2 = bfs_levels(rp, ci, cand, mark, stamp++, last);
      if (l2 <= levels) break;
      start = cand;
      levels = l2;
    }
    size_t head = order.size();
    order.push_back(start);
    done[start] = 1;
    while (head < order.size()) {
      const int v = order[head++];
      nb.clear();
      for (int e = rp[v]; e < rp[v + 1]; e++)
        if (!done[ci[e]]) { done[ci[e]] = 1; nb.push_back(ci[e]); }
      std::stable_sort(nb.begin(), nb.end(), [&](int a, int b) { return deg[a] < deg[b]; });
      order.insert(order.end(), nb.begin(), nb.end());
    }
  }
  std::reverse(order.begin(), order.end());
}

/* one system renumbered: rows in RCM order (by length within each segment), columns
 * renumbered and sorted */
struct sys_csr {
  std::vector<int> rp, ci, perm;   /* perm[new] = old (local) */
  std::vector<float> v;
  int status = 0;
};

void build_system(const int *row_ptr, const int *cols, const float *vals, long long grow, int n, sys_csr &out)
{
  std::vector<int> rp(n + 1), ci;
  ci.reserve((size_t)(row_ptr[grow + n] - row_ptr[grow]));
  rp[0] = 0;
  for (int r = 0; r < n; r++) {
    for (int e = row_ptr[grow + r]; e < row_ptr[grow + r + 1]; e++) {
      const long long c = cols[e] - grow;
      if (c < 0 || c >= n) { out.status = 1; return; }
      ci.push_back((int)c);
    }
    rp[r + 1] = (int)ci.size();
  }
  std::vector<int> order;
  rcm_order(rp, ci, n, order);
  /* rows of one segment sorted by length (longest first): its 64-row slices then hold rows
   * of nearly equal length and little ELL padding, and the segment reaches the same columns */
  for (int b0 = 0; b0 < n; b0 += kSeg * 64) {
    const int b1 = std::min(n, b0 + kSeg * 64);
    std::stable_sort(order.begin() + b0, order.begin() + b1,
                     [&](int a, int b) { return rp[a + 1] - rp[a] > rp[b + 1] - rp[b]; });
  }
  std::vector<int> inv(n);
  for (int i = 0; i < n; i++) inv[order[i]] = i;
  out.perm = order;
  out.rp.assign(n + 1, 0);
  out.ci.resize(ci.size());
  out.v.resize(ci.size());
  std::vector<std::pair<int, float>> row;
  for (int i = 0; i < n; i++) {
    const int o = order[i];
    row.clear();
    for (int e = rp[o]; e < rp[o + 1]; e++) row.push_back({inv[ci[e]], vals[row_ptr[grow] + e]});
    std::stable_sort(row.begin(), row.end(), [](const std::pair<int, float> &a, const std::pair<int, float> &b) { return a.first < b.first; });
    const int b0 = out.rp[i];
    for (size_t k = 0; k < row.size(); k++) { out.ci[b0 + k] = row[k].first; out.v[b0 + k] = row[k].second; }
    out.rp[i + 1] = b0 + (int)row.size();
  }
}

}  // namespace
}  // namespace ispm

using namespace ispm;

struct ispm_batch {
  int n_sys = 0, n_slices = 0, n_segs = 0;
  long long rows_padded = 0;
  long long matrix_bytes = 0;   /* values, columns, slice offsets and widths */
  bool wide_cols = false;
  std::vector<long long> row_off;
  sys_desc *d_sys = nullptr;
  sys_state *d_st = nullptr;
  seg_list *d_list = nullptr;
  int *d_act_sys = nullptr, *d_act_off = nullptr;
  unsigned *d_soff = nullptr;
  unsigned char *d_sw = nullptr;
  void *d_cols = nullptr;
  float *d_vals = nullptr, *d_dinv = nullptr, *d_xi = nullptr;
  int *d_perm = nullptr;
  int *d_win = nullptr;     /* per segment: first local column of its z window, length */
  size_t lds = 0;           /* LDS of the product: the largest window staged */
  float *d_r = nullptr, *d_z = nullptr, *d_p = nullptr, *d_s = nullptr, *d_w = nullptr;
  double *d_parts = nullptr;
  int *d_iters = nullptr;
  double *d_rel = nullptr;
  seg_list *h_list = nullptr;   /* mapped host copy of the list header */
  int grid = 1;
  /* the captured iterations and the arguments they were captured with */
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double g_tol = -1;
  int g_maxit = -1;
};

namespace {

int launch_init(ispm_batch *B, const float *b, double tol2, int maxit)
{
  hipStream_t sm = stream();
  const int ns = B->n_sys;
  hipLaunchKernelGGL(k_list, dim3(1), dim3(1024), 0, sm, (const sys_desc *)B->d_sys, (const sys_state *)B->d_st,
                     ns, B->d_list, B->d_act_sys, B->d_act_off, 1);
  hipLaunchKernelGGL(k_init, dim3(B->grid), dim3(kNT), 0, sm, (const sys_desc *)B->d_sys,
                     (const seg_list *)B->d_list, (const int *)B->d_act_sys, (const int *)B->d_act_off, b,
                     (const int *)B->d_perm, (const float *)B->d_dinv, B->d_xi, B->d_r, B->d_z, B->d_p, B->d_s,
                     B->d_parts, B->n_segs);
  if (B->wide_cols)
    hipLaunchKernelGGL(k_spmv<unsigned>, dim3(B->grid), dim3(kNT), B->lds, sm, (const sys_desc *)B->d_sys,
                       (const seg_list *)B->d_list, (const int *)B->d_act_sys, (const int *)B->d_act_off,
                       (const unsigned *)B->d_soff, (const unsigned char *)B->d_sw, (const unsigned *)B->d_cols,
                       (const float *)B->d_vals, (const int2 *)B->d_win, (const float *)B->d_z, B->d_w,
                       B->d_parts, B->n_segs, 2);
  else
    hipLaunchKernelGGL(k_spmv<unsigned short>, dim3(B->grid), dim3(kNT), B->lds, sm, (const sys_desc *)B->d_sys,
                       (const seg_list *)B->d_list, (const int *)B->d_act_sys, (const int *)B->d_act_off,
                       (const unsigned *)B->d_soff, (const unsigned char *)B->d_sw,
                       (const unsigned short *)B->d_cols, (const float *)B->d_vals, (const int2 *)B->d_win,
                       (const float *)B->d_z, B->d_w, B->d_parts, B->n_segs, 2);
  hipLaunchKernelGGL(k_fin, dim3((ns + 3) / 4), dim3(kNT), 0, sm, (const sys_desc *)B->d_sys, B->d_st,
                     (const seg_list *)B->d_list, (const int *)B->d_act_sys, (const double *)B->d_parts,
                     B->n_segs, tol2, maxit, 1);
  hipLaunchKernelGGL(k_list, dim3(1), dim3(1024), 0, sm, (const sys_desc *)B->d_sys, (const sys_state *)B->d_st,
                     ns, B->d_list, B->d_act_sys, B->d_act_off, 0);
  ISPM_HIP(hipGetLastError());
  return ISPM_OK;
}

int launch_iteration(ispm_batch *B, double tol2, int maxit)
{
  hipStream_t sm = stream();
  const int ns = B->n_sys;
  hipLaunchKernelGGL(k_upd, dim3(B->grid), dim3(kNT), 0, sm, (const sys_desc *)B->d_sys,
                     (const sys_state *)B->d_st, (const seg_list *)B->d_list, (const int *)B->d_act_sys,
                     (const int *)B->d_act_off, (const float *)B->d_dinv, (const float *)B->d_w, B->d_xi, B->d_r,
                     B->d_z, B->d_p, B->d_s, B->d_parts, B->n_segs);
  if (B->wide_cols)
    hipLaunchKernelGGL(k_spmv<unsigned>, dim3(B->grid), dim3(kNT), B->lds, sm, (const sys_desc *)B->d_sys,
                       (const seg_list *)B->d_list, (const int *)B->d_act_sys, (const int *)B->d_act_off,
                       (const unsigned *)B->d_soff, (const unsigned char *)B->d_sw, (const unsigned *)B->d_cols,
                       (const float *)B->d_vals, (const int2 *)B->d_win, (const float *)B->d_z, B->d_w,
                       B->d_parts, B->n_segs, 2);
  else
    hipLaunchKernelGGL(k_spmv<unsigned short>, dim3(B->grid), dim3(kNT), B->lds, sm, (const sys_desc *)B->d_sys,
                       (const seg_list *)B->d_list, (const int *)B->d_act_sys, (const int *)B->d_act_off,
                       (const unsigned *)B->d_soff, (const unsigned char *)B->d_sw,
                       (const unsigned short *)B->d_cols, (const float *)B->d_vals, (const int2 *)B->d_win,
                       (const float *)B->d_z, B->d_w, B->d_parts, B->n_segs, 2);
  hipLaunchKernelGGL(k_fin, dim3((ns + 3) / 4), dim3(kNT), 0, sm, (const sys_desc *)B->d_sys, B->d_st,
                     (const seg_list *)B->d_list, (const int *)B->d_act_sys, (const double *)B->d_parts,
                     B->n_segs, tol2, maxit, 0);
  hipLaunchKernelGGL(k_list, dim3(1), dim3(1024), 0, sm, (const sys_desc *)B->d_sys, (const sys_state *)B->d_st,
                     ns, B->d_list, B->d_act_sys, B->d_act_off, 0);
  ISPM_HIP(hipGetLastError());
  return ISPM_OK;
}

int read_list(ispm_batch *B, seg_list &out)
{
  ISPM_HIP(hipMemcpyAsync(B->h_list, B->d_list, sizeof(seg_list), hipMemcpyDeviceToHost, stream()));
  ISPM_HIP(hipStreamSynchronize(stream()));
  out = *B->h_list;
  return ISPM_OK;
}

}  // namespace

extern "C" {

int ispm_batch_create(ispm_batch **out, int n_sys, const int *sys_rows, const int *row_ptr,
                      const int *cols, const float *vals)
{
  if (!out || n_sys <= 0 || !sys_rows || !row_ptr || !cols || !vals)
    return fail(ISPM_ERR_INVALID, "batch_create: bad argument");
  *out = nullptr;
  int st = ensure_init();
  if (st) return st;
  ispm_batch *B = new ispm_batch;
  B->n_sys = n_sys;
  std::vector<sys_desc> sd(n_sys);
  std::vector<unsigned> soff;
  std::vector<unsigned char> sw;
  std::vector<unsigned> hc;
  std::vector<float> hv, hd;
  int max_n = 0;
  for (int i = 0; i < n_sys; i++) max_n = std::max(max_n, sys_rows[i]);
  B->wide_cols = ((long long)max_n + 63) / 64 * 64 > 65536;
  /* each system in reverse Cuthill-McKee order (neighbours get nearby numbers, so the 64
   * rows of a slice gather from a few cache lines of z), computed on host threads */
  std::vector<long long> grow0(n_sys + 1, 0);
  for (int i = 0; i < n_sys; i++) {
    if (sys_rows[i] <= 0) { delete B; return fail(ISPM_ERR_INVALID, "batch_create: empty system %d", i); }
    grow0[i + 1] = grow0[i] + sys_rows[i];
  }
  std::vector<sys_csr> sc(n_sys);
  {
    const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; t++)
      pool.emplace_back([&]() {
        for (int i; (i = next++) < n_sys;) build_system(row_ptr, cols, vals, grow0[i], sys_rows[i], sc[i]);
      });
    for (auto &th : pool) th.join();
  }
  for (int i = 0; i < n_sys; i++)
    if (sc[i].status) { delete B; return fail(ISPM_ERR_INVALID, "batch_create: system %d couples to another system", i); }
  std::vector<int> hperm, hwin;
  unsigned max_win = 0;
  long long prow = 0;   /* padded row */
  int seg = 0;
  for (int i = 0; i < n_sys; i++) {
    const int n = sys_rows[i];
    const sys_csr &C = sc[i];
    const int ns = (n + 63) / 64;
    const int nseg = (ns + kSeg - 1) / kSeg;
    sd[i] = sys_desc{prow, (int)soff.size(), ns, seg, nseg, n, 0};
    seg += nseg;
    B->row_off.push_back(prow);
    hd.resize(prow + (long long)ns * 64, 1.0f);
    for (int rl = 0; rl < ns * 64; rl++) hperm.push_back((int)(prow + (rl < n ? C.perm[rl] : rl)));
    for (int s = 0; s < ns; s++) {
      int w = 0;
      for (int rl = s * 64; rl < std::min(n, s * 64 + 64); rl++) w = std::max(w, C.rp[rl + 1] - C.rp[rl]);
      if (w == 0) w = 1;
      if (w > 255) { delete B; return fail(ISPM_ERR_UNSUPPORTED, "batch_create: system %d has a row of more than 255 entries", i); }
      const size_t base = hv.size();
      if (base + 64ull * w > 0xffffffffull) { delete B; return fail(ISPM_ERR_UNSUPPORTED, "batch_create: more than 2^32 stored entries"); }
      soff.push_back((unsigned)base);
      sw.push_back((unsigned char)w);
      hv.resize(base + 64 * (size_t)w, 0.0f);
      hc.resize(base + 64 * (size_t)w, 0u);
      for (int l = 0; l < 64; l++) {
        const int rl = s * 64 + l;
        if (rl >= n) {                    /* padding row: identity, column = itself */
          for (int k = 0; k < w; k++) hc[base + l + 64 * (size_t)k] = (unsigned)rl;
          hv[base + l] = 1.0f;
          continue;
        }
        int k = 0;
        float diag = 0.f;
        for (int e = C.rp[rl]; e < C.rp[rl + 1]; e++, k++) {
          hc[base + l + 64 * (size_t)k] = (unsigned)C.ci[e];
          hv[base + l + 64 * (size_t)k] = C.v[e];
          if (C.ci[e] == rl) diag += C.v[e];
        }
        for (; k < w; k++) hc[base + l + 64 * (size_t)k] = (unsigned)rl;
        hd[prow + rl] = diag != 0.f ? 1.0f / diag : 1.0f;
      }
    }
    /* each segment's z window: the columns its stored entries reach (padding included) */
    for (int q = 0; q < nseg; q++) {
      const int s0 = sd[i].slice0 + q * kSeg, s1 = sd[i].slice0 + std::min(ns, (q + 1) * kSeg);
      const size_t e0 = soff[s0], e1 = soff[s1 - 1] + 64 * (size_t)sw[s1 - 1];
      unsigned lo = ~0u, hi = 0;
      for (size_t e = e0; e < e1; e++) { lo = std::min(lo, hc[e]); hi = std::max(hi, hc[e] + 1); }
      /* also the segment's own rows (the dot product reads z there) */
      lo = std::min(lo, (unsigned)(q * kSeg * 64));
      hi = std::max(hi, (unsigned)(std::min(ns, (q + 1) * kSeg) * 64));
      const unsigned lo4 = lo & ~3u, hi4 = (hi + 3) & ~3u;
      const unsigned len = hi4 - lo4 <= kMaxWindow ? hi4 - lo4 : 0u;
      hwin.push_back((int)lo4);
      hwin.push_back((int)len);
      max_win = std::max(max_win, len);
    }
    prow += (long long)ns * 64;
    sc[i] = sys_csr();   /* release the renumbered copy */
  }
  if (prow > 0x7fffffffLL) { delete B; return fail(ISPM_ERR_UNSUPPORTED, "batch_create: more than 2^31 padded rows"); }
  B->rows_padded = prow;
  B->n_slices = (int)soff.size();
  B->n_segs = seg;
  auto up = [&](void **dst, const void *src, size_t bytes) -> int {
    int e = ispm_malloc(dst, bytes ? bytes : 8);
    if (e) return e;
    return bytes ? ispm_memcpy(*dst, src, bytes, ISPM_H2D) : ISPM_OK;
  };
  std::vector<unsigned short> hc16;
  const void *colsrc = hc.data();
  size_t colbytes = sizeof(unsigned) * hc.size();
  if (!B->wide_cols) {
    hc16.assign(hc.begin(), hc.end());
    colsrc = hc16.data();
    colbytes = sizeof(unsigned short) * hc16.size();
  }
  const size_t vb = sizeof(float) * (size_t)prow;
  B->matrix_bytes = (long long)(colbytes + sizeof(float) * hv.size() + sizeof(unsigned) * soff.size() + sw.size());
  void *hl = nullptr;
  if ((st = up((void **)&B->d_sys, sd.data(), sizeof(sys_desc) * n_sys)) ||
      (st = up((void **)&B->d_soff, soff.data(), sizeof(unsigned) * soff.size())) ||
      (st = up((void **)&B->d_sw, sw.data(), sw.size())) ||
      (st = up(&B->d_cols, colsrc, colbytes)) ||
      (st = up((void **)&B->d_vals, hv.data(), sizeof(float) * hv.size())) ||
      (st = up((void **)&B->d_dinv, hd.data(), sizeof(float) * hd.size())) ||
      (st = up((void **)&B->d_perm, hperm.data(), sizeof(int) * hperm.size())) ||
      (st = up((void **)&B->d_win, hwin.data(), sizeof(int) * hwin.size())) ||
      (st = ispm_malloc((void **)&B->d_xi, vb)) ||
      (st = ispm_malloc((void **)&B->d_st, sizeof(sys_state) * n_sys)) ||
      (st = ispm_malloc((void **)&B->d_list, sizeof(seg_list))) ||
      (st = ispm_malloc((void **)&B->d_act_sys, sizeof(int) * n_sys)) ||
      (st = ispm_malloc((void **)&B->d_act_off, sizeof(int) * n_sys)) ||
      (st = ispm_malloc((void **)&B->d_r, vb)) || (st = ispm_malloc((void **)&B->d_z, vb)) ||
      (st = ispm_malloc((void **)&B->d_p, vb)) || (st = ispm_malloc((void **)&B->d_s, vb)) ||
      (st = ispm_malloc((void **)&B->d_w, vb)) ||
      (st = ispm_malloc((void **)&B->d_parts, sizeof(double) * 3 * (size_t)seg)) ||
      (st = ispm_malloc((void **)&B->d_iters, sizeof(int) * n_sys)) ||
      (st = ispm_malloc((void **)&B->d_rel, sizeof(double) * n_sys)) ||
      (st = ispm_host_alloc_mapped(&hl, sizeof(seg_list)))) {
    ispm_batch_destroy(B);
    return st;
  }
  B->h_list = static_cast<seg_list *>(hl);
  B->lds = sizeof(float) * (size_t)max_win;
  /* enough workgroups to fill the part (whole segments each), a multiple of the 8 XCDs */
  B->grid = std::max(8, std::min(num_cus() * 8, (seg + 7) / 8 * 8));
  *out = B;
  return ISPM_OK;
}

int ispm_batch_layout(const ispm_batch *B, long long *rows_padded, long long *row_offsets)
{
  if (!B) return fail(ISPM_ERR_INVALID, "batch_layout: null batch");
  if (rows_padded) *rows_padded = B->rows_padded;
  if (row_offsets) std::copy(B->row_off.begin(), B->row_off.end(), row_offsets);
  return ISPM_OK;
}

int ispm_batch_matrix_bytes(const ispm_batch *B, long long *bytes)
{
  if (!B || !bytes) return fail(ISPM_ERR_INVALID, "batch_matrix_bytes: null argument");
  *bytes = B->matrix_bytes;
  return ISPM_OK;
}

int ispm_batch_solve(ispm_batch *B, const float *b, float *x, double tol, int maxit,
                     int *iters, double *relres)
{
  if (!B || !b || !x) return fail(ISPM_ERR_INVALID, "batch_solve: null argument");
  if (!(tol >= 0)) return fail(ISPM_ERR_INVALID, "batch_solve: bad tolerance");
  if (maxit <= 0) maxit = 10000;
  const double tol2 = tol * tol;
  int st = launch_init(B, b, tol2, maxit);
  if (st) return st;
  hipStream_t sm = stream();
  const bool use_graph = !std::getenv("ISPM_NO_GRAPH");
  if (use_graph && (B->g_tol != tol || B->g_maxit != maxit)) {
    if (B->exec) { (void)hipGraphExecDestroy(B->exec); B->exec = nullptr; }
    if (B->graph) { (void)hipGraphDestroy(B->graph); B->graph = nullptr; }
    ISPM_HIP(hipStreamBeginCapture(sm, hipStreamCaptureModeThreadLocal));
    int e = ISPM_OK;
    for (int i = 0; i < kIterPerGraph && !e; i++) e = launch_iteration(B, tol2, maxit);
    hipGraph_t g = nullptr;
    hipError_t he = hipStreamEndCapture(sm, &g);
    if (e) { if (g) (void)hipGraphDestroy(g); return e; }
    if (he != hipSuccess) return hip_status(he, "hipStreamEndCapture");
    B->graph = g;
    ISPM_HIP(hipGraphInstantiate(&B->exec, g, nullptr, nullptr, 0));
    B->g_tol = tol;
    B->g_maxit = maxit;
  }
  seg_list L;
  if ((st = read_list(B, L))) return st;
  /* every replay advances each iterating system by up to kIterPerGraph steps; systems stop
   * at maxit, so at most maxit / kIterPerGraph + 1 replays */
  for (long long rep = 0; L.n_act > 0 && rep <= maxit / kIterPerGraph + 1; rep++) {
    if (use_graph) {
      ISPM_HIP(hipGraphLaunch(B->exec, sm));
    } else {
      for (int i = 0; i < kIterPerGraph; i++)
        if ((st = launch_iteration(B, tol2, maxit))) return st;
    }
    if ((st = read_list(B, L))) return st;
  }
  hipLaunchKernelGGL(k_unperm, dim3((unsigned)((B->rows_padded + 255) / 256)), dim3(256), 0, sm,
                     (const float *)B->d_xi, (const int *)B->d_perm, x, B->rows_padded);
  ISPM_HIP(hipGetLastError());
  if (iters || relres) {
    hipLaunchKernelGGL(k_results, dim3((B->n_sys + 255) / 256), dim3(256), 0, sm, (const sys_state *)B->d_st,
                       B->n_sys, B->d_iters, B->d_rel);
    ISPM_HIP(hipGetLastError());
    if (iters) ISPM_HIP(hipMemcpyAsync(iters, B->d_iters, sizeof(int) * B->n_sys, hipMemcpyDeviceToHost, sm));
    if (relres) ISPM_HIP(hipMemcpyAsync(relres, B->d_rel, sizeof(double) * B->n_sys, hipMemcpyDeviceToHost, sm));
    ISPM_HIP(hipStreamSynchronize(sm));
  }
  return ISPM_OK;
}

void ispm_batch_destroy(ispm_batch *B)
{
  if (!B) return;
  if (B->exec) (void)hipGraphExecDestroy(B->exec);
  if (B->graph) (void)hipGraphDestroy(B->graph);
  void *ps[] = {B->d_sys, B->d_st, B->d_list, B->d_act_sys, B->d_act_off, B->d_soff,
                B->d_sw, B->d_cols, B->d_vals, B->d_dinv, B->d_xi, B->d_perm, B->d_win, B->d_r, B->d_z, B->d_p, B->d_s, B->d_w,
                B->d_parts, B->d_iters, B->d_rel};
  for (void *p : ps) ispm_free(p);
  if (B->h_list) ispm_host_free(B->h_list);
  delete B;
}

}  // extern "C"
