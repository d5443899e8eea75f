/* plan.cpp -- C-ABI for host format conversion, device upload and SpMV plans.
 *
 * ispm_plan mirrors spmv_plan<T,E> (plan/plan.h:20-161 of the reference):
 *  - MIN_BYTES (no reference counterpart): the format with the fewest stored bytes.
 *  - DEFAULT / ESTIMATE (one kind, plan.h:74-75): the format comes from the guess heuristic
 *    (conversion.h:416-492);
 *    the launch shape is chosen for CDNA4 instead of the reference's fixed 512 threads:
 *    one row per thread and one wave (64 lanes) per 64 rows, so that every plane load of
 *    a wave is a single 512-byte (FP64) run and no cross-wave reduction is needed.
 *  - EXHAUSTIVE: the reference's search over S in {32..512}, V, D, hblock (plan.h:97-144)
 *    with the launch sweep retargeted to wave64 workgroups N in {64..1024}, H in {1,2};
 *    each candidate is timed as the median of several launches on rotating vectors, not one
 *    launch (plan.h:120-124), and the 65535-slice grid limit of sm_35 does not apply.
 *  - update_coefs: new values for the same pattern are written through the CSR-entry ->
 *    slot map recorded at conversion (replaces change_host_matrix + update_device_matrix_coefs,
 *    plan.h:39-56, without re-converting).
 */
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <thread>
#include <vector>
#include <sys/file.h>
#include "convert.h"
#include "internal.h"
#include "ispm.h"
#include "plan_internal.h"

namespace ispm {

static int fail_plain(int code, const char *msg);

size_t tsize(int t) { return t == ISPM_F64 ? 8 : t == ISPM_F32 ? 4 : t == ISPM_BF16 ? 2 : 0; }

static double csr_value(const ispm_csr *A, long long k)
{
  switch (A->elm_type) {
    case ISPM_F64: return static_cast<const double *>(A->elms)[k];
    case ISPM_F32: return static_cast<const float *>(A->elms)[k];
    default: {
      unsigned short u = static_cast<const unsigned short *>(A->elms)[k];
      unsigned v = (unsigned)u << 16;
      float f;
      std::memcpy(&f, &v, 4);
      return f;
    }
  }
}

unsigned short f32_to_bf16(float f)
{
  unsigned u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (unsigned short)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

static void store_value(void *dst, int et, size_t slot, double v)
{
  switch (et) {
    case ISPM_F64: static_cast<double *>(dst)[slot] = v; break;
    case ISPM_F32: static_cast<float *>(dst)[slot] = (float)v; break;
    default: static_cast<unsigned short *>(dst)[slot] = f32_to_bf16((float)v); break;
  }
}

int fill_values(const conv::layout &L, const ispm_csr *A, int et, void *elms)
{
  const long long nnz = A->row_ptr[A->n_rows];
  std::memset(elms, 0, L.n_elms * tsize(et));
  for (long long k = 0; k < nnz; k++) {
    long long s = L.map[k];
    if (s < 0) return -1;
    store_value(elms, et, (size_t)s, csr_value(A, k));
  }
  return 0;
}

/* The conversions assume what the reference's callers provide (conversion.h): row_ptr starts
 * at 0 and never decreases, columns lie in [0, n_cols) and are sorted within each row
 * (repeated columns allowed).  The diagonal encoder of an unsorted row would place entries
 * past the slice's planes, so every conversion checks this first (O(nnz)). */
static int check_csr(const ispm_csr *A)
{
  if (!A || !A->row_ptr) return fail_plain(ISPM_ERR_INVALID, "csr: null array");
  if (A->n_rows <= 0 || A->n_cols <= 0) return fail_plain(ISPM_ERR_INVALID, "csr: empty matrix");
  if ((!A->cols && A->row_ptr[A->n_rows] > 0) || (!A->elms && A->row_ptr[A->n_rows] > 0))
    return fail_plain(ISPM_ERR_INVALID, "csr: null array");
  if (A->elm_type != ISPM_F64 && A->elm_type != ISPM_F32 && A->elm_type != ISPM_BF16)
    return fail_plain(ISPM_ERR_INVALID, "csr: bad element type");
  const int *rp = A->row_ptr, *ci = A->cols;
  if (rp[0] != 0) return fail_plain(ISPM_ERR_INVALID, "csr: row_ptr[0] != 0");
  for (int r = 0; r < A->n_rows; r++) {
    if (rp[r + 1] < rp[r]) return fail_plain(ISPM_ERR_INVALID, "csr: row_ptr decreases");
    for (int k = rp[r]; k < rp[r + 1]; k++) {
      if (ci[k] < 0 || ci[k] >= A->n_cols) return fail_plain(ISPM_ERR_INVALID, "csr: column out of range");
      if (k > rp[r] && ci[k] < ci[k - 1]) return fail_plain(ISPM_ERR_INVALID, "csr: columns of a row not sorted");
    }
  }
  return ISPM_OK;
}

static int to_host(const conv::layout &L, const ispm_csr *A, int et, ispm_slell_host *out)
{
  std::memset(out, 0, sizeof *out);
  out->n_rows = L.n_rows;
  out->n_cols = L.n_cols;
  out->n_nz = L.n_nz;
  out->n_slices = L.n_slices;
  out->slice_height = L.S;
  out->hblock = L.W;
  out->var_height = L.V;
  out->diags = L.D;
  out->elm_type = et;
  out->slice_ptr_len = L.sp.size();
  out->cols_len = L.cols.size();
  out->elms_len = L.n_elms;
  out->slice_ptr = (unsigned *)std::malloc(sizeof(unsigned) * (L.sp.size() ? L.sp.size() : 1));
  out->cols = (unsigned *)std::malloc(sizeof(unsigned) * (L.cols.size() ? L.cols.size() : 1));
  out->elms = std::malloc(tsize(et) * (L.n_elms ? L.n_elms : 1));
  if (!out->slice_ptr || !out->cols || !out->elms) return fail_plain(ISPM_ERR_NOMEM, "host allocation failed");
  std::copy(L.sp.begin(), L.sp.end(), out->slice_ptr);
  std::copy(L.cols.begin(), L.cols.end(), out->cols);
  fill_values(L, A, et, out->elms);
  return ISPM_OK;
}

int upload_layout(const conv::layout &L, const void *elms, int et, ispm_slell_desc *d)
{
  std::memset(d, 0, sizeof *d);
  d->n_rows = L.n_rows;
  d->n_cols = L.n_cols;
  d->n_nz = L.n_nz;
  d->n_slices = L.n_slices;
  d->slice_height = L.S;
  d->hblock = L.W;
  d->var_height = L.V;
  d->diags = L.D;
  d->elm_type = et;
  d->slice_ptr_len = L.sp.size();
  d->cols_len = L.cols.size();
  d->elms_len = L.n_elms;
  void *sp = nullptr, *cl = nullptr, *el = nullptr;
  int st;
  if ((st = ispm_malloc(&sp, sizeof(unsigned) * L.sp.size()))) return st;
  if ((st = ispm_malloc(&cl, sizeof(unsigned) * (L.cols.size() ? L.cols.size() : 1)))) return st;
  if ((st = ispm_malloc(&el, tsize(et) * (L.n_elms ? L.n_elms : 1)))) return st;
  if ((st = ispm_memcpy(sp, L.sp.data(), sizeof(unsigned) * L.sp.size(), ISPM_H2D))) return st;
  if (!L.cols.empty() && (st = ispm_memcpy(cl, L.cols.data(), sizeof(unsigned) * L.cols.size(), ISPM_H2D))) return st;
  if (L.n_elms && (st = ispm_memcpy(el, elms, tsize(et) * L.n_elms, ISPM_H2D))) return st;
  d->slice_ptr = (const unsigned *)sp;
  d->cols = (const unsigned *)cl;
  d->elms = el;
  return ISPM_OK;
}

static int conv_status(int r)
{
  if (r == 0) return ISPM_OK;
  if (r == -3) return fail_plain(ISPM_ERR_UNSUPPORTED, "conversion: more than 2^32 stored elements");
  return fail_plain(ISPM_ERR_INVALID, "conversion failed");
}

/* ---------------------------------------------------------------- plans ------------- */

size_t stored_bytes(const ispm_slell_desc &d, int vt)
{
  return 4 * d.slice_ptr_len + 4 * d.cols_len + tsize(d.elm_type) * d.elms_len +
         tsize(vt) * ((size_t)d.n_cols + (size_t)d.n_rows);
}

size_t csr_bytes(const ispm_csr *A, int et, int vt)
{
  const size_t nnz = (size_t)A->row_ptr[A->n_rows];
  return 4 * ((size_t)A->n_rows + 1) + (4 + tsize(et)) * nnz + tsize(vt) * 2 * (size_t)A->n_rows;
}

/* CDNA4 default launch: one row per thread, wave-multiple workgroups covering a slice
 * (fixed height); one wave per slice for variable-height slices. */
ispm_launch default_launch(const conv::layout &L)
{
  ispm_launch p;
  p.H = 1;
  p.use_tex_sliceptr = p.use_tex_x = 0;
  p.mode = 0;
  if (!L.V && L.S < 64) {
    /* small fixed slices, 4 waves per workgroup: streamed elements for plain planes,
     * packed rows otherwise */
    p.N = 256;
    p.S = L.S;
    p.mode = (L.W == 1 && !L.D) ? 2 : 1;
    return p;
  }
  if (L.V) {
    /* variable-height slices are mostly small: one wave per slice, and each workgroup walks
     * several consecutive slices (see the launch in spmv_kernel.h) */
    p.N = 64;
    p.S = 64;
  } else {
    unsigned S = L.S;
    p.N = std::min(1024u, std::max(64u, S));
    p.S = S;
  }
  return p;
}

static int make_layout(const ispm_csr *A, int et, unsigned S, unsigned W, int V, int D,
                       conv::layout &L)
{
  const int min_height = (!V && S < 4) ? (int)S : 4;
  int r = conv::convert(A->row_ptr, A->n_rows, A->cols, A->n_cols, A->n_nz, (int)tsize(et),
                        (int)S, (int)W, V != 0, D != 0, min_height, L);
  return conv_status(r);
}

/* The diagonal list (positions relative to the slice's first row) that most diagonal slices
 * share, when it covers at least a quarter of them and its windows fit 64 KiB of LDS for
 * FP64 vectors, with one window per run of increasing positions less than S apart (the
 * windows of such neighbours overlap, and one window over both is shorter than two: the
 * near diagonals of a 27-point stencil, -n-1 .. n+1, share one); mask gets one bit per
 * slice that has exactly that list and no ELL planes. */
bool build_pattern(const conv::layout &L, diag_pattern &pat, std::vector<unsigned> &mask)
{
  if (!L.D || L.V || L.W != 1 || L.S < 64 || L.S > 512 || !L.n_slices) return false;
  const unsigned S = L.S, ns = L.n_slices;
  auto rel_of = [&](unsigned s, std::vector<long long> &r) {
    const unsigned *rec = &L.sp[3 * (size_t)s];
    const unsigned e0 = rec[0], c0 = rec[1], d0 = rec[2], e1 = rec[3], c1 = rec[4];
    if (c1 != c0) return false;
    const unsigned nsec = (e1 - e0) / S;
    if (nsec == 0 || nsec > (unsigned)kPatPlanes) return false;
    r.resize(nsec);
    for (unsigned d = 0; d < nsec; d++) r[d] = (long long)L.sp[d0 + d] - (long long)s * S;
    return true;
  };
  std::map<std::vector<long long>, unsigned> count;
  std::vector<long long> rel;
  for (unsigned s = 0; s < ns; s++)
    if (rel_of(s, rel)) count[rel]++;
  const std::vector<long long> *best = nullptr;
  unsigned bc = 0;
  for (const auto &kv : count)
    if (kv.second > bc) { bc = kv.second; best = &kv.first; }
  if (!best || 4ull * bc < ns) return false;
  const std::vector<long long> R = *best;
  pat = diag_pattern{};
  pat.n = (unsigned)R.size();
  unsigned used = 0;
  for (unsigned f = 0; f < pat.n;) {
    unsigned l = f;
    while (l + 1 < pat.n && R[l + 1] >= R[l] && R[l + 1] - R[l] < (long long)S) l++;
    const unsigned span = (unsigned)(R[l] - R[f]);
    pat.rrel[pat.nrun] = (int)(R[f] - (long long)S);
    pat.rlen[pat.nrun] = S + span;
    pat.rbase[pat.nrun] = used;
    for (unsigned d = f; d <= l; d++) pat.poff[d] = used + (unsigned)(R[d] - R[f]);
    used += S + span;
    pat.nrun++;
    f = l + 1;
  }
  pat.lds_len = used;
  if ((size_t)used * 8 > 65536) return false;
  for (unsigned d = 0; d < pat.n; d++)
    if (R[d] == (long long)S) pat.diag_plane1 = d + 1;
  mask.assign((ns + 31) / 32, 0u);
  for (unsigned s = 0; s < ns; s++)
    if (rel_of(s, rel) && rel == R) mask[s >> 5] |= 1u << (s & 31);
  return true;
}

/* device copy of a pattern's slice mask (null when the layout has no pattern) */
static int upload_pattern(const conv::layout &L, diag_pattern &pat, unsigned *&dev_uni, bool &has,
                          unsigned *n_slices = nullptr)
{
  std::vector<unsigned> mask;
  has = false;
  dev_uni = nullptr;
  if (!build_pattern(L, pat, mask)) return ISPM_OK;
  if (n_slices) {
    *n_slices = 0;
    for (unsigned w : mask) *n_slices += (unsigned)__builtin_popcount(w);
  }
  int st = ispm_malloc((void **)&dev_uni, mask.size() * sizeof(unsigned));
  if (st) return st;
  if ((st = ispm_memcpy(dev_uni, mask.data(), mask.size() * sizeof(unsigned), ISPM_H2D))) {
    ispm_free(dev_uni);
    dev_uni = nullptr;
    return st;
  }
  has = true;
  return ISPM_OK;
}

/* Streamed-mapping facts of a plain fixed-height layout: for groups of 256*RPT rows, the
 * largest element range from the group's first element rounded down to a multiple of 8 (the
 * kernel's LDS product buffer holds that range). */
void compute_hints(const conv::layout &L, layout_hints &h)
{
  std::memset(&h, 0, sizeof h);
  if (L.V || L.W != 1 || L.D || L.S == 0 || L.S > 256 || !L.n_slices || L.sp.size() < L.n_slices + 1)
    return;
  auto span_of = [&](unsigned rows) {
    const unsigned spg = std::max(1u, rows / L.S);
    unsigned span = 0;
    for (unsigned s0 = 0; s0 < L.n_slices; s0 += spg) {
      const unsigned s1 = std::min(L.n_slices, s0 + spg);
      span = std::max(span, L.sp[s1] - (L.sp[s0] & ~7u));
    }
    return span;
  };
  for (unsigned k = 0; k < 3; k++) h.stream_span[k] = span_of(256u << k);
  for (unsigned k = 0; k < 2; k++) h.stream_span_small[k] = L.S <= (64u << k) ? span_of(64u << k) : 0u;
}

/* x windows of hblock slices: the columns a slice's entries read span [lo, hi]; a slice is
 * windowed when the span fits 32 KiB of LDS in the plan's vector type and is at most twice
 * the slice's stored slots (staging it costs no more than the gathers it replaces).  Padding
 * slots (value 0) may name columns outside the span: the kernel clamps their index. */
int upload_xwin(const conv::layout &L, const ispm_csr *A, int vt, unsigned *&dev, layout_hints &h)
{
  h.xwin = nullptr;
  h.xwin_max = h.xwin_slices = 0;
  if (L.W != 2 || L.V || L.D || L.S < 64 || L.S > 512 || !L.n_slices) return ISPM_OK;
  /* records of 2 words: element offset, column offset */
  const unsigned cap = 32768u / (unsigned)tsize(vt), S = L.S, P = 2;
  std::vector<unsigned> w(2 * (size_t)L.n_slices, 0u);
  unsigned mx = 0, nw = 0;
  for (unsigned s = 0; s < L.n_slices; s++) {
    const unsigned r0 = s * S, r1 = std::min<unsigned>((unsigned)A->n_rows, r0 + S);
    if (r0 >= r1) continue;
    const int k0 = A->row_ptr[r0], k1 = A->row_ptr[r1];
    if (k1 <= k0) continue;
    int lo = A->cols[k0], hi = A->cols[k0];
    for (int k = k0; k < k1; k++) { lo = std::min(lo, A->cols[k]); hi = std::max(hi, A->cols[k]); }
    /* a pair plane reads columns 2c and 2c + 1: widen to whole pairs */
    lo &= ~1;
    hi = std::min(hi | 1, A->n_cols - 1);
    const unsigned len = (unsigned)(hi - lo + 1);
    const unsigned slots = L.sp[P * (s + 1)] - L.sp[P * s];
    if (len <= cap && len <= 2 * slots) {
      w[2 * (size_t)s] = (unsigned)lo;
      w[2 * (size_t)s + 1] = len;
      mx = std::max(mx, len);
      nw++;
    }
  }
  if (!nw) return ISPM_OK;
  int st = ispm_malloc(reinterpret_cast<void **>(&dev), w.size() * sizeof(unsigned));
  if (st) return st;
  if ((st = ispm_memcpy(dev, w.data(), w.size() * sizeof(unsigned), ISPM_H2D))) {
    ispm_free(dev);
    dev = nullptr;
    return st;
  }
  h.xwin = dev;
  h.xwin_max = mx;
  h.xwin_slices = nw;
  return ISPM_OK;
}

static int finish_plan(ispm_plan *p, const ispm_csr *A, unsigned flags)
{
  compute_hints(p->L, p->hints);
  p->elms.assign(p->L.n_elms * tsize(p->E) + 1, 0);
  if (fill_values(p->L, A, p->E, p->elms.data())) return fail_plain(ISPM_ERR_INVALID, "value map incomplete");
  p->stored = 0;
  p->csr = csr_bytes(A, p->E, p->T);
  if (!(flags & ISPM_SPMV_PLAN_NO_DEVICE)) {
    int st = upload_layout(p->L, p->elms.data(), p->E, &p->dev);
    if (st) return st;
    p->has_dev = true;
    if ((st = upload_pattern(p->L, p->pat, p->dev_uni, p->has_pat, &p->pat_slices))) return st;
    if ((st = upload_xwin(p->L, A, p->T, p->dev_xwin, p->hints))) return st;
  } else {
    std::memset(&p->dev, 0, sizeof p->dev);
    p->dev.n_rows = p->L.n_rows;
    p->dev.n_cols = p->L.n_cols;
    p->dev.slice_ptr_len = p->L.sp.size();
    p->dev.cols_len = p->L.cols.size();
    p->dev.elms_len = p->L.n_elms;
    p->dev.elm_type = p->E;
  }
  p->stored = stored_bytes(p->dev, p->T);
  return ISPM_OK;
}

static int time_candidate(int vt, const ispm_slell_desc &d, const ispm_launch &lp, void *x,
                          void *y, void *x2, void *y2, double &sec,
                          const diag_pattern *pat = nullptr, const unsigned *uni = nullptr,
                          const layout_hints *hints = nullptr)
{
  void *e0 = nullptr, *e1 = nullptr;
  int st;
  /* warm-up; also validity */
  if ((st = spmv_internal(&lp, &d, vt, x, y, nullptr, pat, uni, hints))) return st;
  if ((st = ispm_event_create(&e0)) || (st = ispm_event_create(&e1))) return st;
  const int reps = 5;
  std::vector<float> t;
  for (int r = 0; r < reps; r++) {
    ispm_event_record(e0);
    spmv_internal(&lp, &d, vt, (r & 1) ? x2 : x, (r & 1) ? y2 : y, nullptr, pat, uni, hints);
    ispm_event_record(e1);
    ispm_event_synchronize(e1);
    float ms = 0;
    ispm_event_elapsed_ms(e0, e1, &ms);
    t.push_back(ms);
  }
  ispm_event_destroy(e0);
  ispm_event_destroy(e1);
  std::sort(t.begin(), t.end());
  sec = 1e-3 * t[reps / 2];
  return ISPM_OK;
}

/* PLAN_ESTIMATE: smallest stored bytes.  Plain-ELL bytes are exact and cost O(n) per
 * candidate (row lengths only); the diagonal / hblock variant the sampled statistics
 * suggest is converted once and compared exactly. */
int estimate_format(const ispm_csr *A, int vt, int et, unsigned &bS, unsigned &bW, int &bV, int &bD)
{
  std::vector<int> rp;
  const int n = conv::pad_row_ptr(A->row_ptr, A->n_rows, rp);
  const long long nnz = rp[n];
  const double es = (double)tsize(et);
  double best = INFINITY;
  for (unsigned S = 1; S <= 512; S <<= 1)
    for (int V = 0; V <= (S >= 4 ? 1 : 0); V++) {
      std::vector<int> h;
      if (V) {
        if (conv::var_slice_heights(rp.data(), n, (int)S, nnz, 4, h)) continue;
      } else {
        h.assign(n / S, (int)S);
      }
      double slots = 0;
      int row = 0;
      for (int sz : h) {
        int w = 0;
        for (int r = row; r < row + sz; r++) w = std::max(w, rp[r + 1] - rp[r]);
        slots += (double)sz * w;
        row += sz;
      }
      const double bytes = slots * (4 + es) + 4.0 * (1 + V) * (h.size() + 1);
      if (bytes < best * 0.999) { best = bytes; bS = S; bV = V; bW = 1; bD = 0; }
    }
  conv::guess_info g;
  conv::guess_stats(A->row_ptr, A->n_rows, A->cols, (int)tsize(et), 1, g);
  if (g.diags || g.hblock > 1) {
    conv::layout L;
    if (!make_layout(A, et, bS, (unsigned)g.hblock, bV, g.diags ? 1 : 0, L)) {
      const double bytes = 4.0 * L.sp.size() + 4.0 * L.cols.size() + es * L.n_elms;
      if (bytes < best) { best = bytes; bW = g.hblock; bD = g.diags; }
    }
  }
  (void)vt;
  return std::isfinite(best) ? ISPM_OK : fail_plain(ISPM_ERR_INVALID, "estimate: no format");
}

int tune_exhaustive(const ispm_csr *A, int vt, int et, unsigned &bS, unsigned &bW, int &bV,
                    int &bD, ispm_launch &bl, double &best)
{
  best = INFINITY;
  const size_t nvec = ((size_t)A->n_rows + 511) / 512 * 512;
  const size_t ncols = std::max(nvec, (size_t)A->n_cols);
  void *x = nullptr, *y = nullptr, *x2 = nullptr, *y2 = nullptr;
  int st;
  if ((st = ispm_malloc(&x, tsize(vt) * ncols)) || (st = ispm_malloc(&y, tsize(vt) * nvec)) ||
      (st = ispm_malloc(&x2, tsize(vt) * ncols)) || (st = ispm_malloc(&y2, tsize(vt) * nvec)))
    return st;
  /* the reference's formats (S = 32..512) plus small fixed heights run in packed mode */
  for (unsigned S = 1; S <= 512; S <<= 1)
    for (int V = 0; V <= (S >= 32 ? 1 : 0); V++)
      for (int D = 0; D <= 1; D++)
        for (unsigned W = 1; W <= (D ? 1u : 2u); W <<= 1) {
          conv::layout L;
          if (make_layout(A, et, S, W, V, D, L)) continue;
          std::vector<unsigned char> el(L.n_elms * tsize(et) + 1);
          fill_values(L, A, et, el.data());
          ispm_slell_desc d;
          if ((st = upload_layout(L, el.data(), et, &d))) return st;
          diag_pattern pat;
          unsigned *uni = nullptr;
          bool has_pat = false;
          if ((st = upload_pattern(L, pat, uni, has_pat))) { ispm_slell_device_free(&d); return st; }
          const diag_pattern *pp = has_pat ? &pat : nullptr;
          layout_hints hints;
          compute_hints(L, hints);
          /* packed (S < 64) and streamed (plain ELL planes) mappings */
          for (int mode = 1; mode <= 2; mode++) {
            if (V || (mode == 1 && S >= 64) || (mode == 2 && (W != 1 || D || S > 256))) continue;
            ispm_launch lp{256, S, 1, 0, 0, mode};
            double sec;
            if (!time_candidate(vt, d, lp, x, y, x2, y2, sec, nullptr, nullptr, &hints) && sec < best) {
              best = sec;
              bS = S; bW = W; bV = V; bD = D;
              bl = lp;
            }
          }
          for (unsigned N = 64; N <= 1024; N <<= 1)
            for (unsigned H = 1; H <= 2; H <<= 1) {
              ispm_launch lp{N, V ? N : S, H, 0, 0, 0};
              if (!(lp.N >= lp.S / lp.H && lp.S >= lp.H)) continue;
              double sec;
              if (time_candidate(vt, d, lp, x, y, x2, y2, sec, pp, uni, &hints)) continue;
              if (sec < best) {
                best = sec;
                bS = S; bW = W; bV = V; bD = D;
                bl = lp;
              }
            }
          ispm_slell_device_free(&d);
          if (uni) ispm_free(uni);
        }
  ispm_free(x); ispm_free(y); ispm_free(x2); ispm_free(y2);
  if (!std::isfinite(best)) return fail_plain(ISPM_ERR_UNSUPPORTED, "exhaustive plan: no candidate ran");
  return ISPM_OK;
}

}  // namespace ispm

using namespace ispm;

namespace ispm {
extern "C" const char *ispm_last_error(void);
/* The runtime unit keeps the thread-local message; reuse its formatter through a C shim. */
int set_error_message(int code, const char *msg);
static int fail_plain(int code, const char *msg) { return set_error_message(code, msg); }
}

extern "C" {

int ispm_csr_to_slell(const ispm_csr *A, int S, int hblock, int var_height, int diags,
                      int min_height, int out_elm_type, ispm_slell_host *out)
{
  int st = check_csr(A);
  if (st) return st;
  if (!out) return fail_plain(ISPM_ERR_INVALID, "csr_to_slell: null output");
  if (S <= 0 || hblock <= 0) return fail_plain(ISPM_ERR_INVALID, "csr_to_slell: bad slice height / hblock");
  /* conversion.h:383-384 raises a fixed height below min_height to min_height rows per slice
   * while recording the requested height; such a matrix would not describe its own slices */
  if (!var_height && S < min_height)
    return fail_plain(ISPM_ERR_INVALID, "csr_to_slell: fixed slice height below min_height");
  int et = out_elm_type < 0 ? A->elm_type : out_elm_type;
  if (!tsize(et)) return fail_plain(ISPM_ERR_INVALID, "csr_to_slell: bad output type");
  conv::layout L;
  st = conv_status(conv::convert(A->row_ptr, A->n_rows, A->cols, A->n_cols, A->n_nz, (int)tsize(et),
                                 S, hblock, var_height != 0, diags != 0, min_height, L));
  if (st) return st;
  return to_host(L, A, et, out);
}

int ispm_csr_to_slell_guess(const ispm_csr *A, int examine_ratio, int min_height,
                            int out_elm_type, ispm_slell_host *out)
{
  int st = check_csr(A);
  if (st) return st;
  if (!out) return fail_plain(ISPM_ERR_INVALID, "csr_to_slell_guess: null output");
  int et = out_elm_type < 0 ? A->elm_type : out_elm_type;
  if (!tsize(et)) return fail_plain(ISPM_ERR_INVALID, "csr_to_slell_guess: bad output type");
  conv::layout L;
  st = conv_status(conv::guess(A->row_ptr, A->n_rows, A->cols, A->n_cols, A->n_nz, (int)tsize(et),
                               examine_ratio, min_height, L, nullptr));
  if (st) return st;
  return to_host(L, A, et, out);
}

void ispm_slell_host_free(ispm_slell_host *m)
{
  if (!m) return;
  std::free(m->slice_ptr);
  std::free(m->cols);
  std::free(m->elms);
  m->slice_ptr = nullptr;
  m->cols = nullptr;
  m->elms = nullptr;
}

int ispm_slell_upload(const ispm_slell_host *h, ispm_slell_desc *d)
{
  if (!h || !d) return fail_plain(ISPM_ERR_INVALID, "slell_upload: null argument");
  std::memset(d, 0, sizeof *d);
  void *sp = nullptr, *cl = nullptr, *el = nullptr;
  const size_t es = tsize(h->elm_type);
  int st;
  if ((st = ispm_malloc(&sp, 4 * (h->slice_ptr_len ? h->slice_ptr_len : 1)))) return st;
  if ((st = ispm_malloc(&cl, 4 * (h->cols_len ? h->cols_len : 1)))) return st;
  if ((st = ispm_malloc(&el, es * (h->elms_len ? h->elms_len : 1)))) return st;
  if (h->slice_ptr_len && (st = ispm_memcpy(sp, h->slice_ptr, 4 * h->slice_ptr_len, ISPM_H2D))) return st;
  if (h->cols_len && (st = ispm_memcpy(cl, h->cols, 4 * h->cols_len, ISPM_H2D))) return st;
  if (h->elms_len && (st = ispm_memcpy(el, h->elms, es * h->elms_len, ISPM_H2D))) return st;
  d->n_rows = h->n_rows;
  d->n_cols = h->n_cols;
  d->n_nz = h->n_nz;
  d->n_slices = h->n_slices;
  d->slice_height = h->slice_height;
  d->hblock = h->hblock;
  d->var_height = h->var_height;
  d->diags = h->diags;
  d->slice_ptr = (const unsigned *)sp;
  d->slice_ptr_len = h->slice_ptr_len;
  d->cols = (const unsigned *)cl;
  d->cols_len = h->cols_len;
  d->elms = el;
  d->elms_len = h->elms_len;
  d->elm_type = h->elm_type;
  return ISPM_OK;
}

void ispm_slell_device_free(ispm_slell_desc *d)
{
  if (!d) return;
  ispm_free((void *)d->slice_ptr);
  ispm_free((void *)d->cols);
  ispm_free((void *)d->elms);
  d->slice_ptr = nullptr;
  d->cols = nullptr;
  d->elms = nullptr;
}

namespace ispm {
/* Default and estimated plans of matrices whose diagonal slices share one pattern (stencils):
 * move to 256-row slices, where the pattern kernel stages each x window once for 256 rows
 * (measured on a 27-point 192^3 stencil: FP64 340 -> 332 us).  Results do not depend on
 * the format. */
static void prefer_pattern_layout(const ispm_csr *A, int vt, int et, ispm_plan *p)
{
  if (!p->L.D || p->L.V || p->L.W != 1) return;
  diag_pattern pat;
  std::vector<unsigned> mask;
  if (p->L.S != 256) {
    if (p->L.S < 64 || p->L.S > 256 || !build_pattern(p->L, pat, mask)) return;
    size_t cover = 0;
    for (unsigned w : mask) cover += (unsigned)__builtin_popcount(w);
    if (10 * cover < 9 * (size_t)p->L.n_slices) return;
    conv::layout L2;
    if (make_layout(A, et, 256, 1, 0, 1, L2) || !build_pattern(L2, pat, mask)) return;
    p->L = std::move(L2);
  } else if (!build_pattern(p->L, pat, mask)) {
    return;
  }
  p->launch = default_launch(p->L);
  (void)vt;
}
}  // namespace ispm

namespace ispm {
/* Split-diagonal plan: the main diagonal as a dense vector and the rows with off-diagonal
 * entries as a compressed CSR converted to the plain fixed-height layout (S <= 32, streamed
 * mapping) with the fewest stored bytes. */
static int create_split(const ispm_csr *A, int et, unsigned flags, ispm_plan *p)
{
  if (A->n_rows != A->n_cols) return fail_plain(ISPM_ERR_INVALID, "split-diagonal plan: matrix must be square");
  if (flags & ISPM_SPMV_PLAN_NO_DEVICE) return fail_plain(ISPM_ERR_INVALID, "split-diagonal plan needs the device");
  const int n = A->n_rows;
  const size_t n_pad = ((size_t)n + 511) / 512 * 512;
  std::vector<double> dg(n_pad, 0.0);
  std::vector<int> rp{0}, rc;
  std::vector<double> rv;
  std::vector<unsigned> rmap;
  for (int i = 0; i < n; i++) {
    const int k0 = A->row_ptr[i], k1 = A->row_ptr[i + 1];
    for (int k = k0; k < k1; k++) {
      const double v = csr_value(A, k);
      if (A->cols[k] == i) dg[i] += v;
      else { rc.push_back(A->cols[k]); rv.push_back(v); }
    }
    if ((int)rc.size() != rp.back()) { rp.push_back((int)rc.size()); rmap.push_back((unsigned)i); }
  }
  const unsigned m = (unsigned)rmap.size();
  p->split = true;
  p->split_n = (unsigned)n_pad;
  p->split_m = m;
  int st;
  if (m) {
    const ispm_csr R{(int)m, n, (int)rc.size(), rp.data(), rc.data(), rv.data(), ISPM_F64};
    /* plain layout with the fewest stored bytes: 4 per slice record + (4 + E) per slot */
    unsigned bS = 1;
    size_t best = SIZE_MAX;
    for (unsigned S = 1; S <= 32; S <<= 1) {
      size_t slots = 0, slices = 0;
      for (unsigned r0 = 0; r0 < m; r0 += S, slices++) {
        int w = 0;
        for (unsigned r = r0; r < std::min(m, r0 + S); r++) w = std::max(w, rp[r + 1] - rp[r]);
        slots += (size_t)w * S;
      }
      const size_t b = 4 * slices + (4 + tsize(et)) * slots;
      if (b < best) { best = b; bS = S; }
    }
    if ((st = make_layout(&R, et, bS, 1, 0, 0, p->L))) return st;
    p->launch = default_launch(p->L);
    if ((st = finish_plan(p, &R, flags))) return st;
  } else {
    std::memset(&p->dev, 0, sizeof p->dev);
    p->dev.n_cols = n;
    p->dev.elm_type = et;
    p->L.n_cols = n;
    p->L.S = 1;
    p->launch = default_launch(p->L);
    p->has_dev = true;
    p->stored = 0;
  }
  /* the diagonal in E and the row map (padding rows of R map nowhere) */
  std::vector<unsigned char> dh(n_pad * tsize(et));
  for (size_t i = 0; i < n_pad; i++) store_value(dh.data(), et, i, dg[i]);
  if ((st = ispm_malloc(&p->dev_diag, dh.size()))) return st;
  if ((st = ispm_memcpy(p->dev_diag, dh.data(), dh.size(), ISPM_H2D))) return st;
  const size_t m_pad = m ? (size_t)p->L.n_rows : 0;
  if (m_pad) {
    std::vector<unsigned> mh(m_pad, ~0u), bits(n_pad / 32, 0u);
    std::copy(rmap.begin(), rmap.end(), mh.begin());
    for (unsigned r : rmap) bits[r >> 5] |= 1u << (r & 31);
    if ((st = ispm_malloc((void **)&p->dev_rmap, 4 * m_pad))) return st;
    if ((st = ispm_memcpy(p->dev_rmap, mh.data(), 4 * m_pad, ISPM_H2D))) return st;
    if ((st = ispm_malloc((void **)&p->dev_rmask, 4 * bits.size()))) return st;
    if ((st = ispm_memcpy(p->dev_rmask, bits.data(), 4 * bits.size(), ISPM_H2D))) return st;
  }
  /* bytes one product moves: diagonal, remainder layout, row map and mask, x and y once */
  const size_t rem = m ? 4 * p->dev.slice_ptr_len + 4 * p->dev.cols_len + tsize(et) * p->dev.elms_len : 0;
  p->stored = n_pad * tsize(et) + rem + 4 * m_pad + (m_pad ? n_pad / 8 : 0) + tsize(p->T) * ((size_t)n + n_pad);
  p->csr = csr_bytes(A, p->E, p->T);
  p->dev.n_rows = (int)n_pad;   /* reported as the matrix's padded row count */
  return ISPM_OK;
}
}  // namespace ispm

namespace ispm {
/* Symmetric diagonal plan: the upper diagonals of an exactly symmetric matrix as dense planes
 * of 256-row slices (see sym_desc); the x-window pattern lists all 2K-1 column offsets. */
static int create_sym(const ispm_csr *A, int et, unsigned flags, ispm_plan *p)
{
  if (A->n_rows != A->n_cols) return fail_plain(ISPM_ERR_INVALID, "symmetric plan: matrix must be square");
  if (flags & ISPM_SPMV_PLAN_NO_DEVICE) return fail_plain(ISPM_ERR_INVALID, "symmetric plan needs the device");
  const int n = A->n_rows;
  const size_t n_pad = ((size_t)n + 511) / 512 * 512, ns = n_pad / 256;
  /* the upper diagonal offsets (0 always present) */
  std::vector<long long> d{0};
  for (int i = 0; i < n; i++)
    for (int k = A->row_ptr[i]; k < A->row_ptr[i + 1]; k++) {
      const long long off = std::llabs((long long)A->cols[k] - i);
      if (std::find(d.begin(), d.end(), off) == d.end()) {
        if (d.size() == 16) return fail_plain(ISPM_ERR_INVALID, "symmetric plan: more than 16 upper diagonals");
        d.push_back(off);
      }
    }
  std::sort(d.begin(), d.end());
  const unsigned K = (unsigned)d.size();
  auto idx = [&](long long off) { return (unsigned)(std::lower_bound(d.begin(), d.end(), off) - d.begin()); };
  /* upper entries into the planes, then every lower entry must equal its mirror */
  std::vector<double> v((size_t)K * n_pad, 0.0);
  std::vector<unsigned char> has((size_t)K * n_pad, 0);
  long long nup = 0, nlo = 0;
  for (int i = 0; i < n; i++)
    for (int k = A->row_ptr[i]; k < A->row_ptr[i + 1]; k++) {
      const int c = A->cols[k];
      if (c < i) continue;
      const size_t at = (size_t)idx(c - i) * n_pad + i;
      if (has[at]) {
        char m[96];
        std::snprintf(m, sizeof m, "symmetric plan: duplicate entry (%d, %d)", i, c);
        return fail_plain(ISPM_ERR_INVALID, m);
      }
      has[at] = 1;
      v[at] = csr_value(A, k);
      nup += c > i;
    }
  for (int i = 0; i < n; i++)
    for (int k = A->row_ptr[i]; k < A->row_ptr[i + 1]; k++) {
      const int c = A->cols[k];
      if (c >= i) continue;
      const size_t at = (size_t)idx(i - c) * n_pad + c;
      if (!has[at] || has[at] == 2 || v[at] != csr_value(A, k)) {
        char m[96];
        std::snprintf(m, sizeof m, "symmetric plan: entry (%d, %d) differs from (%d, %d)", i, c, c, i);
        return fail_plain(ISPM_ERR_INVALID, m);
      }
      has[at] = 2;
      nlo++;
    }
  if (nup != nlo) return fail_plain(ISPM_ERR_INVALID, "symmetric plan: the matrix is not symmetric");
  /* the CSR entry -> slot map of the refreshes: an upper entry (i, c >= i) -> its slot
   * (s K + k) kSymPlane + r of row i = 256 s + r; a lower entry -> -(slot of its mirror) - 2 */
  {
    const long long nnz = A->row_ptr[n];
    p->L.map.assign((size_t)nnz, -1);
    for (int i = 0; i < n; i++)
      for (int k = A->row_ptr[i]; k < A->row_ptr[i + 1]; k++) {
        const int c = A->cols[k];
        const int lo = std::min(i, c), hi = std::max(i, c);
        const long long slot = ((long long)(lo / 256) * K + idx(hi - lo)) * kSymPlane + lo % 256;
        p->L.map[k] = c >= i ? slot : -slot - 2;
      }
  }
  /* x-window pattern over the offsets -d[K-1] .. d[K-1] (positions relative to the slice's
   * first row + 256, runs of offsets less than 256 apart, as build_pattern makes them) */
  std::vector<long long> R;
  for (unsigned k = K; k-- > 1;) R.push_back(-d[k] + 256);
  for (unsigned k = 0; k < K; k++) R.push_back(d[k] + 256);
  sym_desc &sd = p->symd;
  sd = sym_desc{};
  diag_pattern &pat = sd.pat;
  pat.n = (unsigned)R.size();
  unsigned used = 0;
  for (unsigned f = 0; f < pat.n;) {
    unsigned l = f;
    while (l + 1 < pat.n && R[l + 1] - R[l] < 256) l++;
    const unsigned span = (unsigned)(R[l] - R[f]);
    pat.rrel[pat.nrun] = (int)(R[f] - 256);
    pat.rlen[pat.nrun] = 256 + span;
    pat.rbase[pat.nrun] = used;
    for (unsigned q = f; q <= l; q++) pat.poff[q] = used + (unsigned)(R[q] - R[f]);
    used += 256 + span;
    pat.nrun++;
    f = l + 1;
  }
  pat.lds_len = used;
  if ((size_t)used * 8 > 65536) return fail_plain(ISPM_ERR_UNSUPPORTED, "symmetric plan: x windows exceed 64 KiB of LDS");
  for (unsigned k = 0; k < K; k++) sd.d[k] = (int)d[k];
  if (ns * K * kSymPlane >= (1ull << 32) || n_pad >= (1ull << 31))
    return fail_plain(ISPM_ERR_UNSUPPORTED, "symmetric plan: more than 2^32 element slots");
  sd.K = K;
  sd.n_slices = (unsigned)ns;
  sd.n_cols = (unsigned)n;
  sd.elm_type = et;
  /* planes slice by slice: element (s K + k) kSymPlane + r; the plane's rows 256 .. 259 repeat
   * the next slice's first rows of the same diagonal (0 after the last slice), so that a
   * mirror read of up to 4 consecutive rows never leaves its plane */
  std::vector<unsigned char> eh(ns * K * kSymPlane * tsize(et));
  for (size_t s0 = 0; s0 < ns; s0++)
    for (unsigned k = 0; k < K; k++)
      for (unsigned r = 0; r < kSymPlane; r++) {
        const size_t row = s0 * 256 + r;
        store_value(eh.data(), et, (s0 * K + k) * kSymPlane + r, row < n_pad ? v[(size_t)k * n_pad + row] : 0.0);
      }
  void *de = nullptr;
  int st = ispm_malloc(&de, eh.size());
  if (st) return st;
  if ((st = ispm_memcpy(de, eh.data(), eh.size(), ISPM_H2D))) { ispm_free(de); return st; }
  sd.elms = de;
  p->sym = true;
  std::memset(&p->dev, 0, sizeof p->dev);
  p->dev.n_rows = (int)n_pad;
  p->dev.n_cols = n;
  p->dev.n_nz = A->n_nz;
  p->dev.n_slices = (unsigned)ns;
  p->dev.slice_height = 256;
  p->dev.diags = 1;
  p->dev.elms = de;
  p->dev.elms_len = ns * K * kSymPlane;
  p->dev.elm_type = et;
  p->has_dev = true;
  p->L.n_rows = (int)n_pad;
  p->L.n_cols = n;
  p->L.S = 256;
  p->L.D = true;
  p->L.n_slices = (unsigned)ns;
  p->L.n_elms = ns * K * kSymPlane;
  p->launch = default_launch(p->L);
  p->stored = tsize(et) * ns * K * 256 + tsize(p->T) * ((size_t)n + n_pad);
  p->csr = csr_bytes(A, p->E, p->T);
  return ISPM_OK;
}
}  // namespace ispm

namespace ispm {
/* New values for a symmetric plan's pattern: upper entries into their slots (and into the
 * repeated rows 256 .. 259 of the previous slice's plane), then every lower entry must equal
 * its mirror, as at creation. */
static int sym_refresh_host(ispm_plan *p, const ispm_csr *A)
{
  const unsigned K = p->symd.K, et = (unsigned)p->E;
  const size_t ks = (size_t)K * kSymPlane;
  std::vector<unsigned char> eh(p->L.n_elms * tsize(et), 0);
  const long long nnz = A->row_ptr[A->n_rows];
  for (long long k = 0; k < nnz; k++) {
    const long long s = p->L.map[k];
    if (s < 0) continue;
    const double v = csr_value(A, k);
    store_value(eh.data(), et, (size_t)s, v);
    if ((size_t)s % kSymPlane < 4 && (size_t)s >= ks) store_value(eh.data(), et, (size_t)s - ks + 256, v);
  }
  for (int i = 0; i < A->n_rows; i++)
    for (int k = A->row_ptr[i]; k < A->row_ptr[i + 1]; k++) {
      const long long s = p->L.map[k];
      if (s >= 0) continue;
      /* the mirror's stored value, compared in the plan's element type */
      std::vector<unsigned char> one(tsize(et));
      store_value(one.data(), et, 0, csr_value(A, k));
      if (std::memcmp(one.data(), eh.data() + (size_t)(-s - 2) * tsize(et), tsize(et))) {
        char m[96];
        std::snprintf(m, sizeof m, "plan_update_coefs: entry (%d, %d) differs from its mirror", i, A->cols[k]);
        return fail_plain(ISPM_ERR_INVALID, m);
      }
    }
  return ispm_memcpy(const_cast<void *>(p->symd.elms), eh.data(), eh.size(), ISPM_H2D);
}
}  // namespace ispm

int ispm_plan_create(ispm_plan **out, int vt, int et, const ispm_csr *A, int kind, unsigned flags)
{
  if (!out) return fail_plain(ISPM_ERR_INVALID, "plan_create: null output");
  *out = nullptr;
  int st = check_csr(A);
  if (st) return st;
  if (et < 0) et = A->elm_type;
  if (!((vt == ISPM_F64 && (et == ISPM_F64 || et == ISPM_F32)) || (vt == ISPM_F32 && (et == ISPM_F32 || et == ISPM_BF16))))
    return fail_plain(ISPM_ERR_UNSUPPORTED, "plan_create: unsupported (T,E) pair");
  if (kind < ISPM_PLAN_DEFAULT || kind > ISPM_PLAN_MIN_BYTES)
    return fail_plain(ISPM_ERR_INVALID, "plan_create: unknown plan kind");
  ispm_plan *p = new ispm_plan;
  p->T = vt;
  p->E = et;
  if (flags & ISPM_SPMV_PLAN_SYMMETRIC) {
    st = create_sym(A, et, flags, p);
    if (st) { ispm_plan_destroy(p); return st; }
    *out = p;
    return ISPM_OK;
  }
  if (flags & ISPM_SPMV_PLAN_SPLIT_DIAG) {
    st = create_split(A, et, flags, p);
    if (st) { ispm_plan_destroy(p); return st; }
    *out = p;
    return ISPM_OK;
  }
  if (kind == ISPM_PLAN_EXHAUSTIVE) {
    if (flags & ISPM_SPMV_PLAN_NO_DEVICE) { delete p; return fail_plain(ISPM_ERR_INVALID, "exhaustive plan needs the device"); }
    unsigned S = 0, W = 1;
    int V = 0, D = 0;
    ispm_launch lp;
    const char *cache = std::getenv("ISPM_PLAN_CACHE");
    unsigned long long fp = 0;
    bool hit = false;
    if (cache && cache[0] && !ispm_csr_fingerprint(A, &fp))
      hit = ispm_plan_cache_lookup(cache, fp, vt, et, &S, &W, &V, &D, &lp, &p->tuned) == ISPM_OK;
    if (!hit) {
      st = tune_exhaustive(A, vt, et, S, W, V, D, lp, p->tuned);
      if (st) { delete p; return st; }
      if (cache && cache[0]) ispm_plan_cache_store(cache, fp, vt, et, S, W, V, D, &lp, p->tuned);
    }
    st = make_layout(A, et, S, W, V, D, p->L);
    if (st) { delete p; return st; }
    p->launch = lp;
  } else if (kind == ISPM_PLAN_MIN_BYTES) {
    unsigned S = 64, W = 1;
    int V = 0, D = 0;
    st = estimate_format(A, vt, et, S, W, V, D);
    if (!st) st = make_layout(A, et, S, W, V, D, p->L);
    if (st) { delete p; return st; }
    p->launch = default_launch(p->L);
    if (!(flags & ISPM_SPMV_PLAN_NO_DEVICE)) prefer_pattern_layout(A, vt, et, p);
  } else {
    conv::guess_info gi;
    st = conv_status(conv::guess(A->row_ptr, A->n_rows, A->cols, A->n_cols, A->n_nz, (int)tsize(et),
                                 1, 4, p->L, &gi));
    if (st) { delete p; return st; }
    p->launch = default_launch(p->L);
    if (!(flags & ISPM_SPMV_PLAN_NO_DEVICE)) prefer_pattern_layout(A, vt, et, p);
  }
  st = finish_plan(p, A, flags);
  if (st) { ispm_plan_destroy(p); return st; }
  if (kind == ISPM_PLAN_EXHAUSTIVE && std::getenv("SPMV_TUNE_RESULT")) {
    /* plan.h:137-142 */
    size_t eff = 4 * p->L.sp.size() + (4 + tsize(et)) * (size_t)A->n_nz + tsize(vt) * 2 * (size_t)p->L.n_rows;
    std::fprintf(stderr, "[spmv tuning] %g ms:  %g (%g eff, %g csr-equiv) GB/s\n", 1e3 * p->tuned,
                 1e-9 * p->stored / p->tuned, 1e-9 * eff / p->tuned, 1e-9 * p->csr / p->tuned);
  }
  *out = p;
  return ISPM_OK;
}

int ispm_plan_create_explicit(ispm_plan **out, int vt, int et, const ispm_csr *A, unsigned S,
                              unsigned hblock, int V, int D, const ispm_launch *launch)
{
  if (!out) return fail_plain(ISPM_ERR_INVALID, "plan_create_explicit: null output");
  *out = nullptr;
  int st = check_csr(A);
  if (st) return st;
  if (et < 0) et = A->elm_type;
  if (!((vt == ISPM_F64 && (et == ISPM_F64 || et == ISPM_F32)) || (vt == ISPM_F32 && (et == ISPM_F32 || et == ISPM_BF16))))
    return fail_plain(ISPM_ERR_UNSUPPORTED, "plan_create_explicit: unsupported (T,E) pair");
  ispm_plan *p = new ispm_plan;
  p->T = vt;
  p->E = et;
  st = make_layout(A, et, S, hblock, V, D, p->L);
  if (st) { delete p; return st; }
  p->launch = launch ? *launch : default_launch(p->L);
  st = finish_plan(p, A, 0);
  if (st) { ispm_plan_destroy(p); return st; }
  *out = p;
  return ISPM_OK;
}

int ispm_plan_default_launch(unsigned slice_height, int var_height, ispm_launch *out)
{
  if (!out) return fail_plain(ISPM_ERR_INVALID, "plan_default_launch: null output");
  conv::layout L;
  L.S = slice_height;
  L.V = var_height != 0;
  *out = default_launch(L);
  return ISPM_OK;
}

int ispm_plan_estimate(int vt, int et, const ispm_csr *A, unsigned *S, unsigned *hblock, int *V,
                       int *D, ispm_launch *launch)
{
  int st = check_csr(A);
  if (st) return st;
  if (!S || !hblock || !V || !D || !launch) return fail_plain(ISPM_ERR_INVALID, "plan_estimate: null output");
  unsigned s = 64, w = 1;
  int v = 0, d = 0;
  st = estimate_format(A, vt, et, s, w, v, d);
  if (st) return st;
  *S = s; *hblock = w; *V = v; *D = d;
  conv::layout L;
  L.S = s;
  L.V = v != 0;
  *launch = default_launch(L);
  return ISPM_OK;
}

/* Exclusive use of the device for timing, shared between processes by an advisory file
 * lock (flock on lock_fd); lock_fd < 0 takes nothing. */
namespace {
struct device_lock {
  int fd;
  explicit device_lock(int f) : fd(f) { if (fd >= 0) while (flock(fd, LOCK_EX) != 0 && errno == EINTR) {} }
  ~device_lock() { if (fd >= 0) flock(fd, LOCK_UN); }
};
}

int ispm_plan_time_formats(int vt, int et, const ispm_csr *A, int count, const unsigned *S,
                           const unsigned *hblock, const int *V, const int *D, double *seconds)
{
  return ispm_plan_time_formats_shared(vt, et, A, count, S, hblock, V, D, seconds, 0, -1);
}

int ispm_plan_time_formats_shared(int vt, int et, const ispm_csr *A, int count, const unsigned *S,
                                  const unsigned *hblock, const int *V, const int *D, double *seconds,
                                  int threads, int lock_fd)
{
  int st = check_csr(A);
  if (st) return st;
  if (count < 0 || (count && (!S || !hblock || !V || !D || !seconds)))
    return fail_plain(ISPM_ERR_INVALID, "plan_time_formats: bad argument");
  if ((st = ensure_init())) return st;
  if (et < 0) et = A->elm_type;
  const size_t nvec = std::max<size_t>((size_t)A->n_rows + 512, (size_t)A->n_cols + 512);
  void *x = nullptr, *y = nullptr, *x2 = nullptr, *y2 = nullptr;
  if ((st = ispm_malloc(&x, tsize(vt) * nvec)) || (st = ispm_malloc(&y, tsize(vt) * nvec)) ||
      (st = ispm_malloc(&x2, tsize(vt) * nvec)) || (st = ispm_malloc(&y2, tsize(vt) * nvec))) {
    ispm_free(x); ispm_free(y); ispm_free(x2); ispm_free(y2);
    return st;
  }
  /* layouts are converted and their values placed on host threads, a batch at a time, then
   * uploaded and timed (only that part holds the device lock) */
  const int nt = threads > 0 ? std::min(threads, 64)
                             : (int)std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
  for (int c0 = 0; c0 < count && !st; c0 += nt) {
    const int nb = std::min(nt, count - c0);
    std::vector<conv::layout> Ls(nb);
    std::vector<std::vector<unsigned char>> els(nb);
    std::vector<int> ok(nb, 0);
    std::vector<std::thread> th;
    for (int k = 0; k < nb; k++)
      th.emplace_back([&, k] {
        ok[k] = !make_layout(A, et, S[c0 + k], hblock[c0 + k], V[c0 + k], D[c0 + k], Ls[k]);
        if (!ok[k]) return;
        els[k].resize(Ls[k].n_elms * tsize(et) + 1);
        fill_values(Ls[k], A, et, els[k].data());
        std::vector<long long>().swap(Ls[k].map);   /* not needed for timing */
      });
    for (auto &t : th) t.join();
    device_lock held(lock_fd);   /* uploads and timings of this batch */
    for (int k = 0; k < nb && !st; k++) {
      const int c = c0 + k;
      seconds[c] = INFINITY;
      if (!ok[k]) continue;   /* format not buildable */
      const conv::layout &L = Ls[k];
      ispm_slell_desc d;
      if ((st = upload_layout(L, els[k].data(), et, &d))) break;
      diag_pattern pat;
      unsigned *uni = nullptr;
      bool has_pat = false;
      if (!(st = upload_pattern(L, pat, uni, has_pat))) {
        const ispm_launch lp = default_launch(L);
        layout_hints hints;
        compute_hints(L, hints);
        double sec;
        if (!time_candidate(vt, d, lp, x, y, x2, y2, sec, has_pat ? &pat : nullptr, uni, &hints)) seconds[c] = sec;
      }
      ispm_slell_device_free(&d);
      if (uni) ispm_free(uni);
    }
  }
  ispm_free(x); ispm_free(y); ispm_free(x2); ispm_free(y2);
  return st;
}

int ispm_plan_tune(int vt, int et, const ispm_csr *A, unsigned *S, unsigned *hblock, int *V,
                   int *D, ispm_launch *launch, double *seconds)
{
  int st = check_csr(A);
  if (st) return st;
  if (!S || !hblock || !V || !D || !launch) return fail_plain(ISPM_ERR_INVALID, "plan_tune: null output");
  double best = 0;
  ispm_launch lp{};
  unsigned s = 0, w = 1;
  int v = 0, d = 0;
  st = tune_exhaustive(A, vt, et, s, w, v, d, lp, best);
  if (st) return st;
  *S = s; *hblock = w; *V = v; *D = d; *launch = lp;
  if (seconds) *seconds = best;
  return ISPM_OK;
}

int ispm_plan_info_get(const ispm_plan *p, ispm_plan_info *info)
{
  if (!p || !info) return fail_plain(ISPM_ERR_INVALID, "plan_info: null argument");
  info->S = p->L.S;
  info->hblock = p->L.W;
  info->V = p->L.V;
  info->D = p->L.D;
  info->launch = p->launch;
  info->tuned_seconds = p->tuned;
  info->stored_bytes = p->stored;
  info->csr_bytes = p->csr;
  info->device = p->dev;
  info->pattern_planes = p->has_pat ? p->pat.n : 0;
  info->pattern_slices = p->has_pat ? p->pat_slices : 0;
  info->window_slices = p->hints.xwin_slices;
  return ISPM_OK;
}

int ispm_plan_execute(ispm_plan *p, const void *x, void *y)
{
  if (!p) return fail_plain(ISPM_ERR_INVALID, "plan_execute: null plan");
  if (!p->has_dev) return fail_plain(ISPM_ERR_INVALID, "plan_execute: plan has no device matrix");
  int st = ensure_init();
  if (st) return st;
  if (p->split) {
    const split_diag sd{p->dev_diag, p->dev_rmap, p->dev_rmask, p->split_n, p->split_m, p->E};
    return spmv_split(sd, &p->launch, &p->dev, p->T, x, y, nullptr, nullptr, nullptr, nullptr, &p->hints);
  }
  if (p->sym) return spmv_sym(p->symd, p->T, x, y, nullptr);
  return spmv_internal(&p->launch, &p->dev, p->T, x, y, nullptr, p->has_pat ? &p->pat : nullptr,
                       p->dev_uni, &p->hints);
}

int ispm_plan_execute_guarded(ispm_plan *p, const void *x, void *y, unsigned x_guard)
{
  if (!p) return fail_plain(ISPM_ERR_INVALID, "plan_execute_guarded: null plan");
  if (!p->sym) return ispm_plan_execute(p, x, y);
  if (!p->has_dev) return fail_plain(ISPM_ERR_INVALID, "plan_execute_guarded: plan has no device matrix");
  int st = ensure_init();
  if (st) return st;
  return spmv_sym(p->symd, p->T, x, y, nullptr, nullptr, nullptr, nullptr, x_guard);
}

unsigned ispm_plan_x_guard(const ispm_plan *p)
{
  return p && p->sym ? sym_x_guard(p->symd) : 0u;
}

int ispm_plan_execute_multi(ispm_plan *p, int k, const void *X, size_t ldx, void *Y, size_t ldy)
{
  if (!p) return fail_plain(ISPM_ERR_INVALID, "plan_execute_multi: null plan");
  if (!p->has_dev) return fail_plain(ISPM_ERR_INVALID, "plan_execute_multi: plan has no device matrix");
  if (p->split) return fail_plain(ISPM_ERR_UNSUPPORTED, "plan_execute_multi: split-diagonal plans multiply one vector");
  if (p->sym) {
    const int st = ensure_init();
    return st ? st : spmm_sym(p->symd, p->T, k, X, ldx, Y, ldy);
  }
  return spmm_internal(&p->launch, &p->dev, p->T, k, X, ldx, Y, ldy, &p->hints,
                       p->has_pat ? &p->pat : nullptr, p->dev_uni);
}

int ispm_plan_update_coefs(ispm_plan *p, const ispm_csr *A)
{
  if (!p) return fail_plain(ISPM_ERR_INVALID, "plan_update_coefs: null plan");
  if (p->split) return fail_plain(ISPM_ERR_UNSUPPORTED, "plan_update_coefs: split-diagonal plans are rebuilt");
  int st = check_csr(A);
  if (st) return st;
  if ((long long)A->row_ptr[A->n_rows] != (long long)p->L.map.size())
    return fail_plain(ISPM_ERR_INVALID, "plan_update_coefs: pattern size differs from the plan's");
  if (p->sym) return sym_refresh_host(p, A);
  if (fill_values(p->L, A, p->E, p->elms.data())) return fail_plain(ISPM_ERR_INVALID, "value map incomplete");
  if (p->has_dev && p->L.n_elms)
    return ispm_memcpy((void *)p->dev.elms, p->elms.data(), tsize(p->E) * p->L.n_elms, ISPM_H2D);
  return ISPM_OK;
}

void ispm_plan_destroy(ispm_plan *p)
{
  if (!p) return;
  if (p->has_dev) ispm_slell_device_free(&p->dev);
  if (p->dev_diag) ispm_free(p->dev_diag);
  if (p->dev_rmap) ispm_free(p->dev_rmap);
  if (p->dev_rmask) ispm_free(p->dev_rmask);
  if (p->dev_map) ispm_free(p->dev_map);
  if (p->dev_uni) ispm_free(p->dev_uni);
  if (p->dev_xwin) ispm_free(p->dev_xwin);
  delete p;
}

}  // extern "C"
