/* fastainv.cpp -- host AINV-type approximate inverse, M = Ainv2 * Ainv1 = W D W^T ~= A^-1.
 *
 * Same algorithm and arithmetic as the reference's fastainv_sym
 * (fastainv/fastainv.cpp:395-566), which the factors must match bit for bit:
 *   for j = 0..n-1:  u = A w_j (k-way heap merge of the rows of A selected by w_j, the unit
 *   diagonal of w_j implicit);  D_jj = u . w_j;  for every i > j with u_i != 0:
 *   w_i <- drop(w_i - (u_i / D_jj) w_j, |e| <= droptol);  D_j = 1 / D_jj (0 on breakdown).
 * The order in which equal columns leave the merge heap fixes the rounding of u, so the
 * heap keeps the reference's discipline (strict comparisons, right child on ties).  Floating
 * point follows the reference's types: u and the merge products in in_real (double), W, D
 * and the column updates in out_real; no fused multiply-add (built with -ffp-contract=off).
 * Storage capacity follows chunky_storage's accounting (fastainv.cpp:13-97) so that
 * truncation under a small max_elts is reproduced, while the columns themselves live in one
 * flat pool (a column is moved to the pool's end when it outgrows its slot).
 */
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>
#include "fastainv_impl.h"

namespace ispm {
namespace {

template<class real>
struct colpool {
  std::vector<int> row;
  std::vector<real> val;
  std::vector<long long> start;
  std::vector<int> len, cap;
  long long chunks, used = 0;          /* reference capacity accounting, in chunks of 8 */
  colpool(int n, long long nchunks) : start(n, 0), len(n, 0), cap(n, 0), chunks(nchunks)
  {
    row.reserve((size_t)n * 4);
    val.reserve((size_t)n * 4);
  }
  static long long nch(long long l) { return (l + 7) >> 3; }
  void replace(int c, const int *nr, const real *nv, int n)
  {
    const long long avail = (chunks - used + nch(len[c])) * 8;
    if ((long long)n > avail) n = (int)avail;
    if (n <= 0) return;                /* empty replacement keeps the old column */
    used += std::max(0LL, nch(n) - nch(len[c]));
    if (n > cap[c]) {
      const int nc = (int)(nch(n) * 8);
      start[c] = (long long)row.size();
      row.resize(row.size() + nc);
      val.resize(val.size() + nc);
      cap[c] = nc;
    }
    std::memcpy(&row[start[c]], nr, sizeof(int) * n);
    std::memcpy(&val[start[c]], nv, sizeof(real) * n);
    len[c] = n;
  }
};

/* Binary min-heap on column with a sentinel slot (rowmerge_heap, fastainv.cpp:99-213). */
struct kheap {
  std::vector<int> col, idx, ofs, lim;
  std::vector<double> fac;
  int sz = 0;
  void reset(int cap)
  {
    if ((int)ofs.size() < cap) {
      col.resize(cap + 1); idx.resize(cap + 1); ofs.resize(cap); lim.resize(cap); fac.resize(cap);
    }
    sz = 0;
  }
  void push(int o, int l, int c, double f)
  {
    col[sz] = c; idx[sz] = sz; ofs[sz] = o; lim[sz] = l; fac[sz] = f; sz++;
  }
  void sink(int i)
  {
    int c = col[i], id = idx[i];
    for (;;) {
      const int l = 2 * i + 1;
      if (l >= sz) break;
      const int r = l + 1;
      const int cl = col[l], cr = col[r];
      const int m = cl < cr ? cl : cr;
      if (!(m < c)) break;
      const int ch = cl < cr ? l : r;
      col[i] = m; idx[i] = idx[ch];
      i = ch;
    }
    col[i] = c; idx[i] = id;
  }
  void build()
  {
    col[sz] = INT_MAX;
    for (int i = sz / 2 - 1; i >= 0; i--) sink(i);
  }
  void pop()
  {
    sz--;
    col[0] = col[sz]; idx[0] = idx[sz];
    col[sz] = INT_MAX;
    if (sz > 0) sink(0);
  }
  void replace_top(int c) { col[0] = c; sink(0); }
};

template<class real>
bool factor(const int n, const int *rp, const int *ci, const double *a, long long max_elts,
            double droptol, void (*progress)(void *), void *pv, std::vector<int> &r1,
            std::vector<int> &c1, std::vector<real> &v1, std::vector<int> &r2,
            std::vector<int> &c2, std::vector<real> &v2)
{
  if (n <= 0) return false;
  const long long nb = ((long long)n + 7) / 8;
  max_elts = std::min(max_elts, nb * (nb + 1) / 2 * 64);
  colpool<real> W(n, max_elts / 8);
  std::vector<real> D(n);
  std::vector<int> uidx(n), mrow(n + 1);
  std::vector<double> uval(n);
  std::vector<real> mval(n + 1);
  kheap h;
  for (int j = 0; j < n; j++) {
    if (progress) progress(pv);
    const int wl = W.len[j];
    const int *wr = W.row.data() + W.start[j];
    const real *wv = W.val.data() + W.start[j];
    const int *ui;
    const double *uv;
    int un, k = 0;
    real Djj = 0;
    if (!wl) {
      ui = ci + rp[j]; uv = a + rp[j]; un = rp[j + 1] - rp[j];
    } else {
      /* u = A w_j */
      h.reset(wl + 1);
      for (int t = 0; t < wl; t++) {
        const int rr = wr[t];
        if (rp[rr] < rp[rr + 1]) h.push(rp[rr], rp[rr + 1], ci[rp[rr]], (double)wv[t]);
      }
      if (rp[j] < rp[j + 1]) h.push(rp[j], rp[j + 1], ci[rp[j]], 1.0);
      h.build();
      int cnt = 0, prev = -1;
      while (h.sz > 0) {
        const int c = h.col[0], id = h.idx[0];
        const double add = a[h.ofs[id]] * h.fac[id];
        if (c != prev) { uidx[cnt] = c; uval[cnt] = add; cnt++; prev = c; }
        else uval[cnt - 1] += add;
        if (++h.ofs[id] < h.lim[id]) h.replace_top(ci[h.ofs[id]]);
        else h.pop();
      }
      ui = uidx.data(); uv = uval.data(); un = cnt;
      /* D_jj = u . w_j */
      int t = 0;
      while (k < un && t < wl) {
        if (ui[k] < wr[t]) k++;
        else if (ui[k] > wr[t]) t++;
        else { Djj = (real)((double)Djj + uv[k] * (double)wv[t]); k++; t++; }
      }
    }
    while (k < un) {
      if (ui[k] < j) k++;
      else if (ui[k] > j) break;
      else { Djj = (real)((double)Djj + uv[k]); k++; }
    }
    if (Djj == 0) { D[j] = 0; continue; }
    Djj = (real)(1.0 / (double)Djj);
    for (; k < un; k++) {
      const int i = ui[k];
      const real f = (real)(-uv[k] * (double)Djj);
      /* merge w_i with f * (w_j + e_j), dropping |e| <= droptol */
      const int il = W.len[i];
      const int *ir = W.row.data() + W.start[i];
      const real *iv = W.val.data() + W.start[i];
      const int *jr = W.row.data() + W.start[j];
      const real *jv = W.val.data() + W.start[j];
      int p = 0, q = 0, m = 0;
      bool diag = false;
      while (p < il || q < wl || !diag) {
        const bool jside = q < wl || !diag;
        const int jrow = q < wl ? jr[q] : j;
        if (p < il && (!jside || ir[p] < jrow)) {
          const real e = iv[p];
          if (std::abs(e) > droptol) { mrow[m] = ir[p]; mval[m] = e; m++; }
          p++;
          continue;
        }
        const real jval = q < wl ? (real)(jv[q] * f) : f;
        real e;
        if (p < il && ir[p] == jrow) { e = (real)(iv[p] + jval); p++; }
        else e = jval;
        if (std::abs(e) > droptol) { mrow[m] = jrow; mval[m] = e; m++; }
        if (q < wl) q++; else diag = true;
      }
      W.replace(i, mrow.data(), mval.data(), m);
    }
    D[j] = Djj;
  }
  /* Ainv2 (fastainv.cpp:489-532) */
  r2.assign(n + 1, 0);
  for (int c = 0; c < n; c++)
    for (int t = 0; t < W.len[c]; t++) r2[W.row[W.start[c] + t] + 1]++;
  for (int i = 1; i <= n; i++) r2[i] += r2[i - 1] + 1;
  c2.assign(r2[n], 0);
  v2.assign(r2[n], 0);
  std::vector<int> cur(n, 0);
  std::vector<real> Ds(n);
  for (int i = 0; i < n; i++) {
    const int c = n - 1 - i;
    const real sc = std::sqrt(std::abs(D[c]));
    Ds[c] = std::signbit(D[c]) ? -sc : sc;
    for (int t = 0; t < W.len[c]; t++) {
      const int rr = W.row[W.start[c] + t];
      const int off = r2[rr] + cur[rr]++;
      c2[off] = i;
      v2[off] = W.val[W.start[c] + t] * sc;
    }
    const int off = r2[c] + cur[c]++;
    c2[off] = i;
    v2[off] = sc;
  }
  /* Ainv1 (fastainv.cpp:534-558) */
  r1.assign(n + 1, 0);
  c1.resize(r2[n]);
  v1.resize(r2[n]);
  long long o = 0;
  for (int i = 0; i < n; i++) {
    const int c = n - 1 - i;
    for (int t = 0; t < W.len[c]; t++) {
      c1[o] = W.row[W.start[c] + t];
      v1[o] = W.val[W.start[c] + t] * Ds[c];
      o++;
    }
    c1[o] = c;
    v1[o] = Ds[c];
    o++;
    r1[i + 1] = (int)o;
  }
  return true;
}

}  // namespace

bool fastainv_f64_f32(int n, const int *rp, const int *ci, const double *a, long long max_elts,
                      double droptol, void (*progress)(void *), void *pv, std::vector<int> &r1,
                      std::vector<int> &c1, std::vector<float> &v1, std::vector<int> &r2,
                      std::vector<int> &c2, std::vector<float> &v2)
{
  return factor<float>(n, rp, ci, a, max_elts, droptol, progress, pv, r1, c1, v1, r2, c2, v2);
}

bool fastainv_f64_f64(int n, const int *rp, const int *ci, const double *a, long long max_elts,
                      double droptol, void (*progress)(void *), void *pv, std::vector<int> &r1,
                      std::vector<int> &c1, std::vector<double> &v1, std::vector<int> &r2,
                      std::vector<int> &c2, std::vector<double> &v2)
{
  return factor<double>(n, rp, ci, a, max_elts, droptol, progress, pv, r1, c1, v1, r2, c2, v2);
}

}  // namespace ispm

/* ------------------------------------------------------------------ C-ABI ---------- */
namespace ispm { int set_error_message(int code, const char *msg); }

template<class real>
static int emit(const std::vector<int> &r, const std::vector<int> &c, const std::vector<real> &v,
                int n, int et, ispm_csr_out *o)
{
  o->n_rows = o->n_cols = n;
  o->n_nz = (int)c.size();
  o->elm_type = et;
  o->row_ptr = (int *)std::malloc(sizeof(int) * (n + 1));
  o->cols = (int *)std::malloc(sizeof(int) * (c.size() + 1));
  o->elms = std::malloc(sizeof(real) * (v.size() + 1));
  if (!o->row_ptr || !o->cols || !o->elms) return ispm::set_error_message(ISPM_ERR_NOMEM, "fastainv: out of host memory");
  std::copy(r.begin(), r.end(), o->row_ptr);
  std::copy(c.begin(), c.end(), o->cols);
  std::copy(v.begin(), v.end(), (real *)o->elms);
  return ISPM_OK;
}

extern "C" int ispm_fastainv_sym(int out_type, const ispm_csr *A, ispm_progress_fn progress,
                                 void *v, long long max_elts, double droptol,
                                 ispm_csr_out *Ainv1, ispm_csr_out *Ainv2)
{
  if (!A || !Ainv1 || !Ainv2) return ispm::set_error_message(ISPM_ERR_INVALID, "fastainv: null argument");
  std::memset(Ainv1, 0, sizeof *Ainv1);
  std::memset(Ainv2, 0, sizeof *Ainv2);
  if (A->n_rows <= 0) return ispm::set_error_message(ISPM_ERR_INVALID, "fastainv: n <= 0");
  if (A->elm_type != ISPM_F64) return ispm::set_error_message(ISPM_ERR_UNSUPPORTED, "fastainv: input must be FP64");
  const double *a = static_cast<const double *>(A->elms);
  std::vector<int> r1, c1, r2, c2;
  if (out_type == ISPM_F32) {
    std::vector<float> v1, v2;
    ispm::fastainv_f64_f32(A->n_rows, A->row_ptr, A->cols, a, max_elts, droptol, progress, v, r1, c1, v1, r2, c2, v2);
    int st = emit(r1, c1, v1, A->n_rows, ISPM_F32, Ainv1);
    if (!st) st = emit(r2, c2, v2, A->n_rows, ISPM_F32, Ainv2);
    return st;
  }
  if (out_type == ISPM_F64) {
    std::vector<double> v1, v2;
    ispm::fastainv_f64_f64(A->n_rows, A->row_ptr, A->cols, a, max_elts, droptol, progress, v, r1, c1, v1, r2, c2, v2);
    int st = emit(r1, c1, v1, A->n_rows, ISPM_F64, Ainv1);
    if (!st) st = emit(r2, c2, v2, A->n_rows, ISPM_F64, Ainv2);
    return st;
  }
  return ispm::set_error_message(ISPM_ERR_UNSUPPORTED, "fastainv: output type must be F32 or F64");
}

extern "C" void ispm_csr_out_free(ispm_csr_out *m)
{
  if (!m) return;
  std::free(m->row_ptr);
  std::free(m->cols);
  std::free(m->elms);
  m->row_ptr = nullptr;
  m->cols = nullptr;
  m->elms = nullptr;
}
