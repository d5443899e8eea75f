/* oracle.cpp -- CPU restatement of the reference's SpMV / pipelined-PCG / approximate-inverse
 * semantics.  TEST INFRASTRUCTURE ONLY: it is the checker that tests/, the smoke check and
 * the benchmark's CPU-baseline leg compare the HIP library against.  Nothing in the product
 * library links, loads or calls this file.
 *
 * Every function names the reference lines it restates (paths relative to the reference
 * tree).  How each function is pinned (golden vectors, closed forms, invariants, brute
 * force) is listed in the header of oracle.py.
 *
 * Build: oracle/Makefile -> oracle/liboracle.so (g++ -O2 -fopenmp -ffp-contract=off).
 * -ffp-contract=off keeps a*b+c as two roundings, as the reference's host build
 * (-O2, no -march, so no FMA) computes it.
 */
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <omp.h>

extern "C" {

/* ---------------------------------------------------------------- threads --------- */
int orc_set_threads(int n)
{
  if (n > 0) omp_set_num_threads(n);
  int t = 1;
#pragma omp parallel
  {
#pragma omp single
    t = omp_get_num_threads();
  }
  return t;
}

/* ------------------------------------------------------------- CSR SpMV ------------- */
/* y = A x in FP64, entries of a row summed in ascending storage order.  The textbook
 * definition the SlELL kernel (cuda/spmv/kernel.h:70-134) must reproduce per row. */
void orc_csr_spmv(int n_rows, const int *rp, const int *ci, const double *a,
                  const double *x, double *y)
{
#pragma omp parallel for schedule(static)
  for (int r = 0; r < n_rows; r++) {
    double s = 0;
    for (int k = rp[r]; k < rp[r + 1]; k++) s += a[k] * x[ci[k]];
    y[r] = s;
  }
}

/* sum_k |a_rk| |x_k| : the per-row scale of the 1e-13 relative SpMV parity bound. */
void orc_csr_spmv_abs(int n_rows, const int *rp, const int *ci, const double *a,
                      const double *x, double *y)
{
#pragma omp parallel for schedule(static)
  for (int r = 0; r < n_rows; r++) {
    double s = 0;
    for (int k = rp[r]; k < rp[r + 1]; k++) s += std::fabs(a[k]) * std::fabs(x[ci[k]]);
    y[r] = s;
  }
}

/* ----------------------------------------------------------- SlELL decode ----------- */
/* Walks a SlELL layout the way the SpMV kernel addresses it (cuda/spmv/kernel.h:13-119,
 * formats/formats.h:39-102, conversion.h:338-414) and lists every stored slot as
 * (row, col, value).  Slot columns of diagonal planes may fall outside [0, n_cols)
 * (texture reads there return 0, texture-wrap.h); such slots are reported with their raw
 * column.  hblock slots report col = W*pair + member.  Returns the slot count, or -1 if
 * cap is too small (then nothing beyond cap is written). */
long long orc_slell_decode(int n_cols, unsigned n_slices, unsigned S, unsigned W, int V,
                           int D, const unsigned *sp, const unsigned *cols,
                           const double *elms, long long cap, int *o_row, long long *o_col,
                           double *o_val)
{
  (void)n_cols;
  const unsigned P = 1 + (W > 1) + (V ? 1 : 0) + (D ? 2 : 0);
  long long n = 0;
  auto emit = [&](long long r, long long c, double v) {
    if (n < cap) { o_row[n] = (int)r; o_col[n] = c; o_val[n] = v; }
    n++;
  };
  for (unsigned s = 0; s < n_slices; s++) {
    const unsigned p0 = P * s;
    unsigned e0 = sp[p0], e1 = sp[p0 + P];
    unsigned y0, sz;
    if (V) { y0 = sp[p0 + 1]; sz = sp[p0 + 1 + P] - y0; }
    else   { y0 = S * s; sz = S; }
    if (D) {
      unsigned c0 = sp[p0 + V + 1], d0 = sp[p0 + V + 2], c1 = sp[p0 + V + 1 + P];
      unsigned d1 = sp[p0 + V + 2 + P];
      unsigned nce = c1 - c0;                     /* ELL part, with column indices */
      for (unsigned k = 0; k < nce; k++)
        emit(y0 + k % sz, cols[c0 + k], elms[e0 + k]);
      unsigned nde = e1 - e0 - nce;               /* diagonal planes */
      if (nde != (d1 - d0) * sz) return -2;       /* inconsistent layout */
      for (unsigned k = 0; k < nde; k++) {
        unsigned plane = k / sz, r = k % sz;
        long long col = (long long)sp[d0 + plane] + r - sz;
        emit(y0 + r, col, elms[e0 + nce + k]);
      }
    } else if (W > 1) {
      unsigned c0 = sp[p0 + V + 1], c1b = sp[p0 + V + 1 + P];
      unsigned nse = (W * (c1b - c0) - (e1 - e0)) / (W - 1);
      for (unsigned k = 0; k < nse; k++)
        emit(y0 + k % sz, cols[c0 + k], elms[e0 + k]);
      unsigned nbc = c1b - c0 - nse;              /* block column slots */
      for (unsigned k = 0; k < nbc; k++) {
        unsigned b = k / sz, r = k % sz;
        unsigned pair = cols[c0 + nse + k];
        for (unsigned m = 0; m < W; m++)
          emit(y0 + r, (long long)pair * W + m, elms[e0 + nse + (b * W + m) * sz + r]);
      }
    } else {
      for (unsigned k = 0; k < e1 - e0; k++)
        emit(y0 + k % sz, cols[e0 + k], elms[e0 + k]);
    }
  }
  return n > cap ? -1 : n;
}

/* ------------------------------------------------------------------ sblas ---------- */
/* sblas::dot (util/cuda/sblas.cu:126-153): s = sum x_i y_i. */
double orc_dot(const double *x, const double *y, int n)
{
  double s = 0;
  for (int i = 0; i < n; i++) s += x[i] * y[i];
  return s;
}

/* sblas::ppcg_update_scalars (util/cuda/sblas.cu:156-174). */
void orc_ppcg_update_scalars(double *alpha, double *beta, double gamma, double gammaold,
                             double delta)
{
  double b = gamma / gammaold;
  *beta = b;
  *alpha = gamma / (delta - b * gamma / *alpha);
}

/* sblas::ppcg_update_vectors (util/cuda/sblas.cu:177-237): Ghysels-Vanroose recurrences
 * with resnorm = sum |r_i| (an L1 norm), gamma = (r,u), delta = (w,u). */
void orc_ppcg_update_vectors(double *resnorm, double *gamma, double *delta, double a,
                             double b, const double *n, const double *m, double *p,
                             double *s, double *q, double *z, double *x, double *r,
                             double *u, double *w, int n_elts)
{
  double rn = 0, g = 0, d = 0;
  for (int i = 0; i < n_elts; i++) {
    double wi = w[i], ui = u[i];
    double zi = n[i] + b * z[i]; z[i] = zi;
    double qi = m[i] + b * q[i]; q[i] = qi;
    double si = wi + b * s[i];   s[i] = si;
    double pi = ui + b * p[i];   p[i] = pi;
    ui = ui - a * qi; u[i] = ui;
    wi = wi - a * zi; w[i] = wi;
    double ri = r[i] - a * si; r[i] = ri;
    x[i] += a * pi;
    rn += std::fabs(ri);
    g += ri * ui;
    d += wi * ui;
  }
  *resnorm = rn; *gamma = g; *delta = d;
}

/* sblas::copy_indexed / add_indexed (util/cuda/sblas.cu:18-51). */
void orc_copy_indexed(double *dst, const double *src, const int *idx, int n)
{
  for (int k = 0; k < n; k++) dst[k] = src[idx[k]];
}
void orc_add_indexed(double *dst, const double *src, const int *idx, int n)
{
  for (int k = 0; k < n; k++) dst[idx[k]] += src[k];
}

/* --------------------------------------------------------------------- PCG ---------- */
struct csr_view { int n; const int *rp, *ci; const double *a; };

static void spmv(const csr_view &A, const double *x, double *y)
{
  orc_csr_spmv(A.n, A.rp, A.ci, A.a, x, y);
}

static double pdot(const double *x, const double *y, int n)
{
  /* fixed-order blocked sum: deterministic for a given thread count */
  int nt = omp_get_max_threads();
  std::vector<double> part(nt, 0.0);
#pragma omp parallel
  {
    int t = omp_get_thread_num(), T = omp_get_num_threads();
    long long lo = (long long)n * t / T, hi = (long long)n * (t + 1) / T;
    double s = 0;
    for (long long i = lo; i < hi; i++) s += x[i] * y[i];
    part[t] = s;
  }
  double s = 0;
  for (double v : part) s += v;
  return s;
}

/* M v = M2 (M1 v) with M = Ainv2 * Ainv1 (fastainv/fastainv.cpp:521-558); identity when
 * both are absent. */
static void apply_prec(const csr_view *M1, const csr_view *M2, const double *v, double *t,
                       double *out, int n)
{
  if (!M1) { std::memcpy(out, v, sizeof(double) * n); return; }
  spmv(*M1, v, t);
  spmv(*M2, t, out);
}

static double true_relres(const csr_view &A, const double *b, const double *x, double *tmp,
                          double bnorm)
{
  spmv(A, x, tmp);
  int n = A.n;
#pragma omp parallel for
  for (int i = 0; i < n; i++) tmp[i] = b[i] - tmp[i];
  return std::sqrt(pdot(tmp, tmp, n)) / bnorm;
}

/* Preconditioned CG (Hestenes-Stiefel) with M = M2*M1.  Stops when the TRUE relative
 * residual ||b - A x||_2 / ||b||_2 <= tol (checked every `check` iterations and whenever
 * the recursive residual says so).  Returns iterations, or -1 if maxit reached. */
int orc_pcg(int n, const int *rp, const int *ci, const double *a,
            const int *m1rp, const int *m1ci, const double *m1a,
            const int *m2rp, const int *m2ci, const double *m2a,
            const double *b, double *x, double tol, int maxit, double *relres)
{
  csr_view A{n, rp, ci, a}, M1{n, m1rp, m1ci, m1a}, M2{n, m2rp, m2ci, m2a};
  const csr_view *pm1 = m1rp ? &M1 : nullptr, *pm2 = m1rp ? &M2 : nullptr;
  std::vector<double> r(n), z(n), p(n), q(n), t(n), tmp(n);
  double bnorm = std::sqrt(pdot(b, b, n));
  if (bnorm == 0) bnorm = 1;
  spmv(A, x, r.data());
  for (int i = 0; i < n; i++) r[i] = b[i] - r[i];
  apply_prec(pm1, pm2, r.data(), t.data(), z.data(), n);
  p = z;
  double rz = pdot(r.data(), z.data(), n);
  for (int it = 1; it <= maxit; it++) {
    spmv(A, p.data(), q.data());
    double alpha = rz / pdot(p.data(), q.data(), n);
#pragma omp parallel for
    for (int i = 0; i < n; i++) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
    double rr = std::sqrt(pdot(r.data(), r.data(), n)) / bnorm;
    if (rr <= tol) {
      double tr = true_relres(A, b, x, tmp.data(), bnorm);
      if (tr <= tol) { *relres = tr; return it; }
    }
    apply_prec(pm1, pm2, r.data(), t.data(), z.data(), n);
    double rz2 = pdot(r.data(), z.data(), n);
    double beta = rz2 / rz;
    rz = rz2;
#pragma omp parallel for
    for (int i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
  }
  *relres = true_relres(A, b, x, tmp.data(), bnorm);
  return -1;
}

/* Pipelined PCG (Ghysels & Vanroose 2014) built from exactly the reference's primitives:
 * ppcg_update_scalars (sblas.cu:156-174) and ppcg_update_vectors (sblas.cu:177-237), with
 * M = M2*M1 applied as two SpMVs and n = A m.  First iteration: gammaold = +inf gives
 * beta = 0 and alpha = gamma/delta.  Stops on the TRUE relative residual as orc_pcg. */
int orc_ppcg(int n, const int *rp, const int *ci, const double *a,
             const int *m1rp, const int *m1ci, const double *m1a,
             const int *m2rp, const int *m2ci, const double *m2a,
             const double *b, double *x, double tol, int maxit, double *relres,
             double *l1_out)
{
  csr_view A{n, rp, ci, a}, M1{n, m1rp, m1ci, m1a}, M2{n, m2rp, m2ci, m2a};
  const csr_view *pm1 = m1rp ? &M1 : nullptr, *pm2 = m1rp ? &M2 : nullptr;
  std::vector<double> r(n), u(n), w(n), m(n), nn(n), z(n, 0.0), q(n, 0.0), s(n, 0.0),
      p(n, 0.0), t(n), tmp(n);
  double bnorm = std::sqrt(pdot(b, b, n));
  if (bnorm == 0) bnorm = 1;
  spmv(A, x, r.data());
  for (int i = 0; i < n; i++) r[i] = b[i] - r[i];
  apply_prec(pm1, pm2, r.data(), t.data(), u.data(), n);
  spmv(A, u.data(), w.data());
  double gamma = pdot(r.data(), u.data(), n), delta = pdot(w.data(), u.data(), n);
  double gammaold = INFINITY, alpha = 1.0, beta = 0.0, l1 = 0;
  for (int it = 1; it <= maxit; it++) {
    apply_prec(pm1, pm2, w.data(), t.data(), m.data(), n);
    spmv(A, m.data(), nn.data());
    orc_ppcg_update_scalars(&alpha, &beta, gamma, gammaold, delta);
    double g2, d2;
    orc_ppcg_update_vectors(&l1, &g2, &d2, alpha, beta, nn.data(), m.data(), p.data(),
                            s.data(), q.data(), z.data(), x, r.data(), u.data(), w.data(), n);
    gammaold = gamma; gamma = g2; delta = d2;
    double rr = std::sqrt(pdot(r.data(), r.data(), n)) / bnorm;
    if (rr <= tol) {
      double tr = true_relres(A, b, x, tmp.data(), bnorm);
      if (tr <= tol) { *relres = tr; if (l1_out) *l1_out = l1; return it; }
    }
  }
  *relres = true_relres(A, b, x, tmp.data(), bnorm);
  if (l1_out) *l1_out = l1;
  return -1;
}

} /* extern "C" */

/* ------------------------------------------------------ approximate inverse -------- */
/* Restatement of fastainv_sym (fastainv/fastainv.cpp:395-566) and its helpers.
 *
 * W is unit upper triangular, stored by column: col[j] lists (row k < j, value) sorted by
 * k, the unit diagonal implicit.  Storage follows chunky_storage's capacity accounting
 * (fastainv.cpp:13-97): max_elts is clamped to maxneed(n, 8) (388-393, 409), the pool holds
 * max_elts/8 chunks of 8, a column owns ceil(len/8) chunks, and a replacement longer than
 * the pool allows is truncated to its first entries; a replacement of length 0 leaves the
 * old column untouched (fastainv.cpp:76-79). */
namespace {

struct merge_heap {
  /* rowmerge_heap (fastainv.cpp:99-213): binary min-heap on column with a sentinel slot;
   * ties never move an element down and pick the right child, which fixes the order in
   * which equal columns are summed. */
  std::vector<int> col, idx, ofs, lim;
  std::vector<double> fac;
  int sz = 0;
  explicit merge_heap(int cap) : col(cap + 1), idx(cap + 1), ofs(cap), lim(cap), fac(cap) {}
  void add(int o, int l, int c, double f)
  {
    col[sz] = c; idx[sz] = sz; fac[sz] = f; ofs[sz] = o; lim[sz] = l; sz++;
  }
  void sift(int i, int c, int id)
  {
    /* place (c, id) at slot i, moving the smaller child up while it is strictly smaller */
    for (;;) {
      int l = 2 * i + 1, r = l + 1;
      if (l >= sz) { col[i] = c; idx[i] = id; return; }
      int mc = std::min(col[l], col[r]);
      if (!(mc < c)) { col[i] = c; idx[i] = id; return; }
      int ch = col[l] < col[r] ? l : r;
      col[i] = mc; idx[i] = idx[ch];
      i = ch;
    }
  }
  void down(int i)
  {
    if (2 * i + 1 >= sz) return;
    sift(i, col[i], idx[i]);
  }
  void build()
  {
    col[sz] = INT_MAX;
    for (int i = sz / 2 - 1; i >= 0; i--) down(i);
  }
  void pop()
  {
    sz--;
    col[0] = col[sz]; idx[0] = idx[sz];
    col[sz] = INT_MAX;
    down(0);
  }
  void bump(int c) { col[0] = c; down(0); }
};

template<class real>
struct wstore {
  std::vector<std::vector<int>> row;
  std::vector<std::vector<real>> val;
  long long pool, used = 0;
  wstore(int n, long long chunks) : row(n), val(n), pool(chunks) {}
  static long long nch(long long len) { return (len + 7) / 8; }
  void replace(int c, std::vector<int> &nr, std::vector<real> &nv)
  {
    long long oldlen = (long long)row[c].size();
    long long avail = (pool - used + nch(oldlen)) * 8;
    long long n = std::min<long long>((long long)nr.size(), avail);
    if (n <= 0) return;
    used += std::max(0LL, nch(n) - nch(oldlen));
    row[c].assign(nr.begin(), nr.begin() + n);
    val[c].assign(nv.begin(), nv.begin() + n);
  }
};

/* u = A w_j by k-way merge of the rows of A selected by w_j (fastainv.cpp:215-264);
 * A symmetric, so A w_j = sum_k w_kj A[k,:].  Products and sums in in_real (double). */
template<class real>
int a_times_wj(const wstore<real> &W, int j, const int *rp, const int *ci, const double *a,
               int *uidx, double *uval)
{
  const std::vector<int> &wr = W.row[j];
  const std::vector<real> &wv = W.val[j];
  merge_heap h((int)wr.size() + 1);
  for (size_t t = 0; t < wr.size(); t++) {
    int k = wr[t];
    if (rp[k] < rp[k + 1]) h.add(rp[k], rp[k + 1], ci[rp[k]], (double)wv[t]);
  }
  if (rp[j] < rp[j + 1]) h.add(rp[j], rp[j + 1], ci[rp[j]], 1.0);
  h.build();
  int cnt = 0, prev = -1;
  while (h.sz > 0) {
    int c = h.col[0], id = h.idx[0];
    double add = a[h.ofs[id]] * h.fac[id];
    if (c != prev) { uidx[cnt] = c; uval[cnt] = add; cnt++; prev = c; }
    else uval[cnt - 1] += add;
    if (++h.ofs[id] < h.lim[id]) h.bump(ci[h.ofs[id]]);
    else h.pop();
  }
  return cnt;
}

/* w_i <- drop(w_i + f * (w_j + e_j), |.| <= droptol) (fastainv.cpp:266-386); float math
 * in `real`, emitted in ascending row order. */
template<class real>
void update_wi(wstore<real> &W, int j, int i, real f, double droptol, std::vector<int> &mr,
               std::vector<real> &mv)
{
  const std::vector<int> &ir = W.row[i], &jr = W.row[j];
  const std::vector<real> &iv = W.val[i], &jv = W.val[j];
  mr.clear(); mv.clear();
  auto put = [&](int r, real e) { if (std::abs(e) > droptol) { mr.push_back(r); mv.push_back(e); } };
  size_t a = 0, b = 0, na = ir.size(), nb = jr.size();
  bool diag_done = false;               /* implicit unit entry of w_j at row j */
  while (a < na || b < nb || !diag_done) {
    /* next row on the w_j side: its stored rows (all < j), then j itself */
    bool jside = b < nb || !diag_done;
    int jrow = b < nb ? jr[b] : j;
    real jval = b < nb ? (real)(jv[b] * f) : f;
    if (a < na && (!jside || ir[a] < jrow)) { put(ir[a], iv[a]); a++; continue; }
    if (a < na && ir[a] == jrow) {
      put(jrow, (real)(iv[a] + jval)); a++;
    } else {
      put(jrow, jval);
    }
    if (b < nb) b++; else diag_done = true;
  }
  W.replace(i, mr, mv);
}

template<class real>
bool fastainv_restated(int n, const int *rp, const int *ci, const double *a,
                       long long max_elts, double droptol,
                       std::vector<int> &r1, std::vector<int> &c1, std::vector<real> &v1,
                       std::vector<int> &r2, std::vector<int> &c2, std::vector<real> &v2)
{
  if (n <= 0) return false;
  long long nb = ((long long)n + 7) / 8;
  long long maxneed = nb * (nb + 1) / 2 * 64;
  max_elts = std::min(max_elts, maxneed);
  wstore<real> W(n, max_elts / 8);
  std::vector<real> D(n);
  std::vector<int> uidx(n), mr;
  std::vector<double> uval(n);
  std::vector<real> mv;
  mr.reserve(n); mv.reserve(n);
  for (int j = 0; j < n; j++) {
    const int *ui_p; const double *uv_p; int un;
    real Djj = 0;
    int ui = 0;
    if (W.row[j].empty()) {
      ui_p = ci + rp[j]; uv_p = a + rp[j]; un = rp[j + 1] - rp[j];
    } else {
      un = a_times_wj(W, j, rp, ci, a, uidx.data(), uval.data());
      ui_p = uidx.data(); uv_p = uval.data();
      /* D_jj = u . w_j over the stored part (fastainv.cpp:441-466) */
      const std::vector<int> &wr = W.row[j];
      const std::vector<real> &wv = W.val[j];
      size_t t = 0;
      while (ui < un && t < wr.size()) {
        if (ui_p[ui] < wr[t]) ui++;
        else if (ui_p[ui] > wr[t]) t++;
        else { Djj = (real)((double)Djj + uv_p[ui] * (double)wv[t]); ui++; t++; }
      }
    }
    /* + u_j (implicit unit diagonal of w_j) */
    while (ui < un) {
      if (ui_p[ui] < j) ui++;
      else if (ui_p[ui] > j) break;
      else { Djj = (real)((double)Djj + uv_p[ui]); ui++; }
    }
    if (Djj == 0) { D[j] = 0; continue; }      /* breakdown (fastainv.cpp:468-472) */
    Djj = (real)(1.0 / (double)Djj);
    for (; ui < un; ui++) {
      int i = ui_p[ui];
      real f = (real)(-uv_p[ui] * (double)Djj);
      update_wi(W, j, i, f, droptol, mr, mv);
    }
    D[j] = Djj;
  }
  /* Ainv2: row r holds W[r,c] sqrt|D_c| at column n-1-c, plus the diagonal (489-532). */
  r2.assign(n + 1, 0);
  for (int c = 0; c < n; c++)
    for (int rr : W.row[c]) r2[rr + 1]++;
  for (int i = 1; i <= n; i++) r2[i] += r2[i - 1] + 1;
  c2.assign(r2[n], 0); v2.assign(r2[n], 0);
  std::vector<int> cur(n, 0);
  std::vector<real> Ds(n);
  for (int i = 0; i < n; i++) {
    int c = n - 1 - i;
    real sc = std::sqrt(std::abs(D[c]));
    Ds[c] = std::signbit(D[c]) ? -sc : sc;
    for (size_t t = 0; t < W.row[c].size(); t++) {
      int rr = W.row[c][t];
      int off = r2[rr] + cur[rr]++;
      c2[off] = i; v2[off] = W.val[c][t] * sc;
    }
    int off = r2[c] + cur[c]++;
    c2[off] = i; v2[off] = sc;
  }
  /* Ainv1: row i = column c = n-1-i of W scaled by sign(D_c) sqrt|D_c|, then the diagonal
   * (534-558). */
  r1.assign(1, 0); c1.clear(); v1.clear();
  for (int i = 0; i < n; i++) {
    int c = n - 1 - i;
    for (size_t t = 0; t < W.row[c].size(); t++) {
      c1.push_back(W.row[c][t]); v1.push_back(W.val[c][t] * Ds[c]);
    }
    c1.push_back(c); v1.push_back(Ds[c]);
    r1.push_back((int)c1.size());
  }
  return true;
}

struct fai_result {
  std::vector<int> r1, c1, r2, c2;
  std::vector<double> v1, v2;
};

} /* namespace */

extern "C" {

/* out_double = 0: fastainv_sym<double,float>; 1: fastainv_sym<double,double>.  Returns an
 * opaque result (NULL if n <= 0). */
void *orc_fastainv_sym(int out_double, int n, const int *rp, const int *ci, const double *a,
                       long long max_elts, double droptol)
{
  fai_result *R = new fai_result;
  bool ok;
  if (out_double) {
    ok = fastainv_restated<double>(n, rp, ci, a, max_elts, droptol, R->r1, R->c1, R->v1,
                                   R->r2, R->c2, R->v2);
  } else {
    std::vector<float> f1, f2;
    ok = fastainv_restated<float>(n, rp, ci, a, max_elts, droptol, R->r1, R->c1, f1, R->r2,
                                  R->c2, f2);
    R->v1.assign(f1.begin(), f1.end());
    R->v2.assign(f2.begin(), f2.end());
  }
  if (!ok) { delete R; return nullptr; }
  return R;
}

int orc_fastainv_nnz(void *h, int which)
{
  fai_result *R = (fai_result *)h;
  return which == 1 ? (int)R->c1.size() : (int)R->c2.size();
}

void orc_fastainv_copy(void *h, int which, int *rp, int *ci, double *v)
{
  fai_result *R = (fai_result *)h;
  const std::vector<int> &r = which == 1 ? R->r1 : R->r2, &c = which == 1 ? R->c1 : R->c2;
  const std::vector<double> &x = which == 1 ? R->v1 : R->v2;
  std::copy(r.begin(), r.end(), rp);
  std::copy(c.begin(), c.end(), ci);
  std::copy(x.begin(), x.end(), v);
}

void orc_fastainv_free(void *h) { delete (fai_result *)h; }

/* ------------------------------------------------------------- DroptolTuner --------- */
/* Restatement of DroptolTuner (fastainv/droptol-tuning.h:4-103): golden-section search on
 * log2(droptol) driven by measured solve times. */
struct orc_dtt {
  int phase;  /* 0 = first probe of m1, 1 = probing m1, 2 = probing m2 */
  double lo, hi, m1, m2, best, t1, t2, tmin;
  int iters;
};
static const double kGolden = 0.3819660112501051518;  /* (3 - sqrt 5) / 2 */

void *orc_dtt_new(double lo, double hi)
{
  orc_dtt *t = new orc_dtt();
  t->phase = 0; t->lo = lo; t->hi = hi;
  t->m1 = lo + (hi - lo) * kGolden;
  t->m2 = lo + (hi - t->m1);
  t->iters = 0;
  return t;
}
double orc_dtt_get(void *h) { orc_dtt *t = (orc_dtt *)h; return std::exp2(t->phase == 2 ? t->m2 : t->m1); }
double orc_dtt_best(void *h) { return std::exp2(((orc_dtt *)h)->best); }
void orc_dtt_note(void *h, double time)
{
  orc_dtt *t = (orc_dtt *)h;
  if (t->phase == 0) {
    t->tmin = time; t->best = t->m1; t->t1 = time; t->phase = 2;
    return;
  }
  if (t->phase == 2) {
    if (t->tmin > time) { t->tmin = time; t->best = t->m2; }
    t->t2 = time;
  } else {
    if (t->tmin > time) { t->tmin = time; t->best = t->m1; }
    t->t1 = time;
  }
  if (t->t1 < t->t2) {        /* minimum lies left of m2 */
    t->hi = t->m2; t->m2 = t->m1; t->m1 = t->lo + (t->hi - t->m2);
    t->t2 = t->t1; t->phase = 1;
  } else {
    t->lo = t->m1; t->m1 = t->m2; t->m2 = t->lo + (t->hi - t->m1);
    t->t1 = t->t2; t->phase = 2;
  }
  t->iters++;
}
void orc_dtt_restart(void *h)
{
  orc_dtt *t = (orc_dtt *)h;
  t->phase = 0; t->lo = t->m1 - 1; t->hi = t->m1 + 1;
  t->m1 = t->lo + (t->hi - t->lo) * kGolden;
  t->m2 = t->lo + (t->hi - t->m1);
  t->iters = 0;
}
void orc_dtt_reset_phase(void *h) { ((orc_dtt *)h)->phase = 0; }
int orc_dtt_iters(void *h) { return ((orc_dtt *)h)->iters; }
void orc_dtt_free(void *h) { delete (orc_dtt *)h; }

} /* extern "C" */
