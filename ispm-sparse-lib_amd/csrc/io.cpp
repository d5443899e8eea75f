/* io.cpp -- on-disk formats: MatrixMarket coordinate files (real/integer/pattern, general
 * or symmetric) to and from CSR, a sparsity-pattern fingerprint, and a persisted plan cache
 * (one text line per tuned matrix: fingerprint, types, format and launch shape) so that an
 * exhaustive plan is tuned once per pattern.  The reference has no persistence; its closest
 * analogue is plan reuse across time steps (plan/plan.h:39-56). */
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>
#include "ispm.h"

namespace ispm { int set_error_message(int code, const char *msg); }
using ispm::set_error_message;

static int mm_read(const char *path, ispm_csr_out *out)
{
  FILE *f = std::fopen(path, "r");
  if (!f) return set_error_message(ISPM_ERR_INVALID, "mm_read: cannot open file");
  char line[1024];
  if (!std::fgets(line, sizeof line, f) || std::strncmp(line, "%%MatrixMarket", 14) != 0) {
    std::fclose(f);
    return set_error_message(ISPM_ERR_INVALID, "mm_read: missing %%MatrixMarket banner");
  }
  std::string banner(line);
  for (auto &c : banner) c = (char)std::tolower((unsigned char)c);
  const bool coord = banner.find("coordinate") != std::string::npos;
  const bool pattern = banner.find("pattern") != std::string::npos;
  const bool symmetric = banner.find("symmetric") != std::string::npos;
  const bool complex_ = banner.find("complex") != std::string::npos;
  if (!coord || complex_) { std::fclose(f); return set_error_message(ISPM_ERR_UNSUPPORTED, "mm_read: only real coordinate files"); }
  long long nr = 0, nc = 0, ne = 0;
  while (std::fgets(line, sizeof line, f)) {
    if (line[0] == '%') continue;
    if (std::sscanf(line, "%lld %lld %lld", &nr, &nc, &ne) == 3) break;
  }
  if (nr <= 0 || nc <= 0 || nr > 0x7fffffff || nc > 0x7fffffff) { std::fclose(f); return set_error_message(ISPM_ERR_INVALID, "mm_read: bad size line"); }
  /* entries are stored with 32-bit row pointers: more than INT_MAX of them (after the
   * mirror entries of a symmetric file) cannot be represented */
  if (ne < 0) { std::fclose(f); return set_error_message(ISPM_ERR_INVALID, "mm_read: negative entry count"); }
  if ((symmetric ? 2 * ne : ne) > 0x7fffffffLL) { std::fclose(f); return set_error_message(ISPM_ERR_UNSUPPORTED, "mm_read: more than 2^31-1 entries"); }
  std::vector<long long> key;
  std::vector<double> val;
  key.reserve(symmetric ? 2 * ne : ne);
  val.reserve(symmetric ? 2 * ne : ne);
  for (long long e = 0; e < ne; e++) {
    long long i, j;
    double v = 1.0;
    int got = pattern ? std::fscanf(f, "%lld %lld", &i, &j) : std::fscanf(f, "%lld %lld %lf", &i, &j, &v);
    if (got != (pattern ? 2 : 3) || i < 1 || j < 1 || i > nr || j > nc) {
      std::fclose(f);
      return set_error_message(ISPM_ERR_INVALID, "mm_read: bad entry");
    }
    key.push_back((i - 1) * nc + (j - 1));
    val.push_back(v);
    if (symmetric && i != j) { key.push_back((j - 1) * nc + (i - 1)); val.push_back(v); }
  }
  std::fclose(f);
  std::vector<size_t> ord(key.size());
  for (size_t k = 0; k < ord.size(); k++) ord[k] = k;
  std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return key[a] < key[b]; });
  std::vector<int> rp(nr + 1, 0), ci;
  std::vector<double> v;
  for (size_t t = 0; t < ord.size(); t++) {
    const long long k = key[ord[t]];
    if (!ci.empty() && t > 0 && key[ord[t - 1]] == k) { v.back() += val[ord[t]]; continue; }   /* duplicates add */
    rp[k / nc + 1]++;
    ci.push_back((int)(k % nc));
    v.push_back(val[ord[t]]);
  }
  for (long long r = 0; r < nr; r++) rp[r + 1] += rp[r];
  out->n_rows = (int)nr;
  out->n_cols = (int)nc;
  out->n_nz = (int)ci.size();
  out->elm_type = ISPM_F64;
  out->row_ptr = (int *)std::malloc(sizeof(int) * (nr + 1));
  out->cols = (int *)std::malloc(sizeof(int) * (ci.size() + 1));
  out->elms = std::malloc(sizeof(double) * (v.size() + 1));
  if (!out->row_ptr || !out->cols || !out->elms) return set_error_message(ISPM_ERR_NOMEM, "mm_read: out of memory");
  std::copy(rp.begin(), rp.end(), out->row_ptr);
  std::copy(ci.begin(), ci.end(), out->cols);
  std::copy(v.begin(), v.end(), (double *)out->elms);
  return ISPM_OK;
}

extern "C" int ispm_mm_read(const char *path, ispm_csr_out *out)
{
  if (!path || !out) return set_error_message(ISPM_ERR_INVALID, "mm_read: null argument");
  std::memset(out, 0, sizeof *out);
  int st;
  try {
    st = mm_read(path, out);
  } catch (const std::bad_alloc &) {
    st = set_error_message(ISPM_ERR_NOMEM, "mm_read: out of memory");
  } catch (const std::exception &) {
    st = set_error_message(ISPM_ERR_INVALID, "mm_read: cannot represent the file's matrix");
  }
  if (st != ISPM_OK) {
    /* nothing half-built reaches the caller */
    std::free(out->row_ptr);
    std::free(out->cols);
    std::free(out->elms);
    std::memset(out, 0, sizeof *out);
  }
  return st;
}

extern "C" int ispm_mm_write(const char *path, const ispm_csr *A)
{
  if (!path || !A || !A->row_ptr) return set_error_message(ISPM_ERR_INVALID, "mm_write: null argument");
  FILE *f = std::fopen(path, "w");
  if (!f) return set_error_message(ISPM_ERR_INVALID, "mm_write: cannot open file");
  const long long nnz = A->row_ptr[A->n_rows];
  std::fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n%d %d %lld\n", A->n_rows, A->n_cols, nnz);
  for (int r = 0; r < A->n_rows; r++)
    for (int k = A->row_ptr[r]; k < A->row_ptr[r + 1]; k++) {
      double v = A->elm_type == ISPM_F64 ? static_cast<const double *>(A->elms)[k]
                                         : (double)static_cast<const float *>(A->elms)[k];
      std::fprintf(f, "%d %d %.17g\n", r + 1, A->cols[k] + 1, v);
    }
  const bool ok = std::fclose(f) == 0;
  return ok ? ISPM_OK : set_error_message(ISPM_ERR_RUNTIME, "mm_write: write failed");
}

/* 64-bit FNV-1a over (n_rows, n_cols, row_ptr, cols): identifies a sparsity pattern. */
extern "C" int ispm_csr_fingerprint(const ispm_csr *A, unsigned long long *fp)
{
  if (!A || !fp || !A->row_ptr) return set_error_message(ISPM_ERR_INVALID, "csr_fingerprint: null argument");
  unsigned long long h = 1469598103934665603ull;
  auto mix = [&](const void *p, size_t n) {
    const unsigned char *b = static_cast<const unsigned char *>(p);
    for (size_t i = 0; i < n; i++) { h ^= b[i]; h *= 1099511628211ull; }
  };
  mix(&A->n_rows, sizeof(int));
  mix(&A->n_cols, sizeof(int));
  mix(A->row_ptr, sizeof(int) * ((size_t)A->n_rows + 1));
  if (A->cols) mix(A->cols, sizeof(int) * (size_t)A->row_ptr[A->n_rows]);
  *fp = h;
  return ISPM_OK;
}

extern "C" int ispm_plan_cache_lookup(const char *path, unsigned long long fp, int vt, int et,
                                      unsigned *S, unsigned *W, int *V, int *D, ispm_launch *lp,
                                      double *seconds)
{
  if (!path || !S || !W || !V || !D || !lp) return set_error_message(ISPM_ERR_INVALID, "plan_cache_lookup: null argument");
  FILE *f = std::fopen(path, "r");
  if (!f) return set_error_message(ISPM_ERR_INVALID, "plan_cache_lookup: no cache file");
  char line[256];
  int found = 0;
  while (std::fgets(line, sizeof line, f)) {
    unsigned long long h;
    int a, b, v, d, mode;
    unsigned s, w, n, ls, hh;
    double sec;
    if (std::sscanf(line, "%llx %d %d %u %u %d %d %u %u %u %d %lf", &h, &a, &b, &s, &w, &v, &d, &n, &ls, &hh, &mode, &sec) != 12) continue;
    if (h == fp && a == vt && b == et) {   /* the last matching line wins */
      *S = s; *W = w; *V = v; *D = d;
      lp->N = n; lp->S = ls; lp->H = hh; lp->use_tex_sliceptr = lp->use_tex_x = 0; lp->mode = mode;
      if (seconds) *seconds = sec;
      found = 1;
    }
  }
  std::fclose(f);
  return found ? ISPM_OK : set_error_message(ISPM_ERR_INVALID, "plan_cache_lookup: pattern not cached");
}

extern "C" int ispm_plan_cache_store(const char *path, unsigned long long fp, int vt, int et,
                                     unsigned S, unsigned W, int V, int D, const ispm_launch *lp,
                                     double seconds)
{
  if (!path || !lp) return set_error_message(ISPM_ERR_INVALID, "plan_cache_store: null argument");
  FILE *f = std::fopen(path, "a");
  if (!f) return set_error_message(ISPM_ERR_INVALID, "plan_cache_store: cannot open cache file");
  std::fprintf(f, "%016llx %d %d %u %u %d %d %u %u %u %d %.9g\n", fp, vt, et, S, W, V, D, lp->N, lp->S, lp->H, lp->mode, seconds);
  return std::fclose(f) == 0 ? ISPM_OK : set_error_message(ISPM_ERR_RUNTIME, "plan_cache_store: write failed");
}
