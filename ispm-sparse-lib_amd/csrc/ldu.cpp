/* ldu.cpp -- OpenFOAM lduMatrix addressing to CSR, and the coefficient scatter that lets a
 * time-stepping caller refresh a plan's values without converting again.
 *
 * No counterpart in the reference library: the conversion lives in its caller (the
 * foam-extend ispm solver, README.md:3-9 of the reference), which hands the library a
 * csr_matrix built from the finite-volume matrix.  lduMatrix stores a diagonal coefficient
 * per cell and, per internal face f between cells l = lower_addr[f] < u = upper_addr[f], the
 * upper coefficient A[l][u] and the lower coefficient A[u][l] (the same array when the
 * matrix is symmetric).  The CSR has, in every row, the diagonal and that row's face
 * coefficients with columns ascending. */
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>
#include "internal.h"
#include "ispm.h"
#include "fastainv/droptol-tuning.h"

using ispm::fail;

extern "C" int ispm_ldu_to_csr(int n, int n_faces, const int *lower_addr, const int *upper_addr,
                               const double *diag, const double *upper, const double *lower,
                               ispm_csr_out *out, int *pos)
{
  if (!out) return fail(ISPM_ERR_INVALID, "ldu_to_csr: null output");
  std::memset(out, 0, sizeof *out);
  if (n < 0 || n_faces < 0) return fail(ISPM_ERR_INVALID, "ldu_to_csr: negative size");
  if ((n && !diag) || (n_faces && (!lower_addr || !upper_addr || !upper)))
    return fail(ISPM_ERR_INVALID, "ldu_to_csr: null argument");
  const long long nnz = (long long)n + 2LL * n_faces;
  if (nnz > 0x7fffffffLL) return fail(ISPM_ERR_UNSUPPORTED, "ldu_to_csr: more than 2^31 - 1 entries");
  for (int f = 0; f < n_faces; f++) {
    const int l = lower_addr[f], u = upper_addr[f];
    if (l < 0 || u >= n || l >= u)
      return fail(ISPM_ERR_INVALID, "ldu_to_csr: face %d joins cells %d and %d (need 0 <= lower < upper < n)", f, l, u);
  }
  /* row counts, then each row's entries as (column, source index) sorted by column; source
   * index i < n is diag[i], n + f is upper[f] (row l, column u), n + n_faces + f is
   * lower[f] (row u, column l) */
  std::vector<int> rp((size_t)n + 1, 0);
  for (int i = 0; i < n; i++) rp[i + 1] = 1;
  for (int f = 0; f < n_faces; f++) { rp[lower_addr[f] + 1]++; rp[upper_addr[f] + 1]++; }
  for (int i = 0; i < n; i++) rp[i + 1] += rp[i];
  std::vector<std::pair<int, int>> ent((size_t)nnz);
  std::vector<int> fill(rp.begin(), rp.end() - 1);
  for (int i = 0; i < n; i++) ent[fill[i]++] = {i, i};
  for (int f = 0; f < n_faces; f++) {
    const int l = lower_addr[f], u = upper_addr[f];
    ent[fill[l]++] = {u, n + f};
    ent[fill[u]++] = {l, n + n_faces + f};
  }
  for (int i = 0; i < n; i++) {
    std::sort(ent.begin() + rp[i], ent.begin() + rp[i + 1]);
    for (int k = rp[i] + 1; k < rp[i + 1]; k++)
      if (ent[k].first == ent[k - 1].first)
        return fail(ISPM_ERR_INVALID, "ldu_to_csr: two faces join cells %d and %d", std::min(i, ent[k].first),
                    std::max(i, ent[k].first));
  }
  out->row_ptr = static_cast<int *>(std::malloc(sizeof(int) * ((size_t)n + 1)));
  out->cols = static_cast<int *>(std::malloc(sizeof(int) * std::max<size_t>((size_t)nnz, 1)));
  out->elms = std::malloc(sizeof(double) * std::max<size_t>((size_t)nnz, 1));
  if (!out->row_ptr || !out->cols || !out->elms) {
    ispm_csr_out_free(out);
    return fail(ISPM_ERR_NOMEM, "ldu_to_csr: out of host memory");
  }
  const double *lo = lower ? lower : upper;
  double *v = static_cast<double *>(out->elms);
  std::copy(rp.begin(), rp.end(), out->row_ptr);
  for (long long k = 0; k < nnz; k++) {
    const int c = ent[k].first, src = ent[k].second;
    out->cols[k] = c;
    v[k] = src < n ? diag[src] : src < n + n_faces ? upper[src - n] : lo[src - n - n_faces];
    if (pos) pos[src] = (int)k;
  }
  out->n_rows = out->n_cols = n;
  out->n_nz = (int)nnz;
  out->elm_type = ISPM_F64;
  return ISPM_OK;
}

extern "C" int ispm_ldu_values(int n, int n_faces, const int *pos, const double *diag, const double *upper,
                               const double *lower, double *vals)
{
  if (n < 0 || n_faces < 0) return fail(ISPM_ERR_INVALID, "ldu_values: negative size");
  if (!vals || !pos || (n && !diag) || (n_faces && !upper)) return fail(ISPM_ERR_INVALID, "ldu_values: null argument");
  const double *lo = lower ? lower : upper;
  const long long nnz = (long long)n + 2LL * n_faces;
  for (long long s = 0; s < nnz; s++)
    if (pos[s] < 0 || pos[s] >= nnz) return fail(ISPM_ERR_INVALID, "ldu_values: position out of range");
  for (int i = 0; i < n; i++) vals[pos[i]] = diag[i];
  for (int f = 0; f < n_faces; f++) {
    vals[pos[n + f]] = upper[f];
    vals[pos[n + n_faces + f]] = lo[f];
  }
  return ISPM_OK;
}

/* ---- DroptolTuner (fastainv/droptol-tuning.h) behind the C-ABI ---- */
struct ispm_droptol_tuner {
  DroptolTuner t;
  ispm_droptol_tuner(double lo, double hi) : t(lo, hi) {}
};

extern "C" int ispm_droptol_tuner_create(double log2_lo, double log2_hi, ispm_droptol_tuner **out)
{
  if (!out) return fail(ISPM_ERR_INVALID, "droptol_tuner_create: null output");
  *out = nullptr;
  if (!(log2_lo < log2_hi)) return fail(ISPM_ERR_INVALID, "droptol_tuner_create: need lo < hi");
  *out = new ispm_droptol_tuner(log2_lo, log2_hi);
  return ISPM_OK;
}
extern "C" double ispm_droptol_tuner_droptol(const ispm_droptol_tuner *t) { return t ? t->t.getDroptol() : 0.0; }
extern "C" double ispm_droptol_tuner_best(const ispm_droptol_tuner *t) { return t ? t->t.getBestDroptol() : 0.0; }
extern "C" void ispm_droptol_tuner_note_time(ispm_droptol_tuner *t, double seconds) { if (t) t->t.noteTime(seconds); }
extern "C" void ispm_droptol_tuner_restart(ispm_droptol_tuner *t) { if (t) t->t.restart(); }
extern "C" int ispm_droptol_tuner_iterations(const ispm_droptol_tuner *t) { return t ? t->t.iterations() : 0; }
extern "C" void ispm_droptol_tuner_destroy(ispm_droptol_tuner *t) { delete t; }
