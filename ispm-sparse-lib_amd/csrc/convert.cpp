/* convert.cpp -- CSR -> sliced-ELLPACK conversion on the host.
 *
 * Produces the same layout as the reference's csr_to_slell / csr_to_slell_guess_format
 * (formats/conversion.h:1-492): slice sizing, the three slice encoders, padding rules and
 * the format-guess heuristic are kept step for step, because the layout is part of the
 * API (slell_matrix members and slice_elt_ptr expose it).  The conversion additionally
 * records, for every CSR entry, the index of the element slot it went to; values are then
 * written through that map, which is also how coefficients are refreshed without
 * re-converting.
 *
 * Kept quirks (they change the layout, never the product y = A x):
 *  - padding slots take the column of the row above in the same plane (row 0: column 0)
 *    and value 0 (conversion.h:73-79, 215-228, 328-334);
 *  - the hblock plane-count ternary search calls its second cost with the hblock width and
 *    the single-slot weight swapped (conversion.h:152-154);
 *  - variable slice heights round down to a power of two and are raised to min_height
 *    (conversion.h:37-44, 383-384).
 * Deliberate difference: the reference pads the caller's CSR in place (conversion.h:343);
 * here the input is read-only and padding is applied to an internal row-pointer copy.  The
 * guess samples rows of that padded view, so matrices with fewer rows than it examines are
 * read in bounds (the reference reads past row_ptr there).
 */
#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstring>
#include <vector>
#include "convert.h"

namespace ispm {
namespace conv {

namespace {

enum { kMinSliceHeight = 32, kMaxSliceHeight = 512, kPadRows = 512 };

int floor_pow2(int v)
{
  if (v <= 0) return 0;
  return 1 << (31 - __builtin_clz((unsigned)v));
}

struct csr_in {
  int n;                          /* padded row count */
  const int *rp;                  /* padded row pointer (n+1) */
  const int *ci;
  int elm_size;                   /* bytes per value, for the hblock cost model */
};

/* estimate_slice_height (conversion.h:6-47) */
int var_slice_rows(const csr_in &A, int row0, int cap, long long n_elms)
{
  float waste_per_row = (float)(0.01 * (double)n_elms * 100 / A.n);
  int maxlen = 0, waste = 0, used = 0, k;
  for (k = 0; k < cap && row0 + k < A.n; k++) {
    int len = A.rp[row0 + k + 1] - A.rp[row0 + k];
    int nu = used + len, nw;
    if (len > maxlen) {
      nw = waste + k * (len - maxlen);
    } else {
      nw = waste + maxlen - len;
      len = maxlen;
    }
    if ((k >= 32 && nw > waste_per_row * k) || (k >= 1 && nu + nw > cap * 4)) break;
    maxlen = len;
    waste = nw;
    used = nu;
  }
  return floor_pow2(k);
}

/* Padding: slot (r, plane) of a region starting at base takes the column of (r-1, plane). */
inline unsigned pad_col(const std::vector<unsigned> &cols, size_t base, int r, int plane, int S)
{
  return r > 0 ? cols[base + (size_t)r - 1 + (size_t)plane * S] : 0u;
}

/* copy_csr_elts_vslell (conversion.h:49-81) */
void enc_ell(const csr_in &A, layout &B, int S, int row0)
{
  int w = 0;
  for (int r = 0; r < S; r++) w = std::max(w, A.rp[row0 + r + 1] - A.rp[row0 + r]);
  const size_t base = B.cols.size();
  B.cols.resize(base + (size_t)w * S);
  for (int r = 0; r < S; r++) {
    const int row = row0 + r;
    int k = 0;
    for (int i = A.rp[row]; i < A.rp[row + 1]; i++, k++) {
      size_t slot = base + r + (size_t)k * S;
      B.cols[slot] = (unsigned)A.ci[i];
      B.map[i] = (long long)slot;
    }
    for (; k < w; k++) B.cols[base + r + (size_t)k * S] = pad_col(B.cols, base, r, k, S);
  }
  B.n_elms = B.cols.size();
}

/* hblk_maxslen / hblk_cost (conversion.h:83-108) */
int blk_maxslen(const int *slen, const int *blen, int S, int nb, int hb)
{
  int m = 0;
  for (int r = 0; r < S; r++) {
    int cb = std::min(nb, blen[r]);
    int cs = slen[r] - (nb - cb) + (blen[r] - cb) * hb;
    m = std::max(m, cs);
  }
  return m;
}
int blk_cost(const int *slen, const int *blen, int S, int nb, int hb, int smul, int bmul)
{
  return smul * blk_maxslen(slen, blen, S, nb, hb) + bmul * nb;
}

inline bool block_start(const csr_in &A, int i, int end, int hb)
{
  return !(A.ci[i] & (hb - 1)) && i + hb - 1 < end && A.ci[i + hb - 1] - A.ci[i] == hb - 1;
}

/* copy_csr_elts_hblocks (conversion.h:110-232).  Element slots here are counted in the
 * element array, which for hblock slices differs from the column array. */
void enc_hblock(const csr_in &A, layout &B, int S, int row0, int hb)
{
  std::vector<int> slen(S), blen(S);
  int minb = INT_MAX, maxb = 0;
  for (int r = 0; r < S; r++) {
    const int row = row0 + r, end = A.rp[row + 1];
    int ns = 0, nb = 0, i;
    for (i = A.rp[row]; i + hb - 1 < end;) {
      if (!(A.ci[i] & (hb - 1)) && A.ci[i + hb - 1] - A.ci[i] == hb - 1) { nb++; i += hb; }
      else { ns++; i++; }
    }
    ns += end - i;
    slen[r] = ns;
    blen[r] = nb;
    maxb = std::max(maxb, nb);
    minb = std::min(minb, nb);
  }
  const int wsingle = (int)sizeof(int) + A.elm_size;          /* bytes per single slot */
  const int wblock = (int)sizeof(int) + hb * A.elm_size;      /* bytes per block slot */
  while (maxb - minb >= 3) {
    int d = (maxb - minb) / 3, b1 = minb + d, b2 = b1 + d;
    int cost1 = blk_cost(slen.data(), blen.data(), S, b1, hb, wsingle, wblock);
    /* argument order as in conversion.h:152-154 */
    int cost2 = blk_cost(slen.data(), blen.data(), S, b2, wsingle, hb, wblock);
    if (cost1 <= cost2) maxb = b2;
    else minb = b1;
  }
  int best = INT_MAX, bestb = INT_MAX;
  for (; minb <= maxb; minb++) {
    int c = blk_cost(slen.data(), blen.data(), S, minb, hb, wsingle, wblock);
    if (best > c) { best = c; bestb = minb; }
  }
  const int nb = bestb;
  const int ns = blk_maxslen(slen.data(), blen.data(), S, nb, hb);
  const size_t cofs = B.cols.size(), eofs = B.n_elms;
  const size_t nse = (size_t)ns * S, nbe = (size_t)nb * S;
  const size_t cbofs = cofs + nse, ebofs = eofs + nse;
  B.cols.resize(cofs + nse + nbe);
  for (int r = 0; r < S; r++) {
    const int row = row0 + r, end = A.rp[row + 1];
    int is = 0, ib = 0;
    for (int i = A.rp[row]; i < end;) {
      if (block_start(A, i, end, hb) && ib < nb) {
        B.cols[cbofs + r + (size_t)ib * S] = (unsigned)(A.ci[i] / hb);
        for (int m = 0; m < hb; m++) B.map[i + m] = (long long)(ebofs + r + (size_t)(ib * hb + m) * S);
        ib++;
        i += hb;
      } else if (is < ns) {
        B.cols[cofs + r + (size_t)is * S] = (unsigned)A.ci[i];
        B.map[i] = (long long)(eofs + r + (size_t)is * S);
        is++;
        i++;
      } else {
        /* a single that does not fit the single planes goes to a block slot, partner 0 */
        B.cols[cbofs + r + (size_t)ib * S] = (unsigned)(A.ci[i] / hb);
        B.map[i] = (long long)(ebofs + r + (size_t)(ib * hb + A.ci[i] % hb) * S);
        ib++;
        i++;
      }
    }
    for (; is < ns; is++) B.cols[cofs + r + (size_t)is * S] = pad_col(B.cols, cofs, r, is, S);
    for (; ib < nb; ib++) B.cols[cbofs + r + (size_t)ib * S] = pad_col(B.cols, cbofs, r, ib, S);
  }
  B.n_elms = eofs + nse + nbe * hb;
}

/* find_diagonals (conversion.h:234-284): accept each offset of the (first) longest row
 * that every row of maximal remaining length also has. */
int find_diags(const csr_in &A, int S, int row0, std::vector<int> &diag_pos)
{
  std::vector<int> len(S), pos(S);
  int maxlen = 0, longrow = row0;
  for (int r = 0; r < S; r++) {
    const int row = row0 + r;
    len[r] = A.rp[row + 1] - A.rp[row];
    pos[r] = A.rp[row];
    if (maxlen < len[r]) { maxlen = len[r]; longrow = row; }
  }
  for (int i = A.rp[longrow]; i < A.rp[longrow + 1]; i++) {
    const int off = A.ci[i] - longrow;
    bool skip = false;
    for (int r = 0; r < S; r++) {
      const int row = row0 + r, end = A.rp[row + 1];
      while (pos[r] < end && A.ci[pos[r]] - row < off) pos[r]++;
      if (!(pos[r] < end && A.ci[pos[r]] - row == off) && len[r] == maxlen) { skip = true; break; }
    }
    if (skip) continue;
    diag_pos.push_back(A.ci[i] - (longrow - row0) + S);
    maxlen--;
    for (int r = 0; r < S; r++) {
      const int row = row0 + r;
      if (pos[r] < A.rp[row + 1] && A.ci[pos[r]] - row == off) len[r]--;
    }
  }
  return maxlen;
}

/* copy_csr_elts_diags (conversion.h:286-336) */
void enc_diag(const csr_in &A, layout &B, int S, int row0, std::vector<int> &diag_pos)
{
  const size_t d_start = diag_pos.size();
  const int w = find_diags(A, S, row0, diag_pos);
  const size_t nd = diag_pos.size() - d_start;
  const size_t ccofs = B.cols.size(), ecofs = B.n_elms;
  const size_t nce = (size_t)w * S, edofs = ecofs + nce;
  B.cols.resize(ccofs + nce);
  for (int r = 0; r < S; r++) {
    const int row = row0 + r;
    size_t di = d_start;
    int nc = 0;
    for (int i = A.rp[row]; i < A.rp[row + 1]; i++) {
      const int c = A.ci[i], d = c - r + S;
      while (di < diag_pos.size() && diag_pos[di] < d) di++;
      if (di < diag_pos.size() && diag_pos[di] == d) {
        B.map[i] = (long long)(edofs + r + (di - d_start) * S);
      } else {
        B.cols[ccofs + r + (size_t)nc * S] = (unsigned)c;
        B.map[i] = (long long)(ecofs + r + (size_t)nc * S);
        nc++;
      }
    }
    for (; nc < w; nc++) B.cols[ccofs + r + (size_t)nc * S] = pad_col(B.cols, ccofs, r, nc, S);
  }
  B.n_elms = ecofs + nce + nd * S;
}

}  // namespace

int var_slice_heights(const int *rp_pad, int n_pad, int cap, long long nnz, int min_height,
                      std::vector<int> &h)
{
  csr_in A{n_pad, rp_pad, nullptr, 8};
  h.clear();
  for (int i = 0; i < n_pad;) {
    int sz = var_slice_rows(A, i, cap, nnz);
    if (sz < min_height) sz = min_height;
    if (sz <= 0 || i + sz > n_pad) return -1;
    h.push_back(sz);
    i += sz;
  }
  return 0;
}

int pad_row_ptr(const int *rp, int n_rows, std::vector<int> &out)
{
  const int n_pad = (n_rows + kPadRows - 1) & ~(kPadRows - 1);
  out.assign(rp, rp + n_rows + 1);
  out.resize(n_pad + 1, rp[n_rows]);
  return n_pad;
}

/* csr_to_slell (conversion.h:338-414) */
int convert(const int *rp_in, int n_rows, const int *ci, int n_cols, int n_nz, int elm_size,
            int req_height, int hb, bool var, bool diags, int min_height, layout &B)
{
  std::vector<int> rp;
  const int n = pad_row_ptr(rp_in, n_rows, rp);
  csr_in A{n, rp.data(), ci, elm_size};
  const long long nnz_stored = rp_in[n_rows];
  if (hb > 1) diags = false;
  B = layout();
  B.n_rows = n;
  B.n_cols = n_cols;
  B.n_nz = n_nz;
  B.S = (unsigned)req_height;
  B.W = (unsigned)hb;
  B.V = var;
  B.D = diags;
  B.map.assign((size_t)nnz_stored, -1);
  B.cols.reserve((size_t)nnz_stored * 3 / 2);
  const unsigned P = 1 + (hb > 1) + (var ? 1 : 0) + (diags ? 2 : 0);
  for (unsigned k = 0; k < P; k++) B.sp.push_back(0);
  std::vector<int> diag_pos;
  for (int i = 0; i < n;) {
    int sz = var ? var_slice_rows(A, i, req_height, nnz_stored) : req_height;
    if (sz < min_height) sz = min_height;
    if (sz <= 0) return -1;
    if (i + sz > n) return -2;   /* cannot happen for power-of-two heights dividing 512 */
    if (diags) enc_diag(A, B, sz, i, diag_pos);
    else if (hb > 1) enc_hblock(A, B, sz, i, hb);
    else enc_ell(A, B, sz, i);
    if (B.n_elms > 0xffffffffull || B.cols.size() > 0xffffffffull) return -3;
    B.sp.push_back((unsigned)B.n_elms);
    if (var) B.sp.push_back((unsigned)(i + sz));
    if (hb > 1) B.sp.push_back((unsigned)B.cols.size());
    if (diags) {
      B.sp.push_back((unsigned)B.cols.size());
      B.sp.push_back((unsigned)diag_pos.size());
    }
    B.n_slices++;
    i += sz;
  }
  if (diags) {
    const unsigned dp_off = var ? 3 : 2;
    const unsigned base = (unsigned)B.sp.size();
    for (unsigned s = 0; s <= B.n_slices; s++) B.sp[dp_off + s * P] += base;
    for (int d : diag_pos) B.sp.push_back((unsigned)d);
  }
  return 0;
}

/* csr_to_slell_guess_format (conversion.h:416-492) */
int guess(const int *rp_in, int n_rows, const int *ci, int n_cols, int n_nz, int elm_size,
          int examine_ratio, int min_height, layout &B, guess_info *gi)
{
  guess_info g;
  guess_stats(rp_in, n_rows, ci, elm_size, examine_ratio, g);
  if (gi) *gi = g;
  return convert(rp_in, n_rows, ci, n_cols, n_nz, elm_size, g.S, g.hblock, g.var_height, g.diags,
                 min_height, B);
}

int guess_stats(const int *rp_in, int n_rows, const int *ci, int elm_size, int examine_ratio,
                guess_info &out)
{
  enum { ELTS_PER_SLICE = 8 * 256, MIN_HBLOCK_RATIO = 25, MIN_DIAGS_RATIO = 50, TRIAL = 32 };
  std::vector<int> rp;
  const int n = pad_row_ptr(rp_in, n_rows, rp);
  csr_in A{n, rp.data(), ci, elm_size};
  int rows = (int)(((long long)n_rows * examine_ratio + 99) / 100);
  rows = (rows + TRIAL - 1) & ~(TRIAL - 1);
  if (rows + 1 > n) rows = n - 1 - ((n - 1) % TRIAL);   /* stay inside the padded view */
  const int elts = std::max(1, rp[rows + 1]);
  long long waste = 0;
  int n_pairs = 0, n_diags = 0;
  for (int i = 0; i < rows; i += TRIAL) {
    int maxlen = 0;
    for (int r = i; r < i + TRIAL; r++) maxlen = std::max(maxlen, rp[r + 1] - rp[r]);
    for (int r = i; r < i + TRIAL; r++) {
      waste += maxlen - (rp[r + 1] - rp[r]);
      for (int k = rp[r]; k + 1 < rp[r + 1]; k++)
        if (!(ci[k] % 2) && ci[k + 1] == ci[k] + 1) { n_pairs++; k++; }
    }
    std::vector<int> dummy;
    find_diags(A, TRIAL, i, dummy);
    n_diags += (int)dummy.size();
  }
  const int waste_ratio = (int)(100 * waste / elts);
  const bool want_var = waste_ratio > 100;
  const float avg = 1.f * elts / rows;
  int S;
  if (want_var) {
    S = 256;
  } else {
    S = (int)(ELTS_PER_SLICE / avg);
    S = floor_pow2(S);
    if (S < kMinSliceHeight) S = kMinSliceHeight;
    else if (S > kMaxSliceHeight) S = kMaxSliceHeight;
  }
  int hb = 1;
  const bool want_diags = (long long)n_diags * TRIAL > (long long)elts * MIN_DIAGS_RATIO / 100;
  if (!want_diags && n_pairs > (long long)elts * MIN_HBLOCK_RATIO / 100) hb = 2;
  out.rows_examined = rows;
  out.elts_examined = elts;
  out.waste = waste;
  out.n_pairs = n_pairs;
  out.n_diags = n_diags;
  out.S = S;
  out.hblock = hb;
  out.var_height = want_var;
  out.diags = want_diags;
  return 0;
}

}  // namespace conv
}  // namespace ispm
