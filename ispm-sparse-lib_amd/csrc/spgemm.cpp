/* spgemm.cpp -- host sparse product C = A * B of two CSR matrices (FP64 accumulation).
 *
 * Used to apply an approximate-inverse preconditioner M = Ainv2 * Ainv1 as one sparse matrix
 * when that is cheaper to stream than its two factors: for factors with few off-diagonal
 * entries per row the product has fewer nonzeros than the two factors together.  No
 * counterpart in the reference, whose solver applies the two factors one after the other.
 *
 * Row-by-row accumulation into a dense scratch row with an occupancy list (Gustavson);
 * rows are split over threads, each counting, then filling its rows.  Columns of each output
 * row are sorted; products that cancel to exactly zero are kept as explicit entries (the
 * structure does not depend on the values). */
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>
#include "internal.h"
#include "ispm.h"

namespace {

double value(const ispm_csr *M, long long k)
{
  switch (M->elm_type) {
    case ISPM_F64: return static_cast<const double *>(M->elms)[k];
    case ISPM_F32: return static_cast<const float *>(M->elms)[k];
    default: {
      const unsigned short h = static_cast<const unsigned short *>(M->elms)[k];
      const unsigned u = (unsigned)h << 16;
      float f;
      std::memcpy(&f, &u, 4);
      return f;
    }
  }
}

struct row_acc {
  std::vector<double> val;
  std::vector<int> mark, list;
  explicit row_acc(int n) : val(n, 0.0), mark(n, -1) { list.reserve(64); }
  /* columns of row i of A*B (sorted); values in val[] */
  void row(const ispm_csr *A, const ispm_csr *B, int i)
  {
    list.clear();
    for (int ka = A->row_ptr[i]; ka < A->row_ptr[i + 1]; ka++) {
      const int j = A->cols[ka];
      const double a = value(A, ka);
      for (int kb = B->row_ptr[j]; kb < B->row_ptr[j + 1]; kb++) {
        const int c = B->cols[kb];
        if (mark[c] != i) {
          mark[c] = i;
          val[c] = 0.0;
          list.push_back(c);
        }
        val[c] += a * value(B, kb);
      }
    }
    std::sort(list.begin(), list.end());
  }
};

bool valid(const ispm_csr *M)
{
  return M && M->n_rows >= 0 && M->n_cols >= 0 && M->row_ptr && (M->n_nz == 0 || (M->cols && M->elms)) &&
         (M->elm_type == ISPM_F64 || M->elm_type == ISPM_F32 || M->elm_type == ISPM_BF16);
}

}  // namespace

extern "C" int ispm_csr_multiply(const ispm_csr *A, const ispm_csr *B, int out_type, ispm_csr_out *C)
{
  using ispm::fail;
  if (!C) return fail(ISPM_ERR_INVALID, "csr_multiply: null output");
  std::memset(C, 0, sizeof *C);
  if (!valid(A) || !valid(B)) return fail(ISPM_ERR_INVALID, "csr_multiply: bad operand");
  if (A->n_cols != B->n_rows) return fail(ISPM_ERR_INVALID, "csr_multiply: inner dimensions differ");
  if (out_type != ISPM_F64 && out_type != ISPM_F32)
    return fail(ISPM_ERR_UNSUPPORTED, "csr_multiply: output type %d", out_type);
  for (int i = 0; i < A->n_nz; i++)
    if (A->cols[i] < 0 || A->cols[i] >= A->n_cols) return fail(ISPM_ERR_INVALID, "csr_multiply: column out of range");
  for (int i = 0; i < B->n_nz; i++)
    if (B->cols[i] < 0 || B->cols[i] >= B->n_cols) return fail(ISPM_ERR_INVALID, "csr_multiply: column out of range");
  const int n = A->n_rows, m = B->n_cols;
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 4096) nt = 1;
  std::vector<long long> cnt((size_t)n + 1, 0);
  auto span = [&](unsigned t, int &r0, int &r1) {
    r0 = (int)((long long)n * t / nt);
    r1 = (int)((long long)n * (t + 1) / nt);
  };
  {
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++)
      th.emplace_back([&, t] {
        int r0, r1;
        span(t, r0, r1);
        row_acc acc(m);
        for (int i = r0; i < r1; i++) {
          acc.row(A, B, i);
          cnt[(size_t)i + 1] = (long long)acc.list.size();
        }
      });
    for (auto &x : th) x.join();
  }
  for (int i = 0; i < n; i++) cnt[(size_t)i + 1] += cnt[i];
  if (cnt[n] > 0x7fffffffLL) return fail(ISPM_ERR_INVALID, "csr_multiply: product has more than 2^31-1 entries");
  const size_t nnz = (size_t)cnt[n], es = out_type == ISPM_F64 ? 8 : 4;
  C->row_ptr = static_cast<int *>(std::malloc(sizeof(int) * ((size_t)n + 1)));
  C->cols = static_cast<int *>(std::malloc(sizeof(int) * std::max<size_t>(nnz, 1)));
  C->elms = std::malloc(es * std::max<size_t>(nnz, 1));
  if (!C->row_ptr || !C->cols || !C->elms) {
    ispm_csr_out_free(C);
    return fail(ISPM_ERR_RUNTIME, "csr_multiply: out of host memory");
  }
  for (int i = 0; i <= n; i++) C->row_ptr[i] = (int)cnt[i];
  {
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++)
      th.emplace_back([&, t] {
        int r0, r1;
        span(t, r0, r1);
        row_acc acc(m);
        for (int i = r0; i < r1; i++) {
          acc.row(A, B, i);
          size_t k = (size_t)cnt[i];
          for (int c : acc.list) {
            C->cols[k] = c;
            if (out_type == ISPM_F64) static_cast<double *>(C->elms)[k] = acc.val[c];
            else static_cast<float *>(C->elms)[k] = (float)acc.val[c];
            k++;
          }
        }
      });
    for (auto &x : th) x.join();
  }
  C->n_rows = n;
  C->n_cols = m;
  C->n_nz = (int)nnz;
  C->elm_type = out_type;
  return ISPM_OK;
}
