/* api_smoke.cpp -- a caller written against the reference's C++ API (plan/plan.h,
 * formats/formats.h, fastainv/fastainv.h, util/cuda/sblas.h, util/cuda/initialize.h),
 * compiled with g++ against the drop-in headers and linked to libispm0.so.  Prints
 * "ok" lines; any failure aborts with a message.  Run by tests/test_gpu_cxx_api.py. */
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "fastainv/droptol-tuning.h"
#include "fastainv/fastainv.h"
#include "plan/plan.h"
#include "util/cuda/initialize.h"
#include "util/cuda/pinner.h"
#include "util/cuda/sblas.h"

#define REQUIRE(c)                                                         \
  do {                                                                     \
    if (!(c)) {                                                            \
      std::fprintf(stderr, "FAILED %s at line %d\n", #c, __LINE__);        \
      std::exit(1);                                                        \
    }                                                                      \
  } while (0)

static csr_matrix<double> poisson(int m)
{
  csr_matrix<double> A;
  const int n = m * m * m;
  A.n_rows = A.n_cols = n;
  A.row_ptr.push_back(0);
  for (int z = 0; z < m; z++)
    for (int y = 0; y < m; y++)
      for (int x = 0; x < m; x++) {
        const int i = x + m * (y + m * z);
        const int nb[7] = {z > 0 ? i - m * m : -1, y > 0 ? i - m : -1, x > 0 ? i - 1 : -1, i,
                           x < m - 1 ? i + 1 : -1, y < m - 1 ? i + m : -1, z < m - 1 ? i + m * m : -1};
        for (int k = 0; k < 7; k++)
          if (nb[k] >= 0) {
            A.cols.push_back(nb[k]);
            A.elms.push_back(nb[k] == i ? 6.001 : -1.0);
          }
        A.row_ptr.push_back((int)A.cols.size());
      }
  A.n_nz = (int)A.cols.size();
  return A;
}

static void count(void *v) { ++*static_cast<long *>(v); }

int main()
{
  std::setvbuf(stdout, nullptr, _IOLBF, 0);
  ispm_initialize();
  std::printf("device %s\n", get_device_string().str);
  csr_matrix<double> A = poisson(12);
  const int n = A.n_rows;
  std::vector<double> x(n);
  for (int i = 0; i < n; i++) x[i] = std::sin(0.37 * i);
  std::vector<double> yref(n, 0.0);
  for (int r = 0; r < n; r++)
    for (int k = A.row_ptr[r]; k < A.row_ptr[r + 1]; k++) yref[r] += A.elms[k] * x[A.cols[k]];

  for (int kind = 0; kind <= 2; kind++) {
    spmv_plan<double> plan(A, (spmv_plan_kind)kind);
    if (kind == PLAN_ESTIMATE) {
      /* PLAN_ESTIMATE is PLAN_DEFAULT (plan/plan.h:74-75 of the reference) */
      spmv_plan<double> def(A, PLAN_DEFAULT);
      REQUIRE(plan.mparams.S == def.mparams.S && plan.mparams.H == def.mparams.H &&
              plan.mparams.V == def.mparams.V && plan.mparams.D == def.mparams.D &&
              plan.params.N == def.params.N && plan.params.S == def.params.S &&
              plan.params.H == def.params.H);
    }
    REQUIRE(A.n_rows % 512 == 0);     /* the plan pads the caller's matrix */
    vector<double, host_memory_space_tag> hx(A.n_rows, 0.0);
    for (int i = 0; i < n; i++) hx[i] = x[i];
    vector<double, device_memory_space_tag> dx(hx), dy(A.n_rows);
    plan.execute_spmv(dx, dy);
    vector<double, host_memory_space_tag> hy(dy);
    double err = 0;
    for (int i = 0; i < n; i++) err = std::fmax(err, std::fabs(hy[i] - yref[i]));
    REQUIRE(err < 1e-12);
    std::printf("ok spmv_plan kind=%d S=%u H=%u V=%u D=%u N=%u err=%.1e\n", kind, plan.mparams.S,
                plan.mparams.H, plan.mparams.V, plan.mparams.D, plan.params.N, err);
  }
  {
    /* coefficient refresh: same pattern, doubled values, re-converted with the saved
     * format parameters and copied to the device */
    spmv_plan<double> plan(A);
    csr_matrix<double> A2 = A;
    for (size_t k = 0; k < A2.elms.size(); k++) A2.elms[k] *= 2.0;
    plan.change_host_matrix(A2);
    plan.update_device_matrix_coefs();
    vector<double, host_memory_space_tag> hx(A.n_rows, 0.0);
    for (int i = 0; i < n; i++) hx[i] = x[i];
    vector<double, device_memory_space_tag> dx(hx), dy(A.n_rows);
    plan.execute_spmv(dx.udata(), dy.udata());
    vector<double, host_memory_space_tag> hy(dy);
    double err = 0;
    for (int i = 0; i < n; i++) err = std::fmax(err, std::fabs(hy[i] - 2.0 * yref[i]));
    REQUIRE(err < 1e-12);
    /* an invalid launch (N < S/H) leaves y unwritten */
    spmv_launch_params bad = plan.params;
    bad.N = 64; bad.S = 512; bad.H = 1;
    vector<double, device_memory_space_tag> dz(A.n_rows);
    launch_spmv_kernel<double, double>(bad, *plan.device_matrix, dx, dz);
    vector<double, host_memory_space_tag> hz(dz);
    for (int i = 0; i < n; i++) REQUIRE(hz[i] == 0.0);
    std::printf("ok coefficient refresh err=%.1e\n", err);
  }
  {
    /* short irregular rows: the exhaustive plan may take a slice height below the
     * conversion's default minimum of 4 (the library times heights from 1); the plan and its
     * coefficient refresh convert with that height */
    csr_matrix<double> R;
    const int nr = 3000;
    R.n_rows = R.n_cols = nr;
    R.row_ptr.push_back(0);
    unsigned long long s = 12345;
    auto next = [&s]() { s = s * 6364136223846793005ULL + 1442695040888963407ULL; return (unsigned)(s >> 33); };
    for (int r = 0; r < nr; r++) {
      /* 1..3 distinct columns, sorted as CSR rows must be */
      const int len = 1 + (int)(next() % 3);
      std::vector<int> c;
      while ((int)c.size() < len) {
        const int v = (int)(next() % nr);
        if (std::find(c.begin(), c.end(), v) == c.end()) c.push_back(v);
      }
      std::sort(c.begin(), c.end());
      for (int v : c) {
        R.cols.push_back(v);
        R.elms.push_back(1.0 + (next() % 7));
      }
      R.row_ptr.push_back((int)R.cols.size());
    }
    R.n_nz = (int)R.cols.size();
    std::vector<double> rx(nr), ry(nr, 0.0);
    for (int i = 0; i < nr; i++) rx[i] = std::cos(0.11 * i);
    for (int r = 0; r < nr; r++)
      for (int k = R.row_ptr[r]; k < R.row_ptr[r + 1]; k++) ry[r] += R.elms[k] * rx[R.cols[k]];
    spmv_plan<double> plan(R, PLAN_EXHAUSTIVE);
    for (int pass = 0; pass < 2; pass++) {
      if (pass) {
        csr_matrix<double> R2 = R;
        for (size_t k = 0; k < R2.elms.size(); k++) R2.elms[k] *= 3.0;
        plan.change_host_matrix(R2);
        plan.update_device_matrix_coefs();
      }
      vector<double, host_memory_space_tag> hx(R.n_rows, 0.0);
      for (int i = 0; i < nr; i++) hx[i] = rx[i];
      vector<double, device_memory_space_tag> dx(hx), dy(R.n_rows);
      plan.execute_spmv(dx, dy);
      vector<double, host_memory_space_tag> hy(dy);
      double err = 0;
      for (int i = 0; i < nr; i++) err = std::fmax(err, std::fabs(hy[i] - (pass ? 3.0 : 1.0) * ry[i]));
      REQUIRE(err < 1e-12);
    }
    std::printf("ok exhaustive plan S=%u with refresh\n", plan.mparams.S);
    /* the reference's timer (util/cuda/timer.h:33), as plan.h:121 uses it */
    vector<double, device_memory_space_tag> dx(R.n_rows), dy(R.n_rows);
    cudatimer t;
    t.start();
    plan.execute_spmv(dx, dy);
    t.stop();
    REQUIRE(t.elapsed_seconds() > 0 && t.elapsed_seconds() < 1.0);
    std::printf("ok cudatimer %.3g ms\n", 1e3 * t.elapsed_seconds());
  }
  {
    csr_matrix<double> B = poisson(6);
    csr_matrix<float> a1, a2;
    long calls = 0;
    const bool ok = fastainv_sym<double, float>(a1, a2, B, count, &calls, 8192000, 0.05);
    REQUIRE(ok);
    REQUIRE(calls == B.n_rows && a1.n_rows == B.n_rows && a1.n_nz == a2.n_nz);
    std::printf("ok fastainv nnz=%d\n", a1.n_nz);
  }
  {
    const int m = 1000;
    std::vector<double> a(m), b(m);
    for (int i = 0; i < m; i++) { a[i] = i * 0.5; b[i] = 1.0 / (1 + i); }
    vector<double, device_memory_space_tag> da(a.begin(), a.end()), db(b.begin(), b.end()), ds(1);
    sblas::dot(da.data(), db.data(), ds.data(), m);
    vector<double, host_memory_space_tag> hs(ds);
    double ref = 0;
    for (int i = 0; i < m; i++) ref += a[i] * b[i];
    REQUIRE(std::fabs(hs[0] - ref) < 1e-12 * std::fabs(ref));
    /* result written through a mapped pinned scalar */
    std::vector<double, cuda_pinned_allocator<double> > pinned(1, 0.0);
    pinmemptr<double> ps(pinned.data());
    sblas::dot(da.data(), db.data(), ps(), m);
    ispm_synchronize();
    REQUIRE(std::fabs(*ps - ref) < 1e-12 * std::fabs(ref));
    std::printf("ok sblas dot %.6f\n", hs[0]);
  }
  {
    /* the reference's device allocator (util/cuda/allocator.h:8): zero-filled memory */
    std::vector<double, cuda_device_allocator<double> > dv(257);
    std::vector<double> back(dv.size(), 1.0);
    copy(back.data(), reinterpret_cast<const devmem<double> *>(dv.data()), dv.size());
    for (double v : back) REQUIRE(v == 0.0);
    std::printf("ok cuda_device_allocator\n");
  }
  {
    DroptolTuner t;
    REQUIRE(std::fabs(t.getDroptol() - 0.006231) < 1e-6);
    std::printf("ok droptol tuner\n");
  }
  std::printf("ok all\n");
  return 0;
}
