/* foam_smoke.cpp -- the foam-extend call pattern written against the C-ABI (include/ispm.h)
 * by a C++ caller that includes no HIP header: an lduMatrix of a 12^3 diffusion mesh to CSR,
 * plans of A and of the approximate-inverse factors, the PCG solve with OpenFOAM's
 * normalised residual (checked against the formula evaluated on the host), then a second
 * time step whose coefficients reach the same A plan through the device scatter.  Prints
 * "ok" lines; any failure exits non-zero.  Run by tests/test_gpu_cxx_api.py. */
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "ispm.h"

#define CHECK(x)                                                                          \
  do {                                                                                    \
    const int st_ = (x);                                                                  \
    if (st_) {                                                                            \
      std::fprintf(stderr, "ispm error %d at line %d: %s\n", st_, __LINE__, ispm_last_error()); \
      std::exit(1);                                                                       \
    }                                                                                     \
  } while (0)
#define REQUIRE(c)                                                  \
  do {                                                              \
    if (!(c)) {                                                     \
      std::fprintf(stderr, "FAILED %s at line %d\n", #c, __LINE__); \
      std::exit(1);                                                 \
    }                                                               \
  } while (0)

struct ldu {
  int n = 0;
  std::vector<int> lo, up;
  std::vector<double> diag, upper;
};

/* faces in owner order (owner < neighbour), conductance c_f, upper = -c_f, diagonal = the
 * cell's face conductances + 1 per boundary face */
static ldu mesh(int m, double scale)
{
  ldu L;
  L.n = m * m * m;
  L.diag.assign(L.n, 0.0);
  std::vector<int> deg(L.n, 0);
  for (int i = 0; i < L.n; i++) {
    const int x = i % m, y = (i / m) % m, z = i / (m * m);
    const int nb[3] = {x + 1 < m ? i + 1 : -1, y + 1 < m ? i + m : -1, z + 1 < m ? i + m * m : -1};
    for (int k = 0; k < 3; k++)
      if (nb[k] >= 0) {
        const double c = scale * (1.0 + 0.5 * std::sin(0.37 * (double)L.lo.size()));
        L.lo.push_back(i);
        L.up.push_back(nb[k]);
        L.upper.push_back(-c);
        L.diag[i] += c;
        L.diag[nb[k]] += c;
        deg[i]++;
        deg[nb[k]]++;
      }
  }
  for (int i = 0; i < L.n; i++) L.diag[i] += 6 - deg[i];
  return L;
}

static std::vector<double> apply(const ldu &L, const std::vector<double> &x)
{
  std::vector<double> y(L.n);
  for (int i = 0; i < L.n; i++) y[i] = L.diag[i] * x[i];
  for (size_t f = 0; f < L.lo.size(); f++) {
    y[L.lo[f]] += L.upper[f] * x[L.up[f]];
    y[L.up[f]] += L.upper[f] * x[L.lo[f]];
  }
  return y;
}

/* OpenFOAM's normalised residual of x with the norm factor of x0 */
static double normalised(const ldu &L, const std::vector<double> &b, const std::vector<double> &x0,
                         const std::vector<double> &x)
{
  std::vector<double> one(L.n, 1.0);
  const std::vector<double> sumA = apply(L, one), ax0 = apply(L, x0), ax = apply(L, x);
  double xbar = 0;
  for (double v : x0) xbar += v;
  xbar /= L.n;
  double nf = 1e-20, r = 0;
  for (int i = 0; i < L.n; i++) {
    nf += std::fabs(ax0[i] - sumA[i] * xbar) + std::fabs(b[i] - sumA[i] * xbar);
    r += std::fabs(b[i] - ax[i]);
  }
  return r / nf;
}

static ispm_csr view(const ispm_csr_out &o)
{
  ispm_csr c;
  c.n_rows = o.n_rows; c.n_cols = o.n_cols; c.n_nz = o.n_nz;
  c.row_ptr = o.row_ptr; c.cols = o.cols; c.elms = o.elms; c.elm_type = o.elm_type;
  return c;
}

template<class V> static void *upload(const std::vector<V> &h)
{
  void *d = nullptr;
  CHECK(ispm_malloc(&d, sizeof(V) * h.size()));
  CHECK(ispm_memcpy(d, h.data(), sizeof(V) * h.size(), ISPM_H2D));
  return d;
}

/* one time step: FSAI factors of the current CSR values, a solver over the A plan, the solve */
static std::vector<double> step(ispm_plan *A, const ispm_csr &csr, int n_pad, const ldu &L,
                                const std::vector<double> &b, const std::vector<double> &x0, double tol)
{
  ispm_csr_out a1, a2;
  CHECK(ispm_fastainv_sym(ISPM_F64, &csr, nullptr, nullptr, 64LL * L.n, 0.125, &a1, &a2));
  const ispm_csr v1 = view(a1), v2 = view(a2);
  ispm_plan *M1 = nullptr, *M2 = nullptr;
  CHECK(ispm_plan_create(&M1, ISPM_F64, ISPM_F64, &v1, ISPM_PLAN_MIN_BYTES, 0));
  CHECK(ispm_plan_create(&M2, ISPM_F64, ISPM_F64, &v2, ISPM_PLAN_MIN_BYTES, 0));
  ispm_solver_opts o;
  o.method = ISPM_SOLVER_PPCG; o.precision = ISPM_PREC_FP64; o.tol = 1e-10; o.maxit = 2000;
  o.check_every = 16; o.use_graph = 1; o.reliable_delta = 0.01;
  ispm_solver *s = nullptr;
  CHECK(ispm_solver_create(&s, A, M1, M2, &o));
  std::vector<double> bp(n_pad, 0.0), xp(n_pad, 0.0);
  for (int i = 0; i < L.n; i++) { bp[i] = b[i]; xp[i] = x0[i]; }
  double *db = static_cast<double *>(upload(bp)), *dx = static_cast<double *>(upload(xp));
  ispm_norm_stats st;
  CHECK(ispm_solver_solve_normalised(s, db, dx, tol, 0.0, &st));
  CHECK(ispm_memcpy(xp.data(), dx, sizeof(double) * n_pad, ISPM_D2H));
  std::vector<double> x(xp.begin(), xp.begin() + L.n);
  const double r = normalised(L, b, x0, x), r0 = normalised(L, b, x0, x0);
  REQUIRE(st.converged && st.final_residual < tol && r < tol);
  REQUIRE(std::fabs(st.final_residual - r) <= 1e-6 * r + 1e-18);
  REQUIRE(std::fabs(st.initial_residual - r0) <= 1e-9 * r0 + 1e-14);
  std::printf("ok step: %d iterations, normalised residual %.2e -> %.2e\n", st.iterations, r0, r);
  CHECK(ispm_free(db));
  CHECK(ispm_free(dx));
  ispm_solver_destroy(s);
  ispm_plan_destroy(M1);
  ispm_plan_destroy(M2);
  ispm_csr_out_free(&a1);
  ispm_csr_out_free(&a2);
  return x;
}

int main()
{
  std::setvbuf(stdout, nullptr, _IOLBF, 0);
  CHECK(ispm_init());
  ldu L = mesh(12, 1.0);
  const int nf = (int)L.lo.size();
  ispm_csr_out co;
  std::vector<int> pos(L.n + 2 * nf);
  CHECK(ispm_ldu_to_csr(L.n, nf, L.lo.data(), L.up.data(), L.diag.data(), L.upper.data(), nullptr, &co, pos.data()));
  REQUIRE(co.n_nz == L.n + 2 * nf);
  ispm_csr csr = view(co);
  ispm_plan *A = nullptr;
  CHECK(ispm_plan_create(&A, ISPM_F64, ISPM_F64, &csr, ISPM_PLAN_DEFAULT, 0));
  ispm_plan_info info;
  CHECK(ispm_plan_info_get(A, &info));
  const int n_pad = info.device.n_rows;
  std::vector<double> xt(L.n);
  for (int i = 0; i < L.n; i++) xt[i] = std::cos(0.1 * i);
  std::vector<double> x = step(A, csr, n_pad, L, apply(L, xt), std::vector<double>(L.n, 0.0), 1e-8);

  /* next time step: new coefficients, same addressing, scattered into the plan on the device */
  ldu L2 = mesh(12, 1.5);
  void *dpos = upload(pos), *dd = upload(L2.diag), *du = upload(L2.upper), *dv = nullptr;
  CHECK(ispm_malloc(&dv, sizeof(double) * co.n_nz));
  CHECK(ispm_ldu_values_device(L2.n, nf, static_cast<const int *>(dpos), static_cast<const double *>(dd),
                               static_cast<const double *>(du), nullptr, static_cast<double *>(dv)));
  CHECK(ispm_plan_update_coefs_device(A, dv, ISPM_F64));
  CHECK(ispm_ldu_values(L2.n, nf, pos.data(), L2.diag.data(), L2.upper.data(), nullptr,
                        static_cast<double *>(co.elms)));   /* host CSR for the factors */
  std::vector<double> xt2(L.n);
  for (int i = 0; i < L.n; i++) xt2[i] = std::sin(0.05 * i) + 0.5;   /* a new field to reach */
  step(A, csr, n_pad, L2, apply(L2, xt2), x, 1e-8);
  for (void *p : {dpos, dd, du, dv}) CHECK(ispm_free(p));

  ispm_droptol_tuner *t = nullptr;
  CHECK(ispm_droptol_tuner_create(-10.0, -3.0, &t));
  REQUIRE(std::fabs(ispm_droptol_tuner_droptol(t) - 0.006231) < 1e-6);
  ispm_droptol_tuner_destroy(t);
  ispm_plan_destroy(A);
  ispm_csr_out_free(&co);
  std::printf("ok all\n");
  return 0;
}
