// cg.hpp -- conjugate_gradient<Precond> and CGResult (cg_crs.cpp:82-162),
// exported in a header (the reference keeps them in its driver TU).  The
// whole CG phase runs on the GPU: b and x are copied in and out once.
#pragma once

#include <algorithm>
#include <cmath>
#include <vector>

#include "ic0.hpp"
#include "precond.hpp"

// cg_crs.cpp:82-86
struct CGResult {
  int iters = 0;
  double rel_resid = NAN;
  bool converged = false;
};

// cg_crs.cpp:89-98.  `flags` (extension): ICCG_SOLVE_TRUE_RESID; `detail`
// (extension) receives the full device-side report.
template <class Precond>
CGResult conjugate_gradient(const CRS &A, const std::vector<double> &b, std::vector<double> &x,
                            const Precond &M, int max_iter, double tol,
                            const ColorSchedule *sched, int flags = 0,
                            iccg_cg_result *detail = nullptr) {
  M.use_schedule(sched);
  iccg_csr_t dA = M.csr();
  iccg_detail::CsrHandle tmp;
  if (M.source() != &A) {
    iccg_csr_t a = nullptr;
    iccg_detail::check(iccg_csr_create(A.n, (long)A.colind.size(), A.rowptr.data(),
                                       A.colind.data(), A.val.data(), &a));
    tmp.reset(a);
    dA = a;
  }
  if (M.precision() == ICCG_MIXED) flags |= ICCG_SOLVE_TRUE_RESID;
  iccg_cg_result r{};
  iccg_detail::check(iccg_cg_solve(dA, M.handle(), b.data(), x.data(), max_iter, tol, flags, &r));
  if (detail) *detail = r;
  CGResult out;
  out.iters = r.iters;
  out.rel_resid = r.rel_resid;
  out.converged = r.converged != 0;
  return out;
}

// conjugate_gradient for several right-hand sides of one system at once
// (extension, no reference counterpart): B[k] and X[k] are right-hand side
// and solution k; each runs its own recurrence and gets the result
// conjugate_gradient gives it alone, bit for bit, while the device reads A
// and the factor once for up to four of them (iccg_cg_solve_multi).
template <class Precond>
std::vector<CGResult> conjugate_gradient_multi(const CRS &A,
                                               const std::vector<std::vector<double>> &B,
                                               std::vector<std::vector<double>> &X,
                                               const Precond &M, int max_iter, double tol,
                                               const ColorSchedule *sched, int flags = 0) {
  M.use_schedule(sched);
  iccg_csr_t dA = M.csr();
  iccg_detail::CsrHandle tmp;
  if (M.source() != &A) {
    iccg_csr_t a = nullptr;
    iccg_detail::check(iccg_csr_create(A.n, (long)A.colind.size(), A.rowptr.data(),
                                       A.colind.data(), A.val.data(), &a));
    tmp.reset(a);
    dA = a;
  }
  if (M.precision() == ICCG_MIXED) flags |= ICCG_SOLVE_TRUE_RESID;
  const size_t n = (size_t)A.n, k = B.size();
  std::vector<double> b(n * k), x(n * k);
  X.resize(k);
  for (size_t j = 0; j < k; ++j) {
    std::copy(B[j].begin(), B[j].begin() + n, b.begin() + j * n);
    X[j].resize(n, 0.0);
    std::copy(X[j].begin(), X[j].end(), x.begin() + j * n);
  }
  std::vector<iccg_cg_result> r(k);
  iccg_detail::check(iccg_cg_solve_multi(dA, M.handle(), (int)k, b.data(), x.data(), max_iter,
                                         tol, flags, r.data()));
  std::vector<CGResult> out(k);
  for (size_t j = 0; j < k; ++j) {
    std::copy(x.begin() + j * n, x.begin() + (j + 1) * n, X[j].begin());
    out[j].iters = r[j].iters;
    out[j].rel_resid = r[j].rel_resid;
    out[j].converged = r[j].converged != 0;
  }
  return out;
}

// y = A x on the device (cg_crs.cpp:54-79), host vectors
inline void spmv(const CRS &A, const std::vector<double> &x, std::vector<double> &y) {
  iccg_csr_t a = nullptr;
  iccg_detail::check(iccg_csr_create(A.n, (long)A.colind.size(), A.rowptr.data(),
                                     A.colind.data(), A.val.data(), &a));
  iccg_detail::CsrHandle h(a);
  y.resize((size_t)A.n);
  iccg_detail::check(iccg_spmv(a, x.data(), y.data()));
}
