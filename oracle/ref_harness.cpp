// ref_harness.cpp -- test harness that compiles the reference ICCG sources in
// place (from the directory given by -I at build time) into oracle/_ref/,
// exposing their routines through a C interface so the tests can compare the
// oracle restatement against the reference itself.
//
// TEST INFRASTRUCTURE ONLY: never linked by the product.  This file holds no
// reference code; it includes the reference's own translation unit
// cg_crs.cpp (with its main renamed) and calls the functions declared there
// and in crs_io.hpp / color.hpp / block.hpp / ic0.hpp.
#include <omp.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>
#include <stdexcept>

#include "crs_io.hpp"
#include "color.hpp"
#include "precond.hpp"
// IC0 keeps its factor private; open it so the harness can read L.
#define private public
#include "ic0.hpp"
#undef private

#define main reference_main
#include "cg_crs.cpp"
#undef main

namespace {
std::string g_err;
CRS g_crs;
IC0 *g_ic0 = nullptr;

CRS make_crs(int n, const int *rp, const int *ci, const double *v) {
  CRS A;
  A.n = n;
  A.rowptr.assign(rp, rp + n + 1);
  long nnz = rp[n];
  A.colind.assign(ci, ci + nnz);
  if (v)
    A.val.assign(v, v + nnz);
  else
    A.val.assign(nnz, 1.0);
  return A;
}

ColorSchedule make_sched(int nc, const int *color_ptr, const int *rows_of_color,
                         int mode, int nb, const int *cpb, const int *boc,
                         const int *bp, const int *rob) {
  ColorSchedule s;
  s.nc = nc;
  s.color_ptr = color_ptr;
  s.rows_of_color = rows_of_color;
  s.mode = mode == 1 ? ColorSchedule::ABMC_BLOCK : ColorSchedule::MC_NODE;
  s.nb = nb;
  s.color_ptr_blk = cpb;
  s.blocks_of_color = boc;
  s.block_ptr = bp;
  s.rows_of_block = rob;
  return s;
}
}  // namespace

extern "C" {

const char *ref_last_error() { return g_err.c_str(); }

// OpenMP threads of the reference's parallel loops.  One thread gives the
// sequential summation order of its long double reductions (the oracle's).
void ref_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

// ---- CRS holder ------------------------------------------------------------
void ref_crs_dims(int *n, long *nnz) {
  *n = g_crs.n;
  *nnz = (long)g_crs.colind.size();
}
void ref_crs_copy(int *rp, int *ci, double *v) {
  std::memcpy(rp, g_crs.rowptr.data(), sizeof(int) * g_crs.rowptr.size());
  std::memcpy(ci, g_crs.colind.data(), sizeof(int) * g_crs.colind.size());
  std::memcpy(v, g_crs.val.data(), sizeof(double) * g_crs.val.size());
}

int ref_read_mm(const char *path) {
  try {
    g_crs = read_matrix_market_crs(path);
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

int ref_expand_lower_to_full(int n, const int *rp, const int *ci,
                             const double *v) {
  g_crs = expand_lower_to_full(make_crs(n, rp, ci, v));
  return 0;
}

int ref_ptap(int n, const int *rp, const int *ci, const double *v,
             const int *new_of_old) {
  try {
    std::vector<int> p(new_of_old, new_of_old + n);
    g_crs = Permute_PtAP_LowerCRS_to_LowerCRS(make_crs(n, rp, ci, v), p);
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

// ---- orderings -------------------------------------------------------------
void ref_perm_by_color(int n, const int *color, int *new_of_old) {
  std::vector<int> c(color, color + n);
  std::vector<int> p = BuildPermutationByColor(c);
  std::memcpy(new_of_old, p.data(), sizeof(int) * n);
}

void ref_perm_by_block_color(int n, const int *block_of_old, int nb,
                             const int *block_color, int *new_of_old) {
  std::vector<int> bo(block_of_old, block_of_old + n);
  std::vector<int> bc(block_color, block_color + nb);
  std::vector<int> p = BuildPermutationByBlockColor(bo, bc);
  std::memcpy(new_of_old, p.data(), sizeof(int) * n);
}

int ref_color_ptr_rows(int n, const int *color, int nc, const int *new_of_old,
                       int *color_ptr, int *rows_of_color) {
  try {
    std::vector<int> c(color, color + n), p(new_of_old, new_of_old + n);
    std::vector<int> cp, rc;
    BuildColorPtrRows_FromPermutationAndColor(c, nc, p, cp, rc);
    std::memcpy(color_ptr, cp.data(), sizeof(int) * cp.size());
    std::memcpy(rows_of_color, rc.data(), sizeof(int) * rc.size());
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

void ref_abmc_schedule(int n, const int *new_of_old, const int *block_of_old,
                       const int *block_color, int nb, int nc, int *cpb,
                       int *boc, int *bp, int *rob) {
  std::vector<int> p(new_of_old, new_of_old + n), bo(block_of_old, block_of_old + n);
  std::vector<int> bc(block_color, block_color + nb);
  std::vector<int> a, b, c, d;
  BuildABMCSchedule(p, bo, bc, nb, nc, a, b, c, d);
  std::memcpy(cpb, a.data(), sizeof(int) * a.size());
  std::memcpy(boc, b.data(), sizeof(int) * b.size());
  std::memcpy(bp, c.data(), sizeof(int) * c.size());
  std::memcpy(rob, d.data(), sizeof(int) * d.size());
}

// File readers: return count through *cnt_out, values into out (size n).
int ref_read_col(const char *path, int n, int *out, int *nc_out) {
  try {
    std::vector<int> c = ReadColorFile_1Based(path, n, *nc_out);
    std::memcpy(out, c.data(), sizeof(int) * n);
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}
int ref_read_blk(const char *path, int n, int *out, int *nb_out) {
  try {
    std::vector<int> c = ReadBlockFile_1Based(path, n, *nb_out);
    std::memcpy(out, c.data(), sizeof(int) * n);
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}
int ref_read_bcol(const char *path, int nb, int *out, int *nc_out) {
  try {
    std::vector<int> c = ReadBlockColorFile_1Based(path, nb, *nc_out);
    std::memcpy(out, c.data(), sizeof(int) * nb);
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

// ---- preconditioners -------------------------------------------------------
void ref_jacobi(int n, const int *rp, const int *ci, const double *v,
                double *inv_diag) {
  Jacobi J(make_crs(n, rp, ci, v));
  std::memcpy(inv_diag, J.inv_diag.data(), sizeof(double) * n);
}

int ref_ic0_create(int n, const int *rp, const int *ci, const double *v,
                   double shift) {
  delete g_ic0;
  g_ic0 = nullptr;
  try {
    g_ic0 = new IC0(make_crs(n, rp, ci, v), shift);
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}
long ref_ic0_nnz() { return g_ic0 ? (long)g_ic0->L_colind.size() : -1; }
void ref_ic0_copy(int *Lrp, int *Lci, double *Lv, int *diag_pos, int *LTcp,
                  int *LTri, double *LTv, int *LT_diag_pos) {
  const IC0 &M = *g_ic0;
  std::memcpy(Lrp, M.L_rowptr.data(), sizeof(int) * M.L_rowptr.size());
  std::memcpy(Lci, M.L_colind.data(), sizeof(int) * M.L_colind.size());
  std::memcpy(Lv, M.L_val.data(), sizeof(double) * M.L_val.size());
  std::memcpy(diag_pos, M.diag_pos.data(), sizeof(int) * M.diag_pos.size());
  std::memcpy(LTcp, M.LT_colptr.data(), sizeof(int) * M.LT_colptr.size());
  std::memcpy(LTri, M.LT_rowind.data(), sizeof(int) * M.LT_rowind.size());
  std::memcpy(LTv, M.LT_val.data(), sizeof(double) * M.LT_val.size());
  std::memcpy(LT_diag_pos, M.LT_diag_pos.data(), sizeof(int) * M.LT_diag_pos.size());
}

void ref_ic0_apply(int n, const double *r, double *z, int use_sched, int nc,
                   const int *color_ptr, const int *rows_of_color, int mode,
                   int nb, const int *cpb, const int *boc, const int *bp,
                   const int *rob) {
  std::vector<double> rv(r, r + n), zv(n, 0.0);
  ColorSchedule s = make_sched(nc, color_ptr, rows_of_color, mode, nb, cpb, boc, bp, rob);
  g_ic0->apply(rv, zv, use_sched ? &s : nullptr);
  std::memcpy(z, zv.data(), sizeof(double) * n);
}

// ---- solver ----------------------------------------------------------------
// kind: 0 Identity, 1 Jacobi, 2 IC0 (the factor built by ref_ic0_create)
int ref_cg(int kind, int n, const int *rp, const int *ci, const double *v,
           const double *b, double *x, int max_iter, double tol, int use_sched,
           int nc, const int *color_ptr, const int *rows_of_color, int mode,
           int nb, const int *cpb, const int *boc, const int *bp,
           const int *rob, double *rel_out, int *conv_out) {
  CRS A = make_crs(n, rp, ci, v);
  std::vector<double> bv(b, b + n), xv(x, x + n);
  ColorSchedule s = make_sched(nc, color_ptr, rows_of_color, mode, nb, cpb, boc, bp, rob);
  const ColorSchedule *sp = use_sched ? &s : nullptr;
  CGResult res;
  if (kind == 0) {
    Identity M(A);
    res = conjugate_gradient(A, bv, xv, M, max_iter, tol, sp);
  } else if (kind == 1) {
    Jacobi M(A);
    res = conjugate_gradient(A, bv, xv, M, max_iter, tol, sp);
  } else {
    res = conjugate_gradient(A, bv, xv, *g_ic0, max_iter, tol, sp);
  }
  std::memcpy(x, xv.data(), sizeof(double) * n);
  *rel_out = res.rel_resid;
  *conv_out = res.converged ? 1 : 0;
  return res.iters;
}

// The reference's CG phase alone, timed like main() (cg_crs.cpp:507-510) but
// in seconds at steady_clock resolution; uses the factor of ref_ic0_create.
int ref_cg_timed_ic0(int n, const int *rp, const int *ci, const double *v,
                     const double *b, double *x, int max_iter, double tol,
                     int use_sched, int nc, const int *color_ptr,
                     const int *rows_of_color, int mode, int nb, const int *cpb,
                     const int *boc, const int *bp, const int *rob,
                     double *rel_out, int *conv_out, double *seconds) {
  CRS A = make_crs(n, rp, ci, v);
  std::vector<double> bv(b, b + n), xv(x, x + n);
  ColorSchedule s = make_sched(nc, color_ptr, rows_of_color, mode, nb, cpb, boc, bp, rob);
  auto t0 = std::chrono::steady_clock::now();
  CGResult res = conjugate_gradient(A, bv, xv, *g_ic0, max_iter, tol, use_sched ? &s : nullptr);
  auto t1 = std::chrono::steady_clock::now();
  *seconds = std::chrono::duration<double>(t1 - t0).count();
  std::memcpy(x, xv.data(), sizeof(double) * n);
  *rel_out = res.rel_resid;
  *conv_out = res.converged ? 1 : 0;
  return res.iters;
}

void ref_spmv(int n, const int *rp, const int *ci, const double *v,
              const double *x, double *y) {
  CRS A = make_crs(n, rp, ci, v);
  std::vector<double> xv(x, x + n), yv;
  spmv(A, xv, yv);
  std::memcpy(y, yv.data(), sizeof(double) * n);
}

double ref_dot(long n, const double *a, const double *b) {
  std::vector<double> av(a, a + n), bv(b, b + n);
  return dot(av, bv);
}

// ---- validators (they print their report to stderr) ------------------------
void ref_check_interblock(int n, const int *rp, const int *ci,
                          const int *block_of_old, int nb,
                          const int *block_color, int nc) {
  std::vector<int> bo(block_of_old, block_of_old + n), bc(block_color, block_color + nb);
  DebugCheck_InterBlockSameColorEdges(make_crs(n, rp, ci, nullptr), bo, bc, nb, nc);
}

void ref_check_hazards(int n, const int *rp, const int *ci, const double *v,
                       const int *new_of_old, const int *color_old,
                       const int *block_of_old, int nb,
                       const int *block_color) {
  std::vector<int> p(new_of_old, new_of_old + n), c(color_old, color_old + n);
  std::vector<int> bo, bc;
  if (block_of_old) {
    bo.assign(block_of_old, block_of_old + n);
    bc.assign(block_color, block_color + nb);
  }
  DebugCheck_ScheduleHazards_OnLower(make_crs(n, rp, ci, v), p, c,
                                     block_of_old ? &bo : nullptr,
                                     block_of_old ? &bc : nullptr);
}

int ref_run_main(int argc, char **argv) { return reference_main(argc, argv); }
}
