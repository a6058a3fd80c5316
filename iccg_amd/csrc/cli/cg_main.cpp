// cg_main.cpp -- the reference's command-line drivers on the GPU solve path.
//
// One source, built five ways like the reference (Makefile:40-53):
//   -DNOPRE            cg_crs    matrix.mtx [tol] [max_iter]
//   -DJAC              dcg_crs   matrix.mtx [tol] [max_iter]
//   -DIC               iccg_crs  matrix.mtx [tol] [max_iter]
//   -DIC -DPIC         piccg_crs / picg_crs  matrix.mtx color.col [tol] [max_iter]
//   -DIC -DPIC -DABMC  abmc_crs  matrix.mtx block.blk bcolor.bcol [tol] [max_iter]
// Same argv, defaults (tol 1e-14, max_iter 10000), right-hand side b = 1,
// x0 = 0, orderings and output line as cg_crs.cpp:338-531.  Options that the
// reference does not have start with "--" and may appear anywhere:
//   --precision=mixed   FP32 factor/triangular solves, FP64 CG, stop on the
//                       FP64 true residual
//   --json              add one JSON line with the full device report
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "iccg/block.hpp"
#include "iccg/cg.hpp"
#include "iccg/color.hpp"
#include "iccg/crs_io.hpp"
#include "iccg/ic0.hpp"
#include "iccg/precond.hpp"

static int run(int argc, char **argv) {
  std::vector<std::string> pos;
  int precision = ICCG_FP64;
  bool json = false;
  for (int k = 1; k < argc; ++k) {
    std::string a = argv[k];
    if (a == "--precision=mixed") precision = ICCG_MIXED;
    else if (a == "--precision=fp64") precision = ICCG_FP64;
    else if (a == "--json") json = true;
    else pos.push_back(a);
  }
#if defined(ABMC)
  const size_t nfiles = 3;
  const char *usage = " matrix.mtx block.blk bcolor.bcol [tol=1e-14] [max_iter=10000]\n";
#elif defined(PIC)
  const size_t nfiles = 2;
  const char *usage = " matrix.mtx color.col [tol=1e-14] [max_iter=10000]\n";
#else
  const size_t nfiles = 1;
  const char *usage = " matrix.mtx [tol=1e-14] [max_iter=10000]\n";
#endif
  if (pos.size() < nfiles) {
    std::cerr << "Usage: " << argv[0] << usage;
    return 1;
  }
  const double tol = pos.size() > nfiles ? std::atof(pos[nfiles].c_str()) : 1e-14;
  const int max_iter = pos.size() > nfiles + 1 ? std::atoi(pos[nfiles + 1].c_str()) : 10000;

  CRS A = read_matrix_market_crs(pos[0]);
  if (A.n <= 0) {
    std::cerr << "Invalid matrix size.\n";
    return 1;
  }

#if defined(PIC)
  int nc = 0;
#if defined(ABMC)
  int nbs = 0, nbc = 0;
  std::vector<int> block_of_old = ReadBlockFile_1Based(pos[1], A.n, nbs);
  std::vector<int> block_color = ReadBlockColorFile_1Based(pos[2], nbs, nbc);
  std::vector<int> color((size_t)A.n);
  for (int i = 0; i < A.n; ++i) color[i] = block_color[block_of_old[i]];
  nc = nbc;
  std::vector<int> new_of_old = BuildPermutationByBlockColor(block_of_old, block_color);
#else
  std::vector<int> color = ReadColorFile_1Based(pos[1], A.n, nc);
  std::vector<int> new_of_old = BuildPermutationByColor(color);
#endif
  CRS Ap = expand_lower_to_full(Permute_PtAP_LowerCRS_to_LowerCRS(A, new_of_old));
  ColorSchedule sched;
#if defined(ABMC)
  std::vector<int> color_ptr_blk, blocks_of_color, block_ptr, rows_of_block;
  BuildABMCSchedule(new_of_old, block_of_old, block_color, nbs, nbc, color_ptr_blk,
                    blocks_of_color, block_ptr, rows_of_block);
  sched.mode = ColorSchedule::ABMC_BLOCK;
  sched.nc = nbc;
  sched.nb = nbs;
  sched.color_ptr_blk = color_ptr_blk.data();
  sched.blocks_of_color = blocks_of_color.data();
  sched.block_ptr = block_ptr.data();
  sched.rows_of_block = rows_of_block.data();
#else
  std::vector<int> color_ptr, rows_of_color;
  BuildColorPtrRows_FromPermutationAndColor(color, nc, new_of_old, color_ptr, rows_of_color);
  sched.nc = nc;
  sched.color_ptr = color_ptr.data();
  sched.rows_of_color = rows_of_color.data();
#endif
  (void)nc;
  const CRS &A_run = Ap;
  const ColorSchedule *sched_ptr = &sched;
#else
  const CRS &A_run = A;
  const ColorSchedule *sched_ptr = nullptr;
#endif

  std::vector<double> x((size_t)A.n, 0.0);
  std::vector<double> b((size_t)A.n, 1.0);

#if defined(NOPRE)
  Identity M(A_run);
#elif defined(JAC)
  Jacobi M(A_run);
#elif defined(IC)
  IC0 M(A_run, 0.0, sched_ptr, precision);
#endif
  (void)precision;

  iccg_cg_result detail{};
  auto start = std::chrono::steady_clock::now();
  CGResult out = conjugate_gradient(A_run, b, x, M, max_iter, tol, sched_ptr, 0, &detail);
  auto end = std::chrono::steady_clock::now();
  auto elapsed = std::chrono::duration_cast<std::chrono::milliseconds>(end - start).count();

  std::cout << "cg_time [ms]=" << elapsed << ", converged=" << out.converged
            << ", iters=" << out.iters << ", rel_resid=" << out.rel_resid;
  // ||A x - b|| / ||b|| (cg_crs.cpp:517-528): device SpMV, host FP64 sums
  std::vector<double> Ax;
  spmv(A_run, x, Ax);
  double nr = 0.0, nb = 0.0;
  for (size_t i = 0; i < Ax.size(); ++i) {
    const double d = Ax[i] - b[i];
    nr = std::fma(d, d, nr);
    nb = std::fma(b[i], b[i], nb);
  }
  std::cout << ", ||Ax-b||/||b|| = " << std::sqrt(nr) / std::sqrt(nb) << "\n";
  if (json) {
    std::printf("{\"iters\": %d, \"rel_resid\": %.17g, \"converged\": %d, \"breakdown\": %d, "
                "\"true_rel_resid\": %.17g, \"restarts\": %d, \"solve_ms\": %.6f}\n",
                detail.iters, detail.rel_resid, detail.converged, detail.breakdown,
                detail.true_rel_resid, detail.restarts, detail.solve_ms);
  }
  return 0;
}

int main(int argc, char **argv) {
  try {
    return run(argc, argv);
  } catch (const std::exception &e) {
    // the reference terminates on an uncaught std::runtime_error
    std::cerr << "terminate called after throwing an instance of 'std::runtime_error'\n"
              << "  what():  " << e.what() << "\n";
    return 134;
  }
}
