/*
 * iccg_oracle.h -- CPU restatement of the stomo0511/ICCG solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is what the HIP product path is
 * compared against: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  Nothing in the product links it.
 *
 * Every function restates one reference routine in plain C, single-threaded,
 * with the same per-row operation order and the same fused multiply-adds that
 * g++ -O3 contracts in the reference build (crs_io.cpp, color.cpp, block.cpp,
 * ic0.cpp, precond.hpp, cg_crs.cpp).  Where the reference uses OpenMP, the
 * oracle runs the single-thread schedule (colors, then blocks, then rows, all
 * ascending), which is the result every valid coloring produces.
 */
#ifndef ICCG_ORACLE_H
#define ICCG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- L0: CRS construction (crs_io.cpp) ---------------------------------- */

/* COO -> CSR: sort by (i,j), sum duplicates (crs_io.cpp:19-70).
 * Returns nnz written (<= nin). rowptr has nrows+1 entries. */
long orc_coo_to_crs(int nrows, long nin, const int *I, const int *J,
                    const double *V, int *rowptr, int *colind, double *val);

/* Lower CSR (j<=i) -> full symmetric CSR, rows column-sorted
 * (crs_io.cpp:163-227). Output capacity 2*nnz_lower. Returns nnz_full. */
long orc_expand_lower_to_full(int n, const int *rp, const int *ci,
                              const double *v, int *frp, int *fci, double *fv);

/* ---- L1: orderings (color.cpp, block.cpp) ------------------------------ */

/* new_of_old = rank under stable sort by (color, old) (color.cpp:50-63). */
void orc_perm_by_color(int n, const int *color, int *new_of_old);

/* new_of_old = rank under stable sort by (block_color[block], block, old)
 * (block.cpp:92-114). */
void orc_perm_by_block_color(int n, const int *block_of_old,
                             const int *block_color, int *new_of_old);

/* P^T A P on the stored lower triangle, folded to c<=r, rows sorted,
 * duplicates summed, exact zeros dropped (color.cpp:66-141).
 * Output capacity = nnz of input lower part. Returns nnz, or -1 on a size
 * mismatch. */
long orc_ptap_lower(int n, const int *rp, const int *ci, const double *v,
                    const int *new_of_old, int *brp, int *bci, double *bv);

/* color_ptr[nc+1], rows_of_color[n] in new index space (color.cpp:164-197).
 * Returns 0, or -1 on an out-of-range color / bad permutation. */
int orc_color_ptr_rows(int n, const int *color, int nc, const int *new_of_old,
                       int *color_ptr, int *rows_of_color);

/* ABMC schedule (block.cpp:116-174). Returns 0. */
int orc_abmc_schedule(int n, const int *new_of_old, const int *block_of_old,
                      const int *block_color, int nb, int nc,
                      int *color_ptr_blk, int *blocks_of_color,
                      int *block_ptr, int *rows_of_block);

/* ---- L2: preconditioners (precond.hpp, ic0.cpp) ------------------------- */

/* inv_diag[i] = 1/a_ii, or 1.0 when a_ii is 0 or absent (precond.hpp:23-35) */
void orc_jacobi_inv_diag(int n, const int *rp, const int *ci, const double *v,
                         double *inv_diag);

typedef struct {
  int n;
  const int *Lrp, *Lci;      /* L = lower+diag CSR (nnz_L) */
  const double *Lv;
  const int *diag_pos;       /* position of L_ii in row i */
  const int *LTcp, *LTri;    /* L^T as CSC of L (column i of L) */
  const double *LTv;
  const int *LT_diag_pos;    /* position of L_ii in column i */
} orc_ic0f;

/* IC(0) factorisation (ic0.cpp:6-151).  Inputs: CSR of A (full or lower).
 * Outputs sized by nnz_L = number of stored entries with j<=i.
 * Returns nnz_L >= 0, or
 *   -1  diagonal missing at row *err_row          (ic0.cpp:52-53)
 *   -2  near-zero pivot L_jj at row *err_row       (ic0.cpp:86-87)
 *   -3  non-positive pivot at row *err_row         (ic0.cpp:104-113)  */
long orc_ic0_factor(int n, const int *rp, const int *ci, const double *v,
                    double shift, int *Lrp, int *Lci, double *Lv,
                    int *diag_pos, int *LTcp, int *LTri, double *LTv,
                    int *LT_diag_pos, int *err_row);

typedef struct {
  int nc;
  const int *color_ptr, *rows_of_color;
  int mode; /* 0 = MC_NODE, 1 = ABMC_BLOCK (color.hpp:15) */
  int nb;
  const int *color_ptr_blk, *blocks_of_color, *block_ptr, *rows_of_block;
} orc_sched;

/* z = (L L^T)^{-1} r with the dispatch of IC0::apply (ic0.cpp:366-393):
 * sched NULL or nc<=0 -> sequential (153-190); ABMC with all four arrays ->
 * ABMC (263-339); color_ptr && rows_of_color -> MC (194-249). */
void orc_ic0_apply(const orc_ic0f *f, const orc_sched *sched, const double *r,
                   double *z, double *y_scratch);

/* The mixed-precision preconditioner apply (no counterpart in the
 * reference, which solves in FP64 only): the factor's values and diagonal
 * rounded to float, r rounded to float, the sequential forward and backward
 * sweeps of ic0.cpp:153-184 in float with fmaf and float division, z
 * widened to double.  Any valid MC / ABMC order gives the same bits. */
void orc_ic0_apply_f32(const orc_ic0f *f, const double *r, double *z);

/* ---- L3: Krylov (cg_crs.cpp) -------------------------------------------- */

/* dot with the reference's 80-bit accumulation (cg_crs.cpp:33-40). */
double orc_dot_ld(long n, const double *a, const double *b);
/* dot with a double-double accumulation, rounded once to double. */
double orc_dot_dd(long n, const double *a, const double *b);

/* y = A x, ascending-k fused multiply-add per row (cg_crs.cpp:54-79). */
void orc_spmv(int n, const int *rp, const int *ci, const double *v,
              const double *x, double *y);

/* ||A x - b|| / ||b|| as printed by main (cg_crs.cpp:517-528). */
double orc_true_resid(int n, const int *rp, const int *ci, const double *v,
                      const double *x, const double *b);
/* The same with double-double sums, each rounded once; ||b|| = 0 -> 1. */
double orc_true_resid_dd(int n, const int *rp, const int *ci, const double *v,
                         const double *x, const double *b);

/* conjugate_gradient<Precond> (cg_crs.cpp:89-162).
 * kind: 0 Identity, 1 Jacobi (inv_diag), 2 IC0 (ic, sched), 3 IC0 applied in
 * FP32 (orc_ic0_apply_f32).
 * dot_mode: 0 = 80-bit accumulation (the reference), 1 = double-double.
 * hist (may be NULL, capacity max_iter+1) receives rel_resid after each
 * update, hist[0] = initial.  Returns iters. */
int orc_cg(int n, const int *rp, const int *ci, const double *v,
           const double *b, double *x, int kind, const double *inv_diag,
           const orc_ic0f *ic, const orc_sched *sched, int max_iter,
           double tol, int dot_mode, double *rel_out, int *converged_out,
           double *hist);

/* Mixed precision CG: kind 3 of orc_cg (the FP32 apply orc_ic0_apply_f32,
 * double-double dots) restarted from the current x while the FP64 true
 * residual is >= tol after the recurrence converged (at most 100 restarts).
 * Returns the total iterations. */
int orc_cg_mixed(int n, const int *rp, const int *ci, const double *v,
                 const double *b, double *x, const orc_ic0f *ic, int max_iter,
                 double tol, double *rel_out, int *converged_out,
                 int *restarts_out, double *true_rel_out);

/* ---- schedule validators (cg_crs.cpp:171-336) --------------------------- */

/* Same-color edges between different blocks, i<j side only
 * (cg_crs.cpp:171-224). conflicts_by_color[nc]. Returns total, -1 on range
 * error. */
long long orc_check_interblock(int n, const int *rp, const int *ci,
                               const int *block_of_old, const int *block_color,
                               int nb, int nc, long long *conflicts_by_color);

/* Same-color j<i dependencies of lower(Ap_full) in the new space, split into
 * same-block / cross-block when block_of_old is given (cg_crs.cpp:259-336).
 * Arrays sized nc_max = max color + 1 (returned through *nc_out).
 * Returns total hazards, -1 on error. */
long long orc_check_hazards(int n, const int *rp, const int *ci,
                            const int *new_of_old, const int *color_old,
                            const int *block_of_old, int *nc_out,
                            long long *hazard_by_color,
                            long long *same_block_by_color,
                            long long *cross_block_by_color, int nc_cap);

#ifdef __cplusplus
}
#endif
#endif
