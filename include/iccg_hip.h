/*
 * iccg_hip.h -- C ABI of the MI355X ICCG solve path (libiccg_hip.so).
 *
 * Plain C: int / long / double / float pointers and sizes, opaque handles,
 * int status codes.  No exception crosses this boundary; on failure a call
 * returns a non-zero ICCG_* status and iccg_last_error() (thread-local) holds
 * the message, worded like the reference's std::runtime_error text.
 *
 * Each entry point names the reference interface it replaces
 * (stomo0511/ICCG, file:line).  The reference's C++ surface (CRS,
 * ColorSchedule, Identity/Jacobi/IC0, conjugate_gradient) is mirrored on top
 * of this header by the headers under include/iccg/.
 */
#ifndef ICCG_HIP_H
#define ICCG_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------- */
enum {
  ICCG_OK = 0,
  ICCG_ERR_ARG = 1,        /* bad argument / size mismatch                   */
  ICCG_ERR_IO = 2,         /* file open / parse failure (crs_io.cpp, color.cpp) */
  ICCG_ERR_PIVOT = 3,      /* IC(0) missing diagonal / zero / non-positive pivot */
  ICCG_ERR_SCHEDULE = 4,   /* coloring leaves a same-color dependency (a race) */
  ICCG_ERR_DEVICE = 5,     /* HIP runtime error                                */
  ICCG_ERR_NOMEM = 6,      /* host or device allocation failure                */
  ICCG_ERR_STATE = 7       /* object used in a state that does not allow it    */
};

/* preconditioner kinds: precond.hpp:8-45 (Identity, Jacobi), ic0.hpp:11 (IC0) */
enum { ICCG_PRECOND_IDENTITY = 0, ICCG_PRECOND_JACOBI = 1, ICCG_PRECOND_IC0 = 2 };

/* precision of the preconditioner (the outer CG recurrence is always FP64):
 *   ICCG_FP64   L values, triangular solves and their vectors in FP64 -- the
 *               reference's arithmetic;
 *   ICCG_MIXED  IC(0) computed in FP64 (pivots FP64), then stored and applied
 *               in FP32 (triangular solves, y/z vectors in FP32).            */
enum { ICCG_FP64 = 0, ICCG_MIXED = 1 };

/* schedule modes: color.hpp:15 ColorSchedule::Mode */
enum { ICCG_SCHED_MC = 0, ICCG_SCHED_ABMC = 1 };

/* solve flags */
enum {
  ICCG_SOLVE_TRUE_RESID = 1, /* stop only when the FP64 true residual
                                ||b - A x|| / ||b|| < tol (restart CG from the
                                true residual otherwise).  Default in MIXED. */
  ICCG_SOLVE_ZERO_X0 = 2      /* start from x0 = 0 (cg_crs.cpp:439), set on the
                                device instead of uploading zeros            */
};

typedef struct iccg_csr_s *iccg_csr_t;
typedef struct iccg_sched_s *iccg_sched_t;
typedef struct iccg_precond_s *iccg_precond_t;
typedef struct iccg_solver_s *iccg_solver_t;
typedef struct iccg_msolver_s *iccg_msolver_t;

/* CGResult (cg_crs.cpp:82-86) plus what the GPU path reports beside it */
typedef struct {
  int iters;              /* iterations taken (cg_crs.cpp:131,144)          */
  double rel_resid;       /* recurrence ||r||/||b|| at return               */
  int converged;          /* cg_crs.cpp:120,147                             */
  int breakdown;          /* pAp <= eps ||p|| ||Ap|| (cg_crs.cpp:127-133)   */
  double true_rel_resid;  /* ||b - A x|| / ||b|| recomputed in FP64 (<0: not computed) */
  int restarts;           /* true-residual restarts (ICCG_SOLVE_TRUE_RESID) */
  double solve_ms;        /* CG phase wall time, device-resident data       */
} iccg_cg_result;

/* ---- errors --------------------------------------------------------------- */
const char *iccg_last_error(void);
const char *iccg_status_string(int status);

/* ---- device --------------------------------------------------------------- */
int iccg_device_count(int *count);
/* Selects the HIP device used by objects created afterwards on this thread. */
int iccg_device_set(int device);
/* Device used by this thread's calls (iccg_device_set, default 0). */
int iccg_device_get(int *device);
/* No counterpart in the reference (its vectors are std::vector).  The
 * library keeps device memory freed by its objects in a pool of its own and
 * reuses it for later objects without a device-wide wait, and keeps up to
 * 4 GiB of pinned host staging the same way.  This hands the idle part of
 * both back (device memory of this thread's device); objects still alive
 * keep theirs.  For callers that share the GPU with other allocators. */
int iccg_device_trim(void);

/* ---- L0: CRS I/O (crs_io.hpp:10-56) ---------------------------------------- */

/* read_matrix_market_crs (crs_io.hpp:52, crs_io.cpp:73-159).  Arrays are
 * allocated by the library; release each with iccg_free_host. */
int iccg_mm_read(const char *path, int *n, long *nnz, int **rowptr,
                 int **colind, double **val);
void iccg_free_host(void *p);

/* coo_to_crs (crs_io.cpp:19-70): sort by (i,j), sum duplicates.
 * rowptr[nrows+1]; colind/val capacity nin; *nnz_out entries written. */
int iccg_coo_to_crs(int nrows, long nin, const int *I, const int *J,
                    const double *V, int *rowptr, int *colind, double *val,
                    long *nnz_out);

/* expand_lower_to_full (crs_io.hpp:56, crs_io.cpp:163-227).  Output capacity
 * 2*nnz_lower. */
int iccg_expand_lower_to_full(int n, const int *rp, const int *ci,
                              const double *v, int *frp, int *fci, double *fv,
                              long *nnz_out);

/* ---- L1: orderings (color.hpp:4-7, block.hpp:10-39) ------------------------ */

/* ReadColorFile_1Based (color.cpp:14-47): color[n] 0-based, *nc. */
int iccg_read_color_file(const char *path, int n, int *color, int *nc);
/* ReadBlockFile_1Based (block.cpp:15-51): block_of_old[n], *nb. */
int iccg_read_block_file(const char *path, int n, int *block_of_old, int *nb);
/* ReadBlockColorFile_1Based (block.cpp:53-89): block_color[nb], *nc. */
int iccg_read_block_color_file(const char *path, int nb, int *block_color,
                               int *nc);
/* Writers of the same 1-based formats (header count line, then "id value"). */
int iccg_write_color_file(const char *path, int n, const int *color, int nc);
int iccg_write_block_file(const char *path, int n, const int *block_of_old,
                          int nb);
int iccg_write_block_color_file(const char *path, int nb,
                                const int *block_color, int nc);

/* BuildPermutationByColor (color.cpp:50-63) */
int iccg_perm_by_color(int n, const int *color, int *new_of_old);
/* BuildPermutationByBlockColor (block.cpp:92-114) */
int iccg_perm_by_block_color(int n, const int *block_of_old, int nb,
                             const int *block_color, int *new_of_old);
/* Permute_PtAP_LowerCRS_to_LowerCRS (color.cpp:66-141).  Output capacity =
 * number of stored entries with j<=i. */
int iccg_ptap_lower(int n, const int *rp, const int *ci, const double *v,
                    const int *new_of_old, int *brp, int *bci, double *bv,
                    long *nnz_out);
/* BuildColorPtrRows_FromPermutationAndColor (color.cpp:164-197) */
int iccg_color_ptr_rows(int n, const int *color, int nc, const int *new_of_old,
                        int *color_ptr, int *rows_of_color);
/* BuildABMCSchedule (block.cpp:116-174) */
int iccg_abmc_schedule(int n, const int *new_of_old, const int *block_of_old,
                       const int *block_color, int nb, int nc,
                       int *color_ptr_blk, int *blocks_of_color,
                       int *block_ptr, int *rows_of_block);

/* DebugCheck_InterBlockSameColorEdges (cg_crs.cpp:171-224): same-color
 * inter-block edges, per color (conflicts[nc]) and in total (*total). */
int iccg_check_interblock(int n, const int *rp, const int *ci,
                          const int *block_of_old, int nb,
                          const int *block_color, int nc,
                          long long *conflicts, long long *total);
/* DebugCheck_ScheduleHazards_OnLower (cg_crs.cpp:259-336).  Arrays sized
 * nc_cap; *nc_out = max color + 1.  block_of_old may be NULL. */
int iccg_check_hazards(int n, const int *rp, const int *ci,
                       const int *new_of_old, const int *color_old,
                       const int *block_of_old, int nc_cap, int *nc_out,
                       long long *hazards, long long *same_block,
                       long long *cross_block, long long *total);

/* ---- ordering generators (no reference counterpart: the reference reads
 *      its .col/.blk/.bcol from an external tool, run_mc.sh:14,
 *      run_abmc.sh:19-20) ------------------------------------------------- */

/* First-fit greedy coloring in index order of the graph of A (i~j when
 * a_ij or a_ji is stored, i != j).  color[n] 0-based, *nc. */
int iccg_greedy_color(int n, const int *rp, const int *ci, int *color, int *nc);
/* ABMC blocking: breadth-first aggregation of up to block_size nodes from the
 * smallest unassigned node (block ids in creation order), then first-fit
 * greedy coloring of the block graph in block-id order.  block_of_old[n],
 * block_color[n] (first *nb used), *nb, *nc. */
int iccg_abmc_blocking(int n, const int *rp, const int *ci, int block_size,
                       int *block_of_old, int *nb, int *block_color, int *nc);
/* ABMC blocking of a structured nx*ny*nz grid (natural index x fastest)
 * into bx*by*bz boxes, boxes colored first-fit in box order. */
int iccg_abmc_blocking_grid(int nx, int ny, int nz, int bx, int by, int bz,
                            int *block_of_old, int *nb, int *block_color,
                            int *nc);

/* ---- synthetic problems (no reference counterpart) ------------------------ */

/* Symmetric 2D 5-point (nz==1) or 3D 7-point finite-volume matrix on an
 * nx*ny*nz Dirichlet grid, natural order.  coef_lo_exp == coef_hi_exp == 0
 * gives the constant-coefficient Laplacian (diag 2d, off-diagonal -1);
 * otherwise every face gets c = 10^(lo + (hi-lo) u), u ~ U[0,1) drawn from
 * splitmix64(seed) in face order (x faces, y faces, z faces; each
 * lexicographic with x fastest), off-diagonal -c, diagonal = sum of the
 * cell's face coefficients + diag_shift.  Pass rowptr == NULL to get the
 * sizes only. */
int iccg_gen_grid(int nx, int ny, int nz, double coef_lo_exp,
                  double coef_hi_exp, double diag_shift, unsigned long long seed,
                  int *n_out, long *nnz_out, int *rowptr, int *colind,
                  double *val);
/* out[i] ~ N(0,1): splitmix64(seed) uniforms, Box-Muller pairs in FP64. */
int iccg_gen_normal(long n, unsigned long long seed, double *out);

/* ---- device objects ------------------------------------------------------- */

/* struct CRS (crs_io.hpp:10-15) uploaded to HBM (int32 rowptr/colind, FP64
 * values).  Arrays are copied; the caller keeps ownership. */
int iccg_csr_create(int n, long nnz, const int *rowptr, const int *colind,
                    const double *val, iccg_csr_t *out);
int iccg_csr_destroy(iccg_csr_t A);
/* spmv (cg_crs.cpp:54-79) on host buffers (copies in/out; test path). */
int iccg_spmv(iccg_csr_t A, const double *x, double *y);

/* struct ColorSchedule (color.hpp:9-23), MC_NODE form. */
int iccg_schedule_create_mc(int n, int nc, const int *color_ptr,
                            const int *rows_of_color, iccg_sched_t *out);
/* struct ColorSchedule, ABMC_BLOCK form (block.hpp:30-39 arrays). */
int iccg_schedule_create_abmc(int n, int nc, int nb, const int *color_ptr_blk,
                              const int *blocks_of_color, const int *block_ptr,
                              const int *rows_of_block, iccg_sched_t *out);
int iccg_schedule_destroy(iccg_sched_t s);

/* Identity(const CRS&) / Jacobi(const CRS&) (precond.hpp:10,23) and
 * IC0(const CRS&, double shift) (ic0.hpp:13) built on the device.  For IC0,
 * sched selects the triangular-solve schedule exactly as IC0::apply does
 * (ic0.cpp:366-393): NULL or nc<=0 -> natural order; ABMC_BLOCK -> ABMC;
 * MC_NODE -> multicolor.  The schedule is checked against A: a same-color
 * dependency the reference would race on is rejected (ICCG_ERR_SCHEDULE).
 * Pivot failures return ICCG_ERR_PIVOT with the reference's message. */
int iccg_precond_create(iccg_csr_t A, int kind, double shift,
                        iccg_sched_t sched, int precision,
                        iccg_precond_t *out);
int iccg_precond_destroy(iccg_precond_t M);
/* Re-plans the triangular solves of an IC0 preconditioner for another
 * schedule (IC0::apply takes the schedule per call, ic0.hpp:34).  The factor
 * is kept: under every valid schedule it is the same, bit for bit.  Not
 * while M is in use by a batch or a solver on another thread. */
int iccg_precond_set_schedule(iccg_precond_t M, iccg_sched_t sched);
/* Which device plan runs the triangular solves (no counterpart in the
 * reference; a diagnostic): plan is one of ICCG_PLAN_*, section_width the
 * in-block section width S of the sectioned ABMC layout, else 0. */
enum {
  ICCG_PLAN_NONE = 0,            /* Identity / Jacobi */
  ICCG_PLAN_LEVELS = 1,          /* natural order, level sets */
  ICCG_PLAN_CSR = 2,             /* MC or ABMC over the CSR factor */
  ICCG_PLAN_MC_SLICED = 3,       /* multicolor, sliced ELL */
  ICCG_PLAN_ABMC_SLICED = 4,     /* ABMC, sliced ELL, masked level loop */
  ICCG_PLAN_ABMC_SECTIONED = 5,  /* ABMC, sectioned layout, mask-free levels */
  ICCG_PLAN_ABMC_SECTIONED_FUSED = 6,  /* the same, the last color forward
                                          and backward in one launch */
  ICCG_PLAN_ABMC_SECTIONED_FUSED_R = 7, /* and, in the solver's iterations,
                                           the sweeps apply the CG's r update
                                           and form r.z */
  ICCG_PLAN_MC_SLICED_R = 8,           /* multicolor, sliced ELL, and, in
                                          the solver's iterations, the
                                          forward sweep applies the r update
                                          and the backward sweep forms r.z */
  ICCG_PLAN_ABMC_BLOCKS = 9            /* ABMC blocks of 65 to 2048 rows:
                                          one workgroup per block, rows in
                                          registers, a barrier per in-block
                                          level */
};
int iccg_precond_get_plan(iccg_precond_t M, int *plan, int *section_width);
/* apply(r, z, sched) (precond.hpp:12,38; ic0.hpp:34) on host buffers. */
int iccg_precond_apply(iccg_precond_t M, const double *r, double *z);
/* Copies the FP64 factor out: L (lower+diag CSR) and L^T (CSC of L), in the
 * layout of IC0's members (ic0.hpp:38-49).  Sizes: n+1 / nnz_L. */
int iccg_precond_factor_nnz(iccg_precond_t M, long *nnz_L);
int iccg_precond_get_factor(iccg_precond_t M, int *Lrp, int *Lci, double *Lv,
                            int *LTcp, int *LTri, double *LTv);

/* conjugate_gradient<Precond> (cg_crs.cpp:89-162) on host buffers:
 * b[n] in, x[n] in/out (x0 in, solution out). */
int iccg_cg_solve(iccg_csr_t A, iccg_precond_t M, const double *b, double *x,
                  int max_iter, double tol, int flags, iccg_cg_result *res);

/* iccg_cg_solve without any allocation.  It has no counterpart in the
 * reference; it is for callers that cannot allocate device memory (code
 * running while other work on the device is under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Every device buffer of
 * the solve (x, r, p, Ap, b, y, z, partials, scalars) is carved from one
 * 64 MiB array per device, which each device the library runs on holds
 * until the process ends, and the iterations are launched kernel by kernel
 * on the matrix's stream (no captured graphs).  Arguments, results and
 * errors are those of iccg_cg_solve; a system that needs more than 64 MiB
 * of scratch (about 56 bytes per row in FP64 with IC(0), 48 mixed, plus
 * 100 KB: some 1.1M rows) is rejected with ICCG_ERR_ARG.  The array keeps
 * the last call's contents; calls that overlap on one device race on it. */
int iccg_cg_solve_noalloc(iccg_csr_t A, iccg_precond_t M, const double *b, double *x,
                          int max_iter, double tol, int flags, iccg_cg_result *res);

/* ---- device-resident solver (timing, repeated solves) --------------------- */
int iccg_solver_create(iccg_csr_t A, iccg_precond_t M, iccg_solver_t *out);
int iccg_solver_destroy(iccg_solver_t S);
/* Uploads b and x0 (NULL x0 -> zeros) into the solver's HBM buffers. */
int iccg_solver_set(iccg_solver_t S, const double *b, const double *x0);
/* Runs the CG phase on the resident buffers (nothing crosses PCIe inside). */
int iccg_solver_run(iccg_solver_t S, int max_iter, double tol, int flags,
                    iccg_cg_result *res);
int iccg_solver_get_x(iccg_solver_t S, double *x);
/* Per-kernel-class event timing inside iccg_solver_run (0 = off).  While
 * timing is on, iterations are launched one kernel at a time. */
int iccg_solver_set_timing(iccg_solver_t S, int on);
/* Replay captured hipGraphs of whole iterations (default on). */
int iccg_solver_set_graphs(iccg_solver_t S, int on);
/* Waits for the device inside iccg_solver_run sleep on a blocking-sync
 * event instead of spinning (default off), for callers that run solves on
 * several host threads beside other host work. */
int iccg_solver_set_blocking(iccg_solver_t S, int on);
/* Kernel classes: 0 SpMV(+dots), 1 forward sweep, 2 backward sweep,
 * 3 vector updates.  Returns launches and summed milliseconds since the last
 * reset. */
int iccg_solver_get_timing(iccg_solver_t S, int cls, long *launches,
                           double *total_ms);
int iccg_solver_reset_timing(iccg_solver_t S);

/* ---- several right-hand sides of one system (no reference counterpart) -----
 * Requests that carry several right-hand sides: nrhs of one matrix and
 * preconditioner solved together.  Each right-hand side keeps its own CG
 * recurrence (cg_crs.cpp:89-162: its own alpha, beta, stop test, iteration
 * count and true-residual restarts) and its iterates are those of
 * iccg_solver_run on it alone, bit for bit; the kernels read each matrix
 * and factor entry once for up to four of them (vectors interleaved in
 * HBM).  The shared kernels run on the sectioned ABMC plan (box and
 * breadth-first blockings of at most 64 rows); on other plans, and for a
 * right-hand side left over from the sets of four (or two), the
 * single-system solver runs them one after another.  b, x0 and x hold nrhs
 * vectors of n entries (b[k*n + i]); res receives nrhs results. */
int iccg_msolver_create(iccg_csr_t A, iccg_precond_t M, int nrhs, iccg_msolver_t *out);
int iccg_msolver_destroy(iccg_msolver_t S);
/* sets: shared-kernel sets; lone: right-hand sides on the single-system solver */
int iccg_msolver_native(iccg_msolver_t S, int *sets, int *lone);
/* Uploads b and x0 (NULL x0 -> zeros). */
int iccg_msolver_set(iccg_msolver_t S, const double *b, const double *x0);
/* The CG phase of every right-hand side (flags as iccg_solver_run). */
int iccg_msolver_run(iccg_msolver_t S, int max_iter, double tol, int flags,
                     iccg_cg_result *res);
int iccg_msolver_get_x(iccg_msolver_t S, double *x);
/* Per-kernel-class event timing of the shared-kernel sets (kernel-by-kernel
 * launches while on; classes as iccg_solver_get_timing, a launch counted
 * once per set iteration in which some member ran). */
int iccg_msolver_set_timing(iccg_msolver_t S, int on);
int iccg_msolver_get_timing(iccg_msolver_t S, int cls, long *launches,
                            double *total_ms);
int iccg_msolver_reset_timing(iccg_msolver_t S);
/* create + set + run + get_x + destroy on host buffers (x: x0 in, solution out) */
int iccg_cg_solve_multi(iccg_csr_t A, iccg_precond_t M, int nrhs, const double *b,
                        double *x, int max_iter, double tol, int flags,
                        iccg_cg_result *res);

/* ---- many systems at once (no reference counterpart) -----------------------
 * Independent systems, each with its own matrix, IC0 preconditioner and
 * vectors resident in HBM, advanced together: every CG phase is one launch
 * over all running systems, each with its own alpha/beta/stop test.  Slots
 * that share a matrix and preconditioner (several right-hand sides) can run
 * as groups (iccg_batch_set_max_group).  The
 * preconditioner needs a sliced plan: a contiguous multicolor schedule (e.g.
 * from iccg_greedy_color) or an ABMC schedule of contiguous blocks of at most
 * 64 rows; rows of A, L and L^T hold at most 16 entries.  Systems may differ
 * in size, ordering kind and precision, and can be added while others run
 * (continuous batching); a system's iterates equal those of iccg_cg_solve on
 * it alone.  The matrix and preconditioner must outlive the slot.  One batch
 * is driven by one thread at a time. */
typedef struct iccg_batch_s *iccg_batch_t;
int iccg_batch_create(int capacity, iccg_batch_t *out);
int iccg_batch_destroy(iccg_batch_t B);
/* Admits a system into a free slot (copies b and x0 to HBM; NULL x0 = 0).
 * flags: ICCG_SOLVE_TRUE_RESID. */
int iccg_batch_add(iccg_batch_t B, iccg_csr_t A, iccg_precond_t M, const double *b,
                   const double *x0, int max_iter, double tol, int flags, int *slot);
/* Starts admitted systems, then runs up to max_steps iterations of every
 * running system; *n_running = systems not yet finished. */
int iccg_batch_run(iccg_batch_t B, int max_steps, int *n_running);
int iccg_batch_result(iccg_batch_t B, int slot, int *finished, iccg_cg_result *res);
int iccg_batch_get_x(iccg_batch_t B, int slot, double *x);
/* Frees a slot (finished or not). */
int iccg_batch_remove(iccg_batch_t B, int slot);
int iccg_batch_set_graphs(iccg_batch_t B, int on);
/* iccg_solver_set_blocking for the batch's waits in iccg_batch_run. */
int iccg_batch_set_blocking(iccg_batch_t B, int on);
/* Work decomposition of the batched kernels: 64-row slices per workgroup of
 * the SpMV / vector kernels (chunk_slices) and 64-row slices (multicolor)
 * or blocks (ABMC) per wavefront of the triangular sweeps (sweep_units).
 * With autotune on (the default), whenever the running systems' total size
 * changes by more than 2x (from 65,536 rows up), the batch runs two real
 * iterations under each candidate (4, 8, 16, 32 slices x 1, 2, 4 per
 * wavefront; 25 iterations in all, counted in max_steps) and keeps the
 * fastest, remembered per power-of-two size class so that a recurring load
 * is tuned once; set_autotune(1) forgets the choices, set_tuning fixes
 * them and turns autotune off.  The arithmetic of
 * every row is the same for every decomposition. */
int iccg_batch_set_autotune(iccg_batch_t B, int on);
int iccg_batch_get_tuning(iccg_batch_t B, int *chunk_slices, int *sweep_units);
int iccg_batch_set_tuning(iccg_batch_t B, int chunk_slices, int sweep_units);
/* Timing mode: kernels launched one by one (no graphs) with HIP events on
 * the batch's stream around every launch group.  Classes: 0 SpMV (+dots),
 * 1 forward sweeps, 2 backward sweeps, 3 vector updates and finish kernels;
 * launches = launch groups, total_ms summed since the last reset. */
int iccg_batch_set_timing(iccg_batch_t B, int on);
int iccg_batch_get_timing(iccg_batch_t B, int cls, long *launches, double *total_ms);
int iccg_batch_reset_timing(iccg_batch_t B);
/* Right-hand sides per group of slots sharing a matrix, 1..8 (default 1:
 * every right-hand side its own work items).  Grouped members read each
 * matrix and factor entry once and step through the ABMC in-block levels
 * together; on gfx950 this measured no faster than separate items (the
 * kernels are bound by per-right-hand-side gathers and occupancy, not by
 * matrix bytes), so it is off by default.  Results do not depend on it. */
int iccg_batch_set_max_group(iccg_batch_t B, int k);
/* Frees the device vectors that free slots keep for reuse (iccg_batch_remove
 * keeps them, so a slot refilled with a system no larger allocates nothing). */
int iccg_batch_trim(iccg_batch_t B);

/* ---- request loop (no reference counterpart) -------------------------------
 * Serves solve requests of different sizes on one GPU.  A request is a
 * symmetric matrix (full CSR, sorted rows, caller's numbering) with nrhs
 * right-hand sides; it is ordered ABMC (bx*by*bz boxes when the caller gives
 * its structured-grid shape, breadth-first blocks of block_size rows
 * otherwise, blocks colored first-fit), permuted, factorized IC(0) with
 * shift 0 and solved to ||r||/||b|| < tol; x comes back in the caller's
 * numbering.
 *   ICCG_SERVE_BATCHED    requests are admitted in arrival order while their
 *                         estimated HBM footprint fits hbm_budget, prepared
 *                         by `workers` host threads (the IC(0) factorization
 *                         runs on the GPU beside the running solves), and
 *                         every right-hand side runs as one slot of a batch
 *                         of max_slots (continuous batching: slots freed by
 *                         finished solves take the next ready request);
 *                         requests of at least multi_rows rows (0: none) go
 *                         instead to multi_runners threads (0 = 6), each
 *                         solving one request at a time with the
 *                         multi-right-hand-side solver (iccg_msolver_*; one
 *                         right-hand side runs on the single-system solver)
 *                         on a stream of its own, beside the batch.
 *   ICCG_SERVE_SEQUENTIAL the plain one-at-a-time path: one request is
 *                         prepared and its right-hand sides solved one after
 *                         another with the single-system solver before the
 *                         next request starts.
 * A right-hand side gives the same x, bit for bit, in either mode and
 * whatever else is being served. */
enum { ICCG_SERVE_BATCHED = 0, ICCG_SERVE_SEQUENTIAL = 1 };
typedef struct {
  int mode;               /* ICCG_SERVE_* */
  int precision;          /* ICCG_FP64 or ICCG_MIXED */
  int max_slots;          /* right-hand sides in flight (batched mode) */
  int workers;            /* host threads preparing requests (0 = 8) */
  int block_size;         /* rows per ABMC block of a general matrix, <= 64 */
  int flags;              /* ICCG_SOLVE_TRUE_RESID */
  long long hbm_budget;   /* bytes of HBM for admitted requests */
  int multi_rows;         /* batched mode: see above (0 = every request batched) */
  int multi_runners;      /* threads running those requests (0 = 6) */
} iccg_server_config;
/* hbm_in_use: the admitted requests' footprint estimates, a finished
 * request's until the reaper thread has freed its device memory. */
typedef struct {
  long long submitted, finished, failed, hbm_in_use;
  int waiting, preparing, ready, running;
  /* engine thread, batched mode (ms): inside iccg_batch_run, admitting
   * requests into slots, and harvesting finished right-hand sides (copy-out,
   * un-permutation, freeing) -- the last two with no GPU work queued */
  double engine_run_ms, engine_admit_ms, engine_harvest_ms;
} iccg_server_stats;
typedef struct iccg_server_s *iccg_server_t;
int iccg_server_create(const iccg_server_config *cfg, iccg_server_t *out);
/* Stops the loop; requests still in flight are dropped. */
int iccg_server_destroy(iccg_server_t S);
/* Queues a request (everything is copied); grid = {nx, ny, nz} of a
 * structured grid in natural order, or NULL.  b[k*n + i] is entry i of
 * right-hand side k. */
int iccg_server_submit(iccg_server_t S, int n, long nnz, const int *rowptr,
                       const int *colind, const double *val, const int *grid,
                       int nrhs, const double *b, int max_iter, double tol,
                       long long *id);
/* Waits up to timeout_ms (< 0: without limit) for request id to finish, or
 * for any finished request not yet reported when id < 0; *done_id is the
 * finished request or -1 on timeout. */
int iccg_server_wait(iccg_server_t S, long long id, int timeout_ms,
                     long long *done_id);
/* x[k*n + i] and res[k] per right-hand side; latency = submit to finish,
 * prepare = submit to ready for the GPU (ms).  A failed request returns its
 * error status and message. */
int iccg_server_result(iccg_server_t S, long long id, double *x,
                       iccg_cg_result *res, double *latency_ms,
                       double *prepare_ms);
/* Forgets a finished request. */
int iccg_server_release(iccg_server_t S, long long id);
int iccg_server_get_stats(iccg_server_t S, iccg_server_stats *st);

/* ---- measurement helpers -------------------------------------------------- */
/* Device copy of nbytes (HBM -> HBM) repeated reps times; average ms per
 * launch through *avg_ms.  Calibrates the hardware byte counters. */
int iccg_bench_copy(long nbytes, int reps, double *avg_ms);

#ifdef __cplusplus
}
#endif
#endif /* ICCG_HIP_H */
