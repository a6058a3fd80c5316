// dev.hpp -- device data layout and kernel launchers of libiccg_hip.so.
//
// HBM layout (one system):
//   A   : CSR, int32 rowptr[n+1], int32 colind[nnz], f64 val[nnz]
//   L   : lower+diag CSR of the IC(0) pattern (rows sorted, diagonal LAST),
//         f64 values (+ f32 copy in mixed precision)
//   U   : L^T as CSR (= CSC of L), diagonal FIRST in each row
//   vectors x, r, p, Ap, b : f64[n];  y, z : f64[n] or f32[n] (mixed)
//   Scal: the solver's scalars (alpha, beta, rz, norms, flags) -- device
//         resident, read back by the host only every few iterations
//   part: per-workgroup double-double partial sums of the dot products
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace iccg {

constexpr int kWG = 256;          // threads per workgroup of streaming kernels
constexpr int kBlockPlanMax = 2048;  // largest ABMC block of the one-workgroup-per-block sweeps
constexpr int kMaxRedBlocks = 2048;  // workgroups of reducing kernels (8 per CU)
// MI355X's Infinity Cache (the memory-side cache in front of HBM): 256 MiB
constexpr double kInfinityCacheBytes = 256.0 * 1024 * 1024;

// Device-resident solver scalars.  Written only by the one-workgroup finish
// kernels that fold the reducing kernels' partials; read by the others at
// kernel start.
struct alignas(128) Scal {
  double normb;    // ||b|| (1.0 when ||b|| == 0, cg_crs.cpp:108-110)
  double rz_old;   // r.z of the previous iteration
  double alpha, beta;
  double rel;      // recurrence ||r|| / ||b||
  double pAp;
  double rr_true;  // ||b - A x||^2 (true-residual check)
  double tol;
  int iters;       // completed iterations
  int done;        // stop flag: converged / breakdown / max_iter
  int converged;
  int breakdown;
  int max_iter;
  int pad0;
  unsigned long long err;  // IC(0) failure key (row << 32 | row_j), ~0 = none
};

enum PrecondKind { PK_IDENT = 0, PK_JAC = 1, PK_IC = 2 };


// Sliced ELL with 64-row slices (one wavefront per slice): entry k of row
// i = 64 s + l sits at off[s] + 64 k + l, so the k-th load of a wavefront is
// one contiguous 512-byte (f64) run.  A row keeps its stored (ascending
// column) order; padding carries column -1 and is skipped, so a row sum is
// the same fma chain as in CSR, bit for bit.
constexpr int kTbl16Max = 256;  // column blocks a slice may reference (16-bit layout)

struct Sell {
  int n = 0;
  int nslice = 0;
  int maxw = 0;          // widest slice
  int *off = nullptr;    // nslice + 1 (entry offsets, multiples of 64)
  int *col = nullptr;    // off[nslice], -1 = padding
  double *v64 = nullptr;
  float *v32 = nullptr;
  // 16-bit columns, when every slice references at most kTbl16Max distinct
  // 64-row column blocks; an entry is (t << 6 | column & 63), padding
  // 0xFFFF, with t the column's block in the slice's table: tbl[16 s + t]
  // when every slice references at most 16 blocks (toff == nullptr), else
  // tbl[toff[s] + t], t < toff[s + 1] - toff[s].  Half the index bytes.
  uint16_t *col16 = nullptr;
  int *tbl = nullptr;
  int *toff = nullptr;
  // sectioned layout (slice = ABMC block b): inb[64 b + lane], byte k = the
  // row within the block of the lane's k-th in-block entry, 64 when absent
  unsigned *inb = nullptr;
};

struct DevCsr {
  int n = 0;
  long nnz = 0;
  int *rp = nullptr, *ci = nullptr;
  double *v = nullptr;
  Sell sell;             // used by the SpMV kernels when sell.off != nullptr
  float *sell_v32 = nullptr;  // f32 copy of sell values when all are exact floats
};

// Triangular-solve / factorization schedule on the device.
//   MC   : colors as contiguous row ranges (rows == nullptr) or row lists
//   ABMC : one workgroup per block; bstart[kb] = first position of block kb
//          (blocks in blocks_of_color order), rows listed by brows (nullptr
//          when positions are the new row ids); levels inside each block
//   LEVEL: natural order, level sets (lvl_ptr / lvl_rows)
struct DevSched {
  int kind = 0;  // 0 = natural levels, 1 = MC, 2 = ABMC
  int nc = 0;
  int *d_rows = nullptr;           // MC: rows_of_color (nullptr if identity)
  int *d_bstart = nullptr;         // ABMC: nb+1
  int *d_brows = nullptr;          // ABMC: n (nullptr if contiguous)
  uint16_t *d_lev_f = nullptr;     // ABMC/levels: per-row forward level
  uint16_t *d_lev_b = nullptr;     // per-row backward level
  int *d_nlev_f = nullptr, *d_nlev_b = nullptr;  // ABMC: per block
  int max_block = 0;
  int nb = 0;
  std::vector<int> h_bstart;       // ABMC: host copy of bstart
  // natural levels: level sets and launch segments of the forward and
  // backward plans.  A segment {0, l0, l1} runs levels [l0, l1) inside one
  // workgroup (narrow levels); {1, l, 0} runs level l over many workgroups.
  std::vector<int> h_lvl_ptr_f, h_lvl_ptr_b;
  int *d_lvl_ptr_f = nullptr, *d_lvl_rows_f = nullptr;
  int *d_lvl_ptr_b = nullptr, *d_lvl_rows_b = nullptr;
  std::vector<int> seg_f, seg_b;  // triples
  // host copies of the range bounds
  std::vector<int> h_color_ptr_v, h_color_blk_v;
};

struct Factor {
  int n = 0;
  long nnz = 0;
  int *lrp = nullptr, *lci = nullptr;
  double *lv = nullptr;
  float *lv32 = nullptr;
  int *urp = nullptr, *uci = nullptr;
  double *uv = nullptr;
  float *uv32 = nullptr;
  int *l2u = nullptr;
  // sliced-ELL copies of the strictly lower (L) and strictly upper (U = L^T)
  // parts and the diagonal, filled from lv after the factorization
  Sell ls, us;
  double *d64 = nullptr;
  float *d32 = nullptr;
  int *l2ls = nullptr;   // L CSR position -> ls entry, or -1 - row (diagonal)
  int *l2us = nullptr;   // L CSR position -> us entry, or -1 (diagonal)
  bool sell = false;
  // ABMC factors: ls / us in the sectioned layout with sec in-block slots
  // and up to seco >= sec outside entries per row (sell_layout_sections), or
  // sec = 0 for the plain one
  int sec = 0, seco = 0;
  // sectioned, and every block of the last color has no outside entries in
  // U (its backward rows have in-block terms only)
  bool fb = false;
  // the sweeps read the sliced entries with nontemporal loads (ldm): set
  // when they and the vectors an iteration reads again exceed the Infinity
  // Cache, so that the vectors stay in it
  bool nt = false;
};

// ---- multi-system engine (batch.hip) --------------------------------------
// One entry per system slot: the system's own HBM arrays.  Kernels take
// lists of work items that name a slot, so one launch per CG phase covers
// every active system whatever its size, and the set can change between
// launches (continuous batching).
struct SysDesc {
  int n = 0, nslice = 0;
  int mixed = 0;                     // factor, y, z in f32
  int max_iter = 0;
  double tol = 0;
  const int *a_off = nullptr, *a_col = nullptr;  // A, sliced ELL
  const double *a_v64 = nullptr;
  int a_paired = 0;                  // a_v64 in the paired layout (pair_slices)
  const float *a_v32 = nullptr;      // set when every value of A is an exact float
  const int *l_off = nullptr, *l_col = nullptr, *u_off = nullptr, *u_col = nullptr;
  const unsigned *l_inb = nullptr, *u_inb = nullptr;  // sectioned layout
  const void *l_v = nullptr, *u_v = nullptr, *d = nullptr;
  // ABMC blocks (nullptr for a multicolor system)
  const int *bstart = nullptr, *nlev_f = nullptr, *nlev_b = nullptr;
  const uint16_t *lev_f = nullptr, *lev_b = nullptr;
  int sec = 0;  // section width of the factor's sectioned layout, 0 = plain
  double *x = nullptr, *r = nullptr, *p = nullptr, *Ap = nullptr;
  const double *b = nullptr;
  void *y = nullptr, *z = nullptr;
};

// per-system solver scalars (the multi-system Scal)
struct alignas(64) SysScal {
  double normb, rz_old, alpha, beta, rel, pAp, rr_true, tol;
  int iters, done, converged, breakdown, max_iter, pad[3];
};

// Right-hand sides of one matrix and preconditioner run as a group: the
// kernels read each matrix entry once for all of them.
constexpr int kMaxRhs = 8;
struct GroupDesc {
  int k, pad;
  int slot[kMaxRhs];  // slot[0] also names the shared matrix and factor
};
// Items name their group and its first slot (slot0): one-member groups
// (the default) reach their system without the group table.
struct ChunkItem {  // rows [64*s0, min(64*s1, n)) of group grp: one workgroup
  int grp, s0, s1, slot0;
};
// One wavefront of a sweep over one color of group grp:
//   multicolor: 64-row slices [a0, a1), rows restricted to [lo, hi);
//   ABMC      : blocks [a0, a1).
struct ColorItem {
  int grp, a0, a1, lo, hi, slot0, pad[2];
};
// Finish work of one slot: its partials are chunks [c0, c1) of plane `plane`.
struct FinItem {
  int slot, c0, c1, plane;
};

// ---- launchers (kernels.hip) ----------------------------------------------
void launch_spmv_plain(const DevCsr &A, const double *x, double *y, hipStream_t st);
// Ap = A p with the partials of p.Ap, ||p||^2, ||Ap||^2; finish = also
// launch the alpha / breakdown kernel (launch_fin_alpha)
void launch_spmv_cg(const DevCsr &A, const double *p, double *Ap, Scal *S, double *part,
                    hipStream_t st, bool finish = true);
void launch_fin_alpha(const DevCsr &A, Scal *S, double *part, hipStream_t st);
// r = b - A x with ||b|| and ||r||.
//   mode 1: start of a solve -- resets S (iters 0), stores tol/max_iter,
//           sets normb, rel and done (rel < tol or max_iter <= 0)
//   mode 2: restart from the true residual -- keeps iters
//   mode 0: true-residual check only: S->rr_true, r not written
void launch_resid(const DevCsr &A, const double *x, const double *b, double *r, Scal *S,
                  double *part, int mode, double tol, int max_iter, hipStream_t st);
void launch_xr(int pk, int n, double *x, const double *p, double *r, const double *Ap,
               const double *inv_diag, double *z, Scal *S, double *part, hipStream_t st);
void launch_precond_diag_init(int pk, int n, const double *r, const double *inv_diag, double *z,
                              double *p, Scal *S, double *part, hipStream_t st);
// IC(0) iterations with the stop test deferred: launch_r_ic (r -= alpha Ap)
// leaves the |r|^2 partials in part's second region, launch_rz_stop folds
// them after the preconditioner, together with r.z, and launch_update_xp
// applies x += alpha p with the p update (x alone in the iteration that
// stops)
void launch_r_ic(int n, double *r, const double *Ap, Scal *S, double *part, hipStream_t st);
void launch_update_xp(int n, double *x, const void *z, bool z32, double *p, const Scal *S,
                      hipStream_t st);
// xr_count: the |r|^2 partials the forward sweep left (launch_fwd's
// ru_count), or 0 for launch_r_ic's
void launch_rz_stop(int n, const double *r, const void *z, bool z32, Scal *S, double *part,
                    hipStream_t st, int xr_count = 0);
void launch_rz(int n, const double *r, const void *z, bool z32, Scal *S, double *part, int init,
               hipStream_t st);
void launch_update_p(int n, const void *z, bool z32, double *p, const Scal *S, int init,
                     hipStream_t st);
void launch_jacobi_apply(int n, const double *inv_diag, const double *r, double *z,
                         hipStream_t st);
void launch_jacobi_setup(const DevCsr &A, double *inv_diag, hipStream_t st);

// sweeps: skip when S != nullptr && S->done.  With rz != 0 the backward
// sweep also forms r.z and, in its last workgroup, beta (rz == 2: rz_old at
// the start of a solve), like launch_rz.
// Sectioned ABMC factors whose last color's backward rows have in-block
// terms only can run that color forward and backward in one launch:
// launch_fwd with z_fused writes z for the last color (and not its y), and
// launch_bwd with last_done then skips that color.
bool sweeps_fused(const DevSched &sc, const Factor &F);
// Sectioned ABMC and multicolor sliced factors (and a partial buffer that
// holds the sweep's workgroups) can also take the CG's r update into the
// forward sweep:
// launch_fwd with Ap applies r -= alpha Ap to every row it starts, writes r
// and leaves *ru_count |r|^2 partials at the start of part.
bool fwd_updates_r(const DevSched &sc, const Factor &F);
// the sweeps that take the r update also form r.z (multicolor: the |r|^2
// partials where launch_r_ic leaves its own, the r.z ones where k_rz does)
bool sweeps_form_rz(const DevSched &sc, const Factor &F);
// ... and then the sweeps also leave the r.z partials (the fused last
// color's launch, and launch_bwd with rz_after_ru): the stop test's finish
// folds both from the layout they leave in part, with no r.z pass between
void launch_fin_sweep_stop(const DevSched &sc, Scal *S, double *part, hipStream_t st);
void launch_fwd(const DevSched &sc, const Factor &F, bool mixed, const double *r, void *y,
                const Scal *S, hipStream_t st, void *z_fused = nullptr,
                const double *Ap = nullptr, double *part = nullptr, int *ru_count = nullptr);
void launch_bwd(const DevSched &sc, const Factor &F, bool mixed, const void *y, void *z,
                const Scal *S, hipStream_t st, const double *r = nullptr, Scal *Sw = nullptr,
                double *part = nullptr, int rz = 0, bool last_done = false,
                bool rz_after_ru = false);
// true when launch_bwd fuses r.z for this schedule (sliced MC / ABMC plans)
// the triangular solves run on the sliced-ELL sweeps (contiguous MC colors,
// or contiguous ABMC blocks of at most 64 rows; rows at most 16 wide)
bool sliced_plan(const DevSched &sc, const Factor &F);
bool bwd_fuses_rz(const DevSched &sc, const Factor &F);
int bwd_partial_slots(const DevSched &sc, const Factor &F);
bool block_plan(const DevSched &sc, const Factor &F);
void launch_scatter_sell(const Factor &F, bool mixed, hipStream_t st);
void launch_ic0_factor(const DevSched &sc, const Factor &F, double shift,
                       unsigned long long *err, hipStream_t st);
void launch_scatter_u(const Factor &F, bool mixed, hipStream_t st);
void launch_cast_f32(const double *a, float *b, long n, hipStream_t st);
void launch_cast_f64(const float *a, double *b, long n, hipStream_t st);
void launch_copy(const double *a, double *b, long n, hipStream_t st);

}  // namespace iccg
