// api.hip -- C ABI objects of libiccg_hip.so: device CSR, schedules,
// preconditioners (Identity / Jacobi / IC(0) with natural, MC and ABMC
// triangular-solve schedules) and the device-resident CG solver.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "dev.hpp"
#include "internal.hpp"
#include "objects.hpp"

namespace iccg {

static thread_local std::string g_last_error;
void set_last_error(const std::string &msg) { g_last_error = msg; }

static thread_local int g_device = 0;
int current_device() { return g_device; }

}  // namespace iccg

using namespace iccg;

struct TimedPair {
  int cls, iter;
  hipEvent_t a, b;
};

struct iccg_solver_s {
  int device = 0;
  iccg_csr_t A = nullptr;
  iccg_precond_t M = nullptr;
  int n = 0;
  hipStream_t st = nullptr;
  double *x = nullptr, *r = nullptr, *p = nullptr, *Ap = nullptr, *b = nullptr;
  void *y = nullptr, *z = nullptr;  // f64 or f32 (mixed IC0)
  double *part = nullptr;
  int part_slots = 0;  // (hi, lo) pairs the partial buffer holds
  Scal *S = nullptr;
  Scal *hS = nullptr;  // pinned mirror
  hipEvent_t wait_ev = nullptr;  // blocking waits (iccg_solver_set_blocking)
  bool timing = false;
  // hipGraphs of k consecutive iterations, replayed while the plan is the
  // one they were captured from
  bool graphs = true;
  int graph_version = -1;
  std::vector<std::pair<int, hipGraphExec_t>> execs;
  std::vector<TimedPair> pending;
  std::vector<hipEvent_t> pool;
  long t_launch[4] = {0, 0, 0, 0};
  double t_ms[4] = {0, 0, 0, 0};
  void drop_graphs() {
    for (auto &e : execs) (void)hipGraphExecDestroy(e.second);
    execs.clear();
  }
  ~iccg_solver_s() {
    (void)hipSetDevice(device);
    if (st) (void)hipStreamSynchronize(st);  // nothing queued still uses its vectors
    drop_graphs();
    dfree(x);
    dfree(r);
    dfree(p);
    dfree(Ap);
    dfree(b);
    dfree(y);
    dfree(z);
    dfree(part);
    dfree(S);
    hfree(hS);
    if (wait_ev) (void)hipEventDestroy(wait_ev);
    for (auto &e : pool) (void)hipEventDestroy(e);
    for (auto &t : pending) {
      (void)hipEventDestroy(t.a);
      (void)hipEventDestroy(t.b);
    }
    if (st) (void)hipStreamDestroy(st);
  }
};

namespace iccg {

// ---------------------------------------------------------------------------
// preconditioner setup (host integer work + device factorization)
// ---------------------------------------------------------------------------

// L pattern (j<=i, diagonal last) and its values from A; U = CSC of L and
// the L->U position map (ic0.cpp:14-38, 118-151).
static void build_patterns(const HostCsr &A, std::vector<int> &lrp, std::vector<int> &lci,
                           std::vector<double> &lv, std::vector<int> &urp,
                           std::vector<int> &uci, std::vector<int> &l2u) {
  const int n = A.n;
  require(rows_sorted_unique(n, A.rowptr.data(), A.colind.data()), ICCG_ERR_ARG,
          "IC(0): matrix rows must be column-sorted without duplicates");
  lrp.assign((size_t)n + 1, 0);
  for (int i = 0; i < n; ++i) {
    int c = 0;
    bool diag = false;
    for (int k = A.rowptr[i]; k < A.rowptr[i + 1]; ++k)
      if (A.colind[k] <= i) {
        ++c;
        if (A.colind[k] == i) diag = true;
      }
    if (!diag) throw Error(ICCG_ERR_PIVOT, "IC(0): diagonal missing at row " + std::to_string(i));
    lrp[i + 1] = lrp[i] + c;
  }
  const long nnz = lrp[n];
  lci.resize((size_t)nnz);
  lv.resize((size_t)nnz);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int i = 0; i < n; ++i) {
    int w = lrp[i];
    for (int k = A.rowptr[i]; k < A.rowptr[i + 1]; ++k)
      if (A.colind[k] <= i) {
        lci[w] = A.colind[k];
        lv[w] = A.val[k];
        ++w;
      }
  }
  urp.assign((size_t)n + 1, 0);
  for (long k = 0; k < nnz; ++k) urp[(size_t)lci[k] + 1]++;
  for (int c = 0; c < n; ++c) urp[c + 1] += urp[c];
  uci.resize((size_t)nnz);
  l2u.resize((size_t)nnz);
  std::vector<int> cur(urp.begin(), urp.end() - 1);
  for (int i = 0; i < n; ++i)
    for (int k = lrp[i]; k < lrp[i + 1]; ++k) {
      int pos = cur[(size_t)lci[k]]++;
      uci[pos] = i;
      l2u[k] = pos;
    }
}

// level sets of a DAG given as "row i depends on deps(i)"; rows of a level
// ascending.  Forward: deps = L off-diagonals; backward: U off-diagonals.
static void level_sets(int n, const std::vector<int> &rp, const std::vector<int> &ci, bool upper,
                       std::vector<int> &lvl_ptr, std::vector<int> &lvl_rows) {
  std::vector<int> lev((size_t)n, 0);
  int nl = 0;
  if (!upper) {
    for (int i = 0; i < n; ++i) {
      int m = 0;
      for (int k = rp[i]; k < rp[i + 1] - 1; ++k) m = std::max(m, lev[ci[k]] + 1);
      lev[i] = m;
      nl = std::max(nl, m + 1);
    }
  } else {
    for (int i = n - 1; i >= 0; --i) {
      int m = 0;
      for (int k = rp[i] + 1; k < rp[i + 1]; ++k) m = std::max(m, lev[ci[k]] + 1);
      lev[i] = m;
      nl = std::max(nl, m + 1);
    }
  }
  lvl_ptr.assign((size_t)nl + 1, 0);
  for (int i = 0; i < n; ++i) lvl_ptr[(size_t)lev[i] + 1]++;
  for (int l = 0; l < nl; ++l) lvl_ptr[l + 1] += lvl_ptr[l];
  lvl_rows.resize((size_t)n);
  std::vector<int> cur(lvl_ptr.begin(), lvl_ptr.end() - 1);
  for (int i = 0; i < n; ++i) lvl_rows[cur[(size_t)lev[i]]++] = i;
}

static std::vector<int> level_segments(const std::vector<int> &lvl_ptr) {
  const int narrow = 4096;
  std::vector<int> seg;
  const int nl = (int)lvl_ptr.size() - 1;
  int run0 = -1;
  for (int l = 0; l < nl; ++l) {
    int w = lvl_ptr[l + 1] - lvl_ptr[l];
    if (w <= narrow) {
      if (run0 < 0) run0 = l;
    } else {
      if (run0 >= 0) {
        seg.insert(seg.end(), {0, run0, l});
        run0 = -1;
      }
      seg.insert(seg.end(), {1, l, 0});
    }
  }
  if (run0 >= 0) seg.insert(seg.end(), {0, run0, nl});
  return seg;
}

static void build_schedule(iccg_precond_s &M, const iccg_sched_s *s, const std::vector<int> &lrp,
                           const std::vector<int> &lci, const std::vector<int> &urp,
                           const std::vector<int> &uci) {
  const int n = M.n;
  DevSched &sc = M.sc;
  hipStream_t st = M.st;
  const bool natural = (s == nullptr) || s->nc <= 0;
  if (natural) {
    sc.kind = 0;
    std::vector<int> rows_f, rows_b;
    level_sets(n, lrp, lci, false, sc.h_lvl_ptr_f, rows_f);
    level_sets(n, urp, uci, true, sc.h_lvl_ptr_b, rows_b);
    sc.seg_f = level_segments(sc.h_lvl_ptr_f);
    sc.seg_b = level_segments(sc.h_lvl_ptr_b);
    sc.d_lvl_ptr_f = dupload(sc.h_lvl_ptr_f.data(), sc.h_lvl_ptr_f.size(), st);
    sc.d_lvl_rows_f = dupload(rows_f.data(), rows_f.size(), st);
    sc.d_lvl_ptr_b = dupload(sc.h_lvl_ptr_b.data(), sc.h_lvl_ptr_b.size(), st);
    sc.d_lvl_rows_b = dupload(rows_b.data(), rows_b.size(), st);
    HIPCHK(hipStreamSynchronize(st));
    return;
  }
  require(s->n == n, ICCG_ERR_ARG, "schedule size does not match the matrix");
  sc.nc = s->nc;
  std::vector<int> color_of((size_t)n, -1);
  if (s->mode == ICCG_SCHED_MC) {
    sc.kind = 1;
    sc.h_color_ptr_v = s->color_ptr;
    bool ident = true;
    for (int c = 0; c < s->nc; ++c)
      for (int k = s->color_ptr[c]; k < s->color_ptr[c + 1]; ++k) {
        int i = s->rows_of_color[k];
        require(i >= 0 && i < n, ICCG_ERR_SCHEDULE, "schedule row out of range");
        require(color_of[i] == -1, ICCG_ERR_SCHEDULE,
                "schedule lists row " + std::to_string(i) + " twice");
        color_of[i] = c;
        if (i != k) ident = false;
      }
    for (int i = 0; i < n; ++i)
      require(color_of[i] >= 0, ICCG_ERR_SCHEDULE,
              "schedule does not cover row " + std::to_string(i));
    // every dependency j<i must be in an earlier color (no same-color race)
    for (int i = 0; i < n; ++i)
      for (int k = lrp[i]; k < lrp[i + 1] - 1; ++k) {
        int j = lci[k];
        if (color_of[j] >= color_of[i])
          throw Error(ICCG_ERR_SCHEDULE,
                      "coloring is not a valid MC schedule: row " + std::to_string(i) +
                          " depends on row " + std::to_string(j) + " of color " +
                          std::to_string(color_of[j]) + " (row color " +
                          std::to_string(color_of[i]) + ")");
      }
    if (!ident) sc.d_rows = dupload(s->rows_of_color.data(), s->rows_of_color.size(), st);
    HIPCHK(hipStreamSynchronize(st));
    return;
  }
  // ABMC
  sc.kind = 2;
  const int nb = s->nb, nc = s->nc;
  require(nb >= 0 && (int)s->blocks_of_color.size() == nb &&
              (int)s->color_ptr_blk.size() == nc + 1 && (int)s->block_ptr.size() == nb + 1 &&
              (int)s->rows_of_block.size() == n,
          ICCG_ERR_SCHEDULE, "ABMC schedule arrays have inconsistent sizes");
  require(s->color_ptr_blk[0] == 0 && s->color_ptr_blk[nc] == nb, ICCG_ERR_SCHEDULE,
          "color_ptr_blk must run from 0 to nb");
  require(s->block_ptr[0] == 0 && s->block_ptr[nb] == n, ICCG_ERR_SCHEDULE,
          "block_ptr must run from 0 to n");
  std::vector<int> block_of((size_t)n, -1), bseen((size_t)nb, 0);
  std::vector<int> bstart((size_t)nb + 1, 0), brows((size_t)n);
  int pos = 0, maxb = 0;
  bool contiguous = true;
  for (int c = 0; c < nc; ++c) {
    require(s->color_ptr_blk[c] <= s->color_ptr_blk[c + 1], ICCG_ERR_SCHEDULE,
            "color_ptr_blk is not monotone");
    for (int kb = s->color_ptr_blk[c]; kb < s->color_ptr_blk[c + 1]; ++kb) {
      int b = s->blocks_of_color[kb];
      require(b >= 0 && b < nb && !bseen[b], ICCG_ERR_SCHEDULE,
              "blocks_of_color is not a permutation of the blocks");
      bseen[b] = 1;
      bstart[kb] = pos;
      int prev = -1;
      require(s->block_ptr[b] <= s->block_ptr[b + 1], ICCG_ERR_SCHEDULE,
              "block_ptr is not monotone");
      for (int t = s->block_ptr[b]; t < s->block_ptr[b + 1]; ++t) {
        int i = s->rows_of_block[t];
        require(i >= 0 && i < n && color_of[i] == -1, ICCG_ERR_SCHEDULE,
                "rows_of_block is not a permutation of the rows");
        require(i > prev, ICCG_ERR_SCHEDULE, "rows of a block must be in ascending order");
        prev = i;
        color_of[i] = c;
        block_of[i] = kb;
        brows[pos] = i;
        if (i != pos) contiguous = false;
        ++pos;
      }
      maxb = std::max(maxb, s->block_ptr[b + 1] - s->block_ptr[b]);
    }
  }
  require(pos == n, ICCG_ERR_SCHEDULE, "ABMC schedule does not cover every row");
  bstart[nb] = n;
  sc.h_color_blk_v = s->color_ptr_blk;
  sc.max_block = maxb;
  // dependency check (the reference's DebugCheck_ScheduleHazards_OnLower,
  // as a hard error): j<i must be in an earlier color or the same block
  for (int i = 0; i < n; ++i)
    for (int k = lrp[i]; k < lrp[i + 1] - 1; ++k) {
      int j = lci[k];
      if (block_of[j] == block_of[i]) continue;
      if (color_of[j] >= color_of[i])
        throw Error(ICCG_ERR_SCHEDULE,
                    "block coloring is not a valid ABMC schedule: row " + std::to_string(i) +
                        " depends on row " + std::to_string(j) + " of another block of color " +
                        std::to_string(color_of[j]));
    }
  // in-block levels
  std::vector<uint16_t> levf((size_t)n, 0), levb((size_t)n, 0);
  std::vector<int> nlf((size_t)nb, 0), nlb((size_t)nb, 0);
  for (int kb = 0; kb < nb; ++kb) {
    int mf = 0, mb = 0;
    for (int t = bstart[kb]; t < bstart[kb + 1]; ++t) {
      int i = brows[t], m = 0;
      for (int k = lrp[i]; k < lrp[i + 1] - 1; ++k)
        if (block_of[lci[k]] == kb) m = std::max(m, (int)levf[lci[k]] + 1);
      require(m < 65535, ICCG_ERR_SCHEDULE, "block too deep");
      levf[i] = (uint16_t)m;
      mf = std::max(mf, m + 1);
    }
    for (int t = bstart[kb + 1] - 1; t >= bstart[kb]; --t) {
      int i = brows[t], m = 0;
      for (int k = urp[i] + 1; k < urp[i + 1]; ++k)
        if (block_of[uci[k]] == kb) m = std::max(m, (int)levb[uci[k]] + 1);
      require(m < 65535, ICCG_ERR_SCHEDULE, "block too deep");
      levb[i] = (uint16_t)m;
      mb = std::max(mb, m + 1);
    }
    nlf[kb] = mf;
    nlb[kb] = mb;
  }
  sc.h_bstart = bstart;
  sc.d_bstart = dupload(bstart.data(), bstart.size(), st);
  if (!contiguous) sc.d_brows = dupload(brows.data(), brows.size(), st);
  sc.d_lev_f = dupload(levf.data(), levf.size(), st);
  sc.d_lev_b = dupload(levb.data(), levb.size(), st);
  sc.d_nlev_f = dupload(nlf.data(), nlf.size(), st);
  sc.d_nlev_b = dupload(nlb.data(), nlb.size(), st);
  HIPCHK(hipStreamSynchronize(st));
}

// Section width for the ABMC sweeps' sectioned layout (sell_layout_sections),
// or 0 for the plain layout: only for contiguous blocks of <= 64 rows, at
// most 3 in-block and 3 outside entries per row (sweep instantiations 2, 3:
// 2D and 3D box blockings of 5- and 7-point stencils),
// and when the sections pad the factor by at most a quarter.  Fills each
// row's block bounds and outside counts of L and U.
static int abmc_sections(const DevSched &sc, int n, const std::vector<int> &lrp,
                         const std::vector<int> &lci, const std::vector<int> &urp,
                         const std::vector<int> &uci, std::vector<int> &blk, std::vector<int> &lout,
                         std::vector<int> &uout, int &so) {
  so = 0;
  if (sc.kind != 2 || sc.d_brows != nullptr || sc.max_block > 64 || n == 0) return 0;
  blk.assign((size_t)n, 0);
  std::vector<int> blo((size_t)n), bhi((size_t)n);
  const int nb = (int)sc.h_bstart.size() - 1;
  for (int kb = 0; kb < nb; ++kb)
    for (int i = sc.h_bstart[kb]; i < sc.h_bstart[kb + 1]; ++i) {
      blk[(size_t)i] = kb;
      blo[(size_t)i] = sc.h_bstart[kb];
      bhi[(size_t)i] = sc.h_bstart[kb + 1];
    }
  std::vector<int> lin((size_t)n), uin((size_t)n);
  lout.assign((size_t)n, 0);
  uout.assign((size_t)n, 0);
  int S = 0, O = 0;
  for (int i = 0; i < n; ++i) {
    for (int k = lrp[i]; k < lrp[i + 1] - 1; ++k) {
      const bool in = lci[k] >= blo[(size_t)i] && lci[k] < bhi[(size_t)i];
      (in ? lin : lout)[(size_t)i]++;
    }
    for (int k = urp[i] + 1; k < urp[i + 1]; ++k) {
      const bool in = uci[k] >= blo[(size_t)i] && uci[k] < bhi[(size_t)i];
      (in ? uin : uout)[(size_t)i]++;
    }
    S = std::max({S, lin[(size_t)i], uin[(size_t)i]});
    O = std::max({O, lout[(size_t)i], uout[(size_t)i]});
  }
  // kernel shapes: 2 or 3 in-block slots, as many or 6 outside ones; one
  // wavefront per block, so blocks must fill most of theirs (40 of 64 lanes
  // on average)
  if (S > 3 || O > 6 || (long)n < 40L * nb) return 0;
  S = std::max(S, 2);
  O = O <= S ? S : 6;
  // bytes per sweep pair (8-byte values, 4-byte columns): a plain row reads
  // its slice's widths of entries, a sectioned row the packed in-block word,
  // S in-block values and its block's outside widths of entries; lanes past
  // a short block's end re-read its last row, which costs no traffic
  long plain = 0, sect = 0;
  for (int s0 = 0; s0 < n; s0 += 64) {
    int wl = 0, wu = 0;
    const int rows = std::min(n, s0 + 64) - s0;
    for (int i = s0; i < s0 + rows; ++i) {
      wl = std::max(wl, lin[(size_t)i] + lout[(size_t)i]);
      wu = std::max(wu, uin[(size_t)i] + uout[(size_t)i]);
    }
    plain += 12L * rows * (wl + wu);
  }
  for (int kb = 0; kb < nb; ++kb) {
    int ol = 0, ou = 0;
    for (int i = sc.h_bstart[kb]; i < sc.h_bstart[kb + 1]; ++i) {
      ol = std::max(ol, lout[(size_t)i]);
      ou = std::max(ou, uout[(size_t)i]);
    }
    sect += (long)(sc.h_bstart[kb + 1] - sc.h_bstart[kb]) * (8 + 16L * S + 12L * (ol + ou));
  }
  if (4 * sect > 5 * plain) return 0;
  so = O;
  return S;
}

// Sliced-ELL copies of the strictly lower / strictly upper factor and its
// diagonal, in the precision the sweeps use, filled on the device from the
// FP64 factor.
static void build_sliced_factor(iccg_precond_s &M, const std::vector<int> &lrp,
                                const std::vector<int> &lci, const std::vector<int> &urp,
                                const std::vector<int> &uci, const std::vector<int> &l2u) {
  Factor &F = M.F;
  const int n = F.n;
  const long nnz = F.nnz;
  hipStream_t st = M.st;
  std::vector<int> loff, lcol, uoff, ucol;
  std::vector<int> l2ls((size_t)nnz), upos((size_t)nnz, -1), l2us((size_t)nnz);
  int lw = 0, uw = 0;
  auto lrange = [&](int i, int &b, int &e) {
    b = lrp[i];
    e = lrp[i + 1] - 1;
  };
  auto urange = [&](int i, int &b, int &e) {
    b = urp[i] + 1;
    e = urp[i + 1];
  };
  std::vector<int> blk, lout, uout;
  int SO = 0;
  const int S = abmc_sections(M.sc, n, lrp, lci, urp, uci, blk, lout, uout, SO);
  std::vector<unsigned> linb, uinb;
  if (S > 0) {
    sell_layout_sections(n, lrange, lci.data(), M.sc.h_bstart, blk.data(), lout, S, loff, lcol,
                         linb, lw, &l2ls);
    sell_layout_sections(n, urange, uci.data(), M.sc.h_bstart, blk.data(), uout, S, uoff, ucol,
                         uinb, uw, &upos);
  } else {
    sell_layout(n, lrange, lci.data(), loff, lcol, lw, &l2ls);
    sell_layout(n, urange, uci.data(), uoff, ucol, uw, &upos);
  }
  if (std::max(lw, uw) > 16) return;  // wide rows: CSR sweeps only
  F.sec = S;
  F.seco = SO;
  F.fb = false;
  if (S > 0 && M.sc.nc > 0) {
    F.fb = true;
    for (int kb = M.sc.h_color_blk_v[M.sc.nc - 1]; kb < M.sc.h_color_blk_v[M.sc.nc]; ++kb)
      if (uoff[(size_t)kb + 1] - uoff[(size_t)kb] != 64 * S) F.fb = false;
  }
  {
    // load policy of the sweeps: the factor's slots (value + column, plus
    // the packed in-block rows) and the vectors an iteration reads again
    // (p, r; y, z and the diagonal in the sweep precision)
    const double vs = M.mixed ? 4.0 : 8.0;
    const double fac = (double)(lcol.size() + ucol.size()) * (vs + 4.0) + (S > 0 ? 8.0 * n : 0.0);
    const double vec = (double)n * (16.0 + 3.0 * vs);
    F.nt = fac + vec > kInfinityCacheBytes;
  }
  for (int i = 0; i < n; ++i) l2ls[(size_t)lrp[i + 1] - 1] = -1 - i;
  for (long k = 0; k < nnz; ++k) l2us[k] = upos[(size_t)l2u[k]];
  auto up = [&](Sell &S, std::vector<int> &off, std::vector<int> &col, int w) {
    S.n = n;
    S.nslice = (int)off.size() - 1;
    S.maxw = w;
    S.off = dupload(off.data(), off.size(), st);
    S.col = dupload(col.data(), col.size(), st);
    const size_t m = std::max<size_t>(col.size(), 1);
    if (M.mixed) {
      S.v32 = dalloc<float>(m);
      HIPCHK(hipMemsetAsync(S.v32, 0, m * sizeof(float), st));
    } else {
      S.v64 = dalloc<double>(m);
      HIPCHK(hipMemsetAsync(S.v64, 0, m * sizeof(double), st));
    }
  };
  up(F.ls, loff, lcol, lw);
  up(F.us, uoff, ucol, uw);
  if (S > 0) {
    F.ls.inb = dupload(linb.data(), linb.size(), st);
    F.us.inb = dupload(uinb.data(), uinb.size(), st);
  }
  if (M.mixed)
    F.d32 = dalloc<float>((size_t)n);
  else
    F.d64 = dalloc<double>((size_t)n);
  F.l2ls = dupload(l2ls.data(), l2ls.size(), st);
  F.l2us = dupload(l2us.data(), l2us.size(), st);
  F.sell = true;
  launch_scatter_sell(F, M.mixed, st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  dfree(F.l2ls);
  dfree(F.l2us);
  F.l2ls = F.l2us = nullptr;
}

static void build_ic0(iccg_precond_s &M, const iccg_sched_s *s) {
  const HostCsr &A = M.A->h;
  std::vector<int> lrp, lci, urp, uci, l2u;
  std::vector<double> lv;
  build_patterns(A, lrp, lci, lv, urp, uci, l2u);
  build_schedule(M, s, lrp, lci, urp, uci);
  hipStream_t st = M.st;
  Factor &F = M.F;
  F.n = A.n;
  F.nnz = (long)lci.size();
  F.lrp = dupload(lrp.data(), lrp.size(), st);
  F.lci = dupload(lci.data(), lci.size(), st);
  F.lv = dupload(lv.data(), lv.size(), st);
  F.urp = dupload(urp.data(), urp.size(), st);
  F.uci = dupload(uci.data(), uci.size(), st);
  F.l2u = dupload(l2u.data(), l2u.size(), st);
  F.uv = dalloc<double>((size_t)F.nnz);
  if (M.mixed) {
    F.lv32 = dalloc<float>((size_t)F.nnz);
    F.uv32 = dalloc<float>((size_t)F.nnz);
  }
  unsigned long long *derr = dalloc<unsigned long long>(1);
  unsigned long long herr = ~0ULL;
  HIPCHK(hipMemcpyAsync(derr, &herr, sizeof(herr), hipMemcpyHostToDevice, st));
  launch_ic0_factor(M.sc, F, M.shift, derr, st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(&herr, derr, sizeof(herr), hipMemcpyDeviceToHost, st));
  // the factorization's waits sleep: preconditioners are often set up on host
  // threads beside running solves (the serving loop's workers)
  struct Ev {
    hipEvent_t e = blocking_event();
    ~Ev() { (void)hipEventDestroy(e); }
  } ev;
  stream_wait(st, ev.e);
  dfree(derr);
  if (herr != ~0ULL) {
    unsigned row = (unsigned)(herr >> 32), pr = (unsigned)(herr & 0xffffffffULL);
    if (pr == row)
      throw Error(ICCG_ERR_PIVOT, "IC(0): non-positive pivot at row " + std::to_string(row));
    throw Error(ICCG_ERR_PIVOT, "IC(0): zero/near-zero pivot at row " + std::to_string(pr));
  }
  launch_scatter_u(F, M.mixed, st);
  HIPCHK(hipGetLastError());
  build_sliced_factor(M, lrp, lci, urp, uci, l2u);
  stream_wait(st, ev.e);
  dfree(F.l2u);
  F.l2u = nullptr;
}

// ---------------------------------------------------------------------------
// solve loop
// ---------------------------------------------------------------------------
static hipEvent_t pool_get(iccg_solver_s &S) {
  if (!S.pool.empty()) {
    hipEvent_t e = S.pool.back();
    S.pool.pop_back();
    return e;
  }
  hipEvent_t e;
  HIPCHK(hipEventCreate(&e));
  return e;
}

struct Timer {
  iccg_solver_s &S;
  int cls, iter;
  hipEvent_t a = nullptr;
  Timer(iccg_solver_s &s, int c, int it) : S(s), cls(c), iter(it) {
    if (S.timing) {
      a = pool_get(S);
      (void)hipEventRecord(a, S.st);
    }
  }
  ~Timer() {
    if (S.timing) {
      hipEvent_t b = pool_get(S);
      (void)hipEventRecord(b, S.st);
      S.pending.push_back({cls, iter, a, b});
    }
  }
};

static void harvest(iccg_solver_s &S, int useful_iters) {
  for (auto &t : S.pending) {
    bool useful = t.iter < 0 || (t.cls == 0 ? t.iter < useful_iters : t.iter < useful_iters - 1);
    if (useful) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
        S.t_ms[t.cls] += ms;
        S.t_launch[t.cls]++;
      }
    }
    S.pool.push_back(t.a);
    S.pool.push_back(t.b);
  }
  S.pending.clear();
}

static void precond_init(iccg_solver_s &S) {
  iccg_precond_s &M = *S.M;
  const int n = S.n;
  if (M.kind == ICCG_PRECOND_IC0) {
    const bool fb = sweeps_fused(M.sc, M.F);
    launch_fwd(M.sc, M.F, M.mixed, S.r, S.y, S.S, S.st, fb ? S.z : nullptr);
    launch_bwd(M.sc, M.F, M.mixed, S.y, S.z, S.S, S.st, S.r, S.S, S.part, 2, fb);
    if (!bwd_fuses_rz(M.sc, M.F)) launch_rz(n, S.r, S.z, M.mixed, S.S, S.part, 1, S.st);
    launch_update_p(n, S.z, M.mixed, S.p, S.S, 1, S.st);
  } else {
    launch_precond_diag_init(M.kind == ICCG_PRECOND_JACOBI ? PK_JAC : PK_IDENT, n, S.r,
                             M.inv_diag, (double *)S.z, S.p, S.S, S.part, S.st);
  }
}

static void enqueue_iteration(iccg_solver_s &S, int it) {
  iccg_precond_s &M = *S.M;
  const int n = S.n;
  const DevCsr &A = S.A->d;
  {
    Timer t(S, 0, it);
    launch_spmv_cg(A, S.p, S.Ap, S.S, S.part, S.st, false);
  }
  if (M.kind == ICCG_PRECOND_IC0) {
    // with r.z in its own pass, the stop test waits for it (one finish
    // kernel fewer per iteration); with r.z fused into the backward sweep
    // the test stays before the preconditioner
    const bool fused = bwd_fuses_rz(M.sc, M.F);
    // sectioned ABMC and multicolor factors: the forward sweep applies
    // r -= alpha Ap itself and the sweeps also form r.z
    const bool ru = !fused && fwd_updates_r(M.sc, M.F);
    const bool rzs = ru && sweeps_form_rz(M.sc, M.F);
    {
      Timer t(S, 3, it);
      launch_fin_alpha(A, S.S, S.part, S.st);
      if (fused)
        launch_xr(PK_IC, n, S.x, S.p, S.r, S.Ap, nullptr, nullptr, S.S, S.part, S.st);
      else if (!ru)
        launch_r_ic(n, S.r, S.Ap, S.S, S.part, S.st);
    }
    // fused: the last color's backward rows run in the forward timer's
    // launch group
    const bool fb = sweeps_fused(M.sc, M.F);
    int xr_count = 0;
    {
      Timer t(S, 1, it);
      launch_fwd(M.sc, M.F, M.mixed, S.r, S.y, S.S, S.st, fb ? S.z : nullptr,
                 ru ? S.Ap : nullptr, S.part, &xr_count);
    }
    {
      Timer t(S, 2, it);
      launch_bwd(M.sc, M.F, M.mixed, S.y, S.z, S.S, S.st, S.r, S.S, S.part, 1, fb, rzs);
    }
    Timer t(S, 3, it);
    if (fused) {
      launch_update_p(n, S.z, M.mixed, S.p, S.S, 0, S.st);
    } else {
      // r.z came out of the sweeps too
      if (rzs)
        launch_fin_sweep_stop(M.sc, S.S, S.part, S.st);
      else
        launch_rz_stop(n, S.r, S.z, M.mixed, S.S, S.part, S.st, xr_count);
      launch_update_xp(n, S.x, S.z, M.mixed, S.p, S.S, S.st);
    }
  } else {
    const int pk = M.kind == ICCG_PRECOND_JACOBI ? PK_JAC : PK_IDENT;
    Timer t(S, 3, it);
    launch_fin_alpha(A, S.S, S.part, S.st);
    launch_xr(pk, n, S.x, S.p, S.r, S.Ap, M.inv_diag, (double *)S.z, S.S, S.part, S.st);
    launch_update_p(n, pk == PK_JAC ? S.z : (const void *)S.r, false, S.p, S.S, 0, S.st);
  }
}

static void sync_scal(iccg_solver_s &S) {
  HIPCHK(hipMemcpyAsync(S.hS, S.S, sizeof(Scal), hipMemcpyDeviceToHost, S.st));
  stream_wait(S.st, S.wait_ev);
}

// Enqueues k iterations: replays a captured hipGraph of k iterations, or
// launches them one by one while event timing is on.
static void enqueue_chunk(iccg_solver_s &S, int first, int k) {
  if (S.timing || !S.graphs) {
    for (int j = 0; j < k; ++j) enqueue_iteration(S, first + j);
    return;
  }
  if (S.graph_version != S.M->plan_version) {
    S.drop_graphs();
    S.graph_version = S.M->plan_version;
  }
  hipGraphExec_t exec = nullptr;
  for (auto &e : S.execs)
    if (e.first == k) exec = e.second;
  if (!exec) {
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(S.st, hipStreamCaptureModeThreadLocal));
    for (int j = 0; j < k; ++j) enqueue_iteration(S, -1);
    HIPCHK(hipStreamEndCapture(S.st, &g));
    hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHK(e);
    S.execs.push_back({k, exec});
  }
  HIPCHK(hipGraphLaunch(exec, S.st));
}

static void run_solver(iccg_solver_s &S, int max_iter, double tol, int flags,
                       iccg_cg_result &res) {
  HIPCHK(hipSetDevice(S.device));
  const DevCsr &A = S.A->d;
  const bool true_resid = (flags & ICCG_SOLVE_TRUE_RESID) != 0;
  // the preconditioner may have been re-planned since the solver was made
  if (S.M->kind == ICCG_PRECOND_IC0 && bwd_partial_slots(S.M->sc, S.M->F) > S.part_slots) {
    const int slots = bwd_partial_slots(S.M->sc, S.M->F);
    double *np = dalloc<double>((size_t)slots * 2);
    HIPCHK(hipStreamSynchronize(S.st));
    dfree(S.part);
    S.part = np;
    S.part_slots = slots;
    S.drop_graphs();
  }
  auto t0 = std::chrono::steady_clock::now();
  if (flags & ICCG_SOLVE_ZERO_X0)
    HIPCHK(hipMemsetAsync(S.x, 0, sizeof(double) * (size_t)S.n, S.st));
  launch_resid(A, S.x, S.b, S.r, S.S, S.part, 1, tol, max_iter, S.st);
  precond_init(S);
  HIPCHK(hipGetLastError());
  int issued = 0, restarts = 0;
  // chunk lengths double up to 64; once the recurrence residual has fallen
  // from one chunk to the next, a chunk is no longer than the iterations its
  // rate of decrease says are left (rounded up to a power of two, so the
  // captured graphs are reused), so the chunk that reaches the stop runs few
  // iterations past it
  int cap = S.n < 100000 ? 8 : 16, chunk = cap;
  double rel_prev = -1.0;
  int it_prev = 0;
  for (;;) {
    const int k = std::min(chunk, max_iter - issued);
    enqueue_chunk(S, issued, k);
    issued += k;
    HIPCHK(hipGetLastError());
    sync_scal(S);
    if (S.timing) harvest(S, S.hS->iters);
    bool stop = S.hS->done || issued >= max_iter;
    if (stop && true_resid && S.hS->converged && !S.hS->breakdown && restarts < 100) {
      launch_resid(A, S.x, S.b, nullptr, S.S, S.part, 0, tol, max_iter, S.st);
      sync_scal(S);
      const double trel = std::sqrt(S.hS->rr_true) / S.hS->normb;
      if (trel >= tol && S.hS->iters < max_iter) {
        ++restarts;
        launch_resid(A, S.x, S.b, S.r, S.S, S.part, 2, tol, max_iter, S.st);
        precond_init(S);
        issued = S.hS->iters;
        rel_prev = -1.0;
        continue;
      }
    }
    if (stop) break;
    if (cap < 64) cap *= 2;
    chunk = cap;
    const double rel = S.hS->rel;
    const int it = S.hS->iters;
    if (rel_prev > 0.0 && rel > 0.0 && rel < rel_prev && it > it_prev && tol > 0.0) {
      const double rate = (std::log(rel) - std::log(rel_prev)) / (it - it_prev);  // < 0
      const double left = (std::log(tol) - std::log(rel)) / rate;
      int want = 4;
      while (want < cap && want < left + 1.0) want *= 2;
      chunk = want;
    }
    rel_prev = rel;
    it_prev = it;
  }
  auto t1 = std::chrono::steady_clock::now();
  res.iters = S.hS->iters;
  res.rel_resid = S.hS->rel;
  res.converged = S.hS->converged;
  res.breakdown = S.hS->breakdown;
  res.restarts = restarts;
  res.solve_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  // ||b - A x|| / ||b|| (cg_crs.cpp:517-528), outside the timed CG phase
  launch_resid(A, S.x, S.b, nullptr, S.S, S.part, 0, tol, max_iter, S.st);
  sync_scal(S);
  res.true_rel_resid = std::sqrt(S.hS->rr_true) / S.hS->normb;
}

}  // namespace iccg

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char *iccg_last_error(void) { return g_last_error.c_str(); }

const char *iccg_status_string(int s) {
  switch (s) {
    case ICCG_OK: return "ok";
    case ICCG_ERR_ARG: return "invalid argument";
    case ICCG_ERR_IO: return "i/o or parse error";
    case ICCG_ERR_PIVOT: return "IC(0) pivot failure";
    case ICCG_ERR_SCHEDULE: return "invalid schedule";
    case ICCG_ERR_DEVICE: return "HIP runtime error";
    case ICCG_ERR_NOMEM: return "out of memory";
    case ICCG_ERR_STATE: return "invalid state";
  }
  return "unknown status";
}

int iccg_device_count(int *count) {
  return guard([&] {
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) c = 0;
    *count = c;
  });
}

int iccg_device_set(int device) {
  return guard([&] {
    HIPCHK(hipSetDevice(device));
    g_device = device;
  });
}

int iccg_device_get(int *device) {
  return guard([&] {
    require(device, ICCG_ERR_ARG, "null argument");
    *device = g_device;
  });
}

int iccg_device_trim(void) {
  return guard([&] {
    HIPCHK(hipSetDevice(g_device));
    const DevicePool P = device_pool();
    HIPCHK(hipStreamSynchronize(P.st));  // frees queued on the pool's stream have run
    HIPCHK(hipMemPoolTrimTo(P.pool, 0));
    pinned_trim();
  });
}

int iccg_csr_create(int n, long nnz, const int *rowptr, const int *colind, const double *val,
                    iccg_csr_t *out) {
  return guard([&] {
    require(out != nullptr, ICCG_ERR_ARG, "null output handle");
    check_csr(n, nnz, rowptr, colind);
    require(nnz == 0 || val != nullptr, ICCG_ERR_ARG, "null value array");
    HIPCHK(hipSetDevice(g_device));
    std::unique_ptr<iccg_csr_s> A(new iccg_csr_s);
    A->device = g_device;
    A->h.n = n;
    A->h.rowptr.assign(rowptr, rowptr + n + 1);
    A->h.colind.assign(colind, colind + nnz);
    A->h.val.assign(val, val + nnz);
    HIPCHK(hipStreamCreateWithFlags(&A->st, hipStreamNonBlocking));
    A->d.n = n;
    A->d.nnz = nnz;
    A->d.rp = dupload(rowptr, (size_t)n + 1, A->st);
    A->d.ci = dupload(colind, (size_t)nnz, A->st);
    A->d.v = dupload(val, (size_t)nnz, A->st);
    {
      std::vector<int> off, col, pos((size_t)nnz);
      int w = 0;
      sell_layout(
          n, [&](int i, int &b, int &e) { b = rowptr[i]; e = rowptr[i + 1]; }, colind, off, col,
          w, &pos);
      const long padded = off.empty() ? 0 : off.back();
      if (w <= 16 && padded <= 2 * nnz + 64L * ((n + 63) / 64)) {
        std::vector<double> v((size_t)std::max<long>(padded, 1), 0.0);
        for (long k = 0; k < nnz; ++k) v[(size_t)pos[k]] = val[k];
        Sell &S = A->d.sell;
        S.n = n;
        S.nslice = (int)off.size() - 1;
        S.maxw = w;
        S.off = dupload(off.data(), off.size(), A->st);
        S.col = dupload(col.data(), col.size(), A->st);
        // values that are all exact floats also get an f32 copy (half the
        // value bytes for the multi-system SpMV, same products bit for bit)
        bool exact = true;
        for (long k = 0; k < nnz && exact; ++k) exact = (double)(float)val[k] == val[k];
        if (exact) {
          std::vector<float> v32(v.begin(), v.end());
          A->d.sell_v32 = dupload(v32.data(), v32.size(), A->st);
        }
        // the single-system SpMV reads 16-bit columns where they apply
        // (measured 146 vs 156 us at 192^3), with the values and columns in
        // the paired layout (two entries of a lane per load); the triangular
        // sweeps and the multi-system SpMV keep 32-bit columns (there the
        // per-slice table load lengthens a latency-bound chain: measured
        // slower), the latter reading the paired FP64 values
        std::vector<uint16_t> c16;
        std::vector<int> tbl, toff;
        if (sell_col16(off, col, c16, tbl, toff)) {
          pair_slices(off, c16);
          pair_slices(off, v);
          if (c16.empty()) c16.push_back((uint16_t)0xFFFF);
          if (tbl.empty()) tbl.push_back(0);
          S.col16 = dupload(c16.data(), c16.size(), A->st);
          S.tbl = dupload(tbl.data(), tbl.size(), A->st);
          if (!toff.empty()) S.toff = dupload(toff.data(), toff.size(), A->st);
        }
        S.v64 = dupload(v.data(), v.size(), A->st);
        HIPCHK(hipStreamSynchronize(A->st));
      }
    }
    HIPCHK(hipStreamSynchronize(A->st));
    *out = A.release();
  });
}

int iccg_csr_destroy(iccg_csr_t A) {
  delete A;
  return ICCG_OK;
}

int iccg_spmv(iccg_csr_t A, const double *x, double *y) {
  return guard([&] {
    require(A && x && y, ICCG_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(A->device));
    const int n = A->d.n;
    double *dx = dupload(x, (size_t)n, A->st);
    double *dy = dalloc<double>((size_t)n);
    launch_spmv_plain(A->d, dx, dy, A->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(y, dy, sizeof(double) * n, hipMemcpyDeviceToHost, A->st));
    HIPCHK(hipStreamSynchronize(A->st));
    dfree(dx);
    dfree(dy);
  });
}

int iccg_schedule_create_mc(int n, int nc, const int *color_ptr, const int *rows_of_color,
                            iccg_sched_t *out) {
  return guard([&] {
    require(out && n >= 0, ICCG_ERR_ARG, "bad argument");
    std::unique_ptr<iccg_sched_s> s(new iccg_sched_s);
    s->n = n;
    s->mode = ICCG_SCHED_MC;
    s->nc = nc;
    if (nc > 0) {
      require(color_ptr && rows_of_color, ICCG_ERR_ARG, "null schedule array");
      s->color_ptr.assign(color_ptr, color_ptr + nc + 1);
      require(s->color_ptr[0] == 0 && s->color_ptr[nc] == n, ICCG_ERR_SCHEDULE,
              "color_ptr must run from 0 to n");
      for (int c = 0; c < nc; ++c)
        require(s->color_ptr[c] <= s->color_ptr[c + 1], ICCG_ERR_SCHEDULE,
                "color_ptr is not monotone");
      s->rows_of_color.assign(rows_of_color, rows_of_color + n);
    }
    *out = s.release();
  });
}

int iccg_schedule_create_abmc(int n, int nc, int nb, const int *color_ptr_blk,
                              const int *blocks_of_color, const int *block_ptr,
                              const int *rows_of_block, iccg_sched_t *out) {
  return guard([&] {
    require(out && n >= 0 && nb >= 0, ICCG_ERR_ARG, "bad argument");
    std::unique_ptr<iccg_sched_s> s(new iccg_sched_s);
    s->n = n;
    s->mode = ICCG_SCHED_ABMC;
    s->nc = nc;
    s->nb = nb;
    if (nc > 0) {
      require(color_ptr_blk && blocks_of_color && block_ptr && rows_of_block, ICCG_ERR_ARG,
              "null schedule array");
      s->color_ptr_blk.assign(color_ptr_blk, color_ptr_blk + nc + 1);
      s->blocks_of_color.assign(blocks_of_color, blocks_of_color + nb);
      s->block_ptr.assign(block_ptr, block_ptr + nb + 1);
      s->rows_of_block.assign(rows_of_block, rows_of_block + n);
    }
    *out = s.release();
  });
}

int iccg_schedule_destroy(iccg_sched_t s) {
  delete s;
  return ICCG_OK;
}

int iccg_precond_create(iccg_csr_t A, int kind, double shift, iccg_sched_t sched,
                        int precision, iccg_precond_t *out) {
  return guard([&] {
    require(A && out, ICCG_ERR_ARG, "null argument");
    require(kind >= ICCG_PRECOND_IDENTITY && kind <= ICCG_PRECOND_IC0, ICCG_ERR_ARG,
            "unknown preconditioner kind");
    require(precision == ICCG_FP64 || precision == ICCG_MIXED, ICCG_ERR_ARG,
            "unknown precision");
    HIPCHK(hipSetDevice(A->device));
    std::unique_ptr<iccg_precond_s> M(new iccg_precond_s);
    M->device = A->device;
    M->kind = kind;
    M->mixed = (kind == ICCG_PRECOND_IC0) && precision == ICCG_MIXED;
    M->shift = shift;
    M->n = A->d.n;
    M->A = A;
    HIPCHK(hipStreamCreateWithFlags(&M->st, hipStreamNonBlocking));
    if (kind == ICCG_PRECOND_JACOBI) {
      M->inv_diag = dalloc<double>((size_t)M->n);
      launch_jacobi_setup(A->d, M->inv_diag, M->st);
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(M->st));
    } else if (kind == ICCG_PRECOND_IC0) {
      build_ic0(*M, sched);
    }
    *out = M.release();
  });
}

int iccg_precond_destroy(iccg_precond_t M) {
  delete M;
  return ICCG_OK;
}

int iccg_precond_set_schedule(iccg_precond_t M, iccg_sched_t sched) {
  return guard([&] {
    require(M, ICCG_ERR_ARG, "null argument");
    if (M->kind != ICCG_PRECOND_IC0) return;  // Identity / Jacobi ignore sched
    HIPCHK(hipSetDevice(M->device));
    std::vector<int> lrp, lci, urp, uci, l2u;
    std::vector<double> lv;
    build_patterns(M->A->h, lrp, lci, lv, urp, uci, l2u);
    // build into a scratch plan, swap on success; the scratch object frees
    // whichever plan it holds at the end, but never M's stream
    iccg_precond_s tmp;
    struct KeepStream {
      iccg_precond_s &t;
      ~KeepStream() { t.st = nullptr; }
    } keep{tmp};
    tmp.device = M->device;
    tmp.n = M->n;
    tmp.st = M->st;
    build_schedule(tmp, sched, lrp, lci, urp, uci);
    HIPCHK(hipStreamSynchronize(M->st));  // before the old plan's memory is freed
    std::swap(M->sc, tmp.sc);
    // an ABMC factor's sliced copies may be laid out by the old schedule's
    // blocks (sectioned layout): lay them out again for the new one
    Factor &F = M->F;
    if (F.sell && (F.sec > 0 || M->sc.kind == 2)) {
      free_sell(F.ls);
      free_sell(F.us);
      dfree(F.d64);
      dfree(F.d32);
      F.d64 = nullptr;
      F.d32 = nullptr;
      F.sell = false;
      F.sec = 0;
      F.seco = 0;
      F.fb = false;
      F.nt = false;
      build_sliced_factor(*M, lrp, lci, urp, uci, l2u);
    }
    M->plan_version++;
  });
}

int iccg_precond_get_plan(iccg_precond_t M, int *plan, int *section_width) {
  return guard([&] {
    require(M && plan && section_width, ICCG_ERR_ARG, "null argument");
    *plan = ICCG_PLAN_NONE;
    *section_width = 0;
    if (M->kind != ICCG_PRECOND_IC0) return;
    const DevSched &sc = M->sc;
    if (!sliced_plan(sc, M->F))
      *plan = sc.kind == 0            ? ICCG_PLAN_LEVELS
              : block_plan(sc, M->F) ? ICCG_PLAN_ABMC_BLOCKS
                                     : ICCG_PLAN_CSR;
    else if (sc.kind == 1)
      *plan = fwd_updates_r(sc, M->F) ? ICCG_PLAN_MC_SLICED_R : ICCG_PLAN_MC_SLICED;
    else if (M->F.sec > 0) {
      *plan = !sweeps_fused(sc, M->F)  ? ICCG_PLAN_ABMC_SECTIONED
              : fwd_updates_r(sc, M->F) ? ICCG_PLAN_ABMC_SECTIONED_FUSED_R
                                        : ICCG_PLAN_ABMC_SECTIONED_FUSED;
      *section_width = M->F.sec;
    } else
      *plan = ICCG_PLAN_ABMC_SLICED;
  });
}

int iccg_precond_apply(iccg_precond_t M, const double *r, double *z) {
  return guard([&] {
    require(M && r && z, ICCG_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(M->device));
    const int n = M->n;
    if (M->kind == ICCG_PRECOND_IDENTITY) {
      std::memcpy(z, r, sizeof(double) * (size_t)n);
      return;
    }
    double *dr = dupload(r, (size_t)n, M->st);
    double *dz = dalloc<double>((size_t)n);
    if (M->kind == ICCG_PRECOND_JACOBI) {
      launch_jacobi_apply(n, M->inv_diag, dr, dz, M->st);
    } else if (!M->mixed) {
      double *dy = dalloc<double>((size_t)n);
      const bool fb = sweeps_fused(M->sc, M->F);
      launch_fwd(M->sc, M->F, false, dr, dy, nullptr, M->st, fb ? dz : nullptr);
      launch_bwd(M->sc, M->F, false, dy, dz, nullptr, M->st, nullptr, nullptr, nullptr, 0, fb);
      HIPCHK(hipStreamSynchronize(M->st));
      dfree(dy);
    } else {
      float *dy = dalloc<float>((size_t)n), *dz32 = dalloc<float>((size_t)n);
      const bool fb = sweeps_fused(M->sc, M->F);
      launch_fwd(M->sc, M->F, true, dr, dy, nullptr, M->st, fb ? dz32 : nullptr);
      launch_bwd(M->sc, M->F, true, dy, dz32, nullptr, M->st, nullptr, nullptr, nullptr, 0, fb);
      launch_cast_f64(dz32, dz, n, M->st);
      HIPCHK(hipStreamSynchronize(M->st));
      dfree(dy);
      dfree(dz32);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(z, dz, sizeof(double) * n, hipMemcpyDeviceToHost, M->st));
    HIPCHK(hipStreamSynchronize(M->st));
    dfree(dr);
    dfree(dz);
  });
}

int iccg_precond_factor_nnz(iccg_precond_t M, long *nnz_L) {
  return guard([&] {
    require(M && nnz_L, ICCG_ERR_ARG, "null argument");
    require(M->kind == ICCG_PRECOND_IC0, ICCG_ERR_STATE, "not an IC(0) preconditioner");
    *nnz_L = M->F.nnz;
  });
}

int iccg_precond_get_factor(iccg_precond_t M, int *Lrp, int *Lci, double *Lv, int *LTcp,
                            int *LTri, double *LTv) {
  return guard([&] {
    require(M, ICCG_ERR_ARG, "null argument");
    require(M->kind == ICCG_PRECOND_IC0, ICCG_ERR_STATE, "not an IC(0) preconditioner");
    HIPCHK(hipSetDevice(M->device));
    const Factor &F = M->F;
    const size_t n1 = (size_t)F.n + 1, nz = (size_t)F.nnz;
    auto cp = [&](void *h, const void *d, size_t bytes) {
      if (h && bytes) HIPCHK(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, M->st));
    };
    cp(Lrp, F.lrp, n1 * sizeof(int));
    cp(Lci, F.lci, nz * sizeof(int));
    cp(Lv, F.lv, nz * sizeof(double));
    cp(LTcp, F.urp, n1 * sizeof(int));
    cp(LTri, F.uci, nz * sizeof(int));
    cp(LTv, F.uv, nz * sizeof(double));
    HIPCHK(hipStreamSynchronize(M->st));
  });
}

int iccg_solver_create(iccg_csr_t A, iccg_precond_t M, iccg_solver_t *out) {
  return guard([&] {
    require(A && M && out, ICCG_ERR_ARG, "null argument");
    require(M->n == A->d.n, ICCG_ERR_ARG, "preconditioner size does not match the matrix");
    HIPCHK(hipSetDevice(A->device));
    std::unique_ptr<iccg_solver_s> S(new iccg_solver_s);
    S->device = A->device;
    S->A = A;
    S->M = M;
    const int n = A->d.n;
    S->n = n;
    HIPCHK(hipStreamCreateWithFlags(&S->st, hipStreamNonBlocking));
    S->x = dalloc<double>((size_t)n);
    S->r = dalloc<double>((size_t)n);
    S->p = dalloc<double>((size_t)n);
    S->Ap = dalloc<double>((size_t)n);
    S->b = dalloc<double>((size_t)n);
    const size_t vb = M->mixed ? sizeof(float) : sizeof(double);
    if (M->kind == ICCG_PRECOND_IC0) S->y = dalloc<char>(vb * (size_t)n);
    if (M->kind != ICCG_PRECOND_IDENTITY) S->z = dalloc<char>(vb * (size_t)n);
    const int slots = std::max(kMaxRedBlocks * 3, M->kind == ICCG_PRECOND_IC0
                                                      ? bwd_partial_slots(M->sc, M->F)
                                                      : 0);
    S->part = dalloc<double>((size_t)slots * 2);
    S->part_slots = slots;
    S->S = dalloc<Scal>(1);
    HIPCHK(hipMemsetAsync(S->S, 0, sizeof(Scal), S->st));
    HIPCHK(hipMemsetAsync(S->x, 0, sizeof(double) * (size_t)n, S->st));
    HIPCHK(hipMemsetAsync(S->b, 0, sizeof(double) * (size_t)n, S->st));
    S->hS = halloc<Scal>(1);
    HIPCHK(hipStreamSynchronize(S->st));
    *out = S.release();
  });
}

int iccg_solver_destroy(iccg_solver_t S) {
  delete S;
  return ICCG_OK;
}

int iccg_solver_set(iccg_solver_t S, const double *b, const double *x0) {
  return guard([&] {
    require(S && b, ICCG_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(S->device));
    const size_t bytes = sizeof(double) * (size_t)S->n;
    HIPCHK(hipMemcpyAsync(S->b, b, bytes, hipMemcpyHostToDevice, S->st));
    if (x0)
      HIPCHK(hipMemcpyAsync(S->x, x0, bytes, hipMemcpyHostToDevice, S->st));
    else
      HIPCHK(hipMemsetAsync(S->x, 0, bytes, S->st));
    HIPCHK(hipStreamSynchronize(S->st));
  });
}

int iccg_solver_run(iccg_solver_t S, int max_iter, double tol, int flags, iccg_cg_result *res) {
  return guard([&] {
    require(S && res, ICCG_ERR_ARG, "null argument");
    std::memset(res, 0, sizeof(*res));
    res->true_rel_resid = -1.0;
    run_solver(*S, max_iter, tol, flags, *res);
  });
}

int iccg_solver_get_x(iccg_solver_t S, double *x) {
  return guard([&] {
    require(S && x, ICCG_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(S->device));
    HIPCHK(hipMemcpyAsync(x, S->x, sizeof(double) * (size_t)S->n, hipMemcpyDeviceToHost, S->st));
    HIPCHK(hipStreamSynchronize(S->st));
  });
}

int iccg_solver_set_timing(iccg_solver_t S, int on) {
  return guard([&] {
    require(S, ICCG_ERR_ARG, "null argument");
    S->timing = on != 0;
  });
}

int iccg_solver_set_blocking(iccg_solver_t S, int on) {
  return guard([&] {
    require(S, ICCG_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(S->device));
    if (on && !S->wait_ev) S->wait_ev = blocking_event();
    if (!on && S->wait_ev) {
      (void)hipEventDestroy(S->wait_ev);
      S->wait_ev = nullptr;
    }
  });
}

int iccg_solver_set_graphs(iccg_solver_t S, int on) {
  return guard([&] {
    require(S, ICCG_ERR_ARG, "null argument");
    S->graphs = on != 0;
  });
}

int iccg_solver_get_timing(iccg_solver_t S, int cls, long *launches, double *total_ms) {
  return guard([&] {
    require(S && cls >= 0 && cls < 4 && launches && total_ms, ICCG_ERR_ARG, "bad argument");
    *launches = S->t_launch[cls];
    *total_ms = S->t_ms[cls];
  });
}

int iccg_solver_reset_timing(iccg_solver_t S) {
  return guard([&] {
    require(S, ICCG_ERR_ARG, "null argument");
    for (int c = 0; c < 4; ++c) {
      S->t_launch[c] = 0;
      S->t_ms[c] = 0.0;
    }
  });
}

int iccg_cg_solve(iccg_csr_t A, iccg_precond_t M, const double *b, double *x, int max_iter,
                  double tol, int flags, iccg_cg_result *res) {
  iccg_solver_t S = nullptr;
  int st = iccg_solver_create(A, M, &S);
  if (st != ICCG_OK) return st;
  st = iccg_solver_set(S, b, x);
  if (st == ICCG_OK) st = iccg_solver_run(S, max_iter, tol, flags, res);
  if (st == ICCG_OK) st = iccg_solver_get_x(S, x);
  iccg_solver_destroy(S);
  return st;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// The CG solve without allocation (iccg_cg_solve_noalloc): every device
// buffer of the solve (x, r, p, Ap, b, y, z, the partials and the scalars)
// is carved from `scratch`, one file-scope array per device, and the
// iterations are launched one by one (instantiating a captured graph
// allocates).  In host code `scratch` names a host-side stand-in, so its
// device address comes from hipGetSymbolAddress, once per device, and only
// that address reaches kernels and copies.
// ---------------------------------------------------------------------------
namespace iccg {

constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB per device
constexpr int kMaxScratchDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
static std::mutex g_scratch_mu;
static char *g_scratch_dev[kMaxScratchDevices];  // device address, per device
static Scal g_scratch_scal[kMaxScratchDevices];  // host mirror of the scalars, per device

// the device address of this device's scratch array (the device is current)
static char *scratch_base(int dev) {
  require(dev >= 0 && dev < kMaxScratchDevices, ICCG_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    require(p != nullptr, ICCG_ERR_DEVICE, "no device address for the scratch array");
    g_scratch_dev[dev] = static_cast<char *>(p);
  }
  return g_scratch_dev[dev];
}

static size_t round256(size_t b) { return (b + 255) & ~size_t(255); }

// bytes of scratch a solve of A with M takes
static size_t noalloc_bytes(const iccg_csr_s &A, const iccg_precond_s &M, int slots) {
  const size_t n = (size_t)A.d.n;
  const size_t vb = M.mixed ? sizeof(float) : sizeof(double);
  size_t need = 5 * round256(sizeof(double) * n);                   // x, r, p, Ap, b
  if (M.kind == ICCG_PRECOND_IC0) need += round256(vb * n);         // y
  if (M.kind != ICCG_PRECOND_IDENTITY) need += round256(vb * n);    // z
  need += round256(sizeof(double) * 2 * (size_t)slots) + round256(sizeof(Scal));
  return need;
}

}  // namespace iccg

extern "C" {

int iccg_cg_solve_noalloc(iccg_csr_t A, iccg_precond_t M, const double *b, double *x,
                          int max_iter, double tol, int flags, iccg_cg_result *res) {
  return guard([&] {
    require(A && M && b && x && res, ICCG_ERR_ARG, "null argument");
    require(M->n == A->d.n, ICCG_ERR_ARG, "preconditioner size does not match the matrix");
    require(M->device == A->device, ICCG_ERR_ARG, "matrix and preconditioner on other devices");
    HIPCHK(hipSetDevice(A->device));
    char *base = scratch_base(A->device);
    const int slots = std::max(kMaxRedBlocks * 3, M->kind == ICCG_PRECOND_IC0
                                                       ? bwd_partial_slots(M->sc, M->F)
                                                       : 0);
    const size_t need = noalloc_bytes(*A, *M, slots);
    require(need <= kScratchBytes, ICCG_ERR_ARG,
            "iccg_cg_solve_noalloc: this system needs " + std::to_string(need) +
                " bytes of device scratch, more than the " + std::to_string(kScratchBytes) +
                " the library holds per device (about 1.1M rows in FP64, 1.3M mixed)");
    const int n = A->d.n;
    const size_t vb = M->mixed ? sizeof(float) : sizeof(double);
    size_t off = 0;
    auto take = [&](size_t bytes) {
      char *p = base + off;
      off += round256(bytes);
      return p;
    };
    // a solver over the scratch memory, launched kernel by kernel on the
    // matrix's stream; the guard hands every pointer back before the
    // solver's destructor could free it
    iccg_solver_s S;
    struct Release {
      iccg_solver_s &s;
      ~Release() {
        s.x = s.r = s.p = s.Ap = s.b = s.part = nullptr;
        s.y = s.z = nullptr;
        s.S = s.hS = nullptr;
        s.st = nullptr;
      }
    } release{S};
    S.device = A->device;
    S.A = A;
    S.M = M;
    S.n = n;
    S.st = A->st;
    S.graphs = false;
    S.x = reinterpret_cast<double *>(take(sizeof(double) * (size_t)n));
    S.r = reinterpret_cast<double *>(take(sizeof(double) * (size_t)n));
    S.p = reinterpret_cast<double *>(take(sizeof(double) * (size_t)n));
    S.Ap = reinterpret_cast<double *>(take(sizeof(double) * (size_t)n));
    S.b = reinterpret_cast<double *>(take(sizeof(double) * (size_t)n));
    if (M->kind == ICCG_PRECOND_IC0) S.y = take(vb * (size_t)n);
    if (M->kind != ICCG_PRECOND_IDENTITY) S.z = take(vb * (size_t)n);
    S.part = reinterpret_cast<double *>(take(sizeof(double) * 2 * (size_t)slots));
    S.part_slots = slots;
    S.S = reinterpret_cast<Scal *>(take(sizeof(Scal)));
    S.hS = &g_scratch_scal[A->device];
    // the scalars start from zero (the last call's values stay in the
    // array), then b and x0 go up, as iccg_cg_solve does
    std::memset(S.hS, 0, sizeof(Scal));
    const size_t bytes = sizeof(double) * (size_t)n;
    HIPCHK(hipMemcpyAsync(S.S, S.hS, sizeof(Scal), hipMemcpyHostToDevice, S.st));
    HIPCHK(hipMemcpyAsync(S.b, b, bytes, hipMemcpyHostToDevice, S.st));
    HIPCHK(hipMemcpyAsync(S.x, x, bytes, hipMemcpyHostToDevice, S.st));
    HIPCHK(hipStreamSynchronize(S.st));
    std::memset(res, 0, sizeof(*res));
    res->true_rel_resid = -1.0;
    run_solver(S, max_iter, tol, flags, *res);
    HIPCHK(hipMemcpyAsync(x, S.x, bytes, hipMemcpyDeviceToHost, S.st));
    HIPCHK(hipStreamSynchronize(S.st));
  });
}

int iccg_bench_copy(long nbytes, int reps, double *avg_ms) {
  return guard([&] {
    require(nbytes >= 8 && reps >= 1 && avg_ms, ICCG_ERR_ARG, "bad argument");
    HIPCHK(hipSetDevice(g_device));
    const long n = nbytes / 8;
    double *a = dalloc<double>((size_t)n), *b = dalloc<double>((size_t)n);
    hipStream_t st;
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    HIPCHK(hipMemsetAsync(a, 0, (size_t)n * 8, st));
    launch_copy(a, b, n, st);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, st));
    for (int k = 0; k < reps; ++k) launch_copy(a, b, n, st);
    HIPCHK(hipEventRecord(e1, st));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    *avg_ms = ms / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    dfree(a);
    dfree(b);
  });
}

}  // extern "C"
