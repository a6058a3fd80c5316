/*
 * iccg_oracle.c -- CPU restatement of the stomo0511/ICCG solve path.
 *
 * TEST INFRASTRUCTURE ONLY (the reference for comparison, never the thing
 * measured or shipped).  Loaded by tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py; the HIP product path never links or calls it.
 *
 * Build: gcc -O2 -std=c11 -ffp-contract=off (see oracle/Makefile).  Every
 * multiply-add that g++ -O3 contracts in the reference build is written here
 * as an explicit fma(); every other operation is a plain IEEE double op, and
 * -ffp-contract=off stops the compiler from fusing anything else.  The dot
 * products keep the reference's x87 80-bit accumulation (long double).
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   coo_to_crs, expand_lower_to_full, ptap_lower, perm_*, color_ptr_rows,
 *   abmc_schedule, jacobi, ic0_factor, ic0_apply, spmv, cg, validators:
 *     bit-exact against oracle/_ref (the reference compiled from its own
 *     sources) on Saad.mm and generated grids, and against closed forms:
 *     scipy permutations/tril, IC(0) == exact Cholesky on tree patterns,
 *     dense triangular solves, numpy.linalg.solve for tiny SPD systems.
 *   orc_dot_dd: against Python fractions (exact rational arithmetic).
 *   orc_cg_mixed: its CG loop is orc_cg and its apply orc_ic0_apply_f32,
 *     both pinned above; the restart rule has no reference counterpart (it
 *     is the precision split's own) and is checked by its invariants: a
 *     solve that converges with fewer than 100 restarts meets tol on the
 *     true residual, no restart happens while the true residual meets tol,
 *     and after each restart the iterates are those of orc_cg started from
 *     the current x (tests/test_oracle.py).
 *   orc_true_resid, orc_true_resid_dd: against exact rational sums of the
 *     same d_i = fl(A x - b)_i (Python fractions), on cases with ||b|| far
 *     from 1 and residuals far from 0, and against the value the reference
 *     binaries print.
 */
#include "iccg_oracle.h"

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* small helpers                                                             */
/* ------------------------------------------------------------------------ */

typedef struct {
  long long k1, k2, k3; /* lexicographic key */
  int idx;
} key3;

static int cmp_key3(const void *a, const void *b) {
  const key3 *x = (const key3 *)a, *y = (const key3 *)b;
  if (x->k1 != y->k1) return x->k1 < y->k1 ? -1 : 1;
  if (x->k2 != y->k2) return x->k2 < y->k2 ? -1 : 1;
  if (x->k3 != y->k3) return x->k3 < y->k3 ? -1 : 1;
  return 0;
}

/* stable insertion sort of (col,val) pairs by col: rows are short */
static void sort_row_pairs(int *c, double *v, long len) {
  for (long a = 1; a < len; ++a) {
    int cc = c[a];
    double vv = v[a];
    long b = a - 1;
    while (b >= 0 && c[b] > cc) {
      c[b + 1] = c[b];
      v[b + 1] = v[b];
      --b;
    }
    c[b + 1] = cc;
    v[b + 1] = vv;
  }
}

/* ------------------------------------------------------------------------ */
/* L0: CRS (crs_io.cpp)                                                      */
/* ------------------------------------------------------------------------ */

/* crs_io.cpp:19-70: order entries by (i,j) (the reference's std::sort; the
 * tie order only matters for duplicate (i,j), summed here in input order),
 * sum runs of equal (i,j) starting from 0.0, then bucket into CSR rows. */
long orc_coo_to_crs(int nrows, long nin, const int *I, const int *J,
                    const double *V, int *rowptr, int *colind, double *val) {
  key3 *ord = (key3 *)malloc(sizeof(key3) * (size_t)(nin > 0 ? nin : 1));
  for (long t = 0; t < nin; ++t) {
    ord[t].k1 = I[t];
    ord[t].k2 = J[t];
    ord[t].k3 = t;
    ord[t].idx = (int)t;
  }
  qsort(ord, (size_t)nin, sizeof(key3), cmp_key3);
  for (int i = 0; i <= nrows; ++i) rowptr[i] = 0;
  long out = 0;
  for (long t = 0; t < nin;) {
    int r = I[ord[t].idx], c = J[ord[t].idx];
    double s = 0.0;
    do {
      s += V[ord[t].idx];
      ++t;
    } while (t < nin && I[ord[t].idx] == r && J[ord[t].idx] == c);
    colind[out] = c;
    val[out] = s;
    rowptr[r + 1]++;
    ++out;
  }
  for (int i = 0; i < nrows; ++i) rowptr[i + 1] += rowptr[i];
  free(ord);
  return out;
}

/* crs_io.cpp:163-227: diagonal once, off-diagonal (i,j) and (j,i), rows then
 * sorted by column. */
long orc_expand_lower_to_full(int n, const int *rp, const int *ci,
                              const double *v, int *frp, int *fci, double *fv) {
  long *cnt = (long *)calloc((size_t)n + 1, sizeof(long));
  long nnzF = 0;
  for (int i = 0; i < n; ++i)
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      int j = ci[k];
      if (j == i) {
        cnt[i]++;
        nnzF++;
      } else {
        cnt[i]++;
        cnt[j]++;
        nnzF += 2;
      }
    }
  frp[0] = 0;
  for (int i = 0; i < n; ++i) frp[i + 1] = frp[i] + (int)cnt[i];
  int *cur = (int *)malloc(sizeof(int) * (size_t)(n + 1));
  memcpy(cur, frp, sizeof(int) * (size_t)(n + 1));
  for (int i = 0; i < n; ++i)
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      int j = ci[k];
      int pos = cur[i]++;
      fci[pos] = j;
      fv[pos] = v[k];
      if (j != i) {
        int pos2 = cur[j]++;
        fci[pos2] = i;
        fv[pos2] = v[k];
      }
    }
  for (int i = 0; i < n; ++i)
    sort_row_pairs(fci + frp[i], fv + frp[i], frp[i + 1] - frp[i]);
  free(cur);
  free(cnt);
  return nnzF;
}

/* ------------------------------------------------------------------------ */
/* L1: orderings (color.cpp, block.cpp)                                      */
/* ------------------------------------------------------------------------ */

/* color.cpp:50-63 */
void orc_perm_by_color(int n, const int *color, int *new_of_old) {
  key3 *ord = (key3 *)malloc(sizeof(key3) * (size_t)(n > 0 ? n : 1));
  for (int a = 0; a < n; ++a) {
    ord[a].k1 = color[a];
    ord[a].k2 = a;
    ord[a].k3 = 0;
    ord[a].idx = a;
  }
  qsort(ord, (size_t)n, sizeof(key3), cmp_key3);
  for (int nw = 0; nw < n; ++nw) new_of_old[ord[nw].idx] = nw;
  free(ord);
}

/* block.cpp:92-114 */
void orc_perm_by_block_color(int n, const int *block_of_old,
                             const int *block_color, int *new_of_old) {
  key3 *ord = (key3 *)malloc(sizeof(key3) * (size_t)(n > 0 ? n : 1));
  for (int a = 0; a < n; ++a) {
    int b = block_of_old[a];
    ord[a].k1 = block_color[b];
    ord[a].k2 = b;
    ord[a].k3 = a;
    ord[a].idx = a;
  }
  qsort(ord, (size_t)n, sizeof(key3), cmp_key3);
  for (int nw = 0; nw < n; ++nw) new_of_old[ord[nw].idx] = nw;
  free(ord);
}

/* color.cpp:66-141: only stored (i,j) with j<=i are read; (new i, new j) is
 * folded so that c<=r and appended to row r in (i,k) scan order; each row
 * is sorted by column, equal columns summed, exact-zero sums dropped. */
long orc_ptap_lower(int n, const int *rp, const int *ci, const double *v,
                    const int *new_of_old, int *brp, int *bci, double *bv) {
  long *cnt = (long *)calloc((size_t)n + 1, sizeof(long));
  long tot = 0;
  for (int i = 0; i < n; ++i)
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      int j = ci[k];
      if (j > i) continue;
      int r = new_of_old[i], c = new_of_old[j];
      if (c > r) r = c;
      cnt[r + 1]++;
      tot++;
    }
  for (int r = 0; r < n; ++r) cnt[r + 1] += cnt[r];
  int *bc = (int *)malloc(sizeof(int) * (size_t)(tot > 0 ? tot : 1));
  double *bvv = (double *)malloc(sizeof(double) * (size_t)(tot > 0 ? tot : 1));
  long *cur = (long *)malloc(sizeof(long) * (size_t)(n + 1));
  memcpy(cur, cnt, sizeof(long) * (size_t)(n + 1));
  for (int i = 0; i < n; ++i) {
    int ip = new_of_old[i];
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      int j = ci[k];
      if (j > i) continue;
      int jp = new_of_old[j];
      int r = ip, c = jp;
      if (c > r) {
        int t = r;
        r = c;
        c = t;
      }
      long pos = cur[r]++;
      bc[pos] = c;
      bvv[pos] = v[k];
    }
  }
  long out = 0;
  brp[0] = 0;
  for (int r = 0; r < n; ++r) {
    long s = cnt[r], e = cnt[r + 1];
    if (e > s) {
      sort_row_pairs(bc + s, bvv + s, e - s);
      int cur_c = bc[s];
      double acc = bvv[s];
      for (long t = s + 1; t < e; ++t) {
        if (bc[t] == cur_c) {
          acc += bvv[t];
        } else {
          if (acc != 0.0) {
            bci[out] = cur_c;
            bv[out] = acc;
            ++out;
          }
          cur_c = bc[t];
          acc = bvv[t];
        }
      }
      if (acc != 0.0) {
        bci[out] = cur_c;
        bv[out] = acc;
        ++out;
      }
    }
    brp[r + 1] = (int)out;
  }
  free(cur);
  free(bvv);
  free(bc);
  free(cnt);
  return out;
}

/* color.cpp:144-153 + 164-197 */
int orc_color_ptr_rows(int n, const int *color, int nc, const int *new_of_old,
                       int *color_ptr, int *rows_of_color) {
  int *old_of_new = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; ++i) old_of_new[i] = -1;
  for (int o = 0; o < n; ++o) {
    int nid = new_of_old[o];
    if (nid < 0 || nid >= n) {
      free(old_of_new);
      return -1;
    }
    old_of_new[nid] = o;
  }
  int *cnt = (int *)calloc((size_t)nc + 1, sizeof(int));
  for (int nw = 0; nw < n; ++nw) {
    int c = color[old_of_new[nw]];
    if ((unsigned)c >= (unsigned)nc) {
      free(cnt);
      free(old_of_new);
      return -1;
    }
    cnt[c]++;
  }
  color_ptr[0] = 0;
  for (int c = 0; c < nc; ++c) color_ptr[c + 1] = color_ptr[c] + cnt[c];
  for (int c = 0; c < nc; ++c) cnt[c] = color_ptr[c];
  for (int nw = 0; nw < n; ++nw) {
    int c = color[old_of_new[nw]];
    rows_of_color[cnt[c]++] = nw;
  }
  free(cnt);
  free(old_of_new);
  return 0;
}

/* block.cpp:116-174 */
int orc_abmc_schedule(int n, const int *new_of_old, const int *block_of_old,
                      const int *block_color, int nb, int nc,
                      int *color_ptr_blk, int *blocks_of_color,
                      int *block_ptr, int *rows_of_block) {
  int *old_of_new = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  for (int o = 0; o < n; ++o) old_of_new[new_of_old[o]] = o;
  char *seen = (char *)calloc((size_t)nb + 1, 1);
  int *cnt = (int *)calloc((size_t)nc + 1, sizeof(int));
  for (int nw = 0; nw < n; ++nw) {
    int b = block_of_old[old_of_new[nw]];
    if (!seen[b]) {
      seen[b] = 1;
      cnt[block_color[b]]++;
    }
  }
  color_ptr_blk[0] = 0;
  for (int c = 0; c < nc; ++c) color_ptr_blk[c + 1] = color_ptr_blk[c] + cnt[c];
  for (int b = 0; b < nb; ++b) blocks_of_color[b] = -1;
  for (int c = 0; c < nc; ++c) cnt[c] = color_ptr_blk[c];
  memset(seen, 0, (size_t)nb + 1);
  for (int nw = 0; nw < n; ++nw) {
    int b = block_of_old[old_of_new[nw]];
    if (!seen[b]) {
      seen[b] = 1;
      blocks_of_color[cnt[block_color[b]]++] = b;
    }
  }
  for (int b = 0; b <= nb; ++b) block_ptr[b] = 0;
  for (int nw = 0; nw < n; ++nw) block_ptr[block_of_old[old_of_new[nw]] + 1]++;
  for (int b = 0; b < nb; ++b) block_ptr[b + 1] += block_ptr[b];
  int *cur = (int *)malloc(sizeof(int) * (size_t)(nb + 1));
  memcpy(cur, block_ptr, sizeof(int) * (size_t)(nb + 1));
  for (int i = 0; i < n; ++i) rows_of_block[i] = -1;
  for (int nw = 0; nw < n; ++nw) {
    int b = block_of_old[old_of_new[nw]];
    rows_of_block[cur[b]++] = nw;
  }
  free(cur);
  free(cnt);
  free(seen);
  free(old_of_new);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* L2: preconditioners                                                       */
/* ------------------------------------------------------------------------ */

/* precond.hpp:23-35: first stored diagonal of row i; 1.0 if 0 or absent */
void orc_jacobi_inv_diag(int n, const int *rp, const int *ci, const double *v,
                         double *inv_diag) {
  for (int i = 0; i < n; ++i) {
    double d = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k)
      if (ci[k] == i) {
        d = v[k];
        break;
      }
    inv_diag[i] = (fabs(d) > 0.0) ? 1.0 / d : 1.0;
  }
}

/* ic0.cpp:6-151 */
long orc_ic0_factor(int n, const int *rp, const int *ci, const double *v,
                    double shift, int *Lrp, int *Lci, double *Lv,
                    int *diag_pos, int *LTcp, int *LTri, double *LTv,
                    int *LT_diag_pos, int *err_row) {
  /* extract_lower_with_diag (ic0.cpp:14-38): keep j<=i in row order */
  long nnz = 0;
  for (int i = 0; i < n; ++i) {
    Lrp[i] = (int)nnz;
    for (int k = rp[i]; k < rp[i + 1]; ++k)
      if (ci[k] <= i) {
        Lci[nnz] = ci[k];
        Lv[nnz] = v[k];
        ++nnz;
      }
  }
  Lrp[n] = (int)nnz;
  /* diagonal positions (ic0.cpp:47-54) */
  for (int i = 0; i < n; ++i) {
    diag_pos[i] = -1;
    for (int k = Lrp[i]; k < Lrp[i + 1]; ++k)
      if (Lci[k] == i) {
        diag_pos[i] = k;
        break;
      }
    if (diag_pos[i] < 0) {
      *err_row = i;
      return -1;
    }
  }
  /* left-looking rows (ic0.cpp:57-115) */
  for (int i = 0; i < n; ++i) {
    for (int kij = Lrp[i]; kij < Lrp[i + 1]; ++kij) {
      int j = Lci[kij];
      if (j >= i) break;
      double s = 0.0;
      int pi = Lrp[i], pj = Lrp[j];
      const int ei = Lrp[i + 1], ej = Lrp[j + 1];
      while (pi < ei && pj < ej) {
        int c1 = Lci[pi], c2 = Lci[pj];
        if (c1 >= j || c2 >= j) break;
        if (c1 == c2) {
          s = fma(Lv[pi], Lv[pj], s);
          ++pi;
          ++pj;
        } else if (c1 < c2) {
          ++pi;
        } else {
          ++pj;
        }
      }
      double Ljj = Lv[diag_pos[j]];
      if (fabs(Ljj) < 1e-30) {
        *err_row = j;
        return -2;
      }
      Lv[kij] = (Lv[kij] - s) / Ljj;
    }
    double s = 0.0;
    for (int k = Lrp[i]; k < Lrp[i + 1]; ++k) {
      int j = Lci[k];
      if (j >= i) break;
      s = fma(Lv[k], Lv[k], s);
    }
    int di = diag_pos[i];
    double diag = Lv[di] - s;
    diag += shift;
    if (diag <= 0.0) {
      *err_row = i;
      return -3;
    }
    Lv[di] = sqrt(diag);
  }
  /* CSC of L = CSR of L^T (ic0.cpp:118-151) */
  for (int c = 0; c <= n; ++c) LTcp[c] = 0;
  for (long k = 0; k < nnz; ++k) LTcp[Lci[k] + 1]++;
  for (int c = 0; c < n; ++c) LTcp[c + 1] += LTcp[c];
  int *ctr = (int *)malloc(sizeof(int) * (size_t)(n + 1));
  memcpy(ctr, LTcp, sizeof(int) * (size_t)(n + 1));
  for (int i = 0; i < n; ++i) LT_diag_pos[i] = -1;
  for (int i = 0; i < n; ++i)
    for (int k = Lrp[i]; k < Lrp[i + 1]; ++k) {
      int c = Lci[k];
      int pos = ctr[c]++;
      LTri[pos] = i;
      LTv[pos] = Lv[k];
      if (i == c) LT_diag_pos[i] = pos;
    }
  free(ctr);
  return nnz;
}

/* one forward row in the MC/ABMC form (ic0.cpp:201-214) */
static inline void fwd_row(const orc_ic0f *f, int i, const double *r,
                           double *y) {
  const double di = f->Lv[f->diag_pos[i]];
  double acc = r[i];
  for (int jj = f->Lrp[i]; jj < f->Lrp[i + 1]; ++jj) {
    int j = f->Lci[jj];
    if (j < i)
      acc = fma(-f->Lv[jj], y[j], acc);
    else if (j == i)
      ;
    else
      break;
  }
  y[i] = acc / di;
}

/* one backward row in the MC/ABMC form (ic0.cpp:229-242) */
static inline void bwd_row(const orc_ic0f *f, int i, const double *y,
                           double *x) {
  const double di = f->LTv[f->LT_diag_pos[i]];
  double acc = y[i];
  for (int jj = f->LTcp[i]; jj < f->LTcp[i + 1]; ++jj) {
    int j = f->LTri[jj];
    if (j > i) acc = fma(-f->LTv[jj], x[j], acc);
  }
  x[i] = acc / di;
}

void orc_ic0_apply(const orc_ic0f *f, const orc_sched *s, const double *r,
                   double *z, double *y) {
  const int n = f->n;
  if (!s || s->nc <= 0) {
    /* forward_solve / backward_solve (ic0.cpp:153-184) */
    for (int i = 0; i < n; ++i) {
      double acc = r[i], di = 1.0;
      for (int k = f->Lrp[i]; k < f->Lrp[i + 1]; ++k) {
        int j = f->Lci[k];
        if (j < i)
          acc = fma(-f->Lv[k], y[j], acc);
        else if (j == i)
          di = f->Lv[k];
        else
          break;
      }
      y[i] = acc / di;
    }
    for (int i = n - 1; i >= 0; --i) {
      double acc = y[i], di = 1.0;
      for (int k = f->LTcp[i]; k < f->LTcp[i + 1]; ++k) {
        int j = f->LTri[k];
        if (j > i)
          acc = fma(-f->LTv[k], z[j], acc);
        else if (j == i)
          di = f->LTv[k];
      }
      z[i] = acc / di;
    }
    return;
  }
  for (int i = 0; i < n; ++i) y[i] = 0.0; /* std::vector<double> y(n) */
  if (s->mode == 1 && s->color_ptr_blk && s->blocks_of_color && s->block_ptr &&
      s->rows_of_block) {
    /* forward_solve_abmc (ic0.cpp:263-300) */
    for (int c = 0; c < s->nc; ++c)
      for (int kb = s->color_ptr_blk[c]; kb < s->color_ptr_blk[c + 1]; ++kb) {
        int b = s->blocks_of_color[kb];
        for (int t = s->block_ptr[b]; t < s->block_ptr[b + 1]; ++t)
          fwd_row(f, s->rows_of_block[t], r, y);
      }
    /* backward_solve_abmc (ic0.cpp:303-339) */
    for (int c = s->nc - 1; c >= 0; --c)
      for (int kb = s->color_ptr_blk[c]; kb < s->color_ptr_blk[c + 1]; ++kb) {
        int b = s->blocks_of_color[kb];
        for (int t = s->block_ptr[b + 1] - 1; t >= s->block_ptr[b]; --t)
          bwd_row(f, s->rows_of_block[t], y, z);
      }
    return;
  }
  if (s->color_ptr && s->rows_of_color) {
    /* forward_solve_mc / backward_solve_mc (ic0.cpp:194-249) */
    for (int c = 0; c < s->nc; ++c)
      for (int k = s->color_ptr[c]; k < s->color_ptr[c + 1]; ++k)
        fwd_row(f, s->rows_of_color[k], r, y);
    for (int c = s->nc - 1; c >= 0; --c)
      for (int k = s->color_ptr[c]; k < s->color_ptr[c + 1]; ++k)
        bwd_row(f, s->rows_of_color[k], y, z);
    return;
  }
  orc_ic0_apply(f, NULL, r, z, y);
}

/* Mixed-precision apply: the FP32 sweeps of the precision split (factor and
 * y, z in float).  Pinned against an exact-rational float32 restatement in
 * tests/test_oracle.py. */
void orc_ic0_apply_f32(const orc_ic0f *f, const double *r, double *z) {
  const int n = f->n;
  float *y = (float *)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
  float *zf = (float *)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; ++i) {
    float acc = (float)r[i], di = 1.0f;
    for (int k = f->Lrp[i]; k < f->Lrp[i + 1]; ++k) {
      int j = f->Lci[k];
      if (j < i)
        acc = fmaf(-(float)f->Lv[k], y[j], acc);
      else if (j == i)
        di = (float)f->Lv[k];
      else
        break;
    }
    y[i] = acc / di;
  }
  for (int i = n - 1; i >= 0; --i) {
    float acc = y[i], di = 1.0f;
    for (int k = f->LTcp[i]; k < f->LTcp[i + 1]; ++k) {
      int j = f->LTri[k];
      if (j > i)
        acc = fmaf(-(float)f->LTv[k], zf[j], acc);
      else if (j == i)
        di = (float)f->LTv[k];
    }
    zf[i] = acc / di;
  }
  for (int i = 0; i < n; ++i) z[i] = (double)zf[i];
  free(y);
  free(zf);
}

/* ------------------------------------------------------------------------ */
/* L3: Krylov                                                                */
/* ------------------------------------------------------------------------ */

/* cg_crs.cpp:33-40 */
double orc_dot_ld(long n, const double *a, const double *b) {
  long double s = 0.0L;
  for (long i = 0; i < n; ++i) s += (long double)a[i] * b[i];
  return (double)s;
}

/* error-free transforms: Knuth TwoSum and the fma TwoProd */
static inline void two_sum(double a, double b, double *s, double *e) {
  double x = a + b;
  double bb = x - a;
  *e = (a - (x - bb)) + (b - bb);
  *s = x;
}

double orc_dot_dd(long n, const double *a, const double *b) {
  double hi = 0.0, lo = 0.0;
  for (long i = 0; i < n; ++i) {
    double p = a[i] * b[i];
    double pe = fma(a[i], b[i], -p);
    double s, e;
    two_sum(hi, p, &s, &e);
    hi = s;
    lo += e + pe;
  }
  double s, e;
  two_sum(hi, lo, &s, &e);
  return s;
}

static double dotm(int mode, long n, const double *a, const double *b) {
  return mode == 0 ? orc_dot_ld(n, a, b) : orc_dot_dd(n, a, b);
}

/* cg_crs.cpp:42-50: sqrt taken in double */
static double nrm2m(int mode, long n, const double *a) {
  return sqrt(dotm(mode, n, a, a));
}

/* cg_crs.cpp:54-79 (OpenMP fallback path, no MKL) */
void orc_spmv(int n, const int *rp, const int *ci, const double *v,
              const double *x, double *y) {
  for (int i = 0; i < n; ++i) {
    double sum = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) sum = fma(v[k], x[ci[k]], sum);
    y[i] = sum;
  }
}

/* cg_crs.cpp:517-528 */
double orc_true_resid(int n, const int *rp, const int *ci, const double *v,
                      const double *x, const double *b) {
  double *Ax = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  orc_spmv(n, rp, ci, v, x, Ax);
  double nr = 0.0, nb = 0.0;
  for (int i = 0; i < n; ++i) {
    double d = Ax[i] - b[i];
    nr = fma(d, d, nr);
    nb = fma(b[i], b[i], nb);
  }
  free(Ax);
  return sqrt(nr) / sqrt(nb);
}

/* The same quotient with both sums accumulated in double-double and each
 * rounded once to double (the form the GPU computes: r = b - A x, whose
 * square equals (A x - b)^2).  ||b|| = 0 is taken as 1, as the solver's
 * normb (cg_crs.cpp:109-110). */
double orc_true_resid_dd(int n, const int *rp, const int *ci, const double *v,
                         const double *x, const double *b) {
  double *Ax = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  orc_spmv(n, rp, ci, v, x, Ax);
  for (int i = 0; i < n; ++i) Ax[i] = Ax[i] - b[i];
  const double nr = orc_dot_dd(n, Ax, Ax), nb2 = orc_dot_dd(n, b, b);
  free(Ax);
  double nb = sqrt(nb2);
  if (nb == 0.0) nb = 1.0;
  return sqrt(nr) / nb;
}

static void precond_apply(int n, int kind, const double *inv_diag,
                          const orc_ic0f *ic, const orc_sched *sched,
                          const double *r, double *z, double *y) {
  if (kind == 0) {
    memcpy(z, r, sizeof(double) * (size_t)n); /* precond.hpp:12-14 */
  } else if (kind == 1) {
    for (int i = 0; i < n; ++i) z[i] = inv_diag[i] * r[i]; /* 38-42 */
  } else if (kind == 2) {
    orc_ic0_apply(ic, sched, r, z, y);
  } else {
    orc_ic0_apply_f32(ic, r, z); /* any valid MC / ABMC order: the same bits */
  }
}

/* cg_crs.cpp:89-162 */
int orc_cg(int n, const int *rp, const int *ci, const double *v,
           const double *b, double *x, int kind, const double *inv_diag,
           const orc_ic0f *ic, const orc_sched *sched, int max_iter,
           double tol, int dot_mode, double *rel_out, int *converged_out,
           double *hist) {
  size_t sz = sizeof(double) * (size_t)(n > 0 ? n : 1);
  double *r = (double *)malloc(sz), *p = (double *)malloc(sz);
  double *z = (double *)malloc(sz), *Ap = (double *)malloc(sz);
  double *Ax = (double *)malloc(sz), *y = (double *)malloc(sz);
  int iters = 0, conv = 0;
  double rel;

  orc_spmv(n, rp, ci, v, x, Ax);
  for (int i = 0; i < n; ++i) r[i] = b[i] - Ax[i];
  double normb = nrm2m(dot_mode, n, b);
  if (normb == 0.0) normb = 1.0;
  precond_apply(n, kind, inv_diag, ic, sched, r, z, y);
  memcpy(p, z, sz);
  double rz_old = dotm(dot_mode, n, r, z);
  rel = nrm2m(dot_mode, n, r) / normb;
  if (hist) hist[0] = rel;
  if (rel < tol) {
    conv = 1;
    iters = 0;
    goto done;
  }
  for (int k = 0; k < max_iter; ++k) {
    orc_spmv(n, rp, ci, v, p, Ap);
    double pAp = dotm(dot_mode, n, p, Ap);
    double eps = DBL_EPSILON;
    double thr = eps * nrm2m(dot_mode, n, p) * nrm2m(dot_mode, n, Ap);
    if (pAp <= thr) {
      iters = k;
      goto done;
    }
    double alpha = rz_old / pAp;
    for (int i = 0; i < n; ++i) {
      x[i] = fma(alpha, p[i], x[i]);
      r[i] = fma(-alpha, Ap[i], r[i]);
    }
    rel = nrm2m(dot_mode, n, r) / normb;
    iters = k + 1;
    if (hist) hist[k + 1] = rel;
    if (rel < tol) {
      conv = 1;
      goto done;
    }
    precond_apply(n, kind, inv_diag, ic, sched, r, z, y);
    double rz_new = dotm(dot_mode, n, r, z);
    double beta = rz_new / rz_old;
    for (int i = 0; i < n; ++i) p[i] = fma(beta, p[i], z[i]);
    rz_old = rz_new;
  }
done:
  *rel_out = rel;
  *converged_out = conv;
  free(r);
  free(p);
  free(z);
  free(Ap);
  free(Ax);
  free(y);
  return iters;
}

/* Mixed precision: the FP32 IC(0) apply (orc_ic0_apply_f32) inside the FP64
 * CG above with double-double dots, stopped on the FP64 true residual.  When
 * the recurrence converges but ||b - A x|| / ||b|| (orc_true_resid_dd) is
 * still >= tol, CG starts again from the current x (r = b - A x, z = M r,
 * p = z), counting iterations on from where it stopped, at most 100 times;
 * a breakdown or max_iter ends the solve.  A restart is orc_cg entered again
 * with x0 = x, so the loop below is exactly that. */
int orc_cg_mixed(int n, const int *rp, const int *ci, const double *v,
                 const double *b, double *x, const orc_ic0f *ic, int max_iter,
                 double tol, double *rel_out, int *converged_out,
                 int *restarts_out, double *true_rel_out) {
  int total = 0, restarts = 0, conv = 0;
  double rel = 0.0;
  for (;;) {
    total += orc_cg(n, rp, ci, v, b, x, 3, NULL, ic, NULL, max_iter - total,
                    tol, 1, &rel, &conv, NULL);
    if (!conv || restarts >= 100) break;
    if (orc_true_resid_dd(n, rp, ci, v, x, b) < tol || total >= max_iter) break;
    ++restarts;
  }
  *rel_out = rel;
  *converged_out = conv;
  *restarts_out = restarts;
  *true_rel_out = orc_true_resid_dd(n, rp, ci, v, x, b);
  return total;
}

/* ------------------------------------------------------------------------ */
/* validators                                                                */
/* ------------------------------------------------------------------------ */

/* cg_crs.cpp:171-224 */
long long orc_check_interblock(int n, const int *rp, const int *ci,
                               const int *block_of_old, const int *block_color,
                               int nb, int nc, long long *conflicts_by_color) {
  long long total = 0;
  for (int c = 0; c < nc; ++c) conflicts_by_color[c] = 0;
  for (int i = 0; i < n; ++i)
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      int j = ci[k];
      if (j <= i) continue;
      int bi = block_of_old[i], bj = block_of_old[j];
      if (bi < 0 || bi >= nb || bj < 0 || bj >= nb) return -1;
      if (bi == bj) continue;
      int c1 = block_color[bi], c2 = block_color[bj];
      if (c1 < 0 || c1 >= nc || c2 < 0 || c2 >= nc) return -1;
      if (c1 == c2) {
        conflicts_by_color[c1]++;
        total++;
      }
    }
  return total;
}

/* cg_crs.cpp:259-336 (lower pattern via extract_lower_with_diag_like,
 * 227-251: stored j<=i entries in row order) */
long long orc_check_hazards(int n, const int *rp, const int *ci,
                            const int *new_of_old, const int *color_old,
                            const int *block_of_old, int *nc_out,
                            long long *hazard_by_color,
                            long long *same_block_by_color,
                            long long *cross_block_by_color, int nc_cap) {
  int *old_of_new = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  for (int o = 0; o < n; ++o) {
    int nid = new_of_old[o];
    if (nid < 0 || nid >= n) {
      free(old_of_new);
      return -1;
    }
    old_of_new[nid] = o;
  }
  int nc = 0;
  for (int i = 0; i < n; ++i)
    if (color_old[old_of_new[i]] + 1 > nc) nc = color_old[old_of_new[i]] + 1;
  if (nc > nc_cap) {
    free(old_of_new);
    return -1;
  }
  for (int c = 0; c < nc; ++c)
    hazard_by_color[c] = same_block_by_color[c] = cross_block_by_color[c] = 0;
  long long total = 0;
  for (int i = 0; i < n; ++i) {
    int c1 = color_old[old_of_new[i]];
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      int j = ci[k];
      if (j > i) continue; /* not in the lower pattern */
      if (j >= i) break;   /* 305: j<i only, rows sorted */
      if (color_old[old_of_new[j]] == c1) {
        hazard_by_color[c1]++;
        total++;
        if (block_of_old) {
          if (block_of_old[old_of_new[i]] == block_of_old[old_of_new[j]])
            same_block_by_color[c1]++;
          else
            cross_block_by_color[c1]++;
        }
      }
    }
  }
  *nc_out = nc;
  free(old_of_new);
  return total;
}
