// kernels.hip -- gfx950 kernels of the ICCG solve path.
//
// Arithmetic contract (checked bit-for-bit against the CPU oracle):
//   * every row sum runs in one lane in ascending stored order with an
//     explicit fma per term, as the reference's loops compile (cg_crs.cpp:
//     68-77 SpMV, ic0.cpp:197-248 / 270-338 triangular solves, ic0.cpp:57-115
//     factorization); division and sqrt are IEEE;
//   * dot products accumulate per lane in double-double (TwoProd by fma,
//     TwoSum), reduce in a fixed tree, and round once to double -- the role
//     of the reference's 80-bit long double accumulation (cg_crs.cpp:33-50);
//   * scalars (alpha, beta, norms, stop flags) stay on the device: the last
//     workgroup to finish a reducing kernel folds all partials in workgroup
//     order and updates Scal, so a CG iteration is a fixed launch sequence
//     with no host round trip.
// Memory layout for HBM bandwidth: matrices in 64-row sliced ELL (one
// wavefront per slice, every load a contiguous 512-byte run); streaming
// vector kernels move 32 bytes per lane per access (two 16-byte loads).
// Compiled with -ffp-contract=off: only the fma() calls written here fuse.
#include <algorithm>
#include <cfloat>
#include <cstdlib>

#include "dev.hpp"
#include "devfun.hpp"

namespace iccg {

// ---------------------------------------------------------------------------
// SpMV (cg_crs.cpp:54-79): one lane per row, ascending fma chain from 0.0
// ---------------------------------------------------------------------------
__device__ __forceinline__ double row_dot(const int *__restrict__ rp, const int *__restrict__ ci,
                                          const double *__restrict__ v,
                                          const double *__restrict__ x, int i) {
  double sum = 0.0;
  const int s = rp[i], e = rp[i + 1];
  for (int k = s; k < e; ++k) sum = fma(v[k], x[ci[k]], sum);
  return sum;
}

// Visits the rows of A: sliced ELL (one wavefront per 64-row slice) or CSR.
// f(i, Ax_i) for every row i < n; lanes past n get no call.
// TV: the 16-bit tables have per-slice lengths (Sell::toff set).
template <int WM, bool TV, class F>
__device__ __forceinline__ void for_rows_Ax(const DevCsr &A, const double *__restrict__ x, F f) {
  if (WM > 0 && A.sell.off && A.sell.col16) {
    // 16-bit columns: the values and columns are in the paired layout
    constexpr int K = WM > 0 ? WM : 2;
    __shared__ int tbs[kWG >> 6][TV ? kTbl16Max : 16];
    const int lane = threadIdx.x & 63;
    int *tb = tbs[threadIdx.x >> 6];
    const WaveRange wr = xcd_waves(0, A.sell.nslice);
    const int first = __builtin_amdgcn_readfirstlane(wr.first);
    const int step = __builtin_amdgcn_readfirstlane(wr.step);
    const int end = __builtin_amdgcn_readfirstlane(wr.end);
    if (first < end) {
      // the next slice's loads go out before the current slice's gathers
      Slice16<K> cur;
      spmv16_load<K, TV>(A.sell.off, A.sell.col16, A.sell.tbl, A.sell.toff, A.sell.v64, first,
                         lane, cur);
      for (int s = first; s < end; s += step) {
        Slice16<K> nxt = cur;
        if (s + step < end)
          spmv16_load<K, TV>(A.sell.off, A.sell.col16, A.sell.tbl, A.sell.toff, A.sell.v64,
                             s + step, lane, nxt);
        const double y = spmv16_finish<K, TV>(cur, tb, A.sell.tbl, x, lane);
        const int i = (s << 6) + lane;
        if (i < A.n) f(i, y);
        cur = nxt;
      }
    }
  } else if (A.sell.off && !A.sell.col16) {
    const int lane = threadIdx.x & 63;
    const WaveRange wr = xcd_waves(0, A.sell.nslice);
    for (int s = wr.first; s < wr.end; s += wr.step) {
      const double y = sell_chain<double, double, double, WM, false, true>(A.sell.off, A.sell.col,
                                                                    A.sell.v64, x, s, lane, 0.0);
      const int i = (s << 6) + lane;
      if (i < A.n) f(i, y);
    }
  } else {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < A.n; i += gridDim.x * blockDim.x)
      f(i, row_dot(A.rp, A.ci, A.v, x, i));
  }
}

template <int WM, bool TV>
__global__ __launch_bounds__(kWG) void k_spmv_plain(DevCsr A, const double *__restrict__ x,
                                                    double *__restrict__ y) {
  for_rows_Ax<WM, TV>(A, x, [&](int i, double v) { y[i] = v; });
}

// Ap = A p with p.Ap, ||p||^2, ||Ap||^2; the last workgroup applies the
// breakdown test (cg_crs.cpp:127-133) and computes alpha (135).
template <int WM, bool TV>
__global__ __launch_bounds__(kWG) void k_spmv_cg(DevCsr A, const double *__restrict__ p,
                                                 double *__restrict__ Ap, Scal *S,
                                                 double *part) {
  if (S->done) return;
  double h[3] = {0, 0, 0}, l[3] = {0, 0, 0};
  for_rows_Ax<WM, TV>(A, p, [&](int i, double y) {
    __builtin_nontemporal_store(y, &Ap[i]);
    const double pi = p[i];
    dd_fma(h[0], l[0], pi, y);
    dd_fma(h[1], l[1], pi, pi);
    dd_fma(h[2], l[2], y, y);
  });
  block_partial<3>(h, l, part, blockIdx.x);
}

// breakdown test (cg_crs.cpp:127-133) and alpha (135).  A solve that
// stopped before this iteration is marked done = 2, so that the x update of
// k_update_xp, which runs for the iteration that stops (done = 1 from the
// stop test), is not repeated by the iterations enqueued after it.
__global__ __launch_bounds__(kFinWG) void k_fin_alpha(const double *__restrict__ part, int count,
                                                   Scal *S) {
  if (S->done) {
    if (threadIdx.x == 0) S->done = 2;
    return;
  }
  double out[3];
  if (finish_reduce<3>(part, count, out)) {
    const double pAp = out[0];
    const double thr = DBL_EPSILON * sqrt(out[1]) * sqrt(out[2]);
    S->pAp = pAp;
    if (pAp <= thr) {
      S->breakdown = 1;
      S->done = 1;
    } else {
      S->alpha = S->rz_old / pAp;
    }
  }
}

// r = b - A x (cg_crs.cpp:102-110, 118-123; true residual 517-528)
template <int WM, bool TV>
__global__ __launch_bounds__(kWG) void k_resid(DevCsr A, const double *__restrict__ x,
                                               const double *__restrict__ b,
                                               double *__restrict__ r, Scal *S, double *part,
                                               int mode, double tol, int max_iter) {
  double h[2] = {0, 0}, l[2] = {0, 0};
  for_rows_Ax<WM, TV>(A, x, [&](int i, double ax) {
    const double bi = b[i];
    const double ri = bi - ax;
    if (mode != 0) r[i] = ri;
    dd_fma(h[0], l[0], bi, bi);
    dd_fma(h[1], l[1], ri, ri);
  });
  block_partial<2>(h, l, part, blockIdx.x);
}

__global__ __launch_bounds__(kFinWG) void k_fin_resid(const double *__restrict__ part, int count,
                                                   Scal *S, int mode, double tol, int max_iter) {
  double out[2];
  if (finish_reduce<2>(part, count, out)) {
    double normb = sqrt(out[0]);
    if (normb == 0.0) normb = 1.0;
    if (mode == 0) {
      S->rr_true = out[1];
      S->normb = normb;
      return;
    }
    const double rel = sqrt(out[1]) / normb;
    if (mode == 1) {
      S->tol = tol;
      S->max_iter = max_iter;
      S->iters = 0;
    }
    S->normb = normb;
    S->rel = rel;
    S->converged = 0;
    S->breakdown = 0;
    S->done = 0;
    if (rel < S->tol) {
      S->converged = 1;
      S->done = 1;
    } else if (S->iters >= S->max_iter) {
      S->done = 1;
    }
  }
}

// ---------------------------------------------------------------------------
// Streaming vector kernels (4 elements per lane per step)
// ---------------------------------------------------------------------------
#define QUAD_LOOP(n)                                                                  \
  for (int q_ = blockIdx.x * kWG + threadIdx.x, nq_ = (n) >> 2; q_ < nq_;            \
       q_ += gridDim.x * kWG)
#define TAIL_LOOP(n) \
  for (int i = (((n) >> 2) << 2) + blockIdx.x * kWG + threadIdx.x; i < (n); i += gridDim.x * kWG)

// x += alpha p; r -= alpha Ap (cg_crs.cpp:137-141) with ||r|| and the stop
// test (143-149).  For Identity / Jacobi the preconditioner (z = r or
// z = D^-1 r, precond.hpp:12-14, 38-42) and r.z / beta (152-153) fuse here.
template <int PK>
__global__ __launch_bounds__(kWG) void k_xr(int n, double *__restrict__ x,
                                            const double *__restrict__ p,
                                            double *__restrict__ r,
                                            const double *__restrict__ Ap,
                                            const double *__restrict__ inv_diag,
                                            double *__restrict__ z, Scal *S, double *part) {
  if (S->done) return;
  const double alpha = S->alpha;
  double h[2] = {0, 0}, l[2] = {0, 0};
  QUAD_LOOP(n) {
    const int i = q_ << 2;
    double xv[4], pv[4], rv[4], av[4];
    ld4_nt(x, i, xv);  // x and Ap: read by no other kernel of the iteration
    ld4(p, i, pv);
    ld4(r, i, rv);
    ld4_nt(Ap, i, av);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      xv[e] = fma(alpha, pv[e], xv[e]);
      rv[e] = fma(-alpha, av[e], rv[e]);
    }
    st4_nt(x, i, xv);
    st4(r, i, rv);
#pragma unroll
    for (int e = 0; e < 4; ++e) dd_fma(h[0], l[0], rv[e], rv[e]);
    if (PK == PK_JAC) {
      double dv[4], zv[4];
      ld4(inv_diag, i, dv);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        zv[e] = dv[e] * rv[e];
        dd_fma(h[1], l[1], rv[e], zv[e]);
      }
      st4(z, i, zv);
    }
  }
  TAIL_LOOP(n) {
    const double xi = fma(alpha, p[i], x[i]);
    const double ri = fma(-alpha, Ap[i], r[i]);
    x[i] = xi;
    r[i] = ri;
    dd_fma(h[0], l[0], ri, ri);
    if (PK == PK_JAC) {
      const double zi = inv_diag[i] * ri;
      z[i] = zi;
      dd_fma(h[1], l[1], ri, zi);
    }
  }
  block_partial<2>(h, l, part, blockIdx.x);
}

// r -= alpha Ap (cg_crs.cpp:141) with the ||r||^2 partials of the deferred
// stop test (IC(0) iterations; x += alpha p, line 139, moves to k_update_xp,
// which reads p anyway)
__global__ __launch_bounds__(kWG) void k_r_ic(int n, double *__restrict__ r,
                                              const double *__restrict__ Ap, const Scal *S,
                                              double *part) {
  if (S->done) return;
  const double alpha = S->alpha;
  double h[1] = {0}, l[1] = {0};
  QUAD_LOOP(n) {
    const int i = q_ << 2;
    double rv[4], av[4];
    ld4(r, i, rv);
    ld4_nt(Ap, i, av);  // Ap: read by no other kernel of the iteration
#pragma unroll
    for (int e = 0; e < 4; ++e) rv[e] = fma(-alpha, av[e], rv[e]);
    st4(r, i, rv);
#pragma unroll
    for (int e = 0; e < 4; ++e) dd_fma(h[0], l[0], rv[e], rv[e]);
  }
  TAIL_LOOP(n) {
    const double ri = fma(-alpha, Ap[i], r[i]);
    r[i] = ri;
    dd_fma(h[0], l[0], ri, ri);
  }
  block_partial<1>(h, l, part, blockIdx.x);
}

// stop test (cg_crs.cpp:143-149); Identity / Jacobi also beta (152-153)
template <int PK>
__global__ __launch_bounds__(kFinWG) void k_fin_xr(const double *__restrict__ part, int count,
                                                Scal *S) {
  if (S->done) return;
  double out[2];
  if (finish_reduce<2>(part, count, out)) {
    const double rel = sqrt(out[0]) / S->normb;
    const int it = S->iters + 1;
    S->iters = it;
    S->rel = rel;
    if (rel < S->tol) {
      S->converged = 1;
      S->done = 1;
    } else if (it >= S->max_iter) {
      S->done = 1;
    } else if (PK != PK_IC) {
      const double rz_new = (PK == PK_IDENT) ? out[0] : out[1];
      S->beta = rz_new / S->rz_old;
      S->rz_old = rz_new;
    }
  }
}

// The IC(0) iteration's stop test (cg_crs.cpp:143-149) taken after the
// preconditioner instead of before it: the same |r| partials, the same
// decision, and when the iteration goes on, beta from r.z (152-153).  In the
// iteration that stops, the preconditioner and r.z ran for nothing and
// k_update_xp applies only the x update (done == 1), so x, r, iters and rel
// are those of the immediate test.  The partials come from k_r_ic and k_rz,
// or from the sweeps that apply the r update and form r.z.
__global__ __launch_bounds__(kFinWG) void k_fin_xr_rz(const double *__restrict__ part_xr,
                                                   int count_xr,
                                                   const double *__restrict__ part_rz,
                                                   int count_rz, Scal *S) {
  if (S->done) return;
  double rr[1], rz[1];  // part_xr: k_r_ic's one-component partials
  if (finish_reduce2<1, 1>(part_xr, count_xr, part_rz, count_rz, rr, rz)) {
    const double rel = sqrt(rr[0]) / S->normb;
    const int it = S->iters + 1;
    S->iters = it;
    S->rel = rel;
    if (rel < S->tol) {
      S->converged = 1;
      S->done = 1;
    } else if (it >= S->max_iter) {
      S->done = 1;
    } else {
      S->beta = rz[0] / S->rz_old;
      S->rz_old = rz[0];
    }
  }
}

// z0 = M^-1 r0, p = z0, rz_old = r.z (cg_crs.cpp:112-115) for Identity/Jacobi
template <int PK>
__global__ __launch_bounds__(kWG) void k_precond_diag_init(int n, const double *__restrict__ r,
                                                           const double *__restrict__ inv_diag,
                                                           double *__restrict__ z,
                                                           double *__restrict__ p, Scal *S,
                                                           double *part) {
  if (S->done) return;
  double h[1] = {0}, l[1] = {0};
  for (int i = blockIdx.x * kWG + threadIdx.x; i < n; i += gridDim.x * kWG) {
    const double ri = r[i];
    const double zi = (PK == PK_JAC) ? inv_diag[i] * ri : ri;
    if (PK == PK_JAC) z[i] = zi;
    p[i] = zi;
    dd_fma(h[0], l[0], ri, zi);
  }
  block_partial<1>(h, l, part, blockIdx.x);
}

// rz_new = r.z; beta = rz_new / rz_old (cg_crs.cpp:152-153), or rz_old at
// the start of a solve (115)
__global__ __launch_bounds__(kFinWG) void k_fin_rz(const double *__restrict__ part, int count,
                                                Scal *S, int init) {
  if (S->done) return;
  double out[1];
  if (finish_reduce<1>(part, count, out)) {
    if (init) {
      S->rz_old = out[0];
    } else {
      S->beta = out[0] / S->rz_old;
      S->rz_old = out[0];
    }
  }
}

// rz_new = r.z; beta = rz_new / rz_old (cg_crs.cpp:152-153), or rz_old at
// start (115)
template <class ZT>
__global__ __launch_bounds__(kWG) void k_rz(int n, const double *__restrict__ r,
                                            const ZT *__restrict__ z, Scal *S, double *part,
                                            int init) {
  if (S->done) return;
  double h[1] = {0}, l[1] = {0};
  QUAD_LOOP(n) {
    const int i = q_ << 2;
    double rv[4], zv[4];
    ld4(r, i, rv);
    ld4(z, i, zv);
#pragma unroll
    for (int e = 0; e < 4; ++e) dd_fma(h[0], l[0], rv[e], zv[e]);
  }
  TAIL_LOOP(n) { dd_fma(h[0], l[0], r[i], (double)z[i]); }
  block_partial<1>(h, l, part, blockIdx.x);
}

// p = z + beta p (cg_crs.cpp:155-157), or p = z at start (114)
template <class ZT>
__global__ __launch_bounds__(kWG) void k_update_p(int n, const ZT *__restrict__ z,
                                                  double *__restrict__ p, const Scal *S,
                                                  int init) {
  if (S->done) return;
  const double beta = S->beta;
  QUAD_LOOP(n) {
    const int i = q_ << 2;
    double zv[4], pv[4];
    ld4(z, i, zv);
    if (init) {
      st4(p, i, zv);
    } else {
      ld4(p, i, pv);
#pragma unroll
      for (int e = 0; e < 4; ++e) pv[e] = fma(beta, pv[e], zv[e]);
      st4(p, i, pv);
    }
  }
  TAIL_LOOP(n) {
    const double zi = (double)z[i];
    p[i] = init ? zi : fma(beta, p[i], zi);
  }
}

// x += alpha p (cg_crs.cpp:139) and p = z + beta p (155-157) of an IC(0)
// iteration, the x update deferred from k_r_ic: it runs for every iteration
// whose x/r update ran (not after a breakdown, not once the solve had
// stopped, done == 2), including the one whose stop test ends the solve
// (done == 1), which updates no p.
template <class ZT>
__global__ __launch_bounds__(kWG) void k_update_xp(int n, double *__restrict__ x,
                                                   const ZT *__restrict__ z,
                                                   double *__restrict__ p, const Scal *S) {
  if (S->done > 1 || S->breakdown) return;
  const double alpha = S->alpha, beta = S->beta;
  const bool pu = !S->done;
  QUAD_LOOP(n) {
    const int i = q_ << 2;
    double xv[4], pv[4];
    ld4_nt(x, i, xv);  // x: read by no other kernel of the iteration
    ld4(p, i, pv);
#pragma unroll
    for (int e = 0; e < 4; ++e) xv[e] = fma(alpha, pv[e], xv[e]);
    st4_nt(x, i, xv);
    if (pu) {
      double zv[4];
      ld4(z, i, zv);
#pragma unroll
      for (int e = 0; e < 4; ++e) pv[e] = fma(beta, pv[e], zv[e]);
      st4(p, i, pv);
    }
  }
  TAIL_LOOP(n) {
    const double pi = p[i];
    x[i] = fma(alpha, pi, x[i]);
    if (pu) p[i] = fma(beta, pi, (double)z[i]);
  }
}

// Jacobi (precond.hpp:23-42)
__global__ void k_jacobi_setup(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                               const double *__restrict__ v, double *__restrict__ inv) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double d = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k)
      if (ci[k] == i) {
        d = v[k];
        break;
      }
    inv[i] = (fabs(d) > 0.0) ? 1.0 / d : 1.0;
  }
}

__global__ void k_jacobi_apply(int n, const double *__restrict__ inv,
                               const double *__restrict__ r, double *__restrict__ z) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    z[i] = inv[i] * r[i];
}

// ---------------------------------------------------------------------------
// CSR row operators (natural-order level schedule, factorization, and the
// general fallbacks)
// ---------------------------------------------------------------------------

// L y = r, one row (ic0.cpp:153-167 / 201-214): j<i terms in stored order,
// diagonal stored last in the row.
template <class VT, class YT>
struct FwdRow {
  const int *rp, *ci;
  const VT *v;
  const double *r;
  YT *y;
  __device__ __forceinline__ void operator()(int i) const {
    const int s = rp[i], e = rp[i + 1] - 1;
    YT acc = (YT)r[i];
    for (int k = s; k < e; ++k) acc = fma_t(-(YT)v[k], y[ci[k]], acc);
    y[i] = acc / (YT)v[e];
  }
};

// L^T z = y, one row (ic0.cpp:169-184 / 229-242): row i of L^T = column i
// of L (CSC order, ascending), diagonal stored first.
template <class VT, class YT>
struct BwdRow {
  const int *rp, *ci;
  const VT *v;
  const YT *y;
  YT *z;
  __device__ __forceinline__ void operator()(int i) const {
    const int s = rp[i], e = rp[i + 1];
    YT acc = y[i];
    for (int k = s + 1; k < e; ++k) acc = fma_t(-(YT)v[k], z[ci[k]], acc);
    z[i] = acc / (YT)v[s];
  }
};

// IC(0), one row, left-looking (ic0.cpp:57-115).  Rows j < i of the pattern
// must be final.  Failures keep the smallest failing row (the first one the
// sequential reference meets): key = row << 32 | pivot row.
struct IC0Row {
  const int *rp, *ci;
  double *v;
  double shift;
  unsigned long long *err;
  __device__ void operator()(int i) const {
    const int si = rp[i], ei = rp[i + 1];
    const int di = ei - 1;
    for (int kij = si; kij < di; ++kij) {
      const int j = ci[kij];
      double s = 0.0;
      int pi = si, pj = rp[j];
      const int ej = rp[j + 1];
      while (pi < ei && pj < ej) {
        const int c1 = ci[pi], c2 = ci[pj];
        if (c1 >= j || c2 >= j) break;
        if (c1 == c2) {
          s = fma(v[pi], v[pj], s);
          ++pi;
          ++pj;
        } else if (c1 < c2) {
          ++pi;
        } else {
          ++pj;
        }
      }
      const double Ljj = v[rp[j + 1] - 1];
      if (fabs(Ljj) < 1e-30) {
        atomicMin(err, ((unsigned long long)(unsigned)i << 32) | (unsigned)j);
        return;
      }
      v[kij] = (v[kij] - s) / Ljj;
    }
    double s = 0.0;
    for (int k = si; k < di; ++k) s = fma(v[k], v[k], s);
    double diag = v[di] - s;
    diag += shift;
    if (diag <= 0.0) {
      atomicMin(err, ((unsigned long long)(unsigned)i << 32) | (unsigned)i);
      return;
    }
    v[di] = sqrt(diag);
  }
};

template <class F>
__global__ __launch_bounds__(kWG) void k_range(int r0, int r1, F f, const Scal *S) {
  if (S && S->done) return;
  for (int i = r0 + blockIdx.x * kWG + threadIdx.x; i < r1; i += gridDim.x * kWG) f(i);
}

template <class F>
__global__ __launch_bounds__(kWG) void k_list(const int *__restrict__ rows, int cnt, F f,
                                              const Scal *S) {
  if (S && S->done) return;
  for (int k = blockIdx.x * kWG + threadIdx.x; k < cnt; k += gridDim.x * kWG) f(rows[k]);
}

// levels [l0, l1) inside one workgroup: a barrier between levels makes the
// previous level's global writes visible to every wave of the workgroup.
template <class F>
__global__ __launch_bounds__(1024) void k_levels_1wg(const int *__restrict__ lvl_ptr,
                                                     const int *__restrict__ lvl_rows, int l0,
                                                     int l1, F f, const Scal *S) {
  if (S && S->done) return;
  for (int lv = l0; lv < l1; ++lv) {
    const int s = lvl_ptr[lv], e = lvl_ptr[lv + 1];
    for (int k = s + threadIdx.x; k < e; k += blockDim.x) f(lvl_rows[k]);
    __syncthreads();
  }
}

// ABMC, general blocks: one workgroup per block; rows of a block follow
// their in-block dependency levels.
template <class F>
__global__ void k_abmc(const int *__restrict__ bstart, const int *__restrict__ brows,
                       const uint16_t *__restrict__ lev, const int *__restrict__ nlev, int kb0,
                       F f, const Scal *S) {
  if (S && S->done) return;
  const int kb = kb0 + blockIdx.x;
  const int s = bstart[kb], e = bstart[kb + 1];
  const int L = nlev[kb];
  for (int lv = 0; lv < L; ++lv) {
    for (int t = s + threadIdx.x; t < e; t += blockDim.x) {
      const int i = brows ? brows[t] : t;
      if ((int)lev[i] == lv) f(i);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// ABMC sweeps for blocks of 65 to 2048 rows (breadth-first blockings, the
// reference's block sizes up to 2047) on the plain sliced factor: one
// workgroup per block, thread t owning rows bs + t + q T (q < RPL).  Each
// row's entries, diagonal and the terms that do not depend on the block's
// own results are loaded once, before the in-block levels run: the forward
// sweep (ic0.cpp:263-300) folds the earlier colors' terms (the columns below
// the block, first in ascending order) into acc = r_i; the backward sweep
// (ic0.cpp:302-339) gathers the later colors' values (the columns past the
// block, last in ascending order).  Level l then applies, for the rows of
// in-block level l, the in-block terms from the block's LDS window (rows of
// lower levels, final), and in the backward sweep the later colors' terms,
// divides by L_ii and writes the window; a barrier separates the levels.
// ---------------------------------------------------------------------------
template <class VT, class YT, int WM, int RPL, bool BWD, bool NT>
__global__ __launch_bounds__(1024) void k_abmc_wg(int kb0, const int *__restrict__ bstart,
                                                  const uint16_t *__restrict__ lev,
                                                  const int *__restrict__ nlev, Sell L,
                                                  const VT *__restrict__ lv,
                                                  const VT *__restrict__ diag,
                                                  const double *__restrict__ r,
                                                  const YT *__restrict__ yin,
                                                  YT *__restrict__ out, const Scal *S) {
  extern __shared__ double wg_win[];  // the block's rows, final once written
  YT *win = reinterpret_cast<YT *>(wg_win);
  if (S && S->done) return;
  const int kb = kb0 + blockIdx.x;
  const int bs = bstart[kb], be = bstart[kb + 1];
  const int nl = nlev[kb];
  const int T = blockDim.x;
  int row[RPL], rl[RPL];
  YT acc[RPL], dd[RPL];
  int cc[RPL][WM];
  VT vv[RPL][WM];
  YT xo[RPL][WM];
#pragma unroll
  for (int q = 0; q < RPL; ++q) {
    const int i = bs + (int)threadIdx.x + q * T;
    row[q] = i < be ? i : -1;
    const int ii = i < be ? i : bs;  // lanes past the block load its first row
    const int o = L.off[ii >> 6];
    const int w = (L.off[(ii >> 6) + 1] - o) >> 6;
    const int base = o + (ii & 63);
#pragma unroll
    for (int k = 0; k < WM; ++k) {
      cc[q][k] = -1;
      vv[q][k] = (VT)0;
      if (k < w) {
        cc[q][k] = ldm<NT>(&L.col[base + 64 * k]);
        vv[q][k] = ldm<NT>(&lv[base + 64 * k]);
      }
    }
    dd[q] = (YT)diag[ii];
    rl[q] = (int)lev[ii];
    acc[q] = BWD ? yin[ii] : (YT)r[ii];
#pragma unroll
    for (int k = 0; k < WM; ++k) {
      const int c = cc[q][k];
      const bool outside = BWD ? c >= be : (c >= 0 && c < bs);
      const YT x = out[outside ? c : ii];
      if (BWD) {
        xo[q][k] = x;
      } else {
        const YT t = fma_t(-(YT)vv[q][k], x, acc[q]);
        acc[q] = outside ? t : acc[q];
      }
    }
  }
  for (int l = 0; l < nl; ++l) {
#pragma unroll
    for (int q = 0; q < RPL; ++q) {
      if (row[q] >= 0 && rl[q] == l) {
        YT a = acc[q];
#pragma unroll
        for (int k = 0; k < WM; ++k) {
          const int c = cc[q][k];
          if (c >= bs && c < be) a = fma_t(-(YT)vv[q][k], win[c - bs], a);
        }
        if (BWD) {
#pragma unroll
          for (int k = 0; k < WM; ++k)
            if (cc[q][k] >= be) a = fma_t(-(YT)vv[q][k], xo[q][k], a);
        }
        const YT y = a / dd[q];
        win[row[q] - bs] = y;
        out[row[q]] = y;
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Sliced-ELL multicolor sweeps (ic0.cpp:194-249): rows [c0, c1) of one
// color, one wavefront per 64-row slice.  Every dependency of a row lies in
// an earlier (forward) / later (backward) color, finished by a previous
// launch.
// ---------------------------------------------------------------------------
// RU (IC(0) iterations): the launch also applies the CG's r update to its
// color's rows, r_i -= alpha (Ap)_i (cg_crs.cpp:141, as k_r_ic), starts them
// from the new r_i and leaves the |r|^2 partials at part[gidx0 + workgroup].
template <class VT, class YT, int WM, bool NT, bool RU>
__global__ __launch_bounds__(kWG) void k_mc_fwd(int c0, int c1, Sell L, const VT *__restrict__ lv,
                                                const VT *__restrict__ diag, double *r,
                                                YT *__restrict__ y, const Scal *S,
                                                const double *__restrict__ Ap, double *part,
                                                int gidx0) {
  if (S && S->done) return;
  const double alpha = RU ? S->alpha : 0.0;
  double h[1] = {0}, l[1] = {0};
  const int lane = threadIdx.x & 63;
  const WaveRange wr = xcd_waves(c0 >> 6, ((c1 - 1) >> 6) + 1);
  // wave-uniform slice loop: the slice offsets load through the scalar unit
  // and the slice width is a scalar branch
  const int first = __builtin_amdgcn_readfirstlane(wr.first);
  const int step = __builtin_amdgcn_readfirstlane(wr.step);
  const int end = __builtin_amdgcn_readfirstlane(wr.end);
  for (int s = first; s < end; s += step) {
    const int i = (s << 6) + lane;
    if (i < c0 || i >= c1) continue;
    double ri = r[i];
    if (RU) {
      ri = fma(-alpha, ldm<true>(&Ap[i]), ri);  // Ap: read once
      r[i] = ri;
      dd_fma(h[0], l[0], ri, ri);
    }
    YT acc = (YT)ri;
    acc = sell_chain_u<VT, YT, YT, WM, true, NT>(L.off, L.col, lv, y, s, lane, acc);
    y[i] = acc / (YT)diag[i];
  }
  if (RU) block_partial<1>(h, l, part, gidx0 + blockIdx.x);
}

template <class VT, class YT, int WM, bool RZ, bool NT>
__global__ __launch_bounds__(kWG) void k_mc_bwd(int c0, int c1, Sell U, const VT *__restrict__ uv,
                                                const VT *__restrict__ diag,
                                                const YT *__restrict__ y, YT *__restrict__ z,
                                                const double *__restrict__ r, Scal *S,
                                                double *part, int gidx0, int total, int init) {
  if (S && S->done) return;
  const int lane = threadIdx.x & 63;
  const WaveRange wr = xcd_waves(c0 >> 6, ((c1 - 1) >> 6) + 1);
  double h[1] = {0}, l[1] = {0};
  const int first = __builtin_amdgcn_readfirstlane(wr.first);
  const int step = __builtin_amdgcn_readfirstlane(wr.step);
  const int end = __builtin_amdgcn_readfirstlane(wr.end);
  for (int s = first; s < end; s += step) {
    const int i = (s << 6) + lane;
    if (i < c0 || i >= c1) continue;
    YT acc = y[i];
    acc = sell_chain_u<VT, YT, YT, WM, true, NT>(U.off, U.col, uv, z, s, lane, acc);
    const YT zi = acc / (YT)diag[i];
    z[i] = zi;
    if (RZ) dd_fma(h[0], l[0], r[i], (double)zi);
  }
  if (RZ) block_partial<1>(h, l, part, gidx0 + blockIdx.x);
}

// ---------------------------------------------------------------------------
// ABMC sweeps for blocks of at most 64 rows: one wavefront per block
// (abmc_fwd_block / abmc_bwd_block), blocks of a color spread over the grid.
// ---------------------------------------------------------------------------
template <class VT, class YT, int WM, bool NT>
__global__ __launch_bounds__(kWG) void k_abmc_fwd_wave(int kb0, int kb1,
                                                       const int *__restrict__ bstart,
                                                       const uint16_t *__restrict__ lev,
                                                       const int *__restrict__ nlev, Sell L,
                                                       const VT *__restrict__ lv,
                                                       const VT *__restrict__ diag,
                                                       const double *__restrict__ r,
                                                       YT *__restrict__ y, const Scal *S) {
  __shared__ YT ys[kWG];
  if (S && S->done) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  YT *my = ys + (w << 6);
  const WaveRange wr = xcd_waves(kb0, kb1);
  for (int kb = wr.first; kb < wr.end; kb += wr.step)
    abmc_fwd_block<VT, YT, WM, NT>(bstart[kb], bstart[kb + 1], nlev[kb], lev, L.off, L.col, lv, diag,
                               r, y, my, lane);
}

template <class VT, class YT, int WM, bool RZ, bool NT>
__global__ __launch_bounds__(kWG) void k_abmc_bwd_wave(
    int kb0, int kb1, const int *__restrict__ bstart, const uint16_t *__restrict__ lev,
    const int *__restrict__ nlev, Sell U, const VT *__restrict__ uv, const VT *__restrict__ diag,
    const YT *__restrict__ y, YT *__restrict__ z, const double *__restrict__ r, Scal *S,
    double *part, int gidx0, int total, int init) {
  __shared__ YT zs[kWG];
  if (S && S->done) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  YT *my = zs + (w << 6);
  double h[1] = {0}, l[1] = {0};
  const WaveRange wr = xcd_waves(kb0, kb1);
  for (int kb = wr.first; kb < wr.end; kb += wr.step) {
    const int bs = bstart[kb], be = bstart[kb + 1];
    double ri;
    const YT zi = abmc_bwd_block<VT, YT, WM, NT>(bs, be, nlev[kb], lev, U.off, U.col, uv, diag, y, z,
                                             RZ ? r : nullptr, ri, my, lane);
    if (RZ && bs + lane < be) dd_fma(h[0], l[0], ri, (double)zi);
  }
  if (RZ) block_partial<1>(h, l, part, gidx0 + blockIdx.x);
}

// ABMC sweeps over the sectioned factor layout (one slice per block): a
// resident grid whose wavefronts loop over blocks, the loop wave-uniform so
// the block bounds and slice headers come from scalar loads.  With PIPE
// (FP32 sweeps) a wavefront issues the next block's loads (sec_fetch, no
// branches) before running the current block's levels, so they arrive
// while the level loop runs; FP64 sweeps fetch each block when they reach
// it (measured at 192^3: FP32 forward / backward 88.7 / 85.3 -> 85.5 /
// 78.5 us pipelined; FP64 121 / 118 us not pipelined vs 132 / 131
// pipelined, where the second block's registers cost occupancy).
//
// With FB (the last color) each block's backward rows follow its forward
// rows at once: the last color has no later colors, so its backward rows
// have in-block terms only, start from the y just computed (kept in
// registers, never stored) and the same diagonal; U's in-block entries are
// loaded while the forward levels run.
struct SecU {  // U's sectioned copy and the backward level counts (FB)
  const int *nlev, *off;
  const unsigned *inb;
  const void *v;
};

//
// With RU (IC(0) iterations) the sweep also applies the CG's r update to its
// rows: r_i -= alpha (Ap)_i (cg_crs.cpp:141, the same fma as k_r_ic), starts
// the row from the new r_i, stores it, and leaves the |r|^2 partials of the
// deferred stop test at part[gidx0 + workgroup] (one launch per color, so
// every row once).  With FB as well, the last color's z_i is final here, and
// the r.z partials of its rows (cg_crs.cpp:152, r_i z_i as k_rz forms it)
// go to part_rz[workgroup].
template <class VT, class YT, int SW, int SO, bool FB, bool NT, bool RU>
__global__ __launch_bounds__(kWG) void k_abmc_fwd_sec(int kb0, int kb1,
                                                      const int *__restrict__ bstart,
                                                      const int *__restrict__ nlev,
                                                      const int *__restrict__ off,
                                                      const unsigned *__restrict__ inb,
                                                      const int *__restrict__ col,
                                                      const VT *__restrict__ lv,
                                                      const VT *__restrict__ diag, double *r,
                                                      YT *__restrict__ y, SecU U,
                                                      YT *__restrict__ z, const Scal *S,
                                                      const double *__restrict__ Ap,
                                                      double *part, int gidx0,
                                                      double *part_rz) {
  constexpr bool PIPE = sizeof(YT) == 4;
  __shared__ YT win[(kWG / 64) * kSecWindow];
  if (S && S->done) return;
  const int lane = threadIdx.x & 63;
  YT *my = win + (threadIdx.x >> 6) * kSecWindow;
  const WaveRange wr = xcd_waves(kb0, kb1);
  const int first = __builtin_amdgcn_readfirstlane(wr.first);
  const int step = __builtin_amdgcn_readfirstlane(wr.step);
  const int end = __builtin_amdgcn_readfirstlane(wr.end);
  if (!RU && first >= end) return;  // RU: every wavefront reaches the partial
  const double alpha = RU ? S->alpha : 0.0;
  double h[1] = {0}, l[1] = {0}, hz[1] = {0}, lz[1] = {0};
  YT none[SO];
#pragma unroll
  for (int k = 0; k < SO; ++k) none[k] = (YT)0;
  const int f0 = min(first, kb1 - 1);
  int bs = bstart[f0], be = bstart[f0 + 1], nl = nlev[f0];
  int o0 = off[f0], w = (off[f0 + 1] - o0) >> 6;
  SecRaw<VT, YT, SW, SO> cur;
  if (PIPE)
    sec_fetch<VT, YT, SW, SO, true, NT, false, RU>(f0, o0, w, bs, be, lane, inb, col, lv, diag, r, y,
                                               cur, Ap);
  for (int kb = first; kb < end; kb += step) {
    if (!PIPE) {
      bs = bstart[kb];
      be = bstart[kb + 1];
      nl = nlev[kb];
      o0 = off[kb];
      w = (off[kb + 1] - o0) >> 6;
      sec_fetch<VT, YT, SW, SO, true, NT, true, RU>(kb, o0, w, bs, be, lane, inb, col, lv, diag, r, y,
                                                cur, Ap);
    }
    const bool act = bs + lane < be;
    int ad[SW], co[SO];
    YT va[SW], vo[SO], d, acc, xo[SO];
    sec_take<VT, YT, SW, SO>(cur, w, ad, va, co, vo, d, acc);
    double rn = 0.0;
    if (RU) {
      rn = fma(-alpha, cur.ap, cur.a);
      acc = (YT)rn;
      if (act) {
        r[bs + lane] = rn;
        dd_fma(h[0], l[0], rn, rn);
      }
    }
#pragma unroll
    for (int k = 0; k < SO; ++k)
      xo[k] = PIPE ? y[co[k] >= 0 ? co[k] : 0] : co[k] >= 0 ? y[co[k]] : (YT)0;
    const int kc = min(kb + step, end - 1);
    const int nbs = PIPE ? bstart[kc] : 0, nbe = PIPE ? bstart[kc + 1] : 0;
    const int nnl = PIPE ? nlev[kc] : 0, no0 = PIPE ? off[kc] : 0;
    const int nw = PIPE ? (off[kc + 1] - no0) >> 6 : 0;
    if (PIPE)
      sec_fetch<VT, YT, SW, SO, true, NT, false, RU>(kc, no0, nw, nbs, nbe, lane, inb, col, lv, diag,
                                                 r, y, cur, Ap);
    int uad[SW];
    YT uva[SW];
    if (FB) {  // this block's U entries, in flight during the forward levels
      const int t = min(lane, be - bs - 1), uo = U.off[kb] + t;
      const unsigned pk = ldm<NT>(&U.inb[64 * kb + t]);
#pragma unroll
      for (int k = 0; k < SW; ++k) {
        uad[k] = (int)((pk >> (8 * k)) & 0xFFu);
        // absent entries hold 0
        uva[k] = (YT)ldm<NT>(&((const VT *)U.v)[uo + 64 * k]);
      }
    }
    // the earlier colors' terms, ascending; absent ones are fma(-0, 0, acc)
#pragma unroll
    for (int k = 0; k < SO; ++k) acc = fma_t(-vo[k], co[k] >= 0 ? xo[k] : (YT)0, acc);
    if (lane == 0) my[64] = (YT)0;
    wave_sync();
    const YT yi = sec_levels<YT, SW, SO, false>(nl, acc, d, ad, va, none, none, my, lane, act);
    if (FB) {
      const YT zi = sec_levels<YT, SW, SO, false>(U.nlev[kb], yi, d, uad, uva, none, none, my, lane,
                                              act);
      if (act) z[bs + lane] = zi;
      if (RU && act) dd_fma(hz[0], lz[0], rn, (double)zi);
    } else if (act) {
      y[bs + lane] = yi;
    }
    if (PIPE) {
      bs = nbs;
      be = nbe;
      nl = nnl;
      o0 = no0;
      w = nw;
    }
  }
  if (RU) block_partial<1>(h, l, part, gidx0 + blockIdx.x);
  if (RU && FB) block_partial<1>(hz, lz, part_rz, blockIdx.x);
}

template <class VT, class YT, int SW, int SO, bool RZ, bool NT>
__global__ __launch_bounds__(kWG) void k_abmc_bwd_sec(
    int kb0, int kb1, const int *__restrict__ bstart, const int *__restrict__ nlev,
    const int *__restrict__ off, const unsigned *__restrict__ inb, const int *__restrict__ col,
    const VT *__restrict__ uv, const VT *__restrict__ diag, const YT *__restrict__ y,
    YT *__restrict__ z, const double *__restrict__ r, Scal *S, double *part, int gidx0) {
  constexpr bool PIPE = sizeof(YT) == 4;
  __shared__ YT win[(kWG / 64) * kSecWindow];
  if (S && S->done) return;
  const int lane = threadIdx.x & 63;
  YT *my = win + (threadIdx.x >> 6) * kSecWindow;
  double h[1] = {0}, l[1] = {0};
  const WaveRange wr = xcd_waves(kb0, kb1);
  const int first = __builtin_amdgcn_readfirstlane(wr.first);
  const int step = __builtin_amdgcn_readfirstlane(wr.step);
  const int end = __builtin_amdgcn_readfirstlane(wr.end);
  const double *rr = RZ ? r : nullptr;
  const int f0 = min(first, kb1 - 1);
  int bs = bstart[f0], be = bstart[f0 + 1], nl = nlev[f0];
  int o0 = off[f0], w = (off[f0 + 1] - o0) >> 6;
  SecRaw<VT, YT, SW, SO> cur;
  if (PIPE) sec_fetch<VT, YT, SW, SO, false, NT>(f0, o0, w, bs, be, lane, inb, col, uv, diag, rr, y, cur);
  for (int kb = first; kb < end; kb += step) {
    if (!PIPE) {
      bs = bstart[kb];
      be = bstart[kb + 1];
      nl = nlev[kb];
      o0 = off[kb];
      w = (off[kb + 1] - o0) >> 6;
      sec_fetch<VT, YT, SW, SO, false, NT, true>(kb, o0, w, bs, be, lane, inb, col, uv, diag, rr, y,
                                         cur);
    }
    const bool act = bs + lane < be;
    int ad[SW], co[SO];
    YT va[SW], vo[SO], d, acc, xo[SO];
    sec_take<VT, YT, SW, SO>(cur, w, ad, va, co, vo, d, acc);
    const double ri = cur.ri;
#pragma unroll
    for (int k = 0; k < SO; ++k)
      xo[k] = PIPE ? z[co[k] >= 0 ? co[k] : 0] : co[k] >= 0 ? z[co[k]] : (YT)0;
    const int kc = min(kb + step, end - 1);
    const int nbs = PIPE ? bstart[kc] : 0, nbe = PIPE ? bstart[kc + 1] : 0;
    const int nnl = PIPE ? nlev[kc] : 0, no0 = PIPE ? off[kc] : 0;
    const int nw = PIPE ? (off[kc + 1] - no0) >> 6 : 0;
    if (PIPE)
      sec_fetch<VT, YT, SW, SO, false, NT>(kc, no0, nw, nbs, nbe, lane, inb, col, uv, diag, rr, y, cur);
#pragma unroll
    for (int k = 0; k < SO; ++k) xo[k] = co[k] >= 0 ? xo[k] : (YT)0;
    if (lane == 0) my[64] = (YT)0;
    wave_sync();
    const YT zi = sec_levels<YT, SW, SO, true>(nl, acc, d, ad, va, vo, xo, my, lane, act);
    if (act) z[bs + lane] = zi;
    if (RZ && act) dd_fma(h[0], l[0], ri, (double)zi);
    if (PIPE) {
      bs = nbs;
      be = nbe;
      nl = nnl;
      o0 = no0;
      w = nw;
    }
  }
  if (RZ) block_partial<1>(h, l, part, gidx0 + blockIdx.x);
}

// ---------------------------------------------------------------------------
// small elementwise kernels
// ---------------------------------------------------------------------------
__global__ void k_scatter_u(long nnz, const int *__restrict__ l2u, const double *__restrict__ lv,
                            double *__restrict__ uv, float *__restrict__ lv32,
                            float *__restrict__ uv32) {
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < nnz;
       k += (long)gridDim.x * blockDim.x) {
    const double a = lv[k];
    const int u = l2u[k];
    uv[u] = a;
    if (lv32) {
      lv32[k] = (float)a;
      uv32[u] = (float)a;
    }
  }
}

// factor values -> sliced-ELL copies of L, U and the diagonal
__global__ void k_scatter_sell(long nnz, const double *__restrict__ lv,
                               const int *__restrict__ l2ls, const int *__restrict__ l2us,
                               double *__restrict__ ls64, float *__restrict__ ls32,
                               double *__restrict__ us64, float *__restrict__ us32,
                               double *__restrict__ d64, float *__restrict__ d32) {
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < nnz;
       k += (long)gridDim.x * blockDim.x) {
    const double a = lv[k];
    const int m = l2ls[k];
    if (m < 0) {
      const int row = -1 - m;
      if (d64) d64[row] = a;
      if (d32) d32[row] = (float)a;
    } else {
      if (ls64) ls64[m] = a;
      if (ls32) ls32[m] = (float)a;
    }
    const int u = l2us[k];
    if (u >= 0) {
      if (us64) us64[u] = a;
      if (us32) us32[u] = (float)a;
    }
  }
}

__global__ void k_cast_f32(long n, const double *__restrict__ a, float *__restrict__ b) {
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n;
       k += (long)gridDim.x * blockDim.x)
    b[k] = (float)a[k];
}
__global__ void k_cast_f64(long n, const float *__restrict__ a, double *__restrict__ b) {
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n;
       k += (long)gridDim.x * blockDim.x)
    b[k] = (double)a[k];
}
__global__ void k_copy(long n, const double *__restrict__ a, double *__restrict__ b) {
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n;
       k += (long)gridDim.x * blockDim.x)
    b[k] = a[k];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static int wm_of(int maxw) { return maxw <= 4 ? 4 : maxw <= 8 ? 8 : maxw <= 16 ? 16 : 0; }

static int ax_grid(const DevCsr &A) {
  return A.sell.off ? slice_grid(A.sell.nslice) : red_grid((A.n + kWG - 1) / kWG);
}

// KERNEL<row width class, 16-bit tables of per-slice lengths> for A
#define DISPATCH_WM(A, KERNEL, ...)                                              \
  do {                                                                           \
    const bool _tv = (A).sell.col16 && (A).sell.toff;                            \
    switch (spmv_wm(A)) {                                                        \
      case 4:                                                                    \
        if (_tv) hipLaunchKernelGGL((KERNEL<4, true>), __VA_ARGS__);             \
        else hipLaunchKernelGGL((KERNEL<4, false>), __VA_ARGS__);                \
        break;                                                                   \
      case 8:                                                                    \
        if (_tv) hipLaunchKernelGGL((KERNEL<8, true>), __VA_ARGS__);             \
        else hipLaunchKernelGGL((KERNEL<8, false>), __VA_ARGS__);                \
        break;                                                                   \
      case 16:                                                                   \
        if (_tv) hipLaunchKernelGGL((KERNEL<16, true>), __VA_ARGS__);            \
        else hipLaunchKernelGGL((KERNEL<16, false>), __VA_ARGS__);               \
        break;                                                                   \
      default: hipLaunchKernelGGL((KERNEL<0, false>), __VA_ARGS__); break;       \
    }                                                                            \
  } while (0)

void launch_spmv_plain(const DevCsr &A, const double *x, double *y, hipStream_t st) {
  if (A.n == 0) return;
  DISPATCH_WM(A, k_spmv_plain, dim3(ax_grid(A)), dim3(kWG), 0, st, A, x, y);
}

void launch_spmv_cg(const DevCsr &A, const double *p, double *Ap, Scal *S, double *part,
                    hipStream_t st, bool finish) {
  const int g = ax_grid(A);
  DISPATCH_WM(A, k_spmv_cg, dim3(g), dim3(kWG), 0, st, A, p, Ap, S, part);
  if (finish) launch_fin_alpha(A, S, part, st);
}

void launch_fin_alpha(const DevCsr &A, Scal *S, double *part, hipStream_t st) {
  hipLaunchKernelGGL(k_fin_alpha, dim3(1), dim3(kFinWG), 0, st, part, ax_grid(A), S);
}

void launch_resid(const DevCsr &A, const double *x, const double *b, double *r, Scal *S,
                  double *part, int mode, double tol, int max_iter, hipStream_t st) {
  const int g = ax_grid(A);
  DISPATCH_WM(A, k_resid, dim3(g), dim3(kWG), 0, st, A, x, b, r, S, part, mode,
              tol, max_iter);
  hipLaunchKernelGGL(k_fin_resid, dim3(1), dim3(kFinWG), 0, st, part, g, S, mode, tol, max_iter);
}

void launch_xr(int pk, int n, double *x, const double *p, double *r, const double *Ap,
               const double *inv_diag, double *z, Scal *S, double *part, hipStream_t st) {
  dim3 g(quad_grid(n)), b(kWG);
  if (pk == PK_IDENT) {
    hipLaunchKernelGGL(k_xr<PK_IDENT>, g, b, 0, st, n, x, p, r, Ap, inv_diag, z, S, part);
    hipLaunchKernelGGL(k_fin_xr<PK_IDENT>, dim3(1), dim3(kFinWG), 0, st, part, (int)g.x, S);
  } else if (pk == PK_JAC) {
    hipLaunchKernelGGL(k_xr<PK_JAC>, g, b, 0, st, n, x, p, r, Ap, inv_diag, z, S, part);
    hipLaunchKernelGGL(k_fin_xr<PK_JAC>, dim3(1), dim3(kFinWG), 0, st, part, (int)g.x, S);
  } else {
    hipLaunchKernelGGL(k_xr<PK_IC>, g, b, 0, st, n, x, p, r, Ap, inv_diag, z, S, part);
    hipLaunchKernelGGL(k_fin_xr<PK_IC>, dim3(1), dim3(kFinWG), 0, st, part, (int)g.x, S);
  }
}

// the |r|^2 partials of the deferred stop test (k_r_ic, at most
// kMaxRedBlocks workgroups): pairs from kMaxRedBlocks on, clear of k_rz's
// r.z pairs below kMaxRedBlocks
static double *xr_parts(double *part) { return part + 2 * kMaxRedBlocks; }

void launch_r_ic(int n, double *r, const double *Ap, Scal *S, double *part, hipStream_t st) {
  hipLaunchKernelGGL(k_r_ic, dim3(quad_grid(n)), dim3(kWG), 0, st, n, r, Ap, S, xr_parts(part));
}

void launch_update_xp(int n, double *x, const void *z, bool z32, double *p, const Scal *S,
                      hipStream_t st) {
  dim3 g(quad_grid(n)), b(kWG);
  if (z32)
    hipLaunchKernelGGL(k_update_xp<float>, g, b, 0, st, n, x, (const float *)z, p, S);
  else
    hipLaunchKernelGGL(k_update_xp<double>, g, b, 0, st, n, x, (const double *)z, p, S);
}

void launch_rz_stop(int n, const double *r, const void *z, bool z32, Scal *S, double *part,
                    hipStream_t st, int xr_count) {
  dim3 g(quad_grid(n)), b(kWG);
  if (z32)
    hipLaunchKernelGGL(k_rz<float>, g, b, 0, st, n, r, (const float *)z, S, part, 0);
  else
    hipLaunchKernelGGL(k_rz<double>, g, b, 0, st, n, r, (const double *)z, S, part, 0);
  hipLaunchKernelGGL(k_fin_xr_rz, dim3(1), dim3(kFinWG), 0, st, xr_parts(part),
                     xr_count > 0 ? xr_count : (int)g.x, part, (int)g.x, S);
}

void launch_precond_diag_init(int pk, int n, const double *r, const double *inv_diag, double *z,
                              double *p, Scal *S, double *part, hipStream_t st) {
  dim3 g(red_grid((n + kWG - 1) / kWG)), b(kWG);
  if (pk == PK_JAC)
    hipLaunchKernelGGL(k_precond_diag_init<PK_JAC>, g, b, 0, st, n, r, inv_diag, z, p, S, part);
  else
    hipLaunchKernelGGL(k_precond_diag_init<PK_IDENT>, g, b, 0, st, n, r, inv_diag, z, p, S,
                       part);
  hipLaunchKernelGGL(k_fin_rz, dim3(1), dim3(kFinWG), 0, st, part, (int)g.x, S, 1);
}

void launch_rz(int n, const double *r, const void *z, bool z32, Scal *S, double *part, int init,
               hipStream_t st) {
  dim3 g(quad_grid(n)), b(kWG);
  if (z32)
    hipLaunchKernelGGL(k_rz<float>, g, b, 0, st, n, r, (const float *)z, S, part, init);
  else
    hipLaunchKernelGGL(k_rz<double>, g, b, 0, st, n, r, (const double *)z, S, part, init);
  hipLaunchKernelGGL(k_fin_rz, dim3(1), dim3(kFinWG), 0, st, part, (int)g.x, S, init);
}

void launch_update_p(int n, const void *z, bool z32, double *p, const Scal *S, int init,
                     hipStream_t st) {
  dim3 g(quad_grid(n)), b(kWG);
  if (z32)
    hipLaunchKernelGGL(k_update_p<float>, g, b, 0, st, n, (const float *)z, p, S, init);
  else
    hipLaunchKernelGGL(k_update_p<double>, g, b, 0, st, n, (const double *)z, p, S, init);
}

void launch_jacobi_setup(const DevCsr &A, double *inv_diag, hipStream_t st) {
  if (A.n == 0) return;
  hipLaunchKernelGGL(k_jacobi_setup, dim3(stream_grid(A.n)), dim3(kWG), 0, st, A.n, A.rp, A.ci,
                     A.v, inv_diag);
}

void launch_jacobi_apply(int n, const double *inv_diag, const double *r, double *z,
                         hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_jacobi_apply, dim3(stream_grid(n)), dim3(kWG), 0, st, n, inv_diag, r, z);
}

// CSR-functor plans: natural-order levels, non-contiguous MC, large ABMC blocks
template <class F>
static void run_csr_plan(const DevSched &sc, F f, const Scal *S, hipStream_t st, bool backward) {
  if (sc.kind == 1) {  // MC
    const std::vector<int> &cp = sc.h_color_ptr_v;
    for (int t = 0; t < sc.nc; ++t) {
      const int c = backward ? sc.nc - 1 - t : t;
      const int r0 = cp[c], r1 = cp[c + 1];
      if (r1 <= r0) continue;
      if (sc.d_rows)
        hipLaunchKernelGGL(k_list<F>, dim3(stream_grid(r1 - r0)), dim3(kWG), 0, st,
                           sc.d_rows + r0, r1 - r0, f, S);
      else
        hipLaunchKernelGGL(k_range<F>, dim3(stream_grid(r1 - r0)), dim3(kWG), 0, st, r0, r1,
                           f, S);
    }
  } else if (sc.kind == 2) {  // ABMC
    const std::vector<int> &cb = sc.h_color_blk_v;
    int wg = 64;
    while (wg < sc.max_block && wg < 256) wg *= 2;
    for (int t = 0; t < sc.nc; ++t) {
      const int c = backward ? sc.nc - 1 - t : t;
      const int b0 = cb[c], b1 = cb[c + 1];
      if (b1 <= b0) continue;
      hipLaunchKernelGGL(k_abmc<F>, dim3(b1 - b0), dim3(wg), 0, st, sc.d_bstart, sc.d_brows,
                         backward ? sc.d_lev_b : sc.d_lev_f,
                         backward ? sc.d_nlev_b : sc.d_nlev_f, b0, f, S);
    }
  } else {  // natural order, level sets
    const std::vector<int> &seg = backward ? sc.seg_b : sc.seg_f;
    const std::vector<int> &lp = backward ? sc.h_lvl_ptr_b : sc.h_lvl_ptr_f;
    const int *dptr = backward ? sc.d_lvl_ptr_b : sc.d_lvl_ptr_f;
    const int *drows = backward ? sc.d_lvl_rows_b : sc.d_lvl_rows_f;
    for (size_t q = 0; q + 2 < seg.size(); q += 3) {
      if (seg[q] == 0) {
        hipLaunchKernelGGL(k_levels_1wg<F>, dim3(1), dim3(1024), 0, st, dptr, drows,
                           seg[q + 1], seg[q + 2], f, S);
      } else {
        const int l = seg[q + 1];
        const int cnt = lp[l + 1] - lp[l];
        hipLaunchKernelGGL(k_list<F>, dim3(stream_grid(cnt)), dim3(kWG), 0, st, drows + lp[l],
                           cnt, f, S);
      }
    }
  }
}

// sliced plans are used for MC with contiguous colors and ABMC with
// contiguous blocks of <= 64 rows, when the factor has sliced copies
bool sliced_plan(const DevSched &sc, const Factor &F) {
  if (!F.sell || std::max(F.ls.maxw, F.us.maxw) > 16) return false;
  if (sc.kind == 1) return sc.d_rows == nullptr;
  if (sc.kind == 2) return sc.d_brows == nullptr && sc.max_block <= 64;
  return false;
}

// r.z comes from a separate streaming kernel after the backward sweep.
// ICCG_FUSE_RZ=1 forms it inside the backward sweep instead (at 192^3 the
// plain-layout sweep lost ~23 us per color launch to it against 15 us for
// the extra pass; the sectioned sweep gains 0.8% per iteration).  The fused
// partials partition the rows differently, so the double-double r.z can
// round to another double when its exact value lies within ~2^-100
// (relative) of a rounding boundary; the separate pass is the default.
static bool fuse_rz_enabled() {
  static const bool on = [] {
    const char *e = std::getenv("ICCG_FUSE_RZ");
    return e && e[0] == '1';
  }();
  return on;
}

bool bwd_fuses_rz(const DevSched &sc, const Factor &F) {
  return sliced_plan(sc, F) && fuse_rz_enabled();
}

// workgroups of the forward sweep's launches on the sectioned plan: as many
// |r|^2 partials of the sweep's r update, and as many r.z partials
static int fwd_sec_groups(const DevSched &sc) {
  int total = 0;
  for (int c = 0; c < sc.nc; ++c) {
    const int b0 = sc.h_color_blk_v[c], b1 = sc.h_color_blk_v[c + 1];
    if (b1 > b0) total += slice_grid(b1 - b0);
  }
  return total;
}

// partial slots of the fused backward sweep (one per workgroup of every
// color launch)
int bwd_partial_slots(const DevSched &sc, const Factor &F) {
  if (!sliced_plan(sc, F)) return 0;
  int total = 0;
  for (int c = 0; c < sc.nc; ++c) {
    if (sc.kind == 1) {
      const int c0 = sc.h_color_ptr_v[c], c1 = sc.h_color_ptr_v[c + 1];
      if (c1 > c0) total += slice_grid(((c1 - 1) >> 6) - (c0 >> 6) + 1);
    } else {
      const int b0 = sc.h_color_blk_v[c], b1 = sc.h_color_blk_v[c + 1];
      if (b1 > b0) total += std::max((b1 - b0 + 3) / 4, slice_grid(b1 - b0));  // either grid
    }
  }
  if (sc.kind == 2 && F.sec > 0) total = std::max(total, 2 * fwd_sec_groups(sc));
  // multicolor: the backward launches' r.z pairs from slot 0, the forward
  // launches' |r|^2 pairs behind them (mc_xr_parts)
  if (sc.kind == 1) total += std::max(kMaxRedBlocks, total);
  return total;
}

static int sweep_wm(const Factor &F) { return wm_of(std::max(F.ls.maxw, F.us.maxw)); }

// Grid of an ABMC sweep over nb blocks: FP32 sweeps run best as a
// resident grid (8 waves per SIMD) looping over blocks, FP64 sweeps with one
// block per wavefront (measured at 192^3: FP32 backward 104 vs 114 us, FP64
// backward 149 vs 158 us).
template <class VT>
static int abmc_grid(int nb) {
  return sizeof(VT) == 8 ? (nb + 3) / 4 : slice_grid(nb);
}

// with Ap (multicolor plan): every launch also applies the r update (r is
// then written: the caller's r is the solver's own vector), the |r|^2
// partials to part[0, *ru_count)
template <class VT, class YT, int WM, bool NT>
static void fwd_sliced(const DevSched &sc, const Factor &F, const VT *lv, const VT *dg,
                       const double *r, YT *y, const Scal *S, hipStream_t st, const double *Ap,
                       double *part, int *ru_count) {
  if (sc.kind == 1) {
    double *rw = const_cast<double *>(r);
    int gidx0 = 0;
    for (int c = 0; c < sc.nc; ++c) {
      const int c0 = sc.h_color_ptr_v[c], c1 = sc.h_color_ptr_v[c + 1];
      if (c1 <= c0) continue;
      const int g = slice_grid(((c1 - 1) >> 6) - (c0 >> 6) + 1);
      if (Ap)
        hipLaunchKernelGGL((k_mc_fwd<VT, YT, WM, NT, true>), dim3(g), dim3(kWG), 0, st, c0, c1,
                           F.ls, lv, dg, rw, y, S, Ap, part, gidx0);
      else
        hipLaunchKernelGGL((k_mc_fwd<VT, YT, WM, NT, false>), dim3(g), dim3(kWG), 0, st, c0, c1,
                           F.ls, lv, dg, rw, y, S, nullptr, nullptr, 0);
      gidx0 += g;
    }
    if (ru_count) *ru_count = Ap ? gidx0 : 0;
  } else {
    if (ru_count) *ru_count = 0;
    for (int c = 0; c < sc.nc; ++c) {
      const int b0 = sc.h_color_blk_v[c], b1 = sc.h_color_blk_v[c + 1];
      if (b1 <= b0) continue;
      hipLaunchKernelGGL((k_abmc_fwd_wave<VT, YT, WM, NT>), dim3(abmc_grid<VT>(b1 - b0)), dim3(kWG), 0,
                         st, b0, b1, sc.d_bstart, sc.d_lev_f, sc.d_nlev_f, F.ls, lv, dg, r, y, S);
    }
  }
}

// rz_slot0 >= 0 (multicolor, after the forward sweep's r update): the r.z
// partials of every launch from part[rz_slot0] on, folded by the stop test
template <class VT, class YT, int WM, bool NT>
static void bwd_sliced(const DevSched &sc, const Factor &F, const VT *uv, const VT *dg,
                       const YT *y, YT *z, const double *r, Scal *S, double *part, int rz,
                       hipStream_t st, int rz_slot0 = -1) {
  // workgroup counts of every color launch, for the cross-launch r.z sum
  std::vector<int> grids;
  for (int t = 0; t < sc.nc; ++t) {
    const int c = sc.nc - 1 - t;
    int g = 0;
    if (sc.kind == 1) {
      const int c0 = sc.h_color_ptr_v[c], c1 = sc.h_color_ptr_v[c + 1];
      if (c1 > c0) g = slice_grid(((c1 - 1) >> 6) - (c0 >> 6) + 1);
    } else {
      const int b0 = sc.h_color_blk_v[c], b1 = sc.h_color_blk_v[c + 1];
      if (b1 > b0) g = abmc_grid<VT>(b1 - b0);
    }
    grids.push_back(g);
  }
  int total = 0;
  for (int g : grids) total += g;
  const bool fin = rz != 0 && r != nullptr && fuse_rz_enabled();
  const bool fuse = fin || (rz_slot0 >= 0 && sc.kind == 1);
  int gidx0 = fin ? 0 : rz_slot0;
  for (int t = 0; t < sc.nc; ++t) {
    const int c = sc.nc - 1 - t;
    const int g = grids[t];
    if (g == 0) continue;
    if (sc.kind == 1) {
      const int c0 = sc.h_color_ptr_v[c], c1 = sc.h_color_ptr_v[c + 1];
      if (fuse)
        hipLaunchKernelGGL((k_mc_bwd<VT, YT, WM, true, NT>), dim3(g), dim3(kWG), 0, st, c0, c1, F.us,
                           uv, dg, y, z, r, S, part, gidx0, total, rz == 2);
      else
        hipLaunchKernelGGL((k_mc_bwd<VT, YT, WM, false, NT>), dim3(g), dim3(kWG), 0, st, c0, c1,
                           F.us, uv, dg, y, z, r, S, part, gidx0, total, 0);
    } else {
      const int b0 = sc.h_color_blk_v[c], b1 = sc.h_color_blk_v[c + 1];
      if (fuse)
        hipLaunchKernelGGL((k_abmc_bwd_wave<VT, YT, WM, true, NT>), dim3(g), dim3(kWG), 0, st, b0,
                           b1, sc.d_bstart, sc.d_lev_b, sc.d_nlev_b, F.us, uv, dg, y, z, r, S,
                           part, gidx0, total, rz == 2);
      else
        hipLaunchKernelGGL((k_abmc_bwd_wave<VT, YT, WM, false, NT>), dim3(g), dim3(kWG), 0, st, b0,
                           b1, sc.d_bstart, sc.d_lev_b, sc.d_nlev_b, F.us, uv, dg, y, z, r, S,
                           part, gidx0, total, 0);
    }
    gidx0 += g;
  }
  if (fin) hipLaunchKernelGGL(k_fin_rz, dim3(1), dim3(kFinWG), 0, st, part, total, S, rz == 2);
}

// forward colors; with uv, the last color through the fused forward +
// backward launch; with Ap, every launch also applies the r update (r is
// then written: the caller's r is the solver's own vector), the |r|^2
// partials go to part[0, G) and the fused launch's r.z partials to
// part[G, ...) (G = fwd_sec_groups, *ru_count)
template <class VT, class YT, int SW, int SO, bool NT>
static void fwd_sec(const DevSched &sc, const Factor &F, const VT *lv, const VT *dg,
                    const double *r, YT *y, const Scal *S, hipStream_t st, const VT *uv,
                    YT *z, const double *Ap, double *part, int *ru_count) {
  const SecU none{nullptr, nullptr, nullptr, nullptr};
  const SecU us{sc.d_nlev_b, F.us.off, F.us.inb, (const void *)uv};
  double *rw = const_cast<double *>(r);
  double *prz = part ? part + 2 * (size_t)fwd_sec_groups(sc) : nullptr;
  int gidx0 = 0;
  for (int c = 0; c < sc.nc; ++c) {
    const int b0 = sc.h_color_blk_v[c], b1 = sc.h_color_blk_v[c + 1];
    if (b1 <= b0) continue;
    const int g = slice_grid(b1 - b0);
    const bool fb = uv && c == sc.nc - 1;
    if (Ap && fb)
      hipLaunchKernelGGL((k_abmc_fwd_sec<VT, YT, SW, SO, true, NT, true>), dim3(g), dim3(kWG), 0, st,
                         b0, b1, sc.d_bstart, sc.d_nlev_f, F.ls.off, F.ls.inb, F.ls.col, lv, dg,
                         rw, y, us, z, S, Ap, part, gidx0, prz);
    else if (Ap)
      hipLaunchKernelGGL((k_abmc_fwd_sec<VT, YT, SW, SO, false, NT, true>), dim3(g), dim3(kWG), 0, st,
                         b0, b1, sc.d_bstart, sc.d_nlev_f, F.ls.off, F.ls.inb, F.ls.col, lv, dg,
                         rw, y, none, nullptr, S, Ap, part, gidx0, nullptr);
    else if (fb)
      hipLaunchKernelGGL((k_abmc_fwd_sec<VT, YT, SW, SO, true, NT, false>), dim3(g), dim3(kWG), 0, st,
                         b0, b1, sc.d_bstart, sc.d_nlev_f, F.ls.off, F.ls.inb, F.ls.col, lv, dg,
                         rw, y, us, z, S, nullptr, nullptr, 0, nullptr);
    else
      hipLaunchKernelGGL((k_abmc_fwd_sec<VT, YT, SW, SO, false, NT, false>), dim3(g), dim3(kWG), 0,
                         st, b0, b1, sc.d_bstart, sc.d_nlev_f, F.ls.off, F.ls.inb, F.ls.col, lv,
                         dg, rw, y, none, nullptr, S, nullptr, nullptr, 0, nullptr);
    gidx0 += g;
  }
  if (ru_count) *ru_count = Ap ? gidx0 : 0;
}

// multicolor launches' workgroups, for the r update's |r|^2 partials
static int mc_fwd_groups(const DevSched &sc) {
  int total = 0;
  for (int c = 0; c < sc.nc; ++c) {
    const int c0 = sc.h_color_ptr_v[c], c1 = sc.h_color_ptr_v[c + 1];
    if (c1 > c0) total += slice_grid(((c1 - 1) >> 6) - (c0 >> 6) + 1);
  }
  return total;
}

// multicolor: the r.z partials of the backward launches go where k_rz leaves
// its own (pairs 0 .. g-1), the |r|^2 partials of the forward ones (as many:
// the same color ranges and grids) in the g pairs from max(kMaxRedBlocks, g)
// on, so that neither region overlaps the other for any g; the buffer holds
// that many pairs (bwd_partial_slots)
static double *mc_xr_parts(const DevSched &sc, double *part) {
  return part + 2 * (size_t)std::max(kMaxRedBlocks, mc_fwd_groups(sc));
}

static bool mc_sweeps_form_r(const DevSched &sc, const Factor &F) {
  return sliced_plan(sc, F) && sc.kind == 1 && !fuse_rz_enabled() &&
         mc_fwd_groups(sc) <= 2 * kMaxRedBlocks;
}

bool sweeps_form_rz(const DevSched &sc, const Factor &F) {
  return (sliced_plan(sc, F) && sc.kind == 2 && F.sec > 0 && !fuse_rz_enabled()) ||
         mc_sweeps_form_r(sc, F);
}

bool fwd_updates_r(const DevSched &sc, const Factor &F) { return sweeps_form_rz(sc, F); }

void launch_fin_sweep_stop(const DevSched &sc, Scal *S, double *part, hipStream_t st) {
  if (sc.kind == 1) {
    const int g = mc_fwd_groups(sc);
    hipLaunchKernelGGL(k_fin_xr_rz, dim3(1), dim3(kFinWG), 0, st, mc_xr_parts(sc, part), g, part,
                       g, S);
    return;
  }
  const int g = fwd_sec_groups(sc);
  hipLaunchKernelGGL(k_fin_xr_rz, dim3(1), dim3(kFinWG), 0, st, part, g, part + 2 * (size_t)g, g,
                     S);
}

template <class VT, class YT, int SW, int SO, bool NT>
static void bwd_sec(const DevSched &sc, const Factor &F, const VT *uv, const VT *dg, const YT *y,
                    YT *z, const double *r, Scal *S, double *part, int rz, hipStream_t st,
                    bool skip_last = false, int rz_slot0 = -1) {
  std::vector<int> grids;  // workgroups of every color launch, for the r.z partials
  int total = 0;
  for (int t = 0; t < sc.nc; ++t) {
    const int c = sc.nc - 1 - t;
    const int b0 = sc.h_color_blk_v[c], b1 = sc.h_color_blk_v[c + 1];
    grids.push_back(b1 > b0 && !(skip_last && t == 0) ? slice_grid(b1 - b0) : 0);
    total += grids.back();
  }
  // rz_slot0 >= 0: the r.z partials of the sweep's rows from that slot on,
  // folded later by the stop test's finish (the forward sweep's r update)
  const bool fuse = rz_slot0 >= 0 || (rz != 0 && r != nullptr && fuse_rz_enabled());
  int gidx0 = rz_slot0 >= 0 ? rz_slot0 : 0;
  for (int t = 0; t < sc.nc; ++t) {
    const int c = sc.nc - 1 - t;
    const int g = grids[t];
    if (g == 0) continue;
    const int b0 = sc.h_color_blk_v[c], b1 = sc.h_color_blk_v[c + 1];
    if (fuse)
      hipLaunchKernelGGL((k_abmc_bwd_sec<VT, YT, SW, SO, true, NT>), dim3(g), dim3(kWG), 0, st, b0, b1,
                         sc.d_bstart, sc.d_nlev_b, F.us.off, F.us.inb, F.us.col, uv, dg, y, z, r, S,
                         part,
                         gidx0);
    else
      hipLaunchKernelGGL((k_abmc_bwd_sec<VT, YT, SW, SO, false, NT>), dim3(g), dim3(kWG), 0, st, b0, b1,
                         sc.d_bstart, sc.d_nlev_b, F.us.off, F.us.inb, F.us.col, uv, dg, y, z, r, S,
                         part,
                         gidx0);
    gidx0 += g;
  }
  if (fuse && rz_slot0 < 0)
    hipLaunchKernelGGL(k_fin_rz, dim3(1), dim3(kFinWG), 0, st, part, total, S, rz == 2);
}

// FN<VT, YT, width, nt> over the section / slot widths and the factor's
// load policy (Factor::nt)
#define DISPATCH_SEC(sw, so, nt, FN, VT, YT, ...)                       \
  do {                                                                  \
    const bool _w = (so) > (sw);                                        \
    if ((sw) == 2 && _w && (nt)) FN<VT, YT, 2, 6, true>(__VA_ARGS__);   \
    else if ((sw) == 2 && _w) FN<VT, YT, 2, 6, false>(__VA_ARGS__);     \
    else if ((sw) == 2 && (nt)) FN<VT, YT, 2, 2, true>(__VA_ARGS__);    \
    else if ((sw) == 2) FN<VT, YT, 2, 2, false>(__VA_ARGS__);           \
    else if (_w && (nt)) FN<VT, YT, 3, 6, true>(__VA_ARGS__);           \
    else if (_w) FN<VT, YT, 3, 6, false>(__VA_ARGS__);                  \
    else if (nt) FN<VT, YT, 3, 3, true>(__VA_ARGS__);                   \
    else FN<VT, YT, 3, 3, false>(__VA_ARGS__);                          \
  } while (0)

#define DISPATCH_SWEEP_WM(wm, nt, FN, VT, YT, ...)               \
  do {                                                           \
    const int _w = (wm) == 4 ? 4 : (wm) == 8 ? 8 : 16;           \
    if (_w == 4 && (nt)) FN<VT, YT, 4, true>(__VA_ARGS__);       \
    else if (_w == 4) FN<VT, YT, 4, false>(__VA_ARGS__);         \
    else if (_w == 8 && (nt)) FN<VT, YT, 8, true>(__VA_ARGS__);  \
    else if (_w == 8) FN<VT, YT, 8, false>(__VA_ARGS__);         \
    else if (nt) FN<VT, YT, 16, true>(__VA_ARGS__);              \
    else FN<VT, YT, 16, false>(__VA_ARGS__);                     \
  } while (0)

// blocks of 65 to 2048 rows on the plain sliced factor (k_abmc_wg)
bool block_plan(const DevSched &sc, const Factor &F) {
  return F.sell && F.sec == 0 && std::max(F.ls.maxw, F.us.maxw) <= 16 && sc.kind == 2 &&
         sc.d_brows == nullptr && sc.max_block > 64 && sc.max_block <= kBlockPlanMax;
}

// one launch per color, forward in color order, backward in reverse
template <class VT, class YT, int WM, int RPL, bool BWD, bool NT>
static void wg_sweep(const DevSched &sc, const Sell &L, const VT *lv, const VT *dg,
                     const double *r, const YT *yin, YT *out, const Scal *S, hipStream_t st) {
  const int per = (sc.max_block + RPL - 1) / RPL;
  const int T = std::min(1024, (per + 63) / 64 * 64);
  const size_t lds = sizeof(YT) * (size_t)sc.max_block;
  for (int t = 0; t < sc.nc; ++t) {
    const int c = BWD ? sc.nc - 1 - t : t;
    const int b0 = sc.h_color_blk_v[c], b1 = sc.h_color_blk_v[c + 1];
    if (b1 <= b0) continue;
    hipLaunchKernelGGL((k_abmc_wg<VT, YT, WM, RPL, BWD, NT>), dim3(b1 - b0), dim3(T), lds, st,
                       b0, sc.d_bstart, BWD ? sc.d_lev_b : sc.d_lev_f,
                       BWD ? sc.d_nlev_b : sc.d_nlev_f, L, lv, dg, r, yin, out, S);
  }
}

template <class VT, class YT, bool BWD>
static void wg_dispatch(const DevSched &sc, const Factor &F, const Sell &L, const VT *lv,
                        const VT *dg, const double *r, const YT *yin, YT *out, const Scal *S,
                        hipStream_t st) {
  const int w = std::max(F.ls.maxw, F.us.maxw);
  const bool two = sc.max_block > 1024;
#define WG_ONE(WMV, RPLV)                                                                  \
  do {                                                                                     \
    if (F.nt)                                                                              \
      wg_sweep<VT, YT, WMV, RPLV, BWD, true>(sc, L, lv, dg, r, yin, out, S, st);           \
    else                                                                                   \
      wg_sweep<VT, YT, WMV, RPLV, BWD, false>(sc, L, lv, dg, r, yin, out, S, st);          \
  } while (0)
  if (w <= 4) {
    if (two) WG_ONE(4, 2); else WG_ONE(4, 1);
  } else if (w <= 8) {
    if (two) WG_ONE(8, 2); else WG_ONE(8, 1);
  } else {
    if (two) WG_ONE(16, 2); else WG_ONE(16, 1);
  }
#undef WG_ONE
}

bool sweeps_fused(const DevSched &sc, const Factor &F) {
  return sliced_plan(sc, F) && sc.kind == 2 && F.sec > 0 && F.fb && !fuse_rz_enabled();
}

void launch_fwd(const DevSched &sc, const Factor &F, bool mixed, const double *r, void *y,
                const Scal *S, hipStream_t st, void *z_fused, const double *Ap, double *part,
                int *ru_count) {
  if (ru_count) *ru_count = 0;
  if (sliced_plan(sc, F) && sc.kind == 2 && F.sec > 0) {
    const bool fb = z_fused && sweeps_fused(sc, F);
    const double *ap = Ap && fwd_updates_r(sc, F) ? Ap : nullptr;
    double *xp = ap ? part : nullptr;
    if (mixed)
      DISPATCH_SEC(F.sec, F.seco, F.nt, fwd_sec, float, float, sc, F, F.ls.v32, F.d32, r, (float *)y, S, st,
                   fb ? F.us.v32 : nullptr, (float *)z_fused, ap, xp, ru_count);
    else
      DISPATCH_SEC(F.sec, F.seco, F.nt, fwd_sec, double, double, sc, F, F.ls.v64, F.d64, r, (double *)y, S,
                   st, fb ? F.us.v64 : nullptr, (double *)z_fused, ap, xp, ru_count);
    return;
  }
  if (sliced_plan(sc, F)) {
    const int wm = sweep_wm(F);
    const double *ap = Ap && sc.kind == 1 && fwd_updates_r(sc, F) ? Ap : nullptr;
    double *xp = ap ? mc_xr_parts(sc, part) : nullptr;
    if (mixed)
      DISPATCH_SWEEP_WM(wm, F.nt, fwd_sliced, float, float, sc, F, F.ls.v32, F.d32, r, (float *)y, S,
                        st, ap, xp, ru_count);
    else
      DISPATCH_SWEEP_WM(wm, F.nt, fwd_sliced, double, double, sc, F, F.ls.v64, F.d64, r, (double *)y,
                        S, st, ap, xp, ru_count);
    return;
  }
  if (block_plan(sc, F)) {
    if (mixed)
      wg_dispatch<float, float, false>(sc, F, F.ls, F.ls.v32, F.d32, r, nullptr, (float *)y, S,
                                       st);
    else
      wg_dispatch<double, double, false>(sc, F, F.ls, F.ls.v64, F.d64, r, nullptr, (double *)y,
                                         S, st);
    return;
  }
  if (mixed) {
    FwdRow<float, float> f{F.lrp, F.lci, F.lv32, r, (float *)y};
    run_csr_plan(sc, f, S, st, false);
  } else {
    FwdRow<double, double> f{F.lrp, F.lci, F.lv, r, (double *)y};
    run_csr_plan(sc, f, S, st, false);
  }
}

void launch_bwd(const DevSched &sc, const Factor &F, bool mixed, const void *y, void *z,
                const Scal *S, hipStream_t st, const double *r, Scal *Sw, double *part, int rz,
                bool last_done, bool rz_after_ru) {
  if (sliced_plan(sc, F) && sc.kind == 2 && F.sec > 0) {
    const bool skip = last_done && sweeps_fused(sc, F);
    // after the forward sweep's r update: r.z partials behind its |r|^2
    // ones, after those of the fused last color
    int slot0 = -1;
    double *pz = part;
    if (rz_after_ru && fwd_updates_r(sc, F)) {
      pz = part + 2 * (size_t)fwd_sec_groups(sc);
      const int cl = sc.nc - 1;
      slot0 = skip ? slice_grid(sc.h_color_blk_v[cl + 1] - sc.h_color_blk_v[cl]) : 0;
    }
    if (mixed)
      DISPATCH_SEC(F.sec, F.seco, F.nt, bwd_sec, float, float, sc, F, F.us.v32, F.d32, (const float *)y,
                   (float *)z, r, Sw, pz, rz, st, skip, slot0);
    else
      DISPATCH_SEC(F.sec, F.seco, F.nt, bwd_sec, double, double, sc, F, F.us.v64, F.d64, (const double *)y,
                   (double *)z, r, Sw, pz, rz, st, skip, slot0);
    return;
  }
  if (sliced_plan(sc, F)) {
    const int wm = sweep_wm(F);
    // multicolor after the forward sweep's r update: r.z partials from part[0]
    const int slot0 = rz_after_ru && sc.kind == 1 && fwd_updates_r(sc, F) ? 0 : -1;
    if (mixed)
      DISPATCH_SWEEP_WM(wm, F.nt, bwd_sliced, float, float, sc, F, F.us.v32, F.d32, (const float *)y,
                        (float *)z, r, Sw, part, rz, st, slot0);
    else
      DISPATCH_SWEEP_WM(wm, F.nt, bwd_sliced, double, double, sc, F, F.us.v64, F.d64,
                        (const double *)y, (double *)z, r, Sw, part, rz, st, slot0);
    return;
  }
  if (block_plan(sc, F)) {
    if (mixed)
      wg_dispatch<float, float, true>(sc, F, F.us, F.us.v32, F.d32, nullptr, (const float *)y,
                                      (float *)z, S, st);
    else
      wg_dispatch<double, double, true>(sc, F, F.us, F.us.v64, F.d64, nullptr,
                                        (const double *)y, (double *)z, S, st);
    return;
  }
  if (mixed) {
    BwdRow<float, float> f{F.urp, F.uci, F.uv32, (const float *)y, (float *)z};
    run_csr_plan(sc, f, S, st, true);
  } else {
    BwdRow<double, double> f{F.urp, F.uci, F.uv, (const double *)y, (double *)z};
    run_csr_plan(sc, f, S, st, true);
  }
}

void launch_ic0_factor(const DevSched &sc, const Factor &F, double shift,
                       unsigned long long *err, hipStream_t st) {
  IC0Row f{F.lrp, F.lci, F.lv, shift, err};
  run_csr_plan(sc, f, (const Scal *)nullptr, st, false);
}

void launch_scatter_u(const Factor &F, bool mixed, hipStream_t st) {
  if (F.nnz == 0) return;
  hipLaunchKernelGGL(k_scatter_u, dim3(stream_grid(F.nnz)), dim3(kWG), 0, st, F.nnz, F.l2u, F.lv,
                     F.uv, mixed ? F.lv32 : nullptr, mixed ? F.uv32 : nullptr);
}

void launch_scatter_sell(const Factor &F, bool mixed, hipStream_t st) {
  if (F.nnz == 0 || !F.sell) return;
  hipLaunchKernelGGL(k_scatter_sell, dim3(stream_grid(F.nnz)), dim3(kWG), 0, st, F.nnz, F.lv,
                     F.l2ls, F.l2us, mixed ? nullptr : F.ls.v64, mixed ? F.ls.v32 : nullptr,
                     mixed ? nullptr : F.us.v64, mixed ? F.us.v32 : nullptr,
                     mixed ? nullptr : F.d64, mixed ? F.d32 : nullptr);
}

void launch_cast_f32(const double *a, float *b, long n, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_cast_f32, dim3(stream_grid(n)), dim3(kWG), 0, st, n, a, b);
}
void launch_cast_f64(const float *a, double *b, long n, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_cast_f64, dim3(stream_grid(n)), dim3(kWG), 0, st, n, a, b);
}
void launch_copy(const double *a, double *b, long n, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_copy, dim3(stream_grid(n)), dim3(kWG), 0, st, n, a, b);
}

}  // namespace iccg
