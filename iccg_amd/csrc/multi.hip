// multi.hip -- several right-hand sides of one system, solved together.
//
// The right-hand sides of a system run as members of a set of K (2 or 4).
// Every member keeps its own CG recurrence -- its own alpha, beta, norms,
// stop test and iteration count -- with exactly the operations the
// single-system solver performs for it alone, so each member's iterates are
// those of iccg_solver_run on that right-hand side, bit for bit.  What the
// set shares is the operator: a member's vectors sit interleaved with the
// others' (v[K i + m]), so one load of a matrix or factor entry serves all
// K members, and the K values one column gathers are one contiguous 16- or
// 32-byte run.  Per right-hand side the SpMV and the sweeps then move the
// matrix and factor bytes once per K, and in the ABMC sweeps every in-block
// level carries K independent rows per lane.
//
// Plan: A in the sliced layout, IC(0) on the sectioned ABMC factor (the
// plan of box and breadth-first blockings); the iteration is the
// single-system sectioned one (the forward sweep applies r -= alpha Ap, the
// last color runs forward and backward in one launch, the sweeps leave the
// |r|^2 and r.z partials and one finish folds both).  Other plans, and a
// member left alone, run on the single-system solver.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <functional>
#include <memory>
#include <thread>
#include <vector>

#include "dev.hpp"
#include "devfun.hpp"
#include "internal.hpp"
#include "objects.hpp"

namespace iccg {

// ---------------------------------------------------------------------------
// interleaved vector access: the K values of one row, 16-byte aligned
// (K = 2 or 4; buffers 256-byte aligned)
// ---------------------------------------------------------------------------
template <int K, bool NT = false, class T>
__device__ __forceinline__ void ldK(const T *__restrict__ p, T (&o)[K]) {
  typedef T vk_t __attribute__((ext_vector_type(K)));
  vk_t u;
  if constexpr (NT)
    u = __builtin_nontemporal_load(reinterpret_cast<const vk_t *>(p));
  else
    u = *reinterpret_cast<const vk_t *>(p);
#pragma unroll
  for (int m = 0; m < K; ++m) o[m] = u[m];
}

template <int K, bool NT = false, class T>
__device__ __forceinline__ void stK(T *__restrict__ p, const T (&o)[K]) {
  typedef T vk_t __attribute__((ext_vector_type(K)));
  vk_t u;
#pragma unroll
  for (int m = 0; m < K; ++m) u[m] = o[m];
  if constexpr (NT)
    __builtin_nontemporal_store(u, reinterpret_cast<vk_t *>(p));
  else
    *reinterpret_cast<vk_t *>(p) = u;
}

// Workgroup partials in component-major order: component c of slot g at
// part[2 (c stride + g)] (hi, lo), so that the one-workgroup fold reads each
// component's slots contiguously (in block_partial's slot-major order a fold
// over the 3K components of K members strides its loads by 6K doubles per
// slot).  The lanes' and the wavefronts' trees are block_partial's, the
// fold's finish_reduce2's.
template <int NC>
__device__ void block_partial_cm(double (&h)[NC], double (&l)[NC], double *part, int gidx,
                                 int stride) {
  __shared__ double sh[NC][2][16];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  wave_dd_reduce<NC>(h, l);
  __syncthreads();  // a previous call's reads of sh are complete
  if (lane == 0) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      sh[c][0][w] = h[c];
      sh[c][1][w] = l[c];
    }
  }
  __syncthreads();
  if (threadIdx.x < NC) {
    const int c = threadIdx.x;
    double ah = sh[c][0][0], al = sh[c][1][0];
    for (int k = 1; k < nw; ++k) dd_add(ah, al, sh[c][0][k], sh[c][1][k]);
    part[((size_t)c * stride + gidx) * 2 + 0] = ah;
    part[((size_t)c * stride + gidx) * 2 + 1] = al;
  }
}

// Adds this thread's slots b = tid, tid + T, tid + 2T, ... (T threads) of NC
// component-major partials, in ascending order, into (h, l).  The loads of
// kFoldAhead slots go out before their adds: one slot at a time, a finish
// over 2 x 4,096 slots waited for 32 memory round trips in a row.
template <int NC>
__device__ __forceinline__ void fold_slots_cm(const double *__restrict__ p, int count,
                                              int stride, double *h, double *l) {
  if (NC == 0) return;
  // twelve components (the SpMV's three per member of four) two slots ahead:
  // four took 252 VGPRs and ran slower (14.6 -> 16.4 us)
  constexpr int U = NC > 8 ? 2 : kFoldAhead;
  const int T = blockDim.x;
  for (int b0 = threadIdx.x; b0 < count; b0 += U * T) {
    double vh[U][NC > 0 ? NC : 1], vl[U][NC > 0 ? NC : 1];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int b = b0 + u * T;
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        vh[u][c] = b < count ? p[((size_t)c * stride + b) * 2] : 0.0;
        vl[u][c] = b < count ? p[((size_t)c * stride + b) * 2 + 1] : 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (b0 + u * T < count) {
#pragma unroll
        for (int c = 0; c < NC; ++c) dd_add(h[c], l[c], vh[u][c], vl[u][c]);
      }
  }
}

// finish_reduce2 over component-major partials (ca slots of stride sa at
// pa, cb of stride sb at pb): the same fold, in the same order
template <int NA, int NB>
__device__ bool finish_reduce2_cm(const double *__restrict__ pa, int ca, int sa,
                                  const double *__restrict__ pb, int cb, int sb, double *oa,
                                  double *ob) {
  constexpr int N = NA + NB;
  __shared__ double sh[N][2][16];
  __shared__ double res[N];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  double h[N], l[N];
#pragma unroll
  for (int c = 0; c < N; ++c) h[c] = l[c] = 0.0;
  fold_slots_cm<NA>(pa, ca, sa, h, l);
  fold_slots_cm<NB>(pb, cb, sb, h + NA, l + NA);
  wave_dd_reduce<N>(h, l);
  if (lane == 0) {
#pragma unroll
    for (int c = 0; c < N; ++c) {
      sh[c][0][w] = h[c];
      sh[c][1][w] = l[c];
    }
  }
  __syncthreads();
  if (threadIdx.x < N) {
    const int c = threadIdx.x;
    double ah = sh[c][0][0], al = sh[c][1][0];
    for (int k = 1; k < nw; ++k) dd_add(ah, al, sh[c][0][k], sh[c][1][k]);
    res[c] = ah + al;
  }
  __syncthreads();
  if (threadIdx.x != 0) return false;
#pragma unroll
  for (int c = 0; c < NA; ++c) oa[c] = res[c];
#pragma unroll
  for (int c = 0; c < NB; ++c) ob[c] = res[NA + c];
  return true;
}

template <int NC>
__device__ bool finish_reduce_cm(const double *__restrict__ part, int count, double (&out)[NC]) {
  return finish_reduce2_cm<NC, 0>(part, count, count, nullptr, 0, 0, out, nullptr);
}

template <int K>
__device__ __forceinline__ bool all_done(const Scal *S) {
  bool d = true;
#pragma unroll
  for (int m = 0; m < K; ++m) d = d && S[m].done != 0;
  return d;
}

// ---------------------------------------------------------------------------
// SpMV over the sliced A for K interleaved vectors: every row's chain is the
// single-system one (ascending stored order, one fma per entry, padding
// dropped by a select), once per member
// ---------------------------------------------------------------------------
template <int WM, bool TV, int K>
__device__ __forceinline__ void spmv16_finish_k(const Slice16<WM> &q, int *tb,
                                                const int *__restrict__ tbl,
                                                const double *__restrict__ X, int lane,
                                                double (&acc)[K]) {
  constexpr int WP = WM / 2;
#pragma unroll
  for (int m = 0; m < K; ++m) acc[m] = 0.0;
  if (q.w == 0) return;  // a slice of empty rows (uniform)
  constexpr unsigned tm = TV ? kTbl16Max - 1u : 15u;
  wave_sync();  // the previous slice's table reads are issued
  if (TV) {
    tb[lane] = q.te;
    for (int k = 64; k < q.tn; k += 64) tb[k + lane] = k + lane < q.tn ? tbl[q.ts + k + lane] : 0;
  } else {
    tb[lane & 15] = q.te;
  }
  wave_sync();
  int cl[WP], ch[WP];
#pragma unroll
  for (int j = 0; j < WP; ++j) {
    const int t0 = tb[(q.cp[j] >> 6) & tm], t1 = tb[(q.cp[j] >> 22) & tm];
    const unsigned lo = q.cp[j] & 0xFFFFu, hi = q.cp[j] >> 16;
    cl[j] = lo == 0xFFFFu ? -1 : (t0 << 6) | (int)(lo & 63u);
    ch[j] = hi == 0xFFFFu ? -1 : (t1 << 6) | (int)(hi & 63u);
  }
  const int tt = tb[(q.ct >> 6) & tm];
  const int ctl = q.ct == 0xFFFFu ? -1 : (tt << 6) | (int)(q.ct & 63u);
#pragma unroll
  for (int j = 0; j < WP; ++j) {
    double xl[K], xh[K];
    ldK<K>(X + (size_t)K * (cl[j] >= 0 ? cl[j] : 0), xl);
    ldK<K>(X + (size_t)K * (ch[j] >= 0 ? ch[j] : 0), xh);
#pragma unroll
    for (int m = 0; m < K; ++m) {
      const double a = fma(q.vp[j].x, xl[m], acc[m]);
      acc[m] = cl[j] >= 0 ? a : acc[m];
      const double b = fma(q.vp[j].y, xh[m], acc[m]);
      acc[m] = ch[j] >= 0 ? b : acc[m];
    }
  }
  double xt[K];
  ldK<K>(X + (size_t)K * (ctl >= 0 ? ctl : 0), xt);
#pragma unroll
  for (int m = 0; m < K; ++m) {
    const double c = fma(q.vt, xt[m], acc[m]);
    acc[m] = ctl >= 0 ? c : acc[m];
  }
}

// f(i, (A x_m)_i for m < K) for every row i < n of A's sliced layout: the
// 16-bit paired layout when A has it, else 32-bit columns
template <int WM, bool TV, int K, bool PF, class F>
__device__ __forceinline__ void for_rows_AXk(const DevCsr &A, const double *__restrict__ X,
                                             F f) {
  const int lane = threadIdx.x & 63;
  const WaveRange wr = xcd_waves(0, A.sell.nslice);
  const int first = __builtin_amdgcn_readfirstlane(wr.first);
  const int step = __builtin_amdgcn_readfirstlane(wr.step);
  const int end = __builtin_amdgcn_readfirstlane(wr.end);
  if (A.sell.col16 && PF) {
    __shared__ int tbs[kWG >> 6][TV ? kTbl16Max : 16];
    int *tb = tbs[threadIdx.x >> 6];
    if (first < end) {
      // the next slice's loads go out before the current slice's gathers
      Slice16<WM> cur;
      spmv16_load<WM, TV>(A.sell.off, A.sell.col16, A.sell.tbl, A.sell.toff, A.sell.v64, first,
                          lane, cur);
      for (int s = first; s < end; s += step) {
        Slice16<WM> nxt = cur;
        if (s + step < end)
          spmv16_load<WM, TV>(A.sell.off, A.sell.col16, A.sell.tbl, A.sell.toff, A.sell.v64,
                              s + step, lane, nxt);
        double y[K];
        spmv16_finish_k<WM, TV, K>(cur, tb, A.sell.tbl, X, lane, y);
        const int i = (s << 6) + lane;
        if (i < A.n) f(i, y);
        cur = nxt;
      }
    }
  } else if (A.sell.col16) {
    __shared__ int tbs[kWG >> 6][TV ? kTbl16Max : 16];
    int *tb = tbs[threadIdx.x >> 6];
    for (int s = first; s < end; s += step) {
      Slice16<WM> q;
      spmv16_load<WM, TV>(A.sell.off, A.sell.col16, A.sell.tbl, A.sell.toff, A.sell.v64, s, lane,
                          q);
      double y[K];
      spmv16_finish_k<WM, TV, K>(q, tb, A.sell.tbl, X, lane, y);
      const int i = (s << 6) + lane;
      if (i < A.n) f(i, y);
    }
  } else {
    for (int s = first; s < end; s += step) {
      const int o0 = A.sell.off[s];
      const int w = (A.sell.off[s + 1] - o0) >> 6;
      const int base = o0 + lane;
      double y[K];
#pragma unroll
      for (int m = 0; m < K; ++m) y[m] = 0.0;
#pragma unroll
      for (int k = 0; k < WM; ++k) {
        if (k < w) {
          const int c = ldm<true>(&A.sell.col[base + 64 * k]);
          const double v = ldm<true>(&A.sell.v64[base + 64 * k]);
          double xv[K];
          ldK<K>(X + (size_t)K * (c >= 0 ? c : 0), xv);
#pragma unroll
          for (int m = 0; m < K; ++m) {
            const double a = fma(v, xv[m], y[m]);
            y[m] = c >= 0 ? a : y[m];
          }
        }
      }
      const int i = (s << 6) + lane;
      if (i < A.n) f(i, y);
    }
  }
}

// AP = A P with p.Ap, ||p||^2, ||Ap||^2 per member (components 3m + c),
// as k_spmv_cg; the next slice's loads go out before the current slice's
// gathers (192^3, four members: 269.7 -> 243.7 us per launch at 166 VGPRs
// and 3 waves per SIMD, against 124 VGPRs and 4 waves without).  Class 4
// (5-point rows) is held to 4 waves per SIMD (130 -> 128 VGPRs, no spills).
template <int WM, bool TV, int K>
__global__ __launch_bounds__(kWG, WM == 4 ? 4 : 1) void km_spmv_cg(DevCsr A, const double *__restrict__ P,
                                                  double *__restrict__ AP, const Scal *S,
                                                  double *part) {
  if (all_done<K>(S)) return;
  double h[3 * K], l[3 * K];
#pragma unroll
  for (int c = 0; c < 3 * K; ++c) h[c] = l[c] = 0.0;
  for_rows_AXk<WM, TV, K, true>(A, P, [&](int i, const double(&y)[K]) {
    stK<K, true>(AP + (size_t)K * i, y);  // read once, by the forward sweep
    double pv[K];
    ldK<K>(P + (size_t)K * i, pv);
#pragma unroll
    for (int m = 0; m < K; ++m) {
      dd_fma(h[3 * m], l[3 * m], pv[m], y[m]);
      dd_fma(h[3 * m + 1], l[3 * m + 1], pv[m], pv[m]);
      dd_fma(h[3 * m + 2], l[3 * m + 2], y[m], y[m]);
    }
  });
  block_partial_cm<3 * K>(h, l, part, blockIdx.x, gridDim.x);
}

// R = B - A X with ||b||^2, ||r||^2 per member (components 2m + c), as
// k_resid; mode 0 writes no R, modes 1 and 2 write the members of mask
template <int WM, bool TV, int K>
__global__ __launch_bounds__(kWG) void km_resid(DevCsr A, const double *__restrict__ X,
                                                const double *__restrict__ B,
                                                double *__restrict__ R, double *part, int mode,
                                                int mask) {
  double h[2 * K], l[2 * K];
#pragma unroll
  for (int c = 0; c < 2 * K; ++c) h[c] = l[c] = 0.0;
  for_rows_AXk<WM, TV, K, false>(A, X, [&](int i, const double(&ax)[K]) {
    double bv[K], rv[K];
    ldK<K>(B + (size_t)K * i, bv);
#pragma unroll
    for (int m = 0; m < K; ++m) rv[m] = bv[m] - ax[m];
    if (mode != 0) {
      if (mask == (1 << K) - 1) {
        stK<K>(R + (size_t)K * i, rv);
      } else {
#pragma unroll
        for (int m = 0; m < K; ++m)
          if ((mask >> m) & 1) R[(size_t)K * i + m] = rv[m];
      }
    }
#pragma unroll
    for (int m = 0; m < K; ++m) {
      dd_fma(h[2 * m], l[2 * m], bv[m], bv[m]);
      dd_fma(h[2 * m + 1], l[2 * m + 1], rv[m], rv[m]);
    }
  });
  block_partial_cm<2 * K>(h, l, part, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// finish kernels: the single-system scalar logic, member by member.  One
// workgroup of four wavefronts: with the 2K-3K components of K members the
// shuffle trees of sixteen wavefronts cost more than they save (192^3, four
// members: the alpha finish 21.9 -> 14.5 us; 1,024 threads with eight slots
// per thread loaded ahead spill, 48 us)
// ---------------------------------------------------------------------------
constexpr int kMFinWG = 256;
// k_fin_resid for the members of mask (mode 0: every member's rr_true)
template <int K>
__global__ __launch_bounds__(kMFinWG) void km_fin_resid(const double *__restrict__ part,
                                                       int count, Scal *S, int mode, double tol,
                                                       int max_iter, int mask) {
  double out[2 * K];
  if (!finish_reduce_cm<2 * K>(part, count, out)) return;
  for (int m = 0; m < K; ++m) {
    if (!((mask >> m) & 1)) continue;
    Scal &s = S[m];
    double normb = sqrt(out[2 * m]);
    if (normb == 0.0) normb = 1.0;
    if (mode == 0) {
      s.rr_true = out[2 * m + 1];
      s.normb = normb;
      continue;
    }
    const double rel = sqrt(out[2 * m + 1]) / normb;
    if (mode == 1) {
      s.tol = tol;
      s.max_iter = max_iter;
      s.iters = 0;
    }
    s.normb = normb;
    s.rel = rel;
    s.converged = 0;
    s.breakdown = 0;
    s.done = 0;
    if (rel < s.tol) {
      s.converged = 1;
      s.done = 1;
    } else if (s.iters >= s.max_iter) {
      s.done = 1;
    }
  }
}

// k_fin_alpha per member
template <int K>
__global__ __launch_bounds__(kMFinWG) void km_fin_alpha(const double *__restrict__ part, int count,
                                                       Scal *S) {
  if (all_done<K>(S)) {
    __syncthreads();  // every thread has read the flags
    if (threadIdx.x == 0)
      for (int m = 0; m < K; ++m) S[m].done = 2;
    return;
  }
  double out[3 * K];
  if (!finish_reduce_cm<3 * K>(part, count, out)) return;
  for (int m = 0; m < K; ++m) {
    Scal &s = S[m];
    if (s.done) {
      s.done = 2;
      continue;
    }
    const double pAp = out[3 * m];
    const double thr = DBL_EPSILON * sqrt(out[3 * m + 1]) * sqrt(out[3 * m + 2]);
    s.pAp = pAp;
    if (pAp <= thr) {
      s.breakdown = 1;
      s.done = 1;
    } else {
      s.alpha = s.rz_old / pAp;
    }
  }
}

// k_fin_xr_rz per member: the stop test and beta from the sweeps' |r|^2 and
// r.z partials
template <int K>
__global__ __launch_bounds__(kMFinWG) void km_fin_stop(const double *__restrict__ pxr, int cxr,
                                                      const double *__restrict__ prz, int crz,
                                                      Scal *S) {
  if (all_done<K>(S)) return;
  double rr[K], rz[K];
  if (!finish_reduce2_cm<K, K>(pxr, cxr, cxr, prz, crz, crz, rr, rz)) return;
  for (int m = 0; m < K; ++m) {
    Scal &s = S[m];
    if (s.done) continue;
    const double rel = sqrt(rr[m]) / s.normb;
    const int it = s.iters + 1;
    s.iters = it;
    s.rel = rel;
    if (rel < s.tol) {
      s.converged = 1;
      s.done = 1;
    } else if (it >= s.max_iter) {
      s.done = 1;
    } else {
      s.beta = rz[m] / s.rz_old;
      s.rz_old = rz[m];
    }
  }
}

// k_fin_rz per member (init: rz_old at the start of a solve)
template <int K>
__global__ __launch_bounds__(kMFinWG) void km_fin_rz(const double *__restrict__ part, int count,
                                                    Scal *S, int init) {
  if (all_done<K>(S)) return;
  double out[K];
  if (!finish_reduce_cm<K>(part, count, out)) return;
  for (int m = 0; m < K; ++m) {
    Scal &s = S[m];
    if (s.done) continue;
    if (init) {
      s.rz_old = out[m];
    } else {
      s.beta = out[m] / s.rz_old;
      s.rz_old = out[m];
    }
  }
}

// ---------------------------------------------------------------------------
// streaming kernels over the n K interleaved entries, 4 per lane per step;
// entry e belongs to member e % K (K divides 4)
// ---------------------------------------------------------------------------
#define KQUAD_LOOP(nk)                                                                \
  for (long q_ = blockIdx.x * (long)kWG + threadIdx.x, nq_ = (nk) >> 2; q_ < nq_;    \
       q_ += (long)gridDim.x * kWG)
#define KTAIL_LOOP(nk)                                                                   \
  for (long i = (((nk) >> 2) << 2) + blockIdx.x * (long)kWG + threadIdx.x; i < (nk); \
       i += (long)gridDim.x * kWG)

template <class T>
__device__ __forceinline__ void ld4v(const T *p, long i, double (&o)[4]) {
  ld4(p + i, 0, o);
}

// r.z per member (k_rz)
template <int K, class ZT>
__global__ __launch_bounds__(kWG) void km_rz(long nk, const double *__restrict__ R,
                                             const ZT *__restrict__ Z, const Scal *S,
                                             double *part) {
  if (all_done<K>(S)) return;
  double h[K], l[K];
#pragma unroll
  for (int m = 0; m < K; ++m) h[m] = l[m] = 0.0;
  KQUAD_LOOP(nk) {
    const long i = q_ << 2;
    double rv[4], zv[4];
    ld4v(R, i, rv);
    ld4v(Z, i, zv);
#pragma unroll
    for (int e = 0; e < 4; ++e) dd_fma(h[e % K], l[e % K], rv[e], zv[e]);
  }
  KTAIL_LOOP(nk) {
#pragma unroll
    for (int m = 0; m < K; ++m)
      if (i % K == m) dd_fma(h[m], l[m], R[i], (double)Z[i]);
  }
  block_partial_cm<K>(h, l, part, blockIdx.x, gridDim.x);
}

// p = z + beta p, or p = z at the start (k_update_p), members not done
template <int K, class ZT>
__global__ __launch_bounds__(kWG) void km_update_p(long nk, const ZT *__restrict__ Z,
                                                   double *__restrict__ P, const Scal *S,
                                                   int init) {
  if (all_done<K>(S)) return;
  double beta[K];
  bool up[K];
#pragma unroll
  for (int m = 0; m < K; ++m) {
    beta[m] = S[m].beta;
    up[m] = S[m].done == 0;
  }
  KQUAD_LOOP(nk) {
    const long i = q_ << 2;
    double zv[4], pv[4];
    ld4v(Z, i, zv);
    ld4v(P, i, pv);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const double pn = init ? zv[e] : fma(beta[e % K], pv[e], zv[e]);
      pv[e] = up[e % K] ? pn : pv[e];
    }
    st4(P + i, 0, pv);
  }
  KTAIL_LOOP(nk) {
#pragma unroll
    for (int m = 0; m < K; ++m)
      if (i % K == m && up[m]) {
        const double zi = (double)Z[i];
        P[i] = init ? zi : fma(beta[m], P[i], zi);
      }
  }
}

// x += alpha p and p = z + beta p per member (k_update_xp): the x update for
// every member whose x/r update ran this iteration (done <= 1, no
// breakdown), the p update for members still running
template <int K, class ZT>
__global__ __launch_bounds__(kWG) void km_update_xp(long nk, double *__restrict__ X,
                                                    const ZT *__restrict__ Z,
                                                    double *__restrict__ P, const Scal *S) {
  double alpha[K], beta[K];
  bool xu[K], pu[K], any = false, anyp = false;
#pragma unroll
  for (int m = 0; m < K; ++m) {
    alpha[m] = S[m].alpha;
    beta[m] = S[m].beta;
    xu[m] = !(S[m].done > 1 || S[m].breakdown);
    pu[m] = xu[m] && !S[m].done;
    any |= xu[m];
    anyp |= pu[m];
  }
  if (!any) return;
  KQUAD_LOOP(nk) {
    const long i = q_ << 2;
    double xv[4], pv[4];
    ld4_nt(X + i, 0, xv);  // x: read by no other kernel of the iteration
    ld4v(P, i, pv);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const double xn = fma(alpha[e % K], pv[e], xv[e]);
      xv[e] = xu[e % K] ? xn : xv[e];
    }
    st4_nt(X + i, 0, xv);
    if (anyp) {
      double zv[4];
      ld4v(Z, i, zv);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const double pn = fma(beta[e % K], pv[e], zv[e]);
        pv[e] = pu[e % K] ? pn : pv[e];
      }
      st4(P + i, 0, pv);
    }
  }
  KTAIL_LOOP(nk) {
#pragma unroll
    for (int m = 0; m < K; ++m)
      if (i % K == m && xu[m]) {
        const double pi = P[i];
        X[i] = fma(alpha[m], pi, X[i]);
        if (pu[m]) P[i] = fma(beta[m], pi, (double)Z[i]);
      }
  }
}

// ---------------------------------------------------------------------------
// Sectioned ABMC sweeps for K members (k_abmc_fwd_sec / k_abmc_bwd_sec): the
// same block layout, the same per-row chains and level loop (sec_levels),
// each lane carrying its row of every member.  The wavefront's LDS window
// holds K windows of kSecWindow slots (member m at m * kSecWindow).
// ---------------------------------------------------------------------------
template <class YT, int SW, int SO, bool OUT, int K>
__device__ __forceinline__ void sec_levels_k(int nl, const YT (&acc)[K], YT d,
                                             const int (&ad)[SW], const YT (&va)[SW],
                                             const YT (&vo)[SO], const YT (&xo)[SO][K], YT *my,
                                             int lane, bool act, YT (&yi)[K]) {
  const YT dy = rcp_refined(d);
  YT a[K];
#pragma unroll
  for (int m = 0; m < K; ++m) {
    yi[m] = (YT)0;
    a[m] = acc[m];
  }
  for (int l = 0; l < nl; ++l) {
    YT t[K][SW];
#pragma unroll
    for (int m = 0; m < K; ++m)
#pragma unroll
      for (int k = 0; k < SW; ++k) t[m][k] = my[m * kSecWindow + ad[k]];
#pragma unroll
    for (int m = 0; m < K; ++m) {
      a[m] = acc[m];
#pragma unroll
      for (int k = 0; k < SW; ++k) a[m] = fma_t(-va[k], t[m][k], a[m]);
      if (OUT) {
#pragma unroll
        for (int k = 0; k < SO; ++k) a[m] = fma_t(-vo[k], xo[k][m], a[m]);
      }
      yi[m] = div_steps(a[m], d, dy);
    }
#pragma unroll
    for (int m = 0; m < K; ++m) my[m * kSecWindow + lane] = yi[m];
    wave_sync();
  }
  bool exact = true;
#pragma unroll
  for (int m = 0; m < K; ++m) exact = exact && div_steps_exact(a[m], d);
  if (!__all(!act || nl == 0 || exact)) {
    for (int l = 0; l < nl; ++l) {
      YT t[K][SW];
#pragma unroll
      for (int m = 0; m < K; ++m)
#pragma unroll
        for (int k = 0; k < SW; ++k) t[m][k] = my[m * kSecWindow + ad[k]];
#pragma unroll
      for (int m = 0; m < K; ++m) {
        a[m] = acc[m];
#pragma unroll
        for (int k = 0; k < SW; ++k) a[m] = fma_t(-va[k], t[m][k], a[m]);
        if (OUT) {
#pragma unroll
          for (int k = 0; k < SO; ++k) a[m] = fma_t(-vo[k], xo[k][m], a[m]);
        }
        yi[m] = a[m] / d;
      }
#pragma unroll
      for (int m = 0; m < K; ++m) my[m * kSecWindow + lane] = yi[m];
      wave_sync();
    }
  }
}

// One lane's entries of block kb in the sectioned layout (sec_fetch with
// LOAD_NOW, sec_take): in-block rows and values, outside columns (-1 when
// absent) and values, the diagonal
template <class VT, class YT, int SW, int SO, bool NT>
__device__ __forceinline__ void sec_entries(int kb, int o0, int w, int t,
                                            const unsigned *__restrict__ inb,
                                            const int *__restrict__ col,
                                            const VT *__restrict__ v, const VT *__restrict__ diag,
                                            int bs, int (&ad)[SW], YT (&va)[SW], int (&co)[SO],
                                            YT (&vo)[SO], YT &d) {
  const int base = o0 + t;
  const unsigned pk = ldm<NT>(&inb[64 * kb + t]);
#pragma unroll
  for (int k = 0; k < SW; ++k) {
    ad[k] = (int)((pk >> (8 * k)) & 0xFFu);
    va[k] = (YT)ldm<NT>(&v[base + 64 * k]);  // absent entries hold 0
  }
#pragma unroll
  for (int k = 0; k < SO; ++k) {
    co[k] = -1;
    vo[k] = (YT)0;
    if (SW + k < w) {
      const int c = ldm<NT>(&col[base + 64 * (SW + k)]);
      const VT vv = ldm<NT>(&v[base + 64 * (SW + k)]);
      co[k] = c;
      vo[k] = c >= 0 ? (YT)vv : (YT)0;
    }
  }
  d = (YT)diag[bs + t];
}

// L y = r for K members over the blocks [kb0, kb1) of one color; FB: the
// last color's backward rows follow (z written, y not); RU: r -= alpha Ap
// first for the members still running, with the |r|^2 partials at
// part[gidx0 + workgroup] (and, with FB, the r.z partials at
// part_rz[workgroup]), K components each
template <int K, class VT, class YT, int SW, int SO, bool FB, bool NT, bool RU>
__global__ __launch_bounds__(kWG) void km_fwd_sec(int kb0, int kb1, const int *__restrict__ bstart,
                                                  const int *__restrict__ nlev,
                                                  const int *__restrict__ off,
                                                  const unsigned *__restrict__ inb,
                                                  const int *__restrict__ col,
                                                  const VT *__restrict__ lv,
                                                  const VT *__restrict__ diag,
                                                  double *__restrict__ R, YT *__restrict__ Y,
                                                  const int *__restrict__ unlev,
                                                  const int *__restrict__ uoff,
                                                  const unsigned *__restrict__ uinb,
                                                  const VT *__restrict__ uv, YT *__restrict__ Z,
                                                  const Scal *S, const double *__restrict__ AP,
                                                  double *part, int gidx0, double *part_rz,
                                                  int nslot) {
  __shared__ YT win[(kWG / 64) * K * kSecWindow];
  if (all_done<K>(S)) return;
  const int lane = threadIdx.x & 63;
  YT *my = win + (threadIdx.x >> 6) * K * kSecWindow;
  const WaveRange wr = xcd_waves(kb0, kb1);
  const int first = __builtin_amdgcn_readfirstlane(wr.first);
  const int step = __builtin_amdgcn_readfirstlane(wr.step);
  const int end = __builtin_amdgcn_readfirstlane(wr.end);
  if (!RU && first >= end) return;  // RU: every wavefront reaches the partials
  double alpha[K];
  bool upd[K];
#pragma unroll
  for (int m = 0; m < K; ++m) {
    alpha[m] = RU ? S[m].alpha : 0.0;
    upd[m] = RU && S[m].done == 0;
  }
  double h[K], l[K], hz[K], lz[K];
#pragma unroll
  for (int m = 0; m < K; ++m) h[m] = l[m] = hz[m] = lz[m] = 0.0;
  YT none[SO], nonek[SO][K];
#pragma unroll
  for (int k = 0; k < SO; ++k) {
    none[k] = (YT)0;
#pragma unroll
    for (int m = 0; m < K; ++m) nonek[k][m] = (YT)0;
  }
  for (int kb = first; kb < end; kb += step) {
    const int bs = bstart[kb], be = bstart[kb + 1], nl = nlev[kb];
    const int o0 = off[kb], w = (off[kb + 1] - o0) >> 6;
    const int t = min(lane, be - bs - 1);  // lanes past the block repeat its last row
    const bool act = bs + lane < be;
    int ad[SW], co[SO];
    YT va[SW], vo[SO], d;
    sec_entries<VT, YT, SW, SO, NT>(kb, o0, w, t, inb, col, lv, diag, bs, ad, va, co, vo, d);
    double a[K];
    ldK<K>(R + (size_t)K * (bs + t), a);
    YT acc[K];
    if (RU) {
      double ap[K];
      ldK<K, true>(AP + (size_t)K * (bs + t), ap);  // Ap: read once
#pragma unroll
      for (int m = 0; m < K; ++m) {
        const double rn = fma(-alpha[m], ap[m], a[m]);
        a[m] = upd[m] ? rn : a[m];
      }
      if (act) {
        stK<K>(R + (size_t)K * (bs + lane), a);
#pragma unroll
        for (int m = 0; m < K; ++m) dd_fma(h[m], l[m], a[m], a[m]);
      }
    }
#pragma unroll
    for (int m = 0; m < K; ++m) acc[m] = (YT)a[m];
    // the earlier colors' terms, ascending; absent ones are fma(-0, 0, acc)
#pragma unroll
    for (int k = 0; k < SO; ++k) {
      YT xo[K];
      ldK<K>(Y + (size_t)K * (co[k] >= 0 ? co[k] : 0), xo);
#pragma unroll
      for (int m = 0; m < K; ++m) acc[m] = fma_t(-vo[k], co[k] >= 0 ? xo[m] : (YT)0, acc[m]);
    }
    int uad[SW];
    YT uva[SW];
    if (FB) {  // this block's U entries
      const int uo = uoff[kb] + t;
      const unsigned pk = ldm<NT>(&uinb[64 * kb + t]);
#pragma unroll
      for (int k = 0; k < SW; ++k) {
        uad[k] = (int)((pk >> (8 * k)) & 0xFFu);
        uva[k] = (YT)ldm<NT>(&uv[uo + 64 * k]);  // absent entries hold 0
      }
    }
    if (lane == 0) {
#pragma unroll
      for (int m = 0; m < K; ++m) my[m * kSecWindow + 64] = (YT)0;
    }
    wave_sync();
    YT yi[K];
    sec_levels_k<YT, SW, SO, false, K>(nl, acc, d, ad, va, none, nonek, my, lane, act, yi);
    if (FB) {
      YT zi[K];
      sec_levels_k<YT, SW, SO, false, K>(unlev[kb], yi, d, uad, uva, none, nonek, my, lane, act,
                                         zi);
      if (act) {
        stK<K>(Z + (size_t)K * (bs + lane), zi);
        if (RU) {
#pragma unroll
          for (int m = 0; m < K; ++m) dd_fma(hz[m], lz[m], a[m], (double)zi[m]);
        }
      }
    } else if (act) {
      stK<K>(Y + (size_t)K * (bs + lane), yi);
    }
  }
  if (RU) block_partial_cm<K>(h, l, part, gidx0 + blockIdx.x, nslot);
  if (RU && FB) block_partial_cm<K>(hz, lz, part_rz, blockIdx.x, nslot);
}

// L^T z = y for K members over the blocks [kb0, kb1) of one color; RZ: the
// r.z partials at part[gidx0 + workgroup], K components
template <int K, class VT, class YT, int SW, int SO, bool RZ, bool NT>
__global__ __launch_bounds__(kWG) void km_bwd_sec(int kb0, int kb1, const int *__restrict__ bstart,
                                                  const int *__restrict__ nlev,
                                                  const int *__restrict__ off,
                                                  const unsigned *__restrict__ inb,
                                                  const int *__restrict__ col,
                                                  const VT *__restrict__ uv,
                                                  const VT *__restrict__ diag,
                                                  const YT *__restrict__ Y, YT *__restrict__ Z,
                                                  const double *__restrict__ R, const Scal *S,
                                                  double *part, int gidx0, int nslot) {
  __shared__ YT win[(kWG / 64) * K * kSecWindow];
  if (all_done<K>(S)) return;
  const int lane = threadIdx.x & 63;
  YT *my = win + (threadIdx.x >> 6) * K * kSecWindow;
  const WaveRange wr = xcd_waves(kb0, kb1);
  const int first = __builtin_amdgcn_readfirstlane(wr.first);
  const int step = __builtin_amdgcn_readfirstlane(wr.step);
  const int end = __builtin_amdgcn_readfirstlane(wr.end);
  double h[K], l[K];
#pragma unroll
  for (int m = 0; m < K; ++m) h[m] = l[m] = 0.0;
  for (int kb = first; kb < end; kb += step) {
    const int bs = bstart[kb], be = bstart[kb + 1], nl = nlev[kb];
    const int o0 = off[kb], w = (off[kb + 1] - o0) >> 6;
    const int t = min(lane, be - bs - 1);
    const bool act = bs + lane < be;
    int ad[SW], co[SO];
    YT va[SW], vo[SO], d;
    sec_entries<VT, YT, SW, SO, NT>(kb, o0, w, t, inb, col, uv, diag, bs, ad, va, co, vo, d);
    YT acc[K];
    ldK<K>(Y + (size_t)K * (bs + t), acc);
    double ri[K];
    if (RZ) ldK<K>(R + (size_t)K * (bs + t), ri);
    // the later colors' z (final), absent ones 0
    YT xo[SO][K];
#pragma unroll
    for (int k = 0; k < SO; ++k) {
      YT zv[K];
      ldK<K>(Z + (size_t)K * (co[k] >= 0 ? co[k] : 0), zv);
#pragma unroll
      for (int m = 0; m < K; ++m) xo[k][m] = co[k] >= 0 ? zv[m] : (YT)0;
    }
    if (lane == 0) {
#pragma unroll
      for (int m = 0; m < K; ++m) my[m * kSecWindow + 64] = (YT)0;
    }
    wave_sync();
    YT zi[K];
    sec_levels_k<YT, SW, SO, true, K>(nl, acc, d, ad, va, vo, xo, my, lane, act, zi);
    if (act) {
      stK<K>(Z + (size_t)K * (bs + lane), zi);
      if (RZ) {
#pragma unroll
        for (int m = 0; m < K; ++m) dd_fma(h[m], l[m], ri[m], (double)zi[m]);
      }
    }
  }
  if (RZ) block_partial_cm<K>(h, l, part, gidx0 + blockIdx.x, nslot);
}

}  // namespace iccg

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
namespace iccg {

// the plan the sets run natively: A sliced with rows of at most 16 entries,
// IC(0) on the sectioned ABMC factor
static bool multi_plan(const iccg_csr_s &A, const iccg_precond_s &M) {
  const Factor &F = M.F;
  const int wm = std::max(A.d.sell.maxw, 1);
  return M.kind == ICCG_PRECOND_IC0 && A.d.sell.off && wm <= 16 && F.sell && F.sec > 0 &&
         M.sc.kind == 2 && M.sc.d_brows == nullptr && M.sc.max_block <= 64 &&
         std::max(F.ls.maxw, F.us.maxw) <= 16;
}

// Row-width class of the SpMV kernels (spmv_wm): 4, 8 or 16, since
// multi_plan takes rows of at most 16 entries
static int mwm(const DevCsr &A) { return spmv_wm(A); }

#define MDISPATCH_A(A, KERNEL, K, ...)                                          \
  do {                                                                          \
    const bool _tv = (A).sell.col16 && (A).sell.toff;                           \
    switch (mwm(A)) {                                                           \
      case 4:                                                                   \
        if (_tv) hipLaunchKernelGGL((KERNEL<4, true, K>), __VA_ARGS__);         \
        else hipLaunchKernelGGL((KERNEL<4, false, K>), __VA_ARGS__);            \
        break;                                                                  \
      case 8:                                                                   \
        if (_tv) hipLaunchKernelGGL((KERNEL<8, true, K>), __VA_ARGS__);         \
        else hipLaunchKernelGGL((KERNEL<8, false, K>), __VA_ARGS__);            \
        break;                                                                  \
      default:                                                                  \
        if (_tv) hipLaunchKernelGGL((KERNEL<16, true, K>), __VA_ARGS__);        \
        else hipLaunchKernelGGL((KERNEL<16, false, K>), __VA_ARGS__);           \
        break;                                                                  \
    }                                                                           \
  } while (0)

// one set of K members: its interleaved vectors, partials and scalars
struct MSet {
  int K = 0, m = 0, first = 0;  // members used of K, first right-hand side
  hipStream_t st = nullptr;     // the set's own stream: sets of one solver run concurrently
  hipEvent_t wait_ev = nullptr;  // blocking waits on st
  double *X = nullptr, *R = nullptr, *P = nullptr, *AP = nullptr, *B = nullptr;
  void *Y = nullptr, *Z = nullptr;
  double *part = nullptr;
  Scal *S = nullptr, *hS = nullptr;
  int graph_version = -1;
  std::vector<std::pair<int, hipGraphExec_t>> execs;
  void drop_graphs() {
    for (auto &e : execs) (void)hipGraphExecDestroy(e.second);
    execs.clear();
  }
  void release() {
    if (st) (void)hipStreamSynchronize(st);  // nothing queued still uses its vectors
    drop_graphs();
    dfree(X);
    dfree(R);
    dfree(P);
    dfree(AP);
    dfree(B);
    dfree(Y);
    dfree(Z);
    dfree(part);
    dfree(S);
    hfree(hS);
    if (wait_ev) (void)hipEventDestroy(wait_ev);
    if (st) (void)hipStreamDestroy(st);
  }
};

// the sweeps' workgroups per color (0 for an empty color) and their sum
static std::vector<int> color_groups(const DevSched &sc, int &total) {
  std::vector<int> g((size_t)sc.nc, 0);
  total = 0;
  for (int c = 0; c < sc.nc; ++c) {
    const int b0 = sc.h_color_blk_v[c], b1 = sc.h_color_blk_v[c + 1];
    g[(size_t)c] = b1 > b0 ? slice_grid(b1 - b0) : 0;
    total += g[(size_t)c];
  }
  return g;
}

// partial pairs a set needs: the SpMV's 3K per workgroup, the residual's 2K,
// the sweeps' |r|^2 and r.z regions of K per workgroup each
static size_t mset_part_pairs(const iccg_csr_s &A, const iccg_precond_s &M, int K) {
  int G = 0;
  (void)color_groups(M.sc, G);
  const size_t ga = (size_t)slice_grid(A.d.sell.nslice);
  const size_t gq = (size_t)kMaxRedBlocks;
  return std::max({ga * 3 * K, 2 * (size_t)G * K, gq * K});
}

template <int K, class VT, class YT, int SW, int SO, bool NT>
struct MSweeps {
  // forward colors; with AP, every launch applies the r update and leaves the
  // |r|^2 partials at part[0, G), the fused last color's r.z ones at
  // part + 2 K G from slot 0
  static void fwd(const DevSched &sc, const Factor &F, bool fb, const VT *lv, const VT *uv,
                  const VT *dg, double *R, YT *Y, YT *Z, const Scal *S, const double *AP,
                  double *part, hipStream_t st) {
    int G = 0;
    const std::vector<int> g = color_groups(sc, G);
    double *prz = AP ? part + 2 * (size_t)K * G : nullptr;
    int gidx0 = 0;
    for (int c = 0; c < sc.nc; ++c) {
      if (g[(size_t)c] == 0) continue;
      const int b0 = sc.h_color_blk_v[c], b1 = sc.h_color_blk_v[c + 1];
      const bool last = fb && c == sc.nc - 1;
      const dim3 grid(g[(size_t)c]), blk(kWG);
#define KM_FWD(FB_, RU_)                                                                      \
  hipLaunchKernelGGL((km_fwd_sec<K, VT, YT, SW, SO, FB_, NT, RU_>), grid, blk, 0, st, b0, b1, \
                     sc.d_bstart, sc.d_nlev_f, F.ls.off, F.ls.inb, F.ls.col, lv, dg, R, Y,     \
                     sc.d_nlev_b, F.us.off, F.us.inb, uv, Z, S, AP, part, gidx0, prz, G)
      if (last && AP)
        KM_FWD(true, true);
      else if (AP)
        KM_FWD(false, true);
      else if (last)
        KM_FWD(true, false);
      else
        KM_FWD(false, false);
#undef KM_FWD
      gidx0 += g[(size_t)c];
    }
  }
  // backward colors (the last one skipped when the forward launch took it);
  // with prz, the r.z partials from slot `slot0` on
  static void bwd(const DevSched &sc, const Factor &F, bool fb, const VT *uv, const VT *dg,
                  const YT *Y, YT *Z, const double *R, const Scal *S, double *prz, int slot0,
                  hipStream_t st) {
    int G = 0;
    const std::vector<int> g = color_groups(sc, G);
    int gidx0 = slot0;
    for (int t = 0; t < sc.nc; ++t) {
      const int c = sc.nc - 1 - t;
      if (g[(size_t)c] == 0 || (fb && t == 0)) continue;
      const int b0 = sc.h_color_blk_v[c], b1 = sc.h_color_blk_v[c + 1];
      const dim3 grid(g[(size_t)c]), blk(kWG);
      if (prz)
        hipLaunchKernelGGL((km_bwd_sec<K, VT, YT, SW, SO, true, NT>), grid, blk, 0, st, b0, b1,
                           sc.d_bstart, sc.d_nlev_b, F.us.off, F.us.inb, F.us.col, uv, dg, Y, Z,
                           R, S, prz, gidx0, G);
      else
        hipLaunchKernelGGL((km_bwd_sec<K, VT, YT, SW, SO, false, NT>), grid, blk, 0, st, b0, b1,
                           sc.d_bstart, sc.d_nlev_b, F.us.off, F.us.inb, F.us.col, uv, dg, Y, Z,
                           R, S, nullptr, 0, G);
      gidx0 += g[(size_t)c];
    }
  }
};

// FN<K, VT, YT, SW, SO, NT>::method(args) over the factor's section shape,
// load policy and precision
#define MDISPATCH_SEC(K, F, mixed, CALL)                                        \
  do {                                                                          \
    const int _sw = (F).sec;                                                    \
    const bool _w = (F).seco > _sw, _nt = (F).nt;                               \
    if (mixed) {                                                                \
      using VT = float;                                                         \
      using YT = float;                                                         \
      if (_sw == 2 && _w && _nt) { using SWP = MSweeps<K, VT, YT, 2, 6, true>; CALL; }        \
      else if (_sw == 2 && _w) { using SWP = MSweeps<K, VT, YT, 2, 6, false>; CALL; }         \
      else if (_sw == 2 && _nt) { using SWP = MSweeps<K, VT, YT, 2, 2, true>; CALL; }         \
      else if (_sw == 2) { using SWP = MSweeps<K, VT, YT, 2, 2, false>; CALL; }               \
      else if (_w && _nt) { using SWP = MSweeps<K, VT, YT, 3, 6, true>; CALL; }               \
      else if (_w) { using SWP = MSweeps<K, VT, YT, 3, 6, false>; CALL; }                     \
      else if (_nt) { using SWP = MSweeps<K, VT, YT, 3, 3, true>; CALL; }                     \
      else { using SWP = MSweeps<K, VT, YT, 3, 3, false>; CALL; }                             \
    } else {                                                                    \
      using VT = double;                                                        \
      using YT = double;                                                        \
      if (_sw == 2 && _w && _nt) { using SWP = MSweeps<K, VT, YT, 2, 6, true>; CALL; }        \
      else if (_sw == 2 && _w) { using SWP = MSweeps<K, VT, YT, 2, 6, false>; CALL; }         \
      else if (_sw == 2 && _nt) { using SWP = MSweeps<K, VT, YT, 2, 2, true>; CALL; }         \
      else if (_sw == 2) { using SWP = MSweeps<K, VT, YT, 2, 2, false>; CALL; }               \
      else if (_w && _nt) { using SWP = MSweeps<K, VT, YT, 3, 6, true>; CALL; }               \
      else if (_w) { using SWP = MSweeps<K, VT, YT, 3, 6, false>; CALL; }                     \
      else if (_nt) { using SWP = MSweeps<K, VT, YT, 3, 3, true>; CALL; }                     \
      else { using SWP = MSweeps<K, VT, YT, 3, 3, false>; CALL; }                             \
    }                                                                           \
  } while (0)

template <int K>
struct MOps {
  static void resid(const iccg_csr_s &A, MSet &s, int mode, double tol, int max_iter, int mask,
                    hipStream_t st) {
    const DevCsr &D = A.d;
    const int g = slice_grid(D.sell.nslice);
    MDISPATCH_A(D, km_resid, K, dim3(g), dim3(kWG), 0, st, D, s.X, s.B, s.R, s.part, mode, mask);
    hipLaunchKernelGGL(km_fin_resid<K>, dim3(1), dim3(kMFinWG), 0, st, s.part, g, s.S, mode, tol,
                       max_iter, mask);
  }

  // z0 = M^-1 r0, rz_old = r.z, p = z0 (cg_crs.cpp:112-115) for the members
  // still running
  static void precond_init(const iccg_precond_s &M, MSet &s, long nk, hipStream_t st) {
    const Factor &F = M.F;
    const bool fb = F.fb;
    MDISPATCH_SEC(K, F, M.mixed, {
      const VT *lv = M.mixed ? (const VT *)F.ls.v32 : (const VT *)F.ls.v64;
      const VT *uv = M.mixed ? (const VT *)F.us.v32 : (const VT *)F.us.v64;
      const VT *dg = M.mixed ? (const VT *)F.d32 : (const VT *)F.d64;
      SWP::fwd(M.sc, F, fb, lv, uv, dg, s.R, (YT *)s.Y, (YT *)s.Z, s.S, nullptr, s.part, st);
      SWP::bwd(M.sc, F, fb, uv, dg, (const YT *)s.Y, (YT *)s.Z, s.R, s.S, nullptr, 0, st);
      const int gq = quad_grid(nk);
      hipLaunchKernelGGL((km_rz<K, YT>), dim3(gq), dim3(kWG), 0, st, nk, s.R, (const YT *)s.Z,
                         s.S, s.part);
      hipLaunchKernelGGL(km_fin_rz<K>, dim3(1), dim3(kMFinWG), 0, st, s.part, gq, s.S, 1);
      hipLaunchKernelGGL((km_update_p<K, YT>), dim3(gq), dim3(kWG), 0, st, nk, (const YT *)s.Z,
                         s.P, s.S, 1);
    });
  }

  // one CG iteration of every member (the single-system sectioned
  // iteration: SpMV + dots, alpha, forward sweep with the r update, backward
  // sweep with r.z, stop test and beta, x and p update)
  // mark(c) before each launch group of kernel class c (0 SpMV, 1 forward
  // sweep, 2 backward sweep, 3 finish and vector work), mark(-1) at the end
  template <class Mark>
  static void iteration(const iccg_csr_s &A, const iccg_precond_s &M, MSet &s, long nk,
                        hipStream_t st, Mark mark) {
    const DevCsr &D = A.d;
    const Factor &F = M.F;
    const int ga = slice_grid(D.sell.nslice);
    mark(0);
    MDISPATCH_A(D, km_spmv_cg, K, dim3(ga), dim3(kWG), 0, st, D, s.P, s.AP, s.S, s.part);
    mark(3);
    hipLaunchKernelGGL(km_fin_alpha<K>, dim3(1), dim3(kMFinWG), 0, st, s.part, ga, s.S);
    int G = 0;
    const std::vector<int> g = color_groups(M.sc, G);
    const bool fb = F.fb;
    const int slot0 = fb ? g[(size_t)M.sc.nc - 1] : 0;
    double *prz = s.part + 2 * (size_t)K * G;
    MDISPATCH_SEC(K, F, M.mixed, {
      const VT *lv = M.mixed ? (const VT *)F.ls.v32 : (const VT *)F.ls.v64;
      const VT *uv = M.mixed ? (const VT *)F.us.v32 : (const VT *)F.us.v64;
      const VT *dg = M.mixed ? (const VT *)F.d32 : (const VT *)F.d64;
      mark(1);
      SWP::fwd(M.sc, F, fb, lv, uv, dg, s.R, (YT *)s.Y, (YT *)s.Z, s.S, s.AP, s.part, st);
      mark(2);
      SWP::bwd(M.sc, F, fb, uv, dg, (const YT *)s.Y, (YT *)s.Z, s.R, s.S, prz, slot0, st);
      mark(3);
      hipLaunchKernelGGL(km_fin_stop<K>, dim3(1), dim3(kMFinWG), 0, st, s.part, G, prz, G, s.S);
      hipLaunchKernelGGL((km_update_xp<K, YT>), dim3(quad_grid(nk)), dim3(kWG), 0, st, nk, s.X,
                         (const YT *)s.Z, s.P, s.S);
    });
    mark(-1);
  }
};

}  // namespace iccg

using namespace iccg;

struct iccg_msolver_s {
  int device = 0;
  iccg_csr_t A = nullptr;
  iccg_precond_t M = nullptr;
  int n = 0, nrhs = 0;
  hipStream_t st = nullptr;
  std::vector<MSet> sets;
  // right-hand sides outside the sets (a lone member, or every member when
  // the plan is not the sets'): the single-system solver, from host copies
  std::vector<int> lone;
  iccg_solver_t single = nullptr;
  std::vector<double> hb, hx;  // nrhs * n: b and x0 / x of the lone members
  double *stage = nullptr;     // pinned staging, 4 n doubles
  // per-class event timing (kernel-by-kernel launches while on)
  bool timing = false;
  long t_launch[4] = {0, 0, 0, 0};
  double t_ms[4] = {0, 0, 0, 0};
  std::vector<hipEvent_t> pool;
  size_t used = 0;  // events of pool recorded since the last harvest
  struct Mark {
    int cls, iter;  // class (-1: end of an iteration), iteration in the chunk
    hipEvent_t ev;
  };
  std::vector<Mark> marks;
  ~iccg_msolver_s() {
    (void)hipSetDevice(device);
    if (st) (void)hipStreamSynchronize(st);
    for (auto e : pool) (void)hipEventDestroy(e);
    for (auto &s : sets) s.release();
    if (single) iccg_solver_destroy(single);
    hfree(stage);
    if (st) (void)hipStreamDestroy(st);
  }
};

namespace iccg {

// the set's scalars to the host; the wait sleeps (the sets of one solver,
// and the serving loop's runners, wait on host threads of their own)
static void mset_sync(iccg_msolver_s &S, MSet &s) {
  HIPCHK(hipMemcpyAsync(s.hS, s.S, sizeof(Scal) * s.K, hipMemcpyDeviceToHost, s.st));
  stream_wait(s.st, s.wait_ev);
}

// k iterations: a replayed hipGraph, or (timing) launched one by one with an
// event at every class boundary (mset_harvest adds the classes' times up)
static void mset_iterations(iccg_msolver_s &S, MSet &s, int k) {
  const long nk = (long)S.n * s.K;
  if (S.timing) {
    for (int j = 0; j < k; ++j) {
      auto mark = [&](int c) {
        if (S.used == S.pool.size()) {
          hipEvent_t e;
          HIPCHK(hipEventCreate(&e));
          S.pool.push_back(e);
        }
        hipEvent_t e = S.pool[S.used++];
        HIPCHK(hipEventRecord(e, s.st));
        S.marks.push_back({c, j, e});
      };
      if (s.K == 4)
        MOps<4>::iteration(*S.A, *S.M, s, nk, s.st, mark);
      else
        MOps<2>::iteration(*S.A, *S.M, s, nk, s.st, mark);
    }
    return;
  }
  auto enqueue = [&] {
    auto none = [](int) {};
    for (int j = 0; j < k; ++j) {
      if (s.K == 4)
        MOps<4>::iteration(*S.A, *S.M, s, nk, s.st, none);
      else
        MOps<2>::iteration(*S.A, *S.M, s, nk, s.st, none);
    }
  };
  if (s.graph_version != S.M->plan_version) {
    s.drop_graphs();
    s.graph_version = S.M->plan_version;
  }
  hipGraphExec_t exec = nullptr;
  for (auto &e : s.execs)
    if (e.first == k) exec = e.second;
  if (!exec) {
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(s.st, hipStreamCaptureModeThreadLocal));
    enqueue();
    HIPCHK(hipStreamEndCapture(s.st, &g));
    hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHK(e);
    s.execs.push_back({k, exec});
  }
  HIPCHK(hipGraphLaunch(exec, s.st));
}

// the timed chunk's class times, for its first `useful` iterations (those in
// which a member still ran); the stream has been synchronized
static void mset_harvest(iccg_msolver_s &S, int useful) {
  for (size_t i = 0; i + 1 < S.marks.size(); ++i) {
    const auto &m = S.marks[i];
    if (m.cls < 0 || m.iter >= useful) continue;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, m.ev, S.marks[i + 1].ev) == hipSuccess) {
      S.t_ms[m.cls] += ms;
      S.t_launch[m.cls]++;
    }
  }
  S.marks.clear();
  S.used = 0;
}

static void mset_resid(iccg_msolver_s &S, MSet &s, int mode, double tol, int max_iter, int mask) {
  if (s.K == 4)
    MOps<4>::resid(*S.A, s, mode, tol, max_iter, mask, s.st);
  else
    MOps<2>::resid(*S.A, s, mode, tol, max_iter, mask, s.st);
}

static void mset_precond_init(iccg_msolver_s &S, MSet &s) {
  const long nk = (long)S.n * s.K;
  if (s.K == 4)
    MOps<4>::precond_init(*S.M, s, nk, s.st);
  else
    MOps<2>::precond_init(*S.M, s, nk, s.st);
}

// The CG phase of one set: run_solver's loop for K members at once.  Chunks
// of captured iterations run until every member has stopped; with
// ICCG_SOLVE_TRUE_RESID the members that stopped converged on the recurrence
// but miss tol on the FP64 true residual restart from it (each member's
// sequence of operations is the one it has alone: a stopped member's kernels
// leave its state untouched until it restarts).
static void mset_run(iccg_msolver_s &S, MSet &s, int max_iter, double tol, int flags,
                     iccg_cg_result *res) {
  const int K = s.K, all = (1 << K) - 1;
  const bool true_resid = (flags & ICCG_SOLVE_TRUE_RESID) != 0;
  auto t0 = std::chrono::steady_clock::now();
  if (flags & ICCG_SOLVE_ZERO_X0)
    HIPCHK(hipMemsetAsync(s.X, 0, sizeof(double) * (size_t)S.n * K, s.st));
  mset_resid(S, s, 1, tol, max_iter, all);
  mset_precond_init(S, s);
  HIPCHK(hipGetLastError());
  if (S.timing) mset_sync(S, s);  // iteration counts before the first chunk
  int restarts[4] = {0, 0, 0, 0};
  int cap = S.n < 100000 ? 8 : 16, chunk = cap, issued = 0;
  double rel_prev = -1.0;
  int it_prev = 0;
  for (;;) {
    int before[4] = {0, 0, 0, 0};
    for (int m = 0; m < K; ++m) before[m] = s.hS[m].iters;
    mset_iterations(S, s, chunk);
    issued += chunk;
    HIPCHK(hipGetLastError());
    mset_sync(S, s);
    if (S.timing) {
      int ran = 0;  // iterations of the chunk in which some member ran
      for (int m = 0; m < K; ++m) ran = std::max(ran, s.hS[m].iters - before[m]);
      mset_harvest(S, ran);
    }
    bool done = true;
    double rel = 0.0;
    int it = 0;
    for (int m = 0; m < K; ++m)
      if (!s.hS[m].done) {
        done = false;
        rel = std::max(rel, s.hS[m].rel);
        it = std::max(it, s.hS[m].iters);
      }
    if (done) {
      int mask = 0;
      if (true_resid) {
        bool any = false;
        for (int m = 0; m < K; ++m)
          any |= s.hS[m].converged && !s.hS[m].breakdown && restarts[m] < 100;
        if (any) {
          mset_resid(S, s, 0, tol, max_iter, all);
          mset_sync(S, s);
          for (int m = 0; m < K; ++m) {
            const Scal &h = s.hS[m];
            if (!(h.converged && !h.breakdown && restarts[m] < 100)) continue;
            const double trel = std::sqrt(h.rr_true) / h.normb;
            if (trel >= tol && h.iters < max_iter) {
              mask |= 1 << m;
              ++restarts[m];
            }
          }
        }
      }
      if (!mask) break;
      mset_resid(S, s, 2, tol, max_iter, mask);
      mset_precond_init(S, s);
      rel_prev = -1.0;
      continue;
    }
    if (cap < 64) cap *= 2;
    chunk = cap;
    // the slowest running member's rate of decrease bounds the next chunk
    if (rel_prev > 0.0 && rel > 0.0 && rel < rel_prev && it > it_prev && tol > 0.0) {
      const double rate = (std::log(rel) - std::log(rel_prev)) / (it - it_prev);
      const double left = (std::log(tol) - std::log(rel)) / rate;
      int want = 4;
      while (want < cap && want < left + 1.0) want *= 2;
      chunk = want;
    }
    rel_prev = rel;
    it_prev = it;
    (void)issued;
  }
  auto t1 = std::chrono::steady_clock::now();
  const double ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  // ||b - A x|| / ||b|| per member (cg_crs.cpp:517-528), outside the CG phase
  mset_resid(S, s, 0, tol, max_iter, all);
  mset_sync(S, s);
  for (int m = 0; m < s.m; ++m) {
    const Scal &h = s.hS[m];
    iccg_cg_result &r = res[s.first + m];
    r.iters = h.iters;
    r.rel_resid = h.rel;
    r.converged = h.converged;
    r.breakdown = h.breakdown;
    r.restarts = restarts[m];
    r.solve_ms = ms;
    r.true_rel_resid = std::sqrt(h.rr_true) / h.normb;
  }
}

// host b (or x) of the set's members -> interleaved device vector
static void mset_upload(iccg_msolver_s &S, MSet &s, const double *h, double *d) {
  const int n = S.n, K = s.K;
  double *st = S.stage;
  for (int i = 0; i < n; ++i)
    for (int m = 0; m < K; ++m)
      st[(size_t)i * K + m] = (m < s.m && h) ? h[(size_t)(s.first + m) * n + i] : 0.0;
  HIPCHK(hipMemcpyAsync(d, st, sizeof(double) * (size_t)n * K, hipMemcpyHostToDevice, S.st));
  HIPCHK(hipStreamSynchronize(S.st));
}

}  // namespace iccg

extern "C" {

int iccg_msolver_create(iccg_csr_t A, iccg_precond_t M, int nrhs, iccg_msolver_t *out) {
  return guard([&] {
    require(A && M && out, ICCG_ERR_ARG, "null argument");
    require(nrhs >= 1, ICCG_ERR_ARG, "nrhs must be at least 1");
    require(M->n == A->d.n, ICCG_ERR_ARG, "preconditioner size does not match the matrix");
    HIPCHK(hipSetDevice(A->device));
    std::unique_ptr<iccg_msolver_s> S(new iccg_msolver_s);
    S->device = A->device;
    S->A = A;
    S->M = M;
    S->nrhs = nrhs;
    const int n = A->d.n;
    S->n = n;
    HIPCHK(hipStreamCreateWithFlags(&S->st, hipStreamNonBlocking));
    // sets of 4, a remainder of 3 as a set of 4 with one idle member, 2 as a
    // set of 2, 1 alone
    const bool native = multi_plan(*A, *M) && nrhs >= 2;
    int k = 0;
    while (native && nrhs - k >= 2) {
      const int left = nrhs - k;
      MSet s;
      s.K = left >= 3 ? 4 : 2;
      s.m = std::min(left, 4);
      s.first = k;
      k += s.m;
      S->sets.push_back(s);
    }
    for (; k < nrhs; ++k) S->lone.push_back(k);
    const size_t vb = M->mixed ? sizeof(float) : sizeof(double);
    for (MSet &s : S->sets) {
      const size_t nk = (size_t)n * s.K;
      HIPCHK(hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking));
      s.wait_ev = blocking_event();
      s.X = dalloc<double>(nk);
      s.R = dalloc<double>(nk);
      s.P = dalloc<double>(nk);
      s.AP = dalloc<double>(nk);
      s.B = dalloc<double>(nk);
      s.Y = dalloc<char>(vb * nk);
      s.Z = dalloc<char>(vb * nk);
      s.part = dalloc<double>(2 * mset_part_pairs(*A, *M, s.K));
      s.S = dalloc<Scal>(4);
      HIPCHK(hipMemsetAsync(s.S, 0, sizeof(Scal) * 4, S->st));
      HIPCHK(hipMemsetAsync(s.X, 0, sizeof(double) * nk, S->st));
      HIPCHK(hipMemsetAsync(s.B, 0, sizeof(double) * nk, S->st));
      s.hS = halloc<Scal>(4);
    }
    if (!S->sets.empty())
      S->stage = halloc<double>(4 * (size_t)std::max(n, 1));
    if (!S->lone.empty()) {
      int st = iccg_solver_create(A, M, &S->single);
      if (st == ICCG_OK) st = iccg_solver_set_blocking(S->single, 1);
      if (st != ICCG_OK) throw Error(st, iccg_last_error());
      S->hb.assign((size_t)nrhs * n, 0.0);
      S->hx.assign((size_t)nrhs * n, 0.0);
    }
    HIPCHK(hipStreamSynchronize(S->st));
    *out = S.release();
  });
}

int iccg_msolver_destroy(iccg_msolver_t S) {
  delete S;
  return ICCG_OK;
}

int iccg_msolver_native(iccg_msolver_t S, int *sets, int *lone) {
  return guard([&] {
    require(S && sets && lone, ICCG_ERR_ARG, "null argument");
    *sets = (int)S->sets.size();
    *lone = (int)S->lone.size();
  });
}

int iccg_msolver_set(iccg_msolver_t S, const double *b, const double *x0) {
  return guard([&] {
    require(S && b, ICCG_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(S->device));
    for (MSet &s : S->sets) {
      mset_upload(*S, s, b, s.B);
      mset_upload(*S, s, x0, s.X);
    }
    const size_t n = (size_t)S->n;
    for (int k : S->lone) {
      std::copy(b + k * n, b + (k + 1) * n, S->hb.begin() + k * n);
      if (x0)
        std::copy(x0 + k * n, x0 + (k + 1) * n, S->hx.begin() + k * n);
      else
        std::fill(S->hx.begin() + k * n, S->hx.begin() + (k + 1) * n, 0.0);
    }
  });
}

int iccg_msolver_run(iccg_msolver_t S, int max_iter, double tol, int flags,
                     iccg_cg_result *res) {
  return guard([&] {
    require(S && res, ICCG_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(S->device));
    require(S->sets.empty() || multi_plan(*S->A, *S->M), ICCG_ERR_STATE,
            "the preconditioner's plan changed since the solver was made: make it again");
    // the sets, and the right-hand sides run alone, are independent: each
    // runs on a host thread of its own (its own stream), so one request's
    // sets share the device instead of queueing behind each other (event
    // timing runs them one after another)
    std::vector<std::function<void()>> units;
    for (MSet &s : S->sets)
      units.push_back([&, ps = &s] { mset_run(*S, *ps, max_iter, tol, flags, res); });
    if (!S->lone.empty())
      units.push_back([&] {
        const size_t n = (size_t)S->n;
        auto ck = [](int st) {
          if (st != ICCG_OK) throw Error(st, iccg_last_error());
        };
        for (int k : S->lone) {
          ck(iccg_solver_set(S->single, S->hb.data() + k * n, S->hx.data() + k * n));
          ck(iccg_solver_run(S->single, max_iter, tol, flags, &res[k]));
          ck(iccg_solver_get_x(S->single, S->hx.data() + k * n));
        }
      });
    if (S->timing || units.size() == 1) {
      for (auto &u : units) u();
    } else {
      std::vector<std::exception_ptr> errs(units.size());
      std::vector<std::thread> th;
      const int dev = S->device;
      for (size_t u = 1; u < units.size(); ++u)
        th.emplace_back([&, u] {
          try {
            HIPCHK(hipSetDevice(dev));
            units[u]();
          } catch (...) {
            errs[u] = std::current_exception();
          }
        });
      try {
        units[0]();
      } catch (...) {
        errs[0] = std::current_exception();
      }
      for (auto &t : th) t.join();
      for (auto &e : errs)
        if (e) std::rethrow_exception(e);
    }
  });
}

int iccg_msolver_set_timing(iccg_msolver_t S, int on) {
  return guard([&] {
    require(S, ICCG_ERR_ARG, "null argument");
    S->timing = on != 0;
  });
}

int iccg_msolver_get_timing(iccg_msolver_t S, int cls, long *launches, double *total_ms) {
  return guard([&] {
    require(S && cls >= 0 && cls < 4 && launches && total_ms, ICCG_ERR_ARG, "bad argument");
    *launches = S->t_launch[cls];
    *total_ms = S->t_ms[cls];
  });
}

int iccg_msolver_reset_timing(iccg_msolver_t S) {
  return guard([&] {
    require(S, ICCG_ERR_ARG, "null argument");
    for (int c = 0; c < 4; ++c) {
      S->t_launch[c] = 0;
      S->t_ms[c] = 0.0;
    }
  });
}

int iccg_msolver_get_x(iccg_msolver_t S, double *x) {
  return guard([&] {
    require(S && x, ICCG_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(S->device));
    const int n = S->n;
    for (MSet &s : S->sets) {
      const int K = s.K;
      HIPCHK(hipMemcpyAsync(S->stage, s.X, sizeof(double) * (size_t)n * K,
                            hipMemcpyDeviceToHost, S->st));
      HIPCHK(hipStreamSynchronize(S->st));
      for (int m = 0; m < s.m; ++m) {
        double *xm = x + (size_t)(s.first + m) * n;
        for (int i = 0; i < n; ++i) xm[i] = S->stage[(size_t)i * K + m];
      }
    }
    for (int k : S->lone)
      std::copy(S->hx.begin() + (size_t)k * n, S->hx.begin() + (size_t)(k + 1) * n,
                x + (size_t)k * n);
  });
}

int iccg_cg_solve_multi(iccg_csr_t A, iccg_precond_t M, int nrhs, const double *b, double *x,
                        int max_iter, double tol, int flags, iccg_cg_result *res) {
  iccg_msolver_t S = nullptr;
  int st = iccg_msolver_create(A, M, nrhs, &S);
  if (st != ICCG_OK) return st;
  st = iccg_msolver_set(S, b, x);
  if (st == ICCG_OK) st = iccg_msolver_run(S, max_iter, tol, flags, res);
  if (st == ICCG_OK) st = iccg_msolver_get_x(S, x);
  iccg_msolver_destroy(S);
  return st;
}

}  // extern "C"
