// batch.hip -- multi-system ICCG: many independent SPD systems, each with
// its own matrix, IC(0) factor and Krylov vectors resident in HBM, advanced
// together so that every CG phase is ONE launch over all active systems.
//
// Every right-hand side is a slot with its own vectors and scalars (SysDesc,
// SysScal).  Slots of the same matrix and preconditioner form a group of up
// to kMaxRhs right-hand sides; the kernels read each matrix and factor entry
// once for the whole group.  Work is described by item lists:
//   ChunkItem (group, s0, s1): 64-row slices [s0, s1), one workgroup --
//     SpMV, vector updates and dot-product partials (one plane per group
//     member);
//   ColorItem (group, a0, a1): one wavefront of a triangular sweep over one
//     color -- 64-row slices (multicolor) or blocks (ABMC);
//   FinItem (slot, chunk range, plane): a finish kernel's workgroup folds
//     one slot's partials and applies its scalar logic.
// Converged slots skip their work, and slots can be admitted or retired
// between launches: the serving loop's continuous batching.  Systems keep
// their own precision (FP64, or FP32 factor with FP64 CG) and ordering
// (multicolor or ABMC); per right-hand side, the arithmetic is the
// single-system solver's -- same row chains, same double-double dots -- so
// it gives the same iterates in a batch as alone.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "dev.hpp"
#include "devfun.hpp"
#include "internal.hpp"
#include "objects.hpp"

namespace iccg {

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__device__ __forceinline__ double *plane(double *part, int j, int nch, int nc) {
  return part + (size_t)j * nch * nc * 2;
}

// the item's group: one-member groups come from the item itself
template <int KM, class Item>
__device__ __forceinline__ GroupDesc group_of(const GroupDesc *__restrict__ groups,
                                              const Item &c) {
  GroupDesc g;
  if (KM == 1) {
    g.k = 1;
    g.slot[0] = c.slot0;
  } else {
    g = groups[c.grp];
  }
  return g;
}

// A x for every member of the group over the item's slices (A's entries
// loaded once per row and pair of members, from the f32 copy when every
// value is an exact float): mode 0/1/2 = true residual / start / restart
// (r = b - A x with b.b, r.r), mode 3 = CG SpMV (Ap = A p with p.Ap, p.p,
// Ap.Ap).  Members are taken two at a time (the double-double partials of
// more would cost occupancy); the later pairs find the item's entries in
// cache.
template <class VT, int WM, int KM, bool PV>
__device__ __forceinline__ void ax_group(const GroupDesc &g, const SysDesc *__restrict__ desc,
                                         const SysScal *__restrict__ sc, const ChunkItem &c,
                                         const VT *__restrict__ av, double *part, int nch,
                                         int mode) {
  constexpr int KS = KM < 2 ? KM : 2;
  const SysDesc &D0 = desc[g.slot[0]];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int j0 = 0; j0 < KM; j0 += KS) {
    double h[KS][3], l[KS][3];
    bool live[KS], any = false;
#pragma unroll
    for (int q = 0; q < KS; ++q) {
      const int j = j0 + q;
      live[q] = j < g.k && (mode != 3 || !sc[g.slot[j]].done);
      any |= live[q];
#pragma unroll
      for (int t = 0; t < 3; ++t) h[q][t] = l[q][t] = 0.0;
    }
    if (!any) continue;  // uniform over the workgroup
    for (int s = c.s0 + w; s < c.s1; s += kWG >> 6) {
      int cc[WM];
      VT vv[WM];
      if (PV)
        sell_load_pairv<VT, WM>(D0.a_off, D0.a_col, av, s, lane, cc, vv);
      else
        sell_load<VT, WM, true>(D0.a_off, D0.a_col, av, s, lane, cc, vv);
      const int i = (s << 6) + lane;
#pragma unroll
      for (int q = 0; q < KS; ++q) {
        if (!live[q]) continue;
        const SysDesc &D = desc[g.slot[j0 + q]];
        if (mode == 3) {
          const double y = sell_apply_sel<double, double, VT, WM, false>(cc, vv, D.p, 0.0);
          if (i < D0.n) {
            __builtin_nontemporal_store(y, &D.Ap[i]);  // read once, by kb_xr
            const double pi = D.p[i];
            dd_fma(h[q][0], l[q][0], pi, y);
            dd_fma(h[q][1], l[q][1], pi, pi);
            dd_fma(h[q][2], l[q][2], y, y);
          }
        } else {
          const double ax = sell_apply_sel<double, double, VT, WM, false>(cc, vv, D.x, 0.0);
          if (i < D0.n) {
            const double bi = D.b[i], ri = bi - ax;
            if (mode != 0) D.r[i] = ri;
            dd_fma(h[q][0], l[q][0], bi, bi);
            dd_fma(h[q][1], l[q][1], ri, ri);
          }
        }
      }
    }
#pragma unroll
    for (int q = 0; q < KS; ++q) {
      if (!live[q]) continue;
      if (mode == 3) {
        block_partial<3>(h[q], l[q], plane(part, j0 + q, nch, 3), blockIdx.x);
      } else {
        double h2[2] = {h[q][0], h[q][1]}, l2[2] = {l[q][0], l[q][1]};
        block_partial<2>(h2, l2, plane(part, j0 + q, nch, 2), blockIdx.x);
      }
    }
  }
}

template <int WM, int KM>
__global__ __launch_bounds__(kWG) void kb_ax(const GroupDesc *__restrict__ groups,
                                             const SysDesc *__restrict__ desc,
                                             const ChunkItem *__restrict__ items,
                                             const SysScal *__restrict__ sc, double *part,
                                             int nch, int mode) {
  const ChunkItem c = items[blockIdx.x];
  const GroupDesc g = group_of<KM>(groups, c);
  if (mode == 3) {
    // A group whose members all stopped is not touched: its matrix may
    // already be freed (the running list keeps retired groups until it is
    // rebuilt).
    bool any = false;
    for (int j = 0; j < g.k; ++j) any |= !sc[g.slot[j]].done;
    if (!any) return;
  }
  const SysDesc &D0 = desc[g.slot[0]];
  if (D0.a_v32)
    ax_group<float, WM, KM, false>(g, desc, sc, c, D0.a_v32, part, nch, mode);
  else if (D0.a_paired)
    ax_group<double, WM, KM, true>(g, desc, sc, c, D0.a_v64, part, nch, mode);
  else
    ax_group<double, WM, KM, false>(g, desc, sc, c, D0.a_v64, part, nch, mode);
}

// r -= alpha Ap (cg_crs.cpp:141); r.r partial -- per member.  x += alpha p
// (line 139) runs in kb_update_xp, which reads p anyway.
template <int KM>
__global__ __launch_bounds__(kWG) void kb_r(const GroupDesc *__restrict__ groups,
                                            const SysDesc *__restrict__ desc,
                                            const ChunkItem *__restrict__ items,
                                            const SysScal *__restrict__ sc, double *part,
                                            int nch) {
  const ChunkItem c = items[blockIdx.x];
  const GroupDesc g = group_of<KM>(groups, c);
  for (int j = 0; j < (KM == 1 ? 1 : g.k); ++j) {
    const int id = g.slot[j];
    if (sc[id].done) continue;
    const SysDesc &D = desc[id];
    const double alpha = sc[id].alpha;
    const int end = min(D.n, c.s1 << 6);
    double h[1] = {0}, l[1] = {0};
    for (int i = (c.s0 << 6) + threadIdx.x; i < end; i += kWG) {
      // Ap: read by no other kernel of the iteration (nontemporal)
      const double ri = fma(-alpha, ldm<true>(&D.Ap[i]), D.r[i]);
      D.r[i] = ri;
      dd_fma(h[0], l[0], ri, ri);
    }
    block_partial<1>(h, l, plane(part, j, nch, 1), blockIdx.x);
  }
}

template <class FT>
__device__ __forceinline__ void rz_rows(const SysDesc &D, int beg, int end, double (&h)[1],
                                        double (&l)[1]) {
  const FT *z = (const FT *)D.z;
  for (int i = beg; i < end; i += kWG) dd_fma(h[0], l[0], D.r[i], (double)z[i]);
}

template <int KM>
__global__ __launch_bounds__(kWG) void kb_rz(const GroupDesc *__restrict__ groups,
                                             const SysDesc *__restrict__ desc,
                                             const ChunkItem *__restrict__ items,
                                             const SysScal *__restrict__ sc, double *part,
                                             int nch) {
  const ChunkItem c = items[blockIdx.x];
  const GroupDesc g = group_of<KM>(groups, c);
  for (int j = 0; j < (KM == 1 ? 1 : g.k); ++j) {
    const int id = g.slot[j];
    if (sc[id].done) continue;
    const SysDesc &D = desc[id];
    const int beg = (c.s0 << 6) + threadIdx.x, end = min(D.n, c.s1 << 6);
    double h[1] = {0}, l[1] = {0};
    if (D.mixed)
      rz_rows<float>(D, beg, end, h, l);
    else
      rz_rows<double>(D, beg, end, h, l);
    block_partial<1>(h, l, plane(part, j, nch, 1), blockIdx.x);
  }
}

template <class FT>
__device__ __forceinline__ void p_rows(const SysDesc &D, int beg, int end, double beta,
                                       int init) {
  const FT *z = (const FT *)D.z;
  for (int i = beg; i < end; i += kWG) {
    const double zi = (double)z[i];
    D.p[i] = init ? zi : fma(beta, D.p[i], zi);
  }
}

template <int KM>
__global__ __launch_bounds__(kWG) void kb_update_p(const GroupDesc *__restrict__ groups,
                                                   const SysDesc *__restrict__ desc,
                                                   const ChunkItem *__restrict__ items,
                                                   const SysScal *__restrict__ sc, int init) {
  const ChunkItem c = items[blockIdx.x];
  const GroupDesc g = group_of<KM>(groups, c);
  for (int j = 0; j < (KM == 1 ? 1 : g.k); ++j) {
    const int id = g.slot[j];
    if (sc[id].done) continue;
    const SysDesc &D = desc[id];
    const double beta = sc[id].beta;
    const int beg = (c.s0 << 6) + threadIdx.x, end = min(D.n, c.s1 << 6);
    if (D.mixed)
      p_rows<float>(D, beg, end, beta, init);
    else
      p_rows<double>(D, beg, end, beta, init);
  }
}

// x += alpha p (cg_crs.cpp:139) and p = z + beta p (155-157) -- per member:
// for every slot whose x/r update ran this iteration (not after a breakdown,
// not once it had stopped, done == 2), including the one whose stop test
// ended it (done == 1), which updates no p
template <class FT>
__device__ __forceinline__ void xp_rows(const SysDesc &D, int beg, int end, double alpha,
                                        double beta, bool pu) {
  const FT *z = (const FT *)D.z;
  for (int i = beg; i < end; i += kWG) {
    const double pi = D.p[i];
    // x: read by no other kernel of the iteration (nontemporal)
    __builtin_nontemporal_store(fma(alpha, pi, ldm<true>(&D.x[i])), &D.x[i]);
    if (pu) D.p[i] = fma(beta, pi, (double)z[i]);
  }
}

template <int KM>
__global__ __launch_bounds__(kWG) void kb_update_xp(const GroupDesc *__restrict__ groups,
                                                    const SysDesc *__restrict__ desc,
                                                    const ChunkItem *__restrict__ items,
                                                    const SysScal *__restrict__ sc) {
  const ChunkItem c = items[blockIdx.x];
  const GroupDesc g = group_of<KM>(groups, c);
  for (int j = 0; j < (KM == 1 ? 1 : g.k); ++j) {
    const int id = g.slot[j];
    const SysScal &S = sc[id];
    if (S.done > 1 || S.breakdown) continue;
    const SysDesc &D = desc[id];
    const int beg = (c.s0 << 6) + threadIdx.x, end = min(D.n, c.s1 << 6);
    if (D.mixed)
      xp_rows<float>(D, beg, end, S.alpha, S.beta, !S.done);
    else
      xp_rows<double>(D, beg, end, S.alpha, S.beta, !S.done);
  }
}

// ---- triangular sweeps (ic0.cpp:194-249 multicolor, 263-339 ABMC) --------
// multicolor: the row's factor entries are loaded once, then every member's
// chain runs on them
template <class FT, int WM, int KM, bool FWD>
__device__ __forceinline__ void mc_item(const GroupDesc &g, const SysDesc *__restrict__ desc,
                                        const bool (&live)[KM], const ColorItem &c, int lane) {
  const SysDesc &D0 = desc[g.slot[0]];
  const int *off = FWD ? D0.l_off : D0.u_off;
  const int *col = FWD ? D0.l_col : D0.u_col;
  const FT *fv = (const FT *)(FWD ? D0.l_v : D0.u_v);
  const FT *dg = (const FT *)D0.d;
  for (int s = c.a0; s < c.a1; ++s) {
    const int i = (s << 6) + lane;
    if (i < c.lo || i >= c.hi) continue;
    int cc[WM];
    FT vv[WM];
    sell_load<FT, WM>(off, col, fv, s, lane, cc, vv);
    const FT d = dg[i];
#pragma unroll
    for (int j = 0; j < KM; ++j) {
      if (!live[j]) continue;
      const SysDesc &D = desc[g.slot[j]];
      if (FWD) {
        FT *y = (FT *)D.y;
        y[i] = sell_apply<FT, FT, FT, WM, true>(cc, vv, y, (FT)D.r[i]) / d;
      } else {
        FT *z = (FT *)D.z;
        z[i] = sell_apply<FT, FT, FT, WM, true>(cc, vv, z, ((const FT *)D.y)[i]) / d;
      }
    }
  }
}

// ABMC blocks of at most 64 rows (contiguous new ids [bs, be), lane t owns
// row bs + t), all live members in lock-step: the block's factor entries
// are loaded once, and every in-block level step advances each member's
// chain, so the level latency is paid once for the group.  Per member and
// row the operations and their order are abmc_fwd_block's /
// abmc_bwd_block's: forward, the terms of earlier colors (columns < bs)
// first, then the in-block terms as their levels complete; backward, the
// in-block terms first, then the later colors' (columns >= be, prefetched).
// my holds one 64-entry LDS window per member.
template <class FT, int WM, int KM>
__device__ __forceinline__ void abmc_fwd_group(int bs, int be, int nl, const SysDesc &D0,
                                               const GroupDesc &g,
                                               const SysDesc *__restrict__ desc,
                                               const bool (&live)[KM], FT *my, int lane) {
  const int i = bs + lane;
  const bool act = i < be;
  int cc[WM];
  FT vv[WM];
  FT acc[KM], d = (FT)1;
  int mylev = -1;
#pragma unroll
  for (int j = 0; j < KM; ++j) acc[j] = (FT)0;
  if (act) {
    sell_load<FT, WM>(D0.l_off, D0.l_col, (const FT *)D0.l_v, i >> 6, i & 63, cc, vv);
    mylev = D0.lev_f[i];
    d = ((const FT *)D0.d)[i];
#pragma unroll
    for (int j = 0; j < KM; ++j) {
      if (!live[j]) continue;
      const SysDesc &D = desc[g.slot[j]];
      const FT *y = (const FT *)D.y;
      FT a = (FT)D.r[i];
      FT xo[WM];
#pragma unroll
      for (int k = 0; k < WM; ++k) xo[k] = (cc[k] >= 0 && cc[k] < bs) ? y[cc[k]] : (FT)0;
#pragma unroll
      for (int k = 0; k < WM; ++k)
        if (cc[k] >= 0 && cc[k] < bs) a = fma_t(-vv[k], xo[k], a);
      acc[j] = a;
    }
  }
  FT yi[KM];
#pragma unroll
  for (int j = 0; j < KM; ++j) yi[j] = (FT)0;
  for (int l = 0; l < nl; ++l) {
    if (mylev == l) {
#pragma unroll
      for (int j = 0; j < KM; ++j) {
        if (!live[j]) continue;
        FT a = acc[j];
        const FT *w = my + 64 * j;
#pragma unroll
        for (int k = 0; k < WM; ++k)
          if (cc[k] >= bs) a = fma_t(-vv[k], w[cc[k] - bs], a);
        yi[j] = a / d;
        my[64 * j + lane] = yi[j];
      }
    }
    wave_sync();
  }
  if (act) {
#pragma unroll
    for (int j = 0; j < KM; ++j)
      if (live[j]) ((FT *)desc[g.slot[j]].y)[i] = yi[j];
  }
}

template <class FT, int WM, int KM>
__device__ __forceinline__ void abmc_bwd_group(int bs, int be, int nl, const SysDesc &D0,
                                               const GroupDesc &g,
                                               const SysDesc *__restrict__ desc,
                                               const bool (&live)[KM], FT *my, int lane) {
  const int i = bs + lane;
  const bool act = i < be;
  int cc[WM];
  FT vv[WM];
  FT acc[KM], xo[KM][WM], d = (FT)1;
  int mylev = -1;
#pragma unroll
  for (int j = 0; j < KM; ++j) {
    acc[j] = (FT)0;
#pragma unroll
    for (int k = 0; k < WM; ++k) xo[j][k] = (FT)0;
  }
  if (act) {
    sell_load<FT, WM>(D0.u_off, D0.u_col, (const FT *)D0.u_v, i >> 6, i & 63, cc, vv);
    mylev = D0.lev_b[i];
    d = ((const FT *)D0.d)[i];
#pragma unroll
    for (int j = 0; j < KM; ++j) {
      if (!live[j]) continue;
      const SysDesc &D = desc[g.slot[j]];
      const FT *z = (const FT *)D.z;
#pragma unroll
      for (int k = 0; k < WM; ++k)
        if (cc[k] >= be) xo[j][k] = z[cc[k]];
      acc[j] = ((const FT *)D.y)[i];
    }
  }
  FT zi[KM];
#pragma unroll
  for (int j = 0; j < KM; ++j) zi[j] = (FT)0;
  for (int lv = 0; lv < nl; ++lv) {
    if (mylev == lv) {
#pragma unroll
      for (int j = 0; j < KM; ++j) {
        if (!live[j]) continue;
        FT a = acc[j];
        const FT *w = my + 64 * j;
#pragma unroll
        for (int k = 0; k < WM; ++k)
          if (cc[k] >= 0 && cc[k] < be) a = fma_t(-vv[k], w[cc[k] - bs], a);
#pragma unroll
        for (int k = 0; k < WM; ++k)
          if (cc[k] >= be) a = fma_t(-vv[k], xo[j][k], a);
        zi[j] = a / d;
        my[64 * j + lane] = zi[j];
      }
    }
    wave_sync();
  }
  if (act) {
#pragma unroll
    for (int j = 0; j < KM; ++j)
      if (live[j]) ((FT *)desc[g.slot[j]].z)[i] = zi[j];
  }
}

// ABMC blocks of a factor in the sectioned layout (sell_layout_sections,
// D0.sec <= SW in-block slots, at most SW outside entries per row): the
// single-system sweeps' mask-free level
// loop (sec_levels) run for every live member on the shared entries.  my
// holds one kSecWindow window per member (row slots and the zero slot).
// FB (forward, blocks of the factor's last color): each member's backward
// rows follow its forward rows at once, from the y just computed (not
// stored); the last color has no later colors, so they have in-block terms
// only, and the result is the single-system sweeps' z bit for bit.
template <class FT, int SW, int KM, bool FWD, bool FB = false>
__device__ __forceinline__ void abmc_group_sec(int kb, int bs, int be, int nl, const SysDesc &D0,
                                               const GroupDesc &g,
                                               const SysDesc *__restrict__ desc,
                                               const bool (&live)[KM], FT *my, int lane) {
  const int i = bs + lane;
  const bool act = i < be;
  const int sec = D0.sec;
  int ad[SW], co[SW], uad[SW];
  FT va[SW], vo[SW], uva[SW];
  FT acc[KM], xo[KM][SW], d = (FT)1;
#pragma unroll
  for (int k = 0; k < SW; ++k) {
    ad[k] = uad[k] = 64;
    co[k] = -1;
    va[k] = vo[k] = uva[k] = (FT)0;
  }
#pragma unroll
  for (int j = 0; j < KM; ++j) {
    acc[j] = (FT)0;
#pragma unroll
    for (int k = 0; k < SW; ++k) xo[j][k] = (FT)0;
  }
  if (act) {
    const int *off = FWD ? D0.l_off : D0.u_off;
    const int *col = FWD ? D0.l_col : D0.u_col;
    const FT *fv = (const FT *)(FWD ? D0.l_v : D0.u_v);
    const unsigned pk = ldm<true>(&(FWD ? D0.l_inb : D0.u_inb)[64 * kb + lane]);
    const int o0 = off[kb];
    const int w = (off[kb + 1] - o0) >> 6;
    const int base = o0 + lane;
#pragma unroll
    for (int k = 0; k < SW; ++k) {
      if (k < sec) {
        ad[k] = (int)((pk >> (8 * k)) & 0xFFu);  // 64 (the zero slot) when absent
        va[k] = ldm<true>(&fv[base + 64 * k]);   // absent entries hold 0
      }
      if (sec + k < w) {
        co[k] = ldm<true>(&col[base + 64 * (sec + k)]);
        vo[k] = ldm<true>(&fv[base + 64 * (sec + k)]);
      }
    }
    d = ((const FT *)D0.d)[i];
    if (FB) {  // U's in-block entries of the row (its only ones)
      const unsigned pu = ldm<true>(&D0.u_inb[64 * kb + lane]);
      const int ub = D0.u_off[kb] + lane;
#pragma unroll
      for (int k = 0; k < SW; ++k)
        if (k < sec) {
          uad[k] = (int)((pu >> (8 * k)) & 0xFFu);
          uva[k] = ldm<true>(&((const FT *)D0.u_v)[ub + 64 * k]);
        }
    }
#pragma unroll
    for (int j = 0; j < KM; ++j) {
      if (!live[j]) continue;
      const SysDesc &D = desc[g.slot[j]];
      if (FWD) {
        const FT *y = (const FT *)D.y;
        FT a = (FT)D.r[i];
        FT t[SW];
#pragma unroll
        for (int k = 0; k < SW; ++k) t[k] = co[k] >= 0 ? y[co[k]] : (FT)0;
#pragma unroll
        for (int k = 0; k < SW; ++k)
          if (co[k] >= 0) a = fma_t(-vo[k], t[k], a);
        acc[j] = a;
      } else {
        const FT *z = (const FT *)D.z;
#pragma unroll
        for (int k = 0; k < SW; ++k)
          if (co[k] >= 0) xo[j][k] = z[co[k]];
        acc[j] = ((const FT *)D.y)[i];
      }
    }
  }
  if (lane < KM) my[kSecWindow * lane + 64] = (FT)0;
  wave_sync();
#pragma unroll
  for (int j = 0; j < KM; ++j) {
    if (!live[j]) continue;
    const FT zi = sec_levels<FT, SW, SW, !FWD>(nl, acc[j], d, ad, va, vo, xo[j],
                                                 my + kSecWindow * j, lane, act);
    if (FB) {
      const FT zb = sec_levels<FT, SW, SW, false>(D0.nlev_b[kb], zi, d, uad, uva, vo, xo[j],
                                              my + kSecWindow * j, lane, act);
      if (act) ((FT *)desc[g.slot[j]].z)[i] = zb;
    } else if (act) {
      ((FT *)(FWD ? desc[g.slot[j]].y : desc[g.slot[j]].z))[i] = zi;
    }
  }
}

template <class FT, int WM, int KM, bool FWD>
__device__ __forceinline__ void abmc_item(const GroupDesc &g, const SysDesc *__restrict__ desc,
                                          const bool (&live)[KM], const ColorItem &c, int lane,
                                          FT *my) {
  const SysDesc &D0 = desc[g.slot[0]];
  for (int kb = c.a0; kb < c.a1; ++kb) {
    const int bs = D0.bstart[kb], be = D0.bstart[kb + 1];
    if (FWD)
      abmc_fwd_group<FT, WM, KM>(bs, be, D0.nlev_f[kb], D0, g, desc, live, my, lane);
    else
      abmc_bwd_group<FT, WM, KM>(bs, be, D0.nlev_b[kb], D0, g, desc, live, my, lane);
  }
}

// one wavefront per color item; every branch below is uniform per wavefront
template <int WM, int KM, bool FWD>
__global__ __launch_bounds__(kWG) void kb_sweep(const GroupDesc *__restrict__ groups,
                                                const SysDesc *__restrict__ desc,
                                                const ColorItem *__restrict__ items, int nitems,
                                                const SysScal *__restrict__ sc) {
  __shared__ double lds[kWG * KM];  // per wavefront: KM windows of 64
  const int wv = blockIdx.x * (kWG >> 6) + (threadIdx.x >> 6);
  if (wv >= nitems) return;
  const ColorItem c = items[wv];
  const GroupDesc g = group_of<KM>(groups, c);
  bool live[KM], any = false;
#pragma unroll
  for (int j = 0; j < KM; ++j) {
    live[j] = j < g.k && !sc[g.slot[j]].done;
    any |= live[j];
  }
  if (!any) return;
  const SysDesc &D0 = desc[g.slot[0]];
  const int lane = threadIdx.x & 63;
  double *my = lds + (threadIdx.x >> 6) * 64 * KM;
  if (D0.bstart) {
    if (D0.mixed)
      abmc_item<float, WM, KM, FWD>(g, desc, live, c, lane, (float *)my);
    else
      abmc_item<double, WM, KM, FWD>(g, desc, live, c, lane, my);
  } else {
    if (D0.mixed)
      mc_item<float, WM, KM, FWD>(g, desc, live, c, lane);
    else
      mc_item<double, WM, KM, FWD>(g, desc, live, c, lane);
  }
}

// one wavefront per color item of a sectioned-layout ABMC group; FB: the
// items of factors whose last color this is, forward and backward
template <int SW, int KM, bool FWD, bool FB>
__global__ __launch_bounds__(kWG) void kb_sweep_sec(const GroupDesc *__restrict__ groups,
                                                    const SysDesc *__restrict__ desc,
                                                    const ColorItem *__restrict__ items,
                                                    int nitems, const SysScal *__restrict__ sc) {
  // per wavefront: KM windows of kSecWindow (64 rows and the zero slot)
  __shared__ double lds[(kWG / 64) * kSecWindow * KM];
  const int wv = blockIdx.x * (kWG >> 6) + (threadIdx.x >> 6);
  if (wv >= nitems) return;
  const ColorItem c = items[wv];
  const GroupDesc g = group_of<KM>(groups, c);
  bool live[KM], any = false;
#pragma unroll
  for (int j = 0; j < KM; ++j) {
    live[j] = j < g.k && !sc[g.slot[j]].done;
    any |= live[j];
  }
  if (!any) return;
  const SysDesc &D0 = desc[g.slot[0]];
  const int lane = threadIdx.x & 63;
  double *my = lds + (threadIdx.x >> 6) * kSecWindow * KM;
  for (int kb = c.a0; kb < c.a1; ++kb) {
    const int bs = D0.bstart[kb], be = D0.bstart[kb + 1];
    const int nl = FWD ? D0.nlev_f[kb] : D0.nlev_b[kb];
    if (D0.mixed)
      abmc_group_sec<float, SW, KM, FWD, FB>(kb, bs, be, nl, D0, g, desc, live, (float *)my,
                                             lane);
    else
      abmc_group_sec<double, SW, KM, FWD, FB>(kb, bs, be, nl, D0, g, desc, live, my, lane);
  }
}

// finish kernels: workgroup k folds the partials of fin[k] and applies the
// single-system scalar logic (cg_crs.cpp:104-160)
template <int NC>
__device__ __forceinline__ bool fold_slot(const FinItem &f, const double *part, int nch,
                                          double (&out)[NC]) {
  return finish_reduce<NC>(part + ((size_t)f.plane * nch + f.c0) * NC * 2, f.c1 - f.c0, out);
}

__global__ __launch_bounds__(kWG) void kbf_resid(const FinItem *__restrict__ fin,
                                                 const double *__restrict__ part, int nch,
                                                 const SysDesc *__restrict__ desc, SysScal *sc,
                                                 int mode) {
  const FinItem f = fin[blockIdx.x];
  double out[2];
  if (!fold_slot<2>(f, part, nch, out)) return;
  SysScal &S = sc[f.slot];
  double normb = sqrt(out[0]);
  if (normb == 0.0) normb = 1.0;
  S.normb = normb;
  if (mode == 0) {
    S.rr_true = out[1];
    return;
  }
  const double rel = sqrt(out[1]) / normb;
  if (mode == 1) {
    S.iters = 0;
    S.tol = desc[f.slot].tol;
    S.max_iter = desc[f.slot].max_iter;
  }
  S.rel = rel;
  S.converged = 0;
  S.breakdown = 0;
  S.done = 0;
  if (rel < S.tol) {
    S.converged = 1;
    S.done = 1;
  } else if (S.iters >= S.max_iter) {
    S.done = 1;
  }
}

__global__ __launch_bounds__(kWG) void kbf_alpha(const FinItem *__restrict__ fin,
                                                 const double *__restrict__ part, int nch,
                                                 SysScal *sc) {
  const FinItem f = fin[blockIdx.x];
  SysScal &S = sc[f.slot];
  if (S.done) {  // stopped before this iteration: no x update from now on
    if (threadIdx.x == 0) S.done = 2;
    return;
  }
  double out[3];
  if (!fold_slot<3>(f, part, nch, out)) return;
  const double pAp = out[0];
  const double thr = DBL_EPSILON * sqrt(out[1]) * sqrt(out[2]);
  S.pAp = pAp;
  if (pAp <= thr) {
    S.breakdown = 1;
    S.done = 1;
  } else {
    S.alpha = S.rz_old / pAp;
  }
}

// The stop test (cg_crs.cpp:143-149) taken after the preconditioner, and
// when the iteration goes on, beta from r.z (152-153): kb_r's r.r partials
// (part) and kb_rz's r.z partials (part_rz) folded in one pass.  In the
// iteration that stops, the preconditioner and r.z ran for nothing and
// kb_update_xp applies only the x update.
__global__ __launch_bounds__(kWG) void kbf_xr_rz(const FinItem *__restrict__ fin,
                                                 const double *__restrict__ part,
                                                 const double *__restrict__ part_rz, int nch,
                                                 SysScal *sc) {
  const FinItem f = fin[blockIdx.x];
  SysScal &S = sc[f.slot];
  if (S.done) return;
  double rr[1], rz[1];
  const size_t o = ((size_t)f.plane * nch + f.c0) * 2;
  if (!finish_reduce2<1, 1>(part + o, f.c1 - f.c0, part_rz + o, f.c1 - f.c0, rr, rz)) return;
  const double rel = sqrt(rr[0]) / S.normb;
  const int it = S.iters + 1;
  S.iters = it;
  S.rel = rel;
  if (rel < S.tol) {
    S.converged = 1;
    S.done = 1;
  } else if (it >= S.max_iter) {
    S.done = 1;
  } else {
    S.beta = rz[0] / S.rz_old;
    S.rz_old = rz[0];
  }
}

__global__ __launch_bounds__(kWG) void kbf_rz(const FinItem *__restrict__ fin,
                                              const double *__restrict__ part, int nch,
                                              SysScal *sc, int init) {
  const FinItem f = fin[blockIdx.x];
  SysScal &S = sc[f.slot];
  if (S.done) return;
  double out[1];
  if (!fold_slot<1>(f, part, nch, out)) return;
  if (init) {
    S.rz_old = out[0];
  } else {
    S.beta = out[0] / S.rz_old;
    S.rz_old = out[0];
  }
}

}  // namespace iccg

using namespace iccg;

// ---------------------------------------------------------------------------
// host engine
// ---------------------------------------------------------------------------
namespace {

constexpr size_t kAlign = 256;
inline size_t align_up(size_t b) { return (b + kAlign - 1) / kAlign * kAlign; }

// Item lists of a set of slots, resident on the device in one buffer.
struct ItemSet {
  int ngroups = 0, nchunks = 0, nfin = 0, ncolors = 0, kmax = 1;
  std::vector<int> color_off;  // ncolors + 1 offsets into d_color
  // per color: the leading items of sectioned-layout factors (kb_sweep_sec),
  // and their largest section width; of those, the leading color_fb items
  // belong to factors whose last color this is, and their forward launch
  // runs their backward rows as well (the backward launch skips them)
  std::vector<int> color_sec, color_fb;
  int secw = 0;
  std::vector<int> work_of;    // per listed slot (fin order): its group's chunks
  char *buf = nullptr;
  GroupDesc *d_groups = nullptr;
  ChunkItem *d_chunks = nullptr;
  ColorItem *d_color = nullptr;
  FinItem *d_fin = nullptr;
  void clear() {
    dfree(buf);
    *this = ItemSet();
  }
};

struct Slot {
  int state = 0;  // 0 free, 1 admitted (needs init), 2 running, 3 finished
  iccg_csr_t A = nullptr;
  iccg_precond_t M = nullptr;
  int n = 0;
  // device vectors x r p Ap b y z in one allocation, kept for reuse
  char *dbuf = nullptr;
  size_t dcap = 0;
  // pinned staging of b (and x0) until admission, kept for reuse
  double *hstage = nullptr;
  size_t hcap = 0;
  bool has_x0 = false;
  double *x = nullptr, *r = nullptr, *p = nullptr, *Ap = nullptr, *b = nullptr;
  void *y = nullptr, *z = nullptr;
  int max_iter = 0, flags = 0, restarts = 0;
  double tol = 0;
  int listed_work = 0;  // chunk items of this slot's group in the running set
  std::chrono::steady_clock::time_point t_add;
  iccg_cg_result res{};
};

int wm_class(int w) { return w <= 4 ? 4 : w <= 5 ? 5 : w <= 8 ? 8 : 16; }

}  // namespace

struct iccg_batch_s {
  int device = 0;
  hipStream_t st = nullptr;
  int cap = 0;
  std::vector<Slot> slots;
  SysDesc *d_desc = nullptr;
  SysDesc *h_desc = nullptr;  // pinned mirror
  SysScal *d_sc = nullptr;
  SysScal *h_sc = nullptr;    // pinned
  double *part = nullptr;
  size_t part_slots = 0;
  int wm = 4;                 // template width of A's rows in the running set (kb_ax)
  int wmf = 4;                // and of its factors' rows (kb_sweep)
  int chunk = 16;             // slices per chunk item
  int cw = 1;                 // slices (MC) or blocks (ABMC) per sweep wavefront
  bool autotune = true;
  long tuned_rows = 0;        // running rows at the last tuning
  int max_group = 1;          // right-hand sides per group (see set_max_group)
  std::map<int, std::pair<int, int>> tuned;  // log2(running rows) -> (chunk, cw)
  ItemSet running;            // items of the running set
  bool running_dirty = true;
  long running_work = 0;      // sum of listed_work over the running set
  long dead_work = 0;         // listed work of slots no longer running
  bool graphs = true;
  std::vector<std::pair<int, hipGraphExec_t>> execs;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // timing mode: HIP events around every launch group, kernels launched
  // one by one (no graphs)
  struct Mark {
    int cls;
    hipEvent_t a, b;
  };
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  hipEvent_t open_ev = nullptr;
  hipEvent_t wait_ev = nullptr;  // blocking waits (iccg_batch_set_blocking)
  std::vector<Mark> marks;
  double t_ms[4] = {0, 0, 0, 0};
  long t_cnt[4] = {0, 0, 0, 0};
  ~iccg_batch_s() {
    (void)hipSetDevice(device);
    if (st) (void)hipStreamSynchronize(st);
    drop_graphs();
    running.clear();
    for (auto &s : slots) {
      dfree(s.dbuf);
      hfree(s.hstage);
    }
    dfree(d_desc);
    dfree(d_sc);
    dfree(part);
    hfree(h_sc);
    hfree(h_desc);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
    if (wait_ev) (void)hipEventDestroy(wait_ev);
    if (st) (void)hipStreamDestroy(st);
  }
  void drop_graphs() {
    for (auto &e : execs) (void)hipGraphExecDestroy(e.second);
    execs.clear();
  }
};

namespace {

void ensure_part(iccg_batch_s &B, size_t slots) {
  if (slots <= B.part_slots) return;
  HIPCHK(hipStreamSynchronize(B.st));
  dfree(B.part);
  B.part = nullptr;
  B.part_slots = 0;
  B.part = dalloc<double>(slots * 2);
  B.part_slots = slots;
  B.drop_graphs();
}

// Item lists for the given slots with chunk / sweep granularities ch, cw.
// Slots of one matrix and preconditioner are grouped, kMaxRhs at most, in
// slot order; fin lists the slots in the order of ids.
void build_items(iccg_batch_s &B, const std::vector<int> &ids, ItemSet &I, int ch, int cw) {
  if (I.buf) HIPCHK(hipStreamSynchronize(B.st));  // launches reading the old lists have run
  I.clear();
  std::vector<GroupDesc> groups;
  std::map<std::pair<const void *, const void *>, int> open;  // (A, M) -> group with room
  std::vector<std::pair<int, int>> where;                      // per id: (group, member)
  for (int id : ids) {
    const Slot &s = B.slots[id];
    const auto key = std::make_pair((const void *)s.A, (const void *)s.M);
    auto it = open.find(key);
    int gi;
    if (it == open.end() || groups[(size_t)it->second].k == B.max_group) {
      gi = (int)groups.size();
      GroupDesc g{};
      groups.push_back(g);
      open[key] = gi;
    } else {
      gi = it->second;
    }
    GroupDesc &g = groups[(size_t)gi];
    where.push_back({gi, g.k});
    g.slot[g.k++] = id;
  }
  std::vector<int> gchunk0, gchunks;
  std::vector<ChunkItem> chunks;
  std::vector<std::vector<ColorItem>> color, color_sec, color_fb;
  int kmax = 1, secw = 0;
  for (int gi = 0; gi < (int)groups.size(); ++gi) {
    const GroupDesc &g = groups[(size_t)gi];
    kmax = std::max(kmax, g.k);
    const Slot &s = B.slots[g.slot[0]];
    const int ns = (s.n + 63) / 64;
    gchunk0.push_back((int)chunks.size());
    for (int a = 0; a < ns; a += ch) chunks.push_back({gi, a, std::min(ns, a + ch), g.slot[0]});
    gchunks.push_back((int)chunks.size() - gchunk0.back());
    const DevSched &sc = s.M->sc;
    const int nc = sc.nc;
    if ((int)color.size() < nc) {
      color.resize(nc);
      color_sec.resize(nc);
      color_fb.resize(nc);
    }
    const int sec = sc.kind == 2 ? s.M->F.sec : 0;
    const bool fb = sec > 0 && s.M->F.fb;
    // the kernels' slot count covers the in-block and the outside entries
    if (sec > 0) secw = std::max({secw, sec, s.M->F.seco});
    for (int c = 0; c < nc; ++c) {
      if (sc.kind == 2) {
        const int b0 = sc.h_color_blk_v[c], b1 = sc.h_color_blk_v[c + 1];
        auto &list = sec == 0 ? color : fb && c == nc - 1 ? color_fb : color_sec;
        for (int kb = b0; kb < b1; kb += cw)
          list[c].push_back({gi, kb, std::min(b1, kb + cw), 0, 0, g.slot[0], {0, 0}});
      } else {
        const int lo = sc.h_color_ptr_v[c], hi = sc.h_color_ptr_v[c + 1];
        if (hi <= lo) continue;
        const int s1 = ((hi - 1) >> 6) + 1;
        for (int sl = lo >> 6; sl < s1; sl += cw)
          color[c].push_back({gi, sl, std::min(s1, sl + cw), lo, hi, g.slot[0], {0, 0}});
      }
    }
  }
  std::vector<FinItem> fin;
  for (size_t k = 0; k < ids.size(); ++k) {
    const int gi = where[k].first;
    fin.push_back({ids[k], gchunk0[(size_t)gi], gchunk0[(size_t)gi] + gchunks[(size_t)gi],
                   where[k].second});
    I.work_of.push_back(gchunks[(size_t)gi]);
  }
  I.ngroups = (int)groups.size();
  I.nchunks = (int)chunks.size();
  I.nfin = (int)fin.size();
  I.ncolors = (int)color.size();
  I.kmax = kmax;
  I.secw = secw;
  I.color_off.assign(1, 0);
  I.color_sec.clear();
  I.color_fb.clear();
  for (size_t c = 0; c < color.size(); ++c) {
    I.color_fb.push_back((int)color_fb[c].size());
    I.color_sec.push_back((int)(color_fb[c].size() + color_sec[c].size()));
    color[c].insert(color[c].begin(), color_sec[c].begin(), color_sec[c].end());
    color[c].insert(color[c].begin(), color_fb[c].begin(), color_fb[c].end());
    I.color_off.push_back(I.color_off.back() + (int)color[c].size());
  }
  if (I.nfin == 0) return;
  // one device buffer: groups | chunks | fin | color items
  const size_t o_chunks = align_up(groups.size() * sizeof(GroupDesc));
  const size_t o_fin = o_chunks + align_up(chunks.size() * sizeof(ChunkItem));
  const size_t o_color = o_fin + align_up(fin.size() * sizeof(FinItem));
  const size_t total = o_color + (size_t)I.color_off.back() * sizeof(ColorItem);
  std::vector<char> h(total);
  std::memcpy(h.data(), groups.data(), groups.size() * sizeof(GroupDesc));
  std::memcpy(h.data() + o_chunks, chunks.data(), chunks.size() * sizeof(ChunkItem));
  std::memcpy(h.data() + o_fin, fin.data(), fin.size() * sizeof(FinItem));
  size_t o = o_color;
  for (auto &v : color) {
    if (!v.empty()) std::memcpy(h.data() + o, v.data(), v.size() * sizeof(ColorItem));
    o += v.size() * sizeof(ColorItem);
  }
  ensure_part(B, (size_t)I.nchunks * 3 * (size_t)kmax);
  I.buf = dalloc<char>(total);
  I.d_groups = (GroupDesc *)I.buf;
  I.d_chunks = (ChunkItem *)(I.buf + o_chunks);
  I.d_fin = (FinItem *)(I.buf + o_fin);
  I.d_color = (ColorItem *)(I.buf + o_color);
  HIPCHK(hipMemcpyAsync(I.buf, h.data(), total, hipMemcpyHostToDevice, B.st));
  HIPCHK(hipStreamSynchronize(B.st));
}

int km_class(int k) { return k <= 1 ? 1 : k <= 2 ? 2 : k <= 4 ? 4 : 8; }

// K<WM, KM, ...extra> over the template widths of the set
#define BATCH_DISPATCH(wm, km, K, EXTRA, ...)                                                  \
  do {                                                                                        \
    const int _w = (wm) <= 4 ? 4 : (wm) <= 5 ? 5 : (wm) <= 8 ? 8 : 16, _k = km_class(km);     \
    if (_w == 4 && _k == 1) hipLaunchKernelGGL((K<4, 1 EXTRA>), __VA_ARGS__);                  \
    else if (_w == 4 && _k == 2) hipLaunchKernelGGL((K<4, 2 EXTRA>), __VA_ARGS__);             \
    else if (_w == 4 && _k == 4) hipLaunchKernelGGL((K<4, 4 EXTRA>), __VA_ARGS__);             \
    else if (_w == 4) hipLaunchKernelGGL((K<4, 8 EXTRA>), __VA_ARGS__);                        \
    else if (_w == 5 && _k == 1) hipLaunchKernelGGL((K<5, 1 EXTRA>), __VA_ARGS__);             \
    else if (_w == 5 && _k == 2) hipLaunchKernelGGL((K<5, 2 EXTRA>), __VA_ARGS__);             \
    else if (_w == 5 && _k == 4) hipLaunchKernelGGL((K<5, 4 EXTRA>), __VA_ARGS__);             \
    else if (_w == 5) hipLaunchKernelGGL((K<5, 8 EXTRA>), __VA_ARGS__);                        \
    else if (_w == 8 && _k == 1) hipLaunchKernelGGL((K<8, 1 EXTRA>), __VA_ARGS__);             \
    else if (_w == 8 && _k == 2) hipLaunchKernelGGL((K<8, 2 EXTRA>), __VA_ARGS__);             \
    else if (_w == 8 && _k == 4) hipLaunchKernelGGL((K<8, 4 EXTRA>), __VA_ARGS__);             \
    else if (_w == 8) hipLaunchKernelGGL((K<8, 8 EXTRA>), __VA_ARGS__);                        \
    else if (_k == 1) hipLaunchKernelGGL((K<16, 1 EXTRA>), __VA_ARGS__);                       \
    else if (_k == 2) hipLaunchKernelGGL((K<16, 2 EXTRA>), __VA_ARGS__);                       \
    else if (_k == 4) hipLaunchKernelGGL((K<16, 4 EXTRA>), __VA_ARGS__);                       \
    else hipLaunchKernelGGL((K<16, 8 EXTRA>), __VA_ARGS__);                                    \
  } while (0)
// vector kernels: a one-member path, and one for groups
#define VEC_LAUNCH(km, K, ...)                                  \
  do {                                                          \
    if ((km) <= 1) hipLaunchKernelGGL((K<1>), __VA_ARGS__);     \
    else hipLaunchKernelGGL((K<kMaxRhs>), __VA_ARGS__);         \
  } while (0)
#define NOEXTRA
#define FWD_T , true
#define BWD_T , false
#define SFB_T , true, true
#define SFWD_T , true, false
#define SBWD_T , false, false

// kernel classes of the timing mode
enum { T_SPMV = 0, T_FWD = 1, T_BWD = 2, T_VEC = 3, T_NCLASS = 4 };

void mark(iccg_batch_s &B, int cls, bool begin);

void launch_ax(iccg_batch_s &B, const ItemSet &I, int mode) {
  BATCH_DISPATCH(B.wm, I.kmax, kb_ax, NOEXTRA, dim3(I.nchunks), dim3(kWG), 0, B.st, I.d_groups,
                 B.d_desc, I.d_chunks, B.d_sc, B.part, I.nchunks, mode);
}

void launch_resid_items(iccg_batch_s &B, const ItemSet &I, int mode) {
  if (I.nfin == 0) return;
  launch_ax(B, I, mode);
  hipLaunchKernelGGL(kbf_resid, dim3(I.nfin), dim3(kWG), 0, B.st, I.d_fin, B.part, I.nchunks,
                     B.d_desc, B.d_sc, mode);
}

// sectioned-layout items: section width class x group size class
#define SEC_DISPATCH(sw, km, K, EXTRA, ...)                                                   \
  do {                                                                                        \
    const int _k = km_class(km);                                                              \
    if ((sw) <= 2 && _k == 1) hipLaunchKernelGGL((K<2, 1 EXTRA>), __VA_ARGS__);               \
    else if ((sw) <= 2) hipLaunchKernelGGL((K<2, kMaxRhs EXTRA>), __VA_ARGS__);               \
    else if ((sw) <= 3 && _k == 1) hipLaunchKernelGGL((K<3, 1 EXTRA>), __VA_ARGS__);          \
    else if ((sw) <= 3) hipLaunchKernelGGL((K<3, kMaxRhs EXTRA>), __VA_ARGS__);               \
    else if (_k == 1) hipLaunchKernelGGL((K<6, 1 EXTRA>), __VA_ARGS__);                       \
    else hipLaunchKernelGGL((K<6, kMaxRhs EXTRA>), __VA_ARGS__);                              \
  } while (0)

template <bool FWD>
static void launch_color(iccg_batch_s &B, const ItemSet &I, int c) {
  const int nf = I.color_fb[c], n0 = I.color_sec[c], ni = I.color_off[c + 1] - I.color_off[c];
  const ColorItem *items = I.d_color + I.color_off[c];
  if (FWD && nf > 0)
    SEC_DISPATCH(I.secw, I.kmax, kb_sweep_sec, SFB_T, dim3((nf + 3) / 4), dim3(kWG), 0, B.st,
                 I.d_groups, B.d_desc, items, nf, B.d_sc);
  if (n0 > nf) {
    if (FWD)
      SEC_DISPATCH(I.secw, I.kmax, kb_sweep_sec, SFWD_T, dim3((n0 - nf + 3) / 4), dim3(kWG), 0,
                   B.st, I.d_groups, B.d_desc, items + nf, n0 - nf, B.d_sc);
    else
      SEC_DISPATCH(I.secw, I.kmax, kb_sweep_sec, SBWD_T, dim3((n0 - nf + 3) / 4), dim3(kWG), 0,
                   B.st, I.d_groups, B.d_desc, items + nf, n0 - nf, B.d_sc);
  }
  if (ni > n0) {
    if (FWD)
      BATCH_DISPATCH(B.wmf, I.kmax, kb_sweep, FWD_T, dim3((ni - n0 + 3) / 4), dim3(kWG), 0, B.st,
                     I.d_groups, B.d_desc, items + n0, ni - n0, B.d_sc);
    else
      BATCH_DISPATCH(B.wmf, I.kmax, kb_sweep, BWD_T, dim3((ni - n0 + 3) / 4), dim3(kWG), 0, B.st,
                     I.d_groups, B.d_desc, items + n0, ni - n0, B.d_sc);
  }
}

void launch_sweeps(iccg_batch_s &B, const ItemSet &I) {
  mark(B, T_FWD, true);
  for (int c = 0; c < I.ncolors; ++c) launch_color<true>(B, I, c);
  mark(B, T_FWD, false);
  mark(B, T_BWD, true);
  for (int c = I.ncolors - 1; c >= 0; --c) launch_color<false>(B, I, c);
  mark(B, T_BWD, false);
}

void launch_rz_p(iccg_batch_s &B, const ItemSet &I, int init) {
  dim3 g(I.nchunks), b(kWG);
  mark(B, T_VEC, true);
  VEC_LAUNCH(I.kmax, kb_rz, g, b, 0, B.st, I.d_groups, B.d_desc, I.d_chunks, B.d_sc, B.part,
                     I.nchunks);
  hipLaunchKernelGGL(kbf_rz, dim3(I.nfin), b, 0, B.st, I.d_fin, B.part, I.nchunks, B.d_sc, init);
  VEC_LAUNCH(I.kmax, kb_update_p, g, b, 0, B.st, I.d_groups, B.d_desc, I.d_chunks, B.d_sc, init);
  mark(B, T_VEC, false);
}

// z0 = M^-1 r0, rz_old, p = z0 (cg_crs.cpp:112-115) for the listed slots
void launch_precond_init(iccg_batch_s &B, const ItemSet &I) {
  if (I.nfin == 0) return;
  launch_sweeps(B, I);
  launch_rz_p(B, I, 1);
}

// one CG iteration (cg_crs.cpp:125-160) of every listed slot
void launch_iteration(iccg_batch_s &B, const ItemSet &I) {
  dim3 g(I.nchunks), b(kWG);
  mark(B, T_SPMV, true);
  launch_ax(B, I, 3);
  mark(B, T_SPMV, false);
  mark(B, T_VEC, true);
  hipLaunchKernelGGL(kbf_alpha, dim3(I.nfin), b, 0, B.st, I.d_fin, B.part, I.nchunks, B.d_sc);
  VEC_LAUNCH(I.kmax, kb_r, g, b, 0, B.st, I.d_groups, B.d_desc, I.d_chunks, B.d_sc, B.part,
                     I.nchunks);
  mark(B, T_VEC, false);
  launch_sweeps(B, I);
  // the stop test after the preconditioner: the r.z partials go behind the
  // r.r ones (which stay for the fold), then x and p together
  double *part_rz = B.part + 2 * (size_t)I.nchunks * I.kmax;
  mark(B, T_VEC, true);
  VEC_LAUNCH(I.kmax, kb_rz, g, b, 0, B.st, I.d_groups, B.d_desc, I.d_chunks, B.d_sc, part_rz,
                     I.nchunks);
  hipLaunchKernelGGL(kbf_xr_rz, dim3(I.nfin), b, 0, B.st, I.d_fin, B.part, part_rz, I.nchunks,
                     B.d_sc);
  VEC_LAUNCH(I.kmax, kb_update_xp, g, b, 0, B.st, I.d_groups, B.d_desc, I.d_chunks, B.d_sc);
  mark(B, T_VEC, false);
}

void mark(iccg_batch_s &B, int cls, bool begin) {
  if (!B.timing) return;
  if (B.ev_used == B.ev_pool.size()) {
    hipEvent_t e = nullptr;
    HIPCHK(hipEventCreate(&e));
    B.ev_pool.push_back(e);
  }
  hipEvent_t e = B.ev_pool[B.ev_used++];
  HIPCHK(hipEventRecord(e, B.st));
  if (begin)
    B.open_ev = e;
  else
    B.marks.push_back({cls, B.open_ev, e});
}

// after a stream synchronization: fold the recorded launch groups
void harvest(iccg_batch_s &B) {
  for (const auto &m : B.marks) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, m.a, m.b));
    B.t_ms[m.cls] += ms;
    B.t_cnt[m.cls]++;
  }
  B.marks.clear();
  B.ev_used = 0;
}

void enqueue_iterations(iccg_batch_s &B, int k) {
  const ItemSet &I = B.running;
  if (!B.graphs || B.timing) {
    for (int j = 0; j < k; ++j) launch_iteration(B, I);
    return;
  }
  hipGraphExec_t exec = nullptr;
  for (auto &e : B.execs)
    if (e.first == k) exec = e.second;
  if (!exec) {
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(B.st, hipStreamCaptureModeThreadLocal));
    for (int j = 0; j < k; ++j) launch_iteration(B, I);
    HIPCHK(hipStreamEndCapture(B.st, &g));
    hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHK(e);
    B.execs.push_back({k, exec});
  }
  HIPCHK(hipGraphLaunch(exec, B.st));
}

void read_scalars(iccg_batch_s &B) {
  HIPCHK(hipMemcpyAsync(B.h_sc, B.d_sc, sizeof(SysScal) * (size_t)B.cap, hipMemcpyDeviceToHost,
                        B.st));
  stream_wait(B.st, B.wait_ev);
  if (B.timing) harvest(B);
}

void retire_listed(iccg_batch_s &B, Slot &s) {
  B.dead_work += s.listed_work;
  s.listed_work = 0;
}

void finish_slot(iccg_batch_s &B, int id) {
  Slot &s = B.slots[id];
  const SysScal &S = B.h_sc[id];
  s.res.iters = S.iters;
  s.res.rel_resid = S.rel;
  s.res.converged = S.converged;
  s.res.breakdown = S.breakdown;
  s.res.restarts = s.restarts;
  s.res.solve_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - s.t_add)
          .count();
  s.state = 3;
  retire_listed(B, s);
}

// Systems whose recurrence stopped get the FP64 true residual; converged
// ones above tol restart from it (ICCG_SOLVE_TRUE_RESID), the others finish.
void settle(iccg_batch_s &B) {
  std::vector<int> done_now;
  for (int id = 0; id < B.cap; ++id) {
    Slot &s = B.slots[id];
    if (s.state == 2 && B.h_sc[id].done) done_now.push_back(id);
  }
  if (done_now.empty()) return;
  ItemSet I;
  build_items(B, done_now, I, B.chunk, B.cw);
  launch_resid_items(B, I, 0);
  HIPCHK(hipGetLastError());
  read_scalars(B);
  I.clear();
  std::vector<int> restart;
  for (int id : done_now) {
    Slot &s = B.slots[id];
    const SysScal &S = B.h_sc[id];
    const double trel = std::sqrt(S.rr_true) / S.normb;
    s.res.true_rel_resid = trel;
    const bool want = (s.flags & ICCG_SOLVE_TRUE_RESID) != 0;
    if (want && S.converged && !S.breakdown && trel >= s.tol && S.iters < s.max_iter &&
        s.restarts < 100) {
      s.restarts++;
      restart.push_back(id);
    } else {
      finish_slot(B, id);
    }
  }
  if (!restart.empty()) {
    ItemSet R;
    build_items(B, restart, R, B.chunk, B.cw);
    launch_resid_items(B, R, 2);
    launch_precond_init(B, R);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(B.st));
    R.clear();
    // restarted systems run again: their scalars are fresh on the device
    read_scalars(B);
  }
}

void admit_pending(iccg_batch_s &B) {
  std::vector<int> ids;
  for (int id = 0; id < B.cap; ++id)
    if (B.slots[id].state == 1) ids.push_back(id);
  if (ids.empty()) return;
  for (int id : ids) {
    const Slot &s = B.slots[id];
    const size_t nb = sizeof(double) * (size_t)s.n;
    HIPCHK(hipMemcpyAsync(s.b, s.hstage, nb, hipMemcpyHostToDevice, B.st));
    if (s.has_x0)
      HIPCHK(hipMemcpyAsync(s.x, s.hstage + s.n, nb, hipMemcpyHostToDevice, B.st));
    else
      HIPCHK(hipMemsetAsync(s.x, 0, nb, B.st));
  }
  HIPCHK(hipMemcpyAsync(B.d_desc, B.h_desc, sizeof(SysDesc) * (size_t)B.cap,
                        hipMemcpyHostToDevice, B.st));
  int wm = 4, wmf = 4;
  for (const Slot &s : B.slots)
    if (s.state == 1 || s.state == 2) {
      wm = std::max(wm, s.A->d.sell.maxw);
      wmf = std::max({wmf, s.M->F.ls.maxw, s.M->F.us.maxw});
    }
  B.wm = wm_class(wm);
  B.wmf = wm_class(wmf);
  ItemSet I;
  build_items(B, ids, I, B.chunk, B.cw);
  launch_resid_items(B, I, 1);
  launch_precond_init(B, I);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(B.st));
  I.clear();
  for (int id : ids) B.slots[id].state = 2;
  B.running_dirty = true;
}

constexpr int kTuneChunks[] = {4, 8, 16, 32};
constexpr int kTuneSweeps[] = {1, 2, 4};
constexpr int kTuneIters = 2;  // timed iterations per candidate
constexpr int kTuneSteps =
    1 + kTuneIters * (int)(sizeof(kTuneChunks) / sizeof(int)) * (int)(sizeof(kTuneSweeps) / sizeof(int));

// the running slots (state 2, not stopped) and their rows
long running_ids(iccg_batch_s &B, std::vector<int> &ids) {
  ids.clear();
  long rows = 0;
  for (int id = 0; id < B.cap; ++id)
    if (B.slots[id].state == 2 && !B.h_sc[id].done) {
      ids.push_back(id);
      rows += B.slots[id].n;
    }
  return rows;
}

bool tune_due(iccg_batch_s &B, long rows) {
  return B.autotune && rows >= (1L << 16) &&
         (B.tuned_rows == 0 || rows > 2 * B.tuned_rows || 2 * rows < B.tuned_rows);
}

// Online autotuner: runs real CG iterations of the running systems under
// every candidate decomposition (kTuneIters each, after one warm-up) with
// HIP events around them and keeps the fastest.  The iterations are the
// solve's own -- the decomposition changes no row's arithmetic -- so tuning
// costs only the launches it makes without graphs.  Returns the iterations
// run (kTuneSteps).
int tune_online(iccg_batch_s &B, const std::vector<int> &ids) {
  if (!B.ev0) {
    HIPCHK(hipEventCreate(&B.ev0));
    HIPCHK(hipEventCreate(&B.ev1));
  }
  const bool timing = B.timing;
  B.timing = false;  // the tuning launches are not counted
  B.drop_graphs();
  ItemSet T;
  build_items(B, ids, T, B.chunk, B.cw);
  launch_iteration(B, T);  // warm-up
  double best = 1e300;
  int best_ch = B.chunk, best_cw = B.cw;
  for (int ch : kTuneChunks)
    for (int cw : kTuneSweeps) {
      build_items(B, ids, T, ch, cw);
      HIPCHK(hipEventRecord(B.ev0, B.st));
      for (int r = 0; r < kTuneIters; ++r) launch_iteration(B, T);
      HIPCHK(hipEventRecord(B.ev1, B.st));
      HIPCHK(hipEventSynchronize(B.ev1));
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, B.ev0, B.ev1));
      if (ms < best) {
        best = ms;
        best_ch = ch;
        best_cw = cw;
      }
    }
  HIPCHK(hipGetLastError());
  T.clear();
  B.chunk = best_ch;
  B.cw = best_cw;
  B.timing = timing;
  B.running_dirty = true;  // rebuild with the chosen decomposition
  return kTuneSteps;
}

void refresh_running(iccg_batch_s &B) {
  if (!B.running_dirty && B.dead_work * 4 <= B.running_work) return;
  std::vector<int> ids;
  running_ids(B, ids);
  B.drop_graphs();
  for (Slot &s : B.slots) s.listed_work = 0;
  build_items(B, ids, B.running, B.chunk, B.cw);
  B.running_work = 0;
  for (size_t k = 0; k < ids.size(); ++k) {
    B.slots[ids[k]].listed_work = B.running.work_of[k];
    B.running_work += B.running.work_of[k];
  }
  B.running_dirty = false;
  B.dead_work = 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int iccg_batch_create(int capacity, iccg_batch_t *out) {
  return guard([&] {
    require(out && capacity > 0, ICCG_ERR_ARG, "bad argument");
    std::unique_ptr<iccg_batch_s> B(new iccg_batch_s);
    B->device = current_device();
    HIPCHK(hipSetDevice(B->device));
    B->cap = capacity;
    B->slots.resize((size_t)capacity);
    HIPCHK(hipStreamCreateWithFlags(&B->st, hipStreamNonBlocking));
    B->d_desc = dalloc<SysDesc>((size_t)capacity);
    B->d_sc = dalloc<SysScal>((size_t)capacity);
    B->h_sc = halloc<SysScal>((size_t)capacity);
    B->h_desc = halloc<SysDesc>((size_t)capacity);
    std::memset((void *)B->h_sc, 0, sizeof(SysScal) * (size_t)capacity);
    for (int k = 0; k < capacity; ++k) B->h_desc[k] = SysDesc();
    HIPCHK(hipMemcpyAsync(B->d_sc, B->h_sc, sizeof(SysScal) * (size_t)capacity,
                          hipMemcpyHostToDevice, B->st));
    HIPCHK(hipMemcpyAsync(B->d_desc, B->h_desc, sizeof(SysDesc) * (size_t)capacity,
                          hipMemcpyHostToDevice, B->st));
    ensure_part(*B, 4096);
    HIPCHK(hipStreamSynchronize(B->st));
    *out = B.release();
  });
}

int iccg_batch_destroy(iccg_batch_t B) {
  delete B;
  return ICCG_OK;
}

int iccg_batch_add(iccg_batch_t B, iccg_csr_t A, iccg_precond_t M, const double *b,
                   const double *x0, int max_iter, double tol, int flags, int *slot) {
  return guard([&] {
    require(B && A && M && b && slot, ICCG_ERR_ARG, "null argument");
    require(M->A == A, ICCG_ERR_ARG, "preconditioner was built for another matrix");
    require(M->kind == ICCG_PRECOND_IC0 && sliced_plan(M->sc, M->F) && A->d.sell.off &&
                A->d.sell.maxw <= 16,
            ICCG_ERR_ARG,
            "batched solves need an IC0 preconditioner with a contiguous multicolor or ABMC "
            "(blocks of at most 64 rows) schedule and rows of at most 16 entries");
    int id = -1;
    for (int k = 0; k < B->cap; ++k)
      if (B->slots[k].state == 0) {
        id = k;
        break;
      }
    require(id >= 0, ICCG_ERR_STATE, "batch is full");
    HIPCHK(hipSetDevice(B->device));
    Slot &s = B->slots[id];
    const int n = A->d.n;
    const size_t vb = M->mixed ? sizeof(float) : sizeof(double);
    const size_t v8 = align_up(sizeof(double) * (size_t)n), vf = align_up(vb * (size_t)n);
    const size_t need = 5 * v8 + 2 * vf;
    if (s.dcap < need) {
      dfree(s.dbuf);
      s.dbuf = nullptr;
      s.dcap = 0;
      s.dbuf = dalloc<char>(need);
      s.dcap = need;
    }
    const size_t hneed = sizeof(double) * 2 * (size_t)n;
    if (s.hcap < hneed) {
      hfree(s.hstage);
      s.hstage = nullptr;
      s.hcap = 0;
      s.hstage = halloc<double>(2 * (size_t)n);
      s.hcap = hneed;
    }
    std::memcpy(s.hstage, b, sizeof(double) * (size_t)n);
    s.has_x0 = x0 != nullptr;
    if (x0) std::memcpy(s.hstage + n, x0, sizeof(double) * (size_t)n);
    char *q = s.dbuf;
    s.x = (double *)q;
    s.r = (double *)(q += v8);
    s.p = (double *)(q += v8);
    s.Ap = (double *)(q += v8);
    s.b = (double *)(q += v8);
    s.y = (void *)(q += v8);
    s.z = (void *)(q += vf);
    s.A = A;
    s.M = M;
    s.n = n;
    s.max_iter = max_iter;
    s.tol = tol;
    s.flags = flags;
    s.restarts = 0;
    s.listed_work = 0;
    s.res = iccg_cg_result{};
    s.res.true_rel_resid = -1.0;
    SysDesc &D = B->h_desc[id];
    D = SysDesc();
    D.n = n;
    D.nslice = (n + 63) / 64;
    D.mixed = M->mixed;
    D.max_iter = max_iter;
    D.tol = tol;
    D.a_off = A->d.sell.off;
    D.a_col = A->d.sell.col;
    D.a_v64 = A->d.sell.v64;
    D.a_paired = A->d.sell.col16 != nullptr;
    D.a_v32 = A->d.sell_v32;
    D.l_off = M->F.ls.off;
    D.l_col = M->F.ls.col;
    D.u_off = M->F.us.off;
    D.u_col = M->F.us.col;
    D.l_v = M->mixed ? (const void *)M->F.ls.v32 : (const void *)M->F.ls.v64;
    D.u_v = M->mixed ? (const void *)M->F.us.v32 : (const void *)M->F.us.v64;
    D.d = M->mixed ? (const void *)M->F.d32 : (const void *)M->F.d64;
    if (M->sc.kind == 2) {
      D.bstart = M->sc.d_bstart;
      D.lev_f = M->sc.d_lev_f;
      D.lev_b = M->sc.d_lev_b;
      D.nlev_f = M->sc.d_nlev_f;
      D.nlev_b = M->sc.d_nlev_b;
      D.sec = M->F.sec;
      D.l_inb = M->F.ls.inb;
      D.u_inb = M->F.us.inb;
    }
    D.x = s.x;
    D.r = s.r;
    D.p = s.p;
    D.Ap = s.Ap;
    D.b = s.b;
    D.y = s.y;
    D.z = s.z;
    s.t_add = std::chrono::steady_clock::now();
    s.state = 1;
    *slot = id;
  });
}

int iccg_batch_run(iccg_batch_t B, int max_steps, int *n_running) {
  return guard([&] {
    require(B && max_steps >= 0, ICCG_ERR_ARG, "bad argument");
    HIPCHK(hipSetDevice(B->device));
    admit_pending(*B);
    read_scalars(*B);
    settle(*B);  // systems already converged at iteration 0
    int steps = 0, chunk = 4;
    std::vector<int> ids;
    for (;;) {
      int running = 0;
      for (auto &s : B->slots) running += s.state == 2;
      if (running == 0 || steps >= max_steps) break;
      const long rows = running_ids(*B, ids);
      if (tune_due(*B, rows)) {
        // a size class seen before keeps its choice; a new one is tuned
        int bucket = 0;
        while ((2L << bucket) <= rows) ++bucket;
        auto it = B->tuned.find(bucket);
        if (it != B->tuned.end()) {
          if (B->chunk != it->second.first || B->cw != it->second.second) {
            B->chunk = it->second.first;
            B->cw = it->second.second;
            B->running_dirty = true;
          }
          B->tuned_rows = rows;
        } else if (max_steps - steps >= kTuneSteps) {
          steps += tune_online(*B, ids);
          B->tuned[bucket] = {B->chunk, B->cw};
          B->tuned_rows = rows;
          read_scalars(*B);
          settle(*B);
          continue;
        }
      }
      refresh_running(*B);
      const int k = std::min(chunk, max_steps - steps);
      enqueue_iterations(*B, k);
      HIPCHK(hipGetLastError());
      steps += k;
      read_scalars(*B);
      settle(*B);
      if (chunk < 32) chunk *= 2;
    }
    int running = 0;
    for (auto &s : B->slots) running += (s.state == 1 || s.state == 2);
    if (n_running) *n_running = running;
  });
}

int iccg_batch_result(iccg_batch_t B, int slot, int *finished, iccg_cg_result *res) {
  return guard([&] {
    require(B && slot >= 0 && slot < B->cap && finished, ICCG_ERR_ARG, "bad argument");
    const Slot &s = B->slots[slot];
    require(s.state != 0, ICCG_ERR_STATE, "empty batch slot");
    *finished = s.state == 3;
    if (res) *res = s.res;
  });
}

int iccg_batch_get_x(iccg_batch_t B, int slot, double *x) {
  return guard([&] {
    require(B && slot >= 0 && slot < B->cap && x, ICCG_ERR_ARG, "bad argument");
    const Slot &s = B->slots[slot];
    require(s.state >= 2, ICCG_ERR_STATE, "batch slot holds no solution");
    HIPCHK(hipSetDevice(B->device));
    HIPCHK(hipMemcpyAsync(x, s.x, sizeof(double) * (size_t)s.n, hipMemcpyDeviceToHost, B->st));
    HIPCHK(hipStreamSynchronize(B->st));
  });
}

int iccg_batch_remove(iccg_batch_t B, int slot) {
  return guard([&] {
    require(B && slot >= 0 && slot < B->cap, ICCG_ERR_ARG, "bad argument");
    Slot &s = B->slots[slot];
    if (s.state == 0) return;
    HIPCHK(hipSetDevice(B->device));
    HIPCHK(hipStreamSynchronize(B->st));
    retire_listed(*B, s);
    // the vector and staging buffers stay with the slot for its next system
    s.state = 0;
    s.A = nullptr;
    s.M = nullptr;
    SysScal S{};
    S.done = 2;  // stopped for good: no kernel touches it, the x update included
    B->h_sc[slot] = S;
    HIPCHK(hipMemcpyAsync(B->d_sc + slot, B->h_sc + slot, sizeof(S), hipMemcpyHostToDevice,
                          B->st));
    HIPCHK(hipStreamSynchronize(B->st));
  });
}

int iccg_batch_trim(iccg_batch_t B) {
  return guard([&] {
    require(B, ICCG_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(B->device));
    HIPCHK(hipStreamSynchronize(B->st));
    for (Slot &s : B->slots)
      if (s.state == 0) {
        dfree(s.dbuf);
        s.dbuf = nullptr;
        s.dcap = 0;
        hfree(s.hstage);
        s.hstage = nullptr;
        s.hcap = 0;
      }
  });
}

int iccg_batch_set_blocking(iccg_batch_t B, int on) {
  return guard([&] {
    require(B, ICCG_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(B->device));
    if (on && !B->wait_ev) B->wait_ev = blocking_event();
    if (!on && B->wait_ev) {
      (void)hipEventDestroy(B->wait_ev);
      B->wait_ev = nullptr;
    }
  });
}

int iccg_batch_set_graphs(iccg_batch_t B, int on) {
  return guard([&] {
    require(B, ICCG_ERR_ARG, "null argument");
    B->graphs = on != 0;
  });
}

int iccg_batch_set_timing(iccg_batch_t B, int on) {
  return guard([&] {
    require(B, ICCG_ERR_ARG, "null argument");
    B->timing = on != 0;
  });
}

int iccg_batch_get_timing(iccg_batch_t B, int cls, long *launches, double *total_ms) {
  return guard([&] {
    require(B && cls >= 0 && cls < 4 && launches && total_ms, ICCG_ERR_ARG, "bad argument");
    *launches = B->t_cnt[cls];
    *total_ms = B->t_ms[cls];
  });
}

int iccg_batch_reset_timing(iccg_batch_t B) {
  return guard([&] {
    require(B, ICCG_ERR_ARG, "null argument");
    for (int c = 0; c < 4; ++c) {
      B->t_ms[c] = 0;
      B->t_cnt[c] = 0;
    }
  });
}

int iccg_batch_set_max_group(iccg_batch_t B, int k) {
  return guard([&] {
    require(B && k >= 1 && k <= kMaxRhs, ICCG_ERR_ARG, "group size must be in [1, 8]");
    B->max_group = k;
    B->running_dirty = true;
  });
}

int iccg_batch_set_autotune(iccg_batch_t B, int on) {
  return guard([&] {
    require(B, ICCG_ERR_ARG, "null argument");
    B->autotune = on != 0;
    B->tuned_rows = 0;
    B->tuned.clear();
  });
}

int iccg_batch_get_tuning(iccg_batch_t B, int *chunk_slices, int *sweep_units) {
  return guard([&] {
    require(B && chunk_slices && sweep_units, ICCG_ERR_ARG, "null argument");
    *chunk_slices = B->chunk;
    *sweep_units = B->cw;
  });
}

int iccg_batch_set_tuning(iccg_batch_t B, int chunk_slices, int sweep_units) {
  return guard([&] {
    require(B && chunk_slices >= 1 && chunk_slices <= 1024 && sweep_units >= 1 &&
                sweep_units <= 64,
            ICCG_ERR_ARG, "bad argument");
    B->chunk = chunk_slices;
    B->cw = sweep_units;
    B->autotune = false;
    B->running_dirty = true;
  });
}

}  // extern "C"
