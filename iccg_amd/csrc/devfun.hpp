// devfun.hpp -- device helpers shared by the solver kernels (kernels.hip)
// and the multi-system kernels (batch.hip): double-double dot products,
// workgroup partials and their one-workgroup fold, XCD-aware work ranges,
// vector loads, and the sliced-ELL row chain.
#pragma once

#include <algorithm>
#include <cfloat>

#include "dev.hpp"

namespace iccg {

// ---------------------------------------------------------------------------
// double-double helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ void dd_add(double &h, double &l, double oh, double ol) {
  double s = h + oh;
  double bb = s - h;
  double e = (h - (s - bb)) + (oh - bb);
  e += l + ol;
  double hi = s + e;
  l = e - (hi - s);
  h = hi;
}

// (h,l) += a*b
__device__ __forceinline__ void dd_fma(double &h, double &l, double a, double b) {
  double p = a * b;
  double pe = fma(a, b, -p);
  double s = h + p;
  double bb = s - h;
  double e = (h - (s - bb)) + (p - bb);
  h = s;
  l += e + pe;
}

template <int NC>
__device__ __forceinline__ void wave_dd_reduce(double (&h)[NC], double (&l)[NC]) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      double oh = __shfl_xor(h[c], off, 64);
      double ol = __shfl_xor(l[c], off, 64);
      dd_add(h[c], l[c], oh, ol);
    }
  }
}

// Workgroup partial of NC double-double sums: the workgroup's lanes are
// folded in a fixed tree and thread 0 stores (hi, lo) pairs at slot gidx.
// No fence: a separate one-workgroup finish kernel reads the partials after
// the kernel boundary (an agent-scope fence per workgroup would write the
// L2 back once per workgroup).  blockDim.x: multiple of 64, at most 1024.
template <int NC>
__device__ void block_partial(double (&h)[NC], double (&l)[NC], double *part, int gidx) {
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
  if (threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      double ah = sh[c][0][0], al = sh[c][1][0];
      for (int k = 1; k < nw; ++k) dd_add(ah, al, sh[c][0][k], sh[c][1][k]);
      part[(gidx * NC + c) * 2 + 0] = ah;
      part[(gidx * NC + c) * 2 + 1] = al;
    }
  }
}

// Adds this thread's slots b = tid, tid + T, tid + 2T, ... (T threads) of
// NC slot-major partials, in ascending order, into (h, l); the loads of
// kFoldAhead slots go out before their adds, instead of one slot's loads
// per memory round trip.
constexpr int kFoldAhead = 4;
template <int NC>
__device__ __forceinline__ void fold_slots(const double *__restrict__ p, int count, double *h,
                                           double *l) {
  if (NC == 0) return;
  const int T = blockDim.x;
  for (int b0 = threadIdx.x; b0 < count; b0 += kFoldAhead * T) {
    double vh[kFoldAhead][NC > 0 ? NC : 1], vl[kFoldAhead][NC > 0 ? NC : 1];
#pragma unroll
    for (int u = 0; u < kFoldAhead; ++u) {
      const int b = b0 + u * T;
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        vh[u][c] = b < count ? p[((size_t)b * NC + c) * 2] : 0.0;
        vl[u][c] = b < count ? p[((size_t)b * NC + c) * 2 + 1] : 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < kFoldAhead; ++u)
      if (b0 + u * T < count) {
#pragma unroll
        for (int c = 0; c < NC; ++c) dd_add(h[c], l[c], vh[u][c], vl[u][c]);
      }
  }
}

// One workgroup folds the partials of one reduction (NA components per
// slot at pa, ca slots) or of two (and NB at pb, cb slots).  Every component
// is folded in the same fixed tree: per thread the slots b = tid, tid +
// blockDim, ... in order; then the wavefront's shuffle tree; then the
// wavefronts in order, each component by a thread of its own; then rounded
// once to double.  True on thread 0, with oa[] and ob[].  It ends with a
// barrier, so a following call may reuse the shared buffers.
template <int NA, int NB>
__device__ bool finish_reduce2(const double *__restrict__ pa, int ca,
                               const double *__restrict__ pb, int cb, double *oa, double *ob) {
  constexpr int N = NA + NB;
  __shared__ double sh[N][2][16];
  __shared__ double res[N];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  double h[N], l[N];
#pragma unroll
  for (int c = 0; c < N; ++c) h[c] = l[c] = 0.0;
  fold_slots<NA>(pa, ca, h, l);
  fold_slots<NB>(pb, cb, h + NA, l + NA);
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
__device__ bool finish_reduce(const double *__restrict__ part, int count, double (&out)[NC]) {
  return finish_reduce2<NC, 0>(part, count, nullptr, 0, out, nullptr);
}

// Makes this wavefront's LDS writes visible to its own lanes (rocPRIM's
// wave_barrier: no s_barrier, only ordering).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// XCD-aware distribution of [lo, hi) over the wavefronts of the grid.
// Workgroups are dealt to the 8 XCDs round-robin (workgroup b -> XCD b % 8)
// and each XCD has its own L2, so each XCD gets one contiguous eighth of the
// items: neighbouring slices / blocks, whose x-vector gathers overlap, then
// share an L2 instead of fetching the same lines on several XCDs.
constexpr int kXcd = 8;
constexpr int kFinWG = 1024;  // threads of the one-workgroup finish kernels
struct WaveRange {
  int first, step, end;
};
__device__ __forceinline__ WaveRange xcd_waves(int lo, int hi) {
  const int g = gridDim.x;
  const int nch = g < kXcd ? g : kXcd;  // chunks: every chunk has a workgroup
  const int x = blockIdx.x % nch, k = blockIdx.x / nch;
  const int wgs = (g - x + nch - 1) / nch;  // workgroups on this chunk
  const int wpg = blockDim.x >> 6;
  const int n = hi - lo;
  const int chunk = (n + nch - 1) / nch;
  const int a = lo + min(n, x * chunk), b = lo + min(n, (x + 1) * chunk);
  WaveRange r;
  r.first = a + k * wpg + (threadIdx.x >> 6);
  r.step = wgs * wpg;
  r.end = b;
  return r;
}

static inline int red_grid(long items_per_thread_block_unit) {
  long g = items_per_thread_block_unit;
  if (g < 1) g = 1;
  if (g > kMaxRedBlocks) g = kMaxRedBlocks;
  return (int)g;
}

static inline int stream_grid(long n) {
  long g = (n + kWG - 1) / kWG;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (int)g;
}

// grid of a streaming kernel that takes 4 elements per lane per step
static inline int quad_grid(long n) { return red_grid((n + 4L * kWG - 1) / (4L * kWG)); }
// grid of a kernel that takes one 64-row slice per wavefront per step
static inline int slice_grid(long nslices) { return red_grid((nslices + 3) / 4); }

// Row-width class WM of the SpMV kernels (for_rows_Ax and its kinds): the
// paired 16-bit layout holds WM / 2 pairs plus an odd last entry, so rows of
// up to WM + 1 entries fit class WM, and 5-point rows take class 4 (8 before:
// 2D 1980^2, four right-hand sides, 131.8 -> 117.7 us per launch); the 32-bit
// columns fit up to WM.  7-point rows stay in class 8: a class 6 measured
// slower (242.5 -> 257.0 us at 192^3, four right-hand sides).  0: CSR rows.
static inline int spmv_wm(const DevCsr &A) {
  const int w = A.sell.maxw;
  if (A.sell.col16) return w <= 5 ? 4 : w <= 9 ? 8 : w <= 17 ? 16 : 0;
  return w <= 4 ? 4 : w <= 8 ? 8 : w <= 16 ? 16 : 0;
}

__device__ __forceinline__ float fma_t(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return fma(a, b, c); }

// ---------------------------------------------------------------------------
// a / d, correctly rounded, split into the part that depends on d alone
// (rcp_refined, once per row) and the part that depends on a (div_steps,
// once per use).  hipcc lowers a / d to: v_div_scale of the divisor and of
// the numerator, v_rcp of the scaled divisor, Newton refinement of that
// reciprocal (one step in f32, two in f64), quotient and remainder
// corrections, v_div_fmas and v_div_fixup.  When |a| and |d| are normal and
// lie in [2^-40, 2^40) (f32) or [2^-300, 2^300) (f64), v_div_scale leaves
// both operands as they are (the exponent gap stays below 96 / 768 and
// nothing approaches the denormal range) and clears the scale flag, so
// v_div_fmas is a plain fma, and v_div_fixup returns the quotient unchanged:
// the sequence is exactly rcp_refined followed by div_steps on the same
// operands, hence the same bits.  div_range_ok tells whether an operand is
// in that range.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool div_range_ok(float a) {
  const unsigned u = __float_as_uint(a) & 0x7fffffffu;
  return u - 0x2B800000u < 0x53800000u - 0x2B800000u;  // 2^-40 <= |a| < 2^40
}
__device__ __forceinline__ bool div_range_ok(double a) {
  const unsigned u = (unsigned)((unsigned long long)__double_as_longlong(a) >> 32) & 0x7fffffffu;
  return u - 0x2D300000u < 0x52B00000u - 0x2D300000u;  // 2^-300 <= |a| < 2^300
}
__device__ __forceinline__ float rcp_refined(float d) {
  const float r = __builtin_amdgcn_rcpf(d);
  return fmaf(fmaf(-d, r, 1.0f), r, r);
}
__device__ __forceinline__ double rcp_refined(double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = fma(r, fma(-d, r, 1.0), r);
  return fma(r, fma(-d, r, 1.0), r);
}
__device__ __forceinline__ float div_steps(float a, float d, float y) {
  const float m = a * y;
  const float q = fmaf(fmaf(-d, m, a), y, m);
  return fmaf(fmaf(-d, q, a), y, q);
}
__device__ __forceinline__ double div_steps(double a, double d, double y) {
  const double m = a * y;
  return fma(fma(-d, m, a), y, m);
}
// a numerator of +0 also gives the quotient's bits: every step above is
// then an exact zero of the right sign
template <class T>
__device__ __forceinline__ bool div_steps_exact(T a, T d) {
  return div_range_ok(d) && (div_range_ok(a) || (a == (T)0 && !signbit(a)));
}

// A load of matrix data, which a kernel reads once.  NT: nontemporal, so
// that it does not displace from the caches (the Infinity Cache in front of
// HBM, above all) the vector lines that are gathered or read again; chosen
// when the matrix and those vectors do not fit in that cache together.
template <bool NT, class T>
__device__ __forceinline__ T ldm(const T *p) {
  if constexpr (NT)
    return __builtin_nontemporal_load(p);
  else
    return *p;
}

// 4-element vector loads/stores (i multiple of 4; HBM buffers 256-B aligned)
__device__ __forceinline__ void ld4(const double *p, int i, double (&o)[4]) {
  const double2 u = *reinterpret_cast<const double2 *>(p + i);
  const double2 w = *reinterpret_cast<const double2 *>(p + i + 2);
  o[0] = u.x; o[1] = u.y; o[2] = w.x; o[3] = w.y;
}
__device__ __forceinline__ void ld4(const float *p, int i, double (&o)[4]) {
  const float4 u = *reinterpret_cast<const float4 *>(p + i);
  o[0] = u.x; o[1] = u.y; o[2] = u.z; o[3] = u.w;
}
__device__ __forceinline__ void st4(double *p, int i, const double (&o)[4]) {
  *reinterpret_cast<double2 *>(p + i) = make_double2(o[0], o[1]);
  *reinterpret_cast<double2 *>(p + i + 2) = make_double2(o[2], o[3]);
}
// The same, nontemporal: for vectors that no later kernel of the iteration
// reads, so that they do not displace the ones that are read again
typedef double dvec2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void ld4_nt(const double *p, int i, double (&o)[4]) {
  const dvec2 u = __builtin_nontemporal_load(reinterpret_cast<const dvec2 *>(p + i));
  const dvec2 w = __builtin_nontemporal_load(reinterpret_cast<const dvec2 *>(p + i + 2));
  o[0] = u.x; o[1] = u.y; o[2] = w.x; o[3] = w.y;
}
__device__ __forceinline__ void st4_nt(double *p, int i, const double (&o)[4]) {
  const dvec2 u = {o[0], o[1]}, w = {o[2], o[3]};
  __builtin_nontemporal_store(u, reinterpret_cast<dvec2 *>(p + i));
  __builtin_nontemporal_store(w, reinterpret_cast<dvec2 *>(p + i + 2));
}

// ---------------------------------------------------------------------------
// Sliced-ELL row chains
// ---------------------------------------------------------------------------
// The lane's row of slice s as registers: columns (padding -1) and values.
// SEL: the loads of a slice with entries are issued without a branch
// (entries k >= w load entry 0 again, in a line already fetched, and become
// padding); faster for the batch SpMV, slower for the batch sweeps.
template <class VT, int WM, bool SEL = false>
__device__ __forceinline__ void sell_load(const int *__restrict__ off, const int *__restrict__ col,
                                          const VT *__restrict__ v, int s, int lane,
                                          int (&cc)[WM], VT (&vv)[WM]) {
  const int o0 = off[s];
  const int w = (off[s + 1] - o0) >> 6;
  const int base = o0 + lane;
#pragma unroll
  for (int k = 0; k < WM; ++k) {
    cc[k] = -1;
    vv[k] = (VT)0;
    if (!SEL && k < w) {
      cc[k] = ldm<true>(&col[base + 64 * k]);
      vv[k] = ldm<true>(&v[base + 64 * k]);
    }
  }
  if (SEL && w > 0) {  // a slice of rows without entries: entry 0 may not exist
#pragma unroll
    for (int k = 0; k < WM; ++k) cc[k] = ldm<true>(&col[k < w ? base + 64 * k : base]);
#pragma unroll
    for (int k = 0; k < WM; ++k) vv[k] = ldm<true>(&v[k < w ? base + 64 * k : base]);
#pragma unroll
    for (int k = 0; k < WM; ++k) cc[k] = k < w ? cc[k] : -1;
  }
}

// acc op= sum_k vv_k * x[cc_k] in ascending k over loaded entries: the same
// operations, in the same order, as sell_chain
template <class AT, class XT, class VT, int WM, bool NEG>
__device__ __forceinline__ AT sell_apply(const int (&cc)[WM], const VT (&vv)[WM],
                                         const XT *__restrict__ x, AT acc) {
  AT xv[WM];
#pragma unroll
  for (int k = 0; k < WM; ++k) xv[k] = cc[k] >= 0 ? (AT)x[cc[k]] : (AT)0;
#pragma unroll
  for (int k = 0; k < WM; ++k)
    if (cc[k] >= 0) acc = NEG ? fma_t(-(AT)vv[k], xv[k], acc) : fma_t((AT)vv[k], xv[k], acc);
  return acc;
}

// sell_apply with the gathers and the sum issued without a per-entry branch:
// a row's padding follows its entries, so a row whose first entry is padding
// has none; the others gather x[0] for padding and drop it by a select.  Same
// result; fewer VGPRs where the caller also carries dot partials (the batch
// SpMV: 82 -> 70 at WM = 8), more in the batch sweeps, which keep sell_apply.
template <class AT, class XT, class VT, int WM, bool NEG>
__device__ __forceinline__ AT sell_apply_sel(const int (&cc)[WM], const VT (&vv)[WM],
                                             const XT *__restrict__ x, AT acc) {
  if (cc[0] < 0) return acc;
  AT xv[WM];
#pragma unroll
  for (int k = 0; k < WM; ++k) xv[k] = (AT)x[cc[k] >= 0 ? cc[k] : 0];
#pragma unroll
  for (int k = 0; k < WM; ++k) {
    const AT t = NEG ? fma_t(-(AT)vv[k], xv[k], acc) : fma_t((AT)vv[k], xv[k], acc);
    acc = cc[k] >= 0 ? t : acc;
  }
  return acc;
}

// The lane's row sum 0 + v_0 x[c_0] + v_1 x[c_1] + ... (ascending stored
// order, one fma per entry) over slice s of A's 16-bit layout, stored
// paired (pair_slices): entries 2j and 2j+1 of a lane sit side by side, so
// one dword load brings both 16-bit columns and one 16-byte load both
// values; a slice of odd width keeps its last entry in place, applied after
// the pairs (the pairs past the width are padding, so the order of the
// terms is unchanged).  spmv16_load issues a slice's loads (for a
// wave-uniform s: the slice header comes from scalar loads); spmv16_finish
// puts the slice's column-block table in LDS, gathers x[c] (x[0] for
// padding) and runs the chain; padding terms are skipped by a select, as
// the reference's CSR loop never sees them.  A wavefront issues the next
// slice's loads before the current slice's gathers (at 192^3: 138.3 ->
// 129.0 us per launch; the next slice's columns alone, 130.9 us at 7 waves
// per SIMD against 5).
typedef double dbl2_t __attribute__((ext_vector_type(2)));
template <int WM>
struct Slice16 {
  int w, te, ts, tn;  // te: table entry `lane` of the tn at tbl + ts
  unsigned cp[WM / 2];
  dbl2_t vp[WM / 2];
  unsigned ct;
  double vt;
};

// tbl / toff: Sell's tables (TV: toff is set; else 16 entries per slice)
template <int WM, bool TV>
__device__ __forceinline__ void spmv16_load(const int *__restrict__ off,
                                            const uint16_t *__restrict__ col16,
                                            const int *__restrict__ tbl,
                                            const int *__restrict__ toff,
                                            const double *__restrict__ v, int s, int lane,
                                            Slice16<WM> &q) {
  constexpr int WP = WM / 2;
  const int o0 = off[s];
  const int w = (off[s + 1] - o0) >> 6;
  const int np = w >> 1;
  q.w = w;
  if (TV) {
    q.ts = toff[s];
    q.tn = toff[s + 1] - q.ts;
    q.te = lane < q.tn ? tbl[q.ts + lane] : 0;
  } else {
    q.ts = 16 * s;
    q.tn = 16;
    q.te = tbl[16 * s + (lane & 15)];
  }
  const unsigned *c32 = reinterpret_cast<const unsigned *>(col16 + o0);  // o0 is a multiple of 64
  const dbl2_t *v2 = reinterpret_cast<const dbl2_t *>(v + o0);
#pragma unroll
  for (int j = 0; j < WP; ++j) {
    q.cp[j] = 0xFFFFFFFFu;
    q.vp[j] = dbl2_t{0.0, 0.0};
    if (j < np) {
      q.cp[j] = ldm<true>(&c32[64 * j + lane]);
      q.vp[j] = ldm<true>(&v2[64 * j + lane]);
    }
  }
  q.ct = 0xFFFFu;
  q.vt = 0.0;
  if (w & 1) {
    q.ct = ldm<true>(&col16[o0 + 128 * np + lane]);
    q.vt = ldm<true>(&v[o0 + 128 * np + lane]);
  }
}

// tb: the wavefront's LDS slots for the table (kTbl16Max with TV, else 16);
// TV: Sell::toff is set (a slice's table beyond its first 64 entries is
// read here)
template <int WM, bool TV>
__device__ __forceinline__ double spmv16_finish(const Slice16<WM> &q, int *tb,
                                                const int *__restrict__ tbl,
                                                const double *__restrict__ x, int lane) {
  constexpr int WP = WM / 2;
  if (q.w == 0) return 0.0;  // a slice of empty rows (uniform)
  // index bits of padding (0xFFFF) select a valid slot; its value is unused
  constexpr unsigned tm = TV ? kTbl16Max - 1u : 15u;
  wave_sync();  // the previous slice's table reads are issued
  if (TV) {
    tb[lane] = q.te;
    for (int k = 64; k < q.tn; k += 64) tb[k + lane] = k + lane < q.tn ? tbl[q.ts + k + lane] : 0;
  } else {
    tb[lane & 15] = q.te;
  }
  wave_sync();
  int cl[WP], ch[WP], t0[WP], t1[WP];
#pragma unroll
  for (int j = 0; j < WP; ++j) {
    t0[j] = tb[(q.cp[j] >> 6) & tm];
    t1[j] = tb[(q.cp[j] >> 22) & tm];
  }
  const int tt = tb[(q.ct >> 6) & tm];
  // every lane reads its table entries, all issued before the first is used
#pragma unroll
  for (int j = 0; j < WP; ++j) asm volatile("" : "+v"(t0[j]), "+v"(t1[j]));
#pragma unroll
  for (int j = 0; j < WP; ++j) {
    const unsigned lo = q.cp[j] & 0xFFFFu, hi = q.cp[j] >> 16;
    cl[j] = lo == 0xFFFFu ? -1 : (t0[j] << 6) | (int)(lo & 63u);
    ch[j] = hi == 0xFFFFu ? -1 : (t1[j] << 6) | (int)(hi & 63u);
  }
  const int ctl = q.ct == 0xFFFFu ? -1 : (tt << 6) | (int)(q.ct & 63u);
  double xl[WP], xh[WP];
#pragma unroll
  for (int j = 0; j < WP; ++j) {
    xl[j] = x[cl[j] >= 0 ? cl[j] : 0];
    xh[j] = x[ch[j] >= 0 ? ch[j] : 0];
  }
  const double xt = x[ctl >= 0 ? ctl : 0];
  double acc = 0.0;
#pragma unroll
  for (int j = 0; j < WP; ++j) {
    const double a = fma(q.vp[j].x, xl[j], acc);
    acc = cl[j] >= 0 ? a : acc;
    const double b = fma(q.vp[j].y, xh[j], acc);
    acc = ch[j] >= 0 ? b : acc;
  }
  const double c = fma(q.vt, xt, acc);
  return ctl >= 0 ? c : acc;
}

// sell_load of A's values in the paired layout (pair_slices): entry k < 2
// (w / 2) of a lane at o0 + 128 (k / 2) + 2 lane + k % 2, an odd last entry
// at o0 + 64 k + lane; columns in the plain layout.  Entries k >= w load
// entry 0 again and get column -1 (the branch-free form of sell_load).
template <class VT, int WM>
__device__ __forceinline__ void sell_load_pairv(const int *__restrict__ off,
                                                const int *__restrict__ col,
                                                const VT *__restrict__ v, int s, int lane,
                                                int (&cc)[WM], VT (&vv)[WM]) {
  const int o0 = off[s];
  const int w = (off[s + 1] - o0) >> 6;
#pragma unroll
  for (int k = 0; k < WM; ++k) {
    cc[k] = -1;
    vv[k] = (VT)0;
  }
  if (w == 0) return;  // a slice of rows without entries: entry 0 may not exist
  const int np2 = (w >> 1) << 1;
  const int base = o0 + lane;
#pragma unroll
  for (int k = 0; k < WM; ++k) cc[k] = ldm<true>(&col[k < w ? base + 64 * k : base]);
#pragma unroll
  for (int k = 0; k < WM; ++k) {
    const int kk = k < w ? k : 0;
    const int q = kk < np2 ? o0 + 128 * (kk >> 1) + 2 * lane + (kk & 1) : base + 64 * kk;
    vv[k] = ldm<true>(&v[q]);
  }
#pragma unroll
  for (int k = 0; k < WM; ++k) cc[k] = k < w ? cc[k] : -1;
}

// sell_chain (WM > 0) for a wave-uniform s, with every load issued without a
// branch: entries k >= w load entry 0 again (a line already fetched) and
// padding gathers x[0], both dropped by a select, as in spmv16_finish.  The
// branchy form, once the kernel also carries an r.z partial, compiled to
// copies of the whole cc/vv arrays on every path (134 VGPRs at WM = 8).
template <class VT, class AT, class XT, int WM, bool NEG, bool NT>
__device__ __forceinline__ AT sell_chain_u(const int *__restrict__ off,
                                           const int *__restrict__ col,
                                           const VT *__restrict__ v, const XT *__restrict__ x,
                                           int s, int lane, AT acc) {
  static_assert(WM > 0, "sell_chain_u needs a fixed slice width bound");
  const int o0 = off[s];
  const int w = (off[s + 1] - o0) >> 6;
  if (w == 0) return acc;  // a slice of rows without entries: entry 0 may not exist
  const int base = o0 + lane;
  int cc[WM];
  VT vv[WM];
#pragma unroll
  for (int k = 0; k < WM; ++k) cc[k] = ldm<NT>(&col[k < w ? base + 64 * k : base]);
#pragma unroll
  for (int k = 0; k < WM; ++k) vv[k] = ldm<NT>(&v[k < w ? base + 64 * k : base]);
#pragma unroll
  for (int k = 0; k < WM; ++k) cc[k] = k < w ? cc[k] : -1;
  AT xv[WM];
#pragma unroll
  for (int k = 0; k < WM; ++k) xv[k] = (AT)x[cc[k] >= 0 ? cc[k] : 0];
#pragma unroll
  for (int k = 0; k < WM; ++k) {
    const AT t = NEG ? fma_t(-(AT)vv[k], xv[k], acc) : fma_t((AT)vv[k], xv[k], acc);
    acc = cc[k] >= 0 ? t : acc;
  }
  return acc;
}

// acc op= sum_k v_k * x[c_k] over the entries of the lane's row in slice s,
// ascending k; NEG selects acc - v*x (triangular solves) over acc + v*x.
template <class VT, class AT, class XT, int WM, bool NEG, bool NT = false>
__device__ __forceinline__ AT sell_chain(const int *__restrict__ off, const int *__restrict__ col,
                                         const VT *__restrict__ v, const XT *__restrict__ x,
                                         int s, int lane, AT acc) {
  const int o0 = off[s];
  const int w = (off[s + 1] - o0) >> 6;
  const int base = o0 + lane;
  if (WM > 0) {
    int cc[WM > 0 ? WM : 1];
    VT vv[WM > 0 ? WM : 1];
#pragma unroll
    for (int k = 0; k < WM; ++k) {
      cc[k] = -1;
      vv[k] = (VT)0;
      if (k < w) {
        cc[k] = ldm<NT>(&col[base + 64 * k]);
        vv[k] = ldm<NT>(&v[base + 64 * k]);
      }
    }
    AT xv[WM > 0 ? WM : 1];
#pragma unroll
    for (int k = 0; k < WM; ++k) xv[k] = cc[k] >= 0 ? (AT)x[cc[k]] : (AT)0;
#pragma unroll
    for (int k = 0; k < WM; ++k)
      if (cc[k] >= 0) acc = NEG ? fma_t(-(AT)vv[k], xv[k], acc) : fma_t((AT)vv[k], xv[k], acc);
  } else {
    for (int k = 0; k < w; ++k) {
      const int c = ldm<NT>(&col[base + 64 * k]);
      if (c >= 0) {
        const AT a = (AT)ldm<NT>(&v[base + 64 * k]);
        acc = NEG ? fma_t(-a, (AT)x[c], acc) : fma_t(a, (AT)x[c], acc);
      }
    }
  }
  return acc;
}

// ---------------------------------------------------------------------------
// ABMC sweeps of one block of at most 64 rows (ic0.cpp:263-339) by one
// wavefront.  The block's rows are the contiguous new ids [bs, be) and lane
// t owns row bs + t; my is the wave's 64-entry LDS window.
//   forward : entries with column < bs belong to earlier colors (final) and
//             come first in the ascending row; they are summed before the
//             level loop.  In-block entries [bs, i) are read from LDS as the
//             in-block levels complete.
//   backward: in-block entries (i, be) come first, then columns >= be of
//             later colors, whose values are prefetched before the levels.
// The per-row operation order equals the sequential reference, bit for bit.
// Both store the block's outputs once, coalesced, after the level loop.
// ---------------------------------------------------------------------------
template <class VT, class YT, int WM, bool NT = false>
__device__ __forceinline__ void abmc_fwd_block(int bs, int be, int nl,
                                               const uint16_t *__restrict__ lev,
                                               const int *__restrict__ off,
                                               const int *__restrict__ col,
                                               const VT *__restrict__ lv,
                                               const VT *__restrict__ diag,
                                               const double *__restrict__ r, YT *__restrict__ y,
                                               YT *my, int lane) {
  const int i = bs + lane;
  const bool act = i < be;
  int cc[WM];
  VT vv[WM];
  YT acc = (YT)0, d = (YT)1;
  int mylev = -1;
  if (act) {
    const int s = i >> 6, ln = i & 63;
    const int o0 = off[s];
    const int wd = (off[s + 1] - o0) >> 6;
    const int base = o0 + ln;
#pragma unroll
    for (int k = 0; k < WM; ++k) {
      cc[k] = -1;
      vv[k] = (VT)0;
      if (k < wd) {
        cc[k] = ldm<NT>(&col[base + 64 * k]);
        vv[k] = ldm<NT>(&lv[base + 64 * k]);
      }
    }
    mylev = ldm<NT>(&lev[i]);
    d = (YT)diag[i];
    acc = (YT)r[i];
    YT xo[WM];
#pragma unroll
    for (int k = 0; k < WM; ++k) xo[k] = (cc[k] >= 0 && cc[k] < bs) ? y[cc[k]] : (YT)0;
#pragma unroll
    for (int k = 0; k < WM; ++k)
      if (cc[k] >= 0 && cc[k] < bs) acc = fma_t(-(YT)vv[k], xo[k], acc);
  }
  YT yi = (YT)0;
  for (int l = 0; l < nl; ++l) {
    if (mylev == l) {
#pragma unroll
      for (int k = 0; k < WM; ++k)
        if (cc[k] >= bs) acc = fma_t(-(YT)vv[k], my[cc[k] - bs], acc);
      yi = acc / d;
      my[lane] = yi;
    }
    wave_sync();
  }
  if (act) y[i] = yi;
}

// returns z_i (0 on lanes past the block); with r, also r_i in ri
template <class VT, class YT, int WM, bool NT = false>
__device__ __forceinline__ YT abmc_bwd_block(int bs, int be, int nl,
                                             const uint16_t *__restrict__ lev,
                                             const int *__restrict__ off,
                                             const int *__restrict__ col,
                                             const VT *__restrict__ uv,
                                             const VT *__restrict__ diag,
                                             const YT *__restrict__ y, YT *__restrict__ z,
                                             const double *__restrict__ r, double &ri, YT *my,
                                             int lane) {
  const int i = bs + lane;
  const bool act = i < be;
  int cc[WM];
  VT vv[WM];
  YT xo[WM];
  YT acc = (YT)0, d = (YT)1, zi = (YT)0;
  int mylev = -1;
  ri = 0.0;
  if (act) {
    const int s = i >> 6, ln = i & 63;
    const int o0 = off[s];
    const int wd = (off[s + 1] - o0) >> 6;
    const int base = o0 + ln;
#pragma unroll
    for (int k = 0; k < WM; ++k) {
      cc[k] = -1;
      vv[k] = (VT)0;
      if (k < wd) {
        cc[k] = ldm<NT>(&col[base + 64 * k]);
        vv[k] = ldm<NT>(&uv[base + 64 * k]);
      }
    }
#pragma unroll
    for (int k = 0; k < WM; ++k) xo[k] = cc[k] >= be ? z[cc[k]] : (YT)0;
    mylev = ldm<NT>(&lev[i]);
    d = (YT)diag[i];
    acc = y[i];
    if (r) ri = r[i];
  }
  for (int lv = 0; lv < nl; ++lv) {
    if (mylev == lv) {
#pragma unroll
      for (int k = 0; k < WM; ++k)
        if (cc[k] >= 0 && cc[k] < be) acc = fma_t(-(YT)vv[k], my[cc[k] - bs], acc);
#pragma unroll
      for (int k = 0; k < WM; ++k)
        if (cc[k] >= be) acc = fma_t(-(YT)vv[k], xo[k], acc);
      zi = acc / d;
      my[lane] = zi;
    }
    wave_sync();
  }
  if (act) z[i] = zi;
  return zi;
}

// ---------------------------------------------------------------------------
// ABMC sweeps over the sectioned factor layout.  Slice b holds block b's
// rows (at most 64, lane = row - bstart[b]).  Entry k < S of a row is its
// k-th in-block entry (columns in [bs, be), ascending; the kernels read their
// rows within the block from the packed word Sell::inb, and the value, 0
// past the last), entry S + j (j < SO, SO >= S outside entries at most) its
// j-th entry outside the block (ascending, column -1 past the last): for L
// the earlier colors' columns (< bs), for U the later colors' (>= be).  A row's chain is the sequential reference's
// (ic0.cpp:263-339): forward, the outside terms first, then the in-block
// terms; backward, the in-block terms first, then the outside terms.
//
// The in-block levels run without masks: at every level every lane
// recomputes its row from the LDS window and stores it.  A term whose entry
// is absent reads the window's zero slot (64) with a zero coefficient, and
// fma(-0, 0, acc) returns acc, so each lane performs the same operations in
// the same order at every level.  After level L the rows of levels <= L are
// final (their in-block inputs were final after level L - 1), so after the
// block's nl levels every row is.  The division runs as rcp_refined once per
// row and div_steps at each level; if a final numerator of a live row is
// outside div_steps_exact's range, the block's levels run again with a / d.
// ---------------------------------------------------------------------------
constexpr int kSecWindow = 65;  // LDS slots per wavefront: 64 rows + the zero slot

template <class YT, int S, int SO, bool OUT>
__device__ __forceinline__ YT sec_levels(int nl, YT acc, YT d, const int (&ad)[S],
                                         const YT (&va)[S], const YT (&vo)[SO],
                                         const YT (&xo)[SO], YT *my, int lane, bool act) {
  const YT dy = rcp_refined(d);
  YT yi = (YT)0, a = acc;
  for (int l = 0; l < nl; ++l) {
    YT t[S];
#pragma unroll
    for (int k = 0; k < S; ++k) t[k] = my[ad[k]];
    a = acc;
#pragma unroll
    for (int k = 0; k < S; ++k) a = fma_t(-va[k], t[k], a);
    if (OUT) {
#pragma unroll
      for (int k = 0; k < SO; ++k) a = fma_t(-vo[k], xo[k], a);
    }
    yi = div_steps(a, d, dy);
    my[lane] = yi;
    wave_sync();
  }
  if (!__all(!act || nl == 0 || div_steps_exact(a, d))) {
    for (int l = 0; l < nl; ++l) {
      YT t[S];
#pragma unroll
      for (int k = 0; k < S; ++k) t[k] = my[ad[k]];
      a = acc;
#pragma unroll
      for (int k = 0; k < S; ++k) a = fma_t(-va[k], t[k], a);
      if (OUT) {
#pragma unroll
        for (int k = 0; k < SO; ++k) a = fma_t(-vo[k], xo[k], a);
      }
      yi = a / d;
      my[lane] = yi;
      wave_sync();
    }
  }
  return yi;
}

// One lane's raw loads for a block, issued without branches so that they
// stay in flight while the previous block runs: lanes past the block end and
// absent outside entries load from in-bounds stand-in positions, and
// sec_take masks them out.
template <class VT, class YT, int S, int SO>
struct SecRaw {
  unsigned inb;  // packed in-block rows (Sell::inb)
  int cout_[SO];
  VT vin[S], vout[SO];
  VT d;
  double a;   // r_i (forward) or y_i widened (backward)
  double ri;  // backward with r
  double ap;  // forward with the r update: (Ap)_i
};

// LOAD_NOW: the caller uses the loads at once, so absent outside entries
// are skipped by (uniform) branches instead of loaded from stand-ins.
// RU (forward): also (Ap)_i, for the r update that the sweep applies.
template <class VT, class YT, int S, int SO, bool FWD, bool NT, bool LOAD_NOW = false,
          bool RU = false>
__device__ __forceinline__ void sec_fetch(int kb, int o0, int w, int bs, int be, int lane,
                                          const unsigned *__restrict__ inb,
                                          const int *__restrict__ col, const VT *__restrict__ v,
                                          const VT *__restrict__ diag,
                                          const double *__restrict__ r,
                                          const YT *__restrict__ yin, SecRaw<VT, YT, S, SO> &e,
                                          const double *__restrict__ Ap = nullptr) {
  const int t = min(lane, be - bs - 1);  // a row of the block
  const int base = o0 + t;
  e.inb = ldm<NT>(&inb[64 * kb + t]);
#pragma unroll
  for (int k = 0; k < S; ++k) e.vin[k] = ldm<NT>(&v[base + 64 * k]);  // absent entries hold 0
#pragma unroll
  for (int k = 0; k < SO; ++k) {
    if (LOAD_NOW) {
      e.cout_[k] = -1;
      e.vout[k] = (VT)0;
      if (S + k < w) {
        e.cout_[k] = ldm<NT>(&col[base + 64 * (S + k)]);
        e.vout[k] = ldm<NT>(&v[base + 64 * (S + k)]);
      }
    } else {
      // absent: stand-ins in lines this fetch reads anyway (the packed word,
      // the first in-block value), so they cost no extra traffic
      const bool here = S + k < w;
      const int *cp = here ? col + base + 64 * (S + k) : (const int *)inb + 64 * kb + t;
      e.cout_[k] = ldm<NT>(cp);
      e.vout[k] = ldm<NT>(&v[here ? base + 64 * (S + k) : base]);
    }
  }
  e.d = diag[bs + t];
  e.a = FWD ? r[bs + t] : (double)yin[bs + t];
  e.ri = (!FWD && r) ? r[bs + t] : 0.0;
  e.ap = RU ? ldm<true>(&Ap[bs + t]) : 0.0;  // Ap: read once
}

// Lanes past the block end hold a copy of the block's last row (sec_fetch
// clamps them): they compute it again into their own window slots, which no
// row of the block reads, and their results are neither stored nor checked.
template <class VT, class YT, int S, int SO>
__device__ __forceinline__ void sec_take(const SecRaw<VT, YT, S, SO> &e, int w, int (&ad)[S],
                                         YT (&va)[S], int (&co)[SO], YT (&vo)[SO], YT &d,
                                         YT &acc) {
#pragma unroll
  for (int k = 0; k < S; ++k) {
    ad[k] = (int)((e.inb >> (8 * k)) & 0xFFu);
    va[k] = (YT)e.vin[k];
  }
#pragma unroll
  for (int k = 0; k < SO; ++k) {
    const bool out = S + k < w && e.cout_[k] >= 0;
    co[k] = out ? e.cout_[k] : -1;
    vo[k] = out ? (YT)e.vout[k] : (YT)0;
  }
  d = (YT)e.d;
  acc = (YT)e.a;
}

}  // namespace iccg
