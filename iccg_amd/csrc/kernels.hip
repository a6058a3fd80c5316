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
// Compiled with -ffp-contract=off: only the fma() calls written here fuse.
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

__device__ __forceinline__ double dd_round(double h, double l) { return h + l; }

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

// Block reduction of NC double-double sums; the workgroup that arrives last
// folds every workgroup's partial in workgroup order.  Returns true on
// thread 0 of that workgroup, with out[] = the rounded totals.
template <int NC>
__device__ bool reduce_last(double (&h)[NC], double (&l)[NC], double *part,
                            unsigned int *arrive, double (&out)[NC]) {
  __shared__ double sh[NC][2][kWG / 64];
  __shared__ int s_last;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  wave_dd_reduce<NC>(h, l);
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
      part[(blockIdx.x * NC + c) * 2 + 0] = ah;
      part[(blockIdx.x * NC + c) * 2 + 1] = al;
    }
    __threadfence();
    unsigned int t = atomicAdd(arrive, 1u);
    s_last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!s_last) return false;
  __threadfence();
  double h2[NC], l2[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) h2[c] = l2[c] = 0.0;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const volatile double *q = part + (b * NC + c) * 2;
      dd_add(h2[c], l2[c], q[0], q[1]);
    }
  }
  wave_dd_reduce<NC>(h2, l2);
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      sh[c][0][w] = h2[c];
      sh[c][1][w] = l2[c];
    }
  }
  __syncthreads();
  if (threadIdx.x != 0) return false;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    double ah = sh[c][0][0], al = sh[c][1][0];
    for (int k = 1; k < nw; ++k) dd_add(ah, al, sh[c][0][k], sh[c][1][k]);
    out[c] = dd_round(ah, al);
  }
  *arrive = 0u;
  return true;
}

static inline int red_grid(long n) {
  long g = (n + kWG - 1) / kWG;
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

__global__ __launch_bounds__(kWG) void k_spmv_plain(int n, const int *__restrict__ rp,
                                                    const int *__restrict__ ci,
                                                    const double *__restrict__ v,
                                                    const double *__restrict__ x,
                                                    double *__restrict__ y) {
  for (int i = blockIdx.x * kWG + threadIdx.x; i < n; i += gridDim.x * kWG)
    y[i] = row_dot(rp, ci, v, x, i);
}

// Ap = A p with p.Ap, ||p||^2, ||Ap||^2; the last workgroup applies the
// breakdown test (cg_crs.cpp:127-133) and computes alpha (135).
__global__ __launch_bounds__(kWG) void k_spmv_cg(int n, const int *__restrict__ rp,
                                                 const int *__restrict__ ci,
                                                 const double *__restrict__ v,
                                                 const double *__restrict__ p,
                                                 double *__restrict__ Ap, Scal *S,
                                                 double *part) {
  if (S->done) return;
  double h[3] = {0, 0, 0}, l[3] = {0, 0, 0};
  for (int i = blockIdx.x * kWG + threadIdx.x; i < n; i += gridDim.x * kWG) {
    double y = row_dot(rp, ci, v, p, i);
    Ap[i] = y;
    double pi = p[i];
    dd_fma(h[0], l[0], pi, y);
    dd_fma(h[1], l[1], pi, pi);
    dd_fma(h[2], l[2], y, y);
  }
  double out[3];
  if (reduce_last<3>(h, l, part, &S->arrive[SLOT_SPMV], out)) {
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
__global__ __launch_bounds__(kWG) void k_resid(int n, const int *__restrict__ rp,
                                               const int *__restrict__ ci,
                                               const double *__restrict__ v,
                                               const double *__restrict__ x,
                                               const double *__restrict__ b,
                                               double *__restrict__ r, Scal *S, double *part,
                                               int mode, double tol, int max_iter) {
  double h[2] = {0, 0}, l[2] = {0, 0};
  for (int i = blockIdx.x * kWG + threadIdx.x; i < n; i += gridDim.x * kWG) {
    double ax = row_dot(rp, ci, v, x, i);
    double bi = b[i];
    double ri = bi - ax;
    if (mode != 0) r[i] = ri;
    dd_fma(h[0], l[0], bi, bi);
    dd_fma(h[1], l[1], ri, ri);
  }
  double out[2];
  if (reduce_last<2>(h, l, part, &S->arrive[SLOT_RESID], out)) {
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
  for (int i = blockIdx.x * kWG + threadIdx.x; i < n; i += gridDim.x * kWG) {
    double xi = fma(alpha, p[i], x[i]);
    double ri = fma(-alpha, Ap[i], r[i]);
    x[i] = xi;
    r[i] = ri;
    dd_fma(h[0], l[0], ri, ri);
    if (PK == PK_JAC) {
      double zi = inv_diag[i] * ri;
      z[i] = zi;
      dd_fma(h[1], l[1], ri, zi);
    }
  }
  double out[2];
  if (reduce_last<2>(h, l, part, &S->arrive[SLOT_XR], out)) {
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
    double ri = r[i];
    double zi = (PK == PK_JAC) ? inv_diag[i] * ri : ri;
    if (PK == PK_JAC) z[i] = zi;
    p[i] = zi;
    dd_fma(h[0], l[0], ri, zi);
  }
  double out[1];
  if (reduce_last<1>(h, l, part, &S->arrive[SLOT_PINIT], out)) S->rz_old = out[0];
}

// rz_new = r.z; beta = rz_new / rz_old (cg_crs.cpp:152-153), or rz_old at
// start (115)
template <class ZT>
__global__ __launch_bounds__(kWG) void k_rz(int n, const double *__restrict__ r,
                                            const ZT *__restrict__ z, Scal *S, double *part,
                                            int init) {
  if (S->done) return;
  double h[1] = {0}, l[1] = {0};
  for (int i = blockIdx.x * kWG + threadIdx.x; i < n; i += gridDim.x * kWG)
    dd_fma(h[0], l[0], r[i], (double)z[i]);
  double out[1];
  if (reduce_last<1>(h, l, part, &S->arrive[SLOT_RZ], out)) {
    if (init) {
      S->rz_old = out[0];
    } else {
      S->beta = out[0] / S->rz_old;
      S->rz_old = out[0];
    }
  }
}

// p = z + beta p (cg_crs.cpp:155-157), or p = z at start (114)
template <class ZT>
__global__ __launch_bounds__(kWG) void k_update_p(int n, const ZT *__restrict__ z,
                                                  double *__restrict__ p, const Scal *S,
                                                  int init) {
  if (S->done) return;
  const double beta = S->beta;
  for (int i = blockIdx.x * kWG + threadIdx.x; i < n; i += gridDim.x * kWG) {
    double zi = (double)z[i];
    p[i] = init ? zi : fma(beta, p[i], zi);
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
// Row operators of the triangular solves and the factorization
// ---------------------------------------------------------------------------
__device__ __forceinline__ float fma_t(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return fma(a, b, c); }

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

// ---------------------------------------------------------------------------
// Schedule drivers
// ---------------------------------------------------------------------------
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

// ABMC: one workgroup per block; blocks of one color run concurrently; rows
// of a block follow their in-block dependency levels.
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

template <class F>
static void run_forward_plan(const DevSched &sc, F f, const Scal *S, hipStream_t st,
                             bool backward) {
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
void launch_spmv_plain(const DevCsr &A, const double *x, double *y, hipStream_t st) {
  if (A.n == 0) return;
  hipLaunchKernelGGL(k_spmv_plain, dim3(stream_grid(A.n)), dim3(kWG), 0, st, A.n, A.rp, A.ci,
                     A.v, x, y);
}

void launch_spmv_cg(const DevCsr &A, const double *p, double *Ap, Scal *S, double *part,
                    hipStream_t st) {
  hipLaunchKernelGGL(k_spmv_cg, dim3(red_grid(A.n)), dim3(kWG), 0, st, A.n, A.rp, A.ci, A.v, p,
                     Ap, S, part);
}

void launch_resid(const DevCsr &A, const double *x, const double *b, double *r, Scal *S,
                  double *part, int mode, double tol, int max_iter, hipStream_t st) {
  hipLaunchKernelGGL(k_resid, dim3(red_grid(A.n)), dim3(kWG), 0, st, A.n, A.rp, A.ci, A.v, x, b,
                     r, S, part, mode, tol, max_iter);
}

void launch_xr(int pk, int n, double *x, const double *p, double *r, const double *Ap,
               const double *inv_diag, double *z, Scal *S, double *part, hipStream_t st) {
  dim3 g(red_grid(n)), b(kWG);
  if (pk == PK_IDENT)
    hipLaunchKernelGGL(k_xr<PK_IDENT>, g, b, 0, st, n, x, p, r, Ap, inv_diag, z, S, part);
  else if (pk == PK_JAC)
    hipLaunchKernelGGL(k_xr<PK_JAC>, g, b, 0, st, n, x, p, r, Ap, inv_diag, z, S, part);
  else
    hipLaunchKernelGGL(k_xr<PK_IC>, g, b, 0, st, n, x, p, r, Ap, inv_diag, z, S, part);
}

void launch_precond_diag_init(int pk, int n, const double *r, const double *inv_diag, double *z,
                              double *p, Scal *S, double *part, hipStream_t st) {
  dim3 g(red_grid(n)), b(kWG);
  if (pk == PK_JAC)
    hipLaunchKernelGGL(k_precond_diag_init<PK_JAC>, g, b, 0, st, n, r, inv_diag, z, p, S, part);
  else
    hipLaunchKernelGGL(k_precond_diag_init<PK_IDENT>, g, b, 0, st, n, r, inv_diag, z, p, S,
                       part);
}

void launch_rz(int n, const double *r, const void *z, bool z32, Scal *S, double *part, int init,
               hipStream_t st) {
  dim3 g(red_grid(n)), b(kWG);
  if (z32)
    hipLaunchKernelGGL(k_rz<float>, g, b, 0, st, n, r, (const float *)z, S, part, init);
  else
    hipLaunchKernelGGL(k_rz<double>, g, b, 0, st, n, r, (const double *)z, S, part, init);
}

void launch_update_p(int n, const void *z, bool z32, double *p, const Scal *S, int init,
                     hipStream_t st) {
  dim3 g(stream_grid(n)), b(kWG);
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

void launch_fwd(const DevSched &sc, const Factor &F, bool mixed, const double *r, void *y,
                const Scal *S, hipStream_t st) {
  if (mixed) {
    FwdRow<float, float> f{F.lrp, F.lci, F.lv32, r, (float *)y};
    run_forward_plan(sc, f, S, st, false);
  } else {
    FwdRow<double, double> f{F.lrp, F.lci, F.lv, r, (double *)y};
    run_forward_plan(sc, f, S, st, false);
  }
}

void launch_bwd(const DevSched &sc, const Factor &F, bool mixed, const void *y, void *z,
                const Scal *S, hipStream_t st) {
  if (mixed) {
    BwdRow<float, float> f{F.urp, F.uci, F.uv32, (const float *)y, (float *)z};
    run_forward_plan(sc, f, S, st, true);
  } else {
    BwdRow<double, double> f{F.urp, F.uci, F.uv, (const double *)y, (double *)z};
    run_forward_plan(sc, f, S, st, true);
  }
}

void launch_ic0_factor(const DevSched &sc, const Factor &F, double shift,
                       unsigned long long *err, hipStream_t st) {
  IC0Row f{F.lrp, F.lci, F.lv, shift, err};
  run_forward_plan(sc, f, (const Scal *)nullptr, st, false);
}

void launch_scatter_u(const Factor &F, bool mixed, hipStream_t st) {
  if (F.nnz == 0) return;
  hipLaunchKernelGGL(k_scatter_u, dim3(stream_grid(F.nnz)), dim3(kWG), 0, st, F.nnz, F.l2u, F.lv,
                     F.uv, mixed ? F.lv32 : nullptr, mixed ? F.uv32 : nullptr);
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
