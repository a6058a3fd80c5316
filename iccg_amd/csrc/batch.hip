// batch.hip -- multi-system ICCG: many independent SPD systems, each with
// its own matrix, IC(0) factor and Krylov vectors resident in HBM, advanced
// together so that every CG phase is ONE launch over all active systems.
//
// Work is described by item lists that name a system slot:
//   ChunkItem (slot, s0, s1): 64-row slices [s0, s1) of one system, one
//     workgroup -- SpMV, vector updates and dot-product partials;
//   ColorItem (slot, a0, a1): one wavefront of a triangular sweep over one
//     color of one system -- 64-row slices (multicolor) or blocks (ABMC).
// Every system keeps its own scalars (SysScal); a finish kernel with one
// workgroup per system folds that system's partials.  Converged systems
// skip their items, and systems can be admitted or retired between launches:
// the serving loop's continuous batching.  Each system keeps its own
// precision (FP64, or FP32 factor with FP64 CG) and its own ordering
// (multicolor or ABMC); per system, the arithmetic is the single-system
// solver's: same row chains, same double-double dots, so a system gives the
// same iterates in a batch as alone.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
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
// (A x)_i of the lane's row in slice s, from the f32 copy of A when every
// value is an exact float (same products, same fma chain)
template <int WM>
__device__ __forceinline__ double a_row(const SysDesc &D, const double *x, int s, int lane) {
  if (D.a_v32)
    return sell_chain<float, double, double, WM, false>(D.a_off, D.a_col, D.a_v32, x, s, lane,
                                                         0.0);
  return sell_chain<double, double, double, WM, false>(D.a_off, D.a_col, D.a_v64, x, s, lane,
                                                        0.0);
}

// r = b - A x with b.b and r.r partials (mode 1 / 2 write r; mode 0 only
// measures the true residual)
template <int WM>
__global__ __launch_bounds__(kWG) void kb_resid(const SysDesc *__restrict__ desc,
                                                const ChunkItem *__restrict__ items,
                                                double *part, int mode) {
  const ChunkItem c = items[blockIdx.x];
  const SysDesc &D = desc[c.sys];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double h[2] = {0, 0}, l[2] = {0, 0};
  for (int s = c.s0 + w; s < c.s1; s += kWG >> 6) {
    const double ax = a_row<WM>(D, D.x, s, lane);
    const int i = (s << 6) + lane;
    if (i < D.n) {
      const double bi = D.b[i], ri = bi - ax;
      if (mode != 0) D.r[i] = ri;
      dd_fma(h[0], l[0], bi, bi);
      dd_fma(h[1], l[1], ri, ri);
    }
  }
  block_partial<2>(h, l, part, blockIdx.x);
}

// Ap = A p with p.Ap, p.p, Ap.Ap partials
template <int WM>
__global__ __launch_bounds__(kWG) void kb_spmv(const SysDesc *__restrict__ desc,
                                               const ChunkItem *__restrict__ items,
                                               const SysScal *__restrict__ sc, double *part) {
  const ChunkItem c = items[blockIdx.x];
  double h[3] = {0, 0, 0}, l[3] = {0, 0, 0};
  if (!sc[c.sys].done) {
    const SysDesc &D = desc[c.sys];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int s = c.s0 + w; s < c.s1; s += kWG >> 6) {
      const double y = a_row<WM>(D, D.p, s, lane);
      const int i = (s << 6) + lane;
      if (i < D.n) {
        D.Ap[i] = y;
        const double pi = D.p[i];
        dd_fma(h[0], l[0], pi, y);
        dd_fma(h[1], l[1], pi, pi);
        dd_fma(h[2], l[2], y, y);
      }
    }
  }
  block_partial<3>(h, l, part, blockIdx.x);
}

// x += alpha p; r -= alpha Ap; r.r partial
__global__ __launch_bounds__(kWG) void kb_xr(const SysDesc *__restrict__ desc,
                                             const ChunkItem *__restrict__ items,
                                             const SysScal *__restrict__ sc, double *part) {
  const ChunkItem c = items[blockIdx.x];
  double h[1] = {0}, l[1] = {0};
  if (!sc[c.sys].done) {
    const SysDesc &D = desc[c.sys];
    const double alpha = sc[c.sys].alpha;
    const int end = min(D.n, c.s1 << 6);
    for (int i = (c.s0 << 6) + threadIdx.x; i < end; i += kWG) {
      const double xi = fma(alpha, D.p[i], D.x[i]);
      const double ri = fma(-alpha, D.Ap[i], D.r[i]);
      D.x[i] = xi;
      D.r[i] = ri;
      dd_fma(h[0], l[0], ri, ri);
    }
  }
  block_partial<1>(h, l, part, blockIdx.x);
}

template <class FT>
__device__ __forceinline__ void rz_rows(const SysDesc &D, int beg, int end, double (&h)[1],
                                        double (&l)[1]) {
  const FT *z = (const FT *)D.z;
  for (int i = beg; i < end; i += kWG) dd_fma(h[0], l[0], D.r[i], (double)z[i]);
}

__global__ __launch_bounds__(kWG) void kb_rz(const SysDesc *__restrict__ desc,
                                             const ChunkItem *__restrict__ items,
                                             const SysScal *__restrict__ sc, double *part) {
  const ChunkItem c = items[blockIdx.x];
  double h[1] = {0}, l[1] = {0};
  if (!sc[c.sys].done) {
    const SysDesc &D = desc[c.sys];
    const int beg = (c.s0 << 6) + threadIdx.x, end = min(D.n, c.s1 << 6);
    if (D.mixed)
      rz_rows<float>(D, beg, end, h, l);
    else
      rz_rows<double>(D, beg, end, h, l);
  }
  block_partial<1>(h, l, part, blockIdx.x);
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

__global__ __launch_bounds__(kWG) void kb_update_p(const SysDesc *__restrict__ desc,
                                                   const ChunkItem *__restrict__ items,
                                                   const SysScal *__restrict__ sc, int init) {
  const ChunkItem c = items[blockIdx.x];
  if (sc[c.sys].done) return;
  const SysDesc &D = desc[c.sys];
  const double beta = sc[c.sys].beta;
  const int beg = (c.s0 << 6) + threadIdx.x, end = min(D.n, c.s1 << 6);
  if (D.mixed)
    p_rows<float>(D, beg, end, beta, init);
  else
    p_rows<double>(D, beg, end, beta, init);
}

// ---- triangular sweeps (ic0.cpp:194-249 multicolor, 263-339 ABMC) --------
template <class FT, int WM>
__device__ __forceinline__ void mc_fwd_item(const SysDesc &D, const ColorItem &c, int lane) {
  FT *y = (FT *)D.y;
  const FT *lv = (const FT *)D.l_v, *dg = (const FT *)D.d;
  for (int s = c.a0; s < c.a1; ++s) {
    const int i = (s << 6) + lane;
    if (i < c.lo || i >= c.hi) continue;
    FT acc = (FT)D.r[i];
    acc = sell_chain<FT, FT, FT, WM, true>(D.l_off, D.l_col, lv, y, s, lane, acc);
    y[i] = acc / dg[i];
  }
}

template <class FT, int WM>
__device__ __forceinline__ void mc_bwd_item(const SysDesc &D, const ColorItem &c, int lane) {
  FT *z = (FT *)D.z;
  const FT *y = (const FT *)D.y, *uv = (const FT *)D.u_v, *dg = (const FT *)D.d;
  for (int s = c.a0; s < c.a1; ++s) {
    const int i = (s << 6) + lane;
    if (i < c.lo || i >= c.hi) continue;
    FT acc = y[i];
    acc = sell_chain<FT, FT, FT, WM, true>(D.u_off, D.u_col, uv, z, s, lane, acc);
    z[i] = acc / dg[i];
  }
}

template <class FT, int WM>
__device__ __forceinline__ void abmc_fwd_item(const SysDesc &D, const ColorItem &c, int lane,
                                              FT *my) {
  for (int kb = c.a0; kb < c.a1; ++kb)
    abmc_fwd_block<FT, FT, WM>(D.bstart[kb], D.bstart[kb + 1], D.nlev_f[kb], D.lev_f, D.l_off,
                               D.l_col, (const FT *)D.l_v, (const FT *)D.d, D.r, (FT *)D.y, my,
                               lane);
}

template <class FT, int WM>
__device__ __forceinline__ void abmc_bwd_item(const SysDesc &D, const ColorItem &c, int lane,
                                              FT *my) {
  double ri;
  for (int kb = c.a0; kb < c.a1; ++kb)
    (void)abmc_bwd_block<FT, FT, WM>(D.bstart[kb], D.bstart[kb + 1], D.nlev_b[kb], D.lev_b,
                                     D.u_off, D.u_col, (const FT *)D.u_v, (const FT *)D.d,
                                     (const FT *)D.y, (FT *)D.z, nullptr, ri, my, lane);
}

// one wavefront per color item; every branch below is uniform per wavefront
template <int WM, bool FWD>
__global__ __launch_bounds__(kWG) void kb_sweep(const SysDesc *__restrict__ desc,
                                                const ColorItem *__restrict__ items, int nitems,
                                                const SysScal *__restrict__ sc) {
  __shared__ double lds[kWG];
  const int wv = blockIdx.x * (kWG >> 6) + (threadIdx.x >> 6);
  if (wv >= nitems) return;
  const ColorItem c = items[wv];
  if (sc[c.sys].done) return;
  const SysDesc &D = desc[c.sys];
  const int lane = threadIdx.x & 63;
  double *my = lds + ((threadIdx.x >> 6) << 6);
  if (D.bstart) {
    if (D.mixed) {
      if (FWD) abmc_fwd_item<float, WM>(D, c, lane, (float *)my);
      else abmc_bwd_item<float, WM>(D, c, lane, (float *)my);
    } else {
      if (FWD) abmc_fwd_item<double, WM>(D, c, lane, my);
      else abmc_bwd_item<double, WM>(D, c, lane, my);
    }
  } else {
    if (D.mixed) {
      if (FWD) mc_fwd_item<float, WM>(D, c, lane);
      else mc_bwd_item<float, WM>(D, c, lane);
    } else {
      if (FWD) mc_fwd_item<double, WM>(D, c, lane);
      else mc_bwd_item<double, WM>(D, c, lane);
    }
  }
}

// finish kernels: workgroup k folds the partials of the k-th listed system
// (chunks [cptr[k], cptr[k+1])) and applies the single-system scalar logic
// (cg_crs.cpp:104-160)
template <int NC>
__device__ __forceinline__ bool fold_system(const int *cptr, const double *part,
                                            double (&out)[NC]) {
  const int k = blockIdx.x;
  const int c0 = cptr[k], c1 = cptr[k + 1];
  return finish_reduce<NC>(part + (size_t)c0 * NC * 2, c1 - c0, out);
}

__global__ __launch_bounds__(kWG) void kbf_resid(const int *__restrict__ act,
                                                 const int *__restrict__ cptr,
                                                 const double *__restrict__ part,
                                                 const SysDesc *__restrict__ desc, SysScal *sc,
                                                 int mode) {
  double out[2];
  if (!fold_system<2>(cptr, part, out)) return;
  const int id = act[blockIdx.x];
  SysScal &S = sc[id];
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
    S.tol = desc[id].tol;
    S.max_iter = desc[id].max_iter;
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

__global__ __launch_bounds__(kWG) void kbf_alpha(const int *__restrict__ act,
                                                 const int *__restrict__ cptr,
                                                 const double *__restrict__ part, SysScal *sc) {
  SysScal &S = sc[act[blockIdx.x]];
  if (S.done) return;
  double out[3];
  if (!fold_system<3>(cptr, part, out)) return;
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

__global__ __launch_bounds__(kWG) void kbf_xr(const int *__restrict__ act,
                                              const int *__restrict__ cptr,
                                              const double *__restrict__ part, SysScal *sc) {
  SysScal &S = sc[act[blockIdx.x]];
  if (S.done) return;
  double out[1];
  if (!fold_system<1>(cptr, part, out)) return;
  const double rel = sqrt(out[0]) / S.normb;
  const int it = S.iters + 1;
  S.iters = it;
  S.rel = rel;
  if (rel < S.tol) {
    S.converged = 1;
    S.done = 1;
  } else if (it >= S.max_iter) {
    S.done = 1;
  }
}

__global__ __launch_bounds__(kWG) void kbf_rz(const int *__restrict__ act,
                                              const int *__restrict__ cptr,
                                              const double *__restrict__ part, SysScal *sc,
                                              int init) {
  SysScal &S = sc[act[blockIdx.x]];
  if (S.done) return;
  double out[1];
  if (!fold_system<1>(cptr, part, out)) return;
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

// Item lists of a set of systems, resident on the device in one buffer.
struct ItemSet {
  int nsys = 0, nchunks = 0, ncolors = 0;
  std::vector<int> color_off;  // ncolors + 1 offsets into d_color
  std::vector<int> chunks_of;  // per listed system
  char *buf = nullptr;
  int *d_act = nullptr, *d_cptr = nullptr;
  ChunkItem *d_chunks = nullptr;
  ColorItem *d_color = nullptr;
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
  int listed_chunks = 0;  // chunk items of this slot in the running set
  std::chrono::steady_clock::time_point t_add;
  iccg_cg_result res{};
};

int wm_class(int w) { return w <= 4 ? 4 : w <= 8 ? 8 : 16; }

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
  int wm = 4;                 // template width of the running set
  int chunk = 16;             // slices per chunk item
  int cw = 1;                 // slices (MC) or blocks (ABMC) per sweep wavefront
  bool autotune = true;
  long tuned_rows = 0;        // running rows at the last tuning
  ItemSet running;            // items of the running set
  bool running_dirty = true;
  int dead_chunks = 0;        // listed chunks of systems no longer running
  bool graphs = true;
  std::vector<std::pair<int, hipGraphExec_t>> execs;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  ~iccg_batch_s() {
    (void)hipSetDevice(device);
    if (st) (void)hipStreamSynchronize(st);
    drop_graphs();
    running.clear();
    for (auto &s : slots) {
      dfree(s.dbuf);
      if (s.hstage) (void)hipHostFree(s.hstage);
    }
    dfree(d_desc);
    dfree(d_sc);
    dfree(part);
    if (h_sc) (void)hipHostFree(h_sc);
    if (h_desc) (void)hipHostFree(h_desc);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
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

// item lists for the given slots with chunk / sweep granularities ch, cw
void build_items(iccg_batch_s &B, const std::vector<int> &ids, ItemSet &I, int ch, int cw) {
  I.clear();
  std::vector<int> act, cptr{0};
  std::vector<ChunkItem> chunks;
  std::vector<std::vector<ColorItem>> color;
  for (int id : ids) {
    const Slot &s = B.slots[id];
    const int ns = (s.n + 63) / 64;
    act.push_back(id);
    const int c_before = (int)chunks.size();
    for (int a = 0; a < ns; a += ch) chunks.push_back({id, a, std::min(ns, a + ch), 0});
    cptr.push_back((int)chunks.size());
    I.chunks_of.push_back((int)chunks.size() - c_before);
    const DevSched &sc = s.M->sc;
    const int nc = sc.nc;
    if ((int)color.size() < nc) color.resize(nc);
    for (int c = 0; c < nc; ++c) {
      if (sc.kind == 2) {
        const int b0 = sc.h_color_blk_v[c], b1 = sc.h_color_blk_v[c + 1];
        for (int kb = b0; kb < b1; kb += cw)
          color[c].push_back({id, kb, std::min(b1, kb + cw), 0, 0, {0, 0, 0}});
      } else {
        const int lo = sc.h_color_ptr_v[c], hi = sc.h_color_ptr_v[c + 1];
        if (hi <= lo) continue;
        const int s1 = ((hi - 1) >> 6) + 1;
        for (int sl = lo >> 6; sl < s1; sl += cw)
          color[c].push_back({id, sl, std::min(s1, sl + cw), lo, hi, {0, 0, 0}});
      }
    }
  }
  I.nsys = (int)act.size();
  I.nchunks = (int)chunks.size();
  I.ncolors = (int)color.size();
  I.color_off.assign(1, 0);
  for (auto &v : color) I.color_off.push_back(I.color_off.back() + (int)v.size());
  if (I.nsys == 0) return;
  // one device buffer: act | cptr | chunks | color items
  const size_t o_cptr = align_up(act.size() * sizeof(int));
  const size_t o_chunks = o_cptr + align_up(cptr.size() * sizeof(int));
  const size_t o_color = o_chunks + align_up(chunks.size() * sizeof(ChunkItem));
  const size_t total = o_color + (size_t)I.color_off.back() * sizeof(ColorItem);
  std::vector<char> h(total);
  std::memcpy(h.data(), act.data(), act.size() * sizeof(int));
  std::memcpy(h.data() + o_cptr, cptr.data(), cptr.size() * sizeof(int));
  std::memcpy(h.data() + o_chunks, chunks.data(), chunks.size() * sizeof(ChunkItem));
  size_t o = o_color;
  for (auto &v : color) {
    if (!v.empty()) std::memcpy(h.data() + o, v.data(), v.size() * sizeof(ColorItem));
    o += v.size() * sizeof(ColorItem);
  }
  ensure_part(B, (size_t)I.nchunks * 3);
  I.buf = dalloc<char>(total);
  I.d_act = (int *)I.buf;
  I.d_cptr = (int *)(I.buf + o_cptr);
  I.d_chunks = (ChunkItem *)(I.buf + o_chunks);
  I.d_color = (ColorItem *)(I.buf + o_color);
  HIPCHK(hipMemcpyAsync(I.buf, h.data(), total, hipMemcpyHostToDevice, B.st));
  HIPCHK(hipStreamSynchronize(B.st));
}

#define BATCH_WM(wm, K, ...)                                          \
  do {                                                                \
    if ((wm) <= 4) hipLaunchKernelGGL((K<4>), __VA_ARGS__);           \
    else if ((wm) <= 8) hipLaunchKernelGGL((K<8>), __VA_ARGS__);      \
    else hipLaunchKernelGGL((K<16>), __VA_ARGS__);                    \
  } while (0)

#define BATCH_WM2(wm, K, F, ...)                                      \
  do {                                                                \
    if ((wm) <= 4) hipLaunchKernelGGL((K<4, F>), __VA_ARGS__);        \
    else if ((wm) <= 8) hipLaunchKernelGGL((K<8, F>), __VA_ARGS__);   \
    else hipLaunchKernelGGL((K<16, F>), __VA_ARGS__);                 \
  } while (0)

void launch_resid_items(iccg_batch_s &B, const ItemSet &I, int mode) {
  if (I.nsys == 0) return;
  dim3 g(I.nchunks), b(kWG);
  BATCH_WM(B.wm, kb_resid, g, b, 0, B.st, B.d_desc, I.d_chunks, B.part, mode);
  hipLaunchKernelGGL(kbf_resid, dim3(I.nsys), b, 0, B.st, I.d_act, I.d_cptr, B.part, B.d_desc,
                     B.d_sc, mode);
}

void launch_sweeps(iccg_batch_s &B, const ItemSet &I) {
  dim3 b(kWG);
  for (int c = 0; c < I.ncolors; ++c) {
    const int ni = I.color_off[c + 1] - I.color_off[c];
    if (ni)
      BATCH_WM2(B.wm, kb_sweep, true, dim3((ni + 3) / 4), b, 0, B.st, B.d_desc,
                I.d_color + I.color_off[c], ni, B.d_sc);
  }
  for (int c = I.ncolors - 1; c >= 0; --c) {
    const int ni = I.color_off[c + 1] - I.color_off[c];
    if (ni)
      BATCH_WM2(B.wm, kb_sweep, false, dim3((ni + 3) / 4), b, 0, B.st, B.d_desc,
                I.d_color + I.color_off[c], ni, B.d_sc);
  }
}

void launch_rz_p(iccg_batch_s &B, const ItemSet &I, int init) {
  dim3 g(I.nchunks), b(kWG);
  hipLaunchKernelGGL(kb_rz, g, b, 0, B.st, B.d_desc, I.d_chunks, B.d_sc, B.part);
  hipLaunchKernelGGL(kbf_rz, dim3(I.nsys), b, 0, B.st, I.d_act, I.d_cptr, B.part, B.d_sc, init);
  hipLaunchKernelGGL(kb_update_p, g, b, 0, B.st, B.d_desc, I.d_chunks, B.d_sc, init);
}

// z0 = M^-1 r0, rz_old, p = z0 (cg_crs.cpp:112-115) for the listed systems
void launch_precond_init(iccg_batch_s &B, const ItemSet &I) {
  if (I.nsys == 0) return;
  launch_sweeps(B, I);
  launch_rz_p(B, I, 1);
}

// one CG iteration (cg_crs.cpp:125-160) of every listed system
void launch_iteration(iccg_batch_s &B, const ItemSet &I) {
  dim3 g(I.nchunks), b(kWG);
  BATCH_WM(B.wm, kb_spmv, g, b, 0, B.st, B.d_desc, I.d_chunks, B.d_sc, B.part);
  hipLaunchKernelGGL(kbf_alpha, dim3(I.nsys), b, 0, B.st, I.d_act, I.d_cptr, B.part, B.d_sc);
  hipLaunchKernelGGL(kb_xr, g, b, 0, B.st, B.d_desc, I.d_chunks, B.d_sc, B.part);
  hipLaunchKernelGGL(kbf_xr, dim3(I.nsys), b, 0, B.st, I.d_act, I.d_cptr, B.part, B.d_sc);
  launch_sweeps(B, I);
  launch_rz_p(B, I, 0);
}

void enqueue_iterations(iccg_batch_s &B, int k) {
  const ItemSet &I = B.running;
  if (!B.graphs) {
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
  HIPCHK(hipStreamSynchronize(B.st));
}

void retire_listed(iccg_batch_s &B, Slot &s) {
  B.dead_chunks += s.listed_chunks;
  s.listed_chunks = 0;
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
  int wm = 4;
  for (const Slot &s : B.slots)
    if (s.state == 1 || s.state == 2)
      wm = std::max({wm, s.A->d.sell.maxw, s.M->F.ls.maxw, s.M->F.us.maxw});
  B.wm = wm_class(wm);
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

double time_ms(iccg_batch_s &B, const std::function<void()> &f, int reps) {
  f();  // warm
  HIPCHK(hipEventRecord(B.ev0, B.st));
  for (int r = 0; r < reps; ++r) f();
  HIPCHK(hipEventRecord(B.ev1, B.st));
  HIPCHK(hipEventSynchronize(B.ev1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, B.ev0, B.ev1));
  return ms / reps;
}

// Autotuner: times the candidate work decompositions on the running systems
// and keeps the fastest.  The probes are side-effect free at an iteration
// boundary: the SpMV rewrites Ap and the sweeps rewrite y and z with the
// values they already hold; partials are scratch.
void tune(iccg_batch_s &B, const std::vector<int> &ids) {
  if (!B.ev0) {
    HIPCHK(hipEventCreate(&B.ev0));
    HIPCHK(hipEventCreate(&B.ev1));
  }
  const int chunks[] = {4, 8, 16, 32};
  const int cws[] = {1, 2, 4};
  double best = 1e300;
  int best_ch = B.chunk;
  ItemSet T;
  for (int ch : chunks) {
    build_items(B, ids, T, ch, B.cw);
    const double ms = time_ms(B, [&] {
      dim3 g(T.nchunks), b(kWG);
      BATCH_WM(B.wm, kb_spmv, g, b, 0, B.st, B.d_desc, T.d_chunks, B.d_sc, B.part);
      hipLaunchKernelGGL(kb_rz, g, b, 0, B.st, B.d_desc, T.d_chunks, B.d_sc, B.part);
    }, 3);
    if (ms < best) {
      best = ms;
      best_ch = ch;
    }
  }
  best = 1e300;
  int best_cw = B.cw;
  for (int cw : cws) {
    build_items(B, ids, T, best_ch, cw);
    const double ms = time_ms(B, [&] { launch_sweeps(B, T); }, 3);
    if (ms < best) {
      best = ms;
      best_cw = cw;
    }
  }
  HIPCHK(hipGetLastError());
  T.clear();
  B.chunk = best_ch;
  B.cw = best_cw;
}

void refresh_running(iccg_batch_s &B) {
  if (!B.running_dirty && B.dead_chunks * 4 <= B.running.nchunks) return;
  std::vector<int> ids;
  long rows = 0;
  for (int id = 0; id < B.cap; ++id)
    if (B.slots[id].state == 2 && !B.h_sc[id].done) {
      ids.push_back(id);
      rows += B.slots[id].n;
    }
  B.drop_graphs();
  if (B.autotune && rows >= (1L << 16) &&
      (B.tuned_rows == 0 || rows > 2 * B.tuned_rows || 2 * rows < B.tuned_rows)) {
    tune(B, ids);
    B.tuned_rows = rows;
  }
  for (Slot &s : B.slots) s.listed_chunks = 0;
  build_items(B, ids, B.running, B.chunk, B.cw);
  for (size_t k = 0; k < ids.size(); ++k) B.slots[ids[k]].listed_chunks = B.running.chunks_of[k];
  B.running_dirty = false;
  B.dead_chunks = 0;
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
    HIPCHK(hipHostMalloc((void **)&B->h_sc, sizeof(SysScal) * (size_t)capacity,
                         hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void **)&B->h_desc, sizeof(SysDesc) * (size_t)capacity,
                         hipHostMallocDefault));
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
      if (s.hstage) (void)hipHostFree(s.hstage);
      s.hstage = nullptr;
      s.hcap = 0;
      HIPCHK(hipHostMalloc((void **)&s.hstage, hneed, hipHostMallocDefault));
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
    s.listed_chunks = 0;
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
    for (;;) {
      int running = 0;
      for (auto &s : B->slots) running += s.state == 2;
      if (running == 0 || steps >= max_steps) break;
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
    S.done = 1;
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
        if (s.hstage) (void)hipHostFree(s.hstage);
        s.hstage = nullptr;
        s.hcap = 0;
      }
  });
}

int iccg_batch_set_graphs(iccg_batch_t B, int on) {
  return guard([&] {
    require(B, ICCG_ERR_ARG, "null argument");
    B->graphs = on != 0;
  });
}

int iccg_batch_set_autotune(iccg_batch_t B, int on) {
  return guard([&] {
    require(B, ICCG_ERR_ARG, "null argument");
    B->autotune = on != 0;
    B->tuned_rows = 0;
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
