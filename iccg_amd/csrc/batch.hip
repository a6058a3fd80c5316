// batch.hip -- multi-system ICCG: many independent SPD systems, each with
// its own matrix, IC(0) factor and Krylov vectors resident in HBM, advanced
// together so that every CG phase is ONE launch over all active systems.
//
// Work is described by item lists that name a system slot:
//   ChunkItem (slot, s0, s1): up to kChunk 64-row slices of one system, one
//     workgroup -- SpMV, vector updates and dot-product partials;
//   ColorItem (slot, slice, lo, hi): rows [lo, hi) of one color inside one
//     slice, one wavefront -- the multicolor triangular sweeps.
// Every system keeps its own scalars (SysScal); a finish kernel with one
// workgroup per system folds that system's partials.  Converged systems
// skip their items, and systems can be admitted or retired between launches:
// the serving loop's continuous batching.  Per system, the arithmetic is
// the single-system solver's: same row chains, same double-double dots, so
// a system gives the same iterates in a batch as alone.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "dev.hpp"
#include "devfun.hpp"
#include "internal.hpp"
#include "objects.hpp"

namespace iccg {

constexpr int kChunk = 16;  // slices per chunk item (1024 rows)

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
template <class AV>
__device__ __forceinline__ const AV *a_vals(const SysDesc &D);
template <>
__device__ __forceinline__ const double *a_vals<double>(const SysDesc &D) { return D.a_v64; }
template <>
__device__ __forceinline__ const float *a_vals<float>(const SysDesc &D) { return D.a_v32; }

// r = b - A x with b.b and r.r partials (mode 1 / 2 write r; mode 0 only
// measures the true residual)
template <class AV, int WM>
__global__ __launch_bounds__(kWG) void kb_resid(const SysDesc *__restrict__ desc,
                                                const ChunkItem *__restrict__ items,
                                                double *part, int mode) {
  const ChunkItem c = items[blockIdx.x];
  const SysDesc &D = desc[c.sys];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double h[2] = {0, 0}, l[2] = {0, 0};
  for (int s = c.s0 + w; s < c.s1; s += kWG >> 6) {
    const double ax = sell_chain<AV, double, double, WM, false>(D.a_off, D.a_col, a_vals<AV>(D),
                                                                D.x, s, lane, 0.0);
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

template <class AV, int WM>
__global__ __launch_bounds__(kWG) void kb_spmv(const SysDesc *__restrict__ desc,
                                               const ChunkItem *__restrict__ items,
                                               const SysScal *__restrict__ sc, double *part) {
  const ChunkItem c = items[blockIdx.x];
  double h[3] = {0, 0, 0}, l[3] = {0, 0, 0};
  if (!sc[c.sys].done) {
    const SysDesc &D = desc[c.sys];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int s = c.s0 + w; s < c.s1; s += kWG >> 6) {
      const double y = sell_chain<AV, double, double, WM, false>(D.a_off, D.a_col, a_vals<AV>(D),
                                                                 D.p, s, lane, 0.0);
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
__global__ __launch_bounds__(kWG) void kb_rz(const SysDesc *__restrict__ desc,
                                             const ChunkItem *__restrict__ items,
                                             const SysScal *__restrict__ sc, double *part) {
  const ChunkItem c = items[blockIdx.x];
  double h[1] = {0}, l[1] = {0};
  if (!sc[c.sys].done) {
    const SysDesc &D = desc[c.sys];
    const FT *z = (const FT *)D.z;
    const int end = min(D.n, c.s1 << 6);
    for (int i = (c.s0 << 6) + threadIdx.x; i < end; i += kWG)
      dd_fma(h[0], l[0], D.r[i], (double)z[i]);
  }
  block_partial<1>(h, l, part, blockIdx.x);
}

template <class FT>
__global__ __launch_bounds__(kWG) void kb_update_p(const SysDesc *__restrict__ desc,
                                                   const ChunkItem *__restrict__ items,
                                                   const SysScal *__restrict__ sc, int init) {
  const ChunkItem c = items[blockIdx.x];
  if (sc[c.sys].done) return;
  const SysDesc &D = desc[c.sys];
  const FT *z = (const FT *)D.z;
  const double beta = sc[c.sys].beta;
  const int end = min(D.n, c.s1 << 6);
  for (int i = (c.s0 << 6) + threadIdx.x; i < end; i += kWG) {
    const double zi = (double)z[i];
    D.p[i] = init ? zi : fma(beta, D.p[i], zi);
  }
}

// multicolor sweeps, one wavefront per color item (ic0.cpp:194-249)
template <class FT, int WM>
__global__ __launch_bounds__(kWG) void kb_mc_fwd(const SysDesc *__restrict__ desc,
                                                 const ColorItem *__restrict__ items, int nitems,
                                                 const SysScal *__restrict__ sc) {
  const int wv = blockIdx.x * (kWG >> 6) + (threadIdx.x >> 6);
  if (wv >= nitems) return;
  const ColorItem c = items[wv];
  if (sc[c.sys].done) return;
  const SysDesc &D = desc[c.sys];
  const int lane = threadIdx.x & 63;
  const int i = (c.slice << 6) + lane;
  if (i < c.lo || i >= c.hi) return;
  FT *y = (FT *)D.y;
  FT acc = (FT)D.r[i];
  acc = sell_chain<FT, FT, FT, WM, true>(D.l_off, D.l_col, (const FT *)D.l_v, y, c.slice, lane,
                                         acc);
  y[i] = acc / ((const FT *)D.d)[i];
}

template <class FT, int WM>
__global__ __launch_bounds__(kWG) void kb_mc_bwd(const SysDesc *__restrict__ desc,
                                                 const ColorItem *__restrict__ items, int nitems,
                                                 const SysScal *__restrict__ sc) {
  const int wv = blockIdx.x * (kWG >> 6) + (threadIdx.x >> 6);
  if (wv >= nitems) return;
  const ColorItem c = items[wv];
  if (sc[c.sys].done) return;
  const SysDesc &D = desc[c.sys];
  const int lane = threadIdx.x & 63;
  const int i = (c.slice << 6) + lane;
  if (i < c.lo || i >= c.hi) return;
  FT *z = (FT *)D.z;
  FT acc = ((const FT *)D.y)[i];
  acc = sell_chain<FT, FT, FT, WM, true>(D.u_off, D.u_col, (const FT *)D.u_v, z, c.slice, lane,
                                         acc);
  z[i] = acc / ((const FT *)D.d)[i];
}

// finish kernels: workgroup k folds the partials of the k-th listed system
// (chunks [cptr[k], cptr[k+1])) and applies the single-system scalar logic
template <int NC>
__device__ __forceinline__ bool fold_system(const int *cptr, const double *part,
                                            double (&out)[NC]) {
  const int k = blockIdx.x;
  const int c0 = cptr[k], c1 = cptr[k + 1];
  return finish_reduce<NC>(part + (size_t)c0 * NC * 2, c1 - c0, out);
}

__global__ __launch_bounds__(kWG) void kbf_resid(const int *__restrict__ act,
                                                 const int *__restrict__ cptr,
                                                 const double *__restrict__ part, SysScal *sc,
                                                 int mode) {
  double out[2];
  if (!fold_system<2>(cptr, part, out)) return;
  SysScal &S = sc[act[blockIdx.x]];
  double normb = sqrt(out[0]);
  if (normb == 0.0) normb = 1.0;
  S.normb = normb;
  if (mode == 0) {
    S.rr_true = out[1];
    return;
  }
  const double rel = sqrt(out[1]) / normb;
  if (mode == 1) S.iters = 0;
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

// Item lists of a set of systems, resident on the device.
struct ItemSet {
  int nsys = 0, nchunks = 0, maxc = 0;
  std::vector<int> ncolor_items;  // per color
  int *d_act = nullptr, *d_cptr = nullptr;
  ChunkItem *d_chunks = nullptr;
  std::vector<ColorItem *> d_color;
  void clear() {
    dfree(d_act);
    dfree(d_cptr);
    dfree(d_chunks);
    for (auto p : d_color) dfree(p);
    *this = ItemSet();
  }
};

struct Slot {
  int state = 0;  // 0 free, 1 admitted (needs init), 2 running, 3 finished
  iccg_csr_t A = nullptr;
  iccg_precond_t M = nullptr;
  int n = 0;
  double *x = nullptr, *r = nullptr, *p = nullptr, *Ap = nullptr, *b = nullptr;
  void *y = nullptr, *z = nullptr;
  int max_iter = 0, flags = 0, restarts = 0;
  double tol = 0;
  bool checked = false;
  std::chrono::steady_clock::time_point t_add;
  iccg_cg_result res{};
};

}  // namespace

struct iccg_batch_s {
  int device = 0;
  hipStream_t st = nullptr;
  int cap = 0;
  std::vector<Slot> slots;
  SysDesc *d_desc = nullptr;
  SysScal *d_sc = nullptr;
  SysScal *h_sc = nullptr;  // pinned
  double *part = nullptr;
  size_t part_slots = 0;
  bool a32 = true, mixed = false, have_kind = false;
  int wm = 4;  // widest sliced row of any admitted system (template width)
  ItemSet running;          // items of the running set
  bool running_dirty = true;
  bool graphs = true;
  std::vector<std::pair<int, hipGraphExec_t>> execs;
  ~iccg_batch_s() {
    (void)hipSetDevice(device);
    for (auto &e : execs) (void)hipGraphExecDestroy(e.second);
    running.clear();
    for (auto &s : slots) free_slot(s);
    dfree(d_desc);
    dfree(d_sc);
    dfree(part);
    if (h_sc) (void)hipHostFree(h_sc);
    if (st) (void)hipStreamDestroy(st);
  }
  static void free_slot(Slot &s) {
    dfree(s.x);
    dfree(s.r);
    dfree(s.p);
    dfree(s.Ap);
    dfree(s.b);
    dfree(s.y);
    dfree(s.z);
    s = Slot();
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
  B.part = dalloc<double>(slots * 2);
  B.part_slots = slots;
  B.drop_graphs();
}

// item lists for the given slots
void build_items(iccg_batch_s &B, const std::vector<int> &ids, ItemSet &I) {
  I.clear();
  std::vector<int> act, cptr{0};
  std::vector<ChunkItem> chunks;
  std::vector<std::vector<ColorItem>> color;
  for (int id : ids) {
    const Slot &s = B.slots[id];
    const int ns = (s.n + 63) / 64;
    act.push_back(id);
    for (int a = 0; a < ns; a += kChunk) chunks.push_back({id, a, std::min(ns, a + kChunk), 0});
    cptr.push_back((int)chunks.size());
    const std::vector<int> &cp = s.M->sc.h_color_ptr_v;
    const int nc = (int)cp.size() - 1;
    if ((int)color.size() < nc) color.resize(nc);
    for (int c = 0; c < nc; ++c) {
      const int lo = cp[c], hi = cp[c + 1];
      if (hi <= lo) continue;
      for (int sl = lo >> 6; sl <= (hi - 1) >> 6; ++sl) color[c].push_back({id, sl, lo, hi});
    }
  }
  I.nsys = (int)act.size();
  I.nchunks = (int)chunks.size();
  I.maxc = (int)color.size();
  if (I.nsys == 0) return;
  I.d_act = dupload(act.data(), act.size(), B.st);
  I.d_cptr = dupload(cptr.data(), cptr.size(), B.st);
  I.d_chunks = dupload(chunks.data(), chunks.size(), B.st);
  for (auto &v : color) {
    I.ncolor_items.push_back((int)v.size());
    I.d_color.push_back(v.empty() ? nullptr : dupload(v.data(), v.size(), B.st));
  }
  ensure_part(B, (size_t)I.nchunks * 3);
  HIPCHK(hipStreamSynchronize(B.st));
}

#define BATCH_WM(wm, K, AV, ...)                                           \
  do {                                                                     \
    if ((wm) <= 4) hipLaunchKernelGGL((K<AV, 4>), __VA_ARGS__);            \
    else if ((wm) <= 8) hipLaunchKernelGGL((K<AV, 8>), __VA_ARGS__);       \
    else hipLaunchKernelGGL((K<AV, 16>), __VA_ARGS__);                     \
  } while (0)

void launch_resid_items(iccg_batch_s &B, const ItemSet &I, int mode) {
  if (I.nsys == 0) return;
  dim3 g(I.nchunks), b(kWG);
  if (B.a32)
    BATCH_WM(B.wm, kb_resid, float, g, b, 0, B.st, B.d_desc, I.d_chunks, B.part, mode);
  else
    BATCH_WM(B.wm, kb_resid, double, g, b, 0, B.st, B.d_desc, I.d_chunks, B.part, mode);
  hipLaunchKernelGGL(kbf_resid, dim3(I.nsys), b, 0, B.st, I.d_act, I.d_cptr, B.part, B.d_sc,
                     mode);
}

void launch_sweeps(iccg_batch_s &B, const ItemSet &I) {
  dim3 b(kWG);
  for (int c = 0; c < I.maxc; ++c) {
    const int ni = I.ncolor_items[c];
    if (!ni) continue;
    dim3 g((ni + 3) / 4);
    if (B.mixed)
      BATCH_WM(B.wm, kb_mc_fwd, float, g, b, 0, B.st, B.d_desc, I.d_color[c], ni, B.d_sc);
    else
      BATCH_WM(B.wm, kb_mc_fwd, double, g, b, 0, B.st, B.d_desc, I.d_color[c], ni, B.d_sc);
  }
  for (int c = I.maxc - 1; c >= 0; --c) {
    const int ni = I.ncolor_items[c];
    if (!ni) continue;
    dim3 g((ni + 3) / 4);
    if (B.mixed)
      BATCH_WM(B.wm, kb_mc_bwd, float, g, b, 0, B.st, B.d_desc, I.d_color[c], ni, B.d_sc);
    else
      BATCH_WM(B.wm, kb_mc_bwd, double, g, b, 0, B.st, B.d_desc, I.d_color[c], ni, B.d_sc);
  }
}

void launch_rz_p(iccg_batch_s &B, const ItemSet &I, int init) {
  dim3 g(I.nchunks), b(kWG);
  if (B.mixed) {
    hipLaunchKernelGGL(kb_rz<float>, g, b, 0, B.st, B.d_desc, I.d_chunks, B.d_sc, B.part);
    hipLaunchKernelGGL(kbf_rz, dim3(I.nsys), b, 0, B.st, I.d_act, I.d_cptr, B.part, B.d_sc, init);
    hipLaunchKernelGGL(kb_update_p<float>, g, b, 0, B.st, B.d_desc, I.d_chunks, B.d_sc, init);
  } else {
    hipLaunchKernelGGL(kb_rz<double>, g, b, 0, B.st, B.d_desc, I.d_chunks, B.d_sc, B.part);
    hipLaunchKernelGGL(kbf_rz, dim3(I.nsys), b, 0, B.st, I.d_act, I.d_cptr, B.part, B.d_sc, init);
    hipLaunchKernelGGL(kb_update_p<double>, g, b, 0, B.st, B.d_desc, I.d_chunks, B.d_sc, init);
  }
}

// z0 = M^-1 r0, rz_old, p = z0 (cg_crs.cpp:112-115) for the listed systems
void launch_precond_init(iccg_batch_s &B, const ItemSet &I) {
  if (I.nsys == 0) return;
  launch_sweeps(B, I);
  launch_rz_p(B, I, 1);
}

void launch_iteration(iccg_batch_s &B, const ItemSet &I) {
  dim3 g(I.nchunks), b(kWG);
  if (B.a32)
    BATCH_WM(B.wm, kb_spmv, float, g, b, 0, B.st, B.d_desc, I.d_chunks, B.d_sc, B.part);
  else
    BATCH_WM(B.wm, kb_spmv, double, g, b, 0, B.st, B.d_desc, I.d_chunks, B.d_sc, B.part);
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
  B.running_dirty = true;
}

// Systems whose recurrence converged get the FP64 true residual; those
// above tol restart from it (ICCG_SOLVE_TRUE_RESID), the others finish.
void settle(iccg_batch_s &B) {
  std::vector<int> check, done_now;
  for (int id = 0; id < B.cap; ++id) {
    Slot &s = B.slots[id];
    if (s.state != 2 || !B.h_sc[id].done) continue;
    done_now.push_back(id);
  }
  if (done_now.empty()) return;
  // true residual of every system that stopped (reported, and tested when asked)
  ItemSet I;
  build_items(B, done_now, I);
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
    build_items(B, restart, R);
    launch_resid_items(B, R, 2);
    launch_precond_init(B, R);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(B.st));
    R.clear();
  }
}

void admit_pending(iccg_batch_s &B) {
  std::vector<int> ids;
  for (int id = 0; id < B.cap; ++id)
    if (B.slots[id].state == 1) ids.push_back(id);
  if (ids.empty()) return;
  ItemSet I;
  build_items(B, ids, I);
  launch_resid_items(B, I, 1);
  launch_precond_init(B, I);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(B.st));
  I.clear();
  for (int id : ids) B.slots[id].state = 2;
  B.running_dirty = true;
}

void refresh_running(iccg_batch_s &B) {
  if (!B.running_dirty) return;
  std::vector<int> ids;
  for (int id = 0; id < B.cap; ++id)
    if (B.slots[id].state == 2) ids.push_back(id);
  B.drop_graphs();
  build_items(B, ids, B.running);
  B.running_dirty = false;
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
    HIPCHK(hipMemsetAsync(B->d_sc, 0, sizeof(SysScal) * (size_t)capacity, B->st));
    HIPCHK(hipHostMalloc((void **)&B->h_sc, sizeof(SysScal) * (size_t)capacity,
                         hipHostMallocDefault));
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
    require(M->kind == ICCG_PRECOND_IC0 && M->sc.kind == 1 && M->sc.d_rows == nullptr &&
                M->F.sell && A->d.sell.off,
            ICCG_ERR_ARG,
            "batched solves need an IC0 preconditioner with a contiguous multicolor schedule");
    const bool a32 = A->d.sell_v32 != nullptr;
    bool any = false;
    for (const Slot &o : B->slots) any |= o.state != 0;
    require(!any || (B->mixed == M->mixed && B->a32 == a32), ICCG_ERR_ARG,
            "all systems of a batch share one precision and one matrix value type");
    const int wm = std::max({any ? B->wm : 0, A->d.sell.maxw, M->F.ls.maxw, M->F.us.maxw});
    require(wm <= 16, ICCG_ERR_ARG, "rows too long for the batched sliced kernels");
    B->mixed = M->mixed;
    B->a32 = a32;
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
    s.A = A;
    s.M = M;
    s.n = n;
    s.max_iter = max_iter;
    s.tol = tol;
    s.flags = flags;
    s.restarts = 0;
    s.res = iccg_cg_result{};
    s.res.true_rel_resid = -1.0;
    const size_t vb = M->mixed ? sizeof(float) : sizeof(double);
    s.x = dalloc<double>((size_t)n);
    s.r = dalloc<double>((size_t)n);
    s.p = dalloc<double>((size_t)n);
    s.Ap = dalloc<double>((size_t)n);
    s.b = dupload(b, (size_t)n, B->st);
    s.y = dalloc<char>(vb * (size_t)n);
    s.z = dalloc<char>(vb * (size_t)n);
    if (x0)
      HIPCHK(hipMemcpyAsync(s.x, x0, sizeof(double) * n, hipMemcpyHostToDevice, B->st));
    else
      HIPCHK(hipMemsetAsync(s.x, 0, sizeof(double) * n, B->st));
    SysDesc D;
    D.n = n;
    D.nslice = (n + 63) / 64;
    D.mixed = M->mixed;
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
    D.x = s.x;
    D.r = s.r;
    D.p = s.p;
    D.Ap = s.Ap;
    D.b = s.b;
    D.y = s.y;
    D.z = s.z;
    B->wm = wm;
    SysScal S{};
    S.tol = tol;
    S.max_iter = max_iter;
    HIPCHK(hipMemcpyAsync(B->d_desc + id, &D, sizeof(D), hipMemcpyHostToDevice, B->st));
    HIPCHK(hipMemcpyAsync(B->d_sc + id, &S, sizeof(S), hipMemcpyHostToDevice, B->st));
    HIPCHK(hipStreamSynchronize(B->st));
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
      refresh_running(*B);
      int running = 0;
      for (auto &s : B->slots) running += s.state == 2;
      if (running == 0 || steps >= max_steps) break;
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
    require(s.state != 0, ICCG_ERR_STATE, "empty batch slot");
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
    if (s.state == 2) B->running_dirty = true;
    iccg_batch_s::free_slot(s);
    SysScal S{};
    S.done = 1;
    HIPCHK(hipMemcpyAsync(B->d_sc + slot, &S, sizeof(S), hipMemcpyHostToDevice, B->st));
    HIPCHK(hipStreamSynchronize(B->st));
  });
}

int iccg_batch_set_graphs(iccg_batch_t B, int on) {
  return guard([&] {
    require(B, ICCG_ERR_ARG, "null argument");
    B->graphs = on != 0;
  });
}

}  // extern "C"
