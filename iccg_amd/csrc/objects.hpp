// objects.hpp -- the C ABI's object types and device-memory helpers,
// shared by api.hip (single-system solver) and batch.hip (multi-system).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <iterator>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "dev.hpp"
#include "internal.hpp"

namespace iccg {

// HIP device selected by iccg_device_set on this thread
int current_device();

#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess)                                                              \
      throw Error(ICCG_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

// Device memory comes from a stream-ordered pool of the device's own,
// allocated and freed on a stream that carries nothing else.  hipFree waits
// until every stream of the device is idle -- the serving loop keeps chunks
// of up to 64 iterations of multi-million-row systems queued on its runner
// streams -- and holds other threads' allocations while it waits; a pool
// free waits for nothing.  The pool keeps the memory freed into it (release
// threshold at its maximum), so the next allocation reuses it rather than
// asking the driver.  A free is therefore no longer a device-wide wait:
// every object waits for its own stream before it frees, and memory is
// freed only once the work that uses it has finished.
struct DevicePool {
  hipMemPool_t pool = nullptr;
  hipStream_t st = nullptr;
};

struct DevicePools {
  std::mutex mu;
  std::vector<DevicePool> pools;
};

inline DevicePool device_pool() {
  // never destroyed: objects may be freed during static destruction at exit
  static DevicePools &all = *new DevicePools;
  std::vector<DevicePool> &pools = all.pools;
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(all.mu);
  if ((int)pools.size() <= dev) pools.resize((size_t)dev + 1);
  DevicePool &P = pools[(size_t)dev];
  if (!P.pool) {
    hipMemPoolProps props{};
    props.allocType = hipMemAllocationTypePinned;
    props.handleTypes = hipMemHandleTypeNone;
    props.location.type = hipMemLocationTypeDevice;
    props.location.id = dev;
    hipMemPool_t pool = nullptr;
    HIPCHK(hipMemPoolCreate(&pool, &props));
    uint64_t keep = ~0ULL;
    HIPCHK(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
    HIPCHK(hipStreamCreateWithFlags(&P.st, hipStreamNonBlocking));
    P.pool = pool;
  }
  return P;
}

template <class T>
inline T *dalloc(size_t count) {
  if (count == 0) count = 1;
  const size_t bytes = count * sizeof(T);
  const DevicePool P = device_pool();
  void *p = nullptr;
  hipError_t e = hipMallocFromPoolAsync(&p, bytes, P.pool, P.st);
  if (e == hipErrorOutOfMemory) {
    // the pool's idle memory goes back to the device, then one more try
    (void)hipGetLastError();
    e = hipStreamSynchronize(P.st);
    if (e == hipSuccess) e = hipMemPoolTrimTo(P.pool, 0);
    if (e == hipSuccess) e = hipMallocFromPoolAsync(&p, bytes, P.pool, P.st);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(P.st);  // usable on every stream
  if (e != hipSuccess)
    throw Error(ICCG_ERR_NOMEM, "device allocation of " + std::to_string(bytes) +
                                    " bytes failed: " + hipGetErrorString(e));
  if (((uintptr_t)p & 255u) != 0) {
    (void)hipFreeAsync(p, P.st);
    throw Error(ICCG_ERR_DEVICE, "device allocation not 256-byte aligned");
  }
  return (T *)p;
}

template <class T>
inline T *dupload(const T *h, size_t count, hipStream_t st) {
  T *d = dalloc<T>(count);
  if (count) HIPCHK(hipMemcpyAsync(d, h, count * sizeof(T), hipMemcpyHostToDevice, st));
  return d;
}

// returns p to its pool at once (the caller has waited for the work using it)
inline void dfree(void *p) {
  if (!p) return;
  try {
    (void)hipFreeAsync(p, device_pool().st);
  } catch (...) {  // destructors free: no exception leaves here
  }
}

// Pinned host staging (hipHostMalloc) from a process-wide cache: a block
// freed here waits, idle, for the next request of at most twice its size,
// and at most kPinnedKeep bytes wait so.  hipHostFree, like hipFree, waits
// for the whole device; pinning fresh pages costs milliseconds per megabyte.
constexpr size_t kPinnedKeep = size_t(4) << 30;

struct PinnedCache {
  std::mutex mu;
  std::multimap<size_t, void *> idle;         // bytes -> block
  std::unordered_map<void *, size_t> blocks;  // every block of the cache, idle or in use
  size_t idle_bytes = 0;
};

inline PinnedCache &pinned_cache() {
  static PinnedCache &c = *new PinnedCache;  // never destroyed, as the device pools
  return c;
}

template <class T>
inline T *halloc(size_t count) {
  const size_t bytes = std::max<size_t>(count * sizeof(T), 1);
  PinnedCache &c = pinned_cache();
  {
    std::lock_guard<std::mutex> lk(c.mu);
    auto it = c.idle.lower_bound(bytes);
    if (it != c.idle.end() && it->first <= 2 * bytes) {
      void *p = it->second;
      c.idle_bytes -= it->first;
      c.idle.erase(it);
      return (T *)p;
    }
  }
  void *p = nullptr;
  hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
  if (e != hipSuccess)
    throw Error(ICCG_ERR_NOMEM, "pinned host allocation of " + std::to_string(bytes) +
                                    " bytes failed: " + hipGetErrorString(e));
  std::lock_guard<std::mutex> lk(c.mu);
  c.blocks[p] = bytes;
  return (T *)p;
}

inline void hfree(void *p) {
  if (!p) return;
  PinnedCache &c = pinned_cache();
  std::vector<void *> drop;
  {
    std::lock_guard<std::mutex> lk(c.mu);
    auto it = c.blocks.find(p);
    if (it == c.blocks.end()) {
      drop.push_back(p);
    } else {
      c.idle.emplace(it->second, p);
      c.idle_bytes += it->second;
      while (c.idle_bytes > kPinnedKeep) {  // the largest idle blocks go first
        auto last = std::prev(c.idle.end());
        c.idle_bytes -= last->first;
        c.blocks.erase(last->second);
        drop.push_back(last->second);
        c.idle.erase(last);
      }
    }
  }
  for (void *q : drop) (void)hipHostFree(q);
}

// frees every idle block of the pinned cache
inline void pinned_trim() {
  PinnedCache &c = pinned_cache();
  std::vector<void *> drop;
  {
    std::lock_guard<std::mutex> lk(c.mu);
    for (auto &kv : c.idle) {
      c.blocks.erase(kv.second);
      drop.push_back(kv.second);
    }
    c.idle.clear();
    c.idle_bytes = 0;
  }
  for (void *q : drop) (void)hipHostFree(q);
}

// Waits for the work queued on st: through a blocking-sync event when ev is
// set (the host thread sleeps until the device signals it), else with
// hipStreamSynchronize, which spins while the device works.  Threads that
// wait beside others on a shared host (the serving loop's engine, runners
// and their sets) use the event, so that their waits leave the cores to the
// threads preparing requests.
inline void stream_wait(hipStream_t st, hipEvent_t ev) {
  if (ev) {
    HIPCHK(hipEventRecord(ev, st));
    HIPCHK(hipEventSynchronize(ev));
  } else {
    HIPCHK(hipStreamSynchronize(st));
  }
}

inline hipEvent_t blocking_event() {
  hipEvent_t e = nullptr;
  HIPCHK(hipEventCreateWithFlags(&e, hipEventBlockingSync | hipEventDisableTiming));
  return e;
}

inline void free_sell(Sell &S) {
  dfree(S.off);
  dfree(S.col);
  dfree(S.v64);
  dfree(S.v32);
  dfree(S.col16);
  dfree(S.tbl);
  dfree(S.toff);
  dfree(S.inb);
  S = Sell();
}

// The 16-bit column layout of a sliced ELL (see Sell): tables of 16 entries
// per slice, or when some slice references more than 16 distinct 64-row
// blocks, of each slice's own length with offsets toff; false, and nothing
// built, when some slice references more than kTbl16Max.
inline bool sell_col16(const std::vector<int> &off, const std::vector<int> &col,
                       std::vector<uint16_t> &c16, std::vector<int> &tbl,
                       std::vector<int> &toff) {
  const int ns = (int)off.size() - 1;
  int maxc = -1;
  for (int c : col) maxc = std::max(maxc, c);
  // a slice's blocks in order of first reference: ids[first[s] ...]; a
  // block's index in the current slice's table: at[b] when seen[b] == s
  std::vector<int> ids, first((size_t)std::max(ns, 0) + 1, 0);
  std::vector<int> seen((size_t)(maxc >> 6) + 1, -1), at((size_t)(maxc >> 6) + 1, 0);
  c16.assign(col.size(), (uint16_t)0xFFFF);
  int most = 0;
  for (int s = 0; s < ns; ++s) {
    first[(size_t)s] = (int)ids.size();
    int nt = 0;
    for (int q = off[s]; q < off[s + 1]; ++q) {
      const int c = col[(size_t)q];
      if (c < 0) continue;
      const int b = c >> 6;
      if (seen[(size_t)b] != s) {
        if (nt == kTbl16Max) {
          c16.clear();
          return false;
        }
        seen[(size_t)b] = s;
        at[(size_t)b] = nt++;
        ids.push_back(b);
      }
      c16[(size_t)q] = (uint16_t)((at[(size_t)b] << 6) | (c & 63));
    }
    most = std::max(most, nt);
  }
  first[(size_t)ns] = (int)ids.size();
  toff.clear();
  if (most > 16) {
    tbl = ids;
    toff = first;
    return true;
  }
  tbl.assign((size_t)16 * std::max(ns, 0), 0);
  for (int s = 0; s < ns; ++s)
    for (int x = first[(size_t)s]; x < first[(size_t)s + 1]; ++x)
      tbl[(size_t)16 * s + (x - first[(size_t)s])] = ids[(size_t)x];
  return true;
}

// The paired layout of a 64-row sliced ELL array (A's values and 16-bit
// columns where the 16-bit layout applies): in a slice of width w at o0,
// entries 2j and 2j+1 of lane l (j < w / 2) move to o0 + 128 j + 2 l and
// o0 + 128 j + 2 l + 1; an odd last entry stays at o0 + 64 (w - 1) + l.
template <class T>
inline void pair_slices(const std::vector<int> &off, std::vector<T> &a) {
  const int ns = (int)off.size() - 1;
#pragma omp parallel
  {
    std::vector<T> tmp;
#pragma omp for schedule(static)
    for (int s = 0; s < ns; ++s) {
      const int o0 = off[(size_t)s], w = (off[(size_t)s + 1] - o0) >> 6, np = w >> 1;
      if (np == 0) continue;
      tmp.assign(a.begin() + o0, a.begin() + o0 + 128 * np);
      for (int j = 0; j < np; ++j)
        for (int l = 0; l < 64; ++l)
          for (int h = 0; h < 2; ++h)
            a[(size_t)o0 + 128 * j + 2 * l + h] = tmp[(size_t)64 * (2 * j + h) + l];
    }
  }
}

// Host layout of a 64-row sliced ELL over row ranges [beg(i), end(i)) of a
// CSR-like index array ci: off[nslice+1], col (padding -1), and for every
// included CSR position k the ELL position (pos[k]).
template <class Range>
inline void sell_layout(int n, Range range, const int *ci, std::vector<int> &off,
                        std::vector<int> &col, int &maxw, std::vector<int> *pos) {
  const int ns = (n + 63) / 64;
  off.assign((size_t)ns + 1, 0);
  maxw = 0;
  long total = 0;
  for (int s = 0; s < ns; ++s) {
    int w = 0;
    for (int i = s * 64; i < std::min(n, s * 64 + 64); ++i) {
      int b, e;
      range(i, b, e);
      w = std::max(w, e - b);
    }
    maxw = std::max(maxw, w);
    total += 64L * w;
    require(total < 0x7fffffffL, ICCG_ERR_ARG, "matrix too large for the sliced layout");
    off[s + 1] = (int)total;
  }
  col.assign((size_t)total, -1);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    int b, e;
    range(i, b, e);
    const int base = off[i >> 6] + (i & 63);
    for (int k = b; k < e; ++k) {
      const int q = base + 64 * (k - b);
      col[q] = ci[k];
      if (pos) (*pos)[k] = q;
    }
  }
}

// Sectioned sliced-ELL layout of a triangular factor for the ABMC sweeps:
// one slice per block (rows [bstart[b], bstart[b+1]), at most 64, lane =
// row - bstart[b]); entry k < S of a row is its k-th column inside the
// block, entry S + j its j-th column outside it, both in ascending order, -1
// elsewhere.  A slice's width is S plus the largest outside count of its
// rows.  blk[i] is row i's block; every in-block and outside count is <= S
// (S <= 3).  inb packs each row's in-block entries as rows within the block,
// one byte per entry (64 = absent), so the sweeps read one word instead of S
// column indices.
template <class Range>
inline void sell_layout_sections(int n, Range range, const int *ci,
                                 const std::vector<int> &bstart, const int *blk,
                                 const std::vector<int> &nout, int S, std::vector<int> &off,
                                 std::vector<int> &col, std::vector<unsigned> &inb, int &maxw,
                                 std::vector<int> *pos) {
  const int nb = (int)bstart.size() - 1;
  off.assign((size_t)nb + 1, 0);
  maxw = 0;
  long total = 0;
  for (int b = 0; b < nb; ++b) {
    int w = 0;
    for (int i = bstart[b]; i < bstart[b + 1]; ++i) w = std::max(w, nout[(size_t)i]);
    w += S;
    maxw = std::max(maxw, w);
    total += 64L * w;
    require(total < 0x7fffffffL, ICCG_ERR_ARG, "matrix too large for the sliced layout");
    off[b + 1] = (int)total;
  }
  col.assign((size_t)total, -1);
  inb.assign((size_t)nb * 64, 0x40404040u);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    int b, e;
    range(i, b, e);
    const int bl = blk[i], lo = bstart[bl], hi = bstart[bl + 1];
    const int base = off[bl] + (i - lo);
    int kin = 0, kout = S;
    unsigned packed = 0x40404040u;
    for (int k = b; k < e; ++k) {
      const int c = ci[k];
      const bool in = c >= lo && c < hi;
      if (in) packed = (packed & ~(0xFFu << (8 * kin))) | ((unsigned)(c - lo) << (8 * kin));
      const int q = base + 64 * (in ? kin++ : kout++);
      col[q] = c;
      if (pos) (*pos)[k] = q;
    }
    inb[(size_t)bl * 64 + (i - lo)] = packed;
  }
}

}  // namespace iccg

using namespace iccg;

// ---------------------------------------------------------------------------
// objects
// ---------------------------------------------------------------------------
struct iccg_csr_s {
  int device = 0;
  HostCsr h;  // host copy of the pattern and values (setup of preconditioners)
  DevCsr d;
  hipStream_t st = nullptr;
  ~iccg_csr_s() {
    (void)hipSetDevice(device);
    if (st) (void)hipStreamSynchronize(st);  // its uploads have landed
    dfree(d.rp);
    dfree(d.ci);
    dfree(d.v);
    free_sell(d.sell);
    dfree(d.sell_v32);
    if (st) (void)hipStreamDestroy(st);
  }
};

struct iccg_sched_s {
  int n = 0, mode = ICCG_SCHED_MC, nc = 0, nb = 0;
  std::vector<int> color_ptr, rows_of_color;                              // MC
  std::vector<int> color_ptr_blk, blocks_of_color, block_ptr, rows_of_block;  // ABMC
};

struct iccg_precond_s {
  int device = 0;
  int kind = ICCG_PRECOND_IDENTITY;
  bool mixed = false;
  double shift = 0.0;
  int n = 0;
  iccg_csr_t A = nullptr;
  hipStream_t st = nullptr;
  double *inv_diag = nullptr;  // Jacobi
  Factor F;                    // IC0
  DevSched sc;
  int plan_version = 0;  // bumped whenever the device plan (sc) is replaced
  ~iccg_precond_s() {
    (void)hipSetDevice(device);
    if (st) (void)hipStreamSynchronize(st);  // no work of its own is left in flight
    dfree(inv_diag);
    dfree(F.lrp);
    dfree(F.lci);
    dfree(F.lv);
    dfree(F.lv32);
    dfree(F.urp);
    dfree(F.uci);
    dfree(F.uv);
    dfree(F.uv32);
    dfree(F.l2u);
    free_sell(F.ls);
    free_sell(F.us);
    dfree(F.d64);
    dfree(F.d32);
    dfree(F.l2ls);
    dfree(F.l2us);
    dfree(sc.d_rows);
    dfree(sc.d_bstart);
    dfree(sc.d_brows);
    dfree(sc.d_lev_f);
    dfree(sc.d_lev_b);
    dfree(sc.d_nlev_f);
    dfree(sc.d_nlev_b);
    dfree(sc.d_lvl_ptr_f);
    dfree(sc.d_lvl_rows_f);
    dfree(sc.d_lvl_ptr_b);
    dfree(sc.d_lvl_rows_b);
    if (st) (void)hipStreamDestroy(st);
  }
};

