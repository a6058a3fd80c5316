// host_order.cpp -- orderings and schedules of the solve path (color.cpp,
// block.cpp), the schedule validators (cg_crs.cpp:171-336), and the
// ordering generators the reference takes from an external tool (greedy
// node coloring, ABMC blocking + block coloring).
//
// Host code, integer only: every array here is compared bit-exactly with
// the reference.  The permutations are built with counting sorts, which
// give exactly the reference's std::stable_sort ranks in O(n).
#include <algorithm>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "internal.hpp"

namespace iccg {

// color.cpp:50-63: rank of each node under a stable sort by (color, old).
std::vector<int> perm_by_color(int n, const int *color) {
  std::vector<int> new_of_old((size_t)n);
  if (n == 0) return new_of_old;
  int lo = *std::min_element(color, color + n), hi = *std::max_element(color, color + n);
  long range = (long)hi - lo + 1;
  if (range <= 4L * n + 1024) {
    std::vector<int> cnt((size_t)range + 1, 0);
    for (int a = 0; a < n; ++a) cnt[(size_t)(color[a] - lo) + 1]++;
    for (long c = 0; c < range; ++c) cnt[c + 1] += cnt[c];
    for (int a = 0; a < n; ++a) new_of_old[a] = cnt[(size_t)(color[a] - lo)]++;
  } else {
    std::vector<int> order((size_t)n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return color[a] < color[b]; });
    for (int nw = 0; nw < n; ++nw) new_of_old[order[nw]] = nw;
  }
  return new_of_old;
}

// block.cpp:92-114: rank under a stable sort by (block_color[b], b, old).
std::vector<int> perm_by_block_color(int n, const int *block_of_old, int nb,
                                     const int *block_color) {
  for (int a = 0; a < n; ++a)
    require(block_of_old[a] >= 0 && block_of_old[a] < nb, ICCG_ERR_ARG,
            "block id out of range");
  int maxc = 0;
  for (int b = 0; b < nb; ++b) {
    require(block_color[b] >= 0, ICCG_ERR_ARG, "color id out of range");
    maxc = std::max(maxc, block_color[b]);
  }
  // order blocks by (color, id), then nodes by (block rank, old)
  std::vector<int> ccnt((size_t)maxc + 2, 0);
  for (int b = 0; b < nb; ++b) ccnt[(size_t)block_color[b] + 1]++;
  for (int c = 0; c <= maxc; ++c) ccnt[c + 1] += ccnt[c];
  std::vector<int> brank((size_t)nb);
  for (int b = 0; b < nb; ++b) brank[b] = ccnt[(size_t)block_color[b]]++;
  std::vector<int> bcnt((size_t)nb + 1, 0);
  for (int a = 0; a < n; ++a) bcnt[(size_t)brank[block_of_old[a]] + 1]++;
  for (int r = 0; r < nb; ++r) bcnt[r + 1] += bcnt[r];
  std::vector<int> new_of_old((size_t)n);
  for (int a = 0; a < n; ++a) new_of_old[a] = bcnt[(size_t)brank[block_of_old[a]]]++;
  return new_of_old;
}

// Every row strictly ascending (sorted, no repeated column); parallel.
static bool rows_strictly_sorted(int n, const int *rp, const int *ci) {
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (int i = 0; i < n; ++i)
    for (int k = rp[i] + 1; k < rp[i + 1]; ++k) bad |= ci[k - 1] >= ci[k];
  return bad == 0;
}

// color.cpp:66-141
HostCsr ptap_lower(int n, const int *rp, const int *ci, const double *v,
                   const int *new_of_old) {
  std::vector<long> cnt((size_t)n + 1, 0);
  // With unique columns in every row, every output entry (r, c) comes from
  // exactly one input entry, so the scatter can run in parallel in any
  // order: the per-row sort by column below fixes the result, and there is
  // nothing to sum.  Otherwise the scatter keeps the input order (duplicates
  // are summed in it).
  const bool par = rows_strictly_sorted(n, rp, ci);
  if (par) {
#pragma omp parallel for schedule(dynamic, 4096)
    for (int i = 0; i < n; ++i)
      for (int k = rp[i]; k < rp[i + 1]; ++k) {
        const int j = ci[k];
        if (j > i) continue;
        const int r = std::max(new_of_old[i], new_of_old[j]);
#pragma omp atomic
        cnt[(size_t)r + 1]++;
      }
  } else {
    for (int i = 0; i < n; ++i)
      for (int k = rp[i]; k < rp[i + 1]; ++k) {
        int j = ci[k];
        if (j > i) continue;
        int r = std::max(new_of_old[i], new_of_old[j]);
        cnt[(size_t)r + 1]++;
      }
  }
  for (int r = 0; r < n; ++r) cnt[r + 1] += cnt[r];
  const long tot = cnt[n];
  std::vector<int> bc((size_t)tot);
  std::vector<double> bv((size_t)tot);
  if (par) {
    std::vector<long> cur(cnt.begin(), cnt.end() - 1);
#pragma omp parallel for schedule(dynamic, 4096)
    for (int i = 0; i < n; ++i) {
      const int ip = new_of_old[i];
      for (int k = rp[i]; k < rp[i + 1]; ++k) {
        const int j = ci[k];
        if (j > i) continue;
        const int jp = new_of_old[j];
        const int r = std::max(ip, jp), c = std::min(ip, jp);
        long pos;
#pragma omp atomic capture
        pos = cur[r]++;
        bc[pos] = c;
        bv[pos] = v[k];
      }
    }
  } else {
    std::vector<long> cur(cnt.begin(), cnt.end() - 1);
    for (int i = 0; i < n; ++i) {
      int ip = new_of_old[i];
      for (int k = rp[i]; k < rp[i + 1]; ++k) {
        int j = ci[k];
        if (j > i) continue;
        int jp = new_of_old[j];
        int r = ip, c = jp;
        if (c > r) std::swap(r, c);
        long pos = cur[r]++;
        bc[pos] = c;
        bv[pos] = v[k];
      }
    }
  }
  // per row: stable sort by column, sum equal columns, drop exact zeros
  std::vector<int> rowlen((size_t)n, 0);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int r = 0; r < n; ++r) {
    long s = cnt[r], e = cnt[r + 1];
    // insertion sort: rows are short
    for (long a = s + 1; a < e; ++a) {
      int cc = bc[a];
      double vv = bv[a];
      long b = a - 1;
      while (b >= s && bc[b] > cc) {
        bc[b + 1] = bc[b];
        bv[b + 1] = bv[b];
        --b;
      }
      bc[b + 1] = cc;
      bv[b + 1] = vv;
    }
    long w = s;
    for (long t = s; t < e;) {
      int c = bc[t];
      double acc = bv[t];
      ++t;
      while (t < e && bc[t] == c) {
        acc += bv[t];
        ++t;
      }
      if (acc != 0.0) {
        bc[w] = c;
        bv[w] = acc;
        ++w;
      }
    }
    rowlen[r] = (int)(w - s);
  }
  HostCsr B;
  B.n = n;
  B.rowptr.assign((size_t)n + 1, 0);
  for (int r = 0; r < n; ++r) B.rowptr[r + 1] = B.rowptr[r] + rowlen[r];
  B.colind.resize((size_t)B.rowptr[n]);
  B.val.resize((size_t)B.rowptr[n]);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int r = 0; r < n; ++r) {
    std::copy(bc.begin() + cnt[r], bc.begin() + cnt[r] + rowlen[r], B.colind.begin() + B.rowptr[r]);
    std::copy(bv.begin() + cnt[r], bv.begin() + cnt[r] + rowlen[r], B.val.begin() + B.rowptr[r]);
  }
  return B;
}

std::vector<int> invert_perm(int n, const int *new_of_old) {
  std::vector<int> old_of_new((size_t)n, -1);
  for (int o = 0; o < n; ++o) {
    int nid = new_of_old[o];
    require(nid >= 0 && nid < n, ICCG_ERR_ARG, "invert_permutation: out of range");
    old_of_new[nid] = o;
  }
  return old_of_new;
}

// ---------------------------------------------------------------------------
// graph helpers for the generators
// ---------------------------------------------------------------------------

// adjacency of A + A^T without the diagonal, rows sorted
static void sym_adjacency(int n, const int *rp, const int *ci, std::vector<long> &ap,
                          std::vector<int> &aj) {
  bool symmetric_pattern = true;
  // quick check: pattern symmetric if every (i,j) has (j,i); test by counting
  std::vector<long> deg((size_t)n + 1, 0);
  for (int i = 0; i < n; ++i)
    for (int k = rp[i]; k < rp[i + 1]; ++k)
      if (ci[k] != i) {
        deg[i + 1]++;
        deg[(size_t)ci[k] + 1]++;
      }
  for (int i = 0; i < n; ++i) deg[i + 1] += deg[i];
  std::vector<int> tmp((size_t)deg[n]);
  std::vector<long> cur(deg.begin(), deg.end() - 1);
  for (int i = 0; i < n; ++i)
    for (int k = rp[i]; k < rp[i + 1]; ++k)
      if (ci[k] != i) {
        tmp[cur[i]++] = ci[k];
        tmp[cur[(size_t)ci[k]]++] = i;
      }
  (void)symmetric_pattern;
  ap.assign((size_t)n + 1, 0);
  std::vector<int> len((size_t)n, 0);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int i = 0; i < n; ++i) {
    int *b = tmp.data() + deg[i], *e = tmp.data() + deg[i + 1];
    std::sort(b, e);
    len[i] = (int)(std::unique(b, e) - b);
  }
  for (int i = 0; i < n; ++i) ap[i + 1] = ap[i] + len[i];
  aj.resize((size_t)ap[n]);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int i = 0; i < n; ++i)
    std::copy(tmp.begin() + deg[i], tmp.begin() + deg[i] + len[i], aj.begin() + ap[i]);
}

// first-fit coloring of graph (ap, aj) in vertex order
static int first_fit(int n, const std::vector<long> &ap, const std::vector<int> &aj,
                     int *color) {
  std::vector<int> mark;
  int nc = 0;
  for (int v = 0; v < n; ++v) color[v] = -1;
  for (int v = 0; v < n; ++v) {
    int deg = (int)(ap[v + 1] - ap[v]);
    if ((int)mark.size() < deg + 2) mark.resize((size_t)deg + 2, -1);
    for (long k = ap[v]; k < ap[v + 1]; ++k) {
      int c = color[aj[k]];
      if (c >= 0 && c <= deg) mark[c] = v;
    }
    int c = 0;
    while (mark[c] == v) ++c;
    color[v] = c;
    nc = std::max(nc, c + 1);
  }
  return nc;
}

}  // namespace iccg

using namespace iccg;

extern "C" {

int iccg_perm_by_color(int n, const int *color, int *new_of_old) {
  return guard([&] {
    require(n >= 0 && (n == 0 || (color && new_of_old)), ICCG_ERR_ARG, "bad argument");
    std::vector<int> p = perm_by_color(n, color);
    std::memcpy(new_of_old, p.data(), sizeof(int) * (size_t)n);
  });
}

int iccg_perm_by_block_color(int n, const int *block_of_old, int nb, const int *block_color,
                             int *new_of_old) {
  return guard([&] {
    require(n >= 0 && nb >= 0, ICCG_ERR_ARG, "bad argument");
    std::vector<int> p = perm_by_block_color(n, block_of_old, nb, block_color);
    std::memcpy(new_of_old, p.data(), sizeof(int) * (size_t)n);
  });
}

int iccg_ptap_lower(int n, const int *rp, const int *ci, const double *v,
                    const int *new_of_old, int *brp, int *bci, double *bv, long *nnz_out) {
  return guard([&] {
    check_csr(n, rp ? rp[n] : 0, rp, ci);
    require(new_of_old != nullptr, ICCG_ERR_ARG, "Permute_PtAP: new_of_old size mismatch");
    invert_perm(n, new_of_old);  // validates the permutation
    HostCsr B = ptap_lower(n, rp, ci, v, new_of_old);
    std::memcpy(brp, B.rowptr.data(), sizeof(int) * B.rowptr.size());
    std::memcpy(bci, B.colind.data(), sizeof(int) * B.colind.size());
    std::memcpy(bv, B.val.data(), sizeof(double) * B.val.size());
    *nnz_out = B.nnz();
  });
}

// color.cpp:164-197
int iccg_color_ptr_rows(int n, const int *color, int nc, const int *new_of_old, int *color_ptr,
                        int *rows_of_color) {
  return guard([&] {
    require(n >= 0 && nc >= 0, ICCG_ERR_ARG, "size mismatch");
    std::vector<int> old_of_new = invert_perm(n, new_of_old);
    std::vector<int> cnt((size_t)nc + 1, 0);
    for (int nw = 0; nw < n; ++nw) {
      int c = color[old_of_new[nw]];
      require((unsigned)c < (unsigned)nc, ICCG_ERR_ARG, "color out of range");
      cnt[c]++;
    }
    color_ptr[0] = 0;
    for (int c = 0; c < nc; ++c) color_ptr[c + 1] = color_ptr[c] + cnt[c];
    for (int c = 0; c < nc; ++c) cnt[c] = color_ptr[c];
    for (int nw = 0; nw < n; ++nw) rows_of_color[cnt[(size_t)color[old_of_new[nw]]]++] = nw;
  });
}

// block.cpp:116-174
int iccg_abmc_schedule(int n, const int *new_of_old, const int *block_of_old,
                       const int *block_color, int nb, int nc, int *color_ptr_blk,
                       int *blocks_of_color, int *block_ptr, int *rows_of_block) {
  return guard([&] {
    require(n >= 0 && nb >= 0 && nc >= 0, ICCG_ERR_ARG, "bad argument");
    std::vector<int> old_of_new = invert_perm(n, new_of_old);
    for (int a = 0; a < n; ++a)
      require(block_of_old[a] >= 0 && block_of_old[a] < nb, ICCG_ERR_ARG, "block id out of range");
    for (int b = 0; b < nb; ++b)
      require(block_color[b] >= 0 && block_color[b] < nc, ICCG_ERR_ARG, "color id out of range");
    std::vector<char> seen((size_t)nb, 0);
    std::vector<int> cnt((size_t)nc + 1, 0);
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
    std::fill(seen.begin(), seen.end(), 0);
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
    std::vector<int> cur(block_ptr, block_ptr + nb + 1);
    for (int i = 0; i < n; ++i) rows_of_block[i] = -1;
    for (int nw = 0; nw < n; ++nw) rows_of_block[cur[(size_t)block_of_old[old_of_new[nw]]]++] = nw;
  });
}

// cg_crs.cpp:171-224
int iccg_check_interblock(int n, const int *rp, const int *ci, const int *block_of_old, int nb,
                          const int *block_color, int nc, long long *conflicts,
                          long long *total) {
  return guard([&] {
    check_csr(n, rp ? rp[n] : 0, rp, ci);
    for (int c = 0; c < nc; ++c) conflicts[c] = 0;
    long long tot = 0;
    for (int i = 0; i < n; ++i)
      for (int k = rp[i]; k < rp[i + 1]; ++k) {
        int j = ci[k];
        if (j <= i) continue;
        int bi = block_of_old[i], bj = block_of_old[j];
        require(bi >= 0 && bi < nb && bj >= 0 && bj < nb, ICCG_ERR_ARG,
                "block id out of range while scanning edges");
        if (bi == bj) continue;
        int c1 = block_color[bi], c2 = block_color[bj];
        require(c1 >= 0 && c1 < nc && c2 >= 0 && c2 < nc, ICCG_ERR_ARG,
                "color id out of range while scanning edges");
        if (c1 == c2) {
          conflicts[c1]++;
          tot++;
        }
      }
    *total = tot;
  });
}

// cg_crs.cpp:259-336
int iccg_check_hazards(int n, const int *rp, const int *ci, const int *new_of_old,
                       const int *color_old, const int *block_of_old, int nc_cap, int *nc_out,
                       long long *hazards, long long *same_block, long long *cross_block,
                       long long *total) {
  return guard([&] {
    check_csr(n, rp ? rp[n] : 0, rp, ci);
    std::vector<int> old_of_new = invert_perm(n, new_of_old);
    int nc = 0;
    for (int i = 0; i < n; ++i) nc = std::max(nc, color_old[old_of_new[i]] + 1);
    require(nc <= nc_cap, ICCG_ERR_ARG, "nc_cap too small");
    for (int c = 0; c < nc; ++c) hazards[c] = same_block[c] = cross_block[c] = 0;
    long long tot = 0;
    for (int i = 0; i < n; ++i) {
      int c1 = color_old[old_of_new[i]];
      for (int k = rp[i]; k < rp[i + 1]; ++k) {
        int j = ci[k];
        if (j > i) continue;
        if (j >= i) break;
        if (color_old[old_of_new[j]] == c1) {
          hazards[c1]++;
          tot++;
          if (block_of_old) {
            if (block_of_old[old_of_new[i]] == block_of_old[old_of_new[j]])
              same_block[c1]++;
            else
              cross_block[c1]++;
          }
        }
      }
    }
    *nc_out = nc;
    *total = tot;
  });
}

int iccg_greedy_color(int n, const int *rp, const int *ci, int *color, int *nc) {
  return guard([&] {
    check_csr(n, rp ? rp[n] : 0, rp, ci);
    std::vector<long> ap;
    std::vector<int> aj;
    sym_adjacency(n, rp, ci, ap, aj);
    *nc = first_fit(n, ap, aj, color);
  });
}

int iccg_abmc_blocking(int n, const int *rp, const int *ci, int block_size, int *block_of_old,
                       int *nb_out, int *block_color, int *nc_out) {
  return guard([&] {
    check_csr(n, rp ? rp[n] : 0, rp, ci);
    require(block_size >= 1, ICCG_ERR_ARG, "block size must be positive");
    std::vector<long> ap;
    std::vector<int> aj;
    sym_adjacency(n, rp, ci, ap, aj);
    for (int i = 0; i < n; ++i) block_of_old[i] = -1;
    std::vector<int> queue;
    queue.reserve((size_t)block_size);
    int nb = 0;
    for (int s = 0; s < n; ++s) {
      if (block_of_old[s] != -1) continue;
      int b = nb++;
      queue.clear();
      queue.push_back(s);
      block_of_old[s] = b;
      size_t head = 0;
      while (head < queue.size() && (int)queue.size() < block_size) {
        int u = queue[head++];
        for (long k = ap[u]; k < ap[u + 1] && (int)queue.size() < block_size; ++k) {
          int w = aj[k];
          if (block_of_old[w] == -1) {
            block_of_old[w] = b;
            queue.push_back(w);
          }
        }
      }
    }
    // block graph (rows = blocks), then first-fit in block id order
    std::vector<long> bdeg((size_t)nb + 1, 0);
    std::vector<int> nodes_of((size_t)n);
    std::vector<long> bptr((size_t)nb + 1, 0);
    for (int i = 0; i < n; ++i) bptr[(size_t)block_of_old[i] + 1]++;
    for (int b = 0; b < nb; ++b) bptr[b + 1] += bptr[b];
    {
      std::vector<long> cur(bptr.begin(), bptr.end() - 1);
      for (int i = 0; i < n; ++i) nodes_of[cur[(size_t)block_of_old[i]]++] = i;
    }
    std::vector<long> gp((size_t)nb + 1, 0);
    std::vector<int> gj;
    std::vector<int> lst;
    for (int b = 0; b < nb; ++b) {
      lst.clear();
      for (long t = bptr[b]; t < bptr[b + 1]; ++t) {
        int u = nodes_of[t];
        for (long k = ap[u]; k < ap[u + 1]; ++k) {
          int ob = block_of_old[aj[k]];
          if (ob != b) lst.push_back(ob);
        }
      }
      std::sort(lst.begin(), lst.end());
      lst.erase(std::unique(lst.begin(), lst.end()), lst.end());
      gj.insert(gj.end(), lst.begin(), lst.end());
      gp[b + 1] = (long)gj.size();
    }
    *nc_out = first_fit(nb, gp, gj, block_color);
    *nb_out = nb;
  });
}

int iccg_abmc_blocking_grid(int nx, int ny, int nz, int bx, int by, int bz, int *block_of_old,
                            int *nb_out, int *block_color, int *nc_out) {
  return guard([&] {
    require(nx > 0 && ny > 0 && nz > 0 && bx > 0 && by > 0 && bz > 0, ICCG_ERR_ARG,
            "bad grid/box size");
    int mx = (nx + bx - 1) / bx, my = (ny + by - 1) / by, mz = (nz + bz - 1) / bz;
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          long id = (long)i + (long)nx * ((long)j + (long)ny * k);
          block_of_old[id] = (i / bx) + mx * ((j / by) + my * (k / bz));
        }
    int nb = mx * my * mz;
    // boxes touch through faces only (5/7-point stencils): first-fit in box
    // order gives the box parity coloring
    int nc = 1;
    for (int c = 0; c < mz; ++c)
      for (int b = 0; b < my; ++b)
        for (int a = 0; a < mx; ++a) {
          int col = (mx > 1 || my > 1 || mz > 1) ? ((a + b + c) & 1) : 0;
          block_color[a + mx * (b + my * c)] = col;
          nc = std::max(nc, col + 1);
        }
    *nb_out = nb;
    *nc_out = nc;
  });
}

}  // extern "C"
