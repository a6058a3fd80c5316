// host_io.cpp -- CRS construction and the text formats of the solve path:
// Matrix Market (crs_io.cpp), .col (color.cpp), .blk/.bcol (block.cpp).
//
// Host code.  The readers parse a whole file from memory with strtoll/strtod
// (the same correctly rounded decimal conversion as the reference's
// std::istream extraction) and keep the reference's acceptance rules and
// error messages, so a file the reference rejects is rejected here with the
// same text.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

#include "internal.hpp"

namespace iccg {

namespace {

std::string lower(std::string s) {
  for (auto &c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

bool slurp(const std::string &path, std::string &out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::ostringstream ss;
  ss << f.rdbuf();
  out = ss.str();
  return true;
}

// Whitespace-separated integer scanner with istream semantics: a token that
// does not start an integer (or overflows int) stops the scan.
struct IntScanner {
  const char *p, *e;
  bool next(long long &v) {
    while (p < e && std::isspace((unsigned char)*p)) ++p;
    if (p >= e) return false;
    char *end = nullptr;
    errno = 0;
    long long x = std::strtoll(p, &end, 10);
    if (end == p || errno == ERANGE) return false;
    p = end;
    v = x;
    return true;
  }
  bool next_int(int &v) {
    long long x;
    if (!next(x)) return false;
    if (x < INT32_MIN || x > INT32_MAX) return false;
    v = (int)x;
    return true;
  }
};

}  // namespace

// crs_io.cpp:19-70 -- entries ordered by (row, column); equal (row, column)
// runs summed from 0.0 in input order; then compressed rows.  Bucketing by
// row followed by a per-row stable sort gives the same order as the
// reference's global (i,j) sort.
HostCsr coo_to_crs(int nrows, std::vector<int> &I, std::vector<int> &J,
                   std::vector<double> &V) {
  const long nin = (long)I.size();
  std::vector<long> cnt((size_t)nrows + 1, 0);
  for (long t = 0; t < nin; ++t) cnt[(size_t)I[t] + 1]++;
  for (int r = 0; r < nrows; ++r) cnt[r + 1] += cnt[r];
  std::vector<int> ord((size_t)nin);
  {
    std::vector<long> cur(cnt.begin(), cnt.end() - 1);
    for (long t = 0; t < nin; ++t) ord[(size_t)cur[(size_t)I[t]]++] = (int)t;
  }
  HostCsr A;
  A.n = nrows;
  A.rowptr.assign((size_t)nrows + 1, 0);
  std::vector<int> rowlen((size_t)nrows, 0);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int r = 0; r < nrows; ++r) {
    int *b = ord.data() + cnt[r], *e = ord.data() + cnt[r + 1];
    std::stable_sort(b, e, [&](int a, int c) { return J[a] < J[c]; });
    int u = 0;
    for (int *q = b; q < e;) {
      int c = J[*q];
      while (q < e && J[*q] == c) ++q;
      ++u;
    }
    rowlen[r] = u;
  }
  for (int r = 0; r < nrows; ++r) A.rowptr[r + 1] = A.rowptr[r] + rowlen[r];
  A.colind.resize((size_t)A.rowptr[nrows]);
  A.val.resize((size_t)A.rowptr[nrows]);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int r = 0; r < nrows; ++r) {
    int *b = ord.data() + cnt[r], *e = ord.data() + cnt[r + 1];
    long pos = A.rowptr[r];
    for (int *q = b; q < e;) {
      int c = J[*q];
      double s = 0.0;
      while (q < e && J[*q] == c) {
        s += V[*q];
        ++q;
      }
      A.colind[pos] = c;
      A.val[pos] = s;
      ++pos;
    }
  }
  return A;
}

// crs_io.cpp:73-159
HostCsr read_matrix_market(const std::string &path) {
  std::string buf;
  if (!slurp(path, buf)) throw Error(ICCG_ERR_IO, "Failed to open: " + path);
  const char *p = buf.data(), *end = buf.data() + buf.size();
  auto getline = [&](std::string &line) -> bool {
    if (p >= end) return false;
    const char *q = (const char *)std::memchr(p, '\n', (size_t)(end - p));
    if (!q) q = end;
    line.assign(p, q);
    p = (q < end) ? q + 1 : end;
    return true;
  };
  std::string line;
  if (!getline(line)) throw Error(ICCG_ERR_IO, "Empty file");
  std::string banner, mtx, storage, field, symmetry;
  {
    std::istringstream iss(line);
    if (!(iss >> banner >> mtx >> storage >> field >> symmetry))
      throw Error(ICCG_ERR_IO, "Invalid MatrixMarket header");
  }
  if (banner != "%%MatrixMarket") throw Error(ICCG_ERR_IO, "Not a MatrixMarket file");
  mtx = lower(mtx);
  storage = lower(storage);
  field = lower(field);
  symmetry = lower(symmetry);
  if (mtx != "matrix" || storage != "coordinate")
    throw Error(ICCG_ERR_IO, "Only 'matrix coordinate' is supported");
  if (!(field == "real" || field == "integer" || field == "pattern"))
    throw Error(ICCG_ERR_IO, "Unsupported field: " + field);
  if (!(symmetry == "general" || symmetry == "symmetric"))
    throw Error(ICCG_ERR_IO, "Unsupported symmetry: " + symmetry);
  // comment lines directly after the banner (crs_io.cpp:99-110)
  while (p < end && *p == '%') {
    if (!getline(line)) break;
  }
  int nrows = 0, ncols = 0, nnz = 0;
  if (!getline(line)) throw Error(ICCG_ERR_IO, "Missing size line");
  {
    std::istringstream iss(line);
    if (!(iss >> nrows >> ncols >> nnz)) throw Error(ICCG_ERR_IO, "Invalid size line");
  }
  const bool sym = symmetry == "symmetric";
  const bool has_val = field != "pattern";
  std::vector<int> I, J;
  std::vector<double> V;
  if (nnz > 0) {
    size_t cap = (size_t)nnz * (sym ? 2 : 1);
    I.reserve(cap);
    J.reserve(cap);
    V.reserve(cap);
  }
  auto push = [&](long long i, long long j, double v) {
    if (i < 0 || i >= nrows || j < 0 || j >= ncols)
      throw Error(ICCG_ERR_IO, "Index out of range");
    I.push_back((int)i);
    J.push_back((int)j);
    V.push_back(v);
  };
  for (int k = 0; k < nnz; ++k) {
    if (p >= end) throw Error(ICCG_ERR_IO, "Unexpected EOF");
    const char *q = (const char *)std::memchr(p, '\n', (size_t)(end - p));
    if (!q) q = end;
    const char *ls = p;
    p = (q < end) ? q + 1 : end;
    if (ls == q || *ls == '%') {
      --k;
      continue;
    }
    // parse within [ls, q)
    std::string tmp(ls, q);  // NUL-terminated copy for strto*
    const char *s = tmp.c_str();
    char *e1 = nullptr;
    errno = 0;
    long long ii = std::strtoll(s, &e1, 10);
    if (e1 == s) throw Error(ICCG_ERR_IO, "Invalid entry line (indices)");
    char *e2 = nullptr;
    long long jj = std::strtoll(e1, &e2, 10);
    if (e2 == e1) throw Error(ICCG_ERR_IO, "Invalid entry line (indices)");
    double vv = 1.0;
    if (has_val) {
      char *e3 = nullptr;
      vv = std::strtod(e2, &e3);
      if (e3 == e2) throw Error(ICCG_ERR_IO, "Invalid entry line (value)");
    }
    long long i = (long long)(int)(ii - 1), j = (long long)(int)(jj - 1);
    push(i, j, vv);
    if (sym && i != j) push(j, i, vv);
  }
  return coo_to_crs(nrows, I, J, V);
}

// crs_io.cpp:163-227
// Lower CSR with strictly ascending rows: row r of the full matrix is its
// own lower entries followed by column r of the strict lower part in
// ascending row order -- built in parallel (atomic cursors for the upper
// entries, then a per-row sort of that part), the same arrays as the
// sequential build below.
static HostCsr expand_sorted_lower(int n, const int *rp, const int *ci, const double *v) {
  HostCsr F;
  F.n = n;
  std::vector<int> cnt((size_t)n, 0);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int i = 0; i < n; ++i) {
#pragma omp atomic
    cnt[i] += rp[i + 1] - rp[i];
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      const int j = ci[k];
      if (j != i) {
#pragma omp atomic
        cnt[j]++;
      }
    }
  }
  F.rowptr.assign((size_t)n + 1, 0);
  for (int i = 0; i < n; ++i) F.rowptr[i + 1] = F.rowptr[i] + cnt[i];
  F.colind.resize((size_t)F.rowptr[n]);
  F.val.resize((size_t)F.rowptr[n]);
  std::vector<int> cur((size_t)n);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int i = 0; i < n; ++i) {
    const int own = rp[i + 1] - rp[i];
    std::copy(ci + rp[i], ci + rp[i + 1], F.colind.begin() + F.rowptr[i]);
    std::copy(v + rp[i], v + rp[i + 1], F.val.begin() + F.rowptr[i]);
    cur[i] = F.rowptr[i] + own;
  }
#pragma omp parallel for schedule(dynamic, 4096)
  for (int i = 0; i < n; ++i)
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      const int j = ci[k];
      if (j == i) continue;
      int pos;
#pragma omp atomic capture
      pos = cur[j]++;
      F.colind[pos] = i;
      F.val[pos] = v[k];
    }
#pragma omp parallel for schedule(dynamic, 4096)
  for (int r = 0; r < n; ++r) {
    const int s = F.rowptr[r] + (rp[r + 1] - rp[r]), e = F.rowptr[r + 1];
    for (int a = s + 1; a < e; ++a) {
      const int cc = F.colind[a];
      const double vv = F.val[a];
      int b = a - 1;
      while (b >= s && F.colind[b] > cc) {
        F.colind[b + 1] = F.colind[b];
        F.val[b + 1] = F.val[b];
        --b;
      }
      F.colind[b + 1] = cc;
      F.val[b + 1] = vv;
    }
  }
  return F;
}

HostCsr expand_lower_to_full(int n, const int *rp, const int *ci,
                             const double *v) {
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (int i = 0; i < n; ++i) {
    for (int k = rp[i] + 1; k < rp[i + 1]; ++k) bad |= ci[k - 1] >= ci[k];
    if (rp[i + 1] > rp[i]) bad |= ci[rp[i + 1] - 1] > i;
  }
  if (!bad) return expand_sorted_lower(n, rp, ci, v);
  HostCsr F;
  F.n = n;
  std::vector<long> cnt((size_t)n + 1, 0);
  for (int i = 0; i < n; ++i)
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      int j = ci[k];
      cnt[i]++;
      if (j != i) cnt[j]++;
    }
  F.rowptr.assign((size_t)n + 1, 0);
  for (int i = 0; i < n; ++i) F.rowptr[i + 1] = F.rowptr[i] + (int)cnt[i];
  F.colind.resize((size_t)F.rowptr[n]);
  F.val.resize((size_t)F.rowptr[n]);
  std::vector<int> cur(F.rowptr.begin(), F.rowptr.end());
  for (int i = 0; i < n; ++i)
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      int j = ci[k];
      int pos = cur[i]++;
      F.colind[pos] = j;
      F.val[pos] = v[k];
      if (j != i) {
        int pos2 = cur[j]++;
        F.colind[pos2] = i;
        F.val[pos2] = v[k];
      }
    }
#pragma omp parallel for schedule(dynamic, 4096)
  for (int i = 0; i < n; ++i) {
    int s = F.rowptr[i], e = F.rowptr[i + 1];
    bool sorted = true;
    for (int k = s + 1; k < e; ++k)
      if (F.colind[k - 1] > F.colind[k]) {
        sorted = false;
        break;
      }
    if (sorted) continue;
    std::vector<int> idx((size_t)(e - s));
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(),
                     [&](int a, int b) { return F.colind[s + a] < F.colind[s + b]; });
    std::vector<int> c2(idx.size());
    std::vector<double> v2(idx.size());
    for (size_t t = 0; t < idx.size(); ++t) {
      c2[t] = F.colind[s + idx[t]];
      v2[t] = F.val[s + idx[t]];
    }
    std::copy(c2.begin(), c2.end(), F.colind.begin() + s);
    std::copy(v2.begin(), v2.end(), F.val.begin() + s);
  }
  return F;
}

void check_csr(int n, long nnz, const int *rp, const int *ci) {
  require(n >= 0 && nnz >= 0, ICCG_ERR_ARG, "negative matrix size");
  require(rp != nullptr && (nnz == 0 || ci != nullptr), ICCG_ERR_ARG, "null CSR array");
  require(rp[0] == 0 && (long)rp[n] == nnz, ICCG_ERR_ARG,
          "rowptr[0] must be 0 and rowptr[n] must equal nnz");
  for (int i = 0; i < n; ++i)
    require(rp[i] <= rp[i + 1], ICCG_ERR_ARG, "rowptr is not monotone");
  for (long k = 0; k < nnz; ++k)
    require(ci[k] >= 0 && ci[k] < n, ICCG_ERR_ARG, "Index out of range");
}

bool rows_sorted_unique(int n, const int *rp, const int *ci) {
  for (int i = 0; i < n; ++i)
    for (int k = rp[i] + 1; k < rp[i + 1]; ++k)
      if (ci[k - 1] >= ci[k]) return false;
  return true;
}

// ---- .col / .blk / .bcol (color.cpp:14-47, block.cpp:15-89) -------------
namespace {
std::vector<int> read_ids(const std::string &path, int n_expected, int &count,
                          const std::string &open_msg, const std::string &prefix,
                          const std::string &first_msg, const std::string &pos_msg,
                          const std::string &id_msg, const std::string &val_msg,
                          const std::string &cover_msg) {
  std::string buf;
  if (!slurp(path, buf)) throw Error(ICCG_ERR_IO, open_msg + path);
  IntScanner sc{buf.data(), buf.data() + buf.size()};
  int cnt = 0;
  if (!sc.next_int(cnt)) throw Error(ICCG_ERR_IO, prefix + first_msg);
  if (cnt <= 0) throw Error(ICCG_ERR_IO, prefix + pos_msg);
  std::vector<int> out((size_t)std::max(n_expected, 0), -1);
  int assigned = 0;
  int id = 0, val = 0;
  while (sc.next_int(id) && sc.next_int(val)) {
    if (id < 1 || id > n_expected) throw Error(ICCG_ERR_IO, prefix + id_msg + std::to_string(id));
    if (val < 1 || val > cnt) throw Error(ICCG_ERR_IO, prefix + val_msg + std::to_string(val));
    if (out[id - 1] == -1) ++assigned;
    out[id - 1] = val - 1;
  }
  if (assigned != n_expected)
    throw Error(ICCG_ERR_IO, prefix + cover_msg + "assigned=" + std::to_string(assigned) +
                                 " expected=" + std::to_string(n_expected));
  count = cnt;
  return out;
}
}  // namespace

std::vector<int> read_color_file(const std::string &path, int n, int &nc) {
  return read_ids(path, n, nc, "cannot open color file: ", "",
                  "failed to read first line (nc) from color file", "nc must be positive",
                  "node_id out of range in color file: ", "color_id out of range in color file: ",
                  "color file does not assign all nodes: ");
}
std::vector<int> read_block_file(const std::string &path, int n, int &nb) {
  const std::string pre = "parse error in '" + path + "': ";
  return read_ids(path, n, nb, "cannot open block file: ", pre, "failed to read first line (nb)",
                  "nb must be positive", "node_id out of range: ", "block_id out of range: ",
                  "not all nodes assigned: ");
}
std::vector<int> read_block_color_file(const std::string &path, int nb, int &nc) {
  const std::string pre = "parse error in '" + path + "': ";
  return read_ids(path, nb, nc, "cannot open block-color file: ", pre,
                  "failed to read first line (nc)", "nc must be positive",
                  "block_id out of range: ", "color_id out of range: ",
                  "not all blocks assigned: ");
}

void write_ids(const std::string &path, int n, const int *vals, int count) {
  FILE *f = std::fopen(path.c_str(), "w");
  if (!f) throw Error(ICCG_ERR_IO, "cannot open for writing: " + path);
  std::fprintf(f, "%d\n", count);
  for (int i = 0; i < n; ++i) std::fprintf(f, "%d %d\n", i + 1, vals[i] + 1);
  if (std::fclose(f) != 0) throw Error(ICCG_ERR_IO, "write failed: " + path);
}

}  // namespace iccg

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
using namespace iccg;

extern "C" {

int iccg_mm_read(const char *path, int *n, long *nnz, int **rowptr, int **colind,
                 double **val) {
  return guard([&] {
    require(path && n && nnz && rowptr && colind && val, ICCG_ERR_ARG, "null argument");
    HostCsr A = read_matrix_market(path);
    size_t m = A.colind.size();
    int *rp = (int *)std::malloc(sizeof(int) * (A.rowptr.size()));
    int *ci = (int *)std::malloc(sizeof(int) * (m ? m : 1));
    double *v = (double *)std::malloc(sizeof(double) * (m ? m : 1));
    if (!rp || !ci || !v) {
      std::free(rp);
      std::free(ci);
      std::free(v);
      throw Error(ICCG_ERR_NOMEM, "out of host memory");
    }
    std::memcpy(rp, A.rowptr.data(), sizeof(int) * A.rowptr.size());
    if (m) {
      std::memcpy(ci, A.colind.data(), sizeof(int) * m);
      std::memcpy(v, A.val.data(), sizeof(double) * m);
    }
    *n = A.n;
    *nnz = (long)m;
    *rowptr = rp;
    *colind = ci;
    *val = v;
  });
}

void iccg_free_host(void *p) { std::free(p); }

int iccg_coo_to_crs(int nrows, long nin, const int *I, const int *J, const double *V,
                    int *rowptr, int *colind, double *val, long *nnz_out) {
  return guard([&] {
    require(nrows >= 0 && nin >= 0 && rowptr && nnz_out, ICCG_ERR_ARG, "bad argument");
    std::vector<int> vi(I, I + nin), vj(J, J + nin);
    std::vector<double> vv(V, V + nin);
    for (long t = 0; t < nin; ++t)
      require(vi[t] >= 0 && vi[t] < nrows, ICCG_ERR_ARG, "Index out of range");
    HostCsr A = coo_to_crs(nrows, vi, vj, vv);
    std::memcpy(rowptr, A.rowptr.data(), sizeof(int) * A.rowptr.size());
    std::memcpy(colind, A.colind.data(), sizeof(int) * A.colind.size());
    std::memcpy(val, A.val.data(), sizeof(double) * A.val.size());
    *nnz_out = A.nnz();
  });
}

int iccg_expand_lower_to_full(int n, const int *rp, const int *ci, const double *v, int *frp,
                              int *fci, double *fv, long *nnz_out) {
  return guard([&] {
    check_csr(n, rp ? rp[n] : 0, rp, ci);
    HostCsr F = expand_lower_to_full(n, rp, ci, v);
    std::memcpy(frp, F.rowptr.data(), sizeof(int) * F.rowptr.size());
    std::memcpy(fci, F.colind.data(), sizeof(int) * F.colind.size());
    std::memcpy(fv, F.val.data(), sizeof(double) * F.val.size());
    *nnz_out = F.nnz();
  });
}

int iccg_read_color_file(const char *path, int n, int *color, int *nc) {
  return guard([&] {
    std::vector<int> c = read_color_file(path, n, *nc);
    std::memcpy(color, c.data(), sizeof(int) * (size_t)n);
  });
}
int iccg_read_block_file(const char *path, int n, int *block_of_old, int *nb) {
  return guard([&] {
    std::vector<int> c = read_block_file(path, n, *nb);
    std::memcpy(block_of_old, c.data(), sizeof(int) * (size_t)n);
  });
}
int iccg_read_block_color_file(const char *path, int nb, int *block_color, int *nc) {
  return guard([&] {
    std::vector<int> c = read_block_color_file(path, nb, *nc);
    std::memcpy(block_color, c.data(), sizeof(int) * (size_t)nb);
  });
}
int iccg_write_color_file(const char *path, int n, const int *color, int nc) {
  return guard([&] { write_ids(path, n, color, nc); });
}
int iccg_write_block_file(const char *path, int n, const int *block_of_old, int nb) {
  return guard([&] { write_ids(path, n, block_of_old, nb); });
}
int iccg_write_block_color_file(const char *path, int nb, const int *block_color, int nc) {
  return guard([&] { write_ids(path, nb, block_color, nc); });
}

}  // extern "C"
