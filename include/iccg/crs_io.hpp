// crs_io.hpp -- drop-in for the reference's crs_io.hpp (crs_io.hpp:10-56):
// struct CRS, read_matrix_market_crs, expand_lower_to_full.  The reader and
// the expansion run in libiccg_hip.so (host code, same results bit for bit).
#pragma once

#include <cstring>
#include <string>
#include <vector>

#include "detail.hpp"

// 0-based CSR with column-sorted rows (crs_io.hpp:10-15)
struct CRS {
  int n = 0;
  std::vector<int> rowptr;  // n+1
  std::vector<int> colind;  // nnz
  std::vector<double> val;  // nnz
};

// crs_io.hpp:52
inline CRS read_matrix_market_crs(const std::string &filepath) {
  int n = 0;
  long nnz = 0;
  int *rp = nullptr, *ci = nullptr;
  double *v = nullptr;
  iccg_detail::check(iccg_mm_read(filepath.c_str(), &n, &nnz, &rp, &ci, &v));
  CRS A;
  A.n = n;
  A.rowptr.assign(rp, rp + n + 1);
  A.colind.assign(ci, ci + nnz);
  A.val.assign(v, v + nnz);
  iccg_free_host(rp);
  iccg_free_host(ci);
  iccg_free_host(v);
  return A;
}

// crs_io.hpp:56
inline CRS expand_lower_to_full(const CRS &L) {
  const long cap = 2 * (long)L.colind.size() + 1;
  CRS F;
  F.n = L.n;
  F.rowptr.resize((size_t)L.n + 1);
  F.colind.resize((size_t)cap);
  F.val.resize((size_t)cap);
  long nnz = 0;
  iccg_detail::check(iccg_expand_lower_to_full(L.n, L.rowptr.data(), L.colind.data(),
                                               L.val.data(), F.rowptr.data(), F.colind.data(),
                                               F.val.data(), &nnz));
  F.colind.resize((size_t)nnz);
  F.val.resize((size_t)nnz);
  return F;
}
