// color.hpp -- drop-in for the reference's color.hpp (color.hpp:1-23):
// multicolor ordering functions and struct ColorSchedule.
#pragma once

#include <string>
#include <vector>

#include "crs_io.hpp"

// color.hpp:4 (color.cpp:14-47)
inline std::vector<int> ReadColorFile_1Based(const std::string &color_file, int n_expected,
                                             int &nc_out) {
  std::vector<int> color((size_t)n_expected > 0 ? (size_t)n_expected : 1);
  iccg_detail::check(iccg_read_color_file(color_file.c_str(), n_expected, color.data(), &nc_out));
  color.resize((size_t)n_expected);
  return color;
}

// color.hpp:5 (color.cpp:50-63)
inline std::vector<int> BuildPermutationByColor(const std::vector<int> &color) {
  std::vector<int> p(color.size());
  iccg_detail::check(iccg_perm_by_color((int)color.size(), color.data(), p.data()));
  return p;
}

// color.hpp:6 (color.cpp:66-141)
inline CRS Permute_PtAP_LowerCRS_to_LowerCRS(const CRS &A, const std::vector<int> &new_of_old) {
  if ((int)new_of_old.size() != A.n)
    throw std::runtime_error("Permute_PtAP: new_of_old size mismatch");
  if ((int)A.rowptr.size() != A.n + 1)
    throw std::runtime_error("Permute_PtAP: A.rowptr size mismatch");
  CRS B;
  B.n = A.n;
  B.rowptr.resize((size_t)A.n + 1);
  B.colind.resize(A.colind.size() + 1);
  B.val.resize(A.colind.size() + 1);
  long nnz = 0;
  iccg_detail::check(iccg_ptap_lower(A.n, A.rowptr.data(), A.colind.data(), A.val.data(),
                                     new_of_old.data(), B.rowptr.data(), B.colind.data(),
                                     B.val.data(), &nnz));
  B.colind.resize((size_t)nnz);
  B.val.resize((size_t)nnz);
  return B;
}

// color.hpp:7 (color.cpp:164-197)
inline void BuildColorPtrRows_FromPermutationAndColor(const std::vector<int> &color, int nc,
                                                      const std::vector<int> &new_of_old,
                                                      std::vector<int> &color_ptr,
                                                      std::vector<int> &rows_of_color) {
  if (new_of_old.size() != color.size()) throw std::runtime_error("size mismatch");
  color_ptr.assign((size_t)nc + 1, 0);
  rows_of_color.assign(color.size(), -1);
  iccg_detail::check(iccg_color_ptr_rows((int)color.size(), color.data(), nc, new_of_old.data(),
                                         color_ptr.data(), rows_of_color.data()));
}

// color.hpp:9-23: a non-owning view of the schedule arrays
struct ColorSchedule {
  int nc = 0;
  const int *color_ptr = nullptr;
  const int *rows_of_color = nullptr;
  enum Mode { MC_NODE, ABMC_BLOCK } mode = MC_NODE;
  int nb = 0;
  const int *color_ptr_blk = nullptr;
  const int *blocks_of_color = nullptr;
  const int *block_ptr = nullptr;
  const int *rows_of_block = nullptr;
};
