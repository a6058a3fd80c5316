// block.hpp -- drop-in for the reference's block.hpp (block.hpp:1-40):
// ABMC block files, permutation and schedule.
#pragma once

#include <string>
#include <vector>

#include "color.hpp"

// block.hpp:10 (block.cpp:15-51)
inline std::vector<int> ReadBlockFile_1Based(const std::string &blk_file, int n_expected,
                                             int &nb_out) {
  std::vector<int> b((size_t)n_expected > 0 ? (size_t)n_expected : 1);
  iccg_detail::check(iccg_read_block_file(blk_file.c_str(), n_expected, b.data(), &nb_out));
  b.resize((size_t)n_expected);
  return b;
}

// block.hpp:18 (block.cpp:53-89)
inline std::vector<int> ReadBlockColorFile_1Based(const std::string &bcol_file, int nb_expected,
                                                  int &nc_out) {
  std::vector<int> c((size_t)nb_expected > 0 ? (size_t)nb_expected : 1);
  iccg_detail::check(
      iccg_read_block_color_file(bcol_file.c_str(), nb_expected, c.data(), &nc_out));
  c.resize((size_t)nb_expected);
  return c;
}

// block.hpp:24 (block.cpp:92-114)
inline std::vector<int> BuildPermutationByBlockColor(const std::vector<int> &block_of_old,
                                                     const std::vector<int> &block_color) {
  std::vector<int> p(block_of_old.size());
  iccg_detail::check(iccg_perm_by_block_color((int)block_of_old.size(), block_of_old.data(),
                                              (int)block_color.size(), block_color.data(),
                                              p.data()));
  return p;
}

// block.hpp:30 (block.cpp:116-174)
inline void BuildABMCSchedule(const std::vector<int> &new_of_old,
                              const std::vector<int> &block_of_old,
                              const std::vector<int> &block_color, int nb, int nc,
                              std::vector<int> &color_ptr_blk, std::vector<int> &blocks_of_color,
                              std::vector<int> &block_ptr, std::vector<int> &rows_of_block) {
  const int n = (int)new_of_old.size();
  color_ptr_blk.assign((size_t)nc + 1, 0);
  blocks_of_color.assign((size_t)nb, -1);
  block_ptr.assign((size_t)nb + 1, 0);
  rows_of_block.assign((size_t)n, -1);
  iccg_detail::check(iccg_abmc_schedule(n, new_of_old.data(), block_of_old.data(),
                                        block_color.data(), nb, nc, color_ptr_blk.data(),
                                        blocks_of_color.data(), block_ptr.data(),
                                        rows_of_block.data()));
}
