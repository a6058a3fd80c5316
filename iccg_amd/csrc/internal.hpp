// internal.hpp -- shared host-side helpers of libiccg_hip.so (not installed).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "iccg_hip.h"

namespace iccg {

// Error carried through host code and turned into a status at the C ABI.
struct Error : std::runtime_error {
  int status;
  Error(int st, const std::string &msg) : std::runtime_error(msg), status(st) {}
};

void set_last_error(const std::string &msg);

// Runs f, maps exceptions to status codes + iccg_last_error().
template <class F>
int guard(F &&f) {
  try {
    f();
    return ICCG_OK;
  } catch (const Error &e) {
    set_last_error(e.what());
    return e.status;
  } catch (const std::bad_alloc &) {
    set_last_error("out of host memory");
    return ICCG_ERR_NOMEM;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return ICCG_ERR_ARG;
  }
}

inline void require(bool ok, int st, const std::string &msg) {
  if (!ok) throw Error(st, msg);
}

// Host CSR owned by the library (int32 indices, FP64 values).
struct HostCsr {
  int n = 0;
  std::vector<int> rowptr, colind;
  std::vector<double> val;
  long nnz() const { return (long)colind.size(); }
};

// ---- host algorithms (host_io.cpp / host_order.cpp) ---------------------
HostCsr coo_to_crs(int nrows, std::vector<int> &I, std::vector<int> &J,
                   std::vector<double> &V);
HostCsr read_matrix_market(const std::string &path);
HostCsr expand_lower_to_full(int n, const int *rp, const int *ci,
                             const double *v);
HostCsr ptap_lower(int n, const int *rp, const int *ci, const double *v,
                   const int *new_of_old);

// Validates a CSR structure (monotone rowptr, indices in range).
void check_csr(int n, long nnz, const int *rp, const int *ci);
bool rows_sorted_unique(int n, const int *rp, const int *ci);

// Level of each row in the lower-triangular dependency DAG restricted to
// rows [lo, hi) and dependencies with col in [lo, hi): used for in-block
// schedules (ABMC) and whole-matrix level sets (natural order).
}  // namespace iccg
