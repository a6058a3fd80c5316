// detail.hpp -- shared helpers of the C++ drop-in headers (not part of the
// reference's interface): status -> std::runtime_error, RAII device handles.
#pragma once

#include <memory>
#include <stdexcept>
#include <string>

#include "../iccg_hip.h"

namespace iccg_detail {

// The reference reports failures as std::runtime_error; the C ABI returns a
// status and keeps the same message in iccg_last_error().
inline void check(int status) {
  if (status != ICCG_OK) throw std::runtime_error(iccg_last_error());
}

template <class H, int (*Destroy)(H)>
struct Deleter {
  void operator()(H h) const {
    if (h) Destroy(h);
  }
};

using CsrHandle =
    std::unique_ptr<iccg_csr_s, Deleter<iccg_csr_t, iccg_csr_destroy>>;
using SchedHandle =
    std::unique_ptr<iccg_sched_s, Deleter<iccg_sched_t, iccg_schedule_destroy>>;
using PrecondHandle =
    std::unique_ptr<iccg_precond_s, Deleter<iccg_precond_t, iccg_precond_destroy>>;

}  // namespace iccg_detail
