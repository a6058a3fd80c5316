// ic0.hpp -- drop-in for the reference's IC0 (ic0.hpp:11-76): IC(0)
// factorization and triangular solves on the GPU.  Pivot failures throw
// std::runtime_error with the reference's messages.
#pragma once

#include <string>
#include <vector>

#include "precond.hpp"

struct IC0 : iccg_detail::DevicePrecond {
  // ic0.hpp:13.  `sched` (extension) plans the factorization and the
  // triangular solves up front; `precision` (extension) ICCG_FP64 or
  // ICCG_MIXED (FP32 factor and triangular solves, FP64 CG).
  explicit IC0(const CRS &A, double shift = 1e-12, const ColorSchedule *sched = nullptr,
               int precision = ICCG_FP64) {
    init(A, ICCG_PRECOND_IC0, shift, sched, precision);
  }

  // ic0.hpp:17: z = (L L^T)^{-1} r, sequential order
  void apply(const std::vector<double> &r, std::vector<double> &z) const {
    use_schedule(nullptr);
    apply_device(r, z);
  }

  // ic0.hpp:34: dispatch on the schedule as ic0.cpp:366-393
  void apply(const std::vector<double> &r, std::vector<double> &z,
             const ColorSchedule *sched) const {
    use_schedule(sched);
    apply_device(r, z);
  }

#ifndef PIC
  std::string name() const { return "IC preconditioner"; }  // ic0.hpp:21
#else
  std::string name() const { return "IC preconditioner with multi-coloring"; }  // ic0.hpp:30
#endif
};
