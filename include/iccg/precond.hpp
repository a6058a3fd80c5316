// precond.hpp -- drop-in for the reference's precond.hpp (precond.hpp:8-45):
// Identity and Jacobi, plus the device-backed base shared with IC0.
// Each preconditioner uploads its matrix once (iccg_csr_create) and keeps
// its state in HBM; apply() is the host-vector compatibility path.
#pragma once

#include <string>
#include <vector>

#include "color.hpp"

namespace iccg_detail {

class DevicePrecond {
 public:
  iccg_precond_t handle() const { return M_.get(); }
  iccg_csr_t csr() const { return A_.get(); }
  const CRS *source() const { return src_; }
  int size() const { return n_; }
  int precision() const { return precision_; }

  // IC0::apply takes the schedule per call (ic0.hpp:34); the device plan of
  // the triangular solves follows the last schedule seen.  Same dispatch as
  // ic0.cpp:366-393: null / nc<=0 -> natural order, ABMC_BLOCK with all four
  // arrays -> ABMC, color_ptr && rows_of_color -> MC.
  void use_schedule(const ColorSchedule *s) const {
    if (kind_ != ICCG_PRECOND_IC0) return;
    if (planned_ && same(s)) return;
    SchedHandle h;
    if (s && s->nc > 0) {
      iccg_sched_t raw = nullptr;
      if (s->mode == ColorSchedule::ABMC_BLOCK && s->color_ptr_blk && s->blocks_of_color &&
          s->block_ptr && s->rows_of_block) {
        check(iccg_schedule_create_abmc(n_, s->nc, s->nb, s->color_ptr_blk, s->blocks_of_color,
                                        s->block_ptr, s->rows_of_block, &raw));
      } else if (s->color_ptr && s->rows_of_color) {
        check(iccg_schedule_create_mc(n_, s->nc, s->color_ptr, s->rows_of_color, &raw));
      }
      h.reset(raw);
    }
    check(iccg_precond_set_schedule(M_.get(), h.get()));
    sched_ = std::move(h);
    snap_ = s ? *s : ColorSchedule();
    had_ = s != nullptr;
    planned_ = true;
  }

  void apply_device(const std::vector<double> &r, std::vector<double> &z) const {
    z.resize(r.size());
    check(iccg_precond_apply(M_.get(), r.data(), z.data()));
  }

 protected:
  void init(const CRS &A, int kind, double shift, const ColorSchedule *s, int precision) {
    iccg_csr_t a = nullptr;
    check(iccg_csr_create(A.n, (long)A.colind.size(), A.rowptr.data(), A.colind.data(),
                          A.val.data(), &a));
    A_.reset(a);
    src_ = &A;
    n_ = A.n;
    kind_ = kind;
    precision_ = precision;
    iccg_precond_t m = nullptr;
    SchedHandle h;
    if (s && s->nc > 0 && kind == ICCG_PRECOND_IC0) {
      iccg_sched_t raw = nullptr;
      if (s->mode == ColorSchedule::ABMC_BLOCK && s->color_ptr_blk && s->blocks_of_color &&
          s->block_ptr && s->rows_of_block)
        check(iccg_schedule_create_abmc(n_, s->nc, s->nb, s->color_ptr_blk, s->blocks_of_color,
                                        s->block_ptr, s->rows_of_block, &raw));
      else if (s->color_ptr && s->rows_of_color)
        check(iccg_schedule_create_mc(n_, s->nc, s->color_ptr, s->rows_of_color, &raw));
      h.reset(raw);
    }
    check(iccg_precond_create(a, kind, shift, h.get(), precision, &m));
    M_.reset(m);
    sched_ = std::move(h);
    snap_ = s ? *s : ColorSchedule();
    had_ = s != nullptr;
    planned_ = kind == ICCG_PRECOND_IC0;
  }

 private:
  bool same(const ColorSchedule *s) const {
    if (!s) return !had_;
    if (!had_) return false;
    return s->nc == snap_.nc && s->color_ptr == snap_.color_ptr &&
           s->rows_of_color == snap_.rows_of_color && s->mode == snap_.mode &&
           s->nb == snap_.nb && s->color_ptr_blk == snap_.color_ptr_blk &&
           s->blocks_of_color == snap_.blocks_of_color && s->block_ptr == snap_.block_ptr &&
           s->rows_of_block == snap_.rows_of_block;
  }
  CsrHandle A_;
  PrecondHandle M_;
  mutable SchedHandle sched_;
  mutable ColorSchedule snap_{};
  mutable bool had_ = false, planned_ = false;
  const CRS *src_ = nullptr;
  int n_ = 0, kind_ = 0, precision_ = ICCG_FP64;
};

}  // namespace iccg_detail

// precond.hpp:8-17
struct Identity : iccg_detail::DevicePrecond {
  Identity() = delete;  // the device path needs the matrix
  explicit Identity(const CRS &A) { init(A, ICCG_PRECOND_IDENTITY, 0.0, nullptr, ICCG_FP64); }
  void apply(const std::vector<double> &r, std::vector<double> &z,
             const ColorSchedule * /*sched*/) const {
    z = r;
  }
  std::string name() const { return "No preconditioner"; }
};

// precond.hpp:20-45
struct Jacobi : iccg_detail::DevicePrecond {
  explicit Jacobi(const CRS &A) { init(A, ICCG_PRECOND_JACOBI, 0.0, nullptr, ICCG_FP64); }
  void apply(const std::vector<double> &r, std::vector<double> &z,
             const ColorSchedule * /*sched*/) const {
    apply_device(r, z);
  }
  std::string name() const { return "Jacobi preconditioner"; }
};
