// server.cpp -- request loop over one GPU: continuous batching of ICCG
// solves of different sizes.
//
// A request carries a symmetric matrix in the caller's (natural) numbering
// and 1..k right-hand sides.  The loop
//   1. admits requests in arrival order while their estimated HBM footprint
//      fits the budget;
//   2. prepares them on host worker threads: ABMC ordering (box blocks for
//      structured grids, breadth-first blocks otherwise), P^T A P, the
//      device matrix and the IC(0) factor (factorized on the GPU on the
//      preconditioner's own stream, overlapping running solves);
//   3. runs every right-hand side as one slot of a Batch (iccg_batch_*),
//      admitting new ones whenever slots free up, so solves that arrive while
//      others run share every kernel launch with them -- except requests of
//      at least cfg.multi_rows rows, which runner threads solve with the
//      multi-right-hand-side solver (iccg_msolver_*: the operator read once
//      per set of up to four right-hand sides; a single right-hand side on
//      the single-system solver) on streams of their own, beside the batch;
//   4. returns x in the caller's numbering and frees the request's device
//      memory as soon as its last right-hand side finishes.
// The sequential mode is the plain one-at-a-time path for comparison: the
// engine thread prepares one request, solves its right-hand sides one after
// another with the single-system solver, and only then takes the next.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <omp.h>

#include "iccg_hip.h"
#include "internal.hpp"

using namespace iccg;
using Clock = std::chrono::steady_clock;

namespace {

void ck(int st) {
  if (st != ICCG_OK) throw Error(st, iccg_last_error());
}

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

enum ReqState { Q_WAIT = 0, Q_PREP, Q_READY, Q_RUN, Q_DONE, Q_FAILED };

struct Request {
  long long id = 0;
  int n = 0;
  std::vector<int> rp, ci;
  std::vector<double> v;
  int grid[3] = {0, 0, 0};
  bool has_grid = false;
  int nrhs = 0;
  std::vector<double> b;  // caller numbering, b[k*n + i]
  int max_iter = 0;
  double tol = 0;
  long long hbm = 0;      // admitted footprint estimate
  // prepared
  std::vector<int> new_of_old;
  std::vector<double> bp;  // permuted right-hand sides
  iccg_csr_t A = nullptr;
  iccg_precond_t M = nullptr;
  // results
  std::vector<double> x;
  std::vector<iccg_cg_result> res;
  int remaining = 0;
  int readers = 0;  // result() calls copying x outside the lock
  int state = Q_WAIT;
  int status = ICCG_OK;
  std::string err;
  Clock::time_point t_submit, t_ready, t_done;
  bool reported = false;
  void free_device() {
    if (M) iccg_precond_destroy(M);
    if (A) iccg_csr_destroy(A);
    M = nullptr;
    A = nullptr;
  }
  ~Request() { free_device(); }
};

// HBM bytes a request holds while resident: the matrix (CSR + sliced ELL,
// FP64 and FP32 values), the IC(0) factor (CSR of L and L^T, sliced copies,
// position maps, diagonal, schedule) and x, r, p, Ap, b, y, z per
// right-hand side.  An upper estimate from n and nnz (padding of the sliced
// layouts taken as 30%).
long long footprint(long long n, long long nnz, int nrhs, bool mixed) {
  const long long nnz_l = (nnz + n) / 2;
  const long long matrix = 4 * (n + 1) + 12 * nnz + (long long)(1.3 * 16 * nnz) + 4 * n;
  const long long factor = 2 * (4 * (n + 1) + 12 * nnz_l) + 4 * nnz_l + 8 * nnz_l +
                           (long long)(1.3 * 2 * 16 * nnz_l) + 8 * nnz_l + 12 * n + 24 * n;
  const long long vec = (5 * 8 + 2 * (mixed ? 4 : 8)) * n + 4 * 256;
  return matrix + factor + (long long)nrhs * vec;
}

}  // namespace

struct iccg_server_s {
  iccg_server_config cfg{};
  int device = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_engine, cv_done, cv_multi;
  std::deque<Request *> waiting, prep, ready, mready;
  std::map<long long, std::unique_ptr<Request>> all;
  long long next_id = 1;
  long long hbm_used = 0;
  long long submitted = 0, finished = 0, failed = 0;
  int in_prep = 0, running_reqs = 0, running_multi = 0;
  double run_ms = 0, admit_ms = 0, harvest_ms = 0;  // engine thread (under mu)
  bool stop = false, stop_reaper = false;
  std::vector<std::thread> workers, runners;
  std::thread engine, reaper;
  // device objects of finished requests, freed by the reaper thread; their
  // HBM budget returns once they are freed
  struct Reap {
    iccg_csr_t A;
    iccg_precond_t M;
    iccg_msolver_t MS;
    long long hbm;
  };
  std::deque<Reap> reaps;
  std::condition_variable cv_reap;
  // engine-thread state
  iccg_batch_t B = nullptr;
  std::vector<std::pair<Request *, int>> owner;  // slot -> (request, rhs)
  int free_slots = 0;

  ~iccg_server_s() {
    {
      std::lock_guard<std::mutex> lk(mu);
      stop = true;
    }
    cv_work.notify_all();
    cv_engine.notify_all();
    cv_multi.notify_all();
    for (auto &t : workers)
      if (t.joinable()) t.join();
    for (auto &t : runners)
      if (t.joinable()) t.join();
    if (engine.joinable()) engine.join();
    {
      std::lock_guard<std::mutex> lk(mu);
      stop_reaper = true;  // after every thread that hands it work
    }
    cv_reap.notify_all();
    if (reaper.joinable()) reaper.join();
    all.clear();
    if (B) iccg_batch_destroy(B);
  }

  // under mu: move waiting requests into preparation while the budget allows
  // (a request larger than the whole budget runs alone)
  void admit_locked() {
    while (!waiting.empty()) {
      Request *r = waiting.front();
      if (hbm_used > 0 && hbm_used + r->hbm > cfg.hbm_budget) break;
      waiting.pop_front();
      hbm_used += r->hbm;
      r->state = Q_PREP;
      prep.push_back(r);
    }
    cv_work.notify_all();
  }

  // under mu: the request's device objects (and the runner's solver of its
  // right-hand sides) go to the reaper thread.  Freeing
  // device memory (hipFree) waits for all work queued on the device, the
  // runners' long captured chunks included: on the reaper that wait delays
  // no engine, runner or worker thread.
  void reap_locked(Request *r, iccg_msolver_t ms = nullptr) {
    reaps.push_back({r->A, r->M, ms, r->hbm});
    r->A = nullptr;
    r->M = nullptr;
    cv_reap.notify_one();
  }

  void reaper_loop() {
    iccg_device_set(device);
    for (;;) {
      Reap t{};
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_reap.wait(lk, [&] { return stop_reaper || !reaps.empty(); });
        if (reaps.empty()) return;
        t = reaps.front();
        reaps.pop_front();
      }
      if (t.MS) iccg_msolver_destroy(t.MS);
      if (t.M) iccg_precond_destroy(t.M);
      if (t.A) iccg_csr_destroy(t.A);
      std::lock_guard<std::mutex> lk(mu);
      hbm_used -= t.hbm;
      admit_locked();
    }
  }

  // under mu; reaped: the request's device objects went to the reaper, which
  // returns its HBM budget
  void complete_locked(Request *r, int status, const std::string &err, bool reaped = false) {
    r->t_done = Clock::now();
    r->status = status;
    r->err = err;
    r->state = status == ICCG_OK ? Q_DONE : Q_FAILED;
    if (status == ICCG_OK)
      finished++;
    else
      failed++;
    if (!reaped) {
      hbm_used -= r->hbm;
      admit_locked();
    }
    cv_done.notify_all();
  }

  // ordering, permuted system, device matrix and IC(0) factor
  void prepare(Request &r) {
    iccg_device_set(device);
    const int n = r.n;
    std::vector<int> bo((size_t)n), bc((size_t)n);
    int nb = 0, nc = 0;
    if (r.has_grid) {
      const int nx = r.grid[0], ny = r.grid[1], nz = r.grid[2];
      const int bx = nz > 1 ? 4 : ny > 1 ? 8 : 64, by = nz > 1 ? 4 : ny > 1 ? 8 : 1,
                bz = nz > 1 ? 4 : 1;
      ck(iccg_abmc_blocking_grid(nx, ny, nz, bx, by, bz, bo.data(), &nb, bc.data(), &nc));
    } else {
      ck(iccg_abmc_blocking(n, r.rp.data(), r.ci.data(), cfg.block_size, bo.data(), &nb,
                            bc.data(), &nc));
    }
    r.new_of_old.resize((size_t)n);
    ck(iccg_perm_by_block_color(n, bo.data(), nb, bc.data(), r.new_of_old.data()));
    const long nnz = (long)r.ci.size();
    std::vector<int> lrp((size_t)n + 1), lci((size_t)nnz);
    std::vector<double> lv((size_t)nnz);
    long nnz_l = 0;
    ck(iccg_ptap_lower(n, r.rp.data(), r.ci.data(), r.v.data(), r.new_of_old.data(),
                       lrp.data(), lci.data(), lv.data(), &nnz_l));
    std::vector<int> frp((size_t)n + 1), fci((size_t)(2 * nnz_l));
    std::vector<double> fv((size_t)(2 * nnz_l));
    long nnz_f = 0;
    ck(iccg_expand_lower_to_full(n, lrp.data(), lci.data(), lv.data(), frp.data(), fci.data(),
                                 fv.data(), &nnz_f));
    std::vector<int> cpb((size_t)nc + 1), boc((size_t)nb), bptr((size_t)nb + 1), rob((size_t)n);
    ck(iccg_abmc_schedule(n, r.new_of_old.data(), bo.data(), bc.data(), nb, nc, cpb.data(),
                          boc.data(), bptr.data(), rob.data()));
    iccg_sched_t sch = nullptr;
    ck(iccg_schedule_create_abmc(n, nc, nb, cpb.data(), boc.data(), bptr.data(), rob.data(),
                                 &sch));
    struct SchedGuard {
      iccg_sched_t s;
      ~SchedGuard() { iccg_schedule_destroy(s); }
    } sg{sch};
    if (cfg.mode == ICCG_SERVE_BATCHED) {
      int wmax = 0;
      for (int i = 0; i < n; ++i) wmax = std::max(wmax, frp[(size_t)i + 1] - frp[(size_t)i]);
      require(wmax <= 16, ICCG_ERR_ARG,
              "the batched request loop takes rows of at most 16 entries (this matrix has " +
                  std::to_string(wmax) + ")");
    }
    ck(iccg_csr_create(n, nnz_f, frp.data(), fci.data(), fv.data(), &r.A));
    ck(iccg_precond_create(r.A, ICCG_PRECOND_IC0, 0.0, sch, cfg.precision, &r.M));
    r.bp.resize(r.b.size());
    for (int k = 0; k < r.nrhs; ++k)
      for (int i = 0; i < n; ++i)
        r.bp[(size_t)k * n + r.new_of_old[i]] = r.b[(size_t)k * n + i];
    std::vector<double>().swap(r.b);
    std::vector<int>().swap(r.rp);
    std::vector<int>().swap(r.ci);
    std::vector<double>().swap(r.v);
    r.x.assign((size_t)r.nrhs * n, 0.0);
    r.res.assign((size_t)r.nrhs, iccg_cg_result{});
    r.remaining = r.nrhs;
  }

  void unpermute(Request &r, int k, const double *xp) {
    double *x = r.x.data() + (size_t)k * r.n;
    for (int i = 0; i < r.n; ++i) x[i] = xp[(size_t)r.new_of_old[i]];
  }

  // requests the runner threads take (batched mode)
  bool multi_request(const Request &r) const {
    return cfg.multi_rows > 0 && r.n >= cfg.multi_rows;
  }

  void worker_loop(int threads) {
    iccg_device_set(device);
    omp_set_num_threads(threads);  // the workers share the host cores
    for (;;) {
      Request *r = nullptr;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_work.wait(lk, [&] { return stop || !prep.empty(); });
        if (stop) return;
        r = prep.front();
        prep.pop_front();
        in_prep++;
      }
      int st = ICCG_OK;
      std::string err;
      try {
        prepare(*r);
      } catch (const Error &e) {
        st = e.status;
        err = e.what();
      } catch (const std::exception &e) {
        st = ICCG_ERR_NOMEM;
        err = e.what();
      }
      std::lock_guard<std::mutex> lk(mu);
      in_prep--;
      if (st != ICCG_OK) {
        r->free_device();
        complete_locked(r, st, err);
      } else {
        r->t_ready = Clock::now();
        r->state = Q_READY;
        if (multi_request(*r)) {
          mready.push_back(r);
          cv_multi.notify_one();
        } else {
          ready.push_back(r);
          cv_engine.notify_all();
        }
      }
    }
  }

  // runner thread: solves a prepared request's right-hand sides together
  // with the multi-right-hand-side solver on a stream of its own
  void multi_loop() {
    iccg_device_set(device);
    std::vector<double> xs;
    for (;;) {
      Request *r = nullptr;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_multi.wait(lk, [&] { return stop || !mready.empty(); });
        if (stop) return;
        r = mready.front();
        mready.pop_front();
        r->state = Q_RUN;
        running_multi++;
      }
      int st = ICCG_OK;
      std::string err;
      iccg_msolver_t M = nullptr;  // freed by the reaper, with the request
      try {
        ck(iccg_msolver_create(r->A, r->M, r->nrhs, &M));
        ck(iccg_msolver_set(M, r->bp.data(), nullptr));
        ck(iccg_msolver_run(M, r->max_iter, r->tol, cfg.flags, r->res.data()));
        xs.resize((size_t)r->nrhs * r->n);
        ck(iccg_msolver_get_x(M, xs.data()));
        for (int k = 0; k < r->nrhs; ++k) unpermute(*r, k, xs.data() + (size_t)k * r->n);
      } catch (const Error &e) {
        st = e.status;
        err = e.what();
      } catch (const std::exception &e) {
        st = ICCG_ERR_NOMEM;
        err = e.what();
      }
      std::lock_guard<std::mutex> lk(mu);
      running_multi--;
      reap_locked(r, M);
      complete_locked(r, st, err, true);
    }
  }

  // batched engine: the only thread that drives the Batch
  void engine_loop() {
    iccg_device_set(device);
    std::vector<double> xp;
    for (;;) {
      std::vector<Request *> take;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_engine.wait(lk, [&] { return stop || !ready.empty() || running_reqs > 0; });
        if (stop) return;
        while (!ready.empty() && free_slots >= ready.front()->nrhs) {
          take.push_back(ready.front());
          ready.pop_front();
          take.back()->state = Q_RUN;  // under mu: read by wait() and result()
          free_slots -= take.back()->nrhs;
          running_reqs++;
        }
      }
      using clk = std::chrono::steady_clock;
      auto ms_since = [](clk::time_point a) {
        return std::chrono::duration<double, std::milli>(clk::now() - a).count();
      };
      const auto t_admit = clk::now();
      // a request the batch does not take fails alone
      for (Request *r : take) {
        try {
          for (int k = 0; k < r->nrhs; ++k) {
            int slot = -1;
            ck(iccg_batch_add(B, r->A, r->M, r->bp.data() + (size_t)k * r->n, nullptr,
                              r->max_iter, r->tol, cfg.flags, &slot));
            owner[(size_t)slot] = {r, k};
          }
        } catch (const std::exception &e) {
          const Error *ie = dynamic_cast<const Error *>(&e);
          for (int s = 0; s < (int)owner.size(); ++s)
            if (owner[(size_t)s].first == r) {
              iccg_batch_remove(B, s);
              owner[(size_t)s] = {nullptr, 0};
            }
          r->free_device();
          std::lock_guard<std::mutex> lk(mu);
          free_slots += r->nrhs;
          running_reqs--;
          complete_locked(r, ie ? ie->status : ICCG_ERR_NOMEM, e.what());
        }
      }
      try {
        int nrun = 0;
        const double a_ms = ms_since(t_admit);
        const auto t_run = clk::now();
        ck(iccg_batch_run(B, 64, &nrun));
        const double r_ms = ms_since(t_run);
        const auto t_harvest = clk::now();
        for (int s = 0; s < (int)owner.size(); ++s) {
          Request *r = owner[(size_t)s].first;
          if (!r) continue;
          int fin = 0;
          iccg_cg_result res{};
          ck(iccg_batch_result(B, s, &fin, &res));
          if (!fin) continue;
          const int k = owner[(size_t)s].second;
          xp.resize((size_t)r->n);
          ck(iccg_batch_get_x(B, s, xp.data()));
          unpermute(*r, k, xp.data());
          r->res[(size_t)k] = res;
          ck(iccg_batch_remove(B, s));
          owner[(size_t)s] = {nullptr, 0};
          if (--r->remaining == 0) {
            // the request's vectors leave HBM with it, once no runner has
            // work queued: a free waits for all the device's queued work,
            // and the batch holds the small requests, whose idle slot
            // vectors are small
            int multi = 0;
            {
              std::lock_guard<std::mutex> lk(mu);
              multi = running_multi;
            }
            if (multi == 0) ck(iccg_batch_trim(B));
            std::lock_guard<std::mutex> lk(mu);
            free_slots += r->nrhs;
            running_reqs--;
            reap_locked(r);
            complete_locked(r, ICCG_OK, "", true);
          }
        }
        const double h_ms = ms_since(t_harvest);
        std::lock_guard<std::mutex> lk(mu);
        admit_ms += a_ms;
        run_ms += r_ms;
        harvest_ms += h_ms;
      } catch (const Error &e) {
        fail_all_running(e.status, e.what());
      } catch (const std::exception &e) {
        fail_all_running(ICCG_ERR_NOMEM, e.what());
      }
    }
  }

  // an engine error ends every request in the batch
  void fail_all_running(int st, const std::string &err) {
    std::lock_guard<std::mutex> lk(mu);
    for (int s = 0; s < (int)owner.size(); ++s) {
      Request *r = owner[(size_t)s].first;
      if (!r) continue;
      iccg_batch_remove(B, s);
      owner[(size_t)s] = {nullptr, 0};
      if (r->state == Q_RUN) {
        r->free_device();
        free_slots += r->nrhs;
        running_reqs--;
        complete_locked(r, st, err);
      }
    }
  }

  // one-at-a-time engine: prepare, then solve every right-hand side alone
  void sequential_loop() {
    iccg_device_set(device);
    std::vector<double> xp;
    for (;;) {
      Request *r = nullptr;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_engine.wait(lk, [&] { return stop || !prep.empty(); });
        if (stop) return;
        r = prep.front();
        prep.pop_front();
      }
      int st = ICCG_OK;
      std::string err;
      try {
        prepare(*r);
        {
          std::lock_guard<std::mutex> lk(mu);  // state is read under mu by other threads
          r->t_ready = Clock::now();
          r->state = Q_RUN;
        }
        iccg_solver_t S = nullptr;
        ck(iccg_solver_create(r->A, r->M, &S));
        struct SolverGuard {
          iccg_solver_t s;
          ~SolverGuard() { iccg_solver_destroy(s); }
        } g{S};
        xp.resize((size_t)r->n);
        for (int k = 0; k < r->nrhs; ++k) {
          ck(iccg_solver_set(S, r->bp.data() + (size_t)k * r->n, nullptr));
          ck(iccg_solver_run(S, r->max_iter, r->tol, cfg.flags, &r->res[(size_t)k]));
          ck(iccg_solver_get_x(S, xp.data()));
          unpermute(*r, k, xp.data());
        }
      } catch (const Error &e) {
        st = e.status;
        err = e.what();
      } catch (const std::exception &e) {
        st = ICCG_ERR_NOMEM;
        err = e.what();
      }
      r->free_device();
      std::lock_guard<std::mutex> lk(mu);
      complete_locked(r, st, err);
    }
  }
};

extern "C" {

int iccg_server_create(const iccg_server_config *cfg, iccg_server_t *out) {
  return guard([&] {
    require(cfg && out, ICCG_ERR_ARG, "null argument");
    require(cfg->mode == ICCG_SERVE_BATCHED || cfg->mode == ICCG_SERVE_SEQUENTIAL, ICCG_ERR_ARG,
            "unknown serving mode");
    require(cfg->precision == ICCG_FP64 || cfg->precision == ICCG_MIXED, ICCG_ERR_ARG,
            "unknown precision");
    require(cfg->hbm_budget > 0 && cfg->max_slots > 0 && cfg->block_size >= 1 &&
                cfg->block_size <= 64 && cfg->workers >= 0 && cfg->multi_rows >= 0 &&
                cfg->multi_runners >= 0,
            ICCG_ERR_ARG, "bad server configuration");
    std::unique_ptr<iccg_server_s> S(new iccg_server_s);
    S->cfg = *cfg;
    int dev_count = 0;
    ck(iccg_device_count(&dev_count));
    require(dev_count > 0, ICCG_ERR_DEVICE, "no GPU visible");
    ck(iccg_device_get(&S->device));
    if (cfg->mode == ICCG_SERVE_BATCHED) {
      ck(iccg_batch_create(cfg->max_slots, &S->B));
      ck(iccg_batch_set_blocking(S->B, 1));  // the engine's waits sleep
      S->owner.assign((size_t)cfg->max_slots, {nullptr, 0});
      S->free_slots = cfg->max_slots;
      const int nw = cfg->workers > 0 ? cfg->workers : 8;
      const int per = std::max(1, omp_get_max_threads() / nw);
      for (int k = 0; k < nw; ++k)
        S->workers.emplace_back([p = S.get(), per] { p->worker_loop(per); });
      S->engine = std::thread([p = S.get()] { p->engine_loop(); });
      S->reaper = std::thread([p = S.get()] { p->reaper_loop(); });
      if (cfg->multi_rows > 0) {
        const int nr = cfg->multi_runners > 0 ? cfg->multi_runners : 6;
        for (int k = 0; k < nr; ++k) S->runners.emplace_back([p = S.get()] { p->multi_loop(); });
      }
    } else {
      S->engine = std::thread([p = S.get()] { p->sequential_loop(); });
    }
    *out = S.release();
  });
}

int iccg_server_destroy(iccg_server_t S) {
  delete S;
  return ICCG_OK;
}

int iccg_server_submit(iccg_server_t S, int n, long nnz, const int *rowptr, const int *colind,
                       const double *val, const int *grid, int nrhs, const double *b,
                       int max_iter, double tol, long long *id) {
  return guard([&] {
    require(S && rowptr && colind && val && b && id, ICCG_ERR_ARG, "null argument");
    require(n > 0 && nrhs > 0 && max_iter >= 0, ICCG_ERR_ARG, "bad request");
    check_csr(n, nnz, rowptr, colind);
    require(rowptr[n] == nnz, ICCG_ERR_ARG, "rowptr[n] != nnz");
    std::unique_ptr<Request> r(new Request);
    r->n = n;
    r->rp.assign(rowptr, rowptr + n + 1);
    r->ci.assign(colind, colind + nnz);
    r->v.assign(val, val + nnz);
    if (grid) {
      require(grid[0] >= 1 && grid[1] >= 1 && grid[2] >= 1 &&
                  (long long)grid[0] * grid[1] * grid[2] == n,
              ICCG_ERR_ARG, "grid dimensions do not match n");
      std::memcpy(r->grid, grid, sizeof(r->grid));
      r->has_grid = true;
    }
    r->nrhs = nrhs;
    r->b.assign(b, b + (size_t)nrhs * n);
    r->max_iter = max_iter;
    r->tol = tol;
    r->hbm = footprint(n, nnz, nrhs, S->cfg.precision == ICCG_MIXED);
    if (S->cfg.mode == ICCG_SERVE_BATCHED)
      require(nrhs <= S->cfg.max_slots, ICCG_ERR_ARG, "more right-hand sides than slots");
    r->t_submit = Clock::now();
    std::lock_guard<std::mutex> lk(S->mu);
    r->id = S->next_id++;
    *id = r->id;
    Request *p = r.get();
    S->all[r->id] = std::move(r);
    S->submitted++;
    S->waiting.push_back(p);
    S->admit_locked();
    S->cv_engine.notify_all();
  });
}

int iccg_server_wait(iccg_server_t S, long long id, int timeout_ms, long long *done_id) {
  return guard([&] {
    require(S && done_id, ICCG_ERR_ARG, "null argument");
    std::unique_lock<std::mutex> lk(S->mu);
    Request *found = nullptr;
    auto ready = [&] {
      if (id >= 0) {
        auto it = S->all.find(id);
        require(it != S->all.end(), ICCG_ERR_ARG, "unknown request id");
        Request *r = it->second.get();
        if (r->state == Q_DONE || r->state == Q_FAILED) found = r;
      } else {
        for (auto &kv : S->all) {
          Request *r = kv.second.get();
          if ((r->state == Q_DONE || r->state == Q_FAILED) && !r->reported) {
            found = r;
            break;
          }
        }
      }
      return found != nullptr;
    };
    if (timeout_ms < 0)
      S->cv_done.wait(lk, ready);
    else
      S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
    if (found) found->reported = true;
    *done_id = found ? found->id : -1;
  });
}

int iccg_server_result(iccg_server_t S, long long id, double *x, iccg_cg_result *res,
                       double *latency_ms, double *prepare_ms) {
  return guard([&] {
    require(S, ICCG_ERR_ARG, "null argument");
    Request *r = nullptr;
    {
      std::lock_guard<std::mutex> lk(S->mu);
      auto it = S->all.find(id);
      require(it != S->all.end(), ICCG_ERR_ARG, "unknown request id");
      r = it->second.get();
      require(r->state == Q_DONE || r->state == Q_FAILED, ICCG_ERR_STATE,
              "request not finished");
      if (r->state == Q_FAILED) throw Error(r->status, r->err);
      if (res) std::memcpy(res, r->res.data(), sizeof(iccg_cg_result) * r->res.size());
      if (latency_ms) *latency_ms = ms_between(r->t_submit, r->t_done);
      if (prepare_ms) *prepare_ms = ms_between(r->t_submit, r->t_ready);
      if (!x) return;
      r->readers++;  // x is final; release() waits for the copy
    }
    // x (up to nrhs n doubles) is copied without the lock, which the
    // engine, runner and worker threads take for every request they move
    std::memcpy(x, r->x.data(), sizeof(double) * r->x.size());
    std::lock_guard<std::mutex> lk(S->mu);
    if (--r->readers == 0) S->cv_done.notify_all();
  });
}

int iccg_server_release(iccg_server_t S, long long id) {
  return guard([&] {
    require(S, ICCG_ERR_ARG, "null argument");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    require(it != S->all.end(), ICCG_ERR_ARG, "unknown request id");
    const int st = it->second->state;
    require(st == Q_DONE || st == Q_FAILED, ICCG_ERR_STATE, "request not finished");
    S->cv_done.wait(lk, [&] {  // a result() copy of x in progress finishes first
      auto j = S->all.find(id);
      return j == S->all.end() || j->second->readers == 0;
    });
    it = S->all.find(id);
    require(it != S->all.end(), ICCG_ERR_ARG, "unknown request id");
    S->all.erase(it);
  });
}

int iccg_server_get_stats(iccg_server_t S, iccg_server_stats *st) {
  return guard([&] {
    require(S && st, ICCG_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    st->submitted = S->submitted;
    st->finished = S->finished;
    st->failed = S->failed;
    st->waiting = (int)S->waiting.size();
    st->preparing = (int)S->prep.size() + S->in_prep;
    st->ready = (int)S->ready.size();
    st->running = S->running_reqs + S->running_multi;
    st->hbm_in_use = S->hbm_used;
    st->engine_run_ms = S->run_ms;
    st->engine_admit_ms = S->admit_ms;
    st->engine_harvest_ms = S->harvest_ms;
  });
}

}  // extern "C"
