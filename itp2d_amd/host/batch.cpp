// Batches of independent ITP problems (include/itp2d_batch.h): worker host
// threads, one context (HIP stream) each, pull instances from a shared queue
// ordered largest first and run the loop of itpsystem.cpp on them.
#include "itp2d_batch.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

struct Instance {
  itps_params p{};
  std::vector<double> eps;
  std::vector<double> V;
  bool has_v = false;
  // results
  int status = ITP_OK;
  int steps = 0;
  size_t converged = 0;
  int error_flag = 0;
  double time_s = 0;
  std::vector<double> E, sigma;
  std::vector<int> flags;
  int mixed[4] = {0, 0, 0, 0};
};

}  // namespace

struct itp_batch {
  int device = 0;
  int workers = 1;
  bool autotune = false;
  std::vector<Instance> inst;
  double wall = 0;
};

namespace {

int run_one(itp_ctx* ctx, Instance& in) {
  in.p.eps_values = in.eps.empty() ? nullptr : in.eps.data();
  in.p.n_eps = in.eps.size();
  itps_system* sys = nullptr;
  int rc = itps_create(ctx, &in.p, in.has_v ? in.V.data() : nullptr, nullptr, &sys);
  if (rc) return rc;
  rc = itps_run(sys);
  const size_t N = in.p.N;
  in.steps = itps_total_steps(sys);
  in.converged = itps_num_converged(sys);
  in.error_flag = itps_error_flag(sys);
  in.time_s = itps_total_time(sys);
  in.E.assign(N, 0.0);
  in.sigma.assign(N, 0.0);
  in.flags.assign(N, 0);
  itps_sorted(sys, in.E.data(), in.sigma.data(), nullptr);
  itps_converged(sys, in.flags.data());
  itps_mixed_stats(sys, in.mixed);
  itps_destroy(sys);
  return rc;
}

}  // namespace

extern "C" {

int itp_batch_create(int device, int workers, itp_batch** out) {
  if (!out || workers < 1 || workers > 64 || device < 0) return ITP_E_INVALID;
  *out = nullptr;
  try {
    std::unique_ptr<itp_batch> b(new itp_batch);
    b->device = device;
    b->workers = workers;
    *out = b.release();
    return ITP_OK;
  } catch (const std::bad_alloc&) {
    return ITP_E_NOMEM;
  }
}

void itp_batch_destroy(itp_batch* b) { delete b; }

int itp_batch_set_autotune(itp_batch* b, int on) {
  if (!b) return ITP_E_INVALID;
  b->autotune = on != 0;
  return ITP_OK;
}

int itp_batch_add(itp_batch* b, const itps_params* p, const double* V, size_t* index) {
  if (!b || !p || p->sx == 0 || p->sy == 0 || p->N == 0) return ITP_E_INVALID;
  try {
    Instance in;
    in.p = *p;
    if (p->n_eps && p->eps_values) in.eps.assign(p->eps_values, p->eps_values + p->n_eps);
    in.p.eps_values = nullptr;
    in.p.n_eps = 0;
    if (V) {
      in.V.assign(V, V + p->sx * p->sy);
      in.has_v = true;
    }
    b->inst.push_back(std::move(in));
    if (index) *index = b->inst.size() - 1;
    return ITP_OK;
  } catch (const std::bad_alloc&) {
    return ITP_E_NOMEM;
  }
}

size_t itp_batch_size(const itp_batch* b) { return b ? b->inst.size() : 0; }

int itp_batch_run(itp_batch* b) {
  if (!b) return ITP_E_INVALID;
  const size_t n = b->inst.size();
  b->wall = 0;
  if (n == 0) return ITP_OK;
  std::vector<size_t> order(n);
  for (size_t i = 0; i < n; i++) order[i] = i;
  auto cost = [&](size_t i) {
    const itps_params& p = b->inst[i].p;
    return (double)p.N * (double)p.sx * (double)p.sy;
  };
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return cost(x) > cost(y); });
  std::atomic<size_t> next(0);
  std::atomic<int> first_err(ITP_OK);
  const int nw = (int)std::min<size_t>((size_t)b->workers, std::max<size_t>(n, 1));
  const auto t0 = std::chrono::steady_clock::now();
  auto worker = [&]() {
    itp_ctx* ctx = nullptr;
    int rc = itp_ctx_create(b->device, &ctx);
    if (!rc && b->autotune) rc = itp_ctx_set_autotune(ctx, 1);
    for (;;) {
      const size_t k = next.fetch_add(1);
      if (k >= n) break;
      Instance& in = b->inst[order[k]];
      in.status = rc ? rc : run_one(ctx, in);
      if (in.status) {
        int expect = ITP_OK;
        first_err.compare_exchange_strong(expect, in.status);
      }
    }
    if (ctx) itp_ctx_destroy(ctx);
  };
  std::vector<std::thread> th;
  try {
    for (int w = 0; w < nw; w++) th.emplace_back(worker);
  } catch (const std::system_error&) {
    // fewer threads than asked: the ones started drain the queue
  }
  if (th.empty()) worker();
  for (auto& t : th) t.join();
  b->wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return first_err.load();
}

double itp_batch_wall_time(const itp_batch* b) { return b ? b->wall : 0.0; }

int itp_batch_result(const itp_batch* b, size_t i, int* status, int* steps, size_t* converged, int* error_flag,
                     double* time_s) {
  if (!b || i >= b->inst.size()) return ITP_E_INVALID;
  const Instance& in = b->inst[i];
  if (status) *status = in.status;
  if (steps) *steps = in.steps;
  if (converged) *converged = in.converged;
  if (error_flag) *error_flag = in.error_flag;
  if (time_s) *time_s = in.time_s;
  return ITP_OK;
}

int itp_batch_mixed_stats(const itp_batch* b, size_t i, int* out) {
  if (!b || !out || i >= b->inst.size()) return ITP_E_INVALID;
  std::copy(b->inst[i].mixed, b->inst[i].mixed + 4, out);
  return ITP_OK;
}

int itp_batch_sorted(const itp_batch* b, size_t i, double* E, double* sigma, int* converged_flags) {
  if (!b || i >= b->inst.size()) return ITP_E_INVALID;
  const Instance& in = b->inst[i];
  if (in.E.size() != in.p.N) return ITP_E_INVALID;  // not run (or failed before its loop)
  if (E) std::copy(in.E.begin(), in.E.end(), E);
  if (sigma) std::copy(in.sigma.begin(), in.sigma.end(), sigma);
  if (converged_flags) std::copy(in.flags.begin(), in.flags.end(), converged_flags);
  return ITP_OK;
}

}  // extern "C"
