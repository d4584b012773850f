// Serving form: a request loop over one context with continuous batching.
//
// Requests (a (T) system each: seeded synthetic or host tensors) wait in a
// FIFO queue.  Each step
//   1. admits queued requests in arrival order while a request slot is free
//      and the resident arrays of the admitted systems fit the HBM budget
//      (the system is generated or uploaded at admission and stays resident
//      until its energy is done);
//   2. picks windows from the active requests (shortest remaining work
//      first, or an equal share each) up to a target iteration time, and
//      runs the windows of small requests in one persistent launch of the
//      batched kernel, and those of large requests (No at or above the
//      context's "serve_plain_no", and every bf16x3 request) through the
//      plain path's class-GEMM launches, so the work of requests admitted at
//      different times shares the device in every step;
//   3. adds each request's window energies to its running compensated sum
//      in canonical window order and retires the requests whose windows are
//      all done, freeing their memory for the queue.
// A request's energy is the same bits alone or in any mix: every window's
// (sum, comp) pair is computed independently of the other windows (by one
// workgroup of the batched kernel, or from the window's tuple energies in
// slot order on the plain path), and the request's sum runs over the same
// window sequence however the windows were split over steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <map>
#include <numeric>
#include <string>
#include <vector>

#include "batch.hpp"

using namespace atrip_amd;

namespace {

using Clock = std::chrono::steady_clock;

enum ReqState { QUEUED = 0, ACTIVE = 1, DONE = 2, FAILED = 3 };

struct Request {
  int64_t id = 0;
  int64_t No = 0, Nv = 0;
  uint64_t seed = 0;
  bool synthetic = true;
  atrip_sys_desc desc{};
  int state = QUEUED;
  int sys = 0;
  uint64_t bytes = 0;
  BatchTune plan;
  std::vector<Win> wins;     // batched windows
  bool plain = false;        // windows run on the plain path
  size_t nwin = 0;           // windows of the request (either kind)
  size_t cursor = 0;         // windows done
  double sum = 0.0, comp = 0.0, energy = 0.0;
  uint64_t tuples = 0;
  double t_submit = 0, t_admit = 0, t_done = 0;  // ms on the server clock
  std::string err;
};

}  // namespace

struct atrip_hip_server {
  atrip_hip_ctx *ctx = nullptr;
  uint64_t budget = 0;
  int max_active = 64;
  double iter_ms = 100.0;
  int policy = 0;  // 0 shortest remaining first, 1 equal share
  int precision = ATRIP_FP64;
  Clock::time_point t0 = Clock::now();
  std::map<int64_t, Request> reqs;
  std::deque<int64_t> queue;
  std::vector<int64_t> active;
  uint64_t resident = 0, peak_resident = 0;
  uint64_t tuples_done = 0, launches = 0;
  double busy_ms = 0.0;
  int64_t next_id = 1;
  double now() const {
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
  }
};

namespace {

void fail_request(atrip_hip_server *srv, Request &r, const std::string &why) {
  r.state = FAILED;
  r.err = why;
  r.t_done = srv->now();
}

// Releases an admitted request's system on a failure path; a failure to
// release is left for the next call to report.
void drop_system(atrip_hip_server *srv, Request &r) {
  if (r.sys) (void)atrip_hip_release(srv->ctx, r.sys);
  r.sys = 0;
  srv->resident -= r.bytes;
}

void retire(atrip_hip_server *srv, Request &r) {
  // src/atrip/Atrip.cxx:1108-1111: E = -sum
  r.energy = -(r.sum + r.comp);
  r.state = DONE;
  r.t_done = srv->now();
  if (r.sys && atrip_hip_release(srv->ctx, r.sys) != 0)
    throw Fail{srv->ctx->err};
  r.sys = 0;
  srv->resident -= r.bytes;
}

// Whether a request's windows run on the plain path: large requests (the
// plain class-GEMM kernels reach a far higher fraction of the MFMA peak once
// No^2 rows fill their tiles), No beyond the batched kernel's limit, and
// every bf16x3 request (a plain-path form only).
bool runs_plain(const atrip_hip_server *srv, int64_t No) {
  return srv->precision == ATRIP_BF16X3 || No > MAX_NO ||
         No >= srv->ctx->serve_plain_no;
}

// Initial estimate of the workgroup-seconds of one plain-path window (every
// workgroup of the grid busy for its GPU time); replaced by the measured
// figure after the request's first step.
double plain_window_estimate(int64_t No, int64_t Nv, int nb, int precision,
                             int grid) {
  const double tflops = precision == ATRIP_FP64         ? 50.0
                        : precision == ATRIP_SPLIT_FP32 ? 70.0
                        : precision == ATRIP_FP32       ? 85.0
                                                        : 120.0;
  const double flop = 12.0 * (double)No * No * No * (double)(No + Nv) *
                      (double)nb * nb * nb;
  return flop / (tflops * 1e12) * grid;
}

// Admit queued requests in arrival order while they fit.
void admit(atrip_hip_server *srv, int grid) {
  atrip_hip_ctx *ctx = srv->ctx;
  while (!srv->queue.empty() && (int)srv->active.size() < srv->max_active) {
    Request &r = srv->reqs.at(srv->queue.front());
    // served systems keep only the operand form the server computes in
    const int store = srv->precision == ATRIP_FP64         ? 0
                      : srv->precision == ATRIP_SPLIT_FP32 ? 1
                      : srv->precision == ATRIP_FP32       ? 2
                                                           : 3;
    r.bytes = system_bytes(r.No, r.Nv, false, store);
    if (r.bytes > srv->budget) {
      fail_request(srv, r, "system larger than the HBM budget");
      srv->queue.pop_front();
      continue;
    }
    if (srv->resident + r.bytes > srv->budget) break;  // wait for memory
    int sys = 0;
    const int store_before = ctx->operand_store;
    ctx->operand_store = store;
    const int rc =
        r.synthetic
            ? atrip_hip_upload_synthetic(ctx, r.No, r.Nv, r.seed, 0, &sys)
            : atrip_hip_upload(ctx, &r.desc, &sys);
    ctx->operand_store = store_before;
    if (rc != 0) {
      // out of device memory with other requests resident: wait for them
      if (!srv->active.empty() &&
          ctx->err.find("out of memory") != std::string::npos)
        break;
      fail_request(srv, r, ctx->err);
      srv->queue.pop_front();
      continue;
    }
    srv->queue.pop_front();
    r.sys = sys;
    r.t_admit = srv->now();
    srv->resident += r.bytes;
    srv->peak_resident = std::max(srv->peak_resident, srv->resident);
    r.plain = runs_plain(srv, r.No);
    try {
      if (r.plain) {
        // the request keeps the window edge it was admitted with, whatever
        // the context's option does later
        r.plan.nb = window_edge(ctx, get_sys(ctx, sys));
        r.nwin = plan_windows(ctx, sys, r.plan.nb);
        r.plan.measured = false;
        r.plan.sec_window = plain_window_estimate(r.No, r.Nv, r.plan.nb,
                                                  srv->precision, grid);
      } else {
        r.plan = plan_for(ctx, get_sys(ctx, sys), grid, srv->precision);
        r.wins = windows_of(r.Nv, r.plan.nb);
        r.nwin = r.wins.size();
      }
    } catch (const std::exception &x) {
      // a request whose plan cannot be made fails alone and gives its
      // memory back, rather than staying resident with no windows queued
      drop_system(srv, r);
      fail_request(srv, r, x.what());
      continue;
    } catch (const Fail &x) {
      drop_system(srv, r);
      fail_request(srv, r, x.msg);
      continue;
    }
    r.state = ACTIVE;
    r.cursor = 0;
    if (r.nwin == 0)
      retire(srv, r);
    else
      srv->active.push_back(r.id);
  }
}

// One step: admission, one persistent launch over the chosen windows,
// accumulation and retirement.  Returns the number of requests finished.
int step(atrip_hip_server *srv) {
  atrip_hip_ctx *ctx = srv->ctx;
  check(hipSetDevice(ctx->device), "hipSetDevice");
  const int grid = ctx->batch_grid ? ctx->batch_grid
                                   : default_grid(ctx, srv->precision);
  const size_t done_before = srv->reqs.size() - srv->queue.size() -
                             srv->active.size();
  admit(srv, grid);
  if (srv->active.empty())
    return (int)(srv->reqs.size() - srv->queue.size() - srv->active.size() -
                 done_before);

  // window selection: target workgroup-seconds of this launch
  std::vector<int64_t> order_req = srv->active;
  auto remaining = [&](int64_t id) {
    const Request &r = srv->reqs.at(id);
    return (double)(r.nwin - r.cursor) * r.plan.sec_window;
  };
  if (srv->policy == 0)
    std::stable_sort(order_req.begin(), order_req.end(),
                     [&](int64_t x, int64_t y) { return remaining(x) < remaining(y); });
  const double target = srv->iter_ms * 1e-3 * grid;
  double left = target;
  std::vector<BSys> sys;
  std::vector<BSplit> split;
  std::vector<BTask> tasks;
  std::vector<double> cost;
  std::vector<std::pair<int64_t, size_t>> took;   // batched: request, windows
  std::vector<std::pair<int64_t, size_t>> plain;  // plain path: request, windows
  const double share = target / (double)order_req.size();
  for (int64_t id : order_req) {
    if (srv->policy == 0 && left <= 0 && (!tasks.empty() || !plain.empty()))
      break;
    Request &r = srv->reqs.at(id);
    const double sw = std::max(r.plan.sec_window, 1e-9);
    const double want = srv->policy == 0 ? left : share;
    size_t k = (size_t)std::ceil(std::max(want, 0.0) / sw);
    k = std::max<size_t>(1, std::min(k, r.nwin - r.cursor));
    left -= (double)k * sw;
    if (r.plain) {
      plain.push_back({id, k});
      continue;
    }
    const int si = (int)sys.size();
    DevSystem &ds = get_sys(ctx, r.sys);
    sys.push_back(make_bsys(ctx, ds, r.plan, srv->precision));
    split.push_back(make_bsplit(ds));
    const double nb3 = (double)r.plan.nb * r.plan.nb * r.plan.nb;
    for (size_t q = 0; q < k; q++) {
      const Win &w = r.wins[r.cursor + q];
      tasks.push_back({si, w.al, w.be, w.ga, 0});
      cost.push_back(r.plan.sec_window * (0.5 + 0.5 * w.valid / nb3));
    }
    took.push_back({id, k});
  }
  if (!tasks.empty()) {
    std::vector<int32_t> order(tasks.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t x, int32_t y) { return cost[x] > cost[y]; });
    const LaunchOut out =
        run_tasks(ctx, srv->precision, sys, split, tasks, order, grid, true);
    srv->busy_ms += out.ms;
    srv->launches++;
    // accumulate in canonical window order
    size_t t = 0;
    for (auto &tk : took) {
      Request &r = srv->reqs.at(tk.first);
      for (size_t q = 0; q < tk.second; q++, t++) {
        neumaier_add(r.sum, r.comp, out.task_e[2 * t], out.task_e[2 * t + 1]);
        r.tuples += (uint64_t)r.wins[r.cursor + q].valid;
      }
      r.cursor += tk.second;
    }
  }
  // the plain-path windows, request by request, in plan order
  std::vector<double> wsum;
  for (auto &pk : plain) {
    Request &r = srv->reqs.at(pk.first);
    const Clock::time_point t0 = Clock::now();
    run_plan_windows(ctx, r.sys, srv->precision, r.plan.nb, r.cursor,
                     r.cursor + pk.second,
                     wsum);
    const double ms =
        std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
    if (wsum.size() != 2 * pk.second) throw Fail{"plain-path window count"};
    for (size_t q = 0; q < pk.second; q++)
      neumaier_add(r.sum, r.comp, wsum[2 * q], wsum[2 * q + 1]);
    r.cursor += pk.second;
    srv->busy_ms += ms;
    srv->launches++;
    // measured workgroup-seconds per window from now on
    r.plan.sec_window = ms * 1e-3 / (double)pk.second * grid;
    r.plan.measured = true;
    took.push_back(pk);
  }
  for (auto &tk : took) {
    Request &r = srv->reqs.at(tk.first);
    if (r.cursor < r.nwin) continue;
    if (r.plain) r.tuples = n_tuples((uint64_t)r.Nv);
    srv->tuples_done += r.tuples;
    retire(srv, r);
    srv->active.erase(std::find(srv->active.begin(), srv->active.end(), r.id));
  }
  return (int)(srv->reqs.size() - srv->queue.size() - srv->active.size() -
               done_before);
}

int64_t submit(atrip_hip_server *srv, Request r, double arrival_ms) {
  r.id = srv->next_id++;
  r.t_submit = arrival_ms >= 0 ? arrival_ms : srv->now();
  r.state = QUEUED;
  const int64_t id = r.id;
  srv->reqs.emplace(id, std::move(r));
  srv->queue.push_back(id);
  return id;
}

}  // namespace

extern "C" {

int atrip_hip_serve_create(atrip_hip_ctx *ctx, uint64_t hbm_budget_bytes,
                           int max_active, double iter_ms, int precision,
                           atrip_hip_server **out) {
  return guarded(ctx, [&] {
    if (!out) throw Fail{"null argument"};
    if (precision != ATRIP_FP64 && precision != ATRIP_SPLIT_FP32 &&
        precision != ATRIP_FP32 && precision != ATRIP_BF16X3)
      throw Fail{"unknown precision " + std::to_string(precision)};
    if (hbm_budget_bytes == 0) throw Fail{"HBM budget must be positive"};
    if (max_active < 1) throw Fail{"max_active must be positive"};
    if (!(iter_ms > 0)) throw Fail{"iter_ms must be positive"};
    auto *s = new atrip_hip_server;
    s->ctx = ctx;
    s->budget = hbm_budget_bytes;
    s->max_active = max_active;
    s->iter_ms = iter_ms;
    s->precision = precision;
    *out = s;
  });
}

int atrip_hip_serve_destroy(atrip_hip_server *srv) {
  if (!srv) return 0;
  for (auto &kv : srv->reqs)
    if (kv.second.sys) (void)atrip_hip_release(srv->ctx, kv.second.sys);
  delete srv;
  return 0;
}

int atrip_hip_serve_set_policy(atrip_hip_server *srv, int policy) {
  if (!srv) return 1;
  return guarded(srv->ctx, [&] {
    if (policy != 0 && policy != 1) throw Fail{"policy must be 0 or 1"};
    srv->policy = policy;
  });
}

double atrip_hip_serve_clock_ms(atrip_hip_server *srv) {
  return srv ? srv->now() : 0.0;
}

int atrip_hip_serve_submit_synthetic(atrip_hip_server *srv, int64_t No,
                                     int64_t Nv, uint64_t seed,
                                     double arrival_ms, int64_t *request) {
  if (!srv) return 1;
  return guarded(srv->ctx, [&] {
    if (!request) throw Fail{"null argument"};
    if (No < 1 || Nv < 1) throw Fail{"No and Nv must be positive"};
    Request r;
    r.No = No;
    r.Nv = Nv;
    r.seed = seed;
    r.synthetic = true;
    *request = submit(srv, std::move(r), arrival_ms);
  });
}

int atrip_hip_serve_submit(atrip_hip_server *srv, const atrip_sys_desc *desc,
                           double arrival_ms, int64_t *request) {
  if (!srv) return 1;
  return guarded(srv->ctx, [&] {
    if (!request || !desc) throw Fail{"null argument"};
    if (desc->Jijka || desc->Jabci)
      throw Fail{"served requests carry no (cT) tensors"};
    Request r;
    r.No = desc->No;
    r.Nv = desc->Nv;
    r.synthetic = false;
    r.desc = *desc;
    *request = submit(srv, std::move(r), arrival_ms);
  });
}

int atrip_hip_serve_step(atrip_hip_server *srv, int *finished) {
  if (!srv) return 1;
  return guarded(srv->ctx, [&] {
    const int n = step(srv);
    if (finished) *finished = n;
  });
}

int atrip_hip_serve_result(atrip_hip_server *srv, int64_t request, int *state,
                           double *energy, double *t_submit_ms,
                           double *t_admit_ms, double *t_done_ms) {
  if (!srv) return 1;
  return guarded(srv->ctx, [&] {
    auto it = srv->reqs.find(request);
    if (it == srv->reqs.end()) throw Fail{"unknown request"};
    const Request &r = it->second;
    if (state) *state = r.state;
    if (energy) *energy = r.energy;
    if (t_submit_ms) *t_submit_ms = r.t_submit;
    if (t_admit_ms) *t_admit_ms = r.t_admit;
    if (t_done_ms) *t_done_ms = r.t_done;
    if (r.state == FAILED) throw Fail{"request failed: " + r.err};
  });
}

int atrip_hip_serve_stats(atrip_hip_server *srv, int *queued, int *active,
                          uint64_t *resident_bytes, uint64_t *peak_bytes,
                          uint64_t *tuples_done, uint64_t *launches,
                          double *busy_ms) {
  if (!srv) return 1;
  return guarded(srv->ctx, [&] {
    if (queued) *queued = (int)srv->queue.size();
    if (active) *active = (int)srv->active.size();
    if (resident_bytes) *resident_bytes = srv->resident;
    if (peak_bytes) *peak_bytes = srv->peak_resident;
    if (tuples_done) *tuples_done = srv->tuples_done;
    if (launches) *launches = srv->launches;
    if (busy_ms) *busy_ms = srv->busy_ms;
  });
}

}  // extern "C"
