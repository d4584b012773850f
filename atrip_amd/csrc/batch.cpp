// Host side of the batched form: the per-shape plan (window edge and tile
// variant of the persistent kernel) from the cost model or from the
// measuring tuner, the window work items of every system, the dispatch
// order, the persistent launch and the per-system reduction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <utility>
#include <vector>

#include "batch.hpp"

using namespace atrip_amd;

namespace atrip_amd {

// Windows (block triples of edge nb) holding at least one tuple, in
// lexicographic order, with their tuple counts.
std::vector<Win> windows_of(int64_t Nv, int nb) {
  if (nb < 1) throw Fail{"window edge must be positive"};
  std::vector<Win> out;
  const int nblk = (int)((Nv + nb - 1) / nb);
  for (int al = 0; al < nblk; al++)
    for (int be = al; be < nblk; be++)
      for (int ga = be; ga < nblk; ga++) {
        int valid = 0;
        for (int ta = 0; ta < nb; ta++)
          for (int tb = 0; tb < nb; tb++)
            for (int tc = 0; tc < nb; tc++)
              valid += tuple_ok((uint64_t)Nv, (uint64_t)al * nb + ta,
                                (uint64_t)be * nb + tb, (uint64_t)ga * nb + tc);
        if (valid) out.push_back({al, be, ga, valid});
      }
  return out;
}

// Number of windows holding a tuple: every block triple except the ones
// made of a single one-element block three times.
static double n_windows(int64_t Nv, int nb) {
  const double nblk = (double)((Nv + nb - 1) / nb);
  const int64_t last = Nv - (int64_t)(nblk - 1) * nb;
  const double all = nblk * (nblk + 1) * (nblk + 2) / 6.0;
  return all - (nb == 1 ? nblk : (last == 1 ? 1.0 : 0.0));
}

// Modelled cycles of one full window on one workgroup of four waves: the
// wave tiles of the three classes (per 16-deep K chunk, 4*fm*fn MFMA of 64
// cycles or the chunk's fm+fn fragment loads, whichever is longer, plus a
// memory latency per K segment) and the energies (one wave per tuple).
double window_cycles(int64_t No, int64_t Nv, int nb, int tile,
                     int precision) {
  const double fm = bfm(tile), fn = bfn(tile);
  // MFMA cycles per 16-deep chunk and 16x16 output fragment: 4 FP64 16x16x4
  // (64 cycles each); 4 FP32 16x16x4 (32 each), plus one bf16 16x16x32 (16)
  // for the split corrections
  const double mf = precision == ATRIP_FP64         ? 4 * 64.0
                    : precision == ATRIP_SPLIT_FP32 ? 4 * 32.0 + 16.0
                                                    : 4 * 32.0;
  // fragment loads per chunk: 32 B/lane FP64, 16 B FP32 (+16 B bf16 pairs)
  const double ld = precision == ATRIP_FP64 ? 180.0
                    : precision == ATRIP_SPLIT_FP32 ? 180.0 : 120.0;
  const double N2 = (double)No * No;
  const double mt = std::ceil(N2 / (16 * fm)), nt = std::ceil(nb * No / (16 * fn));
  const double subs = 3.0 * nb * nb * mt * nt;
  const double chunks = (2.0 * pad16(Nv) + 2.0 * pad16(No)) / 16.0;
  const double per_chunk = std::max(mf * fm * fn, ld * (fm + fn));
  const double sub = 4 * 1200.0 + chunks * per_chunk + 300.0 * fm * fn;
  const double npts = (double)No * (No + 1) * (No + 2) / 6.0;
  const double energy = nb * nb * nb / 4.0 * (std::ceil(npts / 64.0) * 400.0 + 2000.0);
  return subs * sub / 4.0 + energy;
}

bool nb_fits(int64_t No, int nb, int64_t scratch_mb, int grid) {
  const double bytes = (double)nb * nb * nb * No * No * No * 8.0 * grid;
  return nb == 1 || bytes <= (double)scratch_mb * 1048576.0;
}

// Modelled seconds per tuple of the whole system for (nb, tile): partially
// filled windows are charged half their empty slots.
double model_sec_tuple(int64_t No, int64_t Nv, int nb, int tile,
                       int precision) {
  const double nt = (double)n_tuples((uint64_t)Nv);
  const double nwin = n_windows(Nv, nb);
  const double cyc = window_cycles(No, Nv, nb, tile, precision);
  const double total = cyc * (0.5 * nwin + 0.5 * nt / ((double)nb * nb * nb));
  return total / nt / CLOCK_HZ;
}

BatchTune model_pick(int64_t No, int64_t Nv, int64_t scratch_mb, int grid,
                     int precision) {
  BatchTune best;
  // no tuple (Nv = 1): nothing to rank; any valid window edge yields no work
  if (n_tuples((uint64_t)Nv) == 0) return best;
  double best_s = 1e300;
  for (int nb : NB_CAND) {
    if (!nb_fits(No, nb, scratch_mb, grid)) continue;
    for (int t = 0; t < N_BTILES; t++) {
      const double s = model_sec_tuple(No, Nv, nb, t, precision);
      if (s < best_s) {
        best_s = s;
        best.nb = nb;
        best.tile = t;
        best.sec_window = window_cycles(No, Nv, nb, t, precision) / CLOCK_HZ;
      }
    }
  }
  best.measured = false;
  return best;
}

static std::array<int64_t, 3> shape_key(const DevSystem &s, int precision) {
  return {s.No, pad16(s.Nv), precision};
}

void check_batched_precision(int precision) {
  if (precision == ATRIP_BF16X3)
    throw Fail{"the batched form runs fp64, split_fp32 and fp32 (not bf16x3)"};
  if (precision != ATRIP_FP64 && precision != ATRIP_SPLIT_FP32 &&
      precision != ATRIP_FP32)
    throw Fail{"unknown precision " + std::to_string(precision)};
}

int default_grid(atrip_hip_ctx *ctx, int precision) {
  int cus = 0;
  check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                              ctx->device),
        "device attribute");
  return std::max(1, cus) * batched_blocks_per_cu(precision);
}

BatchTune plan_for(atrip_hip_ctx *ctx, const DevSystem &s, int grid,
                   int precision) {
  BatchTune p;
  auto it = ctx->btune.find(shape_key(s, precision));
  if (it != ctx->btune.end() &&
      nb_fits(s.No, it->second.nb, ctx->batch_scratch_mb, grid))
    p = it->second;
  else
    p = model_pick(s.No, s.Nv, ctx->batch_scratch_mb, grid, precision);
  // forced choices (tests, sweeps); the run shrinks the grid if the forced
  // window edge needs more scratch than the bound allows
  if (ctx->batch_nb > 0 && ctx->batch_nb != p.nb) {
    p.nb = ctx->batch_nb;
    p.measured = false;
    p.sec_window = window_cycles(s.No, s.Nv, p.nb, p.tile, precision) / CLOCK_HZ;
  }
  if (ctx->batch_tile >= 0 && ctx->batch_tile != p.tile) {
    p.tile = ctx->batch_tile;
    p.measured = false;
    p.sec_window = window_cycles(s.No, s.Nv, p.nb, p.tile, precision) / CLOCK_HZ;
  }
  return p;
}

BSys make_bsys(atrip_hip_ctx *ctx, DevSystem &s, const BatchTune &p,
               int precision) {
  check_batched_precision(precision);
  check_precision(s, precision);
  if (precision != ATRIP_FP64) ensure_split(ctx, s);
  BSys b;
  b.T3 = s.T3; b.V3 = s.V3; b.H3 = s.H3; b.T2N = s.T2N; b.T2TN = s.T2TN;
  b.VHH = s.VHH; b.Tai = s.Tai; b.ei = s.ei; b.ea = s.ea;
  b.points = s.points;
  b.No = s.No; b.Nv = s.Nv; b.Kv = s.Kv; b.Ko = s.Ko; b.npoints = s.npoints;
  b.nb = p.nb;
  b.tile = p.tile;
  return b;
}

BSplit make_bsplit(const DevSystem &s) {
  BSplit x;
  x.T3h = s.sT3.h; x.V3h = s.sV3.h; x.H3h = s.sH3.h;
  x.T2Nh = s.sT2N.h; x.T2TNh = s.sT2TN.h;
  x.T3l = s.sT3.l; x.V3l = s.sV3.l; x.H3l = s.sH3.l;
  x.T2Nl = s.sT2N.l; x.T2TNl = s.sT2TN.l;
  return x;
}

namespace {
struct DevBuf {
  void *p = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  template <typename T>
  T *put(const std::vector<T> &v, hipStream_t st, const char *what) {
    check(hipMalloc(&p, std::max<size_t>(v.size(), 1) * sizeof(T)), what);
    if (!v.empty())
      check(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T),
                           hipMemcpyHostToDevice, st),
            what);
    return (T *)p;
  }
  template <typename T>
  T *make(size_t n, const char *what) {
    check(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)), what);
    return (T *)p;
  }
};

}  // namespace

// Compensated sum of a system's window energies (sum, comp pairs) in
// canonical window order; the serving loop continues such a sum across its
// iterations, so a system gets the same bits however its windows were split
// over launches.
void neumaier_add(double &sum, double &comp, double s, double c) {
  const double t = sum + s;
  comp += (std::fabs(sum) >= std::fabs(s)) ? (sum - t) + s : (s - t) + sum;
  sum = t;
  comp += c;
}

LaunchOut run_tasks(atrip_hip_ctx *ctx, int precision,
                    const std::vector<BSys> &sys,
                    const std::vector<BSplit> &split,
                    const std::vector<BTask> &tasks,
                    const std::vector<int32_t> &order, int grid, bool timed) {
  LaunchOut out;
  if (tasks.empty()) return out;
  int64_t stride = 1;
  for (const BTask &t : tasks) {
    const BSys &s = sys[t.sys];
    stride = std::max<int64_t>(stride, (int64_t)s.nb * s.nb * s.nb * s.No *
                                           s.No * s.No);
  }
  stride = (stride + 31) / 32 * 32;  // 256-byte aligned workgroup scratch
  const double budget = (double)ctx->batch_scratch_mb * 1048576.0;
  grid = std::min<int64_t>(grid, (int64_t)tasks.size());
  grid = (int)std::max<int64_t>(
      1, std::min<int64_t>(grid, (int64_t)(budget / (stride * 8.0))));
  const size_t need = (size_t)grid * stride * sizeof(double);
  if (ctx->bscratch_bytes < need) {
    if (ctx->bscratch) check(hipFree(ctx->bscratch), "free batch scratch");
    ctx->bscratch = nullptr;
    ctx->bscratch_bytes = 0;
    check(hipMalloc(&ctx->bscratch, need), "alloc batch scratch");
    ctx->bscratch_bytes = need;
  }
  hipStream_t st = ctx->stream;
  DevBuf dsys, dsplit, dtasks, dorder, dtask_e, dcounter;
  BatchedLaunch L;
  L.precision = precision;
  L.sys = dsys.put(sys, st, "upload batch systems");
  L.split = precision == ATRIP_FP64 ? nullptr
                                    : dsplit.put(split, st, "upload batch splits");
  L.tasks = dtasks.put(tasks, st, "upload batch tasks");
  L.order = dorder.put(order, st, "upload batch order");
  L.ntasks = (int)tasks.size();
  L.counter = dcounter.make<unsigned>(1, "alloc counter");
  check(hipMemsetAsync(L.counter, 0, sizeof(unsigned), st), "clear counter");
  L.scratch = ctx->bscratch;
  L.scratch_stride = stride;
  L.task_e = dtask_e.make<double>(2 * tasks.size(), "alloc task energies");
  L.tuple_e = nullptr;
  L.grid = grid;
  struct Events {  // destroyed on every exit, a failed launch included
    hipEvent_t e[2] = {nullptr, nullptr};
    ~Events() {
      for (hipEvent_t x : e)
        if (x) (void)hipEventDestroy(x);
    }
  } ev;
  hipEvent_t &e0 = ev.e[0], &e1 = ev.e[1];
  if (timed) {
    check(hipEventCreate(&e0), "event");
    check(hipEventCreate(&e1), "event");
    check(hipEventRecord(e0, st), "event");
  }
  check(launch_batched(L, st), "batched kernel");
  if (timed) check(hipEventRecord(e1, st), "event");
  out.task_e.resize(2 * tasks.size());
  check(hipMemcpyAsync(out.task_e.data(), L.task_e,
                       out.task_e.size() * sizeof(double),
                       hipMemcpyDeviceToHost, st),
        "download window energies");
  check(hipStreamSynchronize(st), "batched run");
  if (timed) check(hipEventElapsedTime(&out.ms, e0, e1), "elapsed");
  out.grid = grid;
  return out;
}

}  // namespace atrip_amd

extern "C" {

int atrip_hip_batch_model(int64_t No, int64_t Nv, int precision,
                          int64_t scratch_mb, int grid, int *nb, int *fm,
                          int *fn) {
  if (No < 1 || Nv < 1 || grid < 1 || scratch_mb < 1) return 1;
  if (precision != ATRIP_FP64 && precision != ATRIP_SPLIT_FP32 &&
      precision != ATRIP_FP32)
    return 1;
  const BatchTune p = model_pick(No, Nv, scratch_mb, grid, precision);
  if (nb) *nb = p.nb;
  if (fm) *fm = bfm(p.tile);
  if (fn) *fn = bfn(p.tile);
  return 0;
}

int atrip_hip_batch_plan(atrip_hip_ctx *ctx, int sys, int precision, int *nb,
                         int *fm, int *fn, int *measured, double *sec_window) {
  return guarded(ctx, [&] {
    check_batched_precision(precision);
    const DevSystem &s = get_sys(ctx, sys);
    check(hipSetDevice(ctx->device), "hipSetDevice");
    const int grid =
        ctx->batch_grid ? ctx->batch_grid : default_grid(ctx, precision);
    const BatchTune p = plan_for(ctx, s, grid, precision);
    if (nb) *nb = p.nb;
    if (fm) *fm = bfm(p.tile);
    if (fn) *fn = bfn(p.tile);
    if (measured) *measured = p.measured;
    if (sec_window) *sec_window = p.sec_window;
  });
}

int atrip_hip_batch_run(atrip_hip_ctx *ctx, const int *systems, int nsys,
                        int precision, double *energies) {
  return guarded(ctx, [&] {
    if (nsys < 0 || (nsys > 0 && (!systems || !energies)))
      throw Fail{"bad system list"};
    check_batched_precision(precision);
    check(hipSetDevice(ctx->device), "hipSetDevice");
    const int grid =
        ctx->batch_grid ? ctx->batch_grid : default_grid(ctx, precision);
    std::vector<BSys> sys;
    std::vector<BSplit> split;
    std::vector<BTask> tasks;
    std::vector<double> cost;
    std::vector<int64_t> begin{0};
    double flop = 0.0;
    for (int q = 0; q < nsys; q++) {
      DevSystem &s = get_sys(ctx, systems[q]);
      if (s.No > MAX_NO)
        throw Fail{"batched form needs No <= 64 (use atrip_hip_run)"};
      const BatchTune p = plan_for(ctx, s, grid, precision);
      sys.push_back(make_bsys(ctx, s, p, precision));
      split.push_back(make_bsplit(s));
      const int64_t nb3 = (int64_t)p.nb * p.nb * p.nb;
      int64_t off = 0;
      for (const Win &w : windows_of(s.Nv, p.nb)) {
        tasks.push_back({q, w.al, w.be, w.ga, off});
        off += nb3;
        cost.push_back(p.sec_window * (0.5 + 0.5 * w.valid / (double)nb3));
      }
      begin.push_back((int64_t)tasks.size());
      flop += 12.0 * (double)s.No * s.No * s.No * (s.No + s.Nv) *
              (double)n_tuples((uint64_t)s.Nv);
    }
    if (tasks.size() > (size_t)INT32_MAX) throw Fail{"too many windows"};
    std::vector<int32_t> order(tasks.size());
    std::iota(order.begin(), order.end(), 0);
    if (ctx->batch_order == 0)
      std::stable_sort(order.begin(), order.end(),
                       [&](int32_t x, int32_t y) { return cost[x] > cost[y]; });
    LaunchOut r =
        run_tasks(ctx, precision, sys, split, tasks, order, grid, ctx->timing);
    for (int q = 0; q < nsys; q++) {
      double sum = 0.0, comp = 0.0;
      for (int64_t t = begin[q]; t < begin[q + 1]; t++)
        neumaier_add(sum, comp, r.task_e[2 * t], r.task_e[2 * t + 1]);
      // src/atrip/Atrip.cxx:1108-1111: E = -sum
      energies[q] = -(sum + comp);
    }
    if (ctx->timing) {
      KernelStat &k = ctx->stats[precision == ATRIP_FP64 ? "batched"
                                 : precision == ATRIP_SPLIT_FP32
                                     ? "batched_mixed"
                                     : "batched_fp32"];
      k.launches++;
      k.ms += r.ms;
      k.flop += flop;
    }
  });
}

int atrip_hip_batch_tune(atrip_hip_ctx *ctx, const int *systems, int nsys,
                         int precision, int mode) {
  return guarded(ctx, [&] {
    if (nsys < 0 || (nsys > 0 && !systems)) throw Fail{"bad system list"};
    if (mode != 0 && mode != 1) throw Fail{"mode must be 0 or 1"};
    check_batched_precision(precision);
    check(hipSetDevice(ctx->device), "hipSetDevice");
    const int grid =
        ctx->batch_grid ? ctx->batch_grid : default_grid(ctx, precision);
    // one representative system per shape
    std::vector<std::pair<std::array<int64_t, 3>, int>> reps;
    for (int q = 0; q < nsys; q++) {
      const DevSystem &s = get_sys(ctx, systems[q]);
      if (s.No > MAX_NO)
        throw Fail{"batched form needs No <= 64 (use atrip_hip_run)"};
      const auto key = shape_key(s, precision);
      bool seen = false;
      for (auto &r : reps) seen |= r.first == key;
      if (!seen) reps.push_back({key, systems[q]});
    }
    bool warm = false;
    for (auto &rep : reps) {
      DevSystem &s = get_sys(ctx, rep.second);
      if (mode == 0) {
        ctx->btune[rep.first] =
            model_pick(s.No, s.Nv, ctx->batch_scratch_mb, grid, precision);
        continue;
      }
      // candidates: the six best by the model
      std::vector<std::pair<double, std::pair<int, int>>> cand;
      for (int nb : NB_CAND) {
        if (!nb_fits(s.No, nb, ctx->batch_scratch_mb, grid)) continue;
        for (int t = 0; t < N_BTILES; t++)
          cand.push_back(
              {model_sec_tuple(s.No, s.Nv, nb, t, precision), {nb, t}});
      }
      std::stable_sort(cand.begin(), cand.end(),
                       [](const auto &x, const auto &y) { return x.first < y.first; });
      if (cand.size() > 6) cand.resize(6);
      BatchTune best;
      double best_rate = -1.0;
      for (auto &c : cand) {
        BatchTune p;
        p.nb = c.second.first;
        p.tile = c.second.second;
        const std::vector<Win> wins = windows_of(s.Nv, p.nb);
        if (wins.empty()) continue;
        // one window per workgroup, the system's windows taken cyclically
        std::vector<BTask> tasks;
        double tuples = 0.0;
        for (int q = 0; q < grid; q++) {
          const Win &w = wins[q % wins.size()];
          tasks.push_back({0, w.al, w.be, w.ga, 0});
          tuples += w.valid;
        }
        std::vector<int32_t> order(tasks.size());
        std::iota(order.begin(), order.end(), 0);
        const std::vector<BSys> sys{make_bsys(ctx, s, p, precision)};
        const std::vector<BSplit> split{make_bsplit(s)};
        if (!warm) {
          run_tasks(ctx, precision, sys, split, tasks, order, grid, false);
          warm = true;
        }
        const LaunchOut r =
            run_tasks(ctx, precision, sys, split, tasks, order, grid, true);
        const double rate = tuples / std::max(1e-6, (double)r.ms);
        if (rate > best_rate) {
          best_rate = rate;
          best = p;
          best.sec_window = r.ms * 1e-3;
          best.measured = true;
        }
      }
      if (best_rate > 0) ctx->btune[rep.first] = best;
    }
  });
}

}  // extern "C"
