// Host pieces of the batched form shared by the one-shot batched run and
// the serving loop: window lists, the cost model and plans, and the
// persistent launch over a list of window work items.
#pragma once

#include <cstdint>
#include <utility>
#include <vector>

#include "host.hpp"

namespace atrip_amd {

constexpr int NB_CAND[] = {1, 2, 4, 8};
constexpr int MAX_NO = 64;       // singles tables of the batched kernel
constexpr double CLOCK_HZ = 2.4e9;  // MI355X peak engine clock

struct Win {
  int32_t al, be, ga, valid;
};

// Windows (block triples of edge nb) holding at least one tuple, in
// lexicographic order, with their tuple counts.
std::vector<Win> windows_of(int64_t Nv, int nb);
// Modelled cycles of one full window on one workgroup.
double window_cycles(int64_t No, int64_t Nv, int nb, int tile, int precision);
// Window edge nb fits the scratch bound at this grid.
bool nb_fits(int64_t No, int nb, int64_t scratch_mb, int grid);
// Modelled seconds per tuple of a whole system.
double model_sec_tuple(int64_t No, int64_t Nv, int nb, int tile,
                       int precision);
BatchTune model_pick(int64_t No, int64_t Nv, int64_t scratch_mb, int grid,
                     int precision);
// Persistent grid: CUs x resident workgroups per CU.
int default_grid(atrip_hip_ctx *ctx, int precision);
// Plan of a system: tuned shape, else the model; forced options applied.
BatchTune plan_for(atrip_hip_ctx *ctx, const DevSystem &s, int grid,
                   int precision);
// Descriptors of a system for a launch in `precision` (makes the split
// copies of an FP64-resident system for the mixed forms).
BSys make_bsys(atrip_hip_ctx *ctx, DevSystem &s, const BatchTune &p,
               int precision);
BSplit make_bsplit(const DevSystem &s);

struct LaunchOut {
  std::vector<double> task_e;  // 2 per task: (sum, comp)
  float ms = 0.f;              // kernel time when timed
  int grid = 0;
};

// One persistent launch over `tasks` (descriptors sys[task.sys]) taken in
// dispatch `order`; returns the per-window energies.
LaunchOut run_tasks(atrip_hip_ctx *ctx, int precision,
                    const std::vector<BSys> &sys,
                    const std::vector<BSplit> &split,
                    const std::vector<BTask> &tasks,
                    const std::vector<int32_t> &order, int grid, bool timed);

// Check a precision code of the batched form.
void check_batched_precision(int precision);

// Neumaier step adding a window's (sum, comp) to a running (sum, comp).
void neumaier_add(double &sum, double &comp, double s, double c);

}  // namespace atrip_amd
