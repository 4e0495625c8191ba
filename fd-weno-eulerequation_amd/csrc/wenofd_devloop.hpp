// wenofd_devloop.hpp — device-side run loop shared by the batched solver and
// the single solver's wenofd_advance (internal to the library).
//
// Per step, k_batch_dt folds each running solve's CFL partials into its time
// step on the device (clipped to its t_end), the RK stages run for the active
// solves, and k_batch_advance moves the clocks: no host round trip per step.
// Steps are captured once into a HIP graph of graph_steps steps and replayed.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "wenofd_kernels.hpp"

namespace wenofd {

struct DevLoop {
    int scheme = 0, rkmeth = 2;
    int fp = kFpExact;  // FpMode of the sweeps
    Tiles tiles{};
    TileRef* tab_ghost = nullptr;  // block -> (solve, tile) of the three per-stage launches
    TileRef* tab_y = nullptr;
    TileRef* tab_x = nullptr;
    int n_ghost = 0, n_y = 0, n_x = 0;
    bool ghost_out = false;  // every solve has nx, ny >= 5: RK2 / RK3 stages write their outputs' ghosts
    int graph_steps = 16;  // steps per graph replay; 0: plain launches
    hipGraphExec_t exec = nullptr;
    double* graph_U[3] = {nullptr, nullptr, nullptr};  // state buffers the graph was captured with
};

// Block tables for L.tiles; sets each solve's part_off / nparts (the partials
// its last stage writes) and uploads desc to desc_d. Drops the graph.
int devloop_tables(DevLoop& L, std::vector<SolveDesc>& desc, SolveDesc* desc_d, size_t part_cap);
// CFL minimum of every solve's current state into its partials, and the count
// of running solves (the state a run starts from).
int devloop_prime(const BatchView& v, int fp, hipStream_t s);
// Enqueues n steps (whole graph replays, then plain steps for the remainder).
int devloop_steps(DevLoop& L, const BatchView& v, int n, hipStream_t s);
void devloop_drop_graph(DevLoop& L);
void devloop_free(DevLoop& L);

}  // namespace wenofd
