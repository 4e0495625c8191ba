// wenofd_devloop.hpp — device-side run loop shared by the batched solver and
// the single solver's wenofd_advance (internal to the library).
//
// Per step, k_batch_dt folds each running solve's CFL partials into its time
// step on the device (clipped to its t_end), the RK stages run for the active
// solves, and k_batch_advance moves the clocks: no host round trip per step.
// Steps are captured once into a HIP graph of graph_steps steps and replayed.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <vector>

#include "wenofd_kernels.hpp"

namespace wenofd {

// Everything a captured step sequence holds by value: the buffer and table
// addresses its kernels were launched with, the geometry and coefficients of
// the solve (single-solve loop), the kernel variants (scheme, RK method, FP
// mode, tile shape) and the sequence length, as a list of 64-bit words. A
// graph is replayed only while the key it was captured with equals the
// current one; any difference drops it and the next replay captures afresh.
struct GraphKey {
    std::vector<unsigned long long> w;

    void add(const void* p) { w.push_back((unsigned long long)(uintptr_t)p); }
    void add(long long v) { w.push_back((unsigned long long)v); }
    void add(double v) {
        unsigned long long b;
        std::memcpy(&b, &v, sizeof b);
        w.push_back(b);
    }
    void add(const Geom& g) {
        for (long long v : {(long long)g.nx, (long long)g.ny, (long long)g.pitch, (long long)g.plane,
                            (long long)g.fy_pitch, (long long)g.fy_plane})
            add(v);
        for (double v : {g.dx, g.dy, g.rdx, g.rdy, g.gamma, g.zp_lambda}) add(v);
    }
    void add(const BatchView& v) {
        for (const void* p : {(const void*)v.desc, (const void*)v.st, (const void*)v.U[0], (const void*)v.U[1],
                              (const void*)v.U[2], (const void*)v.Fy, (const void*)v.special, (const void*)v.part,
                              (const void*)v.part_nan, (const void*)v.active_count})
            add(p);
        add((long long)v.nsolves);
    }
    bool operator==(const GraphKey& o) const { return w == o.w; }
};

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
    GraphKey key;  // what exec was captured with (valid while exec is set)
};

// The key of the batch loop's step sequence for view v.
GraphKey devloop_key(const DevLoop& L, const BatchView& v);

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
