// wenofd_kernels.hpp — launch interface of the WENO-FD kernels (internal to the library).
#pragma once

#include <hip/hip_runtime.h>

namespace wenofd {

enum BcType { kBcPeriod = 0, kBcSlip = 1, kBcSpecial = 2, kBcSymmetric = 3 };  // utils/equations.hpp:45-51

// Geometry of one solve in device memory (see wenofd_kernels.hip for the layout).
struct Geom {
    int nx, ny;        // interior points along x and y
    int pitch;         // doubles per row of a state plane (>= nx + 8)
    long plane;        // doubles per state plane (rows = ny + 8)
    int fy_pitch;      // doubles per row of a y-face plane (>= nx), ny + 1 rows
    long fy_plane;
    double dx, dy;     // point spacing (xR - xL) / (nx - 1), likewise y
    double gamma;
    double zp_lambda;  // pow(dx, 2/3), used by WENO-ZP in both sweeps
};

struct Bc {
    int left, right, bottom;  // the bottom type also governs the top side
};

enum RkMode { kRkNone = -1, kRk1 = 0, kRkStage = 1 };

struct RkCoef {
    int mode;
    int gravity;       // RTI source term (g = 1)
    double a, b, cdt;  // U = (a Un + b U) + cdt rhs; cdt = c * dt
};

// Tile shapes: faces per block along x (the fused x kernel: 64, 128 or 256)
// and faces per block along y (the y kernel: 16, 32 or 64).
struct Tiles {
    int x = 256, y = 64;
};
bool tiles_valid(const Tiles& t);

constexpr int kProbePhysStride = 82;

struct SolveState;

// ---- one solve --------------------------------------------------------------
// st (optional, device): the solve's time state on the device loop; the
// launches do nothing for an inactive solve and the x sweep takes its step
// as k.cdt * st->dt.
hipError_t launch_ghost_fill(double* U, const double* special, const Geom& g, const Bc& bc, hipStream_t s,
                             const SolveState* st = nullptr);
hipError_t launch_copy_interior(const double* from, double* to, const Geom& g, const SolveState* st, hipStream_t s);
// flags (optional): per-face switch flags of the direction (x faces
// [ny][nx+1], y faces [ny+1][nx]); zero selects the upwind-5 flux.
hipError_t launch_flux_y(int scheme, const Tiles& t, const double* U, double* Fy, const Geom& g, hipStream_t s,
                         const unsigned char* flags = nullptr, const SolveState* st = nullptr);
// x faces + rhs + RK stage. dt_part (optional) receives one partial CFL
// minimum per block: flux_x_rk_blocks(g, t) entries.
hipError_t launch_flux_x_rk(int scheme, const Tiles& t, const double* Uin, const double* Fy, const double* Un,
                            double* Uout, double* rhs_out, const Geom& g, const RkCoef& k, double cfl,
                            double* dt_part, int* dt_nan, hipStream_t s, const unsigned char* flags = nullptr,
                            const SolveState* st = nullptr);
long flux_x_rk_blocks(const Geom& g, const Tiles& t);
long flux_y_blocks(const Geom& g, const Tiles& t);
long ghost_blocks(const Geom& g);
hipError_t launch_dt_final(const double* part, const int* part_nan, int nparts, double* out, hipStream_t s);
hipError_t launch_dt(const double* U, const Geom& g, double cfl, double* part, int* part_nan, int nblocks,
                     double* out, hipStream_t s);
hipError_t launch_probe_weno(int scheme, const double* v5, int n, double zp_lambda, double* out, hipStream_t s);
hipError_t launch_probe_physics(const double* U, int n, double gamma, double* out, hipStream_t s);

// ---- many solves, one launch per phase ----------------------------------------
// Per-solve constants (device array, one entry per solve).
struct SolveDesc {
    Geom g;
    Bc bc;
    int gravity;
    double cfl, t_end;
    long long u_off;     // offset (doubles) of the solve's 4 planes in each batch state buffer
    long long fy_off;    // offset of its y-face planes
    long long sp_off;    // offset of its SPECIAL band [4][2 nx + 2 ny]
    long long part_off;  // first dt partial of the solve
    int nparts;          // dt partials its last stage writes
};

// Per-solve mutable state on the device.
struct SolveState {
    double now, dt;
    long long steps;
    int active;  // the current step advances this solve
    int nan;     // a NaN time step was met (the solve stops)
};

// Block -> solve and tile of one launch: tile = the tile's index within the
// solve (its dt partial slot in the x sweep), (bx, by) = its coordinates
// (x sweep: tile in the row, row; y sweep: column tile, face-row tile).
struct TileRef {
    int solve, tile, bx, by;
};
// Tiles per row of the two sweeps (flux_x_rk_blocks = x_tiles_per_row * ny).
int x_tiles_per_row(const Geom& g, const Tiles& t);
int y_tiles_per_row(const Geom& g, const Tiles& t);

struct BatchView {
    const SolveDesc* desc;
    SolveState* st;
    double* U[3];  // state buffers with the same per-solve offsets
    double* Fy;
    const double* special;
    double* part;
    int* part_nan;
    int* active_count;
    int nsolves;
};

hipError_t launch_batch_ghost(const BatchView& b, int buf, const TileRef* tiles, int nblocks, hipStream_t s);
hipError_t launch_batch_flux_y(int scheme, const Tiles& t, const BatchView& b, int buf, const TileRef* tiles,
                               int nblocks, hipStream_t s);
// Stage in -> out with Un = U[0]; each solve's cdt is c_stage * its dt.
hipError_t launch_batch_flux_x_rk(int scheme, const Tiles& t, const BatchView& b, int in, int out, int mode,
                                  double a, double bcoef, double c_stage, bool last, const TileRef* tiles,
                                  int nblocks, hipStream_t s);
hipError_t launch_batch_dt_partial(const BatchView& b, hipStream_t s);
// dt of every running solve from its partials (clipped to t_end - now) and
// the step's activity flags; after the stages, the time advance of the active
// solves and the count of solves still running (count_only: the count alone).
hipError_t launch_batch_dt(const BatchView& b, hipStream_t s);
hipError_t launch_batch_advance(const BatchView& b, bool count_only, hipStream_t s);
// Interior copy U[from] -> U[to] of the active solves (forward-Euler batches).
// The tile table is the one of the fused x kernel for the same tile shape.
hipError_t launch_batch_copy(const Tiles& t, const BatchView& b, int from, int to, const TileRef* tiles, int nblocks,
                             hipStream_t s);

// ---- learned switch (wenofd_switch.hip) ----------------------------------------
// Flags (and optionally logits [faces][4]) of every face of direction dir from
// the ghost-filled state U; params = MLP parameters (float, device).
hipError_t launch_classify(int dir, const double* U, const Geom& g, const float* params, float thr,
                           unsigned char* flags, float* logits, hipStream_t s);
// Features [faces][4][10] and labels [faces][4] of every face of direction dir.
hipError_t launch_harvest(int dir, const double* U, const Geom& g, float* X, unsigned char* Y, hipStream_t s);
// OR over the (2 rad + 1)^2 neighbourhood of each face in the rows x cols face grid.
hipError_t launch_dilate(const unsigned char* in, unsigned char* out, int rows, int cols, int rad, hipStream_t s);
hipError_t launch_fill_u8(unsigned char* p, long n, unsigned char v, hipStream_t s);
hipError_t launch_mlp_infer(const float* P, const float* X, long n, float* z, hipStream_t s);
hipError_t launch_keep_count(const unsigned char* Y, long n, unsigned long long seed, double p_pos, double p_neg,
                             int* counts, hipStream_t s);
hipError_t launch_keep_write(const float* X, const unsigned char* Y, long n, unsigned long long seed, double p_pos,
                             double p_neg, const long long* offsets, float* Xo, unsigned char* Yo, hipStream_t s);
int mlp_grad_blocks(int n);
// Mean loss and gradient over the n samples idx[0..n) of (X, Y);
// partial: mlp_grad_blocks(n) x 1441 floats, lpart: mlp_grad_blocks(n) floats.
hipError_t launch_mlp_grad(const float* P, const float* X, const unsigned char* Y, const int* idx, int n, float pw,
                           float* partial, float* lpart, float* grad, float* loss, hipStream_t s);
hipError_t launch_adam(float* P, const float* grad, float* m, float* v, float lr, float b1, float b2, float eps,
                       float c1, float c2, hipStream_t s);

}  // namespace wenofd
