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
    double rdx, rdy;   // 1 / dx, 1 / dy (IEEE division on the host)
    double gamma;
    double zp_lambda;  // pow(dx, 2/3), used by WENO-ZP in both sweeps
};

struct Bc {
    int left, right, bottom;  // the bottom type also governs the top side
};

enum RkMode { kRkNone = -1, kRk1 = 0, kRkStage = 1 };

// With on set, an x sweep that updates cells also writes the ghost points of
// its output that the next stage's ghost fill would derive from them (bc, and
// the SPECIAL values for the sides of that type): the output leaves the sweep
// ghost-filled. Valid for nx, ny >= 5, where every ghost copies an interior
// point.
struct GhostOut {
    int on = 0;
    Bc bc{};
    const double* special = nullptr;
};

struct RkCoef {
    int mode;
    int gravity;       // RTI source term (g = 1)
    double a, b, cdt;  // U = (a Un + b U) + cdt rhs; cdt = c * dt
    GhostOut gh{};
};

// Tile shapes: faces per block along x (the fused x kernel: 64, 128 or 256)
// and faces per block along y (the y kernel: 16, 32 or 64).
struct Tiles {
    int x = 256, y = 64;
};

constexpr int kProbePhysStride = 82;

// ---- many solves: per-solve tables on the device ------------------------------
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

// ---- launchers -------------------------------------------------------------------
// The FP64 sweeps exist in two arithmetic modes, built from the same source:
//   exact: no FMA contraction; every product and sum is rounded as the
//          reference writes it (bit-identical to the reference compiled
//          without contraction);
//   fused: the compiler contracts a * b + c into one FMA where it finds one
//          (one rounding instead of two, as the reference's own optimised CPU
//          build does with -march=native) and divides by multiplying with the
//          reciprocal (shared between divisions by one value, two roundings);
//          fewer instructions. The WENO weights' explicit shared-reciprocal
//          divisions stay exact in both builds.
enum FpMode { kFpExact = 0, kFpFused = 1 };

namespace exact {
#include "wenofd_sweeps.inc"
}
namespace fused {
#include "wenofd_sweeps.inc"
}

// Launches whose result contraction cannot change (copies, sign flips,
// minima, clock updates, block counts) and the test probes come from the
// exact build. (Not inside wenofd_kernels.hip itself, where argument-dependent
// lookup would see them beside the build's own definitions.)
#ifndef WENOFD_FP_NS
using exact::tiles_valid;
using exact::launch_ghost_fill;
using exact::launch_copy_interior;
using exact::launch_add_source;
using exact::flux_x_rk_blocks;
using exact::hyb_stage_blocks;
using exact::hyb_stage_segments;
using exact::hyb_waves_per_simd;
using exact::flux_y_blocks;
using exact::ghost_blocks;
using exact::launch_dt_final;
using exact::launch_probe_weno;
using exact::launch_probe_physics;
using exact::x_tiles_per_row;
using exact::y_tiles_per_row;
using exact::launch_batch_ghost;
using exact::launch_batch_dt;
using exact::launch_batch_advance;
using exact::launch_batch_copy;
#endif

// The launchers with FP64 arithmetic (launch_flux_y, launch_flux_x_rk,
// launch_dt, launch_batch_flux_y, launch_batch_flux_x_rk,
// launch_batch_dt_partial) are called in a mode: WENOFD_FP(mode, launch_...(...)).
#define WENOFD_FP(mode, call) ((mode) == ::wenofd::kFpFused ? ::wenofd::fused::call : ::wenofd::exact::call)

// ---- learned switch (wenofd_switch.hip) ----------------------------------------
// Field masks (bit k: field k's logit above thr; optionally the logits
// [faces][4]) of every face of direction dir from the ghost-filled state U;
// params = MLP parameters (float, device).
hipError_t launch_classify(int dir, const double* U, const Geom& g, const float* params, float thr,
                           unsigned char* flags, float* logits, hipStream_t s);
// Features [faces][4][10] and labels [faces][4] of every face of direction dir.
hipError_t launch_harvest(int dir, const double* U, const Geom& g, float* X, unsigned char* Y, hipStream_t s);
// Bitwise OR of the field masks over the (2 rad + 1)^2 neighbourhood of each
// face in the rows x cols face grid.
hipError_t launch_dilate(const unsigned char* in, unsigned char* out, int rows, int cols, int rad, hipStream_t s);
hipError_t launch_fill_u8(unsigned char* p, long n, unsigned char v, hipStream_t s);
// The faces of one direction's masks (F faces) that flag any field: list[]
// grouped by the number of flagged fields (1 .. 4), each group in face order;
// *n = the number of faces (device). cnt / off: 4 list_blocks(F) ints of
// scratch each.
long list_blocks(long F);
hipError_t launch_build_list(const unsigned char* masks, long F, int* cnt, int* off, unsigned* list, int* n,
                             hipStream_t s);
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
