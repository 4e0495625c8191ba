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
    int fx_pitch;      // doubles per row of an x-face plane (>= nx + 1), ny rows
    long fx_plane;
    int fy_pitch;      // doubles per row of a y-face plane (>= nx), ny + 1 rows
    long fy_plane;
    double dx, dy;     // point spacing (xR - xL) / (nx - 1), likewise y
    double gamma;
    double zp_lambda;  // pow(dx, 2/3), used by WENO-ZP in both sweeps
};

struct Bc {
    int left, right, bottom;  // the bottom type also governs the top side
};

enum RkMode { kRkNone = -1, kRk1 = 0, kRkStage = 1, kRkFirst = 2 };

struct RkCoef {
    int mode;
    int gravity;  // RTI source term (g = 1)
    double a, b, cdt;  // U = (a Un + b U) + cdt rhs; cdt = c * dt
};

constexpr int kProbePhysStride = 82;

hipError_t launch_ghost_fill(double* U, const double* special, const Geom& g, const Bc& bc, hipStream_t s);
hipError_t launch_flux_y(int scheme, const double* U, double* Fy, const Geom& g, hipStream_t s);
// x faces + rhs + RK stage (see wenofd_kernels.hip). dt_part (optional) receives
// one partial CFL minimum per block: flux_x_rk_grid(g).x * .y entries.
hipError_t launch_flux_x_rk(int scheme, const double* Uin, const double* Fy, const double* Un, double* Uout,
                            double* rhs_out, const Geom& g, const RkCoef& k, double cfl, double* dt_part, int* dt_nan,
                            hipStream_t s);
dim3 flux_x_rk_grid(const Geom& g);
hipError_t launch_dt_final(const double* part, const int* part_nan, int nparts, double* out, hipStream_t s);
hipError_t launch_dt(const double* U, const Geom& g, double cfl, double* part, int* part_nan, int nblocks,
                     double* out, hipStream_t s);
hipError_t launch_probe_weno(int scheme, const double* v5, int n, double zp_lambda, double* out, hipStream_t s);
hipError_t launch_probe_physics(const double* U, int n, double gamma, double* out, hipStream_t s);

}  // namespace wenofd
