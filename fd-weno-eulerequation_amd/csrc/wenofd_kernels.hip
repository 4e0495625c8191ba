// wenofd_kernels.hip — HIP kernels of the WENO-FD hot path for gfx950 (MI355X).
//
// State layout in HBM (one solve): structure of arrays, x fastest,
//     U[v * plane + j * pitch + i],  v in 0..3 (rho, rho u, rho v, E),
//     i in [0, nx+8) along x, j in [0, ny+8) along y, 4 ghost points per side,
// interior i, j in [4, n+4) (the reference's node numbering, src/WENOFD.cpp:36-70).
// y-face fluxes: Fy[v][f-4][i-4] for faces f in [4, ny+4] of interior columns
// (face f sits between points f-1 and f); x-face fluxes live only in LDS.
//
// One RK stage = ghost fill -> y faces -> x faces fused with rhs + RK update.
// All arithmetic is FP64 in the reference's operation order (wenofd_math.hpp).
//
// Every phase is a tile function shared by two kernels: one solve per launch
// (blockIdx = tile), and a batch of independent solves per launch (a table
// maps each block to a solve and a tile of it, so uneven grids share a launch).
//
// The file is compiled twice: -ffp-contract=off -DWENOFD_FP_NS=exact and
// -ffp-contract=fast -freciprocal-math -DWENOFD_FP_NS=fused -DWENOFD_RECIP_DIV=1
// (FpMode in wenofd_kernels.hpp).

#include <hip/hip_runtime.h>

#include "wenofd_index.hpp"
#include "wenofd_kernels.hpp"
#include "wenofd_math.hpp"
#include "wenofd_switch.hpp"

#ifndef WENOFD_FP_NS
#error "build with -DWENOFD_FP_NS=exact (with -ffp-contract=off) or -DWENOFD_FP_NS=fused (with -ffp-contract=fast)"
#endif

namespace wenofd {
namespace WENOFD_FP_NS {

// The build that divides by the reciprocal: its sweeps are capped where noted
// and recompute a face's IEEE fallback afresh (fewer registers there; more in
// the build without contraction). Its y sweep then fits the folded WENO
// normalisation at 3 waves per SIMD (without the fresh fallback it spilled
// 20-36 dwords and ran 58 % slower).
#if WENOFD_RECIP_DIV
constexpr bool kCapRecip = true;
#else
constexpr bool kCapRecip = false;
#endif

// Blocks are ordered over the XCDs by xcd_block (wenofd_index.hpp): every
// XCD works through one contiguous range of logical blocks, so neighbouring
// tiles, which read the same rows, share one L2 (the x sweep: consecutive rows
// share a y-face row; -0.7 % per step; the y sweep's tiles share only halo
// rows and gain nothing).

// ---------------------------------------------------------------------------
// Ghost fill: exchangeGhostCellsValue (one-rank periodic-x ring) + setBoundary,
// src/WENOFD.cpp:850-929 and :998-1159, restricted to the ghost points the
// sweeps read (x ghosts of interior rows, y ghosts of interior columns; corner
// ghosts are never read). Pure copies and sign flips: bitwise.
// ---------------------------------------------------------------------------
constexpr int kGhostThreads = 256;

__device__ __forceinline__ void ghost_point(double* __restrict__ U, const double* __restrict__ special,
                                            const Geom& g, const Bc& bc, int k) {
    if (k >= ghost_count(g)) return;
    const GhostRef r = ghost_map(g, bc, k);
    const size_t dst = (size_t)r.dj * g.pitch + r.di;
    if (r.spec >= 0) {
        const int ns = 2 * (g.nx + g.ny);
#pragma unroll
        for (int q = 0; q < 4; ++q) U[q * g.plane + dst] = special[q * ns + r.spec];
        return;
    }
    const size_t src = (size_t)r.sj * g.pitch + r.si;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double val = U[q * g.plane + src];
        U[q * g.plane + dst] = (q == r.flip) ? -val : val;
    }
}

// ghost_point inverted (ghost_targets): the ghost points whose source is
// interior point (i, j) of U, holding v, get their values from it (copies, or
// the normal momentum flipped at a slip side); a SPECIAL side's four ghosts
// next to (i, j) take their fixed values. With nx, ny >= 5 every ghost of the
// fill has an interior source, so calling this for every interior point
// leaves U ghost-filled exactly as k_ghost_fill would. xs / ys (uniform over
// the caller's block) enable the x-ghost and the y-ghost parts.
__device__ __forceinline__ void ghost_from_point(double* __restrict__ U, const Geom& g, const GhostOut& gh, int i,
                                                 int j, const double v[4], bool xs, bool ys) {
    ghost_targets(
        g, gh.bc, i, j, xs, ys,
        [&](int di, int dj, int flip) {
            const size_t d = (size_t)dj * g.pitch + di;
#pragma unroll
            for (int r = 0; r < 4; ++r) U[r * g.plane + d] = (r == flip) ? -v[r] : v[r];
        },
        [&](int di, int dj, int spec) {
            const size_t d = (size_t)dj * g.pitch + di;
            const int ns = 2 * (g.nx + g.ny);
#pragma unroll
            for (int r = 0; r < 4; ++r) U[r * g.plane + d] = gh.special[r * ns + spec];
        });
}

// st (optional): the solve's device-side time state; an inactive solve skips
// the launch (device loop, see wenofd_devloop.hpp).
__global__ __launch_bounds__(kGhostThreads) void k_ghost_fill(double* __restrict__ U,
                                                              const double* __restrict__ special, Geom g, Bc bc,
                                                              const SolveState* __restrict__ st) {
    if (st && !st->active) return;
    ghost_point(U, special, g, bc, blockIdx.x * kGhostThreads + threadIdx.x);
}

// Buffer store of v at byte offset off of a row of `pitch` doubles starting at
// row_ptr: an offset past the row (e.g. 0x80000000) drops the store (the
// range check of a raw buffer resource; gfx9-family descriptor word 3).
__device__ __forceinline__ void store_lane(double* row_ptr, int pitch, unsigned off, double v) {
    typedef unsigned U2 __attribute__((ext_vector_type(2)));
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(row_ptr, (short)0, pitch * 8, 0x00020000);
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, v), rs, off, 0, 0);
}

// ---------------------------------------------------------------------------
// y faces. One tile = 64 interior columns x YF faces along y, marched in steps
// of 4 face rows (256 threads = 64 columns x 4 rows). The LLF split of the 9
// point rows one step reads lives in an LDS ring; each step adds 4 new rows, so
// every point is split once per tile (plus the 5-row lead-in). The points
// themselves stay in a second ring of 8 rows for the faces' mean states. Loads
// of U are coalesced along x; each point is read from HBM once per tile.
// ---------------------------------------------------------------------------
// (kYR face rows per step, the kRing-row split ring and the kRawRing-row point
// ring: wenofd_index.hpp, where a host check replays their schedule.)

template <int S, int YF>
__device__ __forceinline__ void flux_y_tile(const double* __restrict__ U, double* __restrict__ Fy, const Geom& g,
                                            int bx, int by) {
    static_assert(YF % kYR == 0, "YF must be a multiple of the step height");
    __shared__ double sP[4][kRing][kYC];
    __shared__ double sM[4][kRing][kYC];
    __shared__ double sU[4][kRawRing][kYC];
    const int tx = threadIdx.x % kYC;
    // one wave = one face row of a step: ty (and every ring row index derived
    // from it) is wave-uniform, which readfirstlane makes scalar arithmetic
    const int ty = __builtin_amdgcn_readfirstlane(threadIdx.x / kYC);
    const int i = 4 + bx * kYC + tx;
    const bool col_ok = i < g.nx + 4;
    const int F0 = y_tile_first_face<YF>(by);   // first face of the tile
    const int F1 = y_tile_end_face<YF>(g, by);  // one past the last face
    const int base = F0 - 3;                    // point row held in ring slot 0

    // U of point row prow in this lane's column, clamped into the array (the
    // values of clamped rows and columns are never used)
    const int il = i < g.nx + 7 ? i : g.nx + 7;
    auto fetch_row = [&](int prow, double u[4]) {
        prow = prow < g.ny + 7 ? prow : g.ny + 7;
        const size_t o = (size_t)prow * g.pitch + il;
#pragma unroll
        for (int r = 0; r < 4; ++r) u[r] = U[r * g.plane + o];
    };
    // the split of point row prow (its U in u) into the rings
    auto put_row = [&](int prow, const double u[4]) {
        const int slot = (prow - base) % kRing;
        if (col_ok && prow < g.ny + 8) {
            double fm[4], fp[4];
            flux_split<1>(u, g.gamma, fm, fp);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                sP[r][slot][tx] = fp[r];
                sM[r][slot][tx] = fm[r];
            }
            // not the lead-in's rows F0-3 and F0-2 (no face reads them, and
            // row F0-3 shares its slot with row F0+5, which the first step
            // writes before any barrier)
            if (y_ring_keeps_point(F0, prow)) {
#pragma unroll
                for (int r = 0; r < 4; ++r) sU[r][(prow - base) % kRawRing][tx] = u[r];
            }
        }
    };
    for (int k = ty; k < 5; k += kYR) {  // lead-in: point rows F0-3 .. F0+1
        double u[4];
        fetch_row(base + k, u);
        put_row(base + k, u);
    }
    // Each step's new row is loaded one step ahead, into registers reloaded in
    // place, so its latency hides behind the previous step's faces; the face
    // stores are buffer stores issued by every lane on every step (lanes and
    // rows past the tile are dropped by the range check), so the wait for the
    // prefetched row at the next step leaves them in flight.
    double un[4];
    fetch_row(F0 + 2 + ty, un);
    // four dropped stores, as every step issues after its prefetch: the loop
    // is then entered with the steady state's memory operations in flight
#pragma unroll
    for (int r = 0; r < 4; ++r) store_lane(Fy, g.fy_pitch, 0x80000000u, 0.0);
    const unsigned lo8 = (unsigned)(i - 4) * 8u;
    for (int f = F0; f < F1; f += kYR) {
        put_row(f + 2 + ty, un);  // rows f+2 .. f+5 complete the window [f-3, f+6)
        fetch_row(f + kYR + 2 + ty, un);
        __syncthreads();
        const int fr = f + ty;
        double Fh[4] = {0.0, 0.0, 0.0, 0.0};
        if (col_ok && fr < F1) {
            double Ul[4], Ur[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {  // points fr-1 and fr
                Ul[r] = sU[r][(fr - 1 - base) % kRawRing][tx];
                Ur[r] = sU[r][(fr - base) % kRawRing][tx];
            }
            const int s0 = fr - 3 - base;  // ring position of point row fr-3
            face_flux_dev<1, S, kCapRecip>(
                Ul, Ur, [&](int s, int r, int z) { return sP[r][(s0 + s) % kRing][tx + z]; },
                [&](int s, int r, int z) { return sM[r][(s0 + s) % kRing][tx + z]; }, g.gamma, g.zp_lambda, Fh);
        }
        const int frc = fr < F1 ? fr : F1 - 1;  // a face row of the array for the resource base
        const unsigned off = (col_ok && fr < F1) ? lo8 : 0x80000000u;
#pragma unroll
        for (int r = 0; r < 4; ++r)
            store_lane(Fy + (size_t)r * g.fy_plane + (size_t)(frc - 4) * g.fy_pitch, g.fy_pitch, off, Fh[r]);
        __syncthreads();
    }
}

// The WENO-Z and WENO-ZP y sweeps want a few more registers than 3 waves
// per SIMD allow (Z: 172 VGPRs); capped there (a 2-dword spill each) they
// run faster than at 2 waves (ZP: 2 %); for WENO-ZPI the spill costs more
// than it gains. The build that divides by the reciprocal caps every scheme:
// JS and Z would otherwise take 246 VGPRs (capped, an 8-dword spill), and
// its ZP and ZPI sweeps run faster capped (34- and 38-dword spills) than at
// 2 waves without a spill (step -1.4 % ZP, -0.7 % ZPI).
template <int S>
constexpr int y_min_waves() { return S == kZ || S == kZP || kCapRecip ? 3 : 1; }

template <int S, int YF>
__global__ __launch_bounds__(kYC* kYR, (y_min_waves<S>())) void k_flux_y(const double* __restrict__ U,
                                                                       double* __restrict__ Fy, Geom g,
                                                                       const SolveState* __restrict__ st) {
    if (st && !st->active) return;
    flux_y_tile<S, YF>(U, Fy, g, blockIdx.x, blockIdx.y);
}

// Minimum over the wave's 64 lanes of values that are never NaN, returned to
// every lane: DPP lane exchanges (quad permutations, row mirrors, then the
// row broadcasts of lanes 15 and 31) leave the minimum in lane 63. All 64
// lanes must be active (every caller reduces after its per-thread work has
// reconverged). No LDS round trips; MIN is exact, so the order of the
// exchanges does not matter.
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ double dpp_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = (int)b, hi = (int)(b >> 32);
    const int l2 = __builtin_amdgcn_update_dpp(lo, lo, CTRL, ROWS, 0xF, false);
    const int h2 = __builtin_amdgcn_update_dpp(hi, hi, CTRL, ROWS, 0xF, false);
    return __longlong_as_double(((long long)h2 << 32) | (unsigned int)l2);
}

__device__ __forceinline__ double wave_min(double v) {
    v = fmin(v, dpp_f64<0xB1>(v));        // quad_perm [1, 0, 3, 2]
    v = fmin(v, dpp_f64<0x4E>(v));        // quad_perm [2, 3, 0, 1]
    v = fmin(v, dpp_f64<0x141>(v));       // row_half_mirror
    v = fmin(v, dpp_f64<0x140>(v));       // row_mirror
    v = fmin(v, dpp_f64<0x142, 0xA>(v));  // row_bcast:15 into rows 1 and 3
    v = fmin(v, dpp_f64<0x143, 0xC>(v));  // row_bcast:31 into rows 2 and 3
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, 63), hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// The cell terms with the spacings (cell_rhs4, local_dt_dev: the build that
// divides by the reciprocal multiplies by the host's 1/dx and 1/dy).
__device__ __forceinline__ double local_dt_g(const double U[4], const Geom& g, double cfl) {
    return local_dt_dev(U, g.gamma, cfl, g.dx, g.dy, g.rdx, g.rdy);
}

// ---------------------------------------------------------------------------
// x faces fused with the rhs and the RK stage update.
//
// One tile = XT-5 consecutive faces of one interior row = the XT-6 cells
// between them (consecutive tiles overlap by one face). The LLF split of the
// XT points the faces read is computed once into LDS, one point per thread
// (component-major: a wave's 64 consecutive slots hit 64 distinct banks), then
// each thread builds one face flux into LDS and completes one cell:
//   rhs = x term + y term (faces read from Fy) [+ RTI source], then
//   kRkNone:  rhs only (rhs_out)           kRk1:   out = in + dt rhs
//   kRkStage: out = (a Un + b in) + (c dt) rhs
// `out` is never `in` (neighbouring tiles still read `in`); it may be `Un`,
// which is read and written only at the thread's own cell. With part set, the
// tile also reduces the local CFL step of its updated cells (the next
// calculateDeltaT) into *part / *part_nan.
// ---------------------------------------------------------------------------
// Cells per x tile of XT threads (x_cells, wenofd_index.hpp): XT - 5 faces
// read exactly XT points, so every thread splits one point.

template <int S, int XT>
__device__ __forceinline__ void flux_x_rk_tile(const double* __restrict__ Uin, const double* __restrict__ Fy,
                                               const double* Un, double* Uout, double* __restrict__ rhs_out,
                                               const Geom& g, const RkCoef& k, double cfl, double* part,
                                               int* part_nan, int bx, int row) {
    constexpr int XC = x_cells<XT>();  // cells per tile
    constexpr int XP = XC + 6;         // points per tile (= XT)
    constexpr int XS = XT + 8;  // LDS row stride (slots)
    constexpr int W = XT / 64;  // waves per tile
    __shared__ double sP[4][XS];
    __shared__ double sM[4][XS];
    __shared__ double sU[4][XS];
    __shared__ double sF[4][XT];
    __shared__ double sMin[W];
    __shared__ int sNan[W];
    const int t = threadIdx.x;
    const int j = row + 4;
    const int f0 = x_tile_first_face<XT>(bx);  // first face = first cell of the tile
    const double* rowp = Uin + (size_t)j * g.pitch;
    // The tile's loads and split at raised wave priority, so that a starting
    // block issues its HBM requests ahead of the face work of the blocks
    // around it (-4 % sweep time); the faces run at the normal priority.
    __builtin_amdgcn_s_setprio(2);
    for (int q = t; q < XP; q += XT) {
        const int p = f0 - 3 + q;  // point p sits in slot p - f0 + 3
        if (p < g.nx + 8) {
            double u[4], fm[4], fp[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) u[r] = rowp[r * g.plane + p];
            flux_split<0>(u, g.gamma, fm, fp);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                sP[r][q] = fp[r];
                sM[r][q] = fm[r];
                sU[r][q] = u[r];
            }
        }
    }
    __syncthreads();
    __builtin_amdgcn_s_setprio(0);
    const int f = f0 + t;
    auto full_face = [&](int tt) {
        double Ul[4], Ur[4], Fh[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            Ul[r] = sU[r][tt + 2];
            Ur[r] = sU[r][tt + 3];
        }
        face_flux_dev<0, S, kCapRecip>(
            Ul, Ur, [&](int s, int r, int z) { return sP[r][tt + s + z]; },
            [&](int s, int r, int z) { return sM[r][tt + s + z]; }, g.gamma, g.zp_lambda, Fh);
#pragma unroll
        for (int r = 0; r < 4; ++r) sF[r][tt] = Fh[r];
    };
    if (t <= XC && f <= g.nx + 4) full_face(t);
    const int i = f0 + t;  // this thread's cell
    const bool cell_ok = t < XC && i < g.nx + 4;
    // ghost points of Uout to write: only the waves whose cells reach within 5
    // of the x edges, and rows within 5 of the y edges, have any (wave-uniform)
    const int wf = __builtin_amdgcn_readfirstlane(f0 + (t & ~63));  // the wave's first cell
    const bool gh_x = k.gh.on && ghost_x_edge(g, wf, wf + 63);
    const bool gh_y = k.gh.on && ghost_y_edge(g, j);
    const size_t o = (size_t)j * g.pitch + i;
    const size_t oy = (size_t)(j - 4) * g.fy_pitch + (i - 4);
    double fyb[4], fyt[4], unb[4];
    if (cell_ok) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            fyb[r] = Fy[r * g.fy_plane + oy];
            fyt[r] = Fy[r * g.fy_plane + oy + g.fy_pitch];
            unb[r] = (k.mode == kRkStage) ? Un[r * g.plane + o] : 0.0;
        }
    }
    __syncthreads();
    double my_min = 1.0;
    int my_nan = 0;
    if (cell_ok) {
        double rhs[4], u[4], fxl[4], fxr[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            fxl[r] = sF[r][t];
            fxr[r] = sF[r][t + 1];
            u[r] = sU[r][t + 3];
        }
        cell_rhs4(fxl, fxr, fyb, fyt, g.dx, g.dy, g.rdx, g.rdy, rhs);
        if (k.gravity) {  // src/WENOFD.cpp:812-813
            rhs[3] = rhs[3] + u[2];
            rhs[2] = rhs[2] + u[0];
        }
        if (rhs_out) {
#pragma unroll
            for (int r = 0; r < 4; ++r) rhs_out[r * g.plane + o] = rhs[r];
        }
        if (k.mode != kRkNone) {
            double nu[4];
#pragma unroll
            for (int r = 0; r < 4; ++r)
                nu[r] = (k.mode == kRk1) ? u[r] + k.cdt * rhs[r] : rk_combine(k.a, unb[r], k.b, u[r], k.cdt, rhs[r]);
#pragma unroll
            for (int r = 0; r < 4; ++r) Uout[r * g.plane + o] = nu[r];
            if (gh_x || gh_y) ghost_from_point(Uout, g, k.gh, i, j, nu, gh_x, gh_y);
            if (part) {
                const double tl = local_dt_g(nu, g, cfl);
                if (tl != tl) my_nan = 1;
                my_min = (tl < my_min) ? tl : my_min;
            }
        }
    }
    if (part) {
        my_min = wave_min(my_min);
        my_nan = __any(my_nan) ? 1 : 0;
        if (t % 64 == 0) {
            sMin[t / 64] = my_min;
            sNan[t / 64] = my_nan;
        }
        __syncthreads();
        if (t == 0) {
            for (int q = 1; q < W; ++q) {
                my_min = (sMin[q] < my_min) ? sMin[q] : my_min;
                my_nan |= sNan[q];
            }
            *part = my_min;
            *part_nan = my_nan;
        }
    }
}

// The contracted WENO-ZPI x sweep would take 170 VGPRs and run at 2 waves per
// SIMD; capped at 3 it fits without spilling (-10 % sweep time). The other
// sweeps of the build without contraction fit by themselves (their code is
// left to the default heuristics). The build that divides by the reciprocal
// caps its x sweeps at 4 waves per SIMD: they spill 33-39 dwords there (at
// 3 waves none) and still run faster (x sweep -7 %, step -1.5 % (ZPI) to
// -3.5 % (JS) at 2048^2).
template <int S, int XT>
constexpr int x_min_waves() { return kCapRecip ? 4 : S == kZPI ? 3 : 1; }

// With st, k.cdt holds the stage coefficient c and the step is c * st->dt.
template <int S, int XT>
__global__ __launch_bounds__(XT, (x_min_waves<S, XT>())) void k_flux_x_rk(const double* __restrict__ Uin,
                                                                          const double* __restrict__ Fy,
                                                  const double* Un, double* Uout, double* __restrict__ rhs_out,
                                                  Geom g, RkCoef k, double cfl, double* __restrict__ dt_part,
                                                  int* __restrict__ dt_nan, const SolveState* __restrict__ st) {
    if (st) {
        if (!st->active) return;
        k.cdt = k.cdt * st->dt;
    }
    const int b = blockIdx.y * gridDim.x + blockIdx.x;
    const int lb = xcd_block(b, gridDim.x * gridDim.y);  // consecutive rows of tiles on one XCD
    flux_x_rk_tile<S, XT>(Uin, Fy, Un, Uout, rhs_out, g, k, cfl, dt_part ? dt_part + b : nullptr,
                          dt_part ? dt_nan + b : nullptr, lb % gridDim.x, lb / gridDim.x);
}

// ---------------------------------------------------------------------------
// calculateDeltaT, src/WENOFD.cpp:398-452 with utils/equations.hpp:137-147:
// per interior point cfl / ((|u|+c)/dx + (|v|+c)/dy), minimum over the domain.
// Partial minima per block (wavefront shuffles, then LDS), folded by one block.
// MIN is exact, so the result does not depend on the order. A NaN anywhere
// raises the flag (the reference aborts on a NaN step).
// ---------------------------------------------------------------------------
template <int NW>
__device__ __forceinline__ void min_nan_reduce_block(double& m, int& nan, double* sm, int* sn) {
    m = wave_min(m);
    nan = __any(nan) ? 1 : 0;
    const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
    if (lane == 0) {
        sm[w] = m;
        sn[w] = nan;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < NW; ++q) {
            m = (sm[q] < m) ? sm[q] : m;
            nan |= sn[q];
        }
    }
}

__device__ __forceinline__ void dt_point(const double* __restrict__ U, const Geom& g, double cfl, long k, double& m,
                                         int& nan) {
    const int jl = (int)(k / g.nx), il = (int)(k % g.nx);
    const size_t o = (size_t)(jl + 4) * g.pitch + (il + 4);
    const double u[4] = {U[o], U[g.plane + o], U[2 * g.plane + o], U[3 * g.plane + o]};
    const double tmp = local_dt_g(u, g, cfl);
    if (tmp != tmp) nan = 1;
    m = (tmp < m) ? tmp : m;
}

__global__ __launch_bounds__(256) void k_dt_partial(const double* __restrict__ U, Geom g, double cfl,
                                                    double* __restrict__ part, int* __restrict__ part_nan) {
    __shared__ double sm[4];
    __shared__ int sn[4];
    double m = 1.0;
    int nan = 0;
    const long total = (long)g.nx * g.ny;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (long)gridDim.x * blockDim.x)
        dt_point(U, g, cfl, k, m, nan);
    min_nan_reduce_block<4>(m, nan, sm, sn);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = m;
        part_nan[blockIdx.x] = nan;
    }
}

// One block folds up to hundreds of thousands of partials: each thread keeps
// kDtFinalLoads independent loads in flight per round.
constexpr int kDtFinalThreads = 1024;
constexpr int kDtFinalLoads = 8;

__device__ __forceinline__ void fold_partials(const double* __restrict__ part, const int* __restrict__ part_nan,
                                              long n, double& m, int& nan) {
    __shared__ double sm[kDtFinalThreads / 64];
    __shared__ int sn[kDtFinalThreads / 64];
    m = 1.0;
    nan = 0;
    for (long k0 = threadIdx.x; k0 < n; k0 += (long)kDtFinalThreads * kDtFinalLoads) {
        double v[kDtFinalLoads];
        int f[kDtFinalLoads];
#pragma unroll
        for (int q = 0; q < kDtFinalLoads; ++q) {
            const long k = k0 + (long)q * kDtFinalThreads;
            v[q] = k < n ? part[k] : 1.0;
            f[q] = k < n ? part_nan[k] : 0;
        }
#pragma unroll
        for (int q = 0; q < kDtFinalLoads; ++q) {
            m = (v[q] < m) ? v[q] : m;
            nan |= f[q];
        }
    }
    min_nan_reduce_block<kDtFinalThreads / 64>(m, nan, sm, sn);
}

__global__ __launch_bounds__(kDtFinalThreads) void k_dt_final(const double* __restrict__ part,
                                                              const int* __restrict__ part_nan, int n,
                                                              double* __restrict__ out) {
    double m;
    int nan;
    fold_partials(part, part_nan, n, m, nan);
    if (threadIdx.x == 0) {
        out[0] = m;
        out[1] = nan ? 1.0 : 0.0;
    }
}

// ---------------------------------------------------------------------------
// Batched phases. Block b works on tile tiles[b].tile of solve tiles[b].solve;
// solves that are not active in this step return at once.
// ---------------------------------------------------------------------------
// The block's table entry, then its solve's descriptor and time state by
// value, issued together behind the one entry load. False for a solve that is
// no longer active.
// idx: the table entry (default: the block id).
__device__ __forceinline__ bool batch_block(const BatchView& b, const TileRef* __restrict__ tiles, TileRef& tr,
                                            SolveDesc& d, SolveState& st, int idx = -1) {
    tr = tiles[idx < 0 ? (int)blockIdx.x : idx];
    d = b.desc[tr.solve];
    st = b.st[tr.solve];
    return st.active;
}

__global__ __launch_bounds__(kGhostThreads) void k_batch_ghost(BatchView b, int buf, const TileRef* __restrict__ tiles) {
    TileRef tr;
    SolveDesc d;
    SolveState st;
    if (!batch_block(b, tiles, tr, d, st)) return;
    ghost_point(b.U[buf] + d.u_off, b.special + d.sp_off, d.g, d.bc, tr.tile * kGhostThreads + threadIdx.x);
}

template <int S, int YF>
__global__ __launch_bounds__(kYC* kYR, (y_min_waves<S>())) void k_batch_flux_y(BatchView b, int buf,
                                                                              const TileRef* __restrict__ tiles) {
    TileRef tr;
    SolveDesc d;
    SolveState st;
    if (!batch_block(b, tiles, tr, d, st)) return;
    flux_y_tile<S, YF>(b.U[buf] + d.u_off, b.Fy + d.fy_off, d.g, tr.bx, tr.by);
}

template <int S, int XT>
__global__ __launch_bounds__(XT, (x_min_waves<S, XT>())) void k_batch_flux_x_rk(BatchView b, int in, int out,
                                                                                        int mode, double a,
                                                        double bcoef, double c_stage, int last, int ghosts,
                                                        const TileRef* __restrict__ tiles) {
    TileRef tr;
    SolveDesc d;
    SolveState st;
    // table entries in XCD-contiguous ranges (rows of one solve meet in one L2)
    if (!batch_block(b, tiles, tr, d, st, xcd_block(blockIdx.x, gridDim.x))) return;
    RkCoef k{mode, d.gravity, a, bcoef, c_stage * st.dt};
    if (ghosts) k.gh = GhostOut{1, d.bc, b.special + d.sp_off};
    double* part = last ? b.part + d.part_off + tr.tile : nullptr;
    int* pnan = last ? b.part_nan + d.part_off + tr.tile : nullptr;
    flux_x_rk_tile<S, XT>(b.U[in] + d.u_off, b.Fy + d.fy_off, b.U[0] + d.u_off, b.U[out] + d.u_off, nullptr, d.g, k,
                          d.cfl, part, pnan, tr.bx, tr.by);
}

// First step: each solve's CFL minima straight from its state. Block
// (bx, solve) of gridDim.x reduces the cells bx, bx + gridDim.x, ... (in units
// of 256) into partial slot bx when bx < nparts, and sets the slots
// >= gridDim.x that fall to it (q = bx mod gridDim.x) neutral.
constexpr int kPrimeBlocks = 256;

__global__ __launch_bounds__(256) void k_batch_dt_partial(BatchView b) {
    __shared__ double sm[4];
    __shared__ int sn[4];
    const SolveDesc& d = b.desc[blockIdx.y];
    const int nb = min((int)gridDim.x, d.nparts);
    if ((int)blockIdx.x >= nb) return;
    const double* U = b.U[0] + d.u_off;
    double m = 1.0;
    int nan = 0;
    const long total = (long)d.g.nx * d.g.ny;
    for (long k = (long)blockIdx.x * 256 + threadIdx.x; k < total; k += (long)nb * 256)
        dt_point(U, d.g, d.cfl, k, m, nan);
    min_nan_reduce_block<4>(m, nan, sm, sn);
    for (int q = nb + blockIdx.x + threadIdx.x * nb; q < d.nparts; q += nb * 256) {
        b.part[d.part_off + q] = 1.0;
        b.part_nan[d.part_off + q] = 0;
    }
    if (threadIdx.x == 0) {
        b.part[d.part_off + blockIdx.x] = m;
        b.part_nan[d.part_off + blockIdx.x] = nan;
    }
}

// Per solve (one block each): the loop test |now - t_end| > 1e-9
// (src/WENOFD.cpp:213), then dt from the partials, clipped to t_end - now
// (:434-435); a NaN step stops the solve.
__global__ __launch_bounds__(kDtFinalThreads) void k_batch_dt(BatchView b) {
    SolveState& st = b.st[blockIdx.x];
    const SolveDesc& d = b.desc[blockIdx.x];
    const bool run = !st.nan && fabs(st.now - d.t_end) > 1e-9;
    if (!run) {
        if (threadIdx.x == 0) {
            st.active = 0;
            st.dt = 0.0;
        }
        return;
    }
    double m;
    int nan;
    fold_partials(b.part + d.part_off, b.part_nan + d.part_off, d.nparts, m, nan);
    if (threadIdx.x == 0) {
        if (nan) {
            st.nan = 1;
            st.active = 0;
            st.dt = 0.0;
        } else {
            double ts = m;
            if (d.t_end - st.now < ts) ts = d.t_end - st.now;
            st.dt = ts;
            st.active = 1;
        }
    }
}

// now += dt for the solves this step advanced (unless count_only); then the
// number of solves the next step will advance.
__global__ __launch_bounds__(256) void k_batch_advance(BatchView b, int count_only) {
    __shared__ int sc[4];
    int cnt = 0;
    for (int s = threadIdx.x; s < b.nsolves; s += 256) {
        SolveState& st = b.st[s];
        if (st.active && !count_only) {
            st.now += st.dt;
            st.steps += 1;
        }
        if (!st.nan && fabs(st.now - b.desc[s].t_end) > 1e-9) ++cnt;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
    if (threadIdx.x % 64 == 0) sc[threadIdx.x / 64] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) *b.active_count = sc[0] + sc[1] + sc[2] + sc[3];
}

template <int XT>
__global__ __launch_bounds__(XT) void k_batch_copy(BatchView b, int from, int to, const TileRef* __restrict__ tiles) {
    TileRef tr;
    SolveDesc d;
    SolveState st;
    if (!batch_block(b, tiles, tr, d, st)) return;
    const int i = 4 + tr.bx * x_cells<XT>() + threadIdx.x;
    const int j = 4 + tr.by;
    if (threadIdx.x >= x_cells<XT>() || i >= d.g.nx + 4) return;
    const size_t o = d.u_off + (size_t)j * d.g.pitch + i;
#pragma unroll
    for (int r = 0; r < 4; ++r) b.U[to][o + r * d.g.plane] = b.U[from][o + r * d.g.plane];
}

// assembleSourceTerm, src/WENOFD.cpp:800-821 (RTI, g = 1): rhs_E += rho v and
// rhs_{rho v} += rho at every interior point.
__global__ __launch_bounds__(256) void k_add_source(const double* __restrict__ U, double* __restrict__ rhs, Geom g) {
    const int i = 4 + blockIdx.x * 256 + threadIdx.x, j = 4 + blockIdx.y;
    if (i >= g.nx + 4) return;
    const size_t o = (size_t)j * g.pitch + i;
    rhs[3 * g.plane + o] = rhs[3 * g.plane + o] + U[2 * g.plane + o];
    rhs[2 * g.plane + o] = rhs[2 * g.plane + o] + U[o];
}

// Interior copy from -> to of one solve (forward Euler on the device loop).
__global__ __launch_bounds__(256) void k_copy_interior(const double* __restrict__ from, double* __restrict__ to, Geom g,
                                                       const SolveState* __restrict__ st) {
    if (st && !st->active) return;
    const int i = 4 + blockIdx.x * 256 + threadIdx.x, j = 4 + blockIdx.y;
    if (i >= g.nx + 4) return;
    const size_t o = (size_t)j * g.pitch + i;
#pragma unroll
    for (int r = 0; r < 4; ++r) to[r * g.plane + o] = from[r * g.plane + o];
}

// ---------------------------------------------------------------------------
// Probes: the same device functions over independent inputs (used by the
// per-function parity tests through the C ABI).
// ---------------------------------------------------------------------------
template <int S>
__global__ void k_probe_weno(const double* __restrict__ v5, int n, double zp_lambda, double* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double* v = v5 + 5 * (size_t)k;
    bool ok = true;
    const double fast = weno5_shared_rcp<S>(v[0], v[1], v[2], v[3], v[4], zp_lambda, ok);
    out[k] = ok ? fast : weno5<S>(v[0], v[1], v[2], v[3], v[4], zp_lambda);
}

// per state: Lx(16) Rx(16) Ly(16) Ry(16) Fm_x Fp_x Fm_y Fp_y (4 each) maxeig_x maxeig_y
__global__ void k_probe_physics(const double* __restrict__ Uin, int n, double gamma, double* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double U[4], L[16], R[16], fm[4], fp[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) U[r] = Uin[4 * (size_t)k + r];
    double* o = out + (size_t)k * kProbePhysStride;
    eigen_matrices<0>(U, gamma, L, R);
    for (int q = 0; q < 16; ++q) { o[q] = L[q]; o[16 + q] = R[q]; }
    eigen_matrices<1>(U, gamma, L, R);
    for (int q = 0; q < 16; ++q) { o[32 + q] = L[q]; o[48 + q] = R[q]; }
    flux_split<0>(U, gamma, fm, fp);
    for (int r = 0; r < 4; ++r) { o[64 + r] = fm[r]; o[68 + r] = fp[r]; }
    flux_split<1>(U, gamma, fm, fp);
    for (int r = 0; r < 4; ++r) { o[72 + r] = fm[r]; o[76 + r] = fp[r]; }
    const Prim<double> q = primitive(U[0], U[1], U[2], U[3], gamma);
    o[80] = fabs(q.u) + q.c;
    o[81] = fabs(q.v) + q.c;
}

// ---------------------------------------------------------------------------
// Launch wrappers: scheme and tile shape become template arguments here.
// ---------------------------------------------------------------------------
bool tiles_valid(const Tiles& t) {
    return (t.x == 64 || t.x == 128 || t.x == 256) && (t.y == 16 || t.y == 32 || t.y == 64);
}

template <class F>
static hipError_t with_scheme(int scheme, F&& f) {
    switch (scheme) {
        case kJS: f(std::integral_constant<int, kJS>()); break;
        case kZ: f(std::integral_constant<int, kZ>()); break;
        case kZP: f(std::integral_constant<int, kZP>()); break;
        default: f(std::integral_constant<int, kZPI>()); break;
    }
    return hipGetLastError();
}

template <class F>
static void with_x_tile(int xt, F&& f) {
    if (xt == 64) f(std::integral_constant<int, 64>());
    else if (xt == 128) f(std::integral_constant<int, 128>());
    else f(std::integral_constant<int, 256>());
}

template <class F>
static void with_y_tile(int yf, F&& f) {
    if (yf == 16) f(std::integral_constant<int, 16>());
    else if (yf == 32) f(std::integral_constant<int, 32>());
    else f(std::integral_constant<int, 64>());
}

long ghost_blocks(const Geom& g) { return (8L * (g.nx + g.ny) + kGhostThreads - 1) / kGhostThreads; }

long flux_y_blocks(const Geom& g, const Tiles& t) {
    long n = 0;
    with_y_tile(t.y, [&](auto YF) { n = (long)y_tiles_x<YF()>(g) * y_tiles_y<YF()>(g); });
    return n;
}

int x_tiles_per_row(const Geom& g, const Tiles& t) {
    int n = 0;
    with_x_tile(t.x, [&](auto XT) { n = x_tiles_x<XT()>(g); });
    return n;
}

int y_tiles_per_row(const Geom& g, const Tiles& t) {
    int n = 0;
    with_y_tile(t.y, [&](auto YF) { n = y_tiles_x<YF()>(g); });
    return n;
}

long flux_x_rk_blocks(const Geom& g, const Tiles& t) {
    long n = 0;
    with_x_tile(t.x, [&](auto XT) { n = (long)x_tiles_x<XT()>(g) * g.ny; });
    return n;
}

hipError_t launch_ghost_fill(double* U, const double* special, const Geom& g, const Bc& bc, hipStream_t s,
                             const SolveState* st) {
    hipLaunchKernelGGL(k_ghost_fill, dim3(ghost_blocks(g)), dim3(kGhostThreads), 0, s, U, special, g, bc, st);
    return hipGetLastError();
}

hipError_t launch_copy_interior(const double* from, double* to, const Geom& g, const SolveState* st, hipStream_t s) {
    hipLaunchKernelGGL(k_copy_interior, dim3((g.nx + 255) / 256, g.ny), dim3(256), 0, s, from, to, g, st);
    return hipGetLastError();
}

hipError_t launch_add_source(const double* U, double* rhs, const Geom& g, hipStream_t s) {
    hipLaunchKernelGGL(k_add_source, dim3((g.nx + 255) / 256, g.ny), dim3(256), 0, s, U, rhs, g);
    return hipGetLastError();
}

hipError_t launch_flux_y(int scheme, const Tiles& t, const double* U, double* Fy, const Geom& g, hipStream_t s,
                         const SolveState* st) {
    return with_scheme(scheme, [&](auto S) {
        with_y_tile(t.y, [&](auto YF) {
            dim3 grid(y_tiles_x<YF()>(g), y_tiles_y<YF()>(g));
            hipLaunchKernelGGL((k_flux_y<S(), YF()>), grid, dim3(kYC * kYR), 0, s, U, Fy, g, st);
        });
    });
}

hipError_t launch_flux_x_rk(int scheme, const Tiles& t, const double* Uin, const double* Fy, const double* Un,
                            double* Uout, double* rhs_out, const Geom& g, const RkCoef& k, double cfl,
                            double* dt_part, int* dt_nan, hipStream_t s, const SolveState* st) {
    return with_scheme(scheme, [&](auto S) {
        with_x_tile(t.x, [&](auto XT) {
            dim3 grid(x_tiles_x<XT()>(g), g.ny);
            hipLaunchKernelGGL((k_flux_x_rk<S(), XT()>), grid, dim3(XT()), 0, s, Uin, Fy, Un, Uout, rhs_out, g, k,
                               cfl, dt_part, dt_nan, st);
        });
    });
}

hipError_t launch_dt_final(const double* part, const int* part_nan, int nparts, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(kDtFinalThreads), 0, s, part, part_nan, nparts, out);
    return hipGetLastError();
}

hipError_t launch_dt(const double* U, const Geom& g, double cfl, double* part, int* part_nan, int nblocks,
                     double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_dt_partial, dim3(nblocks), dim3(256), 0, s, U, g, cfl, part, part_nan);
    hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(kDtFinalThreads), 0, s, part, part_nan, nblocks, out);
    return hipGetLastError();
}

hipError_t launch_probe_weno(int scheme, const double* v5, int n, double zp_lambda, double* out, hipStream_t s) {
    return with_scheme(scheme, [&](auto S) {
        hipLaunchKernelGGL(k_probe_weno<S()>, dim3((n + 255) / 256), dim3(256), 0, s, v5, n, zp_lambda, out);
    });
}

hipError_t launch_probe_physics(const double* U, int n, double gamma, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_probe_physics, dim3((n + 127) / 128), dim3(128), 0, s, U, n, gamma, out);
    return hipGetLastError();
}

hipError_t launch_batch_ghost(const BatchView& b, int buf, const TileRef* tiles, int nblocks, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_ghost, dim3(nblocks), dim3(kGhostThreads), 0, s, b, buf, tiles);
    return hipGetLastError();
}

hipError_t launch_batch_flux_y(int scheme, const Tiles& t, const BatchView& b, int buf, const TileRef* tiles,
                               int nblocks, hipStream_t s) {
    return with_scheme(scheme, [&](auto S) {
        with_y_tile(t.y, [&](auto YF) {
            hipLaunchKernelGGL((k_batch_flux_y<S(), YF()>), dim3(nblocks), dim3(kYC * kYR), 0, s, b, buf, tiles);
        });
    });
}

hipError_t launch_batch_flux_x_rk(int scheme, const Tiles& t, const BatchView& b, int in, int out, int mode,
                                  double a, double bcoef, double c_stage, bool last, bool ghosts,
                                  const TileRef* tiles, int nblocks, hipStream_t s) {
    return with_scheme(scheme, [&](auto S) {
        with_x_tile(t.x, [&](auto XT) {
            hipLaunchKernelGGL((k_batch_flux_x_rk<S(), XT()>), dim3(nblocks), dim3(XT()), 0, s, b, in, out, mode, a,
                               bcoef, c_stage, last ? 1 : 0, ghosts ? 1 : 0, tiles);
        });
    });
}

hipError_t launch_batch_dt_partial(const BatchView& b, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_dt_partial, dim3(kPrimeBlocks, b.nsolves), dim3(256), 0, s, b);
    return hipGetLastError();
}

hipError_t launch_batch_dt(const BatchView& b, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_dt, dim3(b.nsolves), dim3(kDtFinalThreads), 0, s, b);
    return hipGetLastError();
}

hipError_t launch_batch_advance(const BatchView& b, bool count_only, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_advance, dim3(1), dim3(256), 0, s, b, count_only ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_batch_copy(const Tiles& t, const BatchView& b, int from, int to, const TileRef* tiles, int nblocks,
                             hipStream_t s) {
    with_x_tile(t.x, [&](auto XT) {
        hipLaunchKernelGGL(k_batch_copy<XT()>, dim3(nblocks), dim3(XT()), 0, s, b, from, to, tiles);
    });
    return hipGetLastError();
}

}  // namespace WENOFD_FP_NS
}  // namespace wenofd
