// wenofd_index.hpp — index maps of the WENO-FD kernels, host and device.
//
// The kernels (wenofd_kernels.hip) use these maps to place ghost points, to
// order their blocks over the XCDs and to cut the grid into tiles; a host
// program (wenofd_index_check.cpp) brute-forces them over many grid sizes,
// boundary kinds and tile shapes with bounds checks, under the address and
// undefined-behaviour sanitizers.
//
// Point numbering (src/WENOFD.cpp:36-70): i in [0, nx+8), j in [0, ny+8),
// interior i, j in [4, n+4), 4 ghost points per side.
#pragma once

#include <hip/hip_runtime.h>

#include "wenofd_kernels.hpp"

namespace wenofd {

#define WENOFD_IHD __host__ __device__ __forceinline__

// Blocks are dispatched round-robin over the chip's 8 XCDs (block b runs on
// XCD b mod 8), each with its own L2. xcd_block(b, n) maps b to a logical
// block index so that every XCD works through one contiguous range of the n
// logical blocks (a bijection on [0, n)): neighbouring tiles, which read the
// same rows, then share one L2.
constexpr int kXcds = 8;
WENOFD_IHD int xcd_block(int b, int n) {
    const int per = (n + kXcds - 1) / kXcds;              // blocks of the XCDs that take the most
    const int tall = n % kXcds == 0 ? kXcds : n % kXcds;  // how many XCDs take `per`; the rest per - 1
    const int xcd = b % kXcds, local = b / kXcds;
    return xcd < tall ? xcd * per + local : tall * per + (xcd - tall) * (per - 1) + local;
}

// One ghost point of the fill (exchangeGhostCellsValue's one-rank periodic-x
// ring + setBoundary, src/WENOFD.cpp:850-929, 998-1159), restricted to the
// ghost points the sweeps read: x ghosts of interior rows, y ghosts of
// interior columns (corner ghosts are never read).
struct GhostRef {
    int di, dj;  // the ghost point
    int si, sj;  // the point it copies (when spec < 0)
    int flip;    // component negated (1: rho u at an x slip side, 2: rho v at a y slip side), -1 none
    int spec;    // >= 0: index into the SPECIAL band [4][2 nx + 2 ny] instead of a copy
};

constexpr int kGhostPerPoint = 8;  // ghosts per interior row (x) or column (y): 4 each side

// Number of ghost points one fill writes.
WENOFD_IHD long ghost_count(const Geom& g) { return (long)kGhostPerPoint * (g.nx + g.ny); }

// Ghost k of the fill, k in [0, ghost_count(g)): x ghosts first (8 per
// interior row: 4 left, 4 right), then y ghosts (8 per interior column: 4
// bottom, 4 top; the bottom boundary type governs both y sides, :1009).
WENOFD_IHD GhostRef ghost_map(const Geom& g, const Bc& bc, int k) {
    const int nxg = kGhostPerPoint * g.ny;
    const int sx = 4, ex = g.nx + 4, sy = 4, ey = g.ny + 4;
    GhostRef r{0, 0, 0, 0, -1, -1};
    if (k < nxg) {
        const int j = sy + k / 8;
        const int e = k % 8;
        r.dj = r.sj = j;
        if (e < 4) {  // left ghost i = sx - e1
            const int e1 = e + 1;
            r.di = sx - e1;
            if (bc.left == kBcPeriod) r.si = ex - 1 - e1;
            else if (bc.left == kBcSpecial) { r.si = r.di; r.spec = 2 * g.nx + (j - sy); }
            else { r.si = sx + e1; if (bc.left == kBcSlip) r.flip = 1; }
        } else {  // right ghost i = ex - 1 + e1
            const int e1 = e - 3;
            r.di = ex - 1 + e1;
            if (bc.right == kBcPeriod) r.si = sx + e1;
            else if (bc.right == kBcSpecial) { r.si = r.di; r.spec = 2 * g.nx + g.ny + (j - sy); }
            else { r.si = ex - 1 - e1; if (bc.right == kBcSlip) r.flip = 1; }
        }
    } else {
        const int kk = k - nxg;
        const int i = sx + kk / 8;
        const int e = kk % 8;
        r.di = r.si = i;
        if (e < 4) {  // bottom ghost j = sy - e1
            const int e1 = e + 1;
            r.dj = sy - e1;
            if (bc.bottom == kBcPeriod) r.sj = ey - 1 - e1;
            else if (bc.bottom == kBcSpecial) { r.sj = r.dj; r.spec = i - sx; }
            else { r.sj = sy + e1; if (bc.bottom == kBcSlip) r.flip = 2; }
        } else {  // top ghost j = ey - 1 + e1
            const int e1 = e - 3;
            r.dj = ey - 1 + e1;
            if (bc.bottom == kBcPeriod) r.sj = sy + e1;
            else if (bc.bottom == kBcSpecial) { r.sj = r.dj; r.spec = g.nx + (i - sx); }
            else { r.sj = ey - 1 - e1; if (bc.bottom == kBcSlip) r.flip = 2; }
        }
    }
    return r;
}

// ghost_map inverted: the ghost points whose source is interior point (i, j):
// put(di, dj, flip) for each copy (flip as in GhostRef), and for a SPECIAL
// side's four ghosts next to (i, j) put_special(di, dj, spec). With nx, ny >= 5
// every ghost of the fill has an interior source, so enumerating every
// interior point yields every ghost of ghost_map exactly once. xs / ys enable
// the x-ghost part (points within 5 of the left or right edge) and the
// y-ghost part (within 5 of the bottom or top).
template <class Put, class PutSpecial>
WENOFD_IHD void ghost_targets(const Geom& g, const Bc& bc, int i, int j, bool xs, bool ys, Put put,
                              PutSpecial put_special) {
    const int sx = 4, ex = g.nx + 4, sy = 4, ey = g.ny + 4;
    // left ghosts i = sx - e1 and right ghosts i = ex - 1 + e1 of row j
    if (!xs) {
    } else if (bc.left == kBcPeriod) {
        const int e1 = ex - 1 - i;
        if (e1 >= 1 && e1 <= 4) put(sx - e1, j, -1);
    } else if (bc.left == kBcSpecial) {
        if (i == sx)
#pragma unroll 1
            for (int e1 = 1; e1 <= 4; ++e1) put_special(sx - e1, j, 2 * g.nx + (j - sy));
    } else {
        const int e1 = i - sx;
        if (e1 >= 1 && e1 <= 4) put(sx - e1, j, bc.left == kBcSlip ? 1 : -1);
    }
    if (!xs) {
    } else if (bc.right == kBcPeriod) {
        const int e1 = i - sx;
        if (e1 >= 1 && e1 <= 4) put(ex - 1 + e1, j, -1);
    } else if (bc.right == kBcSpecial) {
        if (i == ex - 1)
#pragma unroll 1
            for (int e1 = 1; e1 <= 4; ++e1) put_special(ex - 1 + e1, j, 2 * g.nx + g.ny + (j - sy));
    } else {
        const int e1 = ex - 1 - i;
        if (e1 >= 1 && e1 <= 4) put(ex - 1 + e1, j, bc.right == kBcSlip ? 1 : -1);
    }
    // bottom ghosts j = sy - e1 and top ghosts j = ey - 1 + e1 of column i
    if (!ys) {
    } else if (bc.bottom == kBcPeriod) {
        const int e1 = ey - 1 - j;
        if (e1 >= 1 && e1 <= 4) put(i, sy - e1, -1);
        const int e2 = j - sy;
        if (e2 >= 1 && e2 <= 4) put(i, ey - 1 + e2, -1);
    } else if (bc.bottom == kBcSpecial) {
        if (j == sy)
#pragma unroll 1
            for (int e1 = 1; e1 <= 4; ++e1) put_special(i, sy - e1, i - sx);
        if (j == ey - 1)
#pragma unroll 1
            for (int e1 = 1; e1 <= 4; ++e1) put_special(i, ey - 1 + e1, g.nx + (i - sx));
    } else {
        const int flip = bc.bottom == kBcSlip ? 2 : -1;
        const int e1 = j - sy;
        if (e1 >= 1 && e1 <= 4) put(i, sy - e1, flip);
        const int e2 = ey - 1 - j;
        if (e2 >= 1 && e2 <= 4) put(i, ey - 1 + e2, flip);
    }
}

// Whether interior cells first..last of a row may be sources of x ghosts
// (within 5 of the left or right edge), and whether row j may hold sources of
// y ghosts: the gates of the x sweep's ghost writes (ghost_targets yields
// nothing for the cells they exclude).
WENOFD_IHD bool ghost_x_edge(const Geom& g, int first, int last) { return first <= 4 + 4 || last >= g.nx + 4 - 5; }
WENOFD_IHD bool ghost_y_edge(const Geom& g, int j) { return j <= 4 + 4 || j >= g.ny + 4 - 5; }

// ---- tiles ----------------------------------------------------------------------
// x sweep: one tile = XT-5 consecutive faces of one interior row = the XT-6
// cells between them (consecutive tiles overlap by one face); it reads the XT
// points f0-3 .. f0+XT-4 (f0 = its first face = its first cell).
template <int XT>
constexpr int x_cells() { return XT - 6; }
template <int XT>
WENOFD_IHD int x_tiles_x(const Geom& g) { return (g.nx + x_cells<XT>() - 1) / x_cells<XT>(); }
template <int XT>
WENOFD_IHD int x_tile_first_face(int bx) { return 4 + bx * x_cells<XT>(); }

// y sweep: one tile = 64 interior columns x YF faces along y (the last tile
// of a column stops at face ny + 4), marched kYR face rows per step. The LLF
// split of the 9 point rows one step reads (faces f .. f+3 read point rows
// f-3 .. f+5) lives in a ring of kRing rows; the points themselves (for the
// faces' mean states, rows f-1 .. f+3) in a ring of kRawRing rows. A step
// first loads rows f+2 .. f+5 into both rings, then (after a barrier) builds
// its faces, then waits at a second barrier; before the first step a lead-in
// loads rows F0-3 .. F0+1 (F0 = the tile's first face), the point ring only
// from row F0-1 on.
constexpr int kYC = 64;     // columns per y tile
constexpr int kYR = 4;      // face rows per step
constexpr int kRing = 9;    // split-value ring rows
constexpr int kRawRing = 8; // point ring rows
template <int YF>
WENOFD_IHD int y_tile_first_face(int by) { return 4 + by * YF; }
template <int YF>
WENOFD_IHD int y_tile_end_face(const Geom& g, int by) {  // one past the tile's last face
    const int e = 4 + by * YF + YF;
    return e < g.ny + 5 ? e : g.ny + 5;
}
// Whether the point ring keeps point row prow of the tile starting at face
// F0: the faces read points from row F0-1 on, and the lead-in's rows F0-3 and
// F0-2 would share their slots with rows F0+5 and F0+6, which the first
// steps write before the lead-in's values could be read.
WENOFD_IHD bool y_ring_keeps_point(int F0, int prow) { return prow - (F0 - 3) >= 2; }
template <int YF>
WENOFD_IHD int y_tiles_x(const Geom& g) { return (g.nx + kYC - 1) / kYC; }
template <int YF>
WENOFD_IHD int y_tiles_y(const Geom& g) { return (g.ny + 1 + YF - 1) / YF; }

}  // namespace wenofd
