// wenofd_hybrid.hip — the learned switch's RK stage: one kernel per stage that
// computes every face flux of both directions, the rhs, the RK update, the
// CFL partials of the output and its ghost points. No counterpart in the
// reference (the hybrid flux is this project's; the plain stage is
// src/WENOFD.cpp:481-798 + :352-396).
//
// Flux at each face (hybrid_face in the oracle): the cheap component-wise
// upwind-5 flux, plus, for each characteristic field k flagged in the face's
// mask (bit k), R[:,k] (w_k - l_k) with w_k the scheme's WENO5 reconstruction
// of field k and l_k its linear reconstruction from the same projected
// stencils. Unflagged fields cost nothing beyond the cheap flux.
//
// Two kernels per stage. k_hyb_patch runs first, one face of the list of
// faces with flagged fields per thread (the list is built once per
// classification): the face's six stencil points (L2), their split, the mean
// state's eigen parts, then per flagged field row k of L, column k of R, the
// projections, the WENO and linear reconstructions and the correction.
// k_hyb_stage then does the rest, one wave per block: a block owns a strip of
// 64 loaded columns (56 output columns, 5 lanes left and 3 right of them) and
// a segment of rows; lane l owns column i = c0 - 5 + l and marches up the rows.
//   y faces: each loaded row's y split is added into the partial sums of the
//            six faces it reaches (registers; a face completes once the row
//            two above it is in), so each point is split once per direction;
//   x faces: from a wave-local LDS ring of the last four rows, each lane
//            splits the point two columns to its right and the face sums
//            travel between lanes by DPP shifts;
//   a face with flagged fields adds their summed correction (stored at the
//   face's index), loaded a row ahead once the face's mask, loaded further
//   ahead, says it has one.
// The grid is one round of resident waves (no tail), strips and segments
// ordered so that each XCD takes neighbouring ones.
// Compiled twice, like wenofd_kernels.hip: exact (no contraction) and fused.
#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

#include "wenofd_index.hpp"
#include "wenofd_kernels.hpp"
#include "wenofd_math.hpp"
#include "wenofd_switch.hpp"

#ifndef WENOFD_FP_NS
#error "build with -DWENOFD_FP_NS=exact (with -ffp-contract=off) or -DWENOFD_FP_NS=fused (with -ffp-contract=fast)"
#endif

namespace wenofd {
namespace WENOFD_FP_NS {

constexpr int kHybCols = 64;             // loaded columns per strip = lanes
constexpr int kHybLead = 5;              // lanes left of the first output column
constexpr int kHybOut = kHybCols - 8;    // output columns per strip (lanes 5 .. 60)
constexpr int kHybWaves = 2;             // resident waves per SIMD the kernel is built for
constexpr int kHybMinRows = 8;           // fewest output rows per segment

__host__ __device__ inline int hyb_strips(const Geom& g) { return (g.nx + kHybOut - 1) / kHybOut; }

__device__ __forceinline__ double hyb_wave_min(double v) {
    // every lane active; MIN is exact, so the exchange order does not matter
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = o < v ? o : v;
    }
    return v;
}

// ((L0 f0 + L1 f1) + L2 f2) + L3 f3, every coefficient multiplied (the
// oracle's 4-term sum, src/WENOFD.cpp:552-593).
__device__ __forceinline__ double dot4(const double a[4], const double f[4]) {
    return ((a[0] * f[0] + a[1] * f[1]) + a[2] * f[2]) + a[3] * f[3];
}

// w - l of one field from its projected stencils vp[0..4] (points f-3..f+1 of
// F+) and vm[0..4] (points f-2..f+2 of F-): the scheme's WENO5 (the shared
// reciprocal fast path, the IEEE divisions where a lane's operands leave its
// range) minus the linear reconstruction.
template <int S>
__device__ __forceinline__ double field_departure(const double vp[5], const double vm[5], double zp_lambda) {
    bool ok = true;
    double w = weno5_shared_rcp<S, true>(vp[0], vp[1], vp[2], vp[3], vp[4], zp_lambda, ok) +
               weno5_shared_rcp<S, true>(vm[4], vm[3], vm[2], vm[1], vm[0], zp_lambda, ok);
    if (__any(!ok)) {
        const double we = weno5<S>(vp[0], vp[1], vp[2], vp[3], vp[4], zp_lambda) +
                          weno5<S>(vm[4], vm[3], vm[2], vm[1], vm[0], zp_lambda);
        if (!ok) w = we;
    }
    const double l = linear5(vp[0], vp[1], vp[2], vp[3], vp[4]) + linear5(vm[4], vm[3], vm[2], vm[1], vm[0]);
    return w - l;
}

// Cheap-flux contribution of one point at offset o = p - f (o = -3..2) to
// face f: c+_o F+ + c-_o F- (upwind-5 on F+ and its mirror on F-).
template <int O>
__device__ __forceinline__ double cheap_term(double fp, double fm) {
    if constexpr (O == -3) return 2.0 * fp;
    else if constexpr (O == -2) return -13.0 * fp + -3.0 * fm;
    else if constexpr (O == -1) return 47.0 * fp + 27.0 * fm;
    else if constexpr (O == 0) return 27.0 * fp + 47.0 * fm;
    else if constexpr (O == 1) return -3.0 * fp + -13.0 * fm;
    else return 2.0 * fm;
}

// Whole-wave lane shifts of a double (DPP wave_shr:1 / wave_shl:1, every
// lane active): lane l receives lane l-1's value (lane 0 keeps its own), or
// lane l+1's (lane 63 keeps its own).
__device__ __forceinline__ double lane_from_left(double v) {
    const int lo = __double2loint(v), hi = __double2hiint(v);
    return __hiloint2double(__builtin_amdgcn_update_dpp(hi, hi, 0x138, 0xF, 0xF, false),
                            __builtin_amdgcn_update_dpp(lo, lo, 0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ double lane_from_right(double v) {
    const int lo = __double2loint(v), hi = __double2hiint(v);
    return __hiloint2double(__builtin_amdgcn_update_dpp(hi, hi, 0x130, 0xF, 0xF, false),
                            __builtin_amdgcn_update_dpp(lo, lo, 0x130, 0xF, 0xF, false));
}

// ---------------------------------------------------------------------------
// Corrections of the flagged fields, one face per thread: the face's six
// stencil points in Uin (read from L2: each list group runs in face order)
// split along D once, the mean state's eigen parts once, then for each
// flagged field k in ascending order R[:,k] (w_k - l_k), summed. The list
// groups faces by their number of flagged fields, so the lanes of a wave run
// the same number of fields. Blocks [0, nb) take the x list, [nb, 2 nb) the y
// list (grid-stride, a fixed launch shape).
// ---------------------------------------------------------------------------
constexpr int kPatchThreads = 256;
constexpr int kPatchBlocks = 512;  // per direction

// The face's six stencil points f-3 .. f+2 (rows of Uin along D).
template <int D>
__device__ __forceinline__ void patch_points(const double* __restrict__ Uin, const Geom& g, unsigned face,
                                             double P[6][4]) {
    size_t p0, step;  // index of stencil point f-3 and the stride between points
    if constexpr (D == 0) {
        const int row = (int)(face / (unsigned)(g.nx + 1)), f = (int)(face % (unsigned)(g.nx + 1)) + 4;
        p0 = (size_t)(row + 4) * g.pitch + (f - 3);
        step = 1;
    } else {
        const int fr = (int)(face / (unsigned)g.nx) + 4, i = (int)(face % (unsigned)g.nx) + 4;
        p0 = (size_t)(fr - 3) * g.pitch + i;
        step = (size_t)g.pitch;
    }
#pragma unroll
    for (int s = 0; s < 6; ++s)
#pragma unroll
        for (int r = 0; r < 4; ++r) P[s][r] = Uin[r * g.plane + p0 + (size_t)s * step];
}

// The points' split along D and the face's mean state.
template <int D>
__device__ __forceinline__ void patch_split(const double P[6][4], const Geom& g, double Fp[6][4], double Fm[6][4],
                                            double ave[4]) {
#pragma unroll
    for (int r = 0; r < 4; ++r) ave[r] = 0.5 * (P[2][r] + P[3][r]);
#pragma unroll
    for (int s = 0; s < 6; ++s) flux_split<D>(P[s], g.gamma, Fm[s], Fp[s]);
}

// The summed corrections of the flagged fields m of one face.
template <int S, int D>
__device__ __forceinline__ void patch_fields(const double Fp[6][4], const double Fm[6][4], const double ave[4],
                                             const Geom& g, unsigned m, double* __restrict__ out) {
    const EigenParts<double> e = eigen_parts<D>(ave, g.gamma);
    double corr[4];
    bool first = true;
#pragma unroll 1
    for (unsigned rest = m; rest; rest &= rest - 1) {
        const int kf = __builtin_ctz(rest);
        double Lk[4], Rk[4], vp[5], vm[5];
        eigen_row_col<D>(e, kf, Lk, Rk);
#pragma unroll
        for (int s = 0; s < 5; ++s) {
            vp[s] = dot4(Lk, Fp[s]);
            vm[s] = dot4(Lk, Fm[s + 1]);
        }
        const double d = field_departure<S>(vp, vm, g.zp_lambda);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const double c = Rk[r] * d;
            corr[r] = first ? c : corr[r] + c;
        }
        first = false;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) out[r] = corr[r];
}

// List entries e0, e0 + stride, ... below n of one direction. A face's points
// are split as soon as they are in; the next face's points are then loaded
// into the same registers and arrive while this face's fields are computed
// (the entry of the face after it and its mask one face earlier still). Every
// load reads a valid entry and valid points (the entry index clamped), so no
// load in flight is waited for early.
template <int S, int D>
__device__ __forceinline__ void patch_run(const double* __restrict__ Uin, const Geom& g,
                                          const unsigned* __restrict__ list, const unsigned char* __restrict__ masks,
                                          int n, double* __restrict__ corr, int e0, int stride) {
    if (e0 >= n) return;
    auto entry = [&](int e) { return list[e < n ? e : n - 1]; };
    unsigned fc = entry(e0);
    unsigned mc = masks[fc];
    unsigned fn = entry(e0 + stride);
    double P[6][4];
    patch_points<D>(Uin, g, fc, P);
    for (int e = e0; e < n; e += stride) {
        double Fp[6][4], Fm[6][4], ave[4];
        patch_split<D>(P, g, Fp, Fm, ave);
        const unsigned mn = masks[fn];
        const unsigned fnn = entry(e + 2 * stride);
        patch_points<D>(Uin, g, fn, P);
        patch_fields<S, D>(Fp, Fm, ave, g, mc & 15u, corr + (size_t)fc * 4);
        fc = fn;
        mc = mn;
        fn = fnn;
    }
}

template <int S>
__global__ __launch_bounds__(kPatchThreads) void k_hyb_patch(const double* __restrict__ Uin, Geom g,
                                                             const unsigned* __restrict__ list_x,
                                                             const unsigned* __restrict__ list_y,
                                                             const unsigned char* __restrict__ mx,
                                                             const unsigned char* __restrict__ my,
                                                             const int* __restrict__ n_dev,
                                                             double* __restrict__ corr_x,
                                                             double* __restrict__ corr_y,
                                                             const SolveState* __restrict__ st) {
    if (st && !st->active) return;
    const int nb = gridDim.x / 2;
    const bool xdir = (int)blockIdx.x < nb;
    const int b = xdir ? blockIdx.x : blockIdx.x - nb;
    const int n = n_dev[xdir ? 0 : 1];
    const int e0 = b * kPatchThreads + threadIdx.x, stride = nb * kPatchThreads;
    if (xdir) patch_run<S, 0>(Uin, g, list_x, mx, n, corr_x, e0, stride);
    else patch_run<S, 1>(Uin, g, list_y, my, n, corr_y, e0, stride);
}

// The summed correction of face `face`'s flagged fields (loaded when it has any).
__device__ __forceinline__ void load_correction(bool on, size_t face, const double* __restrict__ corr, double c[4]) {
    if (on) {
#pragma unroll
        for (int r = 0; r < 4; ++r) c[r] = corr[face * 4 + r];
    }
}

// The row march of the stage kernel is unrolled kHybAhead times: row q is
// loaded kHybAhead rows before it is split, with the Un of its cell row, each
// row of an unrolled group into its own registers, which are reloaded in
// place once read, so that no load in flight is ever copied (a copy of a
// register with a load in flight waits for that load). The corrections of the
// faces of row q + 1 are loaded first thing at row q (vector memory counts
// complete in issue order, so nothing issued earlier is waited for early),
// from masks loaded kHybAhead rows before that. Every load reads a valid
// address (columns and rows clamped into the array): a load under a lane
// condition with the other lanes' registers set otherwise would also wait for
// the load.
constexpr int kHybAhead = 2;

// One RK stage of the hybrid solver on one strip of columns and segment of rows:
//   kRkNone:  rhs only (rhs_out)     kRk1: out = in + dt rhs
//   kRkStage: out = (a Un + b in) + (c dt) rhs
// mask_x / mask_y: per face (x faces [ny][nx+1], y faces [ny+1][nx]) the
// flagged fields (bits 0-3), whose summed correction is at the face's index in
// corr_x / corr_y. With part set,
// the block's CFL partial minimum of its updated cells.
// step(j, q0 + j) for j = 0 .. NB-1, j a compile-time constant.
template <class F, int... J>
__device__ __forceinline__ void each_row(F& step, int q0, std::integer_sequence<int, J...>) {
    (step(std::integral_constant<int, J>(), q0 + J), ...);
}

// Keeps the compiler from moving the reload of a register set above the last
// reads of its previous contents: the reload then lands in the same registers
// and the loop carries it without a copy.
__device__ __forceinline__ void reload_fence() { __builtin_amdgcn_sched_barrier(0); }

// Raw buffer resources (stride 0, byte range = num_records): dword 3 of the
// gfx9-family descriptor (32-bit data format); an access at an offset past
// the range is dropped (a store) or reads 0 (a load).
constexpr int kBufferWord3 = 0x00020000;
typedef unsigned U2 __attribute__((ext_vector_type(2)));

template <int NB, int MODE>
__global__ __launch_bounds__(kHybCols, kHybWaves) void k_hyb_stage(const double* __restrict__ Uin, const double* Un,
                                                                   double* Uout, double* __restrict__ rhs_out, Geom g,
                                                                   RkCoef k, double cfl, double* __restrict__ part,
                                                                   int* __restrict__ part_nan,
                                                                   const unsigned char* __restrict__ mask_x,
                                                                   const unsigned char* __restrict__ mask_y,
                                                                   const double* __restrict__ corr_x,
                                                                   const double* __restrict__ corr_y,
                                                                   const SolveState* __restrict__ st) {
    static_assert(NB % 2 == 0, "the corrections alternate between two register sets");
    if (st) {
        if (!st->active) return;
        k.cdt = k.cdt * st->dt;
    }
    __shared__ double su[4][4][kHybCols];  // U of split rows: row q in slot q & 3
    const int lane = threadIdx.x;
    // the block's strip and segment: consecutive blocks (one XCD each, in
    // turn) map to ranges of neighbouring strips and segments per XCD, so the
    // halo columns and rows that neighbours both read share an L2
    const int nblk = gridDim.x * gridDim.y;
    const int lb = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, nblk);
    const int bx = lb % gridDim.x, by = lb / gridDim.x;
    const int c0 = 4 + bx * kHybOut;  // first output column
    const int i = c0 - kHybLead + lane;       // this lane's column
    const int lx2 = lane + 2 < kHybCols ? lane + 2 : kHybCols - 1;  // the lane two columns right
    const bool out_col = lane >= kHybLead && lane < kHybLead + kHybOut && i < g.nx + 4;
    const bool x_face = lane >= kHybLead && lane <= kHybLead + kHybOut && i <= g.nx + 4;  // face i (points i-1, i)
    // segment by of gridDim.y: output rows [r0, r1), the rows split evenly
    const int r0 = 4 + (int)((long)by * g.ny / gridDim.y);
    const int r1 = 4 + (int)((long)(by + 1) * g.ny / gridDim.y);
    const int qa = r0 - 3, qe = r1 + 2;  // the rows split
    const int rmax = g.ny + 7;
    const bool gh_x = k.gh.on && ghost_x_edge(g, c0, c0 + kHybOut - 1);
    // byte offset of this lane's point in a row (clamped into the row), and
    // the index of its y face / x face in a face row (clamped)
    const unsigned lo8 = (unsigned)(i < 0 ? 0 : (i < g.nx + 7 ? i : g.nx + 7)) * 8u;
    const unsigned ly1 = (unsigned)((i < 4 ? 4 : (i > g.nx + 3 ? g.nx + 3 : i)) - 4);
    const unsigned lx1 = (unsigned)((i < 4 ? 4 : (i > g.nx + 4 ? g.nx + 4 : i)) - 4);

    auto lane_d = [&](const double* row_ptr) -> double {
        return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(row_ptr) + lo8);
    };
    // U of row, clamped to the rows the march splits (the prefetches past its
    // last row re-read that row instead of fetching rows it never uses)
    const int qtop = qe < rmax ? qe : rmax;
    auto load_u = [&](int row, double u[4]) {
        row = row < 0 ? 0 : (row < qtop ? row : qtop);
        const double* p = Uin + (size_t)row * g.pitch;
#pragma unroll
        for (int r = 0; r < 4; ++r) u[r] = lane_d(p + (size_t)r * g.plane);
    };
    auto load_un = [&](int row, double u[4]) {  // Un of row (clamped), RK stages only
        if constexpr (MODE == kRkStage) {
            row = row < r0 ? r0 : (row < r1 ? row : r1 - 1);
            const double* p = Un + (size_t)row * g.pitch;
#pragma unroll
            for (int r = 0; r < 4; ++r) u[r] = lane_d(p + (size_t)r * g.plane);
        }
    };
    // buffer stores of w to row `row` of the output planes, dropped (offset
    // past the row) for the lanes without `on`
    double* const dst = MODE == kRkNone ? rhs_out : Uout;
    auto store_row = [&](int row, bool on, const double v[4]) {
        row = row < 0 ? 0 : row;
        const unsigned off = on ? lo8 : 0x80000000u;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                dst + (size_t)r * g.plane + (size_t)row * g.pitch, (short)0, g.pitch * 8, kBufferWord3);
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, v[r]), rs, off, 0, 0);
        }
    };
    // masks of the faces of row q: y face q-2 and the x faces of cell row q-3 (clamped)
    auto load_masks = [&](int q, unsigned& sy, unsigned& sx) {
        int fr = q - 2, cr = q - 3;
        fr = fr < 4 ? 4 : (fr > g.ny + 4 ? g.ny + 4 : fr);
        cr = cr < 4 ? 4 : (cr > g.ny + 3 ? g.ny + 3 : cr);
        sy = mask_y[(size_t)(fr - 4) * g.nx + ly1];
        sx = mask_x[(size_t)(cr - 4) * (g.nx + 1) + lx1];
    };

    double ub[NB][4];     // U of the rows of the group, loaded NB rows ahead
    double unb[NB][4];    // Un of their cell rows
    unsigned syb[NB], sxb[NB];  // masks of the faces of rows q + 1 (loaded NB rows before row q)
    double cyb[2][4], cxb[2][4];  // corrections of the faces of rows q (set q & 1), where flagged
    bool fyb_on[2] = {false, false}, fxb_on[2] = {false, false};  // which lanes have one
    double acc[5][4];   // acc[m]: partial cheap flux of face q-2+m (before row q)
    double fyb[4];      // flux of the bottom face of the next cell row
    double my_min = 1.0;
    int my_nan = 0;
#pragma unroll
    for (int m = 0; m < 5; ++m)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[m][r] = 0.0;
#pragma unroll
    for (int r = 0; r < 4; ++r) fyb[r] = cyb[0][r] = cyb[1][r] = cxb[0][r] = cxb[1][r] = 0.0;
    // The march starts NB rows early, on rows whose only use is to load rows
    // qa .. qa+NB-1 (their own sums feed faces below r0, never used), so that
    // the loop is entered with no load in flight and its waits are those of
    // the steady state.
#pragma unroll
    for (int j = 0; j < NB; ++j) {
#pragma unroll
        for (int r = 0; r < 4; ++r) ub[j][r] = unb[j][r] = 0.0;
        ub[j][0] = ub[j][3] = 1.0;  // a finite state for the early rows' splits
        syb[j] = sxb[j] = 0u;
    }

    auto step = [&](auto J, int q) {
        constexpr int j = decltype(J)::value;
        constexpr int cs = j & 1, cn = cs ^ 1;  // correction sets of rows q and q + 1 (NB is even)
        if (q > qe) return;
        // the corrections of the faces of row q + 1 this block owns
        {
            const int fr = q + 1 - 2, cr = q + 1 - 3;
            fyb_on[cn] = out_col && fr >= r0 && (syb[j] & 15u) != 0u;
            fxb_on[cn] = x_face && cr >= r0 && (sxb[j] & 15u) != 0u;
            load_correction(fyb_on[cn], (size_t)(fr - 4) * g.nx + ly1, corr_y, cyb[cn]);
            load_correction(fxb_on[cn], (size_t)(cr - 4) * (g.nx + 1) + lx1, corr_x, cxb[cn]);
            reload_fence();
            load_masks(q + 1 + NB, syb[j], sxb[j]);
        }
        // ---- row q: y split, partial sums of faces q-2 .. q+3 ----
        double fm[4], fp[4], fyt[4];
        flux_split<1>(ub[j], g.gamma, fm, fp);
#pragma unroll
        for (int r = 0; r < 4; ++r) su[q & 3][r][lane] = ub[j][r];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            fyt[r] = (acc[0][r] + cheap_term<2>(fp[r], fm[r])) * (1.0 / 60.0);  // face q-2 complete
            acc[0][r] = acc[1][r] + cheap_term<1>(fp[r], fm[r]);
            acc[1][r] = acc[2][r] + cheap_term<0>(fp[r], fm[r]);
            acc[2][r] = acc[3][r] + cheap_term<-1>(fp[r], fm[r]);
            acc[3][r] = acc[4][r] + cheap_term<-2>(fp[r], fm[r]);
            acc[4][r] = cheap_term<-3>(fp[r], fm[r]);
        }
        reload_fence();
        load_u(q + NB, ub[j]);
        const int fc = q - 2;  // the face just completed (wave-uniform)
        const int c = fc - 1;  // the cell row updated at this row (when fc > r0)
        double w[4] = {0.0, 0.0, 0.0, 0.0};  // its new state (its rhs with MODE kRkNone)
        if (fc >= r0) {
            if (fyb_on[cs]) {
#pragma unroll
                for (int r = 0; r < 4; ++r) fyt[r] = fyt[r] + cyb[cs][r];
            }
            if (fc > r0) {
                // ---- cell row c: x faces, then the update ----
                // Face i (points i-1, i) sums the cheap-flux terms of points
                // i-3 .. i+2 in that order: each lane splits the point two
                // columns to its right, and the partial sum moves one lane to
                // the right before each next term is added, so lane l ends
                // with its own column's left face (lanes 5 .. 61 whole).
                double uc[4], ux[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    uc[r] = su[c & 3][r][lane];
                    ux[r] = su[c & 3][r][lx2];
                }
                double fxm[4], fxp[4], fl[4], fr[4];
                flux_split<0>(ux, g.gamma, fxm, fxp);
                const bool cx = fxb_on[cs];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    double a = cheap_term<-3>(fxp[r], 0.0);
                    a = lane_from_left(a) + cheap_term<-2>(fxp[r], fxm[r]);
                    a = lane_from_left(a) + cheap_term<-1>(fxp[r], fxm[r]);
                    a = lane_from_left(a) + cheap_term<0>(fxp[r], fxm[r]);
                    a = lane_from_left(a) + cheap_term<1>(fxp[r], fxm[r]);
                    a = lane_from_left(a) + cheap_term<2>(0.0, fxm[r]);
                    a = a * (1.0 / 60.0);
                    fl[r] = cx ? a + cxb[cs][r] : a;
                    fr[r] = lane_from_right(fl[r]);  // the right face: the next lane's left face
                }
                if (out_col) {
                    double rhs[4];
                    cell_rhs4(fl, fr, fyb, fyt, g.dx, g.dy, g.rdx, g.rdy, rhs);
                    if (k.gravity) {  // src/WENOFD.cpp:812-813
                        rhs[3] = rhs[3] + uc[2];
                        rhs[2] = rhs[2] + uc[0];
                    }
                    if constexpr (MODE == kRkNone) {
#pragma unroll
                        for (int r = 0; r < 4; ++r) w[r] = rhs[r];
                    } else {
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            w[r] = (MODE == kRk1) ? uc[r] + k.cdt * rhs[r]
                                                  : rk_combine(k.a, unb[j][r], k.b, uc[r], k.cdt, rhs[r]);
                        const bool gh_y = k.gh.on && ghost_y_edge(g, c);
                        if (gh_x || gh_y) {
                            ghost_targets(
                                g, k.gh.bc, i, c, gh_x, gh_y,
                                [&](int di, int dj, int flip) {
                                    const size_t d = (size_t)dj * g.pitch + di;
#pragma unroll
                                    for (int r = 0; r < 4; ++r) Uout[r * g.plane + d] = (r == flip) ? -w[r] : w[r];
                                },
                                [&](int di, int dj, int spec) {
                                    const size_t d = (size_t)dj * g.pitch + di;
                                    const int ns = 2 * (g.nx + g.ny);
#pragma unroll
                                    for (int r = 0; r < 4; ++r) Uout[r * g.plane + d] = k.gh.special[r * ns + spec];
                                });
                        }
                        if (part) {
                            const double tl = local_dt_dev(w, g.gamma, cfl, g.dx, g.dy, g.rdx, g.rdy);
                            if (tl != tl) my_nan = 1;
                            my_min = (tl < my_min) ? tl : my_min;
                        }
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) fyb[r] = fyt[r];
        }
        // cell row c out (every row issues these stores: lanes and rows that
        // own no cell are dropped by the buffer's range check)
        store_row(c, out_col && fc > r0, w);
        // the set now takes row q + NB
        reload_fence();
        load_un(q + NB - 3, unb[j]);
    };

    for (int q0 = qa - NB; q0 <= qe; q0 += NB) each_row(step, q0, std::make_integer_sequence<int, NB>());
    if (part) {
        my_min = hyb_wave_min(my_min);
        my_nan = __any(my_nan) ? 1 : 0;
        if (lane == 0) {
            const int b = blockIdx.y * gridDim.x + blockIdx.x;
            part[b] = my_min;
            part_nan[b] = my_nan;
        }
    }
}

// Segments per strip: the grid is one round of resident_waves waves (the
// device's CU count x 4 SIMDs x kHybWaves), so every wave marches one segment
// and none waits for a second round; at least kHybMinRows rows per segment.
int hyb_stage_segments(const Geom& g, int resident_waves) {
    const int per = resident_waves / hyb_strips(g);
    const int most = g.ny / kHybMinRows;
    const int segs = per < most ? per : most;
    return segs > 1 ? segs : 1;
}

long hyb_stage_blocks(const Geom& g, int segs) { return (long)hyb_strips(g) * segs; }

int hyb_waves_per_simd() { return kHybWaves; }

hipError_t launch_hyb_patch(int scheme, const double* Uin, const Geom& g, const unsigned* list_x,
                            const unsigned* list_y, const unsigned char* mx, const unsigned char* my,
                            const int* n_dev, double* corr_x, double* corr_y, hipStream_t s, const SolveState* st) {
    auto go = [&](auto S) {
        hipLaunchKernelGGL((k_hyb_patch<decltype(S)::value>), dim3(2 * kPatchBlocks), dim3(kPatchThreads), 0, s,
                           Uin, g, list_x, list_y, mx, my, n_dev, corr_x, corr_y, st);
    };
    switch (scheme) {
        case kJS: go(std::integral_constant<int, kJS>()); break;
        case kZ: go(std::integral_constant<int, kZ>()); break;
        case kZP: go(std::integral_constant<int, kZP>()); break;
        default: go(std::integral_constant<int, kZPI>()); break;
    }
    return hipGetLastError();
}

hipError_t launch_hyb_stage(const double* Uin, const double* Un, double* Uout, double* rhs_out, const Geom& g,
                            const RkCoef& k, double cfl, double* dt_part, int* dt_nan,
                            const unsigned char* mask_x, const unsigned char* mask_y, const double* corr_x,
                            const double* corr_y, int segs,
                            hipStream_t s, const SolveState* st) {
    const dim3 grid(hyb_strips(g), segs);
    auto go = [&](auto M) {
        hipLaunchKernelGGL((k_hyb_stage<kHybAhead, decltype(M)::value>), grid, dim3(kHybCols), 0, s, Uin, Un, Uout,
                           rhs_out, g, k, cfl, dt_part, dt_nan, mask_x, mask_y, corr_x, corr_y, st);
    };
    if (k.mode == kRkNone) go(std::integral_constant<int, kRkNone>());
    else if (k.mode == kRk1) go(std::integral_constant<int, kRk1>());
    else go(std::integral_constant<int, kRkStage>());
    return hipGetLastError();
}

}  // namespace WENOFD_FP_NS
}  // namespace wenofd
