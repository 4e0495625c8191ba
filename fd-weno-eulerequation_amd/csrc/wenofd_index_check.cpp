// wenofd_index_check.cpp — host brute force of the kernels' index maps
// (wenofd_index.hpp), built with the address and undefined-behaviour
// sanitizers. For many grid sizes, every boundary kind and every tile shape
// it checks, against the buffers the library allocates for that size:
//
//   ghost_map       every ghost point the sweeps read (corners excluded) is
//                   written exactly once; each copies an interior point (or a
//                   SPECIAL band entry in range); the periodic maps skip the
//                   duplicated end node (src/WENOFD.cpp:864-865, 1013-1018);
//   ghost_targets   enumerated over every interior point it yields exactly
//                   the writes of ghost_map (nx, ny >= 5), and only at points
//                   the x sweep's gates (ghost_x_edge, ghost_y_edge) admit;
//   xcd_block       a bijection on [0, n);
//   x tiles         cover every interior cell once; every point, LDS slot,
//                   y-face and state access of the fused x sweep stays in its
//                   buffer, and every slot a face reads was loaded;
//   y tiles         cover every face once; replaying the ring schedule, no
//                   two rows written in one phase share a slot and every slot
//                   a face reads holds the row it expects.
//
// Usage: wenofd_index_check  (prints one summary line; exit status 1 on the
// first failure, naming the case).
#include <cstdio>
#include <cstdlib>
#include <map>
#include <set>
#include <tuple>
#include <vector>

#include "wenofd_index.hpp"

using namespace wenofd;

namespace {

long g_checks = 0;

#define CHECK(cond, ...)                                                \
    do {                                                                \
        ++g_checks;                                                     \
        if (!(cond)) {                                                  \
            std::fprintf(stderr, "index check failed: %s: ", #cond);   \
            std::fprintf(stderr, __VA_ARGS__);                          \
            std::fprintf(stderr, "\n");                                 \
            std::exit(1);                                               \
        }                                                               \
    } while (0)

// The geometry the library allocates for an nx x ny solve (wenofd_capi.cpp).
Geom geom(int nx, int ny) {
    auto round_up = [](long v, long m) { return (v + m - 1) / m * m; };
    Geom g{};
    g.nx = nx;
    g.ny = ny;
    g.pitch = (int)round_up(nx + 8, 32);
    g.plane = (long)g.pitch * (ny + 8);
    g.fy_pitch = (int)round_up(nx, 32);
    g.fy_plane = (long)g.fy_pitch * (ny + 1);
    return g;
}

bool interior(const Geom& g, int i, int j) { return i >= 4 && i < g.nx + 4 && j >= 4 && j < g.ny + 4; }
bool corner(const Geom& g, int i, int j) {
    return (i < 4 || i >= g.nx + 4) && (j < 4 || j >= g.ny + 4);
}

using Write = std::tuple<int, int, int, int, int, int>;  // di, dj, si, sj, flip, spec

void check_ghosts(int nx, int ny, const Bc& bc) {
    const Geom g = geom(nx, ny);
    const int ns = 2 * (nx + ny);
    std::map<std::pair<int, int>, Write> fill;
    const long n = ghost_count(g);
    for (int k = 0; k < n; ++k) {
        const GhostRef r = ghost_map(g, bc, k);
        CHECK(r.di >= 0 && r.di < nx + 8 && r.dj >= 0 && r.dj < ny + 8, "nx %d ny %d k %d: ghost out of the array",
              nx, ny, k);
        CHECK(!interior(g, r.di, r.dj) && !corner(g, r.di, r.dj), "nx %d ny %d k %d: not an edge ghost", nx, ny, k);
        CHECK((size_t)r.dj * g.pitch + r.di < (size_t)g.plane, "nx %d ny %d k %d: ghost beyond the plane", nx, ny, k);
        if (r.spec >= 0) {
            CHECK(r.spec < ns, "nx %d ny %d k %d: SPECIAL index %d out of [0, %d)", nx, ny, k, r.spec, ns);
        } else {
            CHECK(interior(g, r.si, r.sj), "nx %d ny %d k %d: source (%d, %d) not interior", nx, ny, k, r.si, r.sj);
            CHECK(r.flip == -1 || r.flip == 1 || r.flip == 2, "flip %d", r.flip);
        }
        const auto key = std::make_pair(r.di, r.dj);
        CHECK(!fill.count(key), "nx %d ny %d: ghost (%d, %d) written twice", nx, ny, r.di, r.dj);
        fill[key] = Write{r.di, r.dj, r.spec >= 0 ? -1 : r.si, r.spec >= 0 ? -1 : r.sj, r.spec >= 0 ? -1 : r.flip,
                          r.spec};
    }
    // every non-corner ghost point
    CHECK((long)fill.size() == 8L * (nx + ny), "nx %d ny %d: %zu ghosts", nx, ny, fill.size());
    // periodic sides copy across the duplicated end node: G[s-e] = U[end-1-e]
    if (bc.left == kBcPeriod) {
        const GhostRef r = ghost_map(g, bc, 0);  // row 4, e1 = 1
        CHECK(r.di == 3 && r.si == nx + 2, "left periodic: ghost 3 copies %d", r.si);
    }
    if (bc.bottom == kBcPeriod) {
        const GhostRef r = ghost_map(g, bc, 8 * ny);  // column 4, e1 = 1
        CHECK(r.dj == 3 && r.sj == ny + 2, "bottom periodic: ghost 3 copies %d", r.sj);
    }
    // the inverted map, enumerated over every interior point
    std::map<std::pair<int, int>, Write> inv;
    for (int j = 4; j < ny + 4; ++j)
        for (int i = 4; i < nx + 4; ++i) {
            bool any_x = false, any_y = false;
            ghost_targets(
                g, bc, i, j, true, true,
                [&](int di, int dj, int flip) {
                    const auto key = std::make_pair(di, dj);
                    CHECK(!inv.count(key), "nx %d ny %d: inverted map writes (%d, %d) twice", nx, ny, di, dj);
                    inv[key] = Write{di, dj, i, j, flip, -1};
                    (dj == j ? any_x : any_y) = true;
                },
                [&](int di, int dj, int spec) {
                    const auto key = std::make_pair(di, dj);
                    CHECK(!inv.count(key), "nx %d ny %d: inverted map writes (%d, %d) twice", nx, ny, di, dj);
                    inv[key] = Write{di, dj, -1, -1, -1, spec};
                    (dj == j ? any_x : any_y) = true;
                });
            if (any_x) CHECK(ghost_x_edge(g, i, i), "nx %d: cell %d has x ghosts but the gate is closed", nx, i);
            if (any_y) CHECK(ghost_y_edge(g, j), "ny %d: row %d has y ghosts but the gate is closed", ny, j);
        }
    CHECK(inv == fill, "nx %d ny %d bc %d %d %d: inverted map differs from the fill", nx, ny, bc.left, bc.right,
          bc.bottom);
}

void check_xcd() {
    for (int n = 1; n <= 5000; ++n) {
        std::vector<char> seen(n, 0);
        for (int b = 0; b < n; ++b) {
            const int l = xcd_block(b, n);
            CHECK(l >= 0 && l < n && !seen[l], "xcd_block(%d, %d) = %d", b, n, l);
            seen[l] = 1;
        }
    }
}

template <int XT>
void check_x_tiles(int nx, int ny) {
    const Geom g = geom(nx, ny);
    constexpr int XC = x_cells<XT>(), XP = XC + 6, XS = XT + 8;
    const int ntx = x_tiles_x<XT>(g);
    std::vector<int> cover(nx + 8, 0);
    for (int bx = 0; bx < ntx; ++bx) {
        const int f0 = x_tile_first_face<XT>(bx);
        std::vector<char> loaded(XS, 0);
        for (int q = 0; q < XP; q += 1) {  // the split loop: q = t, t + XT, ...
            const int p = f0 - 3 + q;
            if (p < nx + 8) {
                CHECK(p >= 0 && q < XS, "XT %d nx %d tile %d: point %d slot %d", XT, nx, bx, p, q);
                loaded[q] = 1;
            }
        }
        for (int t = 0; t < XT; ++t) {
            const int f = f0 + t;
            if (t <= XC && f <= nx + 4) {  // a face: slots t .. t+5 (split), t+2, t+3 (points)
                for (int s = 0; s <= 5; ++s)
                    CHECK(t + s < XS && loaded[t + s], "XT %d nx %d tile %d face %d: slot %d not loaded", XT, nx, bx,
                          f, t + s);
                CHECK(t < XT, "face slot");
            }
            const int i = f0 + t;
            if (t < XC && i < nx + 4) {  // a cell: faces t, t+1; Fy rows j-4, j-3; U at (i, j)
                CHECK(t + 1 <= XC && i + 1 <= nx + 4, "XT %d nx %d: cell %d's right face missing", XT, nx, i);
                cover[i] += 1;
                for (int j = 4; j < ny + 4; j += (ny > 8 ? ny - 1 : 1)) {
                    const size_t oy = (size_t)(j - 4) * g.fy_pitch + (i - 4);
                    CHECK(oy + g.fy_pitch < (size_t)g.fy_plane && (i - 4) < g.fy_pitch, "Fy index");
                    CHECK((size_t)j * g.pitch + i < (size_t)g.plane, "U index");
                }
                // the gate of the wave holding this cell admits it when it has x ghosts
                const int wf = f0 + (t & ~63);
                bool any = false;
                ghost_targets(
                    g, Bc{kBcSymmetric, kBcSymmetric, kBcSymmetric}, i, 4 + ny / 2, true, false,
                    [&](int, int, int) { any = true; }, [&](int, int, int) { any = true; });
                if (any) CHECK(ghost_x_edge(g, wf, wf + 63), "XT %d nx %d: wave at %d misses cell %d", XT, nx, wf, i);
            }
        }
    }
    for (int i = 4; i < nx + 4; ++i) CHECK(cover[i] == 1, "XT %d nx %d: cell %d covered %d times", XT, nx, i, cover[i]);
}

template <int YF>
void check_y_tiles(int ny) {
    const Geom g = geom(64, ny);
    const int nty = y_tiles_y<YF>(g);
    std::vector<int> cover(ny + 8, 0);
    for (int by = 0; by < nty; ++by) {
        const int F0 = y_tile_first_face<YF>(by), F1 = y_tile_end_face<YF>(g, by);
        CHECK(F0 < F1 || F0 > ny + 4, "YF %d ny %d: empty tile %d", YF, ny, by);
        const int base = F0 - 3;
        std::vector<int> split(kRing, -1000), point(kRawRing, -1000);
        // one write phase: the rows it loads (checked for slot clashes), then applied
        auto write_phase = [&](const std::vector<int>& rows) {
            std::set<int> ss, ps;
            for (int prow : rows) {
                if (prow >= ny + 8) continue;
                CHECK(prow >= 0, "YF %d ny %d: row %d", YF, ny, prow);
                const int s = (prow - base) % kRing;
                CHECK(!ss.count(s), "YF %d ny %d tile %d: split slot %d written twice in a phase", YF, ny, by, s);
                ss.insert(s);
                split[s] = prow;
                if (y_ring_keeps_point(F0, prow)) {
                    const int q = (prow - base) % kRawRing;
                    CHECK(!ps.count(q), "YF %d ny %d tile %d: point slot %d written twice in a phase (row %d)", YF,
                          ny, by, q, prow);
                    ps.insert(q);
                    point[q] = prow;
                }
            }
        };
        std::vector<int> rows;
        for (int k = 0; k < 5; ++k) rows.push_back(base + k);  // lead-in, same phase as the first step's loads
        for (int f = F0; f < F1; f += kYR) {
            for (int ty = 0; ty < kYR; ++ty) rows.push_back(f + 2 + ty);
            write_phase(rows);
            rows.clear();
            for (int ty = 0; ty < kYR; ++ty) {
                const int fr = f + ty;
                if (fr >= F1) continue;
                cover[fr] += 1;
                for (int s = 0; s < 6; ++s)
                    CHECK(split[(fr - 3 + s - base) % kRing] == fr - 3 + s,
                          "YF %d ny %d tile %d face %d: split slot holds row %d, not %d", YF, ny, by, fr,
                          split[(fr - 3 + s - base) % kRing], fr - 3 + s);
                for (int s = -1; s <= 0; ++s)
                    CHECK(point[(fr + s - base) % kRawRing] == fr + s,
                          "YF %d ny %d tile %d face %d: point slot holds row %d, not %d", YF, ny, by, fr,
                          point[(fr + s - base) % kRawRing], fr + s);
                CHECK((size_t)(fr - 4) * g.fy_pitch + 63 < (size_t)g.fy_plane, "Fy row %d", fr);
            }
        }
    }
    for (int f = 4; f <= ny + 4; ++f) CHECK(cover[f] == 1, "YF %d ny %d: face %d covered %d times", YF, ny, f, cover[f]);
}

}  // namespace

int main() {
    const int kinds[4] = {kBcPeriod, kBcSlip, kBcSpecial, kBcSymmetric};
    long cases = 0;
    auto all_bcs = [&](int nx, int ny) {
        for (int l : kinds)
            for (int r : kinds)
                for (int b : kinds) {
                    check_ghosts(nx, ny, Bc{l, r, b});
                    ++cases;
                }
    };
    // the library accepts grids of 5 to 65536 points a side
    for (int nx = 5; nx <= 40; ++nx)
        for (int ny = 5; ny <= 40; ++ny) all_bcs(nx, ny);
    for (int nx : {63, 64, 65, 100, 127, 128, 129, 255, 256, 257, 300})
        for (int ny : {5, 64, 129, 300}) all_bcs(nx, ny);
    check_xcd();
    for (int nx = 5; nx <= 300; ++nx)
        for (int ny : {5, 17, 300}) {
            check_x_tiles<64>(nx, ny);
            check_x_tiles<128>(nx, ny);
            check_x_tiles<256>(nx, ny);
            ++cases;
        }
    for (int ny = 5; ny <= 300; ++ny) {
        check_y_tiles<16>(ny);
        check_y_tiles<32>(ny);
        check_y_tiles<64>(ny);
        ++cases;
    }
    std::printf("index_check: %ld cases, %ld checks passed\n", cases, g_checks);
    return 0;
}
