// wenofd_capi.cpp — the C ABI (include/wenofd_amd.h): solver handles, device
// buffers, the RK stage sequence and the run loop, outputs.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>
#include <algorithm>

#include "wenofd_amd.h"
#include "wenofd_devloop.hpp"
#include "wenofd_host.hpp"
#include "wenofd_kernels.hpp"
#include "wenofd_mlp.hpp"
#include "wenofd_switch.hpp"

using namespace wenofd;

static thread_local std::string g_last_error;

int wenofd::set_last_error(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

static int fail(int code, const std::string& msg) { return set_last_error(code, msg); }

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail(WENOFD_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));       \
    } while (0)

namespace {

constexpr int kDtBlocks = 1024;

struct TimedLaunch {
    int kind;
    hipEvent_t start, stop;
};

}  // namespace

struct wenofd_solver {
    Options opt;
    CaseDef cs;
    Geom g{};
    Bc bc{};
    Tiles tiles{};
    int device = 0;
    int cus = 0;  // compute units of the device
    hipStream_t stream = nullptr;
    // Three state buffers (4 planes incl. ghosts each): U is the current
    // state; Ua and Ub hold intermediate RK stages. Stages never copy: RK3 runs
    // U -> Ua -> Ub -> U with U doubling as the RK base state Un.
    double* U = nullptr;
    double* Ua = nullptr;
    double* Ub = nullptr;
    double* Fy = nullptr;       // y-face fluxes
    double* rhs = nullptr;      // rhs in the state layout (allocated on first get_flux)
    double* special = nullptr;  // SPECIAL boundary values [4][2 nx + 2 ny]
    double* dt_part = nullptr;  // partial CFL minima (standalone kernel or last RK stage)
    int* dt_part_nan = nullptr;
    int dt_parts = 0;           // > 0: the partials describe the current state
    double* dt_dev = nullptr;   // {min, nan flag}
    double* dt_host = nullptr;  // pinned copy of dt_dev
    GridSetup grid;
    double now = 0;
    long long steps = 0;
    std::string out_dir;
    bool timing = false;
    std::vector<TimedLaunch> events;
    std::vector<hipEvent_t> event_pool;
    double t_ms[WENOFD_K_COUNT] = {};
    long long t_cnt[WENOFD_K_COUNT] = {};
    // Learned switch: mode, MLP parameters (device), threshold, raw and
    // dilated flags per direction (x faces [ny][nx+1], y faces [ny+1][nx]).
    int sw_mode = WENOFD_SWITCH_OFF;
    float sw_thr = 0.0f;
    int sw_every = 1;           // classify every sw_every steps
    int sw_rad = -1;            // dilation radius (-1: sw_every)
    long long sw_since = 0;     // switched steps since the last classification (0: classify now)
    float* sw_params = nullptr;
    unsigned char* fl_raw_x = nullptr;
    unsigned char* fl_raw_y = nullptr;
    unsigned char* fl_x = nullptr;
    unsigned char* fl_y = nullptr;
    // The faces of the dilated masks that flag any field (rebuilt whenever the
    // masks change): the list (grouped by flagged-field count), the summed
    // corrections of the listed faces at each face's index (4 doubles per
    // face, written per stage), the face counts on the device and the list
    // build's scratch.
    unsigned* ent_x = nullptr;
    unsigned* ent_y = nullptr;
    double* corr_x = nullptr;
    double* corr_y = nullptr;
    int* n_pairs = nullptr;  // {x, y}
    int* list_cnt = nullptr;
    int* list_off = nullptr;
    // After a switched step the state before it survives in *prev_slot (one of
    // Ua / Ub) until the next step; the plain solver redoes the step from it
    // when the switched step produced a NaN time step.
    double** prev_slot = nullptr;
    double prev_dt = 0;
    long long fallbacks = 0;
    // Device-side run loop for wenofd_advance (time step folded and applied on
    // the device, steps replayed from a HIP graph; see wenofd_devloop.hpp).
    DevLoop loop;
    bool loop_on = true;
    bool loop_ready = false;
    std::vector<SolveDesc> loop_desc;
    SolveDesc* loop_desc_d = nullptr;
    SolveState* loop_st_d = nullptr;
    SolveState* loop_st_h = nullptr;  // pinned
    int* loop_active_d = nullptr;
    int* loop_active_h = nullptr;  // pinned
    size_t part_cap = 0;

    size_t state_doubles() const { return 4 * (size_t)g.plane; }
    long faces_x() const { return (long)g.ny * (g.nx + 1); }
    long faces_y() const { return (long)(g.ny + 1) * g.nx; }
};

namespace {

hipEvent_t take_event(wenofd_solver* h) {
    if (!h->event_pool.empty()) {
        hipEvent_t e = h->event_pool.back();
        h->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Brackets one launch with events when timing is enabled.
template <class F>
int timed(wenofd_solver* h, int kind, F&& launch) {
    hipEvent_t a = nullptr, b = nullptr;
    if (h->timing) {
        a = take_event(h);
        b = take_event(h);
        if (a && b) HIP_TRY(hipEventRecord(a, h->stream));
    }
    hipError_t e = launch();
    if (e != hipSuccess) return fail(WENOFD_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    if (h->timing && a && b) {
        HIP_TRY(hipEventRecord(b, h->stream));
        h->events.push_back({kind, a, b});
    }
    return 0;
}

int collect_timing(wenofd_solver* h) {
    if (h->events.empty()) return 0;
    HIP_TRY(hipStreamSynchronize(h->stream));
    for (auto& ev : h->events) {
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, ev.start, ev.stop));
        h->t_ms[ev.kind] += ms;
        h->t_cnt[ev.kind] += 1;
        h->event_pool.push_back(ev.start);
        h->event_pool.push_back(ev.stop);
    }
    h->events.clear();
    return 0;
}

int set_dev(wenofd_solver* h) {
    HIP_TRY(hipSetDevice(h->device));
    return 0;
}


// One RK stage: ghost fill of `in` (unless done), y-face fluxes, then x faces
// fused with the rhs and the update into `out` (which also reduces the next
// CFL step when last_stage is set). With the switch on, one hybrid stage
// kernel does all of it with the field masks fl_x / fl_y.
int stage(wenofd_solver* h, double* in, const double* un, double* out, const RkCoef& k, double* rhs_out,
          bool last_stage, bool ghost = true) {
    int rc;
    const bool sw = h->sw_mode != WENOFD_SWITCH_OFF;
    if (ghost &&
        (rc = timed(h, WENOFD_K_GHOST, [&] { return launch_ghost_fill(in, h->special, h->g, h->bc, h->stream); })))
        return rc;
    if (sw) {
        const int fp = h->opt.fp_contract;
        const int segs = hyb_stage_segments(h->g, h->cus * 4 * hyb_waves_per_simd());
        if (hyb_stage_blocks(h->g, segs) > (long)h->part_cap)
            return fail(WENOFD_E_HIP, "hybrid stage: more blocks than CFL partial slots");
        if ((rc = timed(h, WENOFD_K_PATCH, [&] {
                 return WENOFD_FP(fp, launch_hyb_patch(h->opt.scheme, in, h->g, h->ent_x, h->ent_y, h->fl_x,
                                                       h->fl_y, h->n_pairs, h->corr_x, h->corr_y, h->stream));
             })))
            return rc;
        if ((rc = timed(h, WENOFD_K_HYB, [&] {
                 return WENOFD_FP(fp, launch_hyb_stage(in, un, out, rhs_out, h->g, k, h->opt.cfl,
                                                       last_stage ? h->dt_part : nullptr,
                                                       last_stage ? h->dt_part_nan : nullptr, h->fl_x, h->fl_y,
                                                       h->corr_x, h->corr_y, segs, h->stream));
             })))
            return rc;
        if (last_stage) h->dt_parts = (int)hyb_stage_blocks(h->g, segs);
        return 0;
    }
    if ((rc = timed(h, WENOFD_K_FLUX_Y, [&] {
             return WENOFD_FP(h->opt.fp_contract, launch_flux_y(h->opt.scheme, h->tiles, in, h->Fy, h->g, h->stream));
         })))
        return rc;
    if ((rc = timed(h, WENOFD_K_FLUX_X, [&] {
             return WENOFD_FP(h->opt.fp_contract,
                              launch_flux_x_rk(h->opt.scheme, h->tiles, in, h->Fy, un, out, rhs_out, h->g, k,
                                               h->opt.cfl, last_stage ? h->dt_part : nullptr,
                                               last_stage ? h->dt_part_nan : nullptr, h->stream));
         })))
        return rc;
    if (last_stage) h->dt_parts = (int)flux_x_rk_blocks(h->g, h->tiles);
    return 0;
}

// Ghost fill of U, then (MLP mode) the field masks of every face from the
// network, each field's flags dilated by the radius (default sw_every faces)
// in both directions of the face grid: the masks are computed from the
// initial state of every sw_every-th step and serve all stages of that many
// steps (a wave moves less than one cell per step at CFL <= 1).
// The face lists of the current dilated masks fl_x / fl_y.
hipError_t build_lists(wenofd_solver* h) {
    hipError_t e = launch_build_list(h->fl_x, h->faces_x(), h->list_cnt, h->list_off, h->ent_x,
                                     h->n_pairs, h->stream);
    if (e == hipSuccess)
        e = launch_build_list(h->fl_y, h->faces_y(), h->list_cnt, h->list_off, h->ent_y, h->n_pairs + 1,
                              h->stream);
    return e;
}

int classify(wenofd_solver* h, float* logits_x = nullptr, float* logits_y = nullptr, bool dilate = true,
             bool force = true) {
    int rc;
    if ((rc = timed(h, WENOFD_K_GHOST, [&] { return launch_ghost_fill(h->U, h->special, h->g, h->bc, h->stream); })))
        return rc;
    if (h->sw_mode != WENOFD_SWITCH_MLP) return 0;
    if (!force && h->sw_since > 0 && h->sw_since < h->sw_every) return 0;
    h->sw_since = 0;
    const int rad = h->sw_rad >= 0 ? h->sw_rad : std::min(h->sw_every, 16);
    return timed(h, WENOFD_K_SWITCH, [&] {
        hipError_t e = launch_classify(0, h->U, h->g, h->sw_params, h->sw_thr, h->fl_raw_x, logits_x, h->stream);
        if (e == hipSuccess)
            e = launch_classify(1, h->U, h->g, h->sw_params, h->sw_thr, h->fl_raw_y, logits_y, h->stream);
        if (e == hipSuccess && dilate)
            e = launch_dilate(h->fl_raw_x, h->fl_x, h->g.ny, h->g.nx + 1, rad, h->stream);
        if (e == hipSuccess && dilate)
            e = launch_dilate(h->fl_raw_y, h->fl_y, h->g.ny + 1, h->g.nx, rad, h->stream);
        if (e == hipSuccess && dilate) e = build_lists(h);
        return e;
    });
}

// The ghost points an RK stage's x sweep writes into its output (RK2 / RK3):
// the next stage then needs no ghost fill. Off for grids under 5 points a side.
GhostOut stage_ghosts(const wenofd_solver* h) {
    GhostOut gh;
    gh.on = h->g.nx >= 5 && h->g.ny >= 5;
    gh.bc = h->bc;
    gh.special = h->special;
    return gh;
}

// RK methods with the switch on: masks from the step's initial state (or an
// earlier one, every sw_every steps), and the final stage written outside U so
// that the initial state survives the step. With RK2 / RK3 every stage writes
// its output's ghost points, so only the classification's fill of U runs.
int run_rk_switched(wenofd_solver* h, int rkmeth, double dt) {
    const int grav = h->opt.testcase == kRTI ? 1 : 0;
    const GhostOut gh = stage_ghosts(h);
    int rc;
    h->dt_parts = 0;
    h->prev_slot = nullptr;
    if ((rc = classify(h, nullptr, nullptr, true, false))) return rc;
    h->sw_since += 1;
    if (rkmeth == 0) {
        RkCoef k{kRk1, grav, 0.0, 1.0, dt};
        if ((rc = stage(h, h->U, nullptr, h->Ua, k, nullptr, true, false))) return rc;
        std::swap(h->U, h->Ua);
        h->prev_slot = &h->Ua;
    } else if (rkmeth == 1) {
        RkCoef k1{kRkStage, grav, 1.0, 0.0, 1.0 * dt, gh};
        RkCoef k2{kRkStage, grav, 0.5, 0.5, 0.5 * dt, gh};
        if ((rc = stage(h, h->U, h->U, h->Ua, k1, nullptr, false, false))) return rc;
        if ((rc = stage(h, h->Ua, h->U, h->Ub, k2, nullptr, true, !gh.on))) return rc;
        std::swap(h->U, h->Ub);
        h->prev_slot = &h->Ub;
    } else if (rkmeth == 2) {
        RkCoef k1{kRkStage, grav, 1.0, 0.0, 1.0 * dt, gh};
        RkCoef k2{kRkStage, grav, 3.0 / 4.0, 1.0 / 4.0, (1.0 / 4.0) * dt, gh};
        RkCoef k3{kRkStage, grav, 1.0 / 3.0, 2.0 / 3.0, (2.0 / 3.0) * dt, gh};
        if ((rc = stage(h, h->U, h->U, h->Ua, k1, nullptr, false, false))) return rc;
        if ((rc = stage(h, h->Ua, h->U, h->Ub, k2, nullptr, false, !gh.on))) return rc;
        if ((rc = stage(h, h->Ub, h->U, h->Ua, k3, nullptr, true, !gh.on))) return rc;
        std::swap(h->U, h->Ua);
        h->prev_slot = &h->Ua;
    } else {
        return fail(WENOFD_E_BAD_ENUM, "RKMETH must be 0, 1 or 2");
    }
    h->prev_dt = dt;
    return 0;
}

int run_rk(wenofd_solver* h, int rkmeth, double dt) {
    if (h->sw_mode != WENOFD_SWITCH_OFF) return run_rk_switched(h, rkmeth, dt);
    const int grav = h->opt.testcase == kRTI ? 1 : 0;
    const GhostOut gh = stage_ghosts(h);
    int rc;
    h->dt_parts = 0;
    h->prev_slot = nullptr;
    if (rkmeth == 0) {  // RunRK1, src/WENOFD.cpp:301-311: U += dt rhs
        RkCoef k{kRk1, grav, 0.0, 1.0, dt};
        if ((rc = stage(h, h->U, nullptr, h->Ua, k, nullptr, true))) return rc;
        std::swap(h->U, h->Ua);
        return 0;
    }
    if (rkmeth == 1) {  // RunRK2, :312-351 (Un = U)
        RkCoef k1{kRkStage, grav, 1.0, 0.0, 1.0 * dt, gh};
        RkCoef k2{kRkStage, grav, 0.5, 0.5, 0.5 * dt, gh};
        if ((rc = stage(h, h->U, h->U, h->Ua, k1, nullptr, false))) return rc;
        return stage(h, h->Ua, h->U, h->U, k2, nullptr, true, !gh.on);
    }
    if (rkmeth == 2) {  // RunRK3, :352-396 (Un = U)
        RkCoef k1{kRkStage, grav, 1.0, 0.0, 1.0 * dt, gh};
        RkCoef k2{kRkStage, grav, 3.0 / 4.0, 1.0 / 4.0, (1.0 / 4.0) * dt, gh};
        RkCoef k3{kRkStage, grav, 1.0 / 3.0, 2.0 / 3.0, (2.0 / 3.0) * dt, gh};
        if ((rc = stage(h, h->U, h->U, h->Ua, k1, nullptr, false))) return rc;
        if ((rc = stage(h, h->Ua, h->U, h->Ub, k2, nullptr, false, !gh.on))) return rc;
        return stage(h, h->Ub, h->U, h->U, k3, nullptr, true, !gh.on);
    }
    return fail(WENOFD_E_BAD_ENUM, "RKMETH must be 0, 1 or 2");
}

// calculateDeltaT: folds the partial minima the last RK stage left for the
// current state, or computes them first (after creation or a state upload).
int delta_t(wenofd_solver* h, double* dt) {
    int rc;
    if (h->dt_parts > 0) {
        const int n = h->dt_parts;
        rc = timed(h, WENOFD_K_DT,
                   [&] { return launch_dt_final(h->dt_part, h->dt_part_nan, n, h->dt_dev, h->stream); });
    } else {
        rc = timed(h, WENOFD_K_DT, [&] {
            return WENOFD_FP(h->opt.fp_contract, launch_dt(h->U, h->g, h->opt.cfl, h->dt_part, h->dt_part_nan,
                                                           kDtBlocks, h->dt_dev, h->stream));
        });
    }
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(h->dt_host, h->dt_dev, 2 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    double ts = h->dt_host[0];
    if (h->dt_host[1] != 0.0) {
        *dt = NAN;
        return fail(WENOFD_E_NAN_DT, "Timestep is not a number (non-physical state)");
    }
    if (h->cs.t_end - h->now < ts) ts = h->cs.t_end - h->now;  // :434-435
    *dt = ts;
    return 0;
}

// Host copy of the interior state in the reference order [nx][ny][4].
int download_ref(wenofd_solver* h, std::vector<double>& out) {
    std::vector<double> full(h->state_doubles());
    HIP_TRY(hipMemcpyAsync(full.data(), h->U, full.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    const int nx = h->g.nx, ny = h->g.ny;
    out.resize((size_t)nx * ny * 4);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int r = 0; r < 4; ++r)
                out[((size_t)i * ny + j) * 4 + r] = full[r * h->g.plane + (size_t)(j + 4) * h->g.pitch + (i + 4)];
    return 0;
}

int upload_initial(wenofd_solver* h) {
    std::vector<double> full;
    initial_state(h->cs, h->grid, h->g.pitch, h->g.plane, full);
    for (double* b : {h->U, h->Ua, h->Ub})
        HIP_TRY(hipMemcpyAsync(b, full.data(), full.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->dt_parts = 0;
    h->prev_slot = nullptr;
    h->sw_since = 0;
    h->now = 0;
    h->steps = 0;
    return 0;
}

int create_impl(const std::string& text, int device, wenofd_solver** out) {
    std::map<std::string, std::string> kv;
    parse_config_lines(text, kv);
    Options o;
    std::string err;
    int rc = interpret_options(kv, o, err);
    if (rc) return fail(rc, err);

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(WENOFD_E_HIP, "no HIP device available");
    if (device < 0) HIP_TRY(hipGetDevice(&device));
    if (device >= ndev) return fail(WENOFD_E_ARG, "device index out of range");

    auto* h = new wenofd_solver();
    h->opt = o;
    h->device = device;
    make_case(o, h->cs);
    make_grid(h->cs, o.nx, o.ny, h->grid);
    Geom& g = h->g;
    g.nx = o.nx;
    g.ny = o.ny;
    auto round_up = [](long v, long m) { return (v + m - 1) / m * m; };
    g.pitch = (int)round_up(g.nx + 8, 32);
    g.plane = (long)g.pitch * (g.ny + 8);
    g.fy_pitch = (int)round_up(g.nx, 32);
    g.fy_plane = (long)g.fy_pitch * (g.ny + 1);
    g.dx = h->grid.dx;
    g.dy = h->grid.dy;
    g.rdx = 1.0 / g.dx;
    g.rdy = 1.0 / g.dy;
    g.gamma = h->cs.gamma;
    g.zp_lambda = h->grid.zp_lambda;
    h->bc = Bc{h->cs.bc_left, h->cs.bc_right, h->cs.bc_bottom};

    auto cleanup_fail = [&](int code) {
        wenofd_destroy(h);
        return code;
    };
    if ((rc = set_dev(h))) return cleanup_fail(rc);
    hipError_t e = hipDeviceGetAttribute(&h->cus, hipDeviceAttributeMultiprocessorCount, h->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&h->U, h->state_doubles() * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&h->Ua, h->state_doubles() * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&h->Ub, h->state_doubles() * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&h->Fy, 4 * (size_t)g.fy_plane * sizeof(double));
    const int ns = 2 * (g.nx + g.ny);
    if (e == hipSuccess) e = hipMalloc(&h->special, 4 * (size_t)ns * sizeof(double));
    // room for the partials of the narrowest x tile
    const size_t nparts = std::max<size_t>(kDtBlocks, (size_t)flux_x_rk_blocks(g, Tiles{64, 64}));
    h->part_cap = nparts;
    if (e == hipSuccess) e = hipMalloc(&h->dt_part, nparts * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&h->dt_part_nan, nparts * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&h->dt_dev, 2 * sizeof(double));
    if (e == hipSuccess) e = hipHostMalloc(&h->dt_host, 2 * sizeof(double), hipHostMallocDefault);
    if (e != hipSuccess) return cleanup_fail(fail(WENOFD_E_HIP, std::string("allocation: ") + hipGetErrorString(e)));

    std::vector<double> sp;
    special_band(h->cs, h->grid, sp);
    e = hipMemcpy(h->special, sp.data(), sp.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemsetAsync(h->Fy, 0, 4 * (size_t)g.fy_plane * sizeof(double), h->stream);
    if (e != hipSuccess) return cleanup_fail(fail(WENOFD_E_HIP, std::string("upload: ") + hipGetErrorString(e)));
    if ((rc = upload_initial(h))) return cleanup_fail(rc);
    *out = h;
    return 0;
}

std::string timestamp() {
    // utils/timer.hpp:81-88
    auto now = std::chrono::system_clock::now();
    std::time_t t = std::chrono::system_clock::to_time_t(now);
    std::tm tmv;
    localtime_r(&t, &tmv);
    char buf[32];
    std::strftime(buf, sizeof buf, "%Y%m%d_%H%M%S", &tmv);
    return buf;
}

int mkdirs(const std::string& path) {
    std::string cur;
    for (size_t k = 0; k < path.size(); ++k) {
        cur += path[k];
        if (path[k] == '/' || k + 1 == path.size()) {
            if (cur == "/" || cur == "./" || cur == ".") continue;
            struct stat st;
            if (stat(cur.c_str(), &st) != 0 && mkdir(cur.c_str(), 0755) != 0) return -1;
        }
    }
    return 0;
}

int ensure_out_dir(wenofd_solver* h) {
    if (h->out_dir.empty()) {
        // src/WENOFD.cpp:83-86
        h->out_dir = h->opt.output_dir + timestamp() + "_" + kCaseNames[h->opt.testcase] + "_" +
                     kSchemeNames[h->opt.scheme] + "_" + std::to_string(h->g.nx) + "x" + std::to_string(h->g.ny) +
                     "/";
    }
    if (mkdirs(h->out_dir)) return fail(WENOFD_E_IO, "cannot create " + h->out_dir);
    return 0;
}

BatchView loop_view(wenofd_solver* h) {
    BatchView v;
    v.desc = h->loop_desc_d;
    v.st = h->loop_st_d;
    v.U[0] = h->U;
    v.U[1] = h->Ua;
    v.U[2] = h->Ub;
    v.Fy = h->Fy;
    v.special = h->special;
    v.part = h->dt_part;
    v.part_nan = h->dt_part_nan;
    v.active_count = h->loop_active_d;
    v.nsolves = 1;
    return v;
}

int loop_setup(wenofd_solver* h) {
    if (h->loop_ready && h->loop.tiles.x == h->tiles.x && h->loop.tiles.y == h->tiles.y) return 0;
    if (!h->loop_desc_d) {
        HIP_TRY(hipMalloc(&h->loop_desc_d, sizeof(SolveDesc)));
        HIP_TRY(hipMalloc(&h->loop_st_d, sizeof(SolveState)));
        HIP_TRY(hipMalloc(&h->loop_active_d, sizeof(int)));
        HIP_TRY(hipHostMalloc(&h->loop_st_h, sizeof(SolveState), hipHostMallocDefault));
        HIP_TRY(hipHostMalloc(&h->loop_active_h, sizeof(int), hipHostMallocDefault));
    }
    SolveDesc d{};
    d.g = h->g;
    d.bc = h->bc;
    d.gravity = h->opt.testcase == kRTI ? 1 : 0;
    d.cfl = h->opt.cfl;
    d.t_end = h->cs.t_end;
    d.part_off = 0;
    d.nparts = (int)flux_x_rk_blocks(h->g, h->tiles);  // the partials of the x sweep's last stage
    h->loop_desc.assign(1, d);
    HIP_TRY(hipMemcpy(h->loop_desc_d, &d, sizeof(SolveDesc), hipMemcpyHostToDevice));
    devloop_drop_graph(h->loop);
    h->loop.tiles = h->tiles;
    h->loop.fp = h->opt.fp_contract;
    h->loop_ready = true;
    return 0;
}

// One iteration of the run() loop on the device (src/WENOFD.cpp:213-239):
// k_batch_dt folds the last stage's partials into the step (clipped to t_end,
// the loop test |now - t_end| > 1e-9), the stages run the plain kernels with
// the step read from device memory (nothing runs once the solve is done), and
// k_batch_advance moves the clock. Same kernels and coefficients as run_rk.
// With RK2 / RK3 the stages' x sweeps leave their outputs ghost-filled, so
// only the first step of a launch sequence (first) fills the ghosts of U.
int enqueue_device_step(wenofd_solver* h, const BatchView& v, bool first) {
    const SolveState* st = h->loop_st_d;
    const int grav = h->opt.testcase == kRTI ? 1 : 0;
    const GhostOut gh = stage_ghosts(h);
    auto chk = [](hipError_t e) {
        return e == hipSuccess ? 0 : fail(WENOFD_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    };
    auto stage_dev = [&](double* in, const double* un, double* out, RkCoef k, bool last, bool fill) -> int {
        int rc;
        if (h->opt.rkmeth != 0) k.gh = gh;
        if ((fill || !k.gh.on) && (rc = chk(launch_ghost_fill(in, h->special, h->g, h->bc, h->stream, st))))
            return rc;
        const int fp = h->opt.fp_contract;
        if ((rc = chk(WENOFD_FP(fp, launch_flux_y(h->opt.scheme, h->tiles, in, h->Fy, h->g, h->stream, st)))))
            return rc;
        return chk(WENOFD_FP(fp, launch_flux_x_rk(h->opt.scheme, h->tiles, in, h->Fy, un, out, nullptr, h->g, k,
                                                  h->opt.cfl, last ? h->dt_part : nullptr,
                                                  last ? h->dt_part_nan : nullptr, h->stream, st)));
    };
    int rc;
    if ((rc = chk(launch_batch_dt(v, h->stream)))) return rc;
    // RkCoef.cdt carries the stage coefficient c; the kernels multiply by dt
    if (h->opt.rkmeth == 0) {
        if ((rc = stage_dev(h->U, nullptr, h->Ua, RkCoef{kRk1, grav, 0.0, 1.0, 1.0}, true, true))) return rc;
        if ((rc = chk(launch_copy_interior(h->Ua, h->U, h->g, st, h->stream)))) return rc;
    } else if (h->opt.rkmeth == 1) {
        if ((rc = stage_dev(h->U, h->U, h->Ua, RkCoef{kRkStage, grav, 1.0, 0.0, 1.0}, false, first))) return rc;
        if ((rc = stage_dev(h->Ua, h->U, h->U, RkCoef{kRkStage, grav, 0.5, 0.5, 0.5}, true, false))) return rc;
    } else {
        if ((rc = stage_dev(h->U, h->U, h->Ua, RkCoef{kRkStage, grav, 1.0, 0.0, 1.0}, false, first))) return rc;
        if ((rc = stage_dev(h->Ua, h->U, h->Ub, RkCoef{kRkStage, grav, 3.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0}, false,
                            false)))
            return rc;
        if ((rc = stage_dev(h->Ub, h->U, h->U, RkCoef{kRkStage, grav, 1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}, true,
                            false)))
            return rc;
    }
    return chk(launch_batch_advance(v, false, h->stream));
}

// Everything enqueue_device_step's launches take by value: the buffers (also
// those outside the view: U, Ua, Ub, Fy, SPECIAL band, partials, time state,
// descriptor, active count), the geometry, boundaries, coefficients and
// kernel variants.
GraphKey device_loop_key(const wenofd_solver* h, const BatchView& v) {
    GraphKey k = devloop_key(h->loop, v);
    k.add(h->g);
    for (const void* p : {(const void*)h->U, (const void*)h->Ua, (const void*)h->Ub, (const void*)h->Fy,
                          (const void*)h->special, (const void*)h->dt_part, (const void*)h->dt_part_nan,
                          (const void*)h->loop_st_d})
        k.add(p);
    for (long long n : {(long long)h->bc.left, (long long)h->bc.right, (long long)h->bc.bottom,
                        (long long)h->tiles.x, (long long)h->tiles.y, (long long)h->opt.scheme,
                        (long long)h->opt.rkmeth, (long long)h->opt.fp_contract,
                        (long long)(h->opt.testcase == kRTI ? 1 : 0), (long long)stage_ghosts(h).on})
        k.add(n);
    k.add(h->opt.cfl);
    return k;
}

int enqueue_device_steps(wenofd_solver* h, const BatchView& v, int n) {
    DevLoop& L = h->loop;
    int rc;
    if (L.graph_steps > 0 && n >= L.graph_steps) {
        const GraphKey key = device_loop_key(h, v);
        if (L.exec && !(L.key == key)) devloop_drop_graph(L);  // the graph holds what it was captured with
        if (!L.exec) {
            HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
            rc = 0;
            for (int k = 0; k < L.graph_steps && !rc; ++k) rc = enqueue_device_step(h, v, k == 0);
            hipGraph_t graph = nullptr;
            const hipError_t e = hipStreamEndCapture(h->stream, &graph);
            if (rc) {
                if (graph) (void)hipGraphDestroy(graph);
                return rc;
            }
            if (e != hipSuccess) return fail(WENOFD_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
            const hipError_t ei = hipGraphInstantiateWithFlags(&L.exec, graph, 0);
            (void)hipGraphDestroy(graph);
            if (ei != hipSuccess) {
                L.exec = nullptr;
                return fail(WENOFD_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(ei));
            }
            L.key = key;
        }
        for (; n >= L.graph_steps; n -= L.graph_steps) HIP_TRY(hipGraphLaunch(L.exec, h->stream));
    }
    for (bool first = true; n > 0; --n, first = false)
        if ((rc = enqueue_device_step(h, v, first))) return rc;
    return 0;
}

// wenofd_advance on the device loop: the same steps as the host loop (the
// batch kernels run the same tile functions, the time step is the same exact
// minimum), without a host round trip per step.
int advance_on_device(wenofd_solver* h, int max_steps, int* steps_done) {
    int rc;
    *steps_done = 0;
    if ((rc = loop_setup(h))) return rc;
    const BatchView v = loop_view(h);
    *h->loop_st_h = SolveState{h->now, 0.0, h->steps, 0, 0};
    HIP_TRY(hipMemcpyAsync(h->loop_st_d, h->loop_st_h, sizeof(SolveState), hipMemcpyHostToDevice, h->stream));
    if (h->dt_parts > 0) {  // the last stage's partials describe the current state
        HIP_TRY(launch_batch_advance(v, true, h->stream));
    } else if ((rc = devloop_prime(v, h->opt.fp_contract, h->stream))) {
        return rc;
    }
    const long long steps0 = h->steps;
    int done = 0, active = 0;
    auto read_active = [&]() -> int {
        HIP_TRY(hipMemcpyAsync(h->loop_active_h, h->loop_active_d, sizeof(int), hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
        active = *h->loop_active_h;
        return 0;
    };
    if ((rc = read_active())) return rc;
    while (active > 0 && (max_steps <= 0 || done < max_steps)) {
        int chunk = h->loop.graph_steps > 0 ? h->loop.graph_steps : 64;
        if (max_steps > 0) chunk = std::min(chunk, max_steps - done);
        if ((rc = enqueue_device_steps(h, v, chunk))) return rc;
        done += chunk;
        if ((rc = read_active())) return rc;
    }
    HIP_TRY(hipMemcpyAsync(h->loop_st_h, h->loop_st_d, sizeof(SolveState), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->now = h->loop_st_h->now;
    h->steps = h->loop_st_h->steps;
    h->dt_parts = h->loop_desc[0].nparts;
    h->prev_slot = nullptr;
    *steps_done = (int)(h->steps - steps0);
    if (h->loop_st_h->nan) return fail(WENOFD_E_NAN_DT, "Timestep is not a number (non-physical state)");
    return 0;
}

}  // namespace

extern "C" {

const char* wenofd_last_error(void) { return g_last_error.c_str(); }

int wenofd_version(void) { return 100; }

int wenofd_parse_config(const char* config_text, wenofd_info* info) {
    if (!config_text || !info) return fail(WENOFD_E_ARG, "null argument");
    std::map<std::string, std::string> kv;
    parse_config_lines(config_text, kv);
    Options o;
    std::string err;
    int rc = interpret_options(kv, o, err);
    if (rc) return fail(rc, err);
    CaseDef c;
    make_case(o, c);
    std::memset(info, 0, sizeof *info);
    info->cfl = o.cfl;
    info->xL = c.xL; info->xR = c.xR; info->yL = c.yL; info->yR = c.yR;
    info->dx = (c.xR - c.xL) / (o.nx - 1);
    info->dy = (c.yR - c.yL) / (o.ny - 1);
    info->t_end = c.t_end;
    info->gamma = c.gamma;
    info->zp_lambda = std::pow(info->dx, 2.0 / 3.0);
    info->nx = o.nx; info->ny = o.ny;
    info->testcase = o.testcase; info->scheme = o.scheme; info->rkmeth = o.rkmeth;
    info->device = -1;
    return 0;
}

int wenofd_create(const char* config_text, int device, wenofd_solver** out) {
    if (!config_text || !out) return fail(WENOFD_E_ARG, "null argument");
    *out = nullptr;
    return create_impl(config_text, device, out);
}

int wenofd_create_from_file(const char* path, int device, wenofd_solver** out) {
    if (!path || !out) return fail(WENOFD_E_ARG, "null argument");
    *out = nullptr;
    std::ifstream in(path);
    if (!in.is_open()) return fail(WENOFD_E_CONFIG, std::string("Failed to open file") + path);  // readConfig.hpp:44-48
    std::stringstream ss;
    ss << in.rdbuf();
    return create_impl(ss.str(), device, out);
}

int wenofd_destroy(wenofd_solver* h) {
    if (!h) return 0;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (auto& ev : h->events) {
        (void)hipEventDestroy(ev.start);
        (void)hipEventDestroy(ev.stop);
    }
    for (auto e : h->event_pool) (void)hipEventDestroy(e);
    (void)hipFree(h->U);
    (void)hipFree(h->Ua);
    (void)hipFree(h->Ub);
    (void)hipFree(h->Fy);
    (void)hipFree(h->rhs);
    (void)hipFree(h->special);
    (void)hipFree(h->dt_part);
    (void)hipFree(h->dt_part_nan);
    (void)hipFree(h->dt_dev);
    devloop_drop_graph(h->loop);
    (void)hipFree(h->loop_desc_d);
    (void)hipFree(h->loop_st_d);
    (void)hipFree(h->loop_active_d);
    if (h->loop_st_h) (void)hipHostFree(h->loop_st_h);
    if (h->loop_active_h) (void)hipHostFree(h->loop_active_h);
    (void)hipFree(h->sw_params);
    (void)hipFree(h->fl_raw_x);
    (void)hipFree(h->fl_raw_y);
    (void)hipFree(h->fl_x);
    (void)hipFree(h->fl_y);
    for (void* p : {(void*)h->ent_x, (void*)h->ent_y, (void*)h->corr_x,
                    (void*)h->corr_y, (void*)h->n_pairs, (void*)h->list_cnt, (void*)h->list_off})
        (void)hipFree(p);
    if (h->dt_host) (void)hipHostFree(h->dt_host);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return 0;
}

int wenofd_initialize_ave(wenofd_solver* h) {
    if (!h) return fail(WENOFD_E_ARG, "null handle");
    int rc = set_dev(h);
    return rc ? rc : upload_initial(h);
}

int wenofd_calculate_delta_t(wenofd_solver* h, double* dt) {
    if (!h || !dt) return fail(WENOFD_E_ARG, "null argument");
    int rc = set_dev(h);
    return rc ? rc : delta_t(h, dt);
}

int wenofd_run_rk(wenofd_solver* h, int rkmeth, double dt) {
    if (!h) return fail(WENOFD_E_ARG, "null handle");
    int rc = set_dev(h);
    return rc ? rc : run_rk(h, rkmeth, dt);
}

int wenofd_set_boundary(wenofd_solver* h) {
    if (!h) return fail(WENOFD_E_ARG, "null handle");
    int rc = set_dev(h);
    return rc ? rc
              : timed(h, WENOFD_K_GHOST, [&] { return launch_ghost_fill(h->U, h->special, h->g, h->bc, h->stream); });
}

// The rhs of the current state into h->rhs: ghost fill, both sweeps and,
// with source, the RTI gravity term fused into the x sweep.
static int assemble(wenofd_solver* h, bool source) {
    int rc = set_dev(h);
    if (rc) return rc;
    if (!h->rhs) {
        HIP_TRY(hipMalloc(&h->rhs, h->state_doubles() * sizeof(double)));
        HIP_TRY(hipMemsetAsync(h->rhs, 0, h->state_doubles() * sizeof(double), h->stream));
    }
    RkCoef k{kRkNone, source && h->opt.testcase == kRTI ? 1 : 0, 0.0, 0.0, 0.0};
    if (h->sw_mode != WENOFD_SWITCH_OFF) {
        if ((rc = classify(h))) return rc;
        return stage(h, h->U, nullptr, nullptr, k, h->rhs, false, false);
    }
    return stage(h, h->U, nullptr, nullptr, k, h->rhs, false);
}

int wenofd_get_flux(wenofd_solver* h) {
    if (!h) return fail(WENOFD_E_ARG, "null handle");
    return assemble(h, true);
}

int wenofd_assemble_boundary_term(wenofd_solver* h) {
    if (!h) return fail(WENOFD_E_ARG, "null handle");
    return assemble(h, false);
}

int wenofd_assemble_source_term(wenofd_solver* h) {
    if (!h) return fail(WENOFD_E_ARG, "null handle");
    if (!h->rhs) return fail(WENOFD_E_ARG, "no rhs: call wenofd_assemble_boundary_term first");
    int rc = set_dev(h);
    if (rc || h->opt.testcase != kRTI) return rc;  // src/WENOFD.cpp:803-820: only RTI has a source
    HIP_TRY(launch_add_source(h->U, h->rhs, h->g, h->stream));
    return 0;
}

static int copy_out(wenofd_solver* h, const double* dev, double* host, int layout) {
    std::vector<double> full(h->state_doubles());
    HIP_TRY(hipMemcpyAsync(full.data(), dev, full.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (!unpack_state(full.data(), h->g.nx, h->g.ny, h->g.pitch, h->g.plane, layout, host))
        return fail(WENOFD_E_ARG, "unknown layout");
    return 0;
}

int wenofd_get_rhs(wenofd_solver* h, double* host, int layout) {
    if (!h || !host) return fail(WENOFD_E_ARG, "null argument");
    if (!h->rhs) return fail(WENOFD_E_ARG, "no rhs: call wenofd_get_flux first");
    if (layout == WENOFD_LAYOUT_REF_GHOSTED) return fail(WENOFD_E_ARG, "rhs has no ghost points");
    int rc = set_dev(h);
    return rc ? rc : copy_out(h, h->rhs, host, layout);
}

int wenofd_get_state(wenofd_solver* h, double* host, int layout) {
    if (!h || !host) return fail(WENOFD_E_ARG, "null argument");
    int rc = set_dev(h);
    return rc ? rc : copy_out(h, h->U, host, layout);
}

int wenofd_set_state(wenofd_solver* h, const double* host, int layout) {
    if (!h || !host) return fail(WENOFD_E_ARG, "null argument");
    int rc = set_dev(h);
    if (rc) return rc;
    std::vector<double> full(h->state_doubles());
    HIP_TRY(hipMemcpyAsync(full.data(), h->U, full.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (!pack_state(host, h->g.nx, h->g.ny, h->g.pitch, h->g.plane, layout, full.data()))
        return fail(WENOFD_E_ARG, "unknown layout");
    HIP_TRY(hipMemcpyAsync(h->U, full.data(), full.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->dt_parts = 0;
    h->prev_slot = nullptr;
    h->sw_since = 0;
    return 0;
}

int wenofd_step(wenofd_solver* h, double dt_fixed, double* dt_used) {
    if (!h) return fail(WENOFD_E_ARG, "null handle");
    int rc = set_dev(h);
    if (rc) return rc;
    double dt = dt_fixed;
    if (!(dt_fixed > 0) && (rc = delta_t(h, &dt))) {
        if (rc != WENOFD_E_NAN_DT || !h->prev_slot) return rc;
        // The last (switched) step left a non-physical state: the plain solver
        // redoes it from the state before it, then the step goes on.
        std::swap(h->U, *h->prev_slot);
        h->prev_slot = nullptr;
        h->now -= h->prev_dt;
        h->steps -= 1;
        const int mode = h->sw_mode;
        h->sw_mode = WENOFD_SWITCH_OFF;
        rc = run_rk(h, h->opt.rkmeth, h->prev_dt);
        h->sw_mode = mode;
        if (rc) return rc;
        h->now += h->prev_dt;
        h->steps += 1;
        h->fallbacks += 1;
        h->sw_since = 0;  // the next switched step classifies afresh
        if ((rc = delta_t(h, &dt))) return rc;
    }
    if ((rc = run_rk(h, h->opt.rkmeth, dt))) return rc;
    h->now += dt;
    h->steps += 1;
    if (dt_used) *dt_used = dt;
    return 0;
}

int wenofd_advance(wenofd_solver* h, int max_steps, double dt_fixed, int* steps_done) {
    if (!h) return fail(WENOFD_E_ARG, "null handle");
    int n = 0, rc = 0;
    if (h->loop_on && !(dt_fixed > 0) && h->sw_mode == WENOFD_SWITCH_OFF && !h->timing) {
        if ((rc = set_dev(h))) return rc;
        rc = advance_on_device(h, max_steps, &n);
        if (steps_done) *steps_done = n;
        return rc;
    }
    while (std::fabs(h->now - h->cs.t_end) > 1e-9) {  // src/WENOFD.cpp:213
        if (max_steps > 0 && n >= max_steps) break;
        if ((rc = wenofd_step(h, dt_fixed, nullptr))) break;
        ++n;
    }
    if (steps_done) *steps_done = n;
    if (rc) return rc;
    if (h->timing) {
        // keep the event list bounded during long timed runs
        if (h->events.size() > 4096 && (rc = collect_timing(h))) return rc;
    }
    return 0;
}

int wenofd_get_info(wenofd_solver* h, wenofd_info* info) {
    if (!h || !info) return fail(WENOFD_E_ARG, "null argument");
    std::memset(info, 0, sizeof *info);
    info->cfl = h->opt.cfl;
    info->dx = h->g.dx; info->dy = h->g.dy;
    info->xL = h->cs.xL; info->xR = h->cs.xR; info->yL = h->cs.yL; info->yR = h->cs.yR;
    info->t_end = h->cs.t_end;
    info->now = h->now;
    info->gamma = h->cs.gamma;
    info->zp_lambda = h->g.zp_lambda;
    info->steps = h->steps;
    info->nx = h->g.nx; info->ny = h->g.ny;
    info->testcase = h->opt.testcase; info->scheme = h->opt.scheme; info->rkmeth = h->opt.rkmeth;
    info->device = h->device;
    return 0;
}

int wenofd_set_time(wenofd_solver* h, double now) {
    if (!h) return fail(WENOFD_E_ARG, "null handle");
    h->now = now;
    return 0;
}

int wenofd_output_dir(wenofd_solver* h, char* buf, int buflen) {
    if (!h || !buf || buflen <= 0) return fail(WENOFD_E_ARG, "null argument");
    int rc = ensure_out_dir(h);
    if (rc) return rc;
    std::snprintf(buf, buflen, "%s", h->out_dir.c_str());
    return 0;
}

int wenofd_output_ave(wenofd_solver* h, const char* prefix) {
    // outputAve, src/WENOFD.cpp:1161-1238: Tecplot ASCII POINT, x outer, y
    // inner, default stream formatting (6 significant digits).
    if (!h || !prefix) return fail(WENOFD_E_ARG, "null argument");
    int rc = set_dev(h);
    if (rc || (rc = ensure_out_dir(h))) return rc;
    std::vector<double> U;
    if ((rc = download_ref(h, U))) return rc;
    const std::string fn = h->out_dir + "solution_" + prefix + ".plt";
    FILE* f = std::fopen(fn.c_str(), "w");
    if (!f) return fail(WENOFD_E_IO, "cannot write " + fn);
    std::fprintf(f, "TITLE=DGSolution\nVARIABLES=X Y rho u v pre c E \nZONE T=TA, I=%d, J=%d, DATAPACKING=POINT\n",
                 h->g.ny, h->g.nx);
    const double gm = h->cs.gamma;
    for (int i = 0; i < h->g.nx; ++i)
        for (int j = 0; j < h->g.ny; ++j) {
            const double* q = &U[((size_t)i * h->g.ny + j) * 4];
            const double rho = q[0], u = q[1] / q[0], v = q[2] / q[0], E = q[3];
            const double pre = (gm - 1.0) * (E - 0.5 * rho * (u * u + v * v));  // getVitalVarVal, equations.hpp:683-698
            const double c = std::sqrt(gm * pre / rho);
            std::fprintf(f, "%g %g %g %g %g %g %g %g \n", h->grid.xc[i + 4], h->grid.yc[j + 4], rho, u, v, pre, c, E);
        }
    std::fclose(f);
    return 0;
}

int wenofd_output_error(wenofd_solver* h, const char* prefix) {
    // outputError, src/WENOFD.cpp:1239-1306
    if (!h || !prefix) return fail(WENOFD_E_ARG, "null argument");
    int rc = set_dev(h);
    if (rc || (rc = ensure_out_dir(h))) return rc;
    std::vector<double> U;
    if ((rc = download_ref(h, U))) return rc;
    const std::string fn = h->out_dir + "error_" + prefix + ".plt";
    FILE* f = std::fopen(fn.c_str(), "w");
    if (!f) return fail(WENOFD_E_IO, "cannot write " + fn);
    std::fprintf(f, "TITLE=DGSolution\nVARIABLES=X Y error\nZONE T=TA, I=%d, J=%d, DATAPACKING = POINT\n", h->g.ny,
                 h->g.nx);
    for (int i = 0; i < h->g.nx; ++i)
        for (int j = 0; j < h->g.ny; ++j) {
            const double ex = h->cs.exact_rho(h->grid.xc[i + 4], h->grid.yc[j + 4], h->now);
            std::fprintf(f, "%g %g %g \n", h->grid.xc[i + 4], h->grid.yc[j + 4],
                         std::fabs(ex - U[((size_t)i * h->g.ny + j) * 4]));
        }
    std::fclose(f);
    return 0;
}

int wenofd_accuracy(wenofd_solver* h, double* norms3) {
    if (!h || !norms3) return fail(WENOFD_E_ARG, "null argument");
    int rc = set_dev(h);
    if (rc) return rc;
    std::vector<double> U;
    if ((rc = download_ref(h, U))) return rc;
    rho_error_norms(h->cs, h->grid, h->now, U, norms3);
    return 0;
}

int wenofd_output_accuracy(wenofd_solver* h, const char* prefix, double* norms3) {
    // outputAccuracy, src/WENOFD.cpp:1308-1382
    if (!h || !prefix) return fail(WENOFD_E_ARG, "null argument");
    double nv[3];
    int rc = wenofd_accuracy(h, nv);
    if (rc || (rc = ensure_out_dir(h))) return rc;
    const std::string fn = h->out_dir + "accuracy_" + prefix + ".csv";
    FILE* f = std::fopen(fn.c_str(), "w");
    if (!f) return fail(WENOFD_E_IO, "cannot write " + fn);
    std::fprintf(f, "elemNum, Linf-norm (rho), L1-norm (rho), L2-norm (rho)\n%d, %20.15e, %20.15e, %20.15e", h->g.nx,
                 nv[0], nv[1], nv[2]);
    std::fclose(f);
    if (norms3) std::memcpy(norms3, nv, sizeof nv);
    return 0;
}

int wenofd_run(wenofd_solver* h, int verbose) {
    // CWENOFD::run, src/WENOFD.cpp:189-299
    if (!h) return fail(WENOFD_E_ARG, "null handle");
    int rc;
    if ((rc = wenofd_output_ave(h, "initial"))) return rc;
    if (verbose) std::printf("Start iteration...\n");
    const auto t0 = std::chrono::steady_clock::now();
    int count = 0;
    while (std::fabs(h->now - h->cs.t_end) > 1e-9) {
        double dt = 0;
        if ((rc = wenofd_step(h, 0.0, &dt))) return rc;
        ++count;
        if (verbose && count % 20 == 1) {
            const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            const double progress = h->now / h->cs.t_end;
            const int bar = 50, pos = (int)(progress * bar);
            std::printf("Iteration count: %d, DeltaT: %.2e, Elapsed time: %.2f seconds, Progress: %.2f/%.2f (%.2f%%) [",
                        count, dt, el, h->now, h->cs.t_end, progress * 100.0);
            for (int k = 0; k < bar; ++k) std::putchar(k < pos ? '=' : (k == pos ? '>' : ' '));
            std::printf("]\n");
            std::fflush(stdout);
        }
    }
    if ((rc = wenofd_output_ave(h, "final"))) return rc;
    if (h->cs.has_exact) {
        if ((rc = wenofd_output_accuracy(h, "final", nullptr))) return rc;
        if ((rc = wenofd_output_error(h, "final"))) return rc;
    }
    return 0;
}

int wenofd_set_timing(wenofd_solver* h, int enable) {
    if (!h) return fail(WENOFD_E_ARG, "null handle");
    h->timing = enable != 0;
    return 0;
}

int wenofd_get_timing(wenofd_solver* h, double* ms, long long* counts, int nkinds) {
    if (!h || !ms || !counts || nkinds <= 0) return fail(WENOFD_E_ARG, "null argument");
    int rc = set_dev(h);
    if (rc || (rc = collect_timing(h))) return rc;
    for (int k = 0; k < nkinds && k < WENOFD_K_COUNT; ++k) {
        ms[k] = h->t_ms[k];
        counts[k] = h->t_cnt[k];
        h->t_ms[k] = 0;
        h->t_cnt[k] = 0;
    }
    return 0;
}

int wenofd_set_device_loop(wenofd_solver* h, int graph_steps) {
    if (!h || graph_steps < 0 || graph_steps > 1024) return fail(WENOFD_E_ARG, "graph steps must be in [0, 1024]");
    int rc = set_dev(h);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    devloop_drop_graph(h->loop);
    h->loop_on = graph_steps > 0;
    h->loop.graph_steps = graph_steps;
    return 0;
}

int wenofd_set_tiles(wenofd_solver* h, int x_faces, int y_faces) {
    if (!h) return fail(WENOFD_E_ARG, "null handle");
    Tiles t;
    t.x = x_faces;
    t.y = y_faces;
    if (!tiles_valid(t)) return fail(WENOFD_E_ARG, "tile shape must be x in {64, 128, 256}, y in {16, 32, 64}");
    h->tiles = t;
    h->dt_parts = 0;  // the partial minima are laid out per x tile
    return 0;
}

int wenofd_synchronize(wenofd_solver* h) {
    if (!h) return fail(WENOFD_E_ARG, "null handle");
    int rc = set_dev(h);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    return 0;
}

int wenofd_switch_set(wenofd_solver* h, int mode, const float* params, float threshold) {
    if (!h) return fail(WENOFD_E_ARG, "null handle");
    if (mode < WENOFD_SWITCH_OFF || mode > WENOFD_SWITCH_WENO) return fail(WENOFD_E_BAD_ENUM, "unknown switch mode");
    if (mode == WENOFD_SWITCH_MLP && !params) return fail(WENOFD_E_ARG, "the MLP switch needs parameters");
    int rc = set_dev(h);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (!h->fl_x) {
        const long fx = h->faces_x(), fy = h->faces_y();
        HIP_TRY(hipMalloc(&h->fl_raw_x, fx));
        HIP_TRY(hipMalloc(&h->fl_raw_y, fy));
        HIP_TRY(hipMalloc(&h->fl_x, fx));
        HIP_TRY(hipMalloc(&h->fl_y, fy));
        // up to every face flagged; corrections at the face's index
        HIP_TRY(hipMalloc(&h->ent_x, (size_t)fx * sizeof(unsigned)));
        HIP_TRY(hipMalloc(&h->ent_y, (size_t)fy * sizeof(unsigned)));
        HIP_TRY(hipMalloc(&h->corr_x, 4 * (size_t)fx * sizeof(double)));
        HIP_TRY(hipMalloc(&h->corr_y, 4 * (size_t)fy * sizeof(double)));
        HIP_TRY(hipMalloc(&h->n_pairs, 2 * sizeof(int)));
        const long nb = std::max(list_blocks(fx), list_blocks(fy));
        HIP_TRY(hipMalloc(&h->list_cnt, 4 * nb * sizeof(int)));
        HIP_TRY(hipMalloc(&h->list_off, 4 * nb * sizeof(int)));
    }
    if (mode == WENOFD_SWITCH_MLP) {
        if (!h->sw_params) HIP_TRY(hipMalloc(&h->sw_params, kNParams * sizeof(float)));
        HIP_TRY(hipMemcpyAsync(h->sw_params, params, kNParams * sizeof(float), hipMemcpyHostToDevice, h->stream));
    } else if (mode != WENOFD_SWITCH_OFF) {
        const unsigned char v = mode == WENOFD_SWITCH_WENO ? 15 : 0;  // every field flagged, or none
        HIP_TRY(launch_fill_u8(h->fl_x, h->faces_x(), v, h->stream));
        HIP_TRY(launch_fill_u8(h->fl_y, h->faces_y(), v, h->stream));
        HIP_TRY(build_lists(h));
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->sw_mode = mode;
    h->sw_thr = threshold;
    h->sw_since = 0;
    h->prev_slot = nullptr;
    return 0;
}

int wenofd_switch_interval(wenofd_solver* h, int steps) {
    if (!h) return fail(WENOFD_E_ARG, "null handle");
    if (steps < 1 || steps > 64) return fail(WENOFD_E_ARG, "the classification interval must be in [1, 64]");
    h->sw_every = steps;
    h->sw_since = 0;
    return 0;
}

int wenofd_switch_dilation(wenofd_solver* h, int radius) {
    if (!h) return fail(WENOFD_E_ARG, "null handle");
    if (radius < -1 || radius > 16) return fail(WENOFD_E_ARG, "the dilation radius must be in [0, 16], or -1");
    h->sw_rad = radius;
    h->sw_since = 0;
    return 0;
}

int wenofd_switch_classify(wenofd_solver* h, int dilate, unsigned char* flags_x, unsigned char* flags_y,
                           float* logits_x, float* logits_y) {
    if (!h) return fail(WENOFD_E_ARG, "null handle");
    if (h->sw_mode != WENOFD_SWITCH_MLP) return fail(WENOFD_E_ARG, "the switch is not in MLP mode");
    int rc = set_dev(h);
    if (rc) return rc;
    float *lx = nullptr, *ly = nullptr;
    if (logits_x) HIP_TRY(hipMalloc(&lx, h->faces_x() * 4 * sizeof(float)));
    if (logits_y) {
        hipError_t e = hipMalloc(&ly, h->faces_y() * 4 * sizeof(float));
        if (e != hipSuccess) {
            (void)hipFree(lx);
            return fail(WENOFD_E_HIP, std::string("allocation: ") + hipGetErrorString(e));
        }
    }
    rc = classify(h, lx, ly, dilate != 0);
    const unsigned char* fx = dilate ? h->fl_x : h->fl_raw_x;
    const unsigned char* fy = dilate ? h->fl_y : h->fl_raw_y;
    hipError_t e = hipSuccess;
    if (!rc && flags_x) e = hipMemcpyAsync(flags_x, fx, h->faces_x(), hipMemcpyDeviceToHost, h->stream);
    if (!rc && e == hipSuccess && flags_y) e = hipMemcpyAsync(flags_y, fy, h->faces_y(), hipMemcpyDeviceToHost, h->stream);
    if (!rc && e == hipSuccess && lx)
        e = hipMemcpyAsync(logits_x, lx, h->faces_x() * 4 * sizeof(float), hipMemcpyDeviceToHost, h->stream);
    if (!rc && e == hipSuccess && ly)
        e = hipMemcpyAsync(logits_y, ly, h->faces_y() * 4 * sizeof(float), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(lx);
    (void)hipFree(ly);
    if (rc) return rc;
    if (e != hipSuccess) return fail(WENOFD_E_HIP, std::string("switch_classify: ") + hipGetErrorString(e));
    return 0;
}

int wenofd_switch_flags(wenofd_solver* h, unsigned char* flags_x, unsigned char* flags_y) {
    if (!h || !flags_x || !flags_y) return fail(WENOFD_E_ARG, "null argument");
    if (!h->fl_x) return fail(WENOFD_E_ARG, "the switch was never enabled");
    int rc = set_dev(h);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(flags_x, h->fl_x, h->faces_x(), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipMemcpyAsync(flags_y, h->fl_y, h->faces_y(), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return 0;
}

int wenofd_switch_stats(wenofd_solver* h, long long* fallbacks) {
    if (!h || !fallbacks) return fail(WENOFD_E_ARG, "null argument");
    *fallbacks = h->fallbacks;
    return 0;
}

int wenofd_harvest(wenofd_solver* h, int dir, float* X, unsigned char* Y) {
    if (!h || !X || !Y || dir < 0 || dir > 1) return fail(WENOFD_E_ARG, "bad argument");
    int rc = set_dev(h);
    if (rc) return rc;
    const long nf = dir == 0 ? h->faces_x() : h->faces_y();
    float* dX = nullptr;
    unsigned char* dY = nullptr;
    HIP_TRY(hipMalloc(&dX, nf * 4 * kFeat * sizeof(float)));
    hipError_t e = hipMalloc(&dY, nf * 4);
    if (e == hipSuccess) e = launch_ghost_fill(h->U, h->special, h->g, h->bc, h->stream);
    if (e == hipSuccess) e = launch_harvest(dir, h->U, h->g, dX, dY, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(X, dX, nf * 4 * kFeat * sizeof(float), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(Y, dY, nf * 4, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(dX);
    (void)hipFree(dY);
    if (e != hipSuccess) return fail(WENOFD_E_HIP, std::string("harvest: ") + hipGetErrorString(e));
    return 0;
}

int wenofd_mlp_harvest(wenofd_mlp* m, wenofd_solver* h, double keep_pos, double keep_neg, unsigned long long seed,
                       long long* added) {
    if (!m || !h) return fail(WENOFD_E_ARG, "null handle");
    if (m->device != h->device) return fail(WENOFD_E_ARG, "trainer and solver live on different devices");
    int rc = set_dev(h);
    if (rc) return rc;
    HIP_TRY(launch_ghost_fill(h->U, h->special, h->g, h->bc, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    long long total = 0;
    for (int dir = 0; dir < 2; ++dir) {
        const long nf = dir == 0 ? h->faces_x() : h->faces_y();
        if ((rc = mlp_reserve_dense(m, nf))) return rc;
        HIP_TRY(launch_harvest(dir, h->U, h->g, m->dense_X, m->dense_Y, m->stream));
        long long got = 0;
        if ((rc = mlp_append_dense(m, nf * 4, seed * 2 + dir, keep_pos, keep_neg, &got))) return rc;
        total += got;
    }
    if (added) *added = total;
    return 0;
}

int wenofd_probe_weno5(int device, int scheme, double zp_lambda, int n, const double* v5, double* out) {
    if (!v5 || !out || n < 0) return fail(WENOFD_E_ARG, "bad argument");
    if (scheme < 0 || scheme > 3) return fail(WENOFD_E_BAD_ENUM, "SCHEME must be 0..3");
    if (n == 0) return 0;
    if (device >= 0) HIP_TRY(hipSetDevice(device));
    double *dv = nullptr, *dout = nullptr;
    HIP_TRY(hipMalloc(&dv, 5 * (size_t)n * sizeof(double)));
    hipError_t e = hipMalloc(&dout, (size_t)n * sizeof(double));
    if (e == hipSuccess) e = hipMemcpy(dv, v5, 5 * (size_t)n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_probe_weno(scheme, dv, n, zp_lambda, dout, nullptr);
    if (e == hipSuccess) e = hipMemcpy(out, dout, (size_t)n * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(dv);
    (void)hipFree(dout);
    if (e != hipSuccess) return fail(WENOFD_E_HIP, std::string("probe_weno5: ") + hipGetErrorString(e));
    return 0;
}

int wenofd_probe_physics(int device, double gamma, int n, const double* states, double* out82) {
    if (!states || !out82 || n < 0) return fail(WENOFD_E_ARG, "bad argument");
    if (n == 0) return 0;
    if (device >= 0) HIP_TRY(hipSetDevice(device));
    double *du = nullptr, *dout = nullptr;
    HIP_TRY(hipMalloc(&du, 4 * (size_t)n * sizeof(double)));
    hipError_t e = hipMalloc(&dout, (size_t)n * kProbePhysStride * sizeof(double));
    if (e == hipSuccess) e = hipMemcpy(du, states, 4 * (size_t)n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_probe_physics(du, n, gamma, dout, nullptr);
    if (e == hipSuccess) e = hipMemcpy(out82, dout, (size_t)n * kProbePhysStride * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(du);
    (void)hipFree(dout);
    if (e != hipSuccess) return fail(WENOFD_E_HIP, std::string("probe_physics: ") + hipGetErrorString(e));
    return 0;
}

}  // extern "C"
