/*
 * wenofd_amd.h — C ABI of the MI355X WENO-FD 2-D Euler solver (libwenofd_amd.so).
 *
 * Drop-in for the hot path of infinity-07/FD-WENO-EulerEquation: the CWENOFD
 * class surface (include/WENOFD.hpp:28-124 in the reference) and its
 * config.cfg keys (utils/readConfig.hpp:39-80, src/WENOFD.cpp:24-40), with the
 * MPI slab decomposition replaced by one whole domain on one GPU.
 *
 * Conventions
 *   - Every call returns an int status: WENOFD_OK (0) or a WENOFD_E_* code;
 *     wenofd_last_error() describes the last failure of the calling thread.
 *     Nothing in the library aborts the process (the reference calls
 *     MPI_Abort / exit / throws, src/WENOFD.cpp:234,341,386,441,843).
 *   - A solver handle owns its device buffers and one HIP stream on one
 *     device; it is not thread-safe. Distinct handles may be used from
 *     distinct threads. Host buffers passed in are caller-owned.
 *   - All solver arithmetic is IEEE FP64; with FP_CONTRACT = 0 (the default)
 *     it is done without contraction and results equal a no-contraction
 *     build of the reference bit for bit (FP_CONTRACT below).
 *   - State layouts for get/set: WENOFD_LAYOUT_REF = interior points in the
 *     reference's Array2D<Cvector> order [NX][NY][4] (x outer, y inner, then
 *     rho, rho u, rho v, E); WENOFD_LAYOUT_SOA = interior [4][NY][NX];
 *     WENOFD_LAYOUT_REF_GHOSTED = [(NX+8)][(NY+8)][4] including the 4 ghost
 *     points per side (corner ghosts are never read by the solver; get
 *     returns them as zero). Ghost points are filled before every RK stage
 *     reads them; between steps they hold a fill of the state (RK2 / RK3) or
 *     of an earlier state (RK1), where the reference's hold the fill of its
 *     last stage's input. wenofd_set_boundary() refills them from the state.
 */
#ifndef WENOFD_AMD_H
#define WENOFD_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wenofd_solver wenofd_solver;

enum {
    WENOFD_OK = 0,
    WENOFD_E_CONFIG = 1,   /* missing/unparsable key or unreadable file */
    WENOFD_E_BAD_ENUM = 2, /* TESTCASE, SCHEME or RKMETH out of range */
    WENOFD_E_NAN_DT = 3,   /* a NaN time step (non-physical state) */
    WENOFD_E_HIP = 4,      /* HIP runtime failure or no usable device */
    WENOFD_E_ARG = 5,      /* invalid argument (null pointer, bad layout, size) */
    WENOFD_E_IO = 6        /* output file could not be written */
};

enum { WENOFD_LAYOUT_REF = 0, WENOFD_LAYOUT_SOA = 1, WENOFD_LAYOUT_REF_GHOSTED = 2 };

/* Timing kinds for wenofd_get_timing(). */
enum {
    WENOFD_K_GHOST = 0,   /* ghost fill */
    WENOFD_K_FLUX_X = 1,  /* x-face fluxes fused with rhs and the RK stage update */
    WENOFD_K_FLUX_Y = 2,  /* y-face fluxes */
    WENOFD_K_SWITCH = 3,  /* switch classification (MLP inference + flag dilation) */
    WENOFD_K_DT = 4,      /* CFL time-step reduction */
    WENOFD_K_HYB = 5,     /* the switch's RK stage (both directions' fluxes, rhs, update) */
    WENOFD_K_PATCH = 6,   /* the switch's corrections of the flagged (face, field) pairs */
    WENOFD_K_COUNT = 7
};

/* Learned troubled-stencil switch modes (wenofd_switch_set). */
enum {
    WENOFD_SWITCH_OFF = 0,    /* the reference's characteristic WENO flux at every face */
    WENOFD_SWITCH_MLP = 1,    /* flags from the switch MLP; unflagged faces take upwind-5 */
    WENOFD_SWITCH_LINEAR = 2, /* every face takes the component-wise upwind-5 flux */
    WENOFD_SWITCH_WENO = 3    /* every face flagged (the hybrid kernels, plain results) */
};

typedef struct wenofd_info {
    double cfl, dx, dy, xL, xR, yL, yR, t_end, now, gamma, zp_lambda;
    long long steps; /* completed time steps (run loop iterations) */
    int nx, ny, testcase, scheme, rkmeth, device;
} wenofd_info;

/* Error text of the calling thread's last failed call ("" if none). */
const char* wenofd_last_error(void);
/* ABI version (major * 100 + minor). */
int wenofd_version(void);

/*
 * CWENOFD::initializeSolver + initializeAve (src/WENOFD.cpp:10-188):
 * config_text holds "KEY = value" lines as in input/config.cfg ('#' comment
 * lines, CR-LF tolerated). Required keys TESTCASE (0 SMOOTH, 1 Riemann1,
 * 2 Riemann2, 3 RTI, 4 isentropic vortex, 5 four-quadrant), SCHEME (0 JS, 1 Z,
 * 2 ZP, 3 ZPI), RKMETH (0 RK1, 1 RK2, 2 RK3), CFL, ELEMNUMX, ELEMNUMY.
 * Optional: TEND, VORTEX_EPS/VORTEX_XC/VORTEX_YC, QUAD_STATES (16 numbers),
 * QUAD_SEED, OUTPUTDIR, FP_CONTRACT. device < 0 uses the current device.
 * FP_CONTRACT (no reference counterpart) selects the FP64 arithmetic of the
 * flux sweeps, the RK update and the CFL step: 0 (default) rounds every
 * product and sum as the reference's source writes it, so results equal the
 * reference built without FMA contraction bit for bit; 1 lets the compiler
 * contract a * b + c into FMAs, as the reference's optimised CPU build does,
 * divides by multiplying with the reciprocal and forms the WENO weights over
 * a common denominator (the same weights, one reciprocal per reconstruction),
 * for fewer instructions (results then agree with either reference build to
 * the few-ulp level of those two builds' own difference).
 */
int wenofd_create(const char* config_text, int device, wenofd_solver** out);
/* CConfig(path).read(option) then wenofd_create (utils/readConfig.hpp:39-80). */
int wenofd_create_from_file(const char* path, int device, wenofd_solver** out);
int wenofd_destroy(wenofd_solver* h);

/* Parses config text without touching a GPU; fills info (nx, ny, enums, cfl,
 * domain, t_end, gamma, dx, dy). Host-only: no device needed. */
int wenofd_parse_config(const char* config_text, wenofd_info* info);

/* CWENOFD::initializeAve (src/WENOFD.cpp:165-188): resets the state to the
 * initial condition and the time to 0. */
int wenofd_initialize_ave(wenofd_solver* h);

/* CWENOFD::calculateDeltaT (src/WENOFD.cpp:398-452): CFL step, clipped to
 * t_end - now. WENOFD_E_NAN_DT if any point gives a NaN. */
int wenofd_calculate_delta_t(wenofd_solver* h, double* dt);

/* CWENOFD::RunRK1/RunRK2/RunRK3 (src/WENOFD.cpp:301-396), rkmeth 0/1/2. Does
 * not advance the time (run() does). */
int wenofd_run_rk(wenofd_solver* h, int rkmeth, double dt);

/* exchangeGhostCellsValue + setBoundary (src/WENOFD.cpp:850-929, 998-1159). */
int wenofd_set_boundary(wenofd_solver* h);

/* getFlux: CWENOFD::assembleRHS = ghost fill + assembleBoundaryTerm +
 * assembleSourceTerm (src/WENOFD.cpp:454-465, 481-821) on the current state;
 * the result is read with wenofd_get_rhs. */
int wenofd_get_flux(wenofd_solver* h);
int wenofd_get_rhs(wenofd_solver* h, double* host, int layout);
/* CWENOFD::assembleBoundaryTerm (src/WENOFD.cpp:481-798): ghost fill and the
 * flux terms of both directions into the rhs (which the reference's only
 * caller, assembleRHS, has zeroed first: the rhs is written, not added to). */
int wenofd_assemble_boundary_term(wenofd_solver* h);
/* CWENOFD::assembleSourceTerm (src/WENOFD.cpp:800-821): adds the RTI gravity
 * source of the current state to the rhs (no other case has one). */
int wenofd_assemble_source_term(wenofd_solver* h);

/* One iteration of the run() loop (src/WENOFD.cpp:213-239): dt =
 * calculateDeltaT (or dt_fixed if > 0), RunRK<RKMETH>(dt), now += dt. */
int wenofd_step(wenofd_solver* h, double dt_fixed, double* dt_used);
/* Repeats wenofd_step while |now - t_end| > 1e-9 and fewer than max_steps
 * (max_steps <= 0: no bound) steps were taken. With CFL time steps, the
 * switch off and timing off, the steps run on the device loop: each time step
 * is folded and applied on the device and the steps are replayed from a HIP
 * graph, with no host round trip per step (same results bit for bit). */
int wenofd_advance(wenofd_solver* h, int max_steps, double dt_fixed, int* steps_done);
/* Steps per graph replay of wenofd_advance's device loop (default 16); 0 runs
 * every step from the host (one dt read-back per step, as the reference's loop). */
int wenofd_set_device_loop(wenofd_solver* h, int graph_steps);
/* CWENOFD::run (src/WENOFD.cpp:189-299) after creation: outputAve("initial"),
 * the loop to t_end with progress lines every 20 steps (verbose != 0), then
 * outputAve("final") and, when an exact solution exists, outputAccuracy and
 * outputError("final"). Files go to <OUTPUTDIR>/<stamp>_<CASE>_<SCHEME>_<NX>x<NY>/. */
int wenofd_run(wenofd_solver* h, int verbose);

int wenofd_get_state(wenofd_solver* h, double* host, int layout);
int wenofd_set_state(wenofd_solver* h, const double* host, int layout);
int wenofd_get_info(wenofd_solver* h, wenofd_info* info);
int wenofd_set_time(wenofd_solver* h, double now);
/* The solve's output directory (created on first output). */
int wenofd_output_dir(wenofd_solver* h, char* buf, int buflen);

/* outputAve / outputError / outputAccuracy (src/WENOFD.cpp:1161-1382). */
int wenofd_output_ave(wenofd_solver* h, const char* prefix);
int wenofd_output_error(wenofd_solver* h, const char* prefix);
int wenofd_output_accuracy(wenofd_solver* h, const char* prefix, double* norms3);
/* L_inf, L1, L2 of rho against the exact solution at the current time,
 * without writing a file (norms3 = {Linf, L1, L2}). */
int wenofd_accuracy(wenofd_solver* h, double* norms3);

/*
 * calculateDeltaT of a caller-owned device state, allocating nothing. No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops). Same kernels
 * and result as wenofd_calculate_delta_t (src/WENOFD.cpp:398-452), written to
 * device memory instead of returned to the host:
 *   U      device state in the library's layout, [4][ny + 8][pitch] doubles
 *          (rho, rho u, rho v, E; interior points i, j in [4, n + 4), x
 *          fastest); only interior points are read;
 *   out    device array of 2 doubles: {dt, nan}, dt = the minimum of 1.0 and
 *          cfl / ((|u| + c) / dx + (|v| + c) / dy) over the interior, clipped
 *          to t_left (t_end - now); nan = 1.0 when any point gives a NaN;
 *   stream a hipStream_t (null: the null stream); the call only enqueues.
 * nx, ny in [5, 65536], pitch >= nx + 8; fp_contract as the FP_CONTRACT key.
 * Scratch: a file-scope device array of 1024 partial minima (12 KiB per GPU,
 * held until the process ends) instead of hipMalloc'ed buffers; every call on
 * one GPU shares it, so calls that overlap on one GPU race: serialise them
 * (one stream, or events).
 */
int wenofd_delta_t_noalloc(const double* U, int nx, int ny, int pitch, double gamma, double dx, double dy,
                           double cfl, double t_left, int fp_contract, double* out, void* stream);

/* Per-kernel timing with HIP events on the solver's stream. While enabled,
 * each launch is bracketed by events; wenofd_get_timing synchronises and
 * returns the summed milliseconds and launch counts per WENOFD_K_* kind, then
 * clears them. */
int wenofd_set_timing(wenofd_solver* h, int enable);
int wenofd_get_timing(wenofd_solver* h, double* ms, long long* counts, int nkinds);
/* Blocks until the solver's stream is idle. */
int wenofd_synchronize(wenofd_solver* h);

/* Tile shapes of the two face sweeps: x_faces faces per block of the fused
 * x-face/rhs/RK kernel (64, 128 or 256), y_faces faces per block of the y-face
 * kernel (16, 32 or 64). Results are bitwise independent of the shape. No
 * counterpart in the reference (a launch-shape knob of this implementation). */
int wenofd_set_tiles(wenofd_solver* h, int x_faces, int y_faces);

/*
 * Learned troubled-stencil switch. No counterpart in the reference. Every
 * `interval` steps the switch MLP (10 -> 32 -> 32 -> 1, ReLU; bf16 MFMA
 * inference, FP32 accumulate) reads the normalised 5-point stencils of each
 * characteristic field of each face and flags the fields whose WENO-Z weights
 * it predicts to leave the linear weights (logit > threshold): a 4-bit mask
 * per face, bit k for field k. The masks are dilated (each field's flags on
 * their own, a square of `radius` faces) and serve every stage until the next
 * classification. Each face's flux is the cheap component-wise linear
 * upwind-5 flux plus, for each flagged field k, R[:,k] (w_k - l_k): the
 * characteristic WENO reconstruction of field k minus its linear one. With
 * every field flagged this is the reference's characteristic WENO flux up to
 * rounding. The plain solver is the fallback: a switched step that leads to a
 * NaN time step is redone without the switch.
 * params: 1441 floats, W1[32][10], b1[32], W2[32][32], b2[32], w3[32], b3.
 */
int wenofd_switch_set(wenofd_solver* h, int mode, const float* params, float threshold);
/* Classifies the current state (ghost fill first): flags per face, dilated
 * (dilate != 0) or raw, and the logits [faces][4]; any output may be null.
 * x faces are [NY][NX+1] (row j, face f between points f-1 and f), y faces
 * [NY+1][NX]. */
int wenofd_switch_classify(wenofd_solver* h, int dilate, unsigned char* flags_x, unsigned char* flags_y,
                           float* logits_x, float* logits_y);
/* Classify every `steps` steps (1..64, default 1) and, unless a radius is
 * set (wenofd_switch_dilation), dilate the flags by min(`steps`, 16) faces:
 * a wave moves less than one cell per step at CFL <= 1. */
int wenofd_switch_interval(wenofd_solver* h, int steps);
/* Dilation radius in faces (0..16) in place of the interval; -1 (the
 * default) follows the interval. A radius below the interval trades the
 * guarantee above for fewer WENO faces (the plain solver still redoes a step
 * that fails). No counterpart in the reference. */
int wenofd_switch_dilation(wenofd_solver* h, int radius);
/* The (dilated) flags the last switched step used. */
int wenofd_switch_flags(wenofd_solver* h, unsigned char* flags_x, unsigned char* flags_y);
/* Number of switched steps the plain solver had to redo. */
int wenofd_switch_stats(wenofd_solver* h, long long* fallbacks);
/* Switch samples of every face of one direction of the current state:
 * features [faces][4][10] (per characteristic field: the + stencil and the
 * mirrored - stencil, each centred and scaled) and labels [faces][4]
 * (max |omega - d| of WENO-Z over both sides > 0.05). */
int wenofd_harvest(wenofd_solver* h, int dir, float* X, unsigned char* Y);

/* Switch MLP trainer: parameters, Adam moments and a device-resident sample
 * set; FP32 forward/backward/Adam kernels of this library. */
typedef struct wenofd_mlp wenofd_mlp;
/* Parameters initialised ~ U(-1/sqrt(fan_in), 1/sqrt(fan_in)) from seed. */
int wenofd_mlp_create(int device, unsigned long long seed, wenofd_mlp** out);
int wenofd_mlp_destroy(wenofd_mlp* m);
/* Sets the parameters (and clears the Adam moments) / reads them. */
int wenofd_mlp_set_params(wenofd_mlp* m, const float* params);
int wenofd_mlp_get_params(wenofd_mlp* m, float* params);
/* Appends samples from host arrays X [n][10], Y [n]. */
int wenofd_mlp_dataset_add(wenofd_mlp* m, const float* X, const unsigned char* Y, long long n);
/* Appends the samples of both directions of a solver's current state, each
 * kept with probability keep_pos (label 1) or keep_neg (label 0) from a
 * seeded hash (deterministic). */
int wenofd_mlp_harvest(wenofd_mlp* m, wenofd_solver* h, double keep_pos, double keep_neg, unsigned long long seed,
                       long long* added);
int wenofd_mlp_dataset_size(wenofd_mlp* m, long long* n, long long* positives);
int wenofd_mlp_dataset_get(wenofd_mlp* m, long long start, long long n, float* X, unsigned char* Y);
int wenofd_mlp_dataset_clear(wenofd_mlp* m);
/* Mean loss (weighted BCE with logits: pw y softplus(-z) + (1-y) softplus(z))
 * and its gradient at the current parameters over samples idx[0..n). */
int wenofd_mlp_grad(wenofd_mlp* m, const int* idx, int n, float pos_weight, float* grad, float* loss);
/* Adam (lr, 0.9, 0.999, 1e-8) over shuffled mini-batches; epoch_losses[e] =
 * mean batch loss of epoch e (may be null). */
int wenofd_mlp_train(wenofd_mlp* m, int epochs, int batch, float lr, float pos_weight, unsigned long long seed,
                     float* epoch_losses);
/* Logits of the deployed network (bf16 MFMA, FP32 accumulate) for X [n][10]. */
int wenofd_mlp_infer(wenofd_mlp* m, const float* X, long long n, float* logits);
/* {TP, FP, FN, TN} of the deployed network over samples [start, start+n). */
int wenofd_mlp_evaluate(wenofd_mlp* m, long long start, long long n, float thr, long long* counts4);

/*
 * Batched solves. No counterpart in the reference, which runs one case per
 * process: a batch holds many independent CWENOFD runs (each its own config
 * text: grid size, test case, CFL, TEND, vortex parameters ...; all sharing
 * SCHEME, RKMETH and FP_CONTRACT) and advances them together, one kernel launch per phase
 * for the whole batch. Each solve's time step is computed and applied on the
 * device (calculateDeltaT, src/WENOFD.cpp:398-452, clipped to its own t_end);
 * a solve that reaches its t_end (|now - t_end| <= 1e-9, :213) or meets a NaN
 * step stops while the others go on. The step sequence is captured once in a
 * HIP graph of graph_steps steps and replayed. Each solve's state equals the
 * single-solve path (wenofd_advance) bit for bit.
 */
typedef struct wenofd_batch wenofd_batch;

int wenofd_batch_create(const char* const* config_texts, int nsolves, int device, wenofd_batch** out);
int wenofd_batch_destroy(wenofd_batch* b);
int wenofd_batch_size(wenofd_batch* b, int* nsolves);
/* Every solve back to its initial condition at time 0. */
int wenofd_batch_initialize_ave(wenofd_batch* b);
/* Advances every running solve by up to max_steps run-loop iterations
 * (max_steps <= 0: until all reach t_end). steps_done = iterations launched;
 * active_left = solves still running. WENOFD_E_NAN_DT if a solve met a NaN
 * step (it stopped; wenofd_last_error names it). */
int wenofd_batch_advance(wenofd_batch* b, int max_steps, int* steps_done, int* active_left);
/* Steps per graph replay (default 16; 0 launches each kernel directly). */
int wenofd_batch_set_graph_steps(wenofd_batch* b, int steps);
/* Tile shapes (as wenofd_set_tiles) for every solve of the batch. */
int wenofd_batch_set_tiles(wenofd_batch* b, int x_faces, int y_faces);
int wenofd_batch_get_tiles(wenofd_batch* b, int* x_faces, int* y_faces);
/* Times `steps` steps of the batch under each tile shape, starting each trial
 * from the current state, keeps the fastest and restores the state. */
int wenofd_batch_autotune(wenofd_batch* b, int steps, int* x_faces, int* y_faces, double* ms_per_step);
int wenofd_batch_get_state(wenofd_batch* b, int solve, double* host, int layout);
int wenofd_batch_get_info(wenofd_batch* b, int solve, wenofd_info* info);
/* {Linf, L1, L2} of rho of one solve against its exact solution. */
int wenofd_batch_accuracy(wenofd_batch* b, int solve, double* norms3);
int wenofd_batch_synchronize(wenofd_batch* b);

/* Probes of the device arithmetic on independent inputs (host arrays):
 * WENO5 reconstruction (include/Utility.hpp:415-615) of n 5-tuples, and per
 * state {Lx, Rx, Ly, Ry (16 each, row-major), Fm_x, Fp_x, Fm_y, Fp_y (4
 * each), max eigenvalue x, y} = 82 doubles (utils/equations.hpp:137-393). */
int wenofd_probe_weno5(int device, int scheme, double zp_lambda, int n, const double* v5, double* out);
int wenofd_probe_physics(int device, double gamma, int n, const double* states, double* out82);

#ifdef __cplusplus
}
#endif

#endif /* WENOFD_AMD_H */
