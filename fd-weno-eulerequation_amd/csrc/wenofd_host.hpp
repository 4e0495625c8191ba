// wenofd_host.hpp — host-side pieces of the solver: config keys, test cases,
// grid conventions. Internal to libwenofd_amd.so.
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace wenofd {

enum TestCase { kSmooth = 0, kRiemann1 = 1, kRiemann2 = 2, kRTI = 3, kVortex = 4, kQuadrant = 5 };

// Parsed configuration (utils/readConfig.hpp:39-80 key = value lines; keys of
// src/WENOFD.cpp:24-40 plus optional extensions).
struct Options {
    int testcase = 0, scheme = 0, rkmeth = 2;
    double cfl = 0.5;
    int nx = 0, ny = 0;
    double t_end = -1.0;  // < 0: the case default
    double vortex_eps = 5.0, vortex_xc = 5.0, vortex_yc = 5.0;
    bool has_quad_states = false;
    double quad_states[4][4] = {};
    bool has_quad_seed = false;
    uint64_t quad_seed = 0;
    int fp_contract = 0;  // FP_CONTRACT: FP64 arithmetic mode of the sweeps (FpMode)
    std::string output_dir = "./output/";
};

// Reads key = value lines with CConfig::read's rules: lines that are empty or
// start with '#' or '\r' are skipped, a line without '=' is ignored, keys and
// values are trimmed of spaces and of one trailing '\r'.
void parse_config_lines(const std::string& text, std::map<std::string, std::string>& kv);

// Interprets the keys. Returns 0 or a WENOFD_E_* code with a message in err.
int interpret_options(const std::map<std::string, std::string>& kv, Options& o, std::string& err);

// One test case: domain, boundary types, gamma, end time, initial condition,
// exact solution (utils/equations.hpp:400-662 plus the isentropic vortex and
// the parameterised four-quadrant problem).
struct CaseDef {
    int testcase = 0;
    double xL = 0, xR = 1, yL = 0, yR = 1, t_end = 1, gamma = 1.4;
    int bc_left = 0, bc_right = 0, bc_bottom = 0, bc_top = 0;
    bool has_exact = false;
    double vortex_eps = 5.0, vortex_xc = 5.0, vortex_yc = 5.0;
    double quad[4][4] = {};  // [UR, UL, LL, LR][rho, u, v, p], split at (0.5, 0.5)

    void primitive0(double x, double y, double& rho, double& u, double& v, double& p) const;
    // getU0 (equations.hpp:649-662): conserved variables of the initial condition.
    void conserved0(double x, double y, double U[4]) const;
    double exact_rho(double x, double y, double t) const;
};

void make_case(const Options& o, CaseDef& c);

// splitmix64 stream -> 53-bit uniforms; per quadrant (UR, UL, LL, LR) rho ~
// U[0.1, 2], u ~ U[-1.5, 1.5], v ~ U[-1.5, 1.5], p ~ U[0.05, 2], in that order.
void seeded_quadrant_states(uint64_t seed, double out[4][4]);

// Point coordinates of one solve including 4 ghost points per side, grid
// conventions of src/WENOFD.cpp:73-106 for one rank (procXLeft = xL).
struct GridSetup {
    int nx = 0, ny = 0;
    double dx = 0, dy = 0, zp_lambda = 0;
    std::vector<double> xc, yc;  // nx + 8 and ny + 8 entries
};
void make_grid(const CaseDef& c, int nx, int ny, GridSetup& gs);

// SPECIAL boundary values (time-invariant), [4][2 nx + 2 ny]: bottom row,
// top row, left column, right column (src/WENOFD.cpp:1045-1058, 1092-1102,
// 1138-1149).
void special_band(const CaseDef& c, const GridSetup& gs, std::vector<double>& sp);

// initializeAve (src/WENOFD.cpp:165-188) into a zeroed 4-plane buffer with the
// given row pitch and plane size (interior points only).
void initial_state(const CaseDef& c, const GridSetup& gs, int pitch, long plane, std::vector<double>& full);

// {Linf, L1, L2} of rho against the exact solution at time t (outputAccuracy,
// src/WENOFD.cpp:1308-1360: x outer, y inner). Uref in the [nx][ny][4] order.
void rho_error_norms(const CaseDef& c, const GridSetup& gs, double t, const std::vector<double>& Uref,
                     double out3[3]);

// Conversions between a device-layout state buffer (4 planes, row pitch,
// ghosts) and the C ABI's host layouts (WENOFD_LAYOUT_*). Return false for an
// unknown layout.
bool unpack_state(const double* full, int nx, int ny, int pitch, long plane, int layout, double* host);
bool pack_state(const double* host, int nx, int ny, int pitch, long plane, int layout, double* full);

// Records msg as the calling thread's last error (wenofd_last_error) and
// returns code.
int set_last_error(int code, const std::string& msg);

extern const char* const kCaseNames[6];
extern const char* const kSchemeNames[4];

}  // namespace wenofd
