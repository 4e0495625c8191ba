// wenofd_cases.cpp — configuration keys and test cases (host code).
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <sstream>

#include "wenofd_amd.h"
#include "wenofd_host.hpp"
#include "wenofd_kernels.hpp"

namespace wenofd {

// utils/equations.hpp:23-28, 38-43 (plus the two cases without a reference counterpart)
const char* const kCaseNames[6] = {"SMOOTH", "Riemann1", "Riemann2", "RTI", "VORTEX", "QUADRANT"};
const char* const kSchemeNames[4] = {"WENOJS", "WENOZ", "WENOZP", "WENOZPI"};

static void trim_spaces(std::string& s) {
    // utils/readConfig.hpp:12-28: strip spaces, then one trailing '\r'
    const size_t first = s.find_first_not_of(' ');
    if (first == std::string::npos) {
        s.clear();
        return;
    }
    size_t last = s.find_last_not_of(' ');
    if (s[last] == '\r') {
        if (last == first) {
            s.clear();
            return;
        }
        --last;
    }
    s = s.substr(first, last - first + 1);
}

void parse_config_lines(const std::string& text, std::map<std::string, std::string>& kv) {
    std::istringstream in(text);
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#' || line[0] == '\r') continue;
        const size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        std::string key = line.substr(0, eq), value = line.substr(eq + 1);
        trim_spaces(key);
        trim_spaces(value);
        kv[key] = value;
    }
}

// std::stoi / std::stod semantics: leading whitespace, a numeric prefix, the
// rest ignored; no digits -> error.
static bool leading_int(const std::string& s, long& v) {
    const char* b = s.c_str();
    char* e = nullptr;
    v = std::strtol(b, &e, 10);
    return e != b;
}
static bool leading_double(const std::string& s, double& v) {
    const char* b = s.c_str();
    char* e = nullptr;
    v = std::strtod(b, &e);
    return e != b;
}

int interpret_options(const std::map<std::string, std::string>& kv, Options& o, std::string& err) {
    auto need_int = [&](const char* key, long& v) -> int {
        auto it = kv.find(key);
        if (it == kv.end() || !leading_int(it->second, v)) {
            err = std::string("config key ") + key + (it == kv.end() ? " is missing" : " is not an integer");
            return WENOFD_E_CONFIG;
        }
        return 0;
    };
    long tc, rk, sc, nx, ny;
    int rc;
    if ((rc = need_int("TESTCASE", tc)) || (rc = need_int("RKMETH", rk)) || (rc = need_int("SCHEME", sc)) ||
        (rc = need_int("ELEMNUMX", nx)) || (rc = need_int("ELEMNUMY", ny)))
        return rc;
    auto itc = kv.find("CFL");
    if (itc == kv.end() || !leading_double(itc->second, o.cfl)) {
        err = itc == kv.end() ? "config key CFL is missing" : "config key CFL is not a number";
        return WENOFD_E_CONFIG;
    }
    if (tc < 0 || tc > 5) { err = "TESTCASE must be 0..5"; return WENOFD_E_BAD_ENUM; }
    if (rk < 0 || rk > 2) { err = "RKMETH must be 0 (RK1), 1 (RK2) or 2 (RK3)"; return WENOFD_E_BAD_ENUM; }
    if (sc < 0 || sc > 3) { err = "SCHEME must be 0 (JS), 1 (Z), 2 (ZP) or 3 (ZPI)"; return WENOFD_E_BAD_ENUM; }
    if (nx < 5 || ny < 5 || nx > 65536 || ny > 65536) {
        err = "ELEMNUMX and ELEMNUMY must be in [5, 65536]";
        return WENOFD_E_CONFIG;
    }
    if (!(o.cfl > 0)) { err = "CFL must be positive"; return WENOFD_E_CONFIG; }
    o.testcase = (int)tc; o.rkmeth = (int)rk; o.scheme = (int)sc; o.nx = (int)nx; o.ny = (int)ny;

    auto opt_double = [&](const char* key, double& v) -> int {
        auto it = kv.find(key);
        if (it == kv.end()) return 0;
        if (!leading_double(it->second, v)) {
            err = std::string("config key ") + key + " is not a number";
            return WENOFD_E_CONFIG;
        }
        return 0;
    };
    if ((rc = opt_double("TEND", o.t_end)) || (rc = opt_double("VORTEX_EPS", o.vortex_eps)) ||
        (rc = opt_double("VORTEX_XC", o.vortex_xc)) || (rc = opt_double("VORTEX_YC", o.vortex_yc)))
        return rc;
    auto iq = kv.find("QUAD_STATES");
    if (iq != kv.end()) {
        std::string s = iq->second;
        for (char& ch : s)
            if (ch == ',') ch = ' ';
        std::istringstream in(s);
        for (int k = 0; k < 16; ++k)
            if (!(in >> o.quad_states[k / 4][k % 4])) {
                err = "QUAD_STATES needs 16 numbers (UR, UL, LL, LR x rho, u, v, p)";
                return WENOFD_E_CONFIG;
            }
        o.has_quad_states = true;
    }
    auto is = kv.find("QUAD_SEED");
    if (is != kv.end()) {
        const char* b = is->second.c_str();
        char* e = nullptr;
        o.quad_seed = std::strtoull(b, &e, 10);
        if (e == b) { err = "QUAD_SEED is not an integer"; return WENOFD_E_CONFIG; }
        o.has_quad_seed = true;
    }
    auto ifp = kv.find("FP_CONTRACT");
    if (ifp != kv.end()) {
        long v;
        if (!leading_int(ifp->second, v)) { err = "FP_CONTRACT is not an integer"; return WENOFD_E_CONFIG; }
        if (v != 0 && v != 1) { err = "FP_CONTRACT must be 0 (exact) or 1 (FMA contraction)"; return WENOFD_E_BAD_ENUM; }
        o.fp_contract = (int)v;
    }
    auto io = kv.find("OUTPUTDIR");
    if (io != kv.end() && !io->second.empty()) {
        o.output_dir = io->second;
        if (o.output_dir.back() != '/') o.output_dir += '/';
    }
    return 0;
}

static const double kPi = 3.14159265358979323846;

void CaseDef::primitive0(double x, double y, double& rho, double& u, double& v, double& p) const {
    switch (testcase) {
        case kSmooth:  // Qiu & Shu 2005 smooth density wave, equations.hpp:419-434
            rho = 1 + 0.2 * std::sin(kPi * (x + y));
            u = 0.7;
            v = 0.3;
            p = 1;
            break;
        case kRiemann1:  // equations.hpp:458-513 (split at (1, 1))
            if (x < 1.0 && y < 1.0) { rho = 0.8; u = 0.0; v = 0.0; p = 1.0; }
            else if (x < 1.0) { rho = 1.0; u = 0.7276; v = 0.0; p = 1.0; }
            else if (y < 1.0) { rho = 1.0; u = 0.0; v = 0.7276; p = 1.0; }
            else { rho = 0.5313; u = 0.0; v = 0.0; p = 0.4; }
            break;
        case kRiemann2:  // equations.hpp:537-592 (split at (0.8, 0.8))
            if (x < 0.8 && y < 0.8) { rho = 0.138; u = 1.206; v = 1.206; p = 0.029; }
            else if (x < 0.8) { rho = 0.5323; u = 1.206; v = 0.0; p = 0.3; }
            else if (y < 0.8) { rho = 0.5323; u = 0.0; v = 1.206; p = 0.3; }
            else { rho = 1.5; u = 0.0; v = 0.0; p = 1.5; }
            break;
        case kRTI: {  // Rayleigh-Taylor, equations.hpp:616-641
            rho = (y <= 0.5) ? 2.0 : 1.0;
            u = 0;
            p = (y <= 0.5) ? 1 + 2 * y : y + 1.5;
            const double c = std::sqrt(gamma * p / rho);
            v = -0.025 * c * std::cos(8 * kPi * x);
            break;
        }
        case kVortex: {  // isentropic vortex (Shu 1997), background (1, 1, 1, 1)
            const double xb = x - vortex_xc, yb = y - vortex_yc;
            const double r2 = xb * xb + yb * yb;
            const double amp = vortex_eps / (2.0 * kPi) * std::exp(0.5 * (1.0 - r2));
            const double T =
                1.0 - (gamma - 1.0) * vortex_eps * vortex_eps / (8.0 * gamma * kPi * kPi) * std::exp(1.0 - r2);
            rho = std::pow(T, 1.0 / (gamma - 1.0));
            u = 1.0 - amp * yb;
            v = 1.0 + amp * xb;
            p = std::pow(rho, gamma);
            break;
        }
        default: {  // four-quadrant Riemann problem split at (0.5, 0.5)
            int q;
            if (x < 0.5 && y < 0.5) q = 2;
            else if (x < 0.5) q = 1;
            else if (y < 0.5) q = 3;
            else q = 0;
            rho = quad[q][0]; u = quad[q][1]; v = quad[q][2]; p = quad[q][3];
        }
    }
}

void CaseDef::conserved0(double x, double y, double U[4]) const {
    double rho, u, v, p;
    primitive0(x, y, rho, u, v, p);
    const double E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v);
    U[0] = rho;
    U[1] = rho * u;
    U[2] = rho * v;
    U[3] = E;
}

double CaseDef::exact_rho(double x, double y, double t) const {
    if (testcase == kSmooth) return 1 + 0.2 * std::sin(kPi * (x + y - t));  // equations.hpp:438-439
    if (testcase == kVortex) {  // the initial vortex advected by (1, 1), minimum periodic image
        const double Lx = xR - xL, Ly = yR - yL;
        double xb = x - t - vortex_xc, yb = y - t - vortex_yc;
        xb -= Lx * std::floor(xb / Lx + 0.5);
        yb -= Ly * std::floor(yb / Ly + 0.5);
        const double r2 = xb * xb + yb * yb;
        const double T =
            1.0 - (gamma - 1.0) * vortex_eps * vortex_eps / (8.0 * gamma * kPi * kPi) * std::exp(1.0 - r2);
        return std::pow(T, 1.0 / (gamma - 1.0));
    }
    return 0.0;
}

static inline uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void seeded_quadrant_states(uint64_t seed, double out[4][4]) {
    uint64_t s = seed;
    auto u01 = [&]() { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); };
    for (int q = 0; q < 4; ++q) {
        out[q][0] = 0.1 + 1.9 * u01();
        out[q][1] = -1.5 + 3.0 * u01();
        out[q][2] = -1.5 + 3.0 * u01();
        out[q][3] = 0.05 + 1.95 * u01();
    }
}

void make_case(const Options& o, CaseDef& c) {
    c.testcase = o.testcase;
    switch (o.testcase) {
        case kSmooth:
            c.xL = 0; c.xR = 2; c.yL = 0; c.yR = 2; c.t_end = 1.0; c.gamma = 1.4; c.has_exact = true;
            c.bc_left = c.bc_right = c.bc_bottom = c.bc_top = kBcPeriod;
            break;
        case kRiemann1:
            c.xL = 0; c.xR = 2; c.yL = 0; c.yR = 2; c.t_end = 0.52; c.gamma = 1.4; c.has_exact = false;
            c.bc_left = c.bc_right = c.bc_bottom = c.bc_top = kBcSymmetric;
            break;
        case kRiemann2:
            c.xL = 0; c.xR = 1; c.yL = 0; c.yR = 1; c.t_end = 0.8; c.gamma = 1.4; c.has_exact = false;
            c.bc_left = c.bc_right = c.bc_bottom = c.bc_top = kBcSymmetric;
            break;
        case kRTI:
            c.xL = 0; c.xR = 0.25; c.yL = 0; c.yR = 1; c.t_end = 1.95; c.gamma = 5.0 / 3.0; c.has_exact = false;
            c.bc_left = c.bc_right = kBcSlip;
            c.bc_bottom = c.bc_top = kBcSpecial;
            break;
        case kVortex:
            c.xL = 0; c.xR = 10; c.yL = 0; c.yR = 10; c.t_end = 2.0; c.gamma = 1.4; c.has_exact = true;
            c.bc_left = c.bc_right = c.bc_bottom = c.bc_top = kBcPeriod;
            c.vortex_eps = o.vortex_eps; c.vortex_xc = o.vortex_xc; c.vortex_yc = o.vortex_yc;
            break;
        default: {
            c.xL = 0; c.xR = 1; c.yL = 0; c.yR = 1; c.t_end = 0.3; c.gamma = 1.4; c.has_exact = false;
            c.bc_left = c.bc_right = c.bc_bottom = c.bc_top = kBcSymmetric;
            // Riemann2's four states (equations.hpp:537-592) moved to (0.5, 0.5)
            const double s[4][4] = {{1.5, 0, 0, 1.5}, {0.5323, 1.206, 0, 0.3},
                                    {0.138, 1.206, 1.206, 0.029}, {0.5323, 0, 1.206, 0.3}};
            std::memcpy(c.quad, s, sizeof s);
            if (o.has_quad_seed) seeded_quadrant_states(o.quad_seed, c.quad);
            if (o.has_quad_states) std::memcpy(c.quad, o.quad_states, sizeof c.quad);
        }
    }
    if (o.t_end >= 0) c.t_end = o.t_end;
}

void make_grid(const CaseDef& c, int nx, int ny, GridSetup& gs) {
    gs.nx = nx;
    gs.ny = ny;
    gs.dx = (c.xR - c.xL) / (nx - 1);
    gs.dy = (c.yR - c.yL) / (ny - 1);
    gs.zp_lambda = std::pow(gs.dx, 2.0 / 3.0);  // include/Utility.hpp:531 with m_deltaX (src/WENOFD.cpp:836)
    const double gxl = (c.xL + 0 * gs.dx) - gs.dx * 4;
    const double gyl = c.yL - gs.dy * 4;
    gs.xc.resize(nx + 8);
    gs.yc.resize(ny + 8);
    for (int i = 0; i < nx + 8; ++i) gs.xc[i] = gxl + i * gs.dx;
    for (int j = 0; j < ny + 8; ++j) gs.yc[j] = gyl + j * gs.dy;
}

void special_band(const CaseDef& c, const GridSetup& gs, std::vector<double>& sp) {
    const int nx = gs.nx, ny = gs.ny, ns = 2 * (nx + ny);
    const int sx = 4, ex = nx + 4, sy = 4, ey = ny + 4;
    sp.assign(4 * (size_t)ns, 0.0);
    for (int i = 0; i < nx; ++i) {
        double Ub[4], Ut[4];
        c.conserved0(gs.xc[sx + i], gs.yc[sy], Ub);
        c.conserved0(gs.xc[sx + i], gs.yc[ey - 1], Ut);
        for (int r = 0; r < 4; ++r) {
            sp[r * ns + i] = Ub[r];
            sp[r * ns + nx + i] = Ut[r];
        }
    }
    for (int j = 0; j < ny; ++j) {
        double Ul[4], Ur[4];
        c.conserved0(gs.xc[sx], gs.yc[sy + j], Ul);
        c.conserved0(gs.xc[ex], gs.yc[sy + j], Ur);  // first ghost column's coordinate (:1144)
        for (int r = 0; r < 4; ++r) {
            sp[r * ns + 2 * nx + j] = Ul[r];
            sp[r * ns + 2 * nx + ny + j] = Ur[r];
        }
    }
}

void initial_state(const CaseDef& c, const GridSetup& gs, int pitch, long plane, std::vector<double>& full) {
    full.assign(4 * (size_t)plane, 0.0);
    for (int i = 4; i < gs.nx + 4; ++i)
        for (int j = 4; j < gs.ny + 4; ++j) {
            double Uc[4];
            c.conserved0(gs.xc[i], gs.yc[j], Uc);
            for (int r = 0; r < 4; ++r) full[r * plane + (size_t)j * pitch + i] = Uc[r];
        }
}

void rho_error_norms(const CaseDef& c, const GridSetup& gs, double t, const std::vector<double>& Uref,
                     double out3[3]) {
    const int nx = gs.nx, ny = gs.ny;
    double e1 = 0, e2 = 0, einf = 0;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            const double ex = c.exact_rho(gs.xc[i + 4], gs.yc[j + 4], t);
            const double e = std::fabs(ex - Uref[((size_t)i * ny + j) * 4]);
            e1 += e;
            e2 += e * e;
            if (e > einf) einf = e;
        }
    out3[0] = einf;
    out3[1] = e1 / nx / ny;
    out3[2] = std::sqrt(e2 / nx / ny);
}

bool unpack_state(const double* full, int nx, int ny, int pitch, long plane, int layout, double* host) {
    if (layout == WENOFD_LAYOUT_REF) {
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j)
                for (int r = 0; r < 4; ++r)
                    host[((size_t)i * ny + j) * 4 + r] = full[r * plane + (size_t)(j + 4) * pitch + (i + 4)];
    } else if (layout == WENOFD_LAYOUT_SOA) {
        for (int r = 0; r < 4; ++r)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i)
                    host[((size_t)r * ny + j) * nx + i] = full[r * plane + (size_t)(j + 4) * pitch + (i + 4)];
    } else if (layout == WENOFD_LAYOUT_REF_GHOSTED) {
        for (int i = 0; i < nx + 8; ++i)
            for (int j = 0; j < ny + 8; ++j) {
                const bool corner = (i < 4 || i >= nx + 4) && (j < 4 || j >= ny + 4);
                for (int r = 0; r < 4; ++r)
                    host[((size_t)i * (ny + 8) + j) * 4 + r] = corner ? 0.0 : full[r * plane + (size_t)j * pitch + i];
            }
    } else {
        return false;
    }
    return true;
}

bool pack_state(const double* host, int nx, int ny, int pitch, long plane, int layout, double* full) {
    if (layout == WENOFD_LAYOUT_REF) {
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j)
                for (int r = 0; r < 4; ++r)
                    full[r * plane + (size_t)(j + 4) * pitch + (i + 4)] = host[((size_t)i * ny + j) * 4 + r];
    } else if (layout == WENOFD_LAYOUT_SOA) {
        for (int r = 0; r < 4; ++r)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i)
                    full[r * plane + (size_t)(j + 4) * pitch + (i + 4)] = host[((size_t)r * ny + j) * nx + i];
    } else if (layout == WENOFD_LAYOUT_REF_GHOSTED) {
        for (int i = 0; i < nx + 8; ++i)
            for (int j = 0; j < ny + 8; ++j)
                for (int r = 0; r < 4; ++r)
                    full[r * plane + (size_t)j * pitch + i] = host[((size_t)i * (ny + 8) + j) * 4 + r];
    } else {
        return false;
    }
    return true;
}

}  // namespace wenofd
