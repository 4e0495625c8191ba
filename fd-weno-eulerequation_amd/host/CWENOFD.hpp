// CWENOFD.hpp — C++ host class with the reference solver's surface
// (include/WENOFD.hpp:28-124 of infinity-07/FD-WENO-EulerEquation), running on
// one MI355X through the C ABI of libwenofd_amd.so (include/wenofd_amd.h).
//
// Same method names and meanings as the reference: initializeSolver /
// initializeAve / run / calculateDeltaT / RunRK1-3 / assembleRHS (= getFlux) /
// assembleBoundaryTerm / assembleSourceTerm / setBoundary /
// exchangeGhostCellsValue / useWENO / outputAve / outputError /
// outputAccuracy, the constructor from an EulerEquation pointer, and the same
// config.cfg keys. Differences: the whole domain lives on one GPU (no MPI
// slabs), and failures throw wenofd::Error instead of calling MPI_Abort.
#pragma once

#include <cstdio>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "wenofd_amd.h"

namespace wenofd {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
    if (rc != WENOFD_OK) throw Error(rc, wenofd_last_error());
}

// utils/readConfig.hpp:30-80: key = value file into a map.
class CConfig {
public:
    explicit CConfig(const std::string& path) : m_inputFileName(path) {}
    void read(std::map<std::string, std::string>& option) const {
        FILE* f = std::fopen(m_inputFileName.c_str(), "r");
        if (!f) throw Error(WENOFD_E_CONFIG, "Failed to open file" + m_inputFileName);
        std::string text;
        char buf[4096];
        size_t n;
        while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, n);
        std::fclose(f);
        size_t pos = 0;
        while (pos <= text.size()) {
            size_t nl = text.find('\n', pos);
            std::string line = text.substr(pos, nl == std::string::npos ? std::string::npos : nl - pos);
            pos = (nl == std::string::npos) ? text.size() + 1 : nl + 1;
            if (line.empty() || line[0] == '#' || line[0] == '\r') continue;
            size_t eq = line.find('=');
            if (eq == std::string::npos) continue;
            option[trim(line.substr(0, eq))] = trim(line.substr(eq + 1));
        }
    }
    std::string m_inputFileName;

private:
    static std::string trim(std::string s) {
        size_t first = s.find_first_not_of(' ');
        if (first == std::string::npos) return "";
        size_t last = s.find_last_not_of(' ');
        if (s[last] == '\r' && last > first) --last;
        return s.substr(first, last - first + 1);
    }
};

// Stands in for the reference's EulerEquation (utils/equations.hpp:55-699),
// whose test-case set-up the library performs from the config keys: the
// reference's main (src/main.cpp:13-26) passes one to the solver's
// constructor, and the solver here only keeps the pointer.
struct EulerEquation {};

class CWENOFD {
public:
    explicit CWENOFD(int device = -1) : m_device(device) {}
    // include/WENOFD.hpp:34-37: the equation object is not owned.
    explicit CWENOFD(EulerEquation* eq, int device = -1) : equation(eq), m_device(device) {}
    ~CWENOFD() {
        if (m_h) wenofd_destroy(m_h);
    }
    CWENOFD(const CWENOFD&) = delete;
    CWENOFD& operator=(const CWENOFD&) = delete;

    // src/WENOFD.cpp:10-164 (grid, buffers) + initial state upload.
    void initializeSolver(std::map<std::string, std::string> option) {
        std::string text;
        for (auto& kv : option) text += kv.first + " = " + kv.second + "\n";
        if (m_h) wenofd_destroy(m_h), m_h = nullptr;
        check(wenofd_create(text.c_str(), m_device, &m_h));
        refresh();
    }
    // src/WENOFD.cpp:165-188
    void initializeAve() {
        check(wenofd_initialize_ave(m_h));
        refresh();
    }
    // src/WENOFD.cpp:189-299
    void run(std::map<std::string, std::string> option, bool verbose = true) {
        initializeSolver(option);
        check(wenofd_run(m_h, verbose ? 1 : 0));
        refresh();
    }
    // src/WENOFD.cpp:398-452
    double calculateDeltaT() {
        double dt = 0;
        check(wenofd_calculate_delta_t(m_h, &dt));
        return dt;
    }
    void RunRK1(double deltaT) { check(wenofd_run_rk(m_h, 0, deltaT)); }
    void RunRK2(double deltaT) { check(wenofd_run_rk(m_h, 1, deltaT)); }
    void RunRK3(double deltaT) { check(wenofd_run_rk(m_h, 2, deltaT)); }
    // assembleRHS = ghost fill + getFlux (assembleBoundaryTerm) + source
    // (src/WENOFD.cpp:454-465)
    void assembleRHS() { check(wenofd_get_flux(m_h)); }
    void getFlux() { assembleRHS(); }
    // src/WENOFD.cpp:481-798: the flux terms of both directions (written into
    // the rhs, which the reference's assembleRHS zeroes before calling it)
    void assembleBoundaryTerm() { check(wenofd_assemble_boundary_term(m_h)); }
    // src/WENOFD.cpp:800-821: adds the RTI gravity source to the rhs
    void assembleSourceTerm() { check(wenofd_assemble_source_term(m_h)); }
    // src/WENOFD.cpp:998-1159
    void setBoundary() { check(wenofd_set_boundary(m_h)); }
    // src/WENOFD.cpp:850-929. With the whole domain on one device the halo
    // ring reduces to its self-send, which fills the periodic x ghosts: that
    // fill is part of the library's ghost fill, so this refills every ghost
    // point (a superset of what the reference's call writes, from the same
    // state).
    void exchangeGhostCellsValue() { check(wenofd_set_boundary(m_h)); }
    // src/WENOFD.cpp:823-848: the scalar WENO5 reconstruction of the solve's
    // scheme at j+1/2 from (u_{j-2}, ..., u_{j+2}), run by the device
    // arithmetic the flux kernels use (WENO-ZP with lambda = dx^(2/3)).
    double useWENO(double& uavemm, double& uavem, double& uave, double& uavep, double& uavepp) {
        const double v[5] = {uavemm, uavem, uave, uavep, uavepp};
        double out = 0;
        check(wenofd_probe_weno5(m_info.device, m_info.scheme, m_info.zp_lambda, 1, v, &out));
        return out;
    }
    void outputAve(const std::string& prefix) { check(wenofd_output_ave(m_h, prefix.c_str())); }
    void outputError(const std::string& prefix) { check(wenofd_output_error(m_h, prefix.c_str())); }
    void outputAccuracy(const std::string& prefix) { check(wenofd_output_accuracy(m_h, prefix.c_str(), nullptr)); }

    // One iteration of run()'s loop: dt, RunRK<RKMETH>, m_now += dt.
    double step() {
        double dt = 0;
        check(wenofd_step(m_h, 0.0, &dt));
        refresh();
        return dt;
    }
    std::vector<double> state() const {  // interior, reference order [NX][NY][4]
        std::vector<double> v((size_t)m_info.nx * m_info.ny * 4);
        check(wenofd_get_state(m_h, v.data(), WENOFD_LAYOUT_REF));
        return v;
    }
    void setState(const std::vector<double>& v) { check(wenofd_set_state(m_h, v.data(), WENOFD_LAYOUT_REF)); }
    std::vector<double> rhs() const {
        std::vector<double> v((size_t)m_info.nx * m_info.ny * 4);
        check(wenofd_get_rhs(m_h, v.data(), WENOFD_LAYOUT_REF));
        return v;
    }
    void refresh() {
        check(wenofd_get_info(m_h, &m_info));
        m_now = m_info.now;
        m_deltaX = m_info.dx;
        m_deltaY = m_info.dy;
        m_cfl = m_info.cfl;
        m_outputTime = m_info.t_end;
        m_worldPointNumX = m_info.nx;
        m_worldPointNumY = m_info.ny;
    }
    wenofd_solver* handle() const { return m_h; }

    // Public state mirroring the reference's members (refreshed after calls).
    EulerEquation* equation = nullptr;
    double m_now = 0, m_deltaX = 0, m_deltaY = 0, m_cfl = 0, m_outputTime = 0;
    int m_worldPointNumX = 0, m_worldPointNumY = 0;
    int m_ghostCellNum = 4;

private:
    wenofd_solver* m_h = nullptr;
    wenofd_info m_info{};
    int m_device;
};

}  // namespace wenofd
