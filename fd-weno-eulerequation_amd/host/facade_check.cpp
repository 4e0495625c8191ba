// facade_check.cpp — drives every method of the C++ facade (CWENOFD.hpp) the
// way the reference's own code calls its CWENOFD (src/main.cpp:5-33,
// src/WENOFD.cpp:189-465), and writes what a test needs to check the results:
//
//   facade_check <outdir>
//
// <outdir>/facade.txt holds one "name value ..." line per result (doubles as
// %.17g); <outdir>/*.bin hold raw little-endian doubles in the reference's
// [NX][NY][4] order. Exit status 0 when every call succeeded.
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "CWENOFD.hpp"

namespace {

void dump(const std::string& path, const std::vector<double>& v) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw wenofd::Error(WENOFD_E_IO, "cannot write " + path);
    std::fwrite(v.data(), sizeof(double), v.size(), f);
    std::fclose(f);
}

std::map<std::string, std::string> options(const char* tc, const char* scheme, const char* rk, const char* cfl,
                                           const char* nx, const char* ny, const std::string& outdir) {
    return {{"TESTCASE", tc}, {"SCHEME", scheme}, {"RKMETH", rk},           {"CFL", cfl},
            {"ELEMNUMX", nx}, {"ELEMNUMY", ny},   {"OUTPUTDIR", outdir + "/"}};
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: facade_check <outdir>\n");
        return 2;
    }
    const std::string out = argv[1];
    FILE* rep = std::fopen((out + "/facade.txt").c_str(), "w");
    if (!rep) {
        std::fprintf(stderr, "cannot write %s/facade.txt\n", out.c_str());
        return 2;
    }
    try {
        // RTI 12 x 40, WENO-ZP, RK2: the case with a source term and SPECIAL sides
        wenofd::EulerEquation equation;
        wenofd::CWENOFD rti(&equation);
        rti.initializeSolver(options("3", "2", "1", "0.6", "12", "40", out));
        rti.initializeAve();
        rti.exchangeGhostCellsValue();
        rti.setBoundary();
        rti.assembleRHS();
        dump(out + "/rti_rhs.bin", rti.rhs());
        rti.assembleBoundaryTerm();
        dump(out + "/rti_boundary.bin", rti.rhs());
        rti.assembleSourceTerm();
        dump(out + "/rti_boundary_source.bin", rti.rhs());
        const double dt = rti.calculateDeltaT();
        std::fprintf(rep, "rti_dt %.17g\n", dt);
        rti.RunRK2(dt);
        dump(out + "/rti_rk2.bin", rti.state());
        std::fprintf(rep, "rti_equation_kept %d\n", rti.equation == &equation ? 1 : 0);
        std::fprintf(rep, "rti_deltaX %.17g deltaY %.17g points %d %d ghosts %d\n", rti.m_deltaX, rti.m_deltaY,
                     rti.m_worldPointNumX, rti.m_worldPointNumY, rti.m_ghostCellNum);

        // Riemann2 20 x 18, WENO-Z: RunRK1 and RunRK3 with calculateDeltaT
        wenofd::CWENOFD r2;
        r2.initializeSolver(options("2", "1", "2", "0.5", "20", "18", out));
        double d1 = r2.calculateDeltaT();
        r2.RunRK1(d1);
        double d3 = r2.calculateDeltaT();
        r2.RunRK3(d3);
        std::fprintf(rep, "r2_dts %.17g %.17g\n", d1, d3);
        dump(out + "/r2_rk1_rk3.bin", r2.state());
        r2.initializeAve();
        dump(out + "/r2_initial.bin", r2.state());

        // useWENO of every scheme on one stencil (the solve's scheme and dx)
        double a = 0.3, b = -1.2, c = 2.5, d = 0.7, e = 0.05;
        for (const char* sc : {"0", "1", "2", "3"}) {
            wenofd::CWENOFD w;
            w.initializeSolver(options("0", sc, "2", "0.05", "17", "17", out));
            std::fprintf(rep, "useWENO %s %.17g %.17g\n", sc, w.m_deltaX, w.useWENO(a, b, c, d, e));
        }

        // SMOOTH 16 x 16, WENO-JS: the whole run() with its output files
        wenofd::CWENOFD sm;
        sm.run(options("0", "0", "2", "0.05", "16", "16", out), false);
        std::fprintf(rep, "smooth_now %.17g\n", sm.m_now);
        sm.outputAve("again");
        sm.outputError("again");
        sm.outputAccuracy("again");
        dump(out + "/smooth_final.bin", sm.state());
        std::fprintf(rep, "ok 1\n");
    } catch (const wenofd::Error& err) {
        std::fprintf(rep, "error %d %s\n", err.code, err.what());
        std::fclose(rep);
        std::fprintf(stderr, "Error (%d): %s\n", err.code, err.what());
        return 1;
    }
    std::fclose(rep);
    return 0;
}
