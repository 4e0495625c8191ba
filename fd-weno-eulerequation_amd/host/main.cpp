// main.cpp — drop-in for the reference program (src/main.cpp:5-33): reads
// ./input/config.cfg (or the path given as the first argument) and runs the
// solver on one GPU, writing ./output/<stamp>_<CASE>_<SCHEME>_<NX>x<NY>/.
#include <chrono>
#include <cstdio>
#include <map>
#include <string>

#include "CWENOFD.hpp"

int main(int argc, char** argv) {
    const std::string inputfile = argc > 1 ? argv[1] : "./input/config.cfg";
    try {
        std::printf("Starting to read configuration file...\n");
        wenofd::CConfig config(inputfile);
        std::map<std::string, std::string> option;
        config.read(option);
        wenofd::CWENOFD solver;
        const auto t0 = std::chrono::steady_clock::now();
        solver.run(option);
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("solver: %.3f s, %d x %d points, t = %.6g\n", s, solver.m_worldPointNumX,
                    solver.m_worldPointNumY, solver.m_now);
    } catch (const wenofd::Error& e) {
        std::fprintf(stderr, "Error (%d): %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
