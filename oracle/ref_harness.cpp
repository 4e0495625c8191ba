// ref_harness.cpp — drives the reference solver (infinity-07/FD-WENO-EulerEquation),
// compiled from its own sources where they lie, to write golden vectors.
//
// TEST INFRASTRUCTURE ONLY: built by oracle/build_ref.sh into oracle/_ref/
// (git-ignored), run by tests/golden/make_golden.py in the container that has
// the reference mounted, and by bench.py (bench mode) to time the reference's
// CPU path as the baseline. It uses the reference's public class surface
// (include/WENOFD.hpp:28-124, utils/equations.hpp:55-699) and its inline WENO
// functions (include/Utility.hpp:415-615). All binary files are little-endian
// float64; states are interior-only in the reference's [NX][NY][4] order unless
// stated otherwise.
//
// Modes:
//   weno   <scheme> <deltaX> <in> <out>        in: n x 5, out: n
//   phys   <gamma> <in> <out>                  in: n x 4 (valid states), out: n x 90
//   rhs    <tc> <scheme> <nx> <ny> <in> <out> <ghost_out>
//            sets m_Uh from <in>, assembleRHS(); out = rhs interior (n x 4) then
//            one extra double = calculateDeltaT() with CFL 0.6, now = 0;
//            ghost_out = full m_Uh incl. ghosts [(NX+8)][(NY+8)][4]
//   traj   <tc> <scheme> <rk> <cfl> <nx> <ny> <nsteps> <fixed_dt 0|1> <t_end|-1> <out> [quad16|vortex3...]
//            initializeAve + nsteps x (calculateDeltaT or step-0 dt, RunRK);
//            nsteps < 0 runs the reference loop condition to t_end;
//            out = interior state then the dt sequence. tc 4 = isentropic vortex
//            on [0,10]^2 (periodic, built on SMOOTH's boundary set), tc 5 =
//            four-quadrant on [0,1]^2 split at (0.5,0.5) (built on Riemann2's).
//            Works under mpiexec -n P: rank 0 writes the gathered state.
//   init   same arguments as traj with nsteps = 0 (initial interior state)
//   run    <tc> <scheme> <rk> <cfl> <nx> <ny>   the reference's CWENOFD::run
//   bench  <tc> <scheme> <rk> <cfl> <nx> <ny> <nsteps>
//            as traj from the case's initial state (tc 4 / 5 with the default
//            parameters), timing only the nsteps iterations of the reference's
//            loop body (calculateDeltaT, RunRK, now += dt) between barriers with
//            MPI_Wtime; rank 0 prints "ranks R steps K seconds S" on stdout.
//            Used by bench.py as the CPU baseline (serial and under mpiexec).

#include <mpi.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "WENOFD.hpp"
#include "Utility.hpp"

static std::vector<double> read_bin(const char* path) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) { std::fprintf(stderr, "cannot read %s\n", path); MPI_Abort(MPI_COMM_WORLD, 2); }
    const size_t n = (size_t)f.tellg() / sizeof(double);
    std::vector<double> v(n);
    f.seekg(0);
    f.read(reinterpret_cast<char*>(v.data()), n * sizeof(double));
    return v;
}

static void write_bin(const char* path, const std::vector<double>& v) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(v.data()), v.size() * sizeof(double));
}

static void to_arr(const std::vector<double>& v, size_t off, Array1D<double>& a) {
    for (int r = 0; r < 4; ++r) a[r] = v[off + r];
}

static std::map<std::string, std::string> options(int tc, int scheme, int rk, double cfl, int nx, int ny) {
    std::map<std::string, std::string> o;
    o["TESTCASE"] = std::to_string(tc);
    o["SCHEME"] = std::to_string(scheme);
    o["RKMETH"] = std::to_string(rk);
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.17g", cfl);
    o["CFL"] = buf;
    o["ELEMNUMX"] = std::to_string(nx);
    o["ELEMNUMY"] = std::to_string(ny);
    return o;
}

static const double kPi = M_PI;

// Re-derives the rank-local grid after a domain change, with the same
// expressions as CWENOFD::initializeSolver (WENOFD.cpp:73-106).
static void regrid(CWENOFD& s, EulerEquation* eq) {
    s.m_deltaX = (eq->xR - eq->xL) / (s.m_worldPointNumX - 1);
    s.m_deltaY = (eq->yR - eq->yL) / (s.m_worldPointNumY - 1);
    const double procXLeft = eq->xL + s.m_startPointX_per_rank[s.m_rank] * s.m_deltaX;
    const double procYLeft = eq->yL;
    const double gxl = procXLeft - s.m_deltaX * s.m_ghostCellNum;
    const double gyl = procYLeft - s.m_deltaY * s.m_ghostCellNum;
    for (int ei = 0; ei < s.m_totalPointNumX; ++ei)
        for (int ej = 0; ej < s.m_totalPointNumY; ++ej) {
            s.m_grids[ei][ej].m_xCenter = gxl + ei * s.m_deltaX;
            s.m_grids[ei][ej].m_yCenter = gyl + ej * s.m_deltaY;
        }
    if (s.m_rank == 0)
        for (int ei = 0; ei < s.m_worldPointNumX + 2 * s.m_ghostCellNum; ++ei)
            for (int ej = 0; ej < s.m_worldPointNumY + 2 * s.m_ghostCellNum; ++ej) {
                s.m_worldGrids[ei][ej].m_xCenter = eq->xL - s.m_ghostCellNum * s.m_deltaX + ei * s.m_deltaX;
                s.m_worldGrids[ei][ej].m_yCenter = eq->yL - s.m_ghostCellNum * s.m_deltaY + ej * s.m_deltaY;
            }
}

// Isentropic vortex (Shu 1997) with background (1,1,1,1) on [0,10]^2.
static void make_vortex(CWENOFD& s, EulerEquation* eq, double veps, double xc, double yc) {
    eq->xL = 0; eq->xR = 10; eq->yL = 0; eq->yR = 10;
    const double g = eq->euler_gamma;
    auto T = [=](double x, double y) {
        const double xb = x - xc, yb = y - yc, r2 = xb * xb + yb * yb;
        return 1.0 - (g - 1.0) * veps * veps / (8.0 * g * kPi * kPi) * std::exp(1.0 - r2);
    };
    auto amp = [=](double x, double y) {
        const double xb = x - xc, yb = y - yc, r2 = xb * xb + yb * yb;
        return veps / (2.0 * kPi) * std::exp(0.5 * (1.0 - r2));
    };
    eq->rho0 = [=](double x, double y) { return std::pow(T(x, y), 1.0 / (g - 1.0)); };
    eq->u0 = [=](double x, double y) { return 1.0 - amp(x, y) * (y - yc); };
    eq->v0 = [=](double x, double y) { return 1.0 + amp(x, y) * (x - xc); };
    eq->pre0 = [=](double x, double y) { return std::pow(std::pow(T(x, y), 1.0 / (g - 1.0)), g); };
    regrid(s, eq);
}

static void make_quadrant(CWENOFD& s, EulerEquation* eq, const double q[4][4]) {
    // quadrant order UR, UL, LL, LR; split at (0.5, 0.5)
    std::vector<double> st(&q[0][0], &q[0][0] + 16);
    auto pick = [=](double x, double y, int k) {
        int i;
        if (x < 0.5 && y < 0.5) i = 2;
        else if (x < 0.5 && y >= 0.5) i = 1;
        else if (x >= 0.5 && y < 0.5) i = 3;
        else i = 0;
        return st[i * 4 + k];
    };
    eq->rho0 = [=](double x, double y) { return pick(x, y, 0); };
    eq->u0 = [=](double x, double y) { return pick(x, y, 1); };
    eq->v0 = [=](double x, double y) { return pick(x, y, 2); };
    eq->pre0 = [=](double x, double y) { return pick(x, y, 3); };
    (void)s;
}

static void run_rk(CWENOFD& s, int rk, double dt) {
    if (rk == 0) s.RunRK1(dt);
    else if (rk == 1) s.RunRK2(dt);
    else s.RunRK3(dt);
}

int main(int argc, char** argv) {
    MPI_Init(&argc, &argv);
    int rank;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    if (argc < 2) { std::fprintf(stderr, "usage: see header\n"); MPI_Abort(MPI_COMM_WORLD, 2); }
    const std::string mode = argv[1];

    if (mode == "weno") {
        const int scheme = std::atoi(argv[2]);
        double dx = std::atof(argv[3]);
        std::vector<double> in = read_bin(argv[4]);
        const size_t n = in.size() / 5;
        std::vector<double> out(n);
        for (size_t k = 0; k < n; ++k) {
            double a = in[5 * k], b = in[5 * k + 1], c = in[5 * k + 2], d = in[5 * k + 3], e = in[5 * k + 4];
            switch (scheme) {
                case 0: out[k] = WENO5threconstruction(a, b, c, d, e); break;
                case 1: out[k] = WENO5Zthreconstruction(a, b, c, d, e); break;
                case 2: out[k] = WENO5ZPtheconstruction(a, b, c, d, e, dx); break;
                default: out[k] = WENO5ZPItheconstruction(a, b, c, d, e); break;
            }
        }
        write_bin(argv[5], out);
    } else if (mode == "phys") {
        EulerEquation eq;
        eq.euler_gamma = std::atof(argv[2]);
        std::vector<double> in = read_bin(argv[3]);
        const size_t n = in.size() / 4;
        std::vector<double> out;
        out.reserve(n * 90);
        Array1D<double> U(4), Fm(4), Fp(4), F(4);
        Array2D<double> M(4, 4);
        for (size_t k = 0; k < n; ++k) {
            to_arr(in, 4 * k, U);
            auto dumpM = [&]() { for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) out.push_back(M[i][j]); };
            eq.getLEigenMatrixX(U, M); dumpM();
            eq.getREigenMatrixX(U, M); dumpM();
            eq.getLEigenMatrixY(U, M); dumpM();
            eq.getREigenMatrixY(U, M); dumpM();
            eq.fluxSplitX(U, Fm, Fp);
            for (int r = 0; r < 4; ++r) out.push_back(Fm[r]);
            for (int r = 0; r < 4; ++r) out.push_back(Fp[r]);
            eq.fluxSplitY(U, Fm, Fp);
            for (int r = 0; r < 4; ++r) out.push_back(Fm[r]);
            for (int r = 0; r < 4; ++r) out.push_back(Fp[r]);
            out.push_back(eq.getMaxEigenValue(U, 1, 0));
            out.push_back(eq.getMaxEigenValue(U, 0, 1));
            eq.getPhyFluxX(U, F);
            for (int r = 0; r < 4; ++r) out.push_back(F[r]);
            eq.getPhyFluxY(U, F);
            for (int r = 0; r < 4; ++r) out.push_back(F[r]);
        }
        write_bin(argv[4], out);
    } else if (mode == "rhs") {
        const int tc = std::atoi(argv[2]), scheme = std::atoi(argv[3]);
        const int nx = std::atoi(argv[4]), ny = std::atoi(argv[5]);
        EulerEquation* eq = new EulerEquation();
        CWENOFD s(eq);
        s.initializeSolver(options(tc, scheme, 2, 0.6, nx, ny));
        s.m_solverTimer.start();
        std::vector<double> in = read_bin(argv[6]);
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j)
                for (int r = 0; r < 4; ++r) s.m_Uh[i + 4][j + 4].vector[r] = in[((size_t)i * ny + j) * 4 + r];
        s.m_now = 0;
        s.assembleRHS();
        std::vector<double> out;
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j)
                for (int r = 0; r < 4; ++r) out.push_back(s.m_rhs[i + 4][j + 4].vector[r]);
        out.push_back(s.calculateDeltaT());
        write_bin(argv[7], out);
        std::vector<double> full;
        for (int i = 0; i < nx + 8; ++i)
            for (int j = 0; j < ny + 8; ++j)
                for (int r = 0; r < 4; ++r) full.push_back(s.m_Uh[i][j].vector[r]);
        write_bin(argv[8], full);
    } else if (mode == "bench") {
        const int tc = std::atoi(argv[2]), scheme = std::atoi(argv[3]), rk = std::atoi(argv[4]);
        const double cfl = std::atof(argv[5]);
        const int nx = std::atoi(argv[6]), ny = std::atoi(argv[7]), nsteps = std::atoi(argv[8]);
        EulerEquation* eq = new EulerEquation();
        CWENOFD s(eq);
        const int base_tc = (tc == 4) ? 0 : (tc == 5 ? 2 : tc);
        s.initializeSolver(options(base_tc, scheme, rk, cfl, nx, ny));
        s.m_solverTimer.start();
        if (tc == 4) {
            make_vortex(s, eq, 5.0, 5.0, 5.0);
        } else if (tc == 5) {
            double q[4][4] = {{1.5, 0, 0, 1.5}, {0.5323, 1.206, 0, 0.3}, {0.138, 1.206, 1.206, 0.029}, {0.5323, 0, 1.206, 0.3}};
            make_quadrant(s, eq, q);
        }
        s.m_outputTime = 1e9;
        s.initializeAve();
        s.m_now = 0;
        MPI_Barrier(MPI_COMM_WORLD);
        const double t0 = MPI_Wtime();
        for (int k = 0; k < nsteps; ++k) {
            const double dt = s.calculateDeltaT();
            run_rk(s, rk, dt);
            s.m_now += dt;
        }
        MPI_Barrier(MPI_COMM_WORLD);
        const double t1 = MPI_Wtime();
        int nranks = 1;
        MPI_Comm_size(MPI_COMM_WORLD, &nranks);
        if (rank == 0) std::printf("ranks %d steps %d seconds %.6f\n", nranks, nsteps, t1 - t0);
    } else if (mode == "traj" || mode == "init") {
        const int tc = std::atoi(argv[2]), scheme = std::atoi(argv[3]), rk = std::atoi(argv[4]);
        const double cfl = std::atof(argv[5]);
        const int nx = std::atoi(argv[6]), ny = std::atoi(argv[7]), nsteps = std::atoi(argv[8]);
        const int fixed = std::atoi(argv[9]);
        const double t_end = std::atof(argv[10]);
        const char* outp = argv[11];
        EulerEquation* eq = new EulerEquation();
        CWENOFD s(eq);
        const int base_tc = (tc == 4) ? 0 : (tc == 5 ? 2 : tc);
        s.initializeSolver(options(base_tc, scheme, rk, cfl, nx, ny));
        s.m_solverTimer.start();
        if (tc == 4) {
            double v3[3] = {5.0, 5.0, 5.0};
            for (int k = 0; k < 3 && 12 + k < argc; ++k) v3[k] = std::atof(argv[12 + k]);
            make_vortex(s, eq, v3[0], v3[1], v3[2]);
            s.m_outputTime = 2.0;
        } else if (tc == 5) {
            double q[4][4] = {{1.5, 0, 0, 1.5}, {0.5323, 1.206, 0, 0.3}, {0.138, 1.206, 1.206, 0.029}, {0.5323, 0, 1.206, 0.3}};
            if (argc >= 12 + 16)
                for (int k = 0; k < 16; ++k) q[k / 4][k % 4] = std::atof(argv[12 + k]);
            make_quadrant(s, eq, q);
            s.m_outputTime = 0.3;
        }
        if (t_end >= 0) s.m_outputTime = t_end;
        s.initializeAve();
        s.m_now = 0;
        std::vector<double> dts;
        double dt0 = 0;
        // nsteps < 0: the reference's loop condition (WENOFD.cpp:213) to t_end
        for (int k = 0; nsteps < 0 ? std::fabs(s.m_now - s.m_outputTime) > 1e-9 : k < nsteps; ++k) {
            double dt = (fixed && k > 0) ? dt0 : s.calculateDeltaT();
            if (k == 0) dt0 = dt;
            run_rk(s, rk, dt);
            s.m_now += dt;
            dts.push_back(dt);
        }
        s.GatherAllUhToRank0();
        if (rank == 0) {
            std::vector<double> out;
            out.reserve((size_t)nx * ny * 4 + dts.size());
            for (int i = 0; i < nx; ++i)
                for (int j = 0; j < ny; ++j)
                    for (int r = 0; r < 4; ++r) out.push_back(s.m_worldUh[i + 4][j + 4].vector[r]);
            for (double d : dts) out.push_back(d);
            write_bin(outp, out);
        }
    } else if (mode == "run") {
        // The reference's own driver, end to end (WENOFD.cpp:189-299): writes
        // ./output/<stamp>_<CASE>_<SCHEME>_<NX>x<NY>/ under the working directory.
        const int tc = std::atoi(argv[2]), scheme = std::atoi(argv[3]), rk = std::atoi(argv[4]);
        const double cfl = std::atof(argv[5]);
        const int nx = std::atoi(argv[6]), ny = std::atoi(argv[7]);
        EulerEquation* eq = new EulerEquation();
        CWENOFD s(eq);
        s.run(options(tc, scheme, rk, cfl, nx, ny));
    } else {
        std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
        MPI_Abort(MPI_COMM_WORLD, 2);
    }
    MPI_Finalize();
    return 0;
}
