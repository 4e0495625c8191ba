// wenofd_oracle.cpp — CPU oracle for the WENO-FD 2-D Euler hot path.
//
// TEST INFRASTRUCTURE ONLY. Nothing in the product links, loads or calls this
// file: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
// it, and only as the checker. It is a from-scratch restatement of the
// reference solver's arithmetic (infinity-07/FD-WENO-EulerEquation), written to
// reproduce the reference's floating-point operation ORDER so that it matches a
// no-contraction build of the reference bit for bit. Build with
// -ffp-contract=off, no fast-math (oracle/Makefile).
//
// Pinning (what checks this file, see tests/test_oracle_*.py):
//   * golden vectors written by the compiled reference (oracle/_ref, committed
//     under tests/golden/): WENO JS/Z/ZP/ZPI, eigenmatrices, flux split,
//     max eigenvalue, ghost fill + assembleRHS, calculateDeltaT, RK1/2/3
//     trajectories, full run() accuracy CSV;
//   * closed forms: L.R = I, R.diag(lambda).L = dF/dU, F+ + F- = F,
//     F+ - F- = alpha U, WENO of constant/linear data, the upwind-5 identity,
//     ghost index maps, the reference's published SMOOTH convergence table.
//
// Data layout here mirrors the reference's Array2D<Cvector> index order:
// point (i, j) with i along x (0..NX+7), j along y (0..NY+7), four conserved
// variables (rho, rho u, rho v, E) contiguous per point:
//     idx(i, j, r) = ((i * (NY + 8)) + j) * 4 + r.
//
// Scalar type R is a template parameter so that the same code runs in double
// (the oracle proper) and in OpCount (an operation-counting double used to
// derive the algorithmic FLOP figure per cell-update).

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

namespace orc {

// ---------------------------------------------------------------------------
// Operation-counting double.
// ---------------------------------------------------------------------------
struct Counters {
    unsigned long long add = 0, mul = 0, div = 0, sqrt_ = 0, cmp = 0;
};
static Counters g_cnt;

struct OpCount {
    double x;
    OpCount() : x(0) {}
    OpCount(double v) : x(v) {}
};
static inline OpCount operator+(OpCount a, OpCount b) { g_cnt.add++; return OpCount(a.x + b.x); }
static inline OpCount operator-(OpCount a, OpCount b) { g_cnt.add++; return OpCount(a.x - b.x); }
static inline OpCount operator*(OpCount a, OpCount b) { g_cnt.mul++; return OpCount(a.x * b.x); }
static inline OpCount operator/(OpCount a, OpCount b) { g_cnt.div++; return OpCount(a.x / b.x); }
static inline OpCount operator-(OpCount a) { return OpCount(-a.x); }
static inline bool operator<(OpCount a, OpCount b) { g_cnt.cmp++; return a.x < b.x; }
static inline bool operator>(OpCount a, OpCount b) { g_cnt.cmp++; return a.x > b.x; }
static inline OpCount& operator+=(OpCount& a, OpCount b) { a = a + b; return a; }
static inline OpCount& operator/=(OpCount& a, OpCount b) { a = a / b; return a; }
static inline OpCount r_sqrt(OpCount a) { g_cnt.sqrt_++; return OpCount(std::sqrt(a.x)); }
static inline OpCount r_fabs(OpCount a) { return OpCount(std::fabs(a.x)); }
static inline double val(OpCount a) { return a.x; }

static inline double r_sqrt(double a) { return std::sqrt(a); }
static inline double r_fabs(double a) { return std::fabs(a); }
static inline double val(double a) { return a; }

template <class R> static inline R rmax(R a, R b) { return (a < b) ? b : a; }  // std::max
template <class R> static inline R rmin(R a, R b) { return (b < a) ? b : a; }  // std::min

// ---------------------------------------------------------------------------
// Scalar WENO5 reconstructions at j+1/2 from (f_{j-2}, ..., f_{j+2}).
// include/Utility.hpp:5-11 (constants), :415-456 (JS), :457-504 (Z),
// :505-558 (ZP), :559-615 (ZPI). pow(x, 2) in the reference is x*x (GCC folds
// it at -O3; both are the correctly rounded square).
// ---------------------------------------------------------------------------
static const double C1_3 = 1.0 / 3.0, C1_4 = 1.0 / 4.0, C1_6 = 1.0 / 6.0, C5_6 = 5.0 / 6.0,
                    C7_6 = 7.0 / 6.0, C11_6 = 11.0 / 6.0, C13_12 = 13.0 / 12.0;

template <class R>
static inline void smoothness(R a, R b, R c, R d, R e, R& b0, R& b1, R& b2) {
    // Utility.hpp:427-437 (same forms in Z/ZP/ZPI)
    R t1 = (a - R(2.0) * b) + c;
    R t2 = (a - R(4.0) * b) + R(3.0) * c;
    b0 = (R(C13_12) * t1) * t1 + (R(0.25) * t2) * t2;
    t1 = (b - R(2.0) * c) + d;
    t2 = b - d;
    b1 = (R(C13_12) * t1) * t1 + (R(0.25) * t2) * t2;
    t1 = (c - R(2.0) * d) + e;
    t2 = (R(3.0) * c - R(4.0) * d) + e;
    b2 = (R(C13_12) * t1) * t1 + (R(0.25) * t2) * t2;
}

template <class R>
static inline void candidates(R a, R b, R c, R d, R e, R& p0, R& p1, R& p2) {
    // Utility.hpp:451-453 / 499-501
    p0 = (R(C1_3) * a - R(C7_6) * b) + R(C11_6) * c;
    p1 = (R(-C1_6) * b + R(C5_6) * c) + R(C1_3) * d;
    p2 = (R(C1_3) * c + R(C5_6) * d) - R(C1_6) * e;
}

template <class R>
static inline R weno5_js(R a, R b, R c, R d, R e) {
    R b0, b1, b2, p0, p1, p2;
    smoothness(a, b, c, d, e, b0, b1, b2);
    const R eps(1.0e-6);
    R w0 = R(0.1) / ((eps + b0) * (eps + b0));
    R w1 = R(0.6) / ((eps + b1) * (eps + b1));
    R w2 = R(0.3) / ((eps + b2) * (eps + b2));
    const R sum = (w0 + w1) + w2;
    w0 = w0 / sum;
    w1 = w1 / sum;
    w2 = w2 / sum;
    candidates(a, b, c, d, e, p0, p1, p2);
    return (w0 * p0 + w1 * p1) + w2 * p2;
}

template <class R>
static inline R weno5_z(R a, R b, R c, R d, R e) {
    R b0, b1, b2, p0, p1, p2;
    smoothness(a, b, c, d, e, b0, b1, b2);
    const R eps(1.0e-20);
    const R i0 = R(1.0) / (b0 + eps);
    const R i1 = R(1.0) / (b1 + eps);
    const R i2 = R(1.0) / (b2 + eps);
    const R tau = r_fabs(b0 - b2);
    R w0 = R(0.1) * (R(1.0) + (((tau * i0) * tau) * i0));
    R w1 = R(0.6) * (R(1.0) + (((tau * i1) * tau) * i1));
    R w2 = R(0.3) * (R(1.0) + (((tau * i2) * tau) * i2));
    const R s = (w0 + w1) + w2;
    w0 /= s;
    w1 /= s;
    w2 /= s;
    candidates(a, b, c, d, e, p0, p1, p2);
    return (w0 * p0 + w1 * p1) + w2 * p2;
}

template <class R>
static inline R weno5_zp(R a, R b, R c, R d, R e, double lambda_) {
    R b0, b1, b2, p0, p1, p2;
    smoothness(a, b, c, d, e, b0, b1, b2);
    const R tau = r_fabs(b0 - b2);
    const R lambda(lambda_);
    const R eps(1.0e-40);
    const R q0 = (tau + eps) / (b0 + eps);
    const R q1 = (tau + eps) / (b1 + eps);
    const R q2 = (tau + eps) / (b2 + eps);
    R w0 = R(0.1) * ((R(1.0) + q0 * q0) + (lambda * (b0 + eps)) / (tau + eps));
    R w1 = R(0.6) * ((R(1.0) + q1 * q1) + (lambda * (b1 + eps)) / (tau + eps));
    R w2 = R(0.3) * ((R(1.0) + q2 * q2) + (lambda * (b2 + eps)) / (tau + eps));
    const R sum = (w0 + w1) + w2;
    w0 = w0 / sum;
    w1 = w1 / sum;
    w2 = w2 / sum;
    candidates(a, b, c, d, e, p0, p1, p2);
    return (w0 * p0 + w1 * p1) + w2 * p2;
}

template <class R>
static inline R weno5_zpi(R a, R b, R c, R d, R e) {
    R b0, b1, b2, p0, p1, p2;
    smoothness(a, b, c, d, e, b0, b1, b2);
    const R tau = r_fabs(b0 - b2);
    const R lambda(3.3);
    const R bmax = rmax(b0, rmax(b1, b2));
    const R bmin = rmin(b0, rmin(b1, b2));
    const R alpha = R(1.0) - (bmin / (bmax + R(1e-40)));
    const R eps(1.0e-40);
    const R q0 = tau / (b0 + eps);
    const R q1 = tau / (b1 + eps);
    const R q2 = tau / (b2 + eps);
    R w0 = R(0.1) * ((R(1.0) + q0 * q0) + ((lambda * alpha) * b0) / (bmax + eps));
    R w1 = R(0.6) * ((R(1.0) + q1 * q1) + ((lambda * alpha) * b1) / (bmax + eps));
    R w2 = R(0.3) * ((R(1.0) + q2 * q2) + ((lambda * alpha) * b2) / (bmax + eps));
    const R sum = (w0 + w1) + w2;
    w0 = w0 / sum;
    w1 = w1 / sum;
    w2 = w2 / sum;
    candidates(a, b, c, d, e, p0, p1, p2);
    return (w0 * p0 + w1 * p1) + w2 * p2;
}

enum { JS = 0, Z = 1, ZP = 2, ZPI = 3 };

template <class R>
static inline R weno5(int scheme, R a, R b, R c, R d, R e, double zp_lambda) {
    // src/WENOFD.cpp:823-848 (useWENO)
    switch (scheme) {
        case JS: return weno5_js(a, b, c, d, e);
        case Z: return weno5_z(a, b, c, d, e);
        case ZP: return weno5_zp(a, b, c, d, e, zp_lambda);
        default: return weno5_zpi(a, b, c, d, e);
    }
}

// ---------------------------------------------------------------------------
// Euler physics. utils/equations.hpp:105-171 (flux, max eigenvalue),
// :174-299 (left/right eigenmatrices), :365-393 (LLF split).
// ---------------------------------------------------------------------------
template <class R>
struct Prim {
    R rho, u, v, E, p, c;
};

template <class R>
static inline Prim<R> primitive(const R* U, double gamma) {
    Prim<R> q;
    q.rho = U[0];
    q.u = U[1] / U[0];
    q.v = U[2] / U[0];
    q.E = U[3];
    q.p = (R(gamma) - R(1.0)) * (q.E - (R(0.5) * q.rho) * (q.u * q.u + q.v * q.v));
    q.c = r_sqrt((R(gamma) * q.p) / q.rho);
    return q;
}

template <class R>
static inline R max_eigen(const R* U, double nx, double ny, double gamma) {
    // equations.hpp:137-147
    Prim<R> q = primitive(U, gamma);
    return r_fabs(q.u * R(nx) + q.v * R(ny)) + q.c;
}

template <class R>
static inline void phys_flux(const R* U, int dir, double gamma, R* F) {
    Prim<R> q = primitive<R>(U, gamma);
    if (dir == 0) {  // equations.hpp:105-119
        F[0] = q.rho * q.u;
        F[1] = (q.rho * q.u) * q.u + q.p;
        F[2] = (q.rho * q.u) * q.v;
        F[3] = (q.E + q.p) * q.u;
    } else {  // equations.hpp:121-134
        F[0] = q.rho * q.v;
        F[1] = (q.rho * q.u) * q.v;
        F[2] = (q.rho * q.v) * q.v + q.p;
        F[3] = (q.E + q.p) * q.v;
    }
}

template <class R>
static inline void flux_split(const R* U, int dir, double gamma, R* Fm, R* Fp) {
    // equations.hpp:365-393: alpha = |u_n| + c at this point
    R F[4];
    phys_flux(U, dir, gamma, F);
    Prim<R> q = primitive<R>(U, gamma);
    const R ws = (dir == 0 ? r_fabs(q.u) : r_fabs(q.v)) + q.c;
    for (int r = 0; r < 4; ++r) {
        Fm[r] = R(0.5) * (F[r] - ws * U[r]);
        Fp[r] = R(0.5) * (F[r] + ws * U[r]);
    }
}

// L[row*4+col]
template <class R>
static inline void eigen_left(const R* U, int dir, double gamma, R* L) {
    Prim<R> q = primitive<R>(U, gamma);
    const R B1 = (R(gamma) - R(1.0)) / (q.c * q.c);
    const R B2 = (B1 * (q.u * q.u + q.v * q.v)) / R(2.0);
    const R one_c = R(1.0) / q.c;
    if (dir == 0) {  // equations.hpp:174-204
        const R uc = q.u / q.c;
        L[0] = (B2 + uc) / R(2.0);
        L[4] = q.v;
        L[8] = R(1.0) - B2;
        L[12] = (B2 - uc) / R(2.0);
        L[1] = -(B1 * q.u + one_c) / R(2.0);
        L[5] = R(0.0);
        L[9] = B1 * q.u;
        L[13] = -(B1 * q.u - one_c) / R(2.0);
        L[2] = -(B1 * q.v) / R(2.0);
        L[6] = R(-1.0);
        L[10] = B1 * q.v;
        L[14] = ((-B1) * q.v) / R(2.0);
        L[3] = B1 / R(2.0);
        L[7] = R(0.0);
        L[11] = -B1;
        L[15] = B1 / R(2.0);
    } else {  // equations.hpp:206-236
        const R vc = q.v / q.c;
        L[0] = (B2 + vc) / R(2.0);
        L[4] = -q.u;
        L[8] = R(1.0) - B2;
        L[12] = (B2 - vc) / R(2.0);
        L[1] = -(B1 * q.u) / R(2.0);
        L[5] = R(1.0);
        L[9] = B1 * q.u;
        L[13] = -(B1 * q.u) / R(2.0);
        L[2] = -(B1 * q.v + one_c) / R(2.0);
        L[6] = R(0.0);
        L[10] = B1 * q.v;
        L[14] = -(B1 * q.v - one_c) / R(2.0);
        L[3] = B1 / R(2.0);
        L[7] = R(0.0);
        L[11] = -B1;
        L[15] = B1 / R(2.0);
    }
}

template <class R>
static inline void eigen_right(const R* U, int dir, double gamma, R* Rm) {
    Prim<R> q = primitive<R>(U, gamma);
    const R H = (q.E + q.p) / q.rho;
    const R q2 = (q.u * q.u + q.v * q.v) / R(2.0);
    if (dir == 0) {  // equations.hpp:239-268
        Rm[0] = R(1.0);  Rm[4] = q.u - q.c;  Rm[8] = q.v;   Rm[12] = H - q.c * q.u;
        Rm[1] = R(0.0);  Rm[5] = R(0.0);     Rm[9] = R(-1.0); Rm[13] = -q.v;
        Rm[2] = R(1.0);  Rm[6] = q.u;        Rm[10] = q.v;  Rm[14] = q2;
        Rm[3] = R(1.0);  Rm[7] = q.u + q.c;  Rm[11] = q.v;  Rm[15] = H + q.c * q.u;
    } else {  // equations.hpp:270-299
        Rm[0] = R(1.0);  Rm[4] = q.u;        Rm[8] = q.v - q.c;  Rm[12] = H - q.c * q.v;
        Rm[1] = R(0.0);  Rm[5] = R(1.0);     Rm[9] = R(0.0);     Rm[13] = q.u;
        Rm[2] = R(1.0);  Rm[6] = q.u;        Rm[10] = q.v;       Rm[14] = q2;
        Rm[3] = R(1.0);  Rm[7] = q.u;        Rm[11] = q.v + q.c; Rm[15] = H + q.c * q.v;
    }
}

// One characteristic-wise face flux. Fp5[s] = F+ at points f-3..f+1, Fm5[s] =
// F- at points f-2..f+2 (s = 0..4), Ul/Ur = U at points f-1 and f.
// src/WENOFD.cpp:526-644 (x) and :668-792 (y).
template <class R>
static inline void face_flux(const R* Ul, const R* Ur, const R (*Fp5)[4], const R (*Fm5)[4], int dir,
                             int scheme, double gamma, double zp_lambda, R* Fhat) {
    R ave[4], L[16], Rm[16];
    for (int r = 0; r < 4; ++r) ave[r] = R(0.5) * (Ul[r] + Ur[r]);  // arithmetic mean, :529-530
    eigen_left(ave, dir, gamma, L);
    eigen_right(ave, dir, gamma, Rm);
    R vp[5][4], vm[5][4];
    for (int s = 0; s < 5; ++s) {
        for (int i = 0; i < 4; ++i) {  // left-to-right 4-term sums, :552-593
            R a = L[i * 4 + 0] * Fp5[s][0];
            a += L[i * 4 + 1] * Fp5[s][1];
            a += L[i * 4 + 2] * Fp5[s][2];
            a += L[i * 4 + 3] * Fp5[s][3];
            vp[s][i] = a;
            R b = L[i * 4 + 0] * Fm5[s][0];
            b += L[i * 4 + 1] * Fm5[s][1];
            b += L[i * 4 + 2] * Fm5[s][2];
            b += L[i * 4 + 3] * Fm5[s][3];
            vm[s][i] = b;
        }
    }
    R vh[4];
    for (int k = 0; k < 4; ++k) {
        R hp = weno5<R>(scheme, vp[0][k], vp[1][k], vp[2][k], vp[3][k], vp[4][k], zp_lambda);
        R hm = weno5<R>(scheme, vm[4][k], vm[3][k], vm[2][k], vm[1][k], vm[0][k], zp_lambda);  // mirrored, :604-605
        vh[k] = hp + hm;
    }
    for (int r = 0; r < 4; ++r) {  // :625-643
        R a = Rm[r * 4 + 0] * vh[0];
        a += Rm[r * 4 + 1] * vh[1];
        a += Rm[r * 4 + 2] * vh[2];
        a += Rm[r * 4 + 3] * vh[3];
        Fhat[r] = a;
    }
}

// ---------------------------------------------------------------------------
// Learned troubled-stencil switch (no reference counterpart; this oracle
// defines it). Pinned by closed forms in tests/test_oracle_switch.py: the
// upwind-5 formula equals WENO5 with omega == d and is exact on polynomials of
// degree <= 4; features are invariant to v -> a v + b; labels are 0 on smooth
// data and 1 across a jump; the MLP gradient equals central differences.
// ---------------------------------------------------------------------------
template <class R>
static inline R linear5(R a, R b, R c, R d, R e) {
    R t = R(2.0) * a;
    t = t - R(13.0) * b;
    t = t + R(47.0) * c;
    t = t + R(27.0) * d;
    t = t - R(3.0) * e;
    return t * R(1.0 / 60.0);
}

// Mean-centred 5-point stencil scaled by 1 / (max |deviation| + 1e-10), as floats.
static inline void stencil_features(const double w[5], float* x) {
    double m = w[0];
    for (int j = 1; j < 5; ++j) m = m + w[j];
    m = m * 0.2;
    double d[5], s = 0.0;
    for (int j = 0; j < 5; ++j) {
        d[j] = w[j] - m;
        const double a = std::fabs(d[j]);
        if (j == 0 || a > s) s = a;
    }
    const double inv = 1.0 / (s + 1e-10);
    for (int j = 0; j < 5; ++j) x[j] = (float)(d[j] * inv);
}

// max_j |omega_j - d_j| for WENO-Z on (a..e), include/Utility.hpp:457-496.
static inline double weno_z_departure(double a, double b, double c, double d, double e) {
    double b0, b1, b2;
    smoothness<double>(a, b, c, d, e, b0, b1, b2);
    const double eps = 1e-20;
    const double i0 = 1.0 / (b0 + eps), i1 = 1.0 / (b1 + eps), i2 = 1.0 / (b2 + eps);
    const double tau = std::fabs(b0 - b2);
    const double a0 = 0.1 * (1.0 + (((tau * i0) * tau) * i0));
    const double a1 = 0.6 * (1.0 + (((tau * i1) * tau) * i1));
    const double a2 = 0.3 * (1.0 + (((tau * i2) * tau) * i2));
    const double sum = (a0 + a1) + a2;
    const double e0 = std::fabs(a0 / sum - 0.1), e1 = std::fabs(a1 / sum - 0.6), e2 = std::fabs(a2 / sum - 0.3);
    double m = e0;
    if (e1 > m) m = e1;
    if (e2 > m) m = e2;
    return m;
}

// Characteristic stencils of one face (the first half of face_flux).
static inline void face_char(const double* Ul, const double* Ur, const double (*Fp5)[4], const double (*Fm5)[4],
                             int dir, double gamma, double vp[5][4], double vm[5][4]) {
    double ave[4], L[16];
    for (int r = 0; r < 4; ++r) ave[r] = 0.5 * (Ul[r] + Ur[r]);
    eigen_left(ave, dir, gamma, L);
    for (int s = 0; s < 5; ++s)
        for (int i = 0; i < 4; ++i) {
            double a = L[i * 4 + 0] * Fp5[s][0];
            a += L[i * 4 + 1] * Fp5[s][1];
            a += L[i * 4 + 2] * Fp5[s][2];
            a += L[i * 4 + 3] * Fp5[s][3];
            vp[s][i] = a;
            double b = L[i * 4 + 0] * Fm5[s][0];
            b += L[i * 4 + 1] * Fm5[s][1];
            b += L[i * 4 + 2] * Fm5[s][2];
            b += L[i * 4 + 3] * Fm5[s][3];
            vm[s][i] = b;
        }
}

// Label of a stencil pair: WENO-Z's weights leave the linear ones by more
// than this on either side.
static const double kLabelThreshold = 0.2;

// Features x[k][0..9] (+ stencil, then the mirrored - stencil) and labels of a face.
static inline void face_samples(const double vp[5][4], const double vm[5][4], float* x40, unsigned char* lab4) {
    for (int k = 0; k < 4; ++k) {
        const double wp[5] = {vp[0][k], vp[1][k], vp[2][k], vp[3][k], vp[4][k]};
        const double wm[5] = {vm[4][k], vm[3][k], vm[2][k], vm[1][k], vm[0][k]};
        stencil_features(wp, x40 + 10 * k);
        stencil_features(wm, x40 + 10 * k + 5);
        const double dp = weno_z_departure(wp[0], wp[1], wp[2], wp[3], wp[4]);
        const double dm = weno_z_departure(wm[0], wm[1], wm[2], wm[3], wm[4]);
        lab4[k] = (dp > kLabelThreshold || dm > kLabelThreshold) ? 1 : 0;
    }
}

// Component-wise linear upwind-5 face flux (the switch's cheap branch), as
// the sum of the six points' contributions in point order: with Fp5[s] = F+ at
// point f-3+s and Fm5[s] = F- at point f-2+s, point f+o (o = -3..2) adds
// t_o = c+_o F+ + c-_o F-, c+ = (2, -13, 47, 27, -3, 0) and c- = (0, -3, 27,
// 47, -13, 2) (upwind-5 on F+ and its mirror on F-); the flux is the sum
// times 1/60. The first contribution is assigned, the others added; a zero
// coefficient is left out.
template <class R>
static inline void face_flux_linear(const R (*Fp5)[4], const R (*Fm5)[4], R* Fhat) {
    for (int r = 0; r < 4; ++r) {
        R acc = R(2.0) * Fp5[0][r];                        // o = -3
        acc = acc + (R(-13.0) * Fp5[1][r] + R(-3.0) * Fm5[0][r]);  // o = -2
        acc = acc + (R(47.0) * Fp5[2][r] + R(27.0) * Fm5[1][r]);   // o = -1
        acc = acc + (R(27.0) * Fp5[3][r] + R(47.0) * Fm5[2][r]);   // o = 0
        acc = acc + (R(-3.0) * Fp5[4][r] + R(-13.0) * Fm5[3][r]);  // o = 1
        acc = acc + R(2.0) * Fm5[4][r];                    // o = 2
        Fhat[r] = acc * R(1.0 / 60.0);
    }
}

// Hybrid face flux. Bit k of mask flags characteristic field k: the flux is
// the cheap flux plus, for each flagged field in ascending k, R[:,k] (w_k -
// l_k), where w_k is the scheme's WENO5 reconstruction of field k (the +
// stencil plus the mirrored - stencil, as face_flux) and l_k the linear
// upwind-5 reconstruction of the same projected stencils (WENO with omega =
// d). With every field flagged this is face_flux up to rounding (R L = I);
// with none it is the cheap flux. The corrections are summed first, then
// added to the cheap flux once.
template <class R>
static inline void hybrid_face(int mask, const R* Ul, const R* Ur, const R (*Fp5)[4], const R (*Fm5)[4], int dir,
                               int scheme, double gamma, double zp_lambda, R* Fhat) {
    face_flux_linear<R>(Fp5, Fm5, Fhat);
    if (!mask) return;
    R ave[4], L[16], Rm[16];
    for (int r = 0; r < 4; ++r) ave[r] = R(0.5) * (Ul[r] + Ur[r]);
    eigen_left(ave, dir, gamma, L);
    eigen_right(ave, dir, gamma, Rm);
    R corr[4];
    bool first = true;
    for (int k = 0; k < 4; ++k) {
        if (!(mask >> k & 1)) continue;
        R vp[5], vm[5];
        for (int s = 0; s < 5; ++s) {
            R a = L[k * 4 + 0] * Fp5[s][0];
            a += L[k * 4 + 1] * Fp5[s][1];
            a += L[k * 4 + 2] * Fp5[s][2];
            a += L[k * 4 + 3] * Fp5[s][3];
            vp[s] = a;
            R b = L[k * 4 + 0] * Fm5[s][0];
            b += L[k * 4 + 1] * Fm5[s][1];
            b += L[k * 4 + 2] * Fm5[s][2];
            b += L[k * 4 + 3] * Fm5[s][3];
            vm[s] = b;
        }
        const R w = weno5<R>(scheme, vp[0], vp[1], vp[2], vp[3], vp[4], zp_lambda) +
                    weno5<R>(scheme, vm[4], vm[3], vm[2], vm[1], vm[0], zp_lambda);
        const R l = linear5<R>(vp[0], vp[1], vp[2], vp[3], vp[4]) + linear5<R>(vm[4], vm[3], vm[2], vm[1], vm[0]);
        const R d = w - l;
        for (int r = 0; r < 4; ++r) {
            const R c = Rm[r * 4 + k] * d;
            corr[r] = first ? c : corr[r] + c;
        }
        first = false;
    }
    for (int r = 0; r < 4; ++r) Fhat[r] = Fhat[r] + corr[r];
}

// MLP 10 -> 32 -> 32 -> 1, ReLU, parameters as floats:
// W1[32][10], b1[32], W2[32][32], b2[32], w3[32], b3.
enum { P_W1 = 0, P_B1 = 320, P_W2 = 352, P_B2 = 1376, P_W3 = 1408, P_B3 = 1440, P_N = 1441 };

static inline double relu(double x) { return x > 0.0 ? x : 0.0; }
static inline double relu_grad(double x) { return x > 0.0 ? 1.0 : 0.0; }

struct MlpCache {
    double a1[32], h1[32], a2[32], h2[32], z;
};

static inline double mlp_forward(const double* P, const float* x, MlpCache& c) {
    for (int u = 0; u < 32; ++u) {
        double a = P[P_B1 + u];
        for (int k = 0; k < 10; ++k) a += P[P_W1 + u * 10 + k] * (double)x[k];
        c.a1[u] = a;
        c.h1[u] = relu(a);
    }
    for (int u = 0; u < 32; ++u) {
        double a = P[P_B2 + u];
        for (int k = 0; k < 32; ++k) a += P[P_W2 + u * 32 + k] * c.h1[k];
        c.a2[u] = a;
        c.h2[u] = relu(a);
    }
    double z = P[P_B3];
    for (int k = 0; k < 32; ++k) z += P[P_W3 + k] * c.h2[k];
    c.z = z;
    return z;
}

// Weighted binary cross-entropy with logits: pw y softplus(-z) + (1 - y) softplus(z).
static inline double softplus(double z) { return (z > 0 ? z : 0.0) + std::log1p(std::exp(-std::fabs(z))); }

// Mean loss over n samples and its gradient (accumulated into G, size P_N).
static double mlp_loss_grad(const double* P, const float* X, const unsigned char* Y, int n, double pw, double* G) {
    double loss = 0;
    for (int q = 0; q < P_N; ++q) G[q] = 0;
    for (int s = 0; s < n; ++s) {
        MlpCache c;
        const double z = mlp_forward(P, X + 10 * (size_t)s, c);
        const double y = Y[s] ? 1.0 : 0.0;
        loss += pw * y * softplus(-z) + (1.0 - y) * softplus(z);
        const double sg = 1.0 / (1.0 + std::exp(-z));
        const double dz = (1.0 - y) * sg - pw * y * (1.0 - sg);
        double da2[32], dh1[32] = {0};
        G[P_B3] += dz;
        for (int k = 0; k < 32; ++k) {
            G[P_W3 + k] += dz * c.h2[k];
            da2[k] = dz * P[P_W3 + k] * relu_grad(c.a2[k]);
        }
        for (int u = 0; u < 32; ++u) {
            G[P_B2 + u] += da2[u];
            for (int k = 0; k < 32; ++k) {
                G[P_W2 + u * 32 + k] += da2[u] * c.h1[k];
                dh1[k] += P[P_W2 + u * 32 + k] * da2[u];
            }
        }
        for (int u = 0; u < 32; ++u) {
            const double da1 = dh1[u] * relu_grad(c.a1[u]);
            G[P_B1 + u] += da1;
            for (int k = 0; k < 10; ++k) G[P_W1 + u * 10 + k] += da1 * (double)X[10 * (size_t)s + k];
        }
    }
    for (int q = 0; q < P_N; ++q) G[q] /= n;
    return loss / n;
}

// bfloat16 round-to-nearest-even of a float (finite inputs).
static inline float bf16_round(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}

// Inference as the GPU runs it: features, hidden activations and weights
// rounded to bfloat16, products accumulated in float, biases in float.
static inline float mlp_infer_bf16(const float* P, const float* x) {
    float xb[10], h1[32], h2[32];
    for (int k = 0; k < 10; ++k) xb[k] = bf16_round(x[k]);
    for (int u = 0; u < 32; ++u) {
        float a = 0;
        for (int k = 0; k < 10; ++k) a += bf16_round(P[P_W1 + u * 10 + k]) * xb[k];
        a += P[P_B1 + u];
        h1[u] = bf16_round((float)relu((double)a));
    }
    for (int u = 0; u < 32; ++u) {
        float a = 0;
        for (int k = 0; k < 32; ++k) a += bf16_round(P[P_W2 + u * 32 + k]) * h1[k];
        a += P[P_B2 + u];
        h2[u] = (float)relu((double)a);
    }
    float z = 0;
    for (int k = 0; k < 32; ++k) z += P[P_W3 + k] * h2[k];
    return z + P[P_B3];
}

// ---------------------------------------------------------------------------
// Test cases. utils/equations.hpp:400-662 plus the two cases that have no
// reference counterpart (isentropic vortex, parameterised four-quadrant).
// ---------------------------------------------------------------------------
enum { BC_PERIOD = 0, BC_SLIP = 1, BC_SPECIAL = 2, BC_SYMMETRIC = 3 };
enum { TC_SMOOTH = 0, TC_RIEMANN1 = 1, TC_RIEMANN2 = 2, TC_RTI = 3, TC_VORTEX = 4, TC_QUADRANT = 5 };

struct Case {
    double xL, xR, yL, yR, t_end, gamma;
    int bc_left, bc_right, bc_bottom, bc_top;
    bool has_exact;
    // extension parameters
    double vortex_eps = 5.0, vortex_xc = 5.0, vortex_yc = 5.0;
    double quad_sx = 0.5, quad_sy = 0.5;
    double quad[4][4] = {};  // [UR, UL, LL, LR][rho, u, v, p]
    int testcase = 0;

    void prim0(double x, double y, double& rho, double& u, double& v, double& p) const;
    double exact_rho(double x, double y, double t) const;
};

static const double kPi = 3.14159265358979323846;  // M_PI

void Case::prim0(double x, double y, double& rho, double& u, double& v, double& p) const {
    switch (testcase) {
        case TC_SMOOTH:  // equations.hpp:419-434
            rho = 1 + 0.2 * std::sin(kPi * (x + y));
            u = 0.7;
            v = 0.3;
            p = 1;
            return;
        case TC_RIEMANN1: {  // equations.hpp:458-513
            if (x < 1.0 && y < 1.0) { rho = 0.8; u = 0.0; v = 0.0; p = 1.0; }
            else if (x < 1.0 && y >= 1.0) { rho = 1.0; u = 0.7276; v = 0.0; p = 1.0; }
            else if (x >= 1.0 && y < 1.0) { rho = 1.0; u = 0.0; v = 0.7276; p = 1.0; }
            else { rho = 0.5313; u = 0.0; v = 0.0; p = 0.4; }
            return;
        }
        case TC_RIEMANN2: {  // equations.hpp:537-592
            if (x < 0.8 && y < 0.8) { rho = 0.138; u = 1.206; v = 1.206; p = 0.029; }
            else if (x < 0.8 && y >= 0.8) { rho = 0.5323; u = 1.206; v = 0.0; p = 0.3; }
            else if (x >= 0.8 && y < 0.8) { rho = 0.5323; u = 0.0; v = 1.206; p = 0.3; }
            else { rho = 1.5; u = 0.0; v = 0.0; p = 1.5; }
            return;
        }
        case TC_RTI: {  // equations.hpp:616-641
            rho = (y <= 0.5) ? 2.0 : 1.0;
            u = 0;
            p = (y <= 0.5) ? 1 + 2 * y : y + 1.5;
            const double c = std::sqrt(gamma * p / rho);
            v = -0.025 * c * std::cos(8 * kPi * x);
            return;
        }
        case TC_VORTEX: {  // Shu 1997 isentropic vortex, background (1,1,1,1)
            const double xb = x - vortex_xc, yb = y - vortex_yc;
            const double r2 = xb * xb + yb * yb;
            const double amp = vortex_eps / (2.0 * kPi) * std::exp(0.5 * (1.0 - r2));
            const double T = 1.0 - (gamma - 1.0) * vortex_eps * vortex_eps / (8.0 * gamma * kPi * kPi) *
                                       std::exp(1.0 - r2);
            rho = std::pow(T, 1.0 / (gamma - 1.0));
            u = 1.0 - amp * yb;
            v = 1.0 + amp * xb;
            p = std::pow(rho, gamma);
            return;
        }
        default: {  // TC_QUADRANT: UR (x>=sx,y>=sy), UL, LL, LR
            int q;
            if (x < quad_sx && y < quad_sy) q = 2;
            else if (x < quad_sx && y >= quad_sy) q = 1;
            else if (x >= quad_sx && y < quad_sy) q = 3;
            else q = 0;
            rho = quad[q][0]; u = quad[q][1]; v = quad[q][2]; p = quad[q][3];
            return;
        }
    }
}

double Case::exact_rho(double x, double y, double t) const {
    if (testcase == TC_SMOOTH) return 1 + 0.2 * std::sin(kPi * (x + y - t));  // equations.hpp:438-439
    if (testcase == TC_VORTEX) {
        const double Lx = xR - xL, Ly = yR - yL;
        double xb = x - t - vortex_xc, yb = y - t - vortex_yc;
        xb -= Lx * std::floor(xb / Lx + 0.5);  // minimum periodic image in [-L/2, L/2)
        yb -= Ly * std::floor(yb / Ly + 0.5);
        const double r2 = xb * xb + yb * yb;
        const double T = 1.0 - (gamma - 1.0) * vortex_eps * vortex_eps / (8.0 * gamma * kPi * kPi) *
                                   std::exp(1.0 - r2);
        return std::pow(T, 1.0 / (gamma - 1.0));
    }
    return 0.0;
}

static void setup_case(Case& c, int tc) {
    c.testcase = tc;
    switch (tc) {
        case TC_SMOOTH:
            c.xL = 0; c.xR = 2; c.yL = 0; c.yR = 2; c.t_end = 1.0; c.gamma = 1.4;
            c.bc_left = c.bc_right = c.bc_bottom = c.bc_top = BC_PERIOD; c.has_exact = true; break;
        case TC_RIEMANN1:
            c.xL = 0; c.xR = 2; c.yL = 0; c.yR = 2; c.t_end = 0.52; c.gamma = 1.4;
            c.bc_left = c.bc_right = c.bc_bottom = c.bc_top = BC_SYMMETRIC; c.has_exact = false; break;
        case TC_RIEMANN2:
            c.xL = 0; c.xR = 1; c.yL = 0; c.yR = 1; c.t_end = 0.8; c.gamma = 1.4;
            c.bc_left = c.bc_right = c.bc_bottom = c.bc_top = BC_SYMMETRIC; c.has_exact = false; break;
        case TC_RTI:
            c.xL = 0; c.xR = 0.25; c.yL = 0; c.yR = 1; c.t_end = 1.95; c.gamma = 5.0 / 3.0;
            c.bc_left = c.bc_right = BC_SLIP; c.bc_bottom = c.bc_top = BC_SPECIAL; c.has_exact = false; break;
        case TC_VORTEX:
            c.xL = 0; c.xR = 10; c.yL = 0; c.yR = 10; c.t_end = 2.0; c.gamma = 1.4;
            c.bc_left = c.bc_right = c.bc_bottom = c.bc_top = BC_PERIOD; c.has_exact = true; break;
        default: {
            c.xL = 0; c.xR = 1; c.yL = 0; c.yR = 1; c.t_end = 0.3; c.gamma = 1.4;
            c.bc_left = c.bc_right = c.bc_bottom = c.bc_top = BC_SYMMETRIC; c.has_exact = false;
            const double s[4][4] = {{1.5, 0, 0, 1.5}, {0.5323, 1.206, 0, 0.3},
                                    {0.138, 1.206, 1.206, 0.029}, {0.5323, 0, 1.206, 0.3}};
            std::memcpy(c.quad, s, sizeof(s));
        }
    }
}

// splitmix64 -> uniform double in [0,1) with 53 random bits.
static inline uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static inline double u01(uint64_t& s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

// Seeded four-quadrant states: for q in (UR, UL, LL, LR): rho~U[0.1,2],
// u~U[-1.5,1.5], v~U[-1.5,1.5], p~U[0.05,2], drawn in that order.
static void seeded_quadrant(uint64_t seed, double out[4][4]) {
    uint64_t s = seed;
    for (int q = 0; q < 4; ++q) {
        out[q][0] = 0.1 + 1.9 * u01(s);
        out[q][1] = -1.5 + 3.0 * u01(s);
        out[q][2] = -1.5 + 3.0 * u01(s);
        out[q][3] = 0.05 + 1.95 * u01(s);
    }
}

// ---------------------------------------------------------------------------
// Solver. src/WENOFD.cpp:10-465 restated for one rank holding the whole domain.
// ---------------------------------------------------------------------------
template <class R>
struct Solver {
    Case cs;
    int scheme = 1, rk = 2;
    double cfl = 0.5;
    int NX = 0, NY = 0, TX = 0, TY = 0;
    const int g = 4;
    double dx = 0, dy = 0, zp_lambda = 0, now = 0;
    std::vector<double> xc, yc;  // node coordinates incl. ghosts
    std::vector<R> Uh, Un, rhs;
    // Switch masks (null: every face takes face_flux): x faces [NY][NX+1]
    // (row j-4, face f-4), y faces [NY+1][NX] (face f-4, column i-4); bit k
    // flags characteristic field k (hybrid_face).
    const unsigned char* flags_x = nullptr;
    const unsigned char* flags_y = nullptr;

    size_t idx(int i, int j) const { return ((size_t)i * TY + j) * 4; }

    void init(int nx, int ny) {
        NX = nx; NY = ny; TX = NX + 2 * g; TY = NY + 2 * g;
        dx = (cs.xR - cs.xL) / (NX - 1);  // WENOFD.cpp:73-74
        dy = (cs.yR - cs.yL) / (NY - 1);
        zp_lambda = std::pow(dx, 2.0 / 3.0);  // Utility.hpp:531, with m_deltaX (WENOFD.cpp:836)
        const double gxl = (cs.xL + 0 * dx) - dx * g;  // WENOFD.cpp:76-80 (rank 0)
        const double gyl = cs.yL - dy * g;
        xc.resize(TX); yc.resize(TY);
        for (int i = 0; i < TX; ++i) xc[i] = gxl + i * dx;  // :100-101
        for (int j = 0; j < TY; ++j) yc[j] = gyl + j * dy;
        Uh.assign((size_t)TX * TY * 4, R(0.0));
        Un = Uh;
        rhs = Uh;
        now = 0;
    }

    void u0(double x, double y, double* U) const {  // equations.hpp:649-662
        double rho, u, v, p;
        cs.prim0(x, y, rho, u, v, p);
        const double E = p / (cs.gamma - 1.0) + 0.5 * rho * (u * u + v * v);
        U[0] = rho; U[1] = rho * u; U[2] = rho * v; U[3] = E;
    }

    void init_ave() {  // WENOFD.cpp:165-188, interior points only
        for (int i = g; i < g + NX; ++i)
            for (int j = g; j < g + NY; ++j) {
                double U[4];
                u0(xc[i], yc[j], U);
                for (int r = 0; r < 4; ++r) Uh[idx(i, j) + r] = R(U[r]);
            }
    }

    // Periodic-x halo of the one-rank ring exchange (WENOFD.cpp:850-929):
    // interior rows only, runs for every boundary type.
    void exchange() {
        const int sx = g, ex = g + NX;
        for (int j = g; j < g + NY; ++j)
            for (int e = 0; e < g; ++e)
                for (int r = 0; r < 4; ++r) {
                    Uh[idx(sx - g + e, j) + r] = Uh[idx(ex - g + e - 1, j) + r];
                    Uh[idx(ex + e, j) + r] = Uh[idx(sx + e + 1, j) + r];
                }
    }

    void set_boundary() {  // WENOFD.cpp:998-1159
        const int sx = g, ex = g + NX, sy = g, ey = g + NY;
        double Ub[4];
        for (int i = 0; i < TX; ++i)
            for (int e1 = 1; e1 <= g; ++e1) {
                R* lo = &Uh[idx(i, sy - e1)];
                R* hi = &Uh[idx(i, ey - 1 + e1)];
                switch (cs.bc_bottom) {  // bottom BC governs both y sides (:1009)
                    case BC_PERIOD:
                        for (int r = 0; r < 4; ++r) { lo[r] = Uh[idx(i, ey - 1 - e1) + r]; hi[r] = Uh[idx(i, sy + e1) + r]; }
                        break;
                    case BC_SYMMETRIC:
                    case BC_SLIP:
                        for (int r = 0; r < 4; ++r) { lo[r] = Uh[idx(i, sy + e1) + r]; hi[r] = Uh[idx(i, ey - 1 - e1) + r]; }
                        if (cs.bc_bottom == BC_SLIP) { lo[2] = lo[2] * R(-1.0); hi[2] = hi[2] * R(-1.0); }
                        break;
                    case BC_SPECIAL:
                        u0(xc[i], yc[sy], Ub);
                        for (int r = 0; r < 4; ++r) lo[r] = R(Ub[r]);
                        u0(xc[i], yc[ey - 1], Ub);
                        for (int r = 0; r < 4; ++r) hi[r] = R(Ub[r]);
                        break;
                }
            }
        for (int j = 0; j < TY; ++j)
            for (int e1 = 1; e1 <= g; ++e1) {
                R* lo = &Uh[idx(sx - e1, j)];
                switch (cs.bc_left) {
                    case BC_SYMMETRIC:
                    case BC_SLIP:
                        for (int r = 0; r < 4; ++r) lo[r] = Uh[idx(sx + e1, j) + r];
                        if (cs.bc_left == BC_SLIP) lo[1] = lo[1] * R(-1.0);
                        break;
                    case BC_SPECIAL:
                        u0(xc[sx], yc[j], Ub);
                        for (int r = 0; r < 4; ++r) lo[r] = R(Ub[r]);
                        break;
                    default: break;
                }
            }
        for (int j = 0; j < TY; ++j)
            for (int e1 = 1; e1 <= g; ++e1) {
                R* hi = &Uh[idx(ex - 1 + e1, j)];
                switch (cs.bc_right) {
                    case BC_SYMMETRIC:
                    case BC_SLIP:
                        for (int r = 0; r < 4; ++r) hi[r] = Uh[idx(ex - 1 - e1, j) + r];
                        if (cs.bc_right == BC_SLIP) hi[1] = hi[1] * R(-1.0);
                        break;
                    case BC_SPECIAL:
                        u0(xc[ex], yc[j], Ub);  // first ghost column's coordinate (:1144)
                        for (int r = 0; r < 4; ++r) hi[r] = R(Ub[r]);
                        break;
                    default: break;
                }
            }
    }

    void fill_ghosts() { exchange(); set_boundary(); }


    // assembleBoundaryTerm minus the halo (WENOFD.cpp:481-798); rhs += flux terms.
    void flux_terms() {
        const int sx = g, ex = g + NX, sy = g, ey = g + NY;
        std::vector<R> Fm((size_t)std::max(TX, TY) * 4), Fp(Fm.size()), Fh((size_t)(std::max(TX, TY) + 1) * 4);
        for (int j = sy; j < ey; ++j) {  // x-direction
            for (int i = 0; i < TX; ++i) flux_split<R>(&Uh[idx(i, j)], 0, cs.gamma, &Fm[i * 4], &Fp[i * 4]);
            for (int f = sx; f <= ex; ++f) {
                R fp5[5][4], fm5[5][4];
                for (int s = -2; s <= 2; ++s)
                    for (int r = 0; r < 4; ++r) {
                        fp5[s + 2][r] = Fp[(f + s - 1) * 4 + r];
                        fm5[s + 2][r] = Fm[(f + s) * 4 + r];
                    }
                if (flags_x)
                    hybrid_face<R>(flags_x[(size_t)(j - sy) * (NX + 1) + (f - sx)] & 15, &Uh[idx(f - 1, j)],
                                   &Uh[idx(f, j)], fp5, fm5, 0, scheme, cs.gamma, zp_lambda, &Fh[f * 4]);
                else
                    face_flux<R>(&Uh[idx(f - 1, j)], &Uh[idx(f, j)], fp5, fm5, 0, scheme, cs.gamma, zp_lambda,
                                 &Fh[f * 4]);
            }
            for (int i = sx; i < ex; ++i)
                for (int r = 0; r < 4; ++r) rhs[idx(i, j) + r] += -(Fh[(i + 1) * 4 + r] - Fh[i * 4 + r]) / R(dx);
        }
        for (int i = sx; i < ex; ++i) {  // y-direction
            for (int j = 0; j < TY; ++j) flux_split<R>(&Uh[idx(i, j)], 1, cs.gamma, &Fm[j * 4], &Fp[j * 4]);
            for (int f = sy; f <= ey; ++f) {
                R fp5[5][4], fm5[5][4];
                for (int s = -2; s <= 2; ++s)
                    for (int r = 0; r < 4; ++r) {
                        fp5[s + 2][r] = Fp[(f + s - 1) * 4 + r];
                        fm5[s + 2][r] = Fm[(f + s) * 4 + r];
                    }
                if (flags_y)
                    hybrid_face<R>(flags_y[(size_t)(f - sy) * NX + (i - sx)] & 15, &Uh[idx(i, f - 1)],
                                   &Uh[idx(i, f)], fp5, fm5, 1, scheme, cs.gamma, zp_lambda, &Fh[f * 4]);
                else
                    face_flux<R>(&Uh[idx(i, f - 1)], &Uh[idx(i, f)], fp5, fm5, 1, scheme, cs.gamma, zp_lambda,
                                 &Fh[f * 4]);
            }
            for (int j = sy; j < ey; ++j)
                for (int r = 0; r < 4; ++r) rhs[idx(i, j) + r] += -(Fh[(j + 1) * 4 + r] - Fh[j * 4 + r]) / R(dy);
        }
    }

    // Switch samples of every face of one direction on the current (ghost-filled)
    // state: x faces in [NY][NX+1] order, y faces in [NY+1][NX] order; per face
    // 40 features ([field][10]) and 4 labels.
    void samples(int dir, float* X, unsigned char* Y) const {
        const int sx = g, ex = g + NX, sy = g, ey = g + NY;
        std::vector<double> Fm((size_t)std::max(TX, TY) * 4), Fp(Fm.size());
        const int lines = dir == 0 ? NY : NX;
        for (int l = 0; l < lines; ++l) {
            const int n = dir == 0 ? TX : TY;
            for (int p = 0; p < n; ++p) {
                const double* U = dir == 0 ? &Uh[idx(p, sy + l)] : &Uh[idx(sx + l, p)];
                flux_split<double>(U, dir, cs.gamma, &Fm[p * 4], &Fp[p * 4]);
            }
            const int f0 = dir == 0 ? sx : sy, f1 = dir == 0 ? ex : ey;
            for (int f = f0; f <= f1; ++f) {
                double fp5[5][4], fm5[5][4], vp[5][4], vm[5][4];
                for (int s = -2; s <= 2; ++s)
                    for (int r = 0; r < 4; ++r) {
                        fp5[s + 2][r] = Fp[(f + s - 1) * 4 + r];
                        fm5[s + 2][r] = Fm[(f + s) * 4 + r];
                    }
                const double* Ul = dir == 0 ? &Uh[idx(f - 1, sy + l)] : &Uh[idx(sx + l, f - 1)];
                const double* Ur = dir == 0 ? &Uh[idx(f, sy + l)] : &Uh[idx(sx + l, f)];
                face_char(Ul, Ur, fp5, fm5, dir, cs.gamma, vp, vm);
                const size_t q = dir == 0 ? (size_t)l * (NX + 1) + (f - f0) : (size_t)(f - f0) * NX + l;
                face_samples(vp, vm, X + 40 * q, Y + 4 * q);
            }
        }
    }

    void source_terms() {  // WENOFD.cpp:800-821 (RTI gravity, g = 1)
        if (cs.testcase != TC_RTI) return;
        for (int i = g; i < g + NX; ++i)
            for (int j = g; j < g + NY; ++j) {
                rhs[idx(i, j) + 3] += Uh[idx(i, j) + 2] * R(1.0);
                rhs[idx(i, j) + 2] += Uh[idx(i, j) + 0] * R(1.0);
            }
    }

    void assemble_rhs() {  // WENOFD.cpp:454-465
        for (int i = g; i < g + NX; ++i)
            for (int j = g; j < g + NY; ++j)
                for (int r = 0; r < 4; ++r) rhs[idx(i, j) + r] = R(0.0);
        fill_ghosts();
        flux_terms();
        source_terms();
    }

    void rk_combine(double a, double b, double c, double dt) {  // WENOFD.cpp:391-394
        for (int i = g; i < g + NX; ++i)
            for (int j = g; j < g + NY; ++j)
                for (int r = 0; r < 4; ++r) {
                    const size_t k = idx(i, j) + r;
                    Uh[k] = (R(a) * Un[k] + R(b) * Uh[k]) + (R(c) * R(dt)) * rhs[k];
                }
    }

    void save_un() {
        for (int i = g; i < g + NX; ++i)
            for (int j = g; j < g + NY; ++j)
                for (int r = 0; r < 4; ++r) Un[idx(i, j) + r] = Uh[idx(i, j) + r];
    }

    int step_rk(int method, double dt) {
        if (method == 0) {  // RunRK1, WENOFD.cpp:301-311
            assemble_rhs();
            for (int i = g; i < g + NX; ++i)
                for (int j = g; j < g + NY; ++j)
                    for (int r = 0; r < 4; ++r) Uh[idx(i, j) + r] += R(dt) * rhs[idx(i, j) + r];
            return 0;
        }
        if (method == 1) {  // RunRK2, :312-351
            save_un();
            assemble_rhs(); rk_combine(1.0, 0.0, 1.0, dt);
            assemble_rhs(); rk_combine(0.5, 0.5, 0.5, dt);
            return 0;
        }
        if (method == 2) {  // RunRK3, :352-396
            save_un();
            assemble_rhs(); rk_combine(1.0, 0.0, 1.0, dt);
            assemble_rhs(); rk_combine(3.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0, dt);
            assemble_rhs(); rk_combine(1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0, dt);
            return 0;
        }
        return -1;
    }

    // calculateDeltaT, WENOFD.cpp:398-452. Returns NaN when the reference would abort.
    double delta_t() {
        R ts(1.0);
        for (int i = g; i < g + NX; ++i)
            for (int j = g; j < g + NY; ++j) {
                const R* U = &Uh[idx(i, j)];
                const R ex_ = max_eigen<R>(U, 1, 0, cs.gamma);
                const R ey_ = max_eigen<R>(U, 0, 1, cs.gamma);
                const R tmp = R(cfl) / (ex_ / R(dx) + ey_ / R(dy));
                if (tmp > ts) continue;
                ts = tmp;
            }
        double t = val(ts);
        if (cs.t_end - now < t) t = cs.t_end - now;
        return t;
    }

    // Error norms of rho, WENOFD.cpp:1308-1382: {Linf, L1, L2}.
    void accuracy(double* out) const {
        double e1 = 0, e2 = 0, einf = 0;
        for (int i = g; i < g + NX; ++i)
            for (int j = g; j < g + NY; ++j) {
                const double ref = cs.exact_rho(xc[i], yc[j], now);
                const double e = std::fabs(ref - val(Uh[idx(i, j)]));
                e1 += e;
                e2 += e * e;
                if (e > einf) einf = e;
            }
        out[0] = einf;
        out[1] = e1 / NX / NY;
        out[2] = std::sqrt(e2 / NX / NY);
    }
};

}  // namespace orc

// ---------------------------------------------------------------------------
// C interface for the tests (ctypes).
// ---------------------------------------------------------------------------
using orc::Solver;

struct OracleHandle {
    Solver<double> s;
};

extern "C" {

double oracle_weno5(int scheme, double a, double b, double c, double d, double e, double zp_lambda) {
    return orc::weno5<double>(scheme, a, b, c, d, e, zp_lambda);
}

void oracle_weno5_batch(int scheme, int n, const double* v5, double zp_lambda, double* out) {
    for (int k = 0; k < n; ++k)
        out[k] = orc::weno5<double>(scheme, v5[5 * k], v5[5 * k + 1], v5[5 * k + 2], v5[5 * k + 3], v5[5 * k + 4],
                                    zp_lambda);
}

void oracle_eigen(int dir, const double* U, double gamma, double* L, double* R) {
    orc::eigen_left<double>(U, dir, gamma, L);
    orc::eigen_right<double>(U, dir, gamma, R);
}

void oracle_split(int dir, const double* U, double gamma, double* Fm, double* Fp) {
    orc::flux_split<double>(U, dir, gamma, Fm, Fp);
}

void oracle_phys_flux(int dir, const double* U, double gamma, double* F) { orc::phys_flux<double>(U, dir, gamma, F); }

double oracle_max_eigen(const double* U, double nx, double ny, double gamma) {
    return orc::max_eigen<double>(U, nx, ny, gamma);
}

// Face flux from 6 consecutive points' conserved states U6[6][4] (points f-3..f+2).
void oracle_face_flux(int dir, int scheme, double gamma, double zp_lambda, const double* U6, double* Fhat) {
    double Fm[6][4], Fp[6][4];
    for (int k = 0; k < 6; ++k) orc::flux_split<double>(&U6[4 * k], dir, gamma, Fm[k], Fp[k]);
    double fp5[5][4], fm5[5][4];
    for (int s = 0; s < 5; ++s)
        for (int r = 0; r < 4; ++r) { fp5[s][r] = Fp[s][r]; fm5[s][r] = Fm[s + 1][r]; }
    orc::face_flux<double>(&U6[4 * 2], &U6[4 * 3], fp5, fm5, dir, scheme, gamma, zp_lambda, Fhat);
}

void oracle_seeded_quadrant(unsigned long long seed, double* out16) {
    double q[4][4];
    orc::seeded_quadrant(seed, q);
    std::memcpy(out16, q, sizeof(q));
}

// testcase 0..5; quad16 (optional, may be null) overrides the quadrant states
// for testcase 5; vortex3 (optional) = {eps, xc, yc}; t_end < 0 keeps the case default.
OracleHandle* oracle_create(int testcase, int scheme, int rk, double cfl, int nx, int ny, double t_end,
                            const double* quad16, const double* vortex3) {
    if (testcase < 0 || testcase > 5 || scheme < 0 || scheme > 3 || nx < 5 || ny < 5) return nullptr;
    OracleHandle* h = new OracleHandle();
    orc::setup_case(h->s.cs, testcase);
    if (quad16) std::memcpy(h->s.cs.quad, quad16, sizeof(double) * 16);
    if (vortex3) { h->s.cs.vortex_eps = vortex3[0]; h->s.cs.vortex_xc = vortex3[1]; h->s.cs.vortex_yc = vortex3[2]; }
    if (t_end >= 0) h->s.cs.t_end = t_end;
    h->s.scheme = scheme;
    h->s.rk = rk;
    h->s.cfl = cfl;
    h->s.init(nx, ny);
    h->s.init_ave();
    return h;
}

void oracle_destroy(OracleHandle* h) { delete h; }

void oracle_grid(OracleHandle* h, double* dxdy, double* xc, double* yc) {
    dxdy[0] = h->s.dx; dxdy[1] = h->s.dy; dxdy[2] = h->s.zp_lambda;
    if (xc) std::memcpy(xc, h->s.xc.data(), sizeof(double) * h->s.TX);
    if (yc) std::memcpy(yc, h->s.yc.data(), sizeof(double) * h->s.TY);
}

// Full state incl. ghosts in the oracle's layout [(NX+8)][(NY+8)][4].
void oracle_get_full(OracleHandle* h, double* out) { std::memcpy(out, h->s.Uh.data(), sizeof(double) * h->s.Uh.size()); }
void oracle_set_full(OracleHandle* h, const double* in) { std::memcpy(h->s.Uh.data(), in, sizeof(double) * h->s.Uh.size()); }

// Interior state in reference order [NX][NY][4].
void oracle_get_interior(OracleHandle* h, double* out) {
    Solver<double>& s = h->s;
    for (int i = 0; i < s.NX; ++i)
        for (int j = 0; j < s.NY; ++j)
            for (int r = 0; r < 4; ++r) out[((size_t)i * s.NY + j) * 4 + r] = s.Uh[s.idx(i + 4, j + 4) + r];
}
void oracle_set_interior(OracleHandle* h, const double* in) {
    Solver<double>& s = h->s;
    for (int i = 0; i < s.NX; ++i)
        for (int j = 0; j < s.NY; ++j)
            for (int r = 0; r < 4; ++r) s.Uh[s.idx(i + 4, j + 4) + r] = in[((size_t)i * s.NY + j) * 4 + r];
}

void oracle_fill_ghosts(OracleHandle* h) { h->s.fill_ghosts(); }

// assembleRHS on the current interior state; rhs interior out in [NX][NY][4].
void oracle_assemble_rhs(OracleHandle* h, double* rhs_out) {
    Solver<double>& s = h->s;
    s.assemble_rhs();
    if (rhs_out)
        for (int i = 0; i < s.NX; ++i)
            for (int j = 0; j < s.NY; ++j)
                for (int r = 0; r < 4; ++r) rhs_out[((size_t)i * s.NY + j) * 4 + r] = s.rhs[s.idx(i + 4, j + 4) + r];
}

double oracle_delta_t(OracleHandle* h) { return h->s.delta_t(); }

int oracle_step(OracleHandle* h, int rk, double dt) {
    int rc = h->s.step_rk(rk, dt);
    h->s.now += dt;
    return rc;
}

// Runs the reference time loop (WENOFD.cpp:213-284). max_steps > 0 bounds it;
// dt_fixed > 0 replaces calculateDeltaT. Returns the number of steps, or -1
// on a NaN time step.
int oracle_run(OracleHandle* h, int max_steps, double dt_fixed) {
    Solver<double>& s = h->s;
    int count = 0;
    while (std::fabs(s.now - s.cs.t_end) > 1e-9) {
        if (max_steps > 0 && count >= max_steps) break;
        double dt = dt_fixed > 0 ? dt_fixed : s.delta_t();
        if (dt != dt) return -1;
        s.step_rk(s.rk, dt);
        s.now += dt;
        ++count;
    }
    return count;
}

double oracle_now(OracleHandle* h) { return h->s.now; }

// ---- learned switch -------------------------------------------------------
double oracle_linear5(double a, double b, double c, double d, double e) { return orc::linear5<double>(a, b, c, d, e); }

void oracle_stencil_features(const double* w5, float* out5) { orc::stencil_features(w5, out5); }

double oracle_weno_z_departure(const double* w5) {
    return orc::weno_z_departure(w5[0], w5[1], w5[2], w5[3], w5[4]);
}

// Ghost fill, then features [faces][4][10] and labels [faces][4] of every face
// of direction dir (x: [NY][NX+1] faces, y: [NY+1][NX] faces).
void oracle_samples(OracleHandle* h, int dir, float* X, unsigned char* Y) {
    h->s.fill_ghosts();
    h->s.samples(dir, X, Y);
}

// Hybrid face flux (hybrid_face) from 6 consecutive points' conserved states
// U6[6][4] (points f-3..f+2) and a field mask.
void oracle_hybrid_face(int dir, int scheme, double gamma, double zp_lambda, int mask, const double* U6,
                        double* Fhat) {
    double Fm[6][4], Fp[6][4];
    for (int k = 0; k < 6; ++k) orc::flux_split<double>(&U6[4 * k], dir, gamma, Fm[k], Fp[k]);
    double fp5[5][4], fm5[5][4];
    for (int s = 0; s < 5; ++s)
        for (int r = 0; r < 4; ++r) { fp5[s][r] = Fp[s][r]; fm5[s][r] = Fm[s + 1][r]; }
    orc::hybrid_face<double>(mask, &U6[4 * 2], &U6[4 * 3], fp5, fm5, dir, scheme, gamma, zp_lambda, Fhat);
}

// The cheap flux of one component from split values: Fp6 = F+ at points
// f-3..f+1 (5 values), Fm5 = F- at points f-2..f+2.
double oracle_linear_face(const double* Fp5, const double* Fm5) {
    double fp[5][4] = {}, fm[5][4] = {}, out[4];
    for (int s = 0; s < 5; ++s) { fp[s][0] = Fp5[s]; fm[s][0] = Fm5[s]; }
    orc::face_flux_linear<double>(fp, fm, out);
    return out[0];
}

// Switch masks for the hybrid flux (arrays stay owned by the caller and must
// outlive their use); null pointers restore the plain solver.
void oracle_set_flags(OracleHandle* h, const unsigned char* fx, const unsigned char* fy) {
    h->s.flags_x = fx;
    h->s.flags_y = fy;
}

double oracle_mlp_loss_grad(const float* P, const float* X, const unsigned char* Y, int n, double pos_weight,
                            double* G) {
    double Pd[orc::P_N];
    for (int q = 0; q < orc::P_N; ++q) Pd[q] = P[q];
    return orc::mlp_loss_grad(Pd, X, Y, n, pos_weight, G);
}

// Same as oracle_mlp_loss_grad with double parameters (finite-difference checks).
double oracle_mlp_loss_grad_d(const double* P, const float* X, const unsigned char* Y, int n, double pos_weight,
                              double* G) {
    return orc::mlp_loss_grad(P, X, Y, n, pos_weight, G);
}

void oracle_mlp_forward(const float* P, const float* X, int n, double* z) {
    double Pd[orc::P_N];
    for (int q = 0; q < orc::P_N; ++q) Pd[q] = P[q];
    for (int s = 0; s < n; ++s) {
        orc::MlpCache c;
        z[s] = orc::mlp_forward(Pd, X + 10 * (size_t)s, c);
    }
}

void oracle_mlp_infer_bf16(const float* P, const float* X, int n, float* z) {
    for (int s = 0; s < n; ++s) z[s] = orc::mlp_infer_bf16(P, X + 10 * (size_t)s);
}

void oracle_accuracy(OracleHandle* h, double* out3) { h->s.accuracy(out3); }

// Counts the floating-point operations of one RK step (rk method of the
// handle) on an nx x ny grid of the given case, plus one calculateDeltaT.
// out: {add, mul, div, sqrt, cmp} for the whole step.
void oracle_count_ops(int testcase, int scheme, int rk, int nx, int ny, unsigned long long* out) {
    orc::Solver<orc::OpCount> s;
    orc::setup_case(s.cs, testcase);
    s.scheme = scheme;
    s.rk = rk;
    s.cfl = 0.5;
    s.init(nx, ny);
    s.init_ave();
    orc::g_cnt = orc::Counters();
    double dt = s.delta_t();
    s.step_rk(rk, dt * 0.5);
    out[0] = orc::g_cnt.add; out[1] = orc::g_cnt.mul; out[2] = orc::g_cnt.div;
    out[3] = orc::g_cnt.sqrt_; out[4] = orc::g_cnt.cmp;
}

}  // extern "C"
