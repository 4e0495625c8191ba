// wenofd_math.hpp — FP64 device arithmetic of the WENO-FD hot path.
//
// Every expression keeps the floating-point operation order of the reference
// solver (infinity-07/FD-WENO-EulerEquation) so that, compiled without FMA
// contraction (-ffp-contract=off) and with IEEE division and square root, the
// results equal a no-contraction build of the reference bit for bit. Terms
// whose coefficient is an exact 0 or +-1 in the reference's 4-term sums are
// folded (x + 0*y == x and 1*y == y in IEEE arithmetic up to the sign of a
// zero result), which removes ~35% of the projection work without changing a
// single nonzero result.
#pragma once

#include <hip/hip_runtime.h>

namespace wenofd {

enum Scheme { kJS = 0, kZ = 1, kZP = 2, kZPI = 3 };

// include/Utility.hpp:5-11
constexpr double kC1_3 = 1.0 / 3.0;
constexpr double kC1_6 = 1.0 / 6.0;
constexpr double kC5_6 = 5.0 / 6.0;
constexpr double kC7_6 = 7.0 / 6.0;
constexpr double kC11_6 = 11.0 / 6.0;
constexpr double kC13_12 = 13.0 / 12.0;

// Smoothness indicators beta_0..2 of the three 3-point sub-stencils of
// (a,b,c,d,e) = (f_{j-2},...,f_{j+2}); include/Utility.hpp:427-437.
__device__ __forceinline__ void weno_beta(double a, double b, double c, double d, double e, double& b0, double& b1,
                                          double& b2) {
    double t1 = (a - 2.0 * b) + c;
    double t2 = (a - 4.0 * b) + 3.0 * c;
    b0 = (kC13_12 * t1) * t1 + (0.25 * t2) * t2;
    t1 = (b - 2.0 * c) + d;
    t2 = b - d;
    b1 = (kC13_12 * t1) * t1 + (0.25 * t2) * t2;
    t1 = (c - 2.0 * d) + e;
    t2 = (3.0 * c - 4.0 * d) + e;
    b2 = (kC13_12 * t1) * t1 + (0.25 * t2) * t2;
}

// Third-order candidate values at j+1/2; include/Utility.hpp:451-453.
__device__ __forceinline__ void weno_candidates(double a, double b, double c, double d, double e, double& p0,
                                                double& p1, double& p2) {
    p0 = (kC1_3 * a - kC7_6 * b) + kC11_6 * c;
    p1 = (-kC1_6 * b + kC5_6 * c) + kC1_3 * d;
    p2 = (kC1_3 * c + kC5_6 * d) - kC1_6 * e;
}

// Unnormalised nonlinear weights of one scheme (include/Utility.hpp:439-442,
// 477-491, 530-542, 584-599).
template <int S>
__device__ __forceinline__ void weno_alpha(double b0, double b1, double b2, double zp_lambda, double& w0, double& w1,
                                           double& w2) {
    if constexpr (S == kJS) {
        const double eps = 1.0e-6;
        w0 = 0.1 / ((eps + b0) * (eps + b0));
        w1 = 0.6 / ((eps + b1) * (eps + b1));
        w2 = 0.3 / ((eps + b2) * (eps + b2));
    } else if constexpr (S == kZ) {
        const double eps = 1.0e-20;
        const double i0 = 1.0 / (b0 + eps);
        const double i1 = 1.0 / (b1 + eps);
        const double i2 = 1.0 / (b2 + eps);
        const double tau = fabs(b0 - b2);
        w0 = 0.1 * (1.0 + (((tau * i0) * tau) * i0));
        w1 = 0.6 * (1.0 + (((tau * i1) * tau) * i1));
        w2 = 0.3 * (1.0 + (((tau * i2) * tau) * i2));
    } else if constexpr (S == kZP) {
        const double eps = 1.0e-40;
        const double tau = fabs(b0 - b2);
        const double te = tau + eps;
        const double q0 = te / (b0 + eps);
        const double q1 = te / (b1 + eps);
        const double q2 = te / (b2 + eps);
        w0 = 0.1 * ((1.0 + q0 * q0) + (zp_lambda * (b0 + eps)) / te);
        w1 = 0.6 * ((1.0 + q1 * q1) + (zp_lambda * (b1 + eps)) / te);
        w2 = 0.3 * ((1.0 + q2 * q2) + (zp_lambda * (b2 + eps)) / te);
    } else {
        const double eps = 1.0e-40;
        const double tau = fabs(b0 - b2);
        const double bmax = (b0 < ((b1 < b2) ? b2 : b1)) ? ((b1 < b2) ? b2 : b1) : b0;  // std::max(b0, std::max(b1, b2))
        const double bmn12 = (b2 < b1) ? b2 : b1;
        const double bmin = (bmn12 < b0) ? bmn12 : b0;  // std::min(b0, std::min(b1, b2))
        const double alpha = 1.0 - (bmin / (bmax + 1e-40));
        const double la = 3.3 * alpha;
        const double q0 = tau / (b0 + eps);
        const double q1 = tau / (b1 + eps);
        const double q2 = tau / (b2 + eps);
        w0 = 0.1 * ((1.0 + q0 * q0) + (la * b0) / (bmax + eps));
        w1 = 0.6 * ((1.0 + q1 * q1) + (la * b1) / (bmax + eps));
        w2 = 0.3 * ((1.0 + q2 * q2) + (la * b2) / (bmax + eps));
    }
}

// Fifth-order WENO reconstruction at j+1/2 from (f_{j-2},...,f_{j+2}).
template <int S>
__device__ __forceinline__ double weno5(double a, double b, double c, double d, double e, double zp_lambda) {
    double b0, b1, b2, w0, w1, w2, p0, p1, p2;
    weno_beta(a, b, c, d, e, b0, b1, b2);
    weno_alpha<S>(b0, b1, b2, zp_lambda, w0, w1, w2);
    const double s = (w0 + w1) + w2;
    w0 = w0 / s;
    w1 = w1 / s;
    w2 = w2 / s;
    weno_candidates(a, b, c, d, e, p0, p1, p2);
    return (w0 * p0 + w1 * p1) + w2 * p2;
}

// Primitive variables with the reference's expression forms
// (utils/equations.hpp:139-144): p = (g-1)(E - (0.5 rho)(u^2+v^2)), c = sqrt((g p)/rho).
struct Prim {
    double rho, u, v, E, p, c;
};

__device__ __forceinline__ Prim primitive(double U0, double U1, double U2, double U3, double gamma) {
    Prim q;
    q.rho = U0;
    q.u = U1 / U0;
    q.v = U2 / U0;
    q.E = U3;
    q.p = (gamma - 1.0) * (q.E - (0.5 * q.rho) * (q.u * q.u + q.v * q.v));
    q.c = sqrt((gamma * q.p) / q.rho);
    return q;
}

// Local Lax-Friedrichs split at one point along direction D (0 = x, 1 = y):
// F-+ = 0.5 (F -+ alpha U), alpha = |u_n| + c at this point.
// utils/equations.hpp:105-134 (physical flux), :365-393 (split).
template <int D>
__device__ __forceinline__ void flux_split(const double U[4], double gamma, double Fm[4], double Fp[4]) {
    const Prim q = primitive(U[0], U[1], U[2], U[3], gamma);
    double F[4];
    if constexpr (D == 0) {
        const double ru = q.rho * q.u;
        F[0] = ru;
        F[1] = ru * q.u + q.p;
        F[2] = ru * q.v;
        F[3] = (q.E + q.p) * q.u;
    } else {
        const double rv = q.rho * q.v;
        F[0] = rv;
        F[1] = (q.rho * q.u) * q.v;
        F[2] = rv * q.v + q.p;
        F[3] = (q.E + q.p) * q.v;
    }
    const double ws = (D == 0 ? fabs(q.u) : fabs(q.v)) + q.c;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        Fm[r] = 0.5 * (F[r] - ws * U[r]);
        Fp[r] = 0.5 * (F[r] + ws * U[r]);
    }
}

// Left (rows = left eigenvectors) and right eigenmatrices of dF/dU along D at
// state U, with the reference's normalisation (utils/equations.hpp:174-299).
// Row-major 4x4: L[i*4+j].
template <int D>
__device__ __forceinline__ void eigen_matrices(const double U[4], double gamma, double L[16], double R[16]) {
    const Prim q = primitive(U[0], U[1], U[2], U[3], gamma);
    const double B1 = (gamma - 1.0) / (q.c * q.c);
    const double uu = q.u * q.u + q.v * q.v;
    const double B2 = (B1 * uu) / 2.0;
    const double ic = 1.0 / q.c;
    const double H = (q.E + q.p) / q.rho;
    const double B1u = B1 * q.u, B1v = B1 * q.v;
    if constexpr (D == 0) {
        const double uc = q.u / q.c;
        L[0] = (B2 + uc) / 2.0;  L[1] = -(B1u + ic) / 2.0;  L[2] = -B1v / 2.0;  L[3] = B1 / 2.0;
        L[4] = q.v;              L[5] = 0.0;                L[6] = -1.0;        L[7] = 0.0;
        L[8] = 1.0 - B2;         L[9] = B1u;                L[10] = B1v;        L[11] = -B1;
        L[12] = (B2 - uc) / 2.0; L[13] = -(B1u - ic) / 2.0; L[14] = -B1v / 2.0; L[15] = B1 / 2.0;
        R[0] = 1.0;              R[1] = 0.0;   R[2] = 1.0;        R[3] = 1.0;
        R[4] = q.u - q.c;        R[5] = 0.0;   R[6] = q.u;        R[7] = q.u + q.c;
        R[8] = q.v;              R[9] = -1.0;  R[10] = q.v;       R[11] = q.v;
        R[12] = H - q.c * q.u;   R[13] = -q.v; R[14] = uu / 2.0;  R[15] = H + q.c * q.u;
    } else {
        const double vc = q.v / q.c;
        L[0] = (B2 + vc) / 2.0;  L[1] = -B1u / 2.0; L[2] = -(B1v + ic) / 2.0;  L[3] = B1 / 2.0;
        L[4] = -q.u;             L[5] = 1.0;        L[6] = 0.0;                L[7] = 0.0;
        L[8] = 1.0 - B2;         L[9] = B1u;        L[10] = B1v;               L[11] = -B1;
        L[12] = (B2 - vc) / 2.0; L[13] = -B1u / 2.0; L[14] = -(B1v - ic) / 2.0; L[15] = B1 / 2.0;
        R[0] = 1.0;              R[1] = 0.0;   R[2] = 1.0;        R[3] = 1.0;
        R[4] = q.u;              R[5] = 1.0;   R[6] = q.u;        R[7] = q.u;
        R[8] = q.v - q.c;        R[9] = 0.0;   R[10] = q.v;       R[11] = q.v + q.c;
        R[12] = H - q.c * q.v;   R[13] = q.u;  R[14] = uu / 2.0;  R[15] = H + q.c * q.v;
    }
}

// Projection of one split-flux vector onto the characteristic fields:
// v_i = ((L_i0 f0 + L_i1 f1) + L_i2 f2) + L_i3 f3 (src/WENOFD.cpp:552-593),
// with the exact-zero / unit coefficients of row 1 folded.
template <int D>
__device__ __forceinline__ void project_left(const double L[16], const double f[4], double v[4]) {
    v[0] = ((L[0] * f[0] + L[1] * f[1]) + L[2] * f[2]) + L[3] * f[3];
    if constexpr (D == 0)
        v[1] = L[4] * f[0] - f[2];  // row (v, 0, -1, 0)
    else
        v[1] = L[4] * f[0] + f[1];  // row (-u, 1, 0, 0)
    v[2] = ((L[8] * f[0] + L[9] * f[1]) + L[10] * f[2]) + L[11] * f[3];
    v[3] = ((L[12] * f[0] + L[13] * f[1]) + L[14] * f[2]) + L[15] * f[3];
}

// Back-projection F = R vh (src/WENOFD.cpp:625-643), unit/zero coefficients folded.
template <int D>
__device__ __forceinline__ void project_right(const double R[16], const double vh[4], double F[4]) {
    F[0] = (vh[0] + vh[2]) + vh[3];  // row (1, 0, 1, 1)
    if constexpr (D == 0) {
        F[1] = (R[4] * vh[0] + R[6] * vh[2]) + R[7] * vh[3];        // (u-c, 0, u, u+c)
        F[2] = ((R[8] * vh[0] - vh[1]) + R[10] * vh[2]) + R[11] * vh[3];  // (v, -1, v, v)
    } else {
        F[1] = ((R[4] * vh[0] + vh[1]) + R[6] * vh[2]) + R[7] * vh[3];    // (u, 1, u, u)
        F[2] = (R[8] * vh[0] + R[10] * vh[2]) + R[11] * vh[3];       // (v-c, 0, v, v+c)
    }
    F[3] = ((R[12] * vh[0] + R[13] * vh[1]) + R[14] * vh[2]) + R[15] * vh[3];
}

}  // namespace wenofd
