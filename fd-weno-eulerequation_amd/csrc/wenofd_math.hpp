// wenofd_math.hpp — FP64 device arithmetic of the WENO-FD hot path.
//
// Every expression keeps the floating-point operation order of the reference
// solver (infinity-07/FD-WENO-EulerEquation) so that, compiled without FMA
// contraction (-ffp-contract=off) and with IEEE division and square root, the
// results equal a no-contraction build of the reference bit for bit. Terms
// whose coefficient is an exact 0 or +-1 in the reference's 4-term sums are
// folded (x + 0*y == x and 1*y == y in IEEE arithmetic up to the sign of a
// zero result), which removes ~35% of the projection work without changing a
// single nonzero result; a difference a - k b with k = 2 or 4 is one FMA (k b
// is exact, sub_pow2).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace wenofd {

// Every function is a template on the scalar type T: the kernels instantiate
// T = double; a host tool instantiates an operation-counting type to derive
// the algorithmic FLOPs per face, per point and per cell from this same code.
#define WENOFD_HD __host__ __device__ __forceinline__

enum Scheme { kJS = 0, kZ = 1, kZP = 2, kZPI = 3 };

// include/Utility.hpp:5-11
constexpr double kC1_3 = 1.0 / 3.0;
constexpr double kC1_6 = 1.0 / 6.0;
constexpr double kC5_6 = 5.0 / 6.0;
constexpr double kC7_6 = 7.0 / 6.0;
constexpr double kC11_6 = 11.0 / 6.0;
constexpr double kC13_12 = 13.0 / 12.0;

// a - k b for k = 2 or 4: k b is exact (subnormal b included) unless it
// overflows, so one FMA rounds the difference exactly as the subtraction of
// the product does; device code takes the FMA (one instruction instead of
// two), with bit-identical results for |b| < 2^1021.
template <class T>
WENOFD_HD T sub_pow2(T a, double k, T b) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (std::is_same<T, double>::value) return __builtin_fma(-k, b, a);
#endif
    return a - k * b;
}

// Smoothness indicators beta_0..2 of the three 3-point sub-stencils of
// (a,b,c,d,e) = (f_{j-2},...,f_{j+2}); include/Utility.hpp:427-437.
template <class T>
WENOFD_HD void weno_beta(T a, T b, T c, T d, T e, T& b0, T& b1, T& b2) {
    T t1 = sub_pow2(a, 2.0, b) + c;
    T t2 = sub_pow2(a, 4.0, b) + 3.0 * c;
    b0 = (kC13_12 * t1) * t1 + (0.25 * t2) * t2;
    t1 = sub_pow2(b, 2.0, c) + d;
    t2 = b - d;
    b1 = (kC13_12 * t1) * t1 + (0.25 * t2) * t2;
    t1 = sub_pow2(c, 2.0, d) + e;
    t2 = sub_pow2(3.0 * c, 4.0, d) + e;
    b2 = (kC13_12 * t1) * t1 + (0.25 * t2) * t2;
}

// Third-order candidate values at j+1/2; include/Utility.hpp:451-453.
template <class T>
WENOFD_HD void weno_candidates(T a, T b, T c, T d, T e, T& p0, T& p1, T& p2) {
    p0 = (kC1_3 * a - kC7_6 * b) + kC11_6 * c;
    p1 = (-kC1_6 * b + kC5_6 * c) + kC1_3 * d;
    p2 = (kC1_3 * c + kC5_6 * d) - kC1_6 * e;
}

// Unnormalised nonlinear weights of one scheme (include/Utility.hpp:439-442,
// 477-491, 530-542, 584-599).
template <int S, class T>
WENOFD_HD void weno_alpha(T b0, T b1, T b2, double zp_lambda, T& w0, T& w1, T& w2) {
    if constexpr (S == kJS) {
        const double eps = 1.0e-6;
        w0 = 0.1 / ((eps + b0) * (eps + b0));
        w1 = 0.6 / ((eps + b1) * (eps + b1));
        w2 = 0.3 / ((eps + b2) * (eps + b2));
    } else if constexpr (S == kZ) {
        const double eps = 1.0e-20;
        const T i0 = 1.0 / (b0 + eps);
        const T i1 = 1.0 / (b1 + eps);
        const T i2 = 1.0 / (b2 + eps);
        const T tau = fabs(b0 - b2);
        w0 = 0.1 * (1.0 + (((tau * i0) * tau) * i0));
        w1 = 0.6 * (1.0 + (((tau * i1) * tau) * i1));
        w2 = 0.3 * (1.0 + (((tau * i2) * tau) * i2));
    } else if constexpr (S == kZP) {
        const double eps = 1.0e-40;
        const T tau = fabs(b0 - b2);
        const T te = tau + eps;
        const T q0 = te / (b0 + eps);
        const T q1 = te / (b1 + eps);
        const T q2 = te / (b2 + eps);
        w0 = 0.1 * ((1.0 + q0 * q0) + (zp_lambda * (b0 + eps)) / te);
        w1 = 0.6 * ((1.0 + q1 * q1) + (zp_lambda * (b1 + eps)) / te);
        w2 = 0.3 * ((1.0 + q2 * q2) + (zp_lambda * (b2 + eps)) / te);
    } else {
        const double eps = 1.0e-40;
        const T tau = fabs(b0 - b2);
        const T bmx12 = (b1 < b2) ? b2 : b1;
        const T bmax = (b0 < bmx12) ? bmx12 : b0;  // std::max(b0, std::max(b1, b2))
        const T bmn12 = (b2 < b1) ? b2 : b1;
        const T bmin = (bmn12 < b0) ? bmn12 : b0;  // std::min(b0, std::min(b1, b2))
        const T alpha = 1.0 - (bmin / (bmax + 1e-40));
        const T la = 3.3 * alpha;
        const T q0 = tau / (b0 + eps);
        const T q1 = tau / (b1 + eps);
        const T q2 = tau / (b2 + eps);
        w0 = 0.1 * ((1.0 + q0 * q0) + (la * b0) / (bmax + eps));
        w1 = 0.6 * ((1.0 + q1 * q1) + (la * b1) / (bmax + eps));
        w2 = 0.3 * ((1.0 + q2 * q2) + (la * b2) / (bmax + eps));
    }
}

// Fifth-order WENO reconstruction at j+1/2 from (f_{j-2},...,f_{j+2}).
template <int S, class T>
WENOFD_HD T weno5(T a, T b, T c, T d, T e, double zp_lambda) {
    T b0, b1, b2, w0, w1, w2, p0, p1, p2;
    weno_beta(a, b, c, d, e, b0, b1, b2);
    weno_alpha<S>(b0, b1, b2, zp_lambda, w0, w1, w2);
    const T s = (w0 + w1) + w2;
    w0 = w0 / s;
    w1 = w1 / s;
    w2 = w2 / s;
    weno_candidates(a, b, c, d, e, p0, p1, p2);
    return (w0 * p0 + w1 * p1) + w2 * p2;
}

// ---------------------------------------------------------------------------
// Division without the scaling steps.
//
// gfx950 divides in FP64 with: d' = div_scale(d), r = rcp(d'), two Newton
// refinements r <- fma(r, fma(-d', r, 1), r), n' = div_scale(n), q = n' r,
// q <- div_fmas(fma(-d', q, n'), r, q), div_fixup(q, d, n). When both operands
// lie in [2^-300, 2^300] neither div_scale rescales (their exponents differ by
// less than 768, no operand or quotient is denormal, the numerator is not tiny),
// div_fmas adds no scale and div_fixup returns q: the sequence reduces to
// rcp + 4 FMAs that depend on d only, then a mul and two FMAs per numerator,
// with bit-identical results. The refined reciprocal is therefore shared by
// divisions with a common denominator. Callers check the range and fall back
// to the full division outside it.
// ---------------------------------------------------------------------------
constexpr double kDivSafeMax = 0x1p300;

__device__ __forceinline__ double rcp_refined(double d) {
    const double r0 = __builtin_amdgcn_rcp(d);
    const double r1 = __builtin_fma(r0, __builtin_fma(-d, r0, 1.0), r0);
    return __builtin_fma(r1, __builtin_fma(-d, r1, 1.0), r1);
}

__device__ __forceinline__ double div_refined(double n, double d, double r) {
    const double q = n * r;
    return __builtin_fma(__builtin_fma(-d, q, n), r, q);
}

// n / d given r = rcp_refined(d): the correctly rounded quotient, or in the
// build that divides by the reciprocal (WENOFD_RECIP_DIV) the product n r
// (one rounding more, as every other division there).
__device__ __forceinline__ double div_shared(double n, double d, double r) {
#if WENOFD_RECIP_DIV
    (void)d;
    return n * r;
#else
    return div_refined(n, d, r);
#endif
}

// The build without contraction divides by the density, the sound speed,
// its square and the spacings through shared refined reciprocals as well
// (WENOFD_SHARED_DIV): div_exact is bit-identical to the IEEE division when
// the denominator d lies in [2^-300, 2^300] (checked once per denominator,
// den_ok) and the numerator is +-0 or its magnitude lies in [2^-600, 2^301)
// (its biased exponent in [423, 1323]): no div_scale then rescales, and
// div_fixup returns the refined quotient, or for a zero numerator the zero of
// the quotient's sign, which n r already is (the two FMAs would turn -0 into
// +0). ok is cleared otherwise (NaN and infinity included), and the caller
// redoes the lane's divisions in IEEE arithmetic.
#if !WENOFD_RECIP_DIV && defined(__HIP_DEVICE_COMPILE__)
#define WENOFD_SHARED_DIV 1
#else
#define WENOFD_SHARED_DIV 0
#endif

__device__ __forceinline__ bool den_ok(double d) { return d >= 0x1p-300 && d <= kDivSafeMax; }

__device__ __forceinline__ double div_exact(double n, double d, double r, bool& ok) {
    const double q0 = n * r;
    const double q = __builtin_fma(__builtin_fma(-d, q0, n), r, q0);
    const unsigned e = ((unsigned)__double2hiint(n) >> 20) & 0x7FFu;
    ok = ok && (e - 423u <= 900u || n == 0.0);
    return n == 0.0 ? q0 : q;
}

// weno5 with the weight normalisation (three divisions by the common sum s)
// and WENO-Z's reciprocals 1/(beta + 1e-20) done by div_refined. ok is cleared
// when an operand leaves the exact range: s >= 1 and 0.1 <= w_k <= s always
// hold, so s <= 2^300 covers the normalisation; beta + 1e-20 >= 1e-20 always,
// so an upper bound covers the reciprocals. NaN fails both checks.
//
// The build that divides by the reciprocal (WENOFD_RECIP_DIV) forms the JS, Z
// and ZP weights over a common denominator instead: with B_k = (beta_k + eps),
// D_k = B_k^2 and te = tau + eps, JS's d_k / D_k, Z's d_k (1 + tau^2 / D_k) and
// ZP's d_k (1 + te^2 / D_k + lambda B_k / te) are proportional to d_k D_i D_j,
// d_k (D_k + tau^2) D_i D_j and d_k (D_k (te + lambda B_k) + te^3) D_i D_j
// ({i, j} the other two indices), and the normalisation cancels the common
// factor (with FOLD, the linear weights and the normalisation fold into the
// candidate polynomials, integer coefficients, so the reconstruction is one
// weighted sum times one reciprocal). The weights are sums and products of
// positive terms (no cancellation; a few roundings each), and the only
// reciprocal left is the one of their sum s. Every term and partial product is >= 1e-282 (normal), so
// s <= 2^1000 (which NaN and an overflowed term fail) keeps 1/s normal. ZPI's
// weights are first scaled by 1/M (M = bmax + eps, one reciprocal more) and
// stay within [0, 6.3] per factor; where a beta ratio underflows, s >= 2^-900
// is required, which leaves every term that weighs more than 2^-60 of s normal.
// In the build that divides by the reciprocal the reciprocal of the weights'
// sum takes one Newton step (<= 11 ulp from 1/x; two steps are exact), and
// FOLD folds the linear weights and the normalisation into the candidates, 3
// instructions fewer per reconstruction (the sweeps' faces and the hybrid's
// corrections use it; the probe keeps the unfolded form).
template <int S, bool FOLD = false>
__device__ __forceinline__ double weno5_shared_rcp(double a, double b, double c, double d, double e,
                                                   double zp_lambda, bool& ok) {
    double b0, b1, b2, w0, w1, w2, p0, p1, p2;
#if WENOFD_RECIP_DIV
    constexpr bool kShared = FOLD;
#else
    constexpr bool kShared = false;  // the reference's expression forms
#endif
    // first differences of the stencil (the shared form below)
    const double d0 = a - b, d1 = b - c, d2 = c - d, d3 = d - e;
    // kShared: the smoothness indicators times 4, 13/3 s_k^2 + u_k^2 (three
    // instructions each), from shared differences: the second differences s_k
    // and the first-order terms u_k (a - 4b + 3c, b - d, 3c - 4d + e) follow
    // from d0 .. d3. Every scheme's weights are invariant when the betas and
    // eps scale together (eps is scaled by the same exact power of 2).
    constexpr double kBetaScale = kShared ? 4.0 : 1.0;
    (void)kBetaScale;
    if constexpr (kShared) {
        const double s0 = d0 - d1, s1 = d1 - d2, s2 = d2 - d3;
        const double u0 = __builtin_fma(-2.0, d1, s0), u1 = d1 + d2, u2 = __builtin_fma(2.0, d2, s2);
        constexpr double k13_3 = 13.0 / 3.0;
        if constexpr (S == kZPI) {
            b0 = __builtin_fma(k13_3 * s0, s0, u0 * u0);
            b1 = __builtin_fma(k13_3 * s1, s1, u1 * u1);
            b2 = __builtin_fma(k13_3 * s2, s2, u2 * u2);
        } else {
            // JS, Z and ZP use the betas only as B_k = beta_k + eps and through
            // B0 - B2: eps is added inside the indicators' FMAs
            constexpr double eps4 = (S == kJS ? 1.0e-6 : S == kZ ? 1.0e-20 : 1.0e-40) * kBetaScale;
            b0 = __builtin_fma(k13_3 * s0, s0, __builtin_fma(u0, u0, eps4));
            b1 = __builtin_fma(k13_3 * s1, s1, __builtin_fma(u1, u1, eps4));
            b2 = __builtin_fma(k13_3 * s2, s2, __builtin_fma(u2, u2, eps4));
        }
    } else {
        (void)d0;
        (void)d1;
        (void)d2;
        (void)d3;
        weno_beta(a, b, c, d, e, b0, b1, b2);
    }
#if WENOFD_RECIP_DIV
    {
        // per scheme: the factors D_k (squared, shifted betas) and E_k
        double D0, D1, D2, E0 = 1.0, E1 = 1.0, E2 = 1.0;
        double tz2 = 0.0;  // WENO-Z: tau^2 (E_k = D_k + tau^2)
        if constexpr (S == kZPI) {
            // ZPI: d_k (1 + tau^2 / B_k^2 + 3.3 (1 - bmin / M) b_k / M) with
            // M = bmax + eps, scaled by rM^2 B_k^2 (rM = 1/M): every ratio lies
            // in [0, 1], so the scaled terms stay near 1 for any beta scale
            const double eps = 1.0e-40 * kBetaScale;
            const double tau = fabs(b0 - b2);
            const double bmx12 = (b1 < b2) ? b2 : b1;
            const double bmax = (b0 < bmx12) ? bmx12 : b0;
            const double bmn12 = (b2 < b1) ? b2 : b1;
            const double bmin = (bmn12 < b0) ? bmn12 : b0;
            const double M = bmax + eps;
            ok = ok && M <= 0x1p1000;
            const double rM = rcp_refined(M);
            const double x0 = (b0 + eps) * rM, x1 = (b1 + eps) * rM, x2 = (b2 + eps) * rM;
            D0 = x0 * x0;
            D1 = x1 * x1;
            D2 = x2 * x2;
            const double t = tau * rM;
            const double t2 = t * t;
            const double la = 3.3 * (1.0 - bmin * rM);
            E0 = D0 * (1.0 + la * (b0 * rM)) + t2;
            E1 = D1 * (1.0 + la * (b1 * rM)) + t2;
            E2 = D2 * (1.0 + la * (b2 * rM)) + t2;
        } else {
            const double eps = (S == kJS ? 1.0e-6 : S == kZ ? 1.0e-20 : 1.0e-40) * kBetaScale;
            const double B0 = kShared ? b0 : S == kJS ? eps + b0 : b0 + eps;
            const double B1 = kShared ? b1 : S == kJS ? eps + b1 : b1 + eps;
            const double B2 = kShared ? b2 : S == kJS ? eps + b2 : b2 + eps;
            D0 = B0 * B0;
            D1 = B1 * B1;
            D2 = B2 * B2;
            if constexpr (S == kZ) {
                const double tau = fabs(b0 - b2);
                const double t2 = tau * tau;
                tz2 = t2;
                E0 = D0 + t2;
                E1 = D1 + t2;
                E2 = D2 + t2;
            } else if constexpr (S == kZP) {
                // 1 + te^2 / D_k + lambda B_k / te, times te: D_k (te + lambda B_k) + te^3
                // (>= 1e-120; lambda = dx^(2/3) > 0 keeps every term positive)
                const double te = fabs(b0 - b2) + eps;
                const double te3 = (te * te) * te;
                E0 = D0 * (te + zp_lambda * B0) + te3;
                E1 = D1 * (te + zp_lambda * B1) + te3;
                E2 = D2 * (te + zp_lambda * B2) + te3;
                ok = ok && zp_lambda >= 0.0;
            }
        }
        if constexpr (FOLD) {
            // d_k scaled by 10: (1, 6, 3), folded with the normalisation into
            // the candidates (q_k = 6 p_k, integer coefficients): the
            // reconstruction is (W0 q0 + W1 6 q1 + W2 3 q2) / (6 (W0 + 6 W1 + 3 W2))
            const double D12 = D1 * D2, D02 = D0 * D2, D01 = D0 * D1;
            double W0, W1, W2;
            if constexpr (S == kZ) {
                // (D_k + tau^2) D_i D_j = D0 D1 D2 + tau^2 D_i D_j: one FMA each
                const double P = D0 * D12;
                W0 = __builtin_fma(tz2, D12, P);
                W1 = __builtin_fma(tz2, D02, P);
                W2 = __builtin_fma(tz2, D01, P);
            } else {
                W0 = E0 * D12;
                W1 = E1 * D02;
                W2 = E2 * D01;
            }
            const double s = (W0 + 6.0 * W1) + 3.0 * W2;
            ok = ok && s <= 0x1p1000 && (S != kZPI || s >= 0x1p-900);
            const double r0 = __builtin_amdgcn_rcp(s);
            const double r = __builtin_fma(r0, __builtin_fma(-s, r0, 1.0), r0);  // one Newton step: <= 11 ulp
            // q_k = 6 p_k = m_k c + r_k with (m_k) = 6 (1, 6, 3), so the
            // weighted sum of the m_k c is 6 s c: the reconstruction is
            // c + (W0 r0 + W1 r1 + W2 r2) / (6 s), the r_k / 6 from the differences
            const double t = -2.0 * d2;
            const double q0 = __builtin_fma(kC1_3, d0, -kC5_6 * d1);  // (2a - 7b + 5c) / 6
            const double q1 = t - d1;                                 // -b - c + 2d
            const double q2 = __builtin_fma(0.5, d3, t);              // -2c + 5d/2 - e/2
            return __builtin_fma((W0 * q0 + W1 * q1) + W2 * q2, r, c);
        }
        // d_k scaled by 10: (1, 6, 3)
        w0 = E0 * (D1 * D2);
        w1 = 6.0 * (E1 * (D0 * D2));
        w2 = 3.0 * (E2 * (D0 * D1));
        const double s = (w0 + w1) + w2;
        ok = ok && s <= 0x1p1000 && (S != kZPI || s >= 0x1p-900);
        const double r0 = __builtin_amdgcn_rcp(s);
        const double r = __builtin_fma(r0, __builtin_fma(-s, r0, 1.0), r0);  // one Newton step: <= 11 ulp
        w0 = w0 * r;
        w1 = w1 * r;
        w2 = w2 * r;
        weno_candidates(a, b, c, d, e, p0, p1, p2);
        return (w0 * p0 + w1 * p1) + w2 * p2;
    }
#endif
    if constexpr (S == kJS) {
        // d_k / (eps + beta_k)^2: denominators >= 1e-12
        const double eps = 1.0e-6;
        const double d0 = (eps + b0) * (eps + b0), d1 = (eps + b1) * (eps + b1), d2 = (eps + b2) * (eps + b2);
        ok = ok && d0 <= kDivSafeMax && d1 <= kDivSafeMax && d2 <= kDivSafeMax;
        w0 = div_shared(0.1, d0, rcp_refined(d0));
        w1 = div_shared(0.6, d1, rcp_refined(d1));
        w2 = div_shared(0.3, d2, rcp_refined(d2));
    } else if constexpr (S == kZ) {
        const double eps = 1.0e-20;
        const double d0 = b0 + eps, d1 = b1 + eps, d2 = b2 + eps;
        ok = ok && d0 <= kDivSafeMax && d1 <= kDivSafeMax && d2 <= kDivSafeMax;
        const double i0 = div_shared(1.0, d0, rcp_refined(d0));
        const double i1 = div_shared(1.0, d1, rcp_refined(d1));
        const double i2 = div_shared(1.0, d2, rcp_refined(d2));
        const double tau = fabs(b0 - b2);
        w0 = 0.1 * (1.0 + (((tau * i0) * tau) * i0));
        w1 = 0.6 * (1.0 + (((tau * i1) * tau) * i1));
        w2 = 0.3 * (1.0 + (((tau * i2) * tau) * i2));
    } else {
        // ZP / ZPI keep the IEEE divisions: their shared-reciprocal form costs
        // more registers than it saves instructions (3 waves per SIMD lost)
        weno_alpha<S>(b0, b1, b2, zp_lambda, w0, w1, w2);
    }
    const double s = (w0 + w1) + w2;
    ok = ok && s <= kDivSafeMax;
    const double r = rcp_refined(s);
    w0 = div_shared(w0, s, r);
    w1 = div_shared(w1, s, r);
    w2 = div_shared(w2, s, r);
    weno_candidates(a, b, c, d, e, p0, p1, p2);
    return (w0 * p0 + w1 * p1) + w2 * p2;
}

// Device code of the build that divides by the reciprocal: divisions by the
// density and the sound speed multiply by rcp_one (v_rcp_f64 and one Newton
// step) instead of the compiler's reciprocal, a full IEEE 1/x, 1/c^2 is
// (1/c)^2, and the sound speed's square root is sqrt_gs (at the faces, c and
// 1/(2c) come from sqrt_half_rsqrt_gs). A zero, infinite or NaN density or
// sound speed (non-physical states) gives NaN where the division gives inf
// or NaN.
#if WENOFD_RECIP_DIV && defined(__HIP_DEVICE_COMPILE__)
#define WENOFD_RCP_STATE 1
#else
#define WENOFD_RCP_STATE 0
#endif

#if WENOFD_RECIP_DIV
// The state's reciprocals in the build that divides by the reciprocal: one
// Newton step after v_rcp_f64 (<= 11 ulp from 1/x on 4.2 M sampled normal
// operands; v_rcp_f64 alone is 2^-24 accurate).
__device__ __forceinline__ double rcp_one(double d) {
    const double r0 = __builtin_amdgcn_rcp(d);
    return __builtin_fma(r0, __builtin_fma(-d, r0, 1.0), r0);
}

// The sound speed's square root there: Goldschmidt's iteration from v_rsq_f64
// (g = x y, h = y / 2, r = 1/2 - g h, g += g r, h += h r) and one Newton
// correction g += (x - g^2) h: 8 instructions against the compiler's 17 (its
// scaling of tiny operands and second correction), correctly rounded on 4.2 M
// sampled operands in [2^-60, 2^61]; 0 for 0, NaN for a negative, infinite
// or NaN operand (non-physical states).
__device__ __forceinline__ double sqrt_gs(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double r = __builtin_fma(-g, h, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    g = __builtin_fma(__builtin_fma(-g, g, x), h, g);
    return x == 0.0 ? x : g;
}

// sqrt(x) and 1 / (2 sqrt(x)) together: two coupled Goldschmidt iterations
// from v_rsq_f64 (no final correction), both within a few ulps; 9
// instructions against sqrt_gs's 8 and a refined reciprocal's 4.
__device__ __forceinline__ double sqrt_half_rsqrt_gs(double x, double& hr) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    double r = __builtin_fma(-g, h, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    r = __builtin_fma(-g, h, 0.5);
    hr = __builtin_fma(h, r, h);
    return __builtin_fma(g, r, g);
}
#endif

// Primitive variables with the reference's expression forms
// (utils/equations.hpp:139-144): p = (g-1)(E - (0.5 rho)(u^2+v^2)), c = sqrt((g p)/rho).
// irho = 1/rho is set only where WENOFD_RCP_STATE computes it, or (the
// refined reciprocal) where WENOFD_SHARED_DIV does.
template <class T>
struct Prim {
    T rho, u, v, E, p, c, irho;
};

template <class T>
WENOFD_HD Prim<T> primitive(T U0, T U1, T U2, T U3, double gamma) {
    Prim<T> q;
#if WENOFD_RCP_STATE
    if constexpr (std::is_same<T, double>::value) {
        const double r = rcp_one(U0);
        q.rho = U0;
        q.irho = r;
        q.u = U1 * r;
        q.v = U2 * r;
        q.E = U3;
        q.p = (gamma - 1.0) * (q.E - (0.5 * q.rho) * (q.u * q.u + q.v * q.v));
        q.c = sqrt_gs((gamma * q.p) * r);
        return q;
    }
#endif
    q.rho = U0;
    q.u = U1 / U0;
    q.v = U2 / U0;
    q.E = U3;
    q.p = (gamma - 1.0) * (q.E - (0.5 * q.rho) * (q.u * q.u + q.v * q.v));
    q.c = sqrt((gamma * q.p) / q.rho);
    return q;
}

#if WENOFD_SHARED_DIV
// primitive through the density's refined reciprocal (div_exact; irho = the
// reciprocal); ok is cleared where a division leaves the exact range, and the
// caller then takes primitive's IEEE result for that lane.
__device__ __forceinline__ Prim<double> primitive_sd(double U0, double U1, double U2, double U3, double gamma,
                                                     bool& ok) {
    Prim<double> q;
    ok = ok && den_ok(U0);
    const double r = rcp_refined(U0);
    q.rho = U0;
    q.irho = r;
    q.u = div_exact(U1, U0, r, ok);
    q.v = div_exact(U2, U0, r, ok);
    q.E = U3;
    q.p = (gamma - 1.0) * (q.E - (0.5 * q.rho) * (q.u * q.u + q.v * q.v));
    q.c = sqrt(div_exact(gamma * q.p, U0, r, ok));
    return q;
}
#endif

// Local Lax-Friedrichs split at one point along direction D (0 = x, 1 = y):
// F-+ = 0.5 (F -+ alpha U), alpha = |u_n| + c at this point.
// utils/equations.hpp:105-134 (physical flux), :365-393 (split).
template <int D, class T>
WENOFD_HD void flux_split(const T U[4], double gamma, T Fm[4], T Fp[4]) {
#if WENOFD_RCP_STATE
    if constexpr (std::is_same<T, double>::value) {
        // formed halved: p/2 from rho (u^2 + v^2) = U1 u + U2 v, the flux's
        // halves Fh from the normal momentum, w = (|u_n| + c)/2, and
        // F-+ = Fh -+ w U (one FMA each)
        const double r = rcp_one(U[0]);
        const double u = U[1] * r, v = U[2] * r;
        const double ph = (0.5 * (gamma - 1.0)) * __builtin_fma(-0.5, __builtin_fma(U[1], u, U[2] * v), U[3]);
        const double ch = sqrt_gs(((0.5 * gamma) * ph) * r);  // c / 2
        const double un = D == 0 ? u : v;
        const double mh = 0.5 * U[1 + D];
        double Fh[4];
        Fh[0] = mh;
        Fh[1] = D == 0 ? __builtin_fma(mh, u, ph) : mh * u;
        Fh[2] = D == 0 ? mh * v : __builtin_fma(mh, v, ph);
        Fh[3] = __builtin_fma(0.5, U[3], ph) * un;
        const double w = __builtin_fma(0.5, fabs(un), ch);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            Fm[k] = __builtin_fma(-w, U[k], Fh[k]);
            Fp[k] = __builtin_fma(w, U[k], Fh[k]);
        }
        return;
    }
#endif
#if WENOFD_SHARED_DIV
    Prim<T> q;
    if constexpr (std::is_same<T, double>::value) {
        bool ok = true;
        q = primitive_sd(U[0], U[1], U[2], U[3], gamma, ok);
        if (__any(!ok)) {
            if (!ok) q = primitive(U[0], U[1], U[2], U[3], gamma);
        }
    } else {
        q = primitive(U[0], U[1], U[2], U[3], gamma);
    }
#else
    const Prim<T> q = primitive(U[0], U[1], U[2], U[3], gamma);
#endif
    T F[4];
    if constexpr (D == 0) {
        const T ru = q.rho * q.u;
        F[0] = ru;
        F[1] = ru * q.u + q.p;
        F[2] = ru * q.v;
        F[3] = (q.E + q.p) * q.u;
    } else {
        const T rv = q.rho * q.v;
        F[0] = rv;
        F[1] = (q.rho * q.u) * q.v;
        F[2] = rv * q.v + q.p;
        F[3] = (q.E + q.p) * q.v;
    }
    const T ws = (D == 0 ? fabs(q.u) : fabs(q.v)) + q.c;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        Fm[r] = 0.5 * (F[r] - ws * U[r]);
        Fp[r] = 0.5 * (F[r] + ws * U[r]);
    }
}

// The quantities both eigenmatrices of dF/dU along D are built from, at state
// U, with the reference's expression forms (utils/equations.hpp:174-299):
// B1 = (g-1)/c^2, B2 = B1 (u^2+v^2)/2, H = (E+p)/rho, 1/c and u_n/c.
template <class T>
struct EigenParts {
    T u, v, c, H, B1, B2, B1u, B1v, ic, nc, uu;  // nc = u/c (x) or v/c (y)
};

template <int D, class T>
WENOFD_HD EigenParts<T> eigen_parts(const T U[4], double gamma) {
    const Prim<T> q = primitive(U[0], U[1], U[2], U[3], gamma);
    EigenParts<T> e;
    e.u = q.u;
    e.v = q.v;
    e.c = q.c;
#if WENOFD_RCP_STATE
    e.ic = rcp_one(q.c);
    e.B1 = (gamma - 1.0) * (e.ic * e.ic);
    e.H = (q.E + q.p) * q.irho;
    e.nc = (D == 0 ? q.u : q.v) * e.ic;
#else
    e.B1 = (gamma - 1.0) / (q.c * q.c);
    e.ic = 1.0 / q.c;
    e.H = (q.E + q.p) / q.rho;
    e.nc = (D == 0 ? q.u : q.v) / q.c;
#endif
    e.uu = q.u * q.u + q.v * q.v;
    e.B2 = (e.B1 * e.uu) / 2.0;
    e.B1u = e.B1 * q.u;
    e.B1v = e.B1 * q.v;
    return e;
}

// Left (rows = left eigenvectors) and right eigenmatrices of dF/dU along D at
// state U, with the reference's normalisation (utils/equations.hpp:174-299).
// Row-major 4x4: L[i*4+j].
template <int D, class T>
WENOFD_HD void eigen_matrices_from(const EigenParts<T>& e, T L[16], T R[16]) {
    const T B1 = e.B1, B2 = e.B2, B1u = e.B1u, B1v = e.B1v, ic = e.ic, H = e.H, uu = e.uu;
    if constexpr (D == 0) {
        const T uc = e.nc;
        L[0] = (B2 + uc) / 2.0;  L[1] = -(B1u + ic) / 2.0;  L[2] = -B1v / 2.0;  L[3] = B1 / 2.0;
        L[4] = e.v;              L[5] = 0.0;                L[6] = -1.0;        L[7] = 0.0;
        L[8] = 1.0 - B2;         L[9] = B1u;                L[10] = B1v;        L[11] = -B1;
        L[12] = (B2 - uc) / 2.0; L[13] = -(B1u - ic) / 2.0; L[14] = -B1v / 2.0; L[15] = B1 / 2.0;
        R[0] = 1.0;              R[1] = 0.0;   R[2] = 1.0;        R[3] = 1.0;
        R[4] = e.u - e.c;        R[5] = 0.0;   R[6] = e.u;        R[7] = e.u + e.c;
        R[8] = e.v;              R[9] = -1.0;  R[10] = e.v;       R[11] = e.v;
        R[12] = H - e.c * e.u;   R[13] = -e.v; R[14] = uu / 2.0;  R[15] = H + e.c * e.u;
    } else {
        const T vc = e.nc;
        L[0] = (B2 + vc) / 2.0;  L[1] = -B1u / 2.0; L[2] = -(B1v + ic) / 2.0;  L[3] = B1 / 2.0;
        L[4] = -e.u;             L[5] = 1.0;        L[6] = 0.0;                L[7] = 0.0;
        L[8] = 1.0 - B2;         L[9] = B1u;        L[10] = B1v;               L[11] = -B1;
        L[12] = (B2 - vc) / 2.0; L[13] = -B1u / 2.0; L[14] = -(B1v - ic) / 2.0; L[15] = B1 / 2.0;
        R[0] = 1.0;              R[1] = 0.0;   R[2] = 1.0;        R[3] = 1.0;
        R[4] = e.u;              R[5] = 1.0;   R[6] = e.u;        R[7] = e.u;
        R[8] = e.v - e.c;        R[9] = 0.0;   R[10] = e.v;       R[11] = e.v + e.c;
        R[12] = H - e.c * e.v;   R[13] = e.u;  R[14] = uu / 2.0;  R[15] = H + e.c * e.v;
    }
}

template <int D, class T>
WENOFD_HD void eigen_matrices(const T U[4], double gamma, T L[16], T R[16]) {
    eigen_matrices_from<D>(eigen_parts<D>(U, gamma), L, R);
}

// Row k of L and column k of R of eigen_matrices<D> from its parts (the same
// entries, the same expressions; k in 0..3 at run time).
template <int D, class T>
WENOFD_HD void eigen_row_col(const EigenParts<T>& e, int k, T Lk[4], T Rk[4]) {
    const T B1 = e.B1, B2 = e.B2, B1u = e.B1u, B1v = e.B1v, ic = e.ic, H = e.H;
    if (k == 1) {
        if constexpr (D == 0) {
            Lk[0] = e.v; Lk[1] = 0.0; Lk[2] = -1.0; Lk[3] = 0.0;
            Rk[0] = 0.0; Rk[1] = 0.0; Rk[2] = -1.0; Rk[3] = -e.v;
        } else {
            Lk[0] = -e.u; Lk[1] = 1.0; Lk[2] = 0.0; Lk[3] = 0.0;
            Rk[0] = 0.0;  Rk[1] = 1.0; Rk[2] = 0.0; Rk[3] = e.u;
        }
    } else if (k == 2) {
        Lk[0] = 1.0 - B2; Lk[1] = B1u; Lk[2] = B1v; Lk[3] = -B1;
        Rk[0] = 1.0;      Rk[1] = e.u; Rk[2] = e.v; Rk[3] = e.uu / 2.0;
    } else {  // the acoustic fields k = 0 (s = -1) and k = 3 (s = +1)
        const bool minus = k == 0;
        const T cn = e.c * (D == 0 ? e.u : e.v);
        Lk[0] = minus ? (B2 + e.nc) / 2.0 : (B2 - e.nc) / 2.0;
        if constexpr (D == 0) {
            Lk[1] = minus ? -(B1u + ic) / 2.0 : -(B1u - ic) / 2.0;
            Lk[2] = -B1v / 2.0;
            Rk[1] = minus ? e.u - e.c : e.u + e.c;
            Rk[2] = e.v;
        } else {
            Lk[1] = -B1u / 2.0;
            Lk[2] = minus ? -(B1v + ic) / 2.0 : -(B1v - ic) / 2.0;
            Rk[1] = e.u;
            Rk[2] = minus ? e.v - e.c : e.v + e.c;
        }
        Lk[3] = B1 / 2.0;
        Rk[0] = 1.0;
        Rk[3] = minus ? H - cn : H + cn;
    }
}

template <int D, class T>
WENOFD_HD void eigen_row_col(const T U[4], double gamma, int k, T Lk[4], T Rk[4]) {
    eigen_row_col<D>(eigen_parts<D>(U, gamma), k, Lk, Rk);
}

// Projection of one split-flux vector onto the characteristic fields:
// v_i = ((L_i0 f0 + L_i1 f1) + L_i2 f2) + L_i3 f3 (src/WENOFD.cpp:552-593),
// with the exact-zero / unit coefficients of row 1 folded.
template <int D, class T>
WENOFD_HD void project_left(const T L[16], const T f[4], T v[4]) {
    v[0] = ((L[0] * f[0] + L[1] * f[1]) + L[2] * f[2]) + L[3] * f[3];
    if constexpr (D == 0)
        v[1] = L[4] * f[0] - f[2];  // row (v, 0, -1, 0)
    else
        v[1] = L[4] * f[0] + f[1];  // row (-u, 1, 0, 0)
    v[2] = ((L[8] * f[0] + L[9] * f[1]) + L[10] * f[2]) + L[11] * f[3];
    v[3] = ((L[12] * f[0] + L[13] * f[1]) + L[14] * f[2]) + L[15] * f[3];
}

// Back-projection F = R vh (src/WENOFD.cpp:625-643), unit/zero coefficients folded.
template <int D, class T>
WENOFD_HD void project_right(const T R[16], const T vh[4], T F[4]) {
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

// Characteristic-wise LLF-split WENO5 flux at one face (src/WENOFD.cpp:526-644):
// arithmetic-mean state -> L, R -> project the 5-point F+ (points f-3..f+1) and
// F- (points f-2..f+2) stencils -> 2 x 4 WENO5 -> R (v+ + v-).
// sp(s, r) / sm(s, r) return component r of F+ / F- at stencil slot s
// (slot 0 = point f-3, slot 5 = point f+2).
template <int D, int S, class T, class GetP, class GetM>
WENOFD_HD void face_flux_ave(const T ave[4], GetP sp, GetM sm, double gamma, double zp_lambda, T Fh[4]) {
    T L[16], R[16];
    eigen_matrices<D>(ave, gamma, L, R);
    T vp[5][4], vm[5][4];
#pragma unroll
    for (int s = 0; s < 5; ++s) {
        T fp[4], fm[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            fp[r] = sp(s, r);
            fm[r] = sm(s + 1, r);
        }
        project_left<D>(L, fp, vp[s]);
        project_left<D>(L, fm, vm[s]);
    }
    T vh[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const T hp = weno5<S>(vp[0][k], vp[1][k], vp[2][k], vp[3][k], vp[4][k], zp_lambda);
        const T hm = weno5<S>(vm[4][k], vm[3][k], vm[2][k], vm[1][k], vm[0][k], zp_lambda);  // mirrored (:604-605)
        vh[k] = hp + hm;
    }
    project_right<D>(R, vh, Fh);
}

template <int D, int S, class T, class GetP, class GetM>
WENOFD_HD void face_flux(const T Ul[4], const T Ur[4], GetP sp, GetM sm, double gamma, double zp_lambda, T Fh[4]) {
    T ave[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) ave[r] = 0.5 * (Ul[r] + Ur[r]);
    face_flux_ave<D, S>(ave, sp, sm, gamma, zp_lambda, Fh);
}

// face_flux as the kernels run it: the 8 reconstructions use the shared
// refined reciprocals (weno5_shared_rcp). If an operand of one of them left the
// exact range in any lane of the wave, the wave recomputes the face with full
// IEEE divisions and those lanes take that result; the check is wave-uniform
// and placed after the fast path; the recomputation starts from the mean
// state, so only that (not the face's two points) stays live across the fast
// path.
template <int D, int S, bool FRESH, class GetP, class GetM>
__device__ __forceinline__ void face_flux_dev(const double Ul[4], const double Ur[4], GetP sp, GetM sm, double gamma,
                                              double zp_lambda, double Fh[4]) {
    double L[16], R[16];
    // (the mean state's divisions stay IEEE in the build without contraction:
    // through refined reciprocals they would cost the x sweep its third wave
    // per SIMD)
#if WENOFD_RECIP_DIV
    // The eigen parts (velocities, sound speed, enthalpy) are ratios of the
    // state's components, so the sum of the two points serves for their mean
    // (the scaling by 2 is exact); the mean state itself is formed only for
    // the IEEE fallback.
    double sum[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) sum[r] = Ul[r] + Ur[r];
    // from the sum S of the two points: u = S1 / S0, v = S2 / S0 and, with
    // X = S3 - S0 (u^2 + v^2) / 2 = 2 (E - rho (u^2 + v^2) / 2),
    // H = (S3 + (gamma - 1) X) / S0 and c^2 = gamma (gamma - 1) X / S0
    const double rs = rcp_one(sum[0]);
    const double qu = sum[1] * rs, qv = sum[2] * rs;
    const double uuh = 0.5 * __builtin_fma(qu, qu, qv * qv);
    const double X = __builtin_fma(-sum[0], uuh, sum[3]);
    double ich;  // 1 / (2 c)
    const double qc = sqrt_half_rsqrt_gs(((gamma * (gamma - 1.0)) * X) * rs, ich);
    struct {
        double u, v, c, H;
    } e{qu, qv, qc, __builtin_fma(gamma - 1.0, X, sum[3]) * rs};
    // Rows 0, 2 and 3 of L share their B terms: with g = B1u f1 + B1v f2 - B1 f3,
    // h = B2 f0 - g and k = (u_n/c) f0 - (1/c) f_n, they are (h + k)/2,
    // f0 - h and (h - k)/2 (the reference's rows, reassociated: 10 instructions
    // per vector instead of 13); B1 = (gamma - 1)/c^2, B2 = B1 (u^2 + v^2)/2,
    // formed halved.
    const double B1h = (2.0 * (gamma - 1.0)) * (ich * ich);
    const double B2h = B1h * uuh, B1uh = B1h * qu, B1vh = B1h * qv;
    const double nch = (D == 0 ? qu : qv) * ich;
    (void)L;
    (void)R;
    auto project = [&](const double f[4], double v[4]) {
        const double gh = __builtin_fma(-B1h, f[3], __builtin_fma(B1vh, f[2], B1uh * f[1]));
        const double hh = __builtin_fma(B2h, f[0], -gh);
        const double kh = __builtin_fma(nch, f[0], -(ich * f[D == 0 ? 1 : 2]));
        v[0] = hh + kh;
        v[1] = D == 0 ? __builtin_fma(e.v, f[0], -f[2]) : __builtin_fma(-e.u, f[0], f[1]);
        v[2] = __builtin_fma(-2.0, hh, f[0]);
        v[3] = hh - kh;
    };
#else
    double ave[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) ave[r] = 0.5 * (Ul[r] + Ur[r]);
    eigen_matrices<D>(ave, gamma, L, R);
    auto project = [&](const double f[4], double v[4]) { project_left<D>(L, f, v); };
#endif
    bool ok = true;  // cleared where a reconstruction's division leaves its exact range
    double vp[5][4], vm[5][4];
#pragma unroll
    for (int s = 0; s < 5; ++s) {
        double fp[4], fm[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            fp[r] = sp(s, r, 0);
            fm[r] = sm(s + 1, r, 0);
        }
        project(fp, vp[s]);
        project(fm, vm[s]);
    }
    double vh[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double hp =
            weno5_shared_rcp<S, true>(vp[0][k], vp[1][k], vp[2][k], vp[3][k], vp[4][k], zp_lambda, ok);
        const double hm =
            weno5_shared_rcp<S, true>(vm[4][k], vm[3][k], vm[2][k], vm[1][k], vm[0][k], zp_lambda, ok);
        vh[k] = hp + hm;
    }
#if WENOFD_RECIP_DIV
    {
        // R vh with R's shared terms (reassociated): F0 = (vh0 + vh3) + vh2,
        // the normal momentum u_n F0 + c (vh3 - vh0), the tangential one
        // u_t F0 -+ vh1, the energy H (vh0 + vh3) + c u_n (vh3 - vh0) + the
        // shear and entropy columns
        const double s03 = vh[0] + vh[3], dd = vh[3] - vh[0];
        const double un = D == 0 ? e.u : e.v, ut = D == 0 ? e.v : e.u;
        const double F0 = s03 + vh[2];
        const double Fn = __builtin_fma(un, F0, e.c * dd);
        const double Ft = __builtin_fma(ut, F0, D == 0 ? -vh[1] : vh[1]);
        Fh[0] = F0;
        Fh[1] = D == 0 ? Fn : Ft;
        Fh[2] = D == 0 ? Ft : Fn;
        Fh[3] = __builtin_fma(e.H, s03, __builtin_fma(e.c * un, dd, __builtin_fma(D == 0 ? -e.v : e.u, vh[1], uuh * vh[2])));
    }
#else
    project_right<D>(R, vh, Fh);
#endif
    if (__any(!ok)) {
        // with FRESH, recompute from the mean state and fresh LDS reads: the
        // empty asm hides that they equal the fast path's, so none of its
        // values (loaded points, projections, betas) is kept live across it for
        // reuse here (fewer registers in some sweeps, more in others)
        double av[4], Fe[4];
        int k0 = 0;
#pragma unroll
#if WENOFD_RECIP_DIV
        for (int r = 0; r < 4; ++r) av[r] = 0.5 * sum[r];
#else
        for (int r = 0; r < 4; ++r) av[r] = ave[r];
#endif
        if constexpr (FRESH) {
            asm volatile("" : "+v"(k0));
#pragma unroll
            for (int r = 0; r < 4; ++r) asm volatile("" : "+v"(av[r]));
        }
        face_flux_ave<D, S>(
            av, [&](int s, int r) { return sp(s, r, k0); }, [&](int s, int r) { return sm(s, r, k0); }, gamma,
            zp_lambda, Fe);
        if (!ok) {
#pragma unroll
            for (int r = 0; r < 4; ++r) Fh[r] = Fe[r];
        }
    }
}

// Local CFL step of one point (src/WENOFD.cpp:419-423, utils/equations.hpp:137-147):
// cfl / ((|u| + c)/dx + (|v| + c)/dy); |u*1 + v*0| == |u| for every finite v.
template <class T>
WENOFD_HD T local_dt(const T U[4], double gamma, double cfl, double dx, double dy) {
    const Prim<T> q = primitive(U[0], U[1], U[2], U[3], gamma);
    const T ex = fabs(q.u) + q.c;
    const T ey = fabs(q.v) + q.c;
    return cfl / (ex / dx + ey / dy);
}

// rhs of one conserved variable at one cell: -(Fx[i+1]-Fx[i])/dx - (Fy[j+1]-Fy[j])/dy
// (src/WENOFD.cpp:646-648, 794-796; the x term is added to a zeroed rhs first).
template <class T>
WENOFD_HD T cell_rhs(T fxl, T fxr, T fyb, T fyt, double dx, double dy) {
    const T tx = -(fxr - fxl) / dx;
    const T ty = -(fyt - fyb) / dy;
    return tx + ty;
}

// The device forms of local_dt and of cell_rhs for the four conserved
// variables, given the spacings and their reciprocals rdx = 1/dx, rdy = 1/dy
// from the host. The build that divides by the reciprocal multiplies by
// them; the build without contraction divides through refined reciprocals
// (div_exact: the same quotients), and takes the IEEE divisions in a lane
// whose operands leave that range.
__device__ __forceinline__ double local_dt_dev(const double U[4], double gamma, double cfl, double dx, double dy,
                                               double rdx, double rdy) {
#if WENOFD_SHARED_DIV
    bool ok = den_ok(dx) && den_ok(dy);
    const Prim<double> q = primitive_sd(U[0], U[1], U[2], U[3], gamma, ok);
#else
    const Prim<double> q = primitive(U[0], U[1], U[2], U[3], gamma);
#endif
    const double ex = fabs(q.u) + q.c;
    const double ey = fabs(q.v) + q.c;
#if WENOFD_RECIP_DIV
    (void)dx;
    (void)dy;
    return cfl / (ex * rdx + ey * rdy);
#elif WENOFD_SHARED_DIV
    (void)rdx;
    (void)rdy;
    const double s = div_exact(ex, dx, rcp_refined(dx), ok) + div_exact(ey, dy, rcp_refined(dy), ok);
    ok = ok && den_ok(s);
    double t = div_exact(cfl, s, rcp_refined(s), ok);
    if (__any(!ok)) {
        if (!ok) t = local_dt(U, gamma, cfl, dx, dy);
    }
    return t;
#else
    (void)rdx;
    (void)rdy;
    return cfl / (ex / dx + ey / dy);
#endif
}

__device__ __forceinline__ void cell_rhs4(const double fxl[4], const double fxr[4], const double fyb[4],
                                          const double fyt[4], double dx, double dy, double rdx, double rdy,
                                          double rhs[4]) {
#if WENOFD_RECIP_DIV
    (void)dx;
    (void)dy;
#pragma unroll
    for (int r = 0; r < 4; ++r) rhs[r] = -(fxr[r] - fxl[r]) * rdx + -(fyt[r] - fyb[r]) * rdy;
#elif WENOFD_SHARED_DIV
    (void)rdx;
    (void)rdy;
    bool ok = den_ok(dx) && den_ok(dy);
    const double ex = rcp_refined(dx), ey = rcp_refined(dy);
#pragma unroll
    for (int r = 0; r < 4; ++r)
        rhs[r] = div_exact(-(fxr[r] - fxl[r]), dx, ex, ok) + div_exact(-(fyt[r] - fyb[r]), dy, ey, ok);
    if (__any(!ok)) {
        if (!ok) {
#pragma unroll
            for (int r = 0; r < 4; ++r) rhs[r] = cell_rhs(fxl[r], fxr[r], fyb[r], fyt[r], dx, dy);
        }
    }
#else
    (void)rdx;
    (void)rdy;
#pragma unroll
    for (int r = 0; r < 4; ++r) rhs[r] = cell_rhs(fxl[r], fxr[r], fyb[r], fyt[r], dx, dy);
#endif
}

// RK stage combination (src/WENOFD.cpp:394): (a Un + b U) + (c dt) rhs.
template <class T>
WENOFD_HD T rk_combine(double a, T un, double b, T u, double cdt, T rhs) {
    return (a * un + b * u) + cdt * rhs;
}

}  // namespace wenofd
