// wenofd_switch.hpp — arithmetic of the learned troubled-stencil switch.
//
// No counterpart in the reference. The switch decides, per face, between the
// full characteristic WENO-Z flux of the reference (face_flux) and a cheap
// component-wise linear upwind-5 flux. Its inputs are the 5-point stencils of
// the characteristic fields that face_flux reconstructs; its training labels
// say where WENO-Z's nonlinear weights leave the linear ones.
//
//   linear upwind-5 (WENO5 with omega == d, in closed form):
//       h(a,b,c,d,e) = ((((2a - 13b) + 47c) + 27d) - 3e) * (1/60)
//   per face, per conserved component r:
//       F_r = h(F+_r[f-3..f+1]) + h(F-_r[f+2..f-2])
//   features of characteristic field k at a face (10 floats): the + stencil
//   v+_k[0..4] and the mirrored - stencil v-_k[4..0], each centred on its mean
//   and scaled by 1 / (max |deviation| + 1e-10), rounded to float
//   label: max over both sides and j of |omega_j - d_j| > 0.2 (WENO-Z weights)
//   MLP 10 -> 32 -> 32 -> 1 with ReLU, one logit per (face, field);
//   a face is flagged when any of its 4 fields has a logit above the threshold.
#pragma once

#include <hip/hip_runtime.h>

#include "wenofd_math.hpp"

namespace wenofd {

constexpr int kFeat = 10;
constexpr int kHid = 32;
// Parameter vector layout (float): W1[32][10], b1[32], W2[32][32], b2[32], w3[32], b3.
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kHid * kFeat;  // 320
constexpr int kOffW2 = kOffB1 + kHid;          // 352
constexpr int kOffB2 = kOffW2 + kHid * kHid;   // 1376
constexpr int kOffW3 = kOffB2 + kHid;          // 1408
constexpr int kOffB3 = kOffW3 + kHid;          // 1440
constexpr int kNParams = kOffB3 + 1;           // 1441
constexpr double kLabelThreshold = 0.2;
constexpr double kFeatScaleEps = 1e-10;

template <class T>
WENOFD_HD T linear5(T a, T b, T c, T d, T e) {
    return ((((2.0 * a - 13.0 * b) + 47.0 * c) + 27.0 * d) - 3.0 * e) * (1.0 / 60.0);
}

// Normalised 5-point stencil (see the file comment).
template <class T>
WENOFD_HD void stencil_features(T w0, T w1, T w2, T w3, T w4, float* x) {
    const T m = ((((w0 + w1) + w2) + w3) + w4) * 0.2;
    const T d[5] = {w0 - m, w1 - m, w2 - m, w3 - m, w4 - m};
    T s = fabs(d[0]);
#pragma unroll
    for (int j = 1; j < 5; ++j) s = fabs(d[j]) > s ? fabs(d[j]) : s;
    const T inv = 1.0 / (s + kFeatScaleEps);
#pragma unroll
    for (int j = 0; j < 5; ++j) x[j] = (float)(d[j] * inv);
}

// max_j |omega_j - d_j| of WENO-Z on (a..e), weights as the reference
// normalises them (include/Utility.hpp:477-496).
template <class T>
WENOFD_HD T weno_z_departure(T a, T b, T c, T d, T e) {
    T b0, b1, b2, w0, w1, w2;
    weno_beta(a, b, c, d, e, b0, b1, b2);
    weno_alpha<kZ>(b0, b1, b2, 0.0, w0, w1, w2);
    const T s = (w0 + w1) + w2;
    const T e0 = fabs(w0 / s - 0.1), e1 = fabs(w1 / s - 0.6), e2 = fabs(w2 / s - 0.3);
    const T m01 = e0 > e1 ? e0 : e1;
    return m01 > e2 ? m01 : e2;
}

// Characteristic stencils of one face: vp[s] = L F+(point f-3+s), vm[s] =
// L F-(point f-2+s), s = 0..4, with L from the arithmetic-mean state (the
// first half of face_flux, src/WENOFD.cpp:526-593).
template <int D, class T, class GetP, class GetM>
WENOFD_HD void face_char_stencils(const T Ul[4], const T Ur[4], GetP sp, GetM sm, double gamma, T vp[5][4],
                                  T vm[5][4], T R[16]) {
    T ave[4], L[16];
#pragma unroll
    for (int r = 0; r < 4; ++r) ave[r] = 0.5 * (Ul[r] + Ur[r]);
    eigen_matrices<D>(ave, gamma, L, R);
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
}

// Features (x[k][0..9]) and labels (lab[k]) of the 4 fields of one face.
template <class T>
WENOFD_HD void face_samples(const T vp[5][4], const T vm[5][4], float x[4][kFeat], int lab[4]) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        stencil_features(vp[0][k], vp[1][k], vp[2][k], vp[3][k], vp[4][k], x[k]);
        stencil_features(vm[4][k], vm[3][k], vm[2][k], vm[1][k], vm[0][k], x[k] + 5);
        const T dp = weno_z_departure(vp[0][k], vp[1][k], vp[2][k], vp[3][k], vp[4][k]);
        const T dm = weno_z_departure(vm[4][k], vm[3][k], vm[2][k], vm[1][k], vm[0][k]);
        lab[k] = (dp > kLabelThreshold || dm > kLabelThreshold) ? 1 : 0;
    }
}

// Component-wise linear upwind-5 flux of one face from the 6 split points
// (slot 0 = point f-3 ... slot 5 = point f+2).
template <class T, class GetP, class GetM>
WENOFD_HD void face_flux_linear(GetP sp, GetM sm, T Fh[4]) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const T hp = linear5(sp(0, r), sp(1, r), sp(2, r), sp(3, r), sp(4, r));
        const T hm = linear5(sm(5, r), sm(4, r), sm(3, r), sm(2, r), sm(1, r));
        Fh[r] = hp + hm;
    }
}

// The hidden activation, ReLU, and its derivative (0 at the kink). ReLU as
// an integer max with 0 on the float's bits (negative exactly when the sign
// bit is set): one instruction, the same result as x > 0 ? x : 0 for every
// non-NaN x (-0 gives +0).
__device__ __forceinline__ float relu_f(float x) {
    const int b = __float_as_int(x);
    return __int_as_float(b > 0 ? b : 0);
}
__device__ __forceinline__ float relu_grad_f(float x) { return x > 0.0f ? 1.0f : 0.0f; }

}  // namespace wenofd
