// wenofd_flops.cpp — counts the floating-point operations of the solver's
// arithmetic by running the kernels' own templated math (wenofd_math.hpp) on an
// operation-counting scalar. Prints one JSON object:
//   per x/y face and scheme, per x/y point split, per cell rhs + RK stage
//   update, per cell local time step; each as {add, mul, div, sqrt, cmp}.
// add counts additions and subtractions; negation and fabs are free (sign-bit
// operations); div and sqrt count once each in "flops".
#include <cstdio>

#include "wenofd_math.hpp"

namespace cnt {

struct Ops {
    unsigned long long add = 0, mul = 0, div = 0, sqrt_ = 0, cmp = 0;
};
static Ops g_ops;

struct F {
    double x;
    F() : x(0) {}
    F(double v) : x(v) {}
};
inline F operator+(F a, F b) { ++g_ops.add; return F(a.x + b.x); }
inline F operator-(F a, F b) { ++g_ops.add; return F(a.x - b.x); }
inline F operator*(F a, F b) { ++g_ops.mul; return F(a.x * b.x); }
inline F operator/(F a, F b) { ++g_ops.div; return F(a.x / b.x); }
inline F operator+(double a, F b) { return F(a) + b; }
inline F operator-(double a, F b) { return F(a) - b; }
inline F operator*(double a, F b) { return F(a) * b; }
inline F operator/(double a, F b) { return F(a) / b; }
inline F operator+(F a, double b) { return a + F(b); }
inline F operator-(F a, double b) { return a - F(b); }
inline F operator*(F a, double b) { return a * F(b); }
inline F operator/(F a, double b) { return a / F(b); }
inline F operator-(F a) { return F(-a.x); }
inline bool operator<(F a, F b) { ++g_ops.cmp; return a.x < b.x; }
inline F fabs(F a) { return F(a.x < 0 ? -a.x : a.x); }
inline F sqrt(F a) { ++g_ops.sqrt_; return F(__builtin_sqrt(a.x)); }

}  // namespace cnt

using cnt::F;

static void print(const char* name, const cnt::Ops& o, bool last = false) {
    std::printf("  \"%s\": {\"add\": %llu, \"mul\": %llu, \"div\": %llu, \"sqrt\": %llu, \"cmp\": %llu, \"flops\": %llu}%s\n",
                name, o.add, o.mul, o.div, o.sqrt_, o.cmp, o.add + o.mul + o.div + o.sqrt_, last ? "" : ",");
}

template <int D, int S>
static cnt::Ops count_face() {
    // a generic state: nonzero, distinct values everywhere
    F Ul[4] = {1.1, 0.3, -0.2, 2.9}, Ur[4] = {1.0, 0.25, -0.15, 2.7};
    F P[6][4], M[6][4];
    for (int s = 0; s < 6; ++s)
        for (int r = 0; r < 4; ++r) {
            P[s][r] = F(0.5 + 0.1 * s + 0.01 * r);
            M[s][r] = F(-0.3 + 0.07 * s - 0.02 * r);
        }
    F Fh[4];
    cnt::g_ops = cnt::Ops();
    wenofd::face_flux<D, S>(
        Ul, Ur, [&](int s, int r) { return P[s][r]; }, [&](int s, int r) { return M[s][r]; }, 1.4, 0.02, Fh);
    return cnt::g_ops;
}

template <int D>
static cnt::Ops count_split() {
    F U[4] = {1.1, 0.3, -0.2, 2.9}, Fm[4], Fp[4];
    cnt::g_ops = cnt::Ops();
    wenofd::flux_split<D>(U, 1.4, Fm, Fp);
    return cnt::g_ops;
}

int main() {
    std::printf("{\n");
    print("face_x_js", count_face<0, wenofd::kJS>());
    print("face_x_z", count_face<0, wenofd::kZ>());
    print("face_x_zp", count_face<0, wenofd::kZP>());
    print("face_x_zpi", count_face<0, wenofd::kZPI>());
    print("face_y_js", count_face<1, wenofd::kJS>());
    print("face_y_z", count_face<1, wenofd::kZ>());
    print("face_y_zp", count_face<1, wenofd::kZP>());
    print("face_y_zpi", count_face<1, wenofd::kZPI>());
    print("split_x", count_split<0>());
    print("split_y", count_split<1>());
    {
        cnt::g_ops = cnt::Ops();
        F r = wenofd::cell_rhs(F(1.0), F(1.5), F(0.2), F(0.1), 0.01, 0.02);
        F u = wenofd::rk_combine(0.75, F(1.0), 0.25, F(1.1), 0.001, r);
        (void)u;
        print("cell_var_stage", cnt::g_ops);  // one conserved variable, one stage
    }
    {
        F U[4] = {1.1, 0.3, -0.2, 2.9};
        cnt::g_ops = cnt::Ops();
        F t = wenofd::local_dt(U, 1.4, 0.6, 0.01, 0.01);
        (void)t;
        print("cell_dt", cnt::g_ops, true);
    }
    std::printf("}\n");
    return 0;
}
