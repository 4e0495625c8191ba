// oracle_san_check.cpp — drives every entry point of the CPU oracle
// (wenofd_oracle.cpp) over every test case, scheme and RK method on small
// non-square grids, built with the address and undefined-behaviour sanitizers
// (oracle/Makefile: bin/oracle_san_check). Test infrastructure like the oracle
// itself; run by tests/test_oracle_sanitized.py. Exit status 0 when every
// call returned without a sanitizer report and every result is finite where
// the mathematics says it must be.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

extern "C" {
struct OracleHandle;
double oracle_weno5(int scheme, double a, double b, double c, double d, double e, double zp_lambda);
void oracle_weno5_batch(int scheme, int n, const double* v5, double zp_lambda, double* out);
void oracle_eigen(int dir, const double* U, double gamma, double* L, double* R);
void oracle_split(int dir, const double* U, double gamma, double* Fm, double* Fp);
void oracle_phys_flux(int dir, const double* U, double gamma, double* F);
double oracle_max_eigen(const double* U, double nx, double ny, double gamma);
void oracle_face_flux(int dir, int scheme, double gamma, double zp_lambda, const double* U6, double* Fhat);
void oracle_seeded_quadrant(unsigned long long seed, double* out16);
OracleHandle* oracle_create(int testcase, int scheme, int rk, double cfl, int nx, int ny, double t_end,
                            const double* quad16, const double* vortex3);
void oracle_destroy(OracleHandle* h);
void oracle_grid(OracleHandle* h, double* dxdy, double* xc, double* yc);
void oracle_get_full(OracleHandle* h, double* out);
void oracle_set_full(OracleHandle* h, const double* in);
void oracle_get_interior(OracleHandle* h, double* out);
void oracle_set_interior(OracleHandle* h, const double* in);
void oracle_fill_ghosts(OracleHandle* h);
void oracle_assemble_rhs(OracleHandle* h, double* rhs_out);
double oracle_delta_t(OracleHandle* h);
int oracle_step(OracleHandle* h, int rk, double dt);
int oracle_run(OracleHandle* h, int max_steps, double dt_fixed);
double oracle_now(OracleHandle* h);
void oracle_stencil_features(const double* w5, float* out5);
double oracle_weno_z_departure(const double* w5);
void oracle_samples(OracleHandle* h, int dir, float* X, unsigned char* Y);
void oracle_set_flags(OracleHandle* h, const unsigned char* fx, const unsigned char* fy);
double oracle_mlp_loss_grad(const float* P, const float* X, const unsigned char* Y, int n, double pos_weight,
                            double* G);
void oracle_mlp_forward(const float* P, const float* X, int n, double* z);
void oracle_mlp_infer_bf16(const float* P, const float* X, int n, float* z);
void oracle_accuracy(OracleHandle* h, double* out3);
void oracle_count_ops(int testcase, int scheme, int rk, int nx, int ny, unsigned long long* out);
}

namespace {

int g_fail = 0;

void expect(bool ok, const char* what, int tc, int sc, int rk) {
    if (!ok) {
        std::fprintf(stderr, "oracle_san_check: %s (case %d scheme %d rk %d)\n", what, tc, sc, rk);
        g_fail = 1;
    }
}

bool finite(const std::vector<double>& v) {
    for (double x : v)
        if (!std::isfinite(x)) return false;
    return true;
}

}  // namespace

int main() {
    const int nx = 13, ny = 11;
    long calls = 0;
    for (int tc = 0; tc <= 5; ++tc)
        for (int sc = 0; sc <= 3; ++sc)
            for (int rk = 0; rk <= 2; ++rk) {
                OracleHandle* h = oracle_create(tc, sc, rk, 0.4, nx, ny, 100.0, nullptr, nullptr);
                expect(h != nullptr, "create", tc, sc, rk);
                if (!h) continue;
                std::vector<double> dxdy(3), xc(nx + 8), yc(ny + 8), full((nx + 8) * (ny + 8) * 4),
                    in(nx * ny * 4), rhs(nx * ny * 4), acc(3);
                oracle_grid(h, dxdy.data(), xc.data(), yc.data());
                oracle_fill_ghosts(h);
                oracle_get_full(h, full.data());
                oracle_set_full(h, full.data());
                oracle_assemble_rhs(h, rhs.data());
                expect(finite(rhs), "rhs finite", tc, sc, rk);
                for (int k = 0; k < 2; ++k) {
                    const double dt = oracle_delta_t(h);
                    expect(dt > 0 && std::isfinite(dt), "dt", tc, sc, rk);
                    expect(oracle_step(h, rk, dt) == 0, "step", tc, sc, rk);
                }
                oracle_get_interior(h, in.data());
                expect(finite(in), "state finite", tc, sc, rk);
                oracle_set_interior(h, in.data());
                const int fx = ny * (nx + 1), fy = (ny + 1) * nx;
                std::vector<float> X((size_t)std::max(fx, fy) * 40);
                std::vector<unsigned char> Y((size_t)std::max(fx, fy) * 4);
                oracle_samples(h, 0, X.data(), Y.data());
                oracle_samples(h, 1, X.data(), Y.data());
                std::vector<unsigned char> flx(fx), fly(fy);
                for (int q = 0; q < fx; ++q) flx[q] = (unsigned char)((q * 7 + tc) % 3 == 0);
                for (int q = 0; q < fy; ++q) fly[q] = (unsigned char)((q * 5 + sc) % 3 == 0);
                oracle_set_flags(h, flx.data(), fly.data());
                expect(oracle_step(h, rk, 0.5 * oracle_delta_t(h)) == 0, "switched step", tc, sc, rk);
                oracle_set_flags(h, nullptr, nullptr);
                expect(oracle_run(h, 2, 0.0) >= 0, "run", tc, sc, rk);
                oracle_accuracy(h, acc.data());
                (void)oracle_now(h);
                oracle_destroy(h);
                calls += 20;
            }
    // the per-function entry points
    double U[4] = {1.2, 0.3, -0.4, 2.9}, U6[24], L[16], R[16], Fm[4], Fp[4], F[4], Fh[4];
    for (int k = 0; k < 6; ++k)
        for (int r = 0; r < 4; ++r) U6[4 * k + r] = U[r] * (1.0 + 0.01 * k * (r + 1));
    for (int dir = 0; dir < 2; ++dir) {
        oracle_eigen(dir, U, 1.4, L, R);
        oracle_split(dir, U, 1.4, Fm, Fp);
        oracle_phys_flux(dir, U, 1.4, F);
        for (int sc = 0; sc <= 3; ++sc) oracle_face_flux(dir, sc, 1.4, 0.1, U6, Fh);
    }
    (void)oracle_max_eigen(U, 1.0, 0.0, 1.4);
    std::vector<double> v5(5 * 64), out(64);
    for (int k = 0; k < 5 * 64; ++k) v5[k] = std::sin(0.37 * k) + (k % 11 == 0 ? 3.0 : 0.0);
    for (int sc = 0; sc <= 3; ++sc) {
        oracle_weno5_batch(sc, 64, v5.data(), 0.2, out.data());
        expect(finite(out), "weno5 finite", -1, sc, -1);
        (void)oracle_weno5(sc, 1, 2, 3, 4, 5, 0.2);
    }
    float f5[5];
    oracle_stencil_features(v5.data(), f5);
    (void)oracle_weno_z_departure(v5.data());
    double q16[16];
    oracle_seeded_quadrant(3, q16);
    std::vector<float> P(1441), Xs(40 * 10);
    std::vector<unsigned char> Ys(40);
    for (int k = 0; k < 1441; ++k) P[k] = 0.1f * (float)std::sin(0.11 * k);
    for (int k = 0; k < 400; ++k) Xs[k] = (float)std::cos(0.07 * k);
    for (int k = 0; k < 40; ++k) Ys[k] = (unsigned char)(k % 3 == 0);
    std::vector<double> G(1441), z(40);
    std::vector<float> zb(40);
    (void)oracle_mlp_loss_grad(P.data(), Xs.data(), Ys.data(), 40, 2.0, G.data());
    oracle_mlp_forward(P.data(), Xs.data(), 40, z.data());
    oracle_mlp_infer_bf16(P.data(), Xs.data(), 40, zb.data());
    expect(finite(G) && finite(z), "mlp finite", -1, -1, -1);
    unsigned long long ops[5];
    oracle_count_ops(5, 1, 2, 9, 8, ops);
    expect(ops[0] > 0 && ops[1] > 0, "op counts", 5, 1, 2);
    std::printf("oracle_san_check: %ld solver calls, %s\n", calls, g_fail ? "FAILED" : "ok");
    return g_fail;
}
