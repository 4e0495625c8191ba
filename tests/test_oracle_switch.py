"""The oracle's learned-switch arithmetic pinned by closed forms (CPU only).

The switch has no counterpart in the reference; these checks fix what the
oracle computes: the upwind-5 flux equals WENO5 with omega == d and its flux
difference is exact on polynomials of degree <= 5 (not 6), features are
invariant to affine maps of the stencil, labels are 0 on smooth data and 1 at
a jump, the hybrid flux with every characteristic field flagged is the
characteristic WENO flux to rounding and its per-field corrections lie along
their right eigenvectors, and the MLP gradient equals central differences of
the loss.
"""
from fractions import Fraction

import numpy as np
import pytest

import oracle as O
from conftest import same


def test_linear5_is_the_exact_rational_formula():
    rng = np.random.default_rng(1)
    for _ in range(200):
        v = rng.normal(size=5) * 10.0 ** rng.integers(-3, 4)
        exact = sum(Fraction(c) * Fraction(float(x)) for c, x in zip((2, -13, 47, 27, -3), v)) / 60
        got = O.linear5(*v)
        assert abs(Fraction(got) - exact) <= 4e-16 * max(abs(float(exact)), np.abs(v).max())


def test_linear5_equals_weno_with_linear_weights():
    """On linear data beta_0 = beta_1 = beta_2, so WENO-JS/Z/ZP weights are d."""
    rng = np.random.default_rng(2)
    for _ in range(50):
        a, s = rng.normal(), rng.normal()
        v = [a + s * k for k in range(5)]
        for scheme in (0, 1, 2):
            assert abs(O.weno5(scheme, v) - O.linear5(*v)) <= 1e-13 * (abs(a) + abs(s) * 5)


@pytest.mark.parametrize("m", range(7))
def test_flux_difference_exact_to_degree_five(m):
    x0 = 0.3125  # exact binary fraction: the samples below are exact doubles for small m
    f = [(x0 + k) ** m for k in range(-3, 3)]  # points j-3 .. j+2
    hp = O.linear5(*f[1:6])   # h_{j+1/2} from j-2 .. j+2
    hm = O.linear5(*f[0:5])   # h_{j-1/2} from j-3 .. j+1
    exact = m * x0 ** (m - 1) if m else 0.0
    err = abs((hp - hm) - exact)
    if m <= 5:
        assert err < 1e-12
    else:
        assert err > 1.0  # degree 6 is not reproduced: the check discriminates


def test_features_are_affine_invariant():
    rng = np.random.default_rng(3)
    for _ in range(100):
        w = rng.normal(size=5)
        x = O.stencil_features(w)
        assert np.isclose(np.abs(x).max(), 1.0, atol=1e-6) and abs(x.sum()) < 1e-5
        a, b = rng.uniform(0.5, 20.0), rng.normal() * 5
        assert np.allclose(O.stencil_features(a * w + b), x, atol=2e-6)
        assert np.allclose(O.stencil_features(-a * w + b), -x, atol=2e-6)
    # a flat stencil (below the 1e-10 scale) maps to ~0
    assert np.abs(O.stencil_features(np.full(5, 3.0) + np.array([0, 1e-15, 0, -1e-15, 0]))).max() < 1e-4


def test_labels_smooth_zero_jump_one():
    # the label is departure > 0.2 (wenofd_oracle.cpp face_samples)
    x = np.linspace(0, 1, 5) * 0.01
    assert O.weno_z_departure(np.sin(x)) < 0.2            # smooth, well resolved
    for w in ([1, 1, 1, 0, 0.0], [1, 1, 0, 0, 0.0], [0, 0, 0, 0, 1.0]):  # jumps inside the stencil
        assert O.weno_z_departure(np.array(w)) > 0.2
    assert O.weno_z_departure(np.full(5, 2.0)) == 0.0        # constant: omega == d exactly? beta = 0 -> tau = 0


def test_samples_on_uniform_flow_are_unflagged_and_smooth():
    s = O.Solver(4, 1, 2, 0.5, 12, 10, vortex=[0.0, 5.0, 5.0])  # zero-strength vortex = uniform flow
    for d in (0, 1):
        X, Y = s.samples(d)
        assert X.shape[0] == (10 * 13 if d == 0 else 11 * 12)
        assert not Y.any()
        assert np.abs(X).max() < 1e-3  # differences at round-off level are scaled by 1e-10 at most


def test_samples_flag_the_quadrant_interfaces():
    n = 24
    s = O.Solver(5, 1, 2, 0.5, n, n)
    X, Y = s.samples(0)
    Y = Y.reshape(n, n + 1, 4).any(axis=2)  # [row j][face f]
    # the x faces next to x = 0.5 (between points 11 and 12 -> face 12) see the jump
    assert Y[:, 12].all()
    # faces far from both interfaces in a constant quadrant are clean
    assert not Y[2:6, 2:6].any()


def test_cheap_flux_is_the_exact_rational_formula():
    """The cheap flux in its point-accumulation order is upwind-5 on F+ plus
    its mirror on F-: (2a - 13b + 47c + 27d - 3e) / 60 over F+ at f-3..f+1 and
    over F- at f+2..f-2, to within the roundings of the accumulation."""
    rng = np.random.default_rng(11)
    for _ in range(200):
        fp = rng.normal(size=5) * 10.0 ** rng.integers(-3, 4)
        fm = rng.normal(size=5) * 10.0 ** rng.integers(-3, 4)
        exact = (sum(Fraction(c) * Fraction(float(x)) for c, x in zip((2, -13, 47, 27, -3), fp)) +
                 sum(Fraction(c) * Fraction(float(x)) for c, x in zip((2, -13, 47, 27, -3), fm[::-1]))) / 60
        got = O.linear_face(fp, fm)
        scale = max(np.abs(fp).max(), np.abs(fm).max())
        assert abs(Fraction(got) - exact) <= 8e-16 * scale * 2.5, (got, float(exact))


@pytest.mark.parametrize("m", range(7))
def test_cheap_flux_difference_exact_to_degree_five(m):
    """With F- = 0 the accumulation form is upwind-5 on F+: its flux
    difference reproduces d/dx of polynomials of degree <= 5, not 6."""
    x0 = 0.3125
    f = [(x0 + k) ** m for k in range(-3, 3)]  # points j-3 .. j+2
    z = np.zeros(5)
    hp = O.linear_face(np.array(f[1:6]), z)  # face j+1/2
    hm = O.linear_face(np.array(f[0:5]), z)  # face j-1/2
    exact = m * x0 ** (m - 1) if m else 0.0
    err = abs((hp - hm) - exact)
    if m <= 5:
        assert err < 1e-12
    else:
        assert err > 1.0


def _face_states(rng, jump):
    rho = 1.0 + 0.2 * rng.uniform(size=6)
    u = 0.3 * rng.normal(size=6)
    v = 0.3 * rng.normal(size=6)
    p = 1.0 + 0.2 * rng.uniform(size=6)
    if jump:
        rho[3:] *= 4.0
        p[3:] *= 6.0
    E = p / 0.4 + 0.5 * rho * (u * u + v * v)
    return np.stack([rho, rho * u, rho * v, E], -1)


@pytest.mark.parametrize("direction", [0, 1])
@pytest.mark.parametrize("scheme", [0, 1, 2, 3])
def test_hybrid_face_fields(direction, scheme):
    """Per face: with every field flagged the hybrid flux is the
    characteristic WENO flux up to rounding (R L = I); with none it is the
    cheap flux; the corrections of single fields add up to the correction of
    all fields; and the correction of field k lies along the k-th right
    eigenvector: the left eigenvectors of the other fields annihilate it."""
    rng = np.random.default_rng(12 + direction + 2 * scheme)
    lam = 0.07
    for jump in (False, True):
        for _ in range(20):
            U6 = _face_states(rng, jump)
            plain = O.face_flux(direction, scheme, U6, zp_lambda=lam)
            lin = O.hybrid_face(direction, scheme, 0, U6, zp_lambda=lam)
            full = O.hybrid_face(direction, scheme, 15, U6, zp_lambda=lam)
            scale = np.abs(plain).max()
            assert np.abs(full - plain).max() <= 1e-13 * scale
            ave = 0.5 * (U6[2] + U6[3])
            L, _ = O.eigen(direction, ave)
            parts = []
            for k in range(4):
                dk = O.hybrid_face(direction, scheme, 1 << k, U6, zp_lambda=lam) - lin
                parts.append(dk)
                proj = L @ dk
                others = np.delete(proj, k)
                assert np.abs(others).max() <= 1e-12 * (np.abs(proj[k]) + scale), (k, proj)
            assert np.abs(sum(parts) - (full - lin)).max() <= 1e-14 * scale * 10


def test_hybrid_all_fields_is_the_plain_solver_to_rounding_and_none_differs():
    a = O.Solver(5, 1, 2, 0.5, 20, 18)
    b = O.Solver(5, 1, 2, 0.5, 20, 18)
    c = O.Solver(5, 1, 2, 0.5, 20, 18)
    b.set_flags(np.full((18, 21), 15, np.uint8), np.full((19, 20), 15, np.uint8))
    c.set_flags(np.zeros((18, 21), np.uint8), np.zeros((19, 20), np.uint8))
    for _ in range(3):
        dt = a.delta_t()
        a.step(dt)
        b.step(dt)
        c.step(dt)
    A, B, C = a.get_state(), b.get_state(), c.get_state()
    rel = np.abs(A - B).max(axis=(0, 1)) / np.abs(A).max(axis=(0, 1))
    assert rel.max() <= 1e-12
    assert not same(A, C)  # the cheap flux alone (NaN here: it oscillates at the quadrant's jumps)


def test_linear_scheme_converges_on_the_vortex():
    """The cheap branch alone is a fifth-order scheme on smooth flow."""
    errs = []
    for n in (24, 48):
        s = O.Solver(4, 1, 2, 0.5, n, n, t_end=0.2)
        s.set_flags(np.zeros((n, n + 1), np.uint8), np.zeros((n + 1, n), np.uint8))
        s.run()
        errs.append(s.accuracy()[1])
    assert errs[1] < errs[0] / 8  # well above first order at these coarse grids


def _random_params(rng):
    P = np.empty(O.N_PARAMS)
    P[0:320] = rng.uniform(-1, 1, 320) / np.sqrt(10)
    P[320:352] = rng.uniform(-1, 1, 32) / np.sqrt(10)
    P[352:1376] = rng.uniform(-1, 1, 1024) / np.sqrt(32)
    P[1376:1408] = rng.uniform(-1, 1, 32) / np.sqrt(32)
    P[1408:1441] = rng.uniform(-1, 1, 33) / np.sqrt(32)
    return P


def test_mlp_gradient_equals_central_differences():
    rng = np.random.default_rng(4)
    P = _random_params(rng)
    X = rng.normal(size=(16, 10)).astype(np.float32)
    Y = (rng.uniform(size=16) < 0.3).astype(np.uint8)
    loss, G = O.mlp_loss_grad(P, X, Y, pos_weight=3.0)
    assert loss > 0
    idx = np.concatenate([rng.choice(320, 12, replace=False), 320 + rng.choice(32, 4, replace=False),
                          352 + rng.choice(1024, 12, replace=False), 1376 + rng.choice(32, 4, replace=False),
                          1408 + rng.choice(32, 4, replace=False), [1440]])
    h = 1e-6
    for q in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[q] += h
        Pm[q] -= h
        fd = (O.mlp_loss_grad(Pp, X, Y, 3.0)[0] - O.mlp_loss_grad(Pm, X, Y, 3.0)[0]) / (2 * h)
        assert abs(fd - G[q]) <= 1e-7 + 1e-5 * abs(G[q]), (q, fd, G[q])


def test_mlp_forward_paths_agree():
    rng = np.random.default_rng(5)
    P = _random_params(rng).astype(np.float32)
    X = rng.normal(size=(64, 10)).astype(np.float32)
    z64 = O.mlp_forward(P, X)
    zb = O.mlp_infer_bf16(P, X)
    assert np.abs(z64 - zb).max() < 0.05 * (1 + np.abs(z64).max())  # bf16 rounding only
    loss, _ = O.mlp_loss_grad(P, X, np.zeros(64, np.uint8), 1.0)
    sp = np.logaddexp(0, z64)
    assert np.isclose(loss, sp.mean(), rtol=1e-12)
