"""The learned troubled-stencil switch on the GPU against the CPU oracle.

Bars: harvested features and labels equal the oracle's bit for bit (float ==);
a switched step equals the oracle's hybrid step with the same field masks bit
for bit; with every field of every face flagged the hybrid stage gives the
plain solver's state to rounding (1e-12 relative max-norm: the corrections
rebuild R (WENO) from the cheap flux, R L = I to rounding); with none it is the
cheap flux; the bf16 MFMA inference matches the oracle's bf16 emulation to
2e-2 (absolute, on logits of order 1-10; the MFMA summation order differs) and
the FP64 forward to 0.15; the FP32 training gradient matches the oracle's FP64
gradient to 1e-4 relative (norm) — the oracle's gradient is itself pinned to
central differences in test_oracle_switch.py.
"""
import numpy as np
import pytest

import oracle as O
import wenofd as W
from conftest import relmax, same

pytestmark = pytest.mark.gpu

QUAD_STRONG = np.array([1, 0, 0, 1000, 1, 0, 0, 0.01, 1, 0, 0, 0.01, 1, 0, 0, 0.01], float)


def cfg(tc, sc, rk, cfl, nx, ny, **extra):
    keys = dict(TESTCASE=tc, SCHEME=sc, RKMETH=rk, CFL="%.17g" % cfl, ELEMNUMX=nx, ELEMNUMY=ny, TEND=100)
    keys.update(extra)
    return W.config_text(**keys)


def pair(tc, nx, ny, rk=2, cfl=0.5, seed=None, quad=None):
    extra = {}
    q = None
    if seed is not None:
        extra["QUAD_SEED"] = seed
        q = O.seeded_quadrant(seed)
    if quad is not None:
        extra["QUAD_STATES"] = " ".join("%.17g" % v for v in quad)
        q = np.asarray(quad, float).reshape(4, 4)
    print(f"[gpu] switch case tc={tc} {nx}x{ny} rk={rk} seed={seed}", flush=True)
    g = W.Solver(cfg(tc, 1, rk, cfl, nx, ny, **extra))
    o = O.Solver(tc, 1, rk, cfl, nx, ny, t_end=100.0, quad=q)
    return g, o


def random_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    P = np.empty(W.N_PARAMS, np.float32)
    P[0:320] = rng.uniform(-1, 1, 320) / np.sqrt(10) * scale
    P[320:352] = rng.uniform(-1, 1, 32) / np.sqrt(10)
    P[352:1376] = rng.uniform(-1, 1, 1024) / np.sqrt(32) * scale
    P[1376:1408] = rng.uniform(-1, 1, 32) / np.sqrt(32)
    P[1408:1441] = rng.uniform(-1, 1, 33) / np.sqrt(32) * scale
    return P


def top_threshold(lx, ly, frac=0.005):
    """A logit threshold that flags about frac of the faces of each direction
    (the lower of the two directions' quantiles). Faces of a constant region
    share one logit, so a quantile may sit on a tie that flags nothing: then
    the next distinct value below it is taken."""
    thr = []
    for lg in (lx, ly):
        v = lg.max(axis=2).ravel()
        t = float(np.quantile(v, 1.0 - frac))
        if not (v > t).any():
            below = np.unique(v[v < t])
            assert below.size, "every face has the same logit"
            t = float(below[-1])
        thr.append(t)
    return min(thr)


@pytest.mark.parametrize("tc,seed", [(5, 7), (4, None), (3, None)])
def test_harvest_matches_oracle(tc, seed):
    g, o = pair(tc, 22, 19, seed=seed)
    for _ in range(2):  # a non-trivial state: two steps on both sides
        dt = o.delta_t()
        o.step(dt)
        g.step(dt)
    assert same(g.get_state(), o.get_state())
    for d in (0, 1):
        Xg, Yg = g.harvest(d)
        Xo, Yo = o.samples(d)
        assert same(Xg, Xo), f"features differ in direction {d}"
        assert same(Yg, Yo)
    g.close()


def dil_masks(f, r):
    """Each field's flags dilated over the (2r+1)^2 face neighbourhood: the
    bitwise OR of the masks."""
    p = np.pad(f, r)
    out = np.zeros_like(f)
    for a in range(-r, r + 1):
        for b in range(-r, r + 1):
            out |= p[r + a:r + a + f.shape[0], r + b:r + b + f.shape[1]]
    return out


def masks_of(z, thr):
    """Field masks from logits [..., 4]: bit k set where field k's logit > thr."""
    return sum(((z[..., k] > thr).astype(np.uint8) << k) for k in range(4)).astype(np.uint8)


@pytest.mark.parametrize("rk", [0, 1, 2])
def test_all_fields_is_plain_to_rounding_and_the_oracle_bitwise(rk):
    a, o = pair(5, 20, 17, rk=rk, seed=3)
    o15 = O.Solver(5, 1, rk, 0.5, 20, 17, t_end=100.0, quad=O.seeded_quadrant(3))
    b = W.Solver(cfg(5, 1, rk, 0.5, 20, 17, QUAD_SEED=3))
    c = W.Solver(cfg(5, 1, rk, 0.5, 20, 17, QUAD_SEED=3))
    b.switch_set(W.SWITCH_WENO)
    c.switch_set(W.SWITCH_LINEAR)
    o.set_flags(np.zeros((17, 21), np.uint8), np.zeros((18, 20), np.uint8))
    o15.set_flags(np.full((17, 21), 15, np.uint8), np.full((18, 20), 15, np.uint8))
    for _ in range(3):
        a.step()
        dt_b = o15.delta_t()
        assert b.step() == dt_b
        o15.step(dt_b)
        dt = o.delta_t()
        assert c.step() == dt
        o.step(dt)
    assert same(b.get_state(), o15.get_state())
    assert relmax(b.get_state(), a.get_state()) <= 1e-12
    assert same(c.get_state(), o.get_state())
    assert not same(a.get_state(), c.get_state())
    for s in (a, b, c):
        s.close()


@pytest.mark.parametrize("fp", [0, 1])
def test_all_fields_is_plain_to_rounding_full_size(fp):
    """At 2048^2 (every strip and segment of the hybrid stage in use), every
    field of every face flagged gives the plain solver's state to rounding."""
    keys = dict(TESTCASE=5, SCHEME=1, RKMETH=2, CFL="0.6", ELEMNUMX=2048, ELEMNUMY=2048, QUAD_SEED=1000,
                TEND=100, FP_CONTRACT=fp)
    print(f"[gpu] all fields flagged vs plain 2048x2048 fp={fp}", flush=True)
    with W.Solver(**keys) as a, W.Solver(**keys) as b:
        b.switch_set(W.SWITCH_WENO)
        assert a.advance(4) == 4 and b.advance(4) == 4
        assert relmax(b.get_state(), a.get_state()) <= 1e-12


def test_switched_step_full_size_matches_oracle():
    """One switched RK3 step at 2048^2 with the trained network's masks (after
    30 steps, so that the masks follow the developing waves) equals the
    oracle's hybrid step from the same state with the same masks."""
    n, seed = 2048, 1000
    print(f"[gpu] switched step vs oracle {n}x{n} seed={seed}", flush=True)
    g = W.Solver(cfg(5, 1, 2, 0.6, n, n, QUAD_SEED=seed))
    P, thr = W.load_switch()
    g.switch_set(W.SWITCH_MLP, P, thr)
    assert g.advance(30) == 30
    o = O.Solver(5, 1, 2, 0.6, n, n, t_end=100.0, quad=O.seeded_quadrant(seed))
    o.set_state(g.get_state())
    dt = g.step()
    fx, fy = g.switch_flags()
    assert 0 < (fx != 0).mean() < 1 and 0 < (fy != 0).mean() < 1
    o.set_flags(fx, fy)
    assert o.delta_t() == dt
    o.step(dt)
    assert same(g.get_state(), o.get_state())
    g.close()


def test_mlp_inference_matches_oracle():
    g, _ = pair(5, 40, 36, seed=11)
    X = np.concatenate([g.harvest(0)[0].reshape(-1, 10), g.harvest(1)[0].reshape(-1, 10)])
    rng = np.random.default_rng(0)
    X = np.concatenate([X[rng.choice(len(X), 3000, replace=False)], rng.normal(size=(97, 10)).astype(np.float32)])
    with W.Mlp(seed=1) as m:
        P = random_params(2, scale=2.0)
        m.params = P
        assert same(m.params, P)
        z = m.infer(X)
    zb = O.mlp_infer_bf16(P, X)
    z64 = O.mlp_forward(P, X)
    assert np.abs(z - zb).max() < 2e-2, np.abs(z - zb).max()
    assert np.abs(z - z64).max() < 0.15
    assert np.std(z64) > 0.1  # the check is not vacuous
    g.close()


def test_classifier_logits_flags_and_dilation():
    g, _ = pair(5, 45, 38, seed=5)
    g.step()
    P = random_params(3, scale=2.0)
    with W.Mlp() as m:
        m.params = P
        Xx, _ = g.harvest(0)
        Xy, _ = g.harvest(1)
        zx = m.infer(Xx.reshape(-1, 10)).reshape(38, 46, 4)
        zy = m.infer(Xy.reshape(-1, 10)).reshape(39, 45, 4)
    thr = float(np.median(zx))
    g.switch_set(W.SWITCH_MLP, P, thr)
    fx, fy, lx, ly = g.switch_classify(dilate=False, logits=True)
    assert same(lx, zx) and same(ly, zy)  # the same device arithmetic on the same features
    assert same(fx, masks_of(zx, thr)) and same(fy, masks_of(zy, thr))
    dx, dy, _, _ = g.switch_classify(dilate=True)
    assert same(dx, dil_masks(fx, 1)) and same(dy, dil_masks(fy, 1))
    assert 0 < (fx != 0).mean() < 1 and len(np.unique(fx)) > 4  # several field combinations
    g.close()


@pytest.mark.parametrize("rk", [0, 1, 2])
def test_switched_steps_match_oracle_with_the_same_flags(rk):
    g, o = pair(5, 34, 29, rk=rk, seed=9)
    P = random_params(4, scale=2.0)
    g.switch_set(W.SWITCH_MLP, P, 0.0)
    _, _, lx, ly = g.switch_classify(dilate=False, logits=True)
    thr = top_threshold(lx, ly)  # about 0.5 % of the faces flagged before dilation
    g.switch_set(W.SWITCH_MLP, P, thr)
    mixed_x, mixed_y = [], []
    for _ in range(3):
        dt = g.step()
        fx, fy = g.switch_flags()
        mixed_x.append(0 < (fx != 0).mean() < 1)
        mixed_y.append(0 < (fy != 0).mean() < 1)
        o.set_flags(fx, fy)
        assert o.delta_t() == dt
        o.step(dt)
        assert same(g.get_state(), o.get_state())
    assert any(mixed_x) and any(mixed_y)  # both fluxes ran in both directions
    g.close()


@pytest.mark.parametrize("tc,nx,ny", [(0, 70, 29), (3, 30, 72), (4, 70, 33)])
def test_switched_steps_of_every_boundary_kind_match_oracle(tc, nx, ny):
    """Switched RK3 steps on periodic (SMOOTH, VORTEX) and slip + SPECIAL
    (RTI) boundaries equal the oracle's hybrid steps bit for bit, given the
    same flags (random network, threshold flagging about 0.5 % of the faces)."""
    g, o = pair(tc, nx, ny, rk=2)
    P = random_params(6, scale=2.0)
    g.switch_set(W.SWITCH_MLP, P, 0.0)
    _, _, lx, ly = g.switch_classify(dilate=False, logits=True)
    thr = top_threshold(lx, ly)
    g.switch_set(W.SWITCH_MLP, P, thr)
    mixed_x, mixed_y = [], []
    for _ in range(3):
        dt = g.step()
        fx, fy = g.switch_flags()
        mixed_x.append(0 < (fx != 0).mean() < 1)
        mixed_y.append(0 < (fy != 0).mean() < 1)
        o.set_flags(fx, fy)
        assert o.delta_t() == dt
        o.step(dt)
        assert same(g.get_state(), o.get_state()), tc
    assert any(mixed_x) and any(mixed_y)
    g.close()


@pytest.mark.parametrize("scheme", [0, 2, 3])
def test_switched_steps_of_every_scheme_match_oracle(scheme):
    """The switched sweeps of WENO-JS, -ZP and -ZPI (flagged faces take that
    scheme's characteristic flux, the others upwind-5) equal the oracle's
    hybrid steps bit for bit, given the same flags."""
    seed = 9
    print(f"[gpu] switched steps scheme={scheme} 34x29 seed={seed}", flush=True)
    g = W.Solver(cfg(5, scheme, 2, 0.5, 34, 29, QUAD_SEED=seed))
    o = O.Solver(5, scheme, 2, 0.5, 34, 29, t_end=100.0, quad=O.seeded_quadrant(seed))
    P, thr = W.load_switch()
    g.switch_set(W.SWITCH_MLP, P, thr)
    mixed_x, mixed_y = [], []
    for _ in range(3):
        dt = g.step()
        fx, fy = g.switch_flags()
        mixed_x.append(0 < (fx != 0).mean() < 1)
        mixed_y.append(0 < (fy != 0).mean() < 1)
        o.set_flags(fx, fy)
        assert o.delta_t() == dt
        o.step(dt)
        assert same(g.get_state(), o.get_state()), scheme
    assert any(mixed_x) and any(mixed_y)
    g.close()


def nan_time_step(U, gamma=1.4):
    """Whether any point's local CFL step is NaN (the library's NaN test; the
    reference's own minimum loop can lose a NaN that a finite value follows)."""
    with np.errstate(all="ignore"):
        rho = U[..., 0]
        u, v = U[..., 1] / rho, U[..., 2] / rho
        p = (gamma - 1.0) * (U[..., 3] - (0.5 * rho) * (u * u + v * v))
        c = np.sqrt((gamma * p) / rho)
        return bool(np.isnan(np.abs(u) + c).any() or np.isnan(np.abs(v) + c).any())


def test_fallback_redoes_a_failed_switched_step_with_the_plain_solver():
    n = 24
    g, o = pair(5, n, n, cfl=0.6, quad=QUAD_STRONG)
    g.switch_set(W.SWITCH_LINEAR)
    zeros = (np.zeros((n, n + 1), np.uint8), np.zeros((n + 1, n), np.uint8))
    o.set_flags(*zeros)
    prev = None
    fallbacks = 0
    for _ in range(6):
        dt = o.delta_t()
        if nan_time_step(o.get_state()):  # the step before left a NaN time step: redo it without the switch
            o.set_state(prev[0])
            o.set_flags(None, None)
            o.step(prev[1])
            o.set_flags(*zeros)
            fallbacks += 1
            dt = o.delta_t()
            assert not nan_time_step(o.get_state())
        prev = (o.get_state(), dt)
        o.step(dt)
        assert g.step() == dt
        assert same(g.get_state(), o.get_state())
    assert fallbacks >= 1 and g.switch_fallbacks() == fallbacks
    g.close()


def test_training_gradient_matches_oracle():
    g, _ = pair(5, 40, 40, seed=13)
    g.step()
    with W.Mlp(seed=3) as m:
        m.harvest(g, keep_pos=1.0, keep_neg=0.3, seed=1)
        n, pos = m.size()
        assert n > 1000 and 0 < pos < n
        X, Y = m.samples()
        rng = np.random.default_rng(1)
        idx = rng.choice(n, 700, replace=False).astype(np.int32)
        loss, G = m.grad(idx, pos_weight=2.5)
        lo, Go = O.mlp_loss_grad(m.params, X[idx], Y[idx], pos_weight=2.5)
    assert abs(loss - lo) <= 1e-5 * max(1.0, abs(lo))
    assert np.linalg.norm(G - Go) <= 1e-4 * np.linalg.norm(Go)
    assert np.abs(G - Go).max() <= 1e-4 * np.abs(Go).max()
    g.close()


def test_adam_steps_follow_kingma_ba():
    """k_adam is Adam (Kingma & Ba 2015, Algorithm 1) with lr, beta1 0.9,
    beta2 0.999, eps 1e-8. From zero moments its first step is, per
    parameter, -lr g / (|g| + eps): lr against the sign of the gradient
    wherever |g| >> eps. The second step is the published update recomputed
    in float64 from the two full-batch gradients (each taken at the
    parameters the step starts from)."""
    rng = np.random.default_rng(11)
    n, lr = 512, 1e-3
    X = rng.uniform(-1, 1, (n, 10)).astype(np.float32)
    Y = (X[:, 2] - X[:, 7] > 0.1).astype(np.uint8)
    idx = np.arange(n, dtype=np.int32)
    with W.Mlp(seed=5) as m:
        m.add(X, Y)
        P0 = m.params.astype(np.float64)
        _, g1 = m.grad(idx)
        m.train(epochs=1, batch=n, lr=lr, seed=1)
        P1 = m.params.astype(np.float64)
        _, g2 = m.grad(idx)
        m.train(epochs=1, batch=n, lr=lr, seed=2)
        P2 = m.params.astype(np.float64)
    g1, g2 = g1.astype(np.float64), g2.astype(np.float64)
    big = np.abs(g1) > 1e-4
    assert big.sum() > 600  # ReLU units that no sample activates have zero gradient
    step1 = P1 - P0
    assert np.all(np.abs(step1) <= lr * (1 + 1e-5))
    assert np.all(np.abs(step1[big] + lr * np.sign(g1[big])) <= 2e-3 * lr)
    m2 = 0.9 * (0.1 * g1) + 0.1 * g2
    v2 = 0.999 * (0.001 * g1 * g1) + 0.001 * g2 * g2
    c1, c2 = 1 - 0.9 ** 2, 1 - 0.999 ** 2
    want = -lr * (m2 / c1) / (np.sqrt(v2 / c2) + 1e-8)
    ok = big & (np.abs(g2) > 1e-4)
    assert np.all(np.abs((P2 - P1)[ok] - want[ok]) <= 5e-3 * lr)


def test_harvest_keeps_a_deterministic_subset():
    g, o = pair(5, 30, 26, seed=21)
    Xall = np.concatenate([g.harvest(0)[0].reshape(-1, 10), g.harvest(1)[0].reshape(-1, 10)])
    Yall = np.concatenate([g.harvest(0)[1].reshape(-1), g.harvest(1)[1].reshape(-1)])
    with W.Mlp() as a, W.Mlp() as b:
        assert a.harvest(g, 1.0, 1.0, seed=5) == len(Yall)
        Xa, Ya = a.samples()
        assert same(Xa, Xall) and same(Ya, Yall)  # keep-all preserves the face order
        na = b.harvest(g, 0.5, 0.1, seed=7)
        Xb, Yb = b.samples()
        b.clear()
        assert b.harvest(g, 0.5, 0.1, seed=7) == na
        assert same(b.samples()[0], Xb)
        assert 0 < na < len(Yall)
    g.close()


def test_training_is_deterministic_and_lowers_the_loss():
    g, _ = pair(5, 48, 48, seed=17)
    for _ in range(3):
        g.step()
    runs = []
    for _ in range(2):
        with W.Mlp(seed=4) as m:
            m.harvest(g, keep_pos=1.0, keep_neg=0.5, seed=2)
            losses = m.train(epochs=4, batch=1024, lr=3e-3, pos_weight=2.0, seed=9)
            runs.append((losses, m.params))
    assert same(runs[0][0], runs[1][0]) and same(runs[0][1], runs[1][1])
    assert runs[0][0][-1] < runs[0][0][0]
    g.close()


def test_ablation_trained_switch_beats_random_flags_at_the_same_rate():
    """What the trained network gains over the method without it: at the same
    fraction of flagged (face, field) pairs, masks placed by the network keep
    the solution much closer to plain WENO-Z than masks placed at random, and
    the cheap flux alone (no network) cannot run the problem without the plain
    fallback."""
    n, seed, t_end = 128, 1003, 0.1
    P, thr = W.load_switch()
    cfg_ = cfg(5, 1, 2, 0.6, n, n, QUAD_SEED=seed, TEND=t_end)
    print(f"[gpu] ablation {n}^2 seed {seed}", flush=True)
    with W.Solver(cfg_) as s:
        s.advance(0)
        plain = s.get_state()[..., 0]
    with W.Solver(cfg_) as s:
        s.switch_set(W.SWITCH_MLP, P, thr)
        steps = s.advance(0)
        trained = s.get_state()[..., 0]
        assert s.switch_fallbacks() == 0
    # random flags at the trained run's final flag rate, drawn every step
    rng = np.random.default_rng(0)
    with W.Solver(cfg_) as s:
        s.switch_set(W.SWITCH_MLP, P, thr)
        s.advance(0)
        fx, fy = s.switch_flags()
        bits = np.unpackbits(np.concatenate([fx.ravel(), fy.ravel()])[:, None], axis=1)[:, 4:]
        rate = bits.mean()  # flagged (face, field) pairs
    o = O.Solver(5, 1, 2, 0.6, n, n, t_end=t_end, quad=O.seeded_quadrant(seed))

    def random_masks(shape):
        return masks_of(rng.uniform(size=shape + (4,)), 1.0 - rate)

    while abs(o.now() - t_end) > 1e-9:
        o.set_flags(random_masks(fx.shape), random_masks(fy.shape))
        dt = o.delta_t()
        if nan_time_step(o.get_state()):
            break
        o.step(dt)
    rand = o.get_state()[..., 0]
    with W.Solver(cfg_) as s:
        s.switch_set(W.SWITCH_LINEAR)
        try:
            s.advance(0)
            linear_failed = False
        except W.WenofdError as e:  # even the plain redo of a step cannot recover the state
            assert e.code == W.E_NAN_DT
            linear_failed = True
        linear_fallbacks = s.switch_fallbacks()
        linear = s.get_state()[..., 0]
    e_trained = np.abs(trained - plain).mean()
    e_rand = np.abs(rand - plain).mean()
    e_lin = np.abs(linear - plain).mean()
    print(f"[gpu] ablation: steps {steps}, flag rate {rate:.3f}, L1 delta trained {e_trained:.3e} "
          f"random {e_rand:.3e} linear {e_lin:.3e} (linear fallbacks {linear_fallbacks}, failed {linear_failed})",
          flush=True)
    assert 0.05 < rate < 0.95
    assert e_trained * 20 < e_rand
    assert linear_failed or linear_fallbacks > 0 or e_trained * 20 < e_lin


def test_dilation_radius_below_the_interval():
    """Interval 4 with radius 1: the first step's flags are the raw flags
    dilated by one face, reused unchanged for four steps, and every switched
    step still equals the oracle's with those flags."""
    g, o = pair(5, 40, 34, seed=19)
    P, thr = W.load_switch()
    g.switch_set(W.SWITCH_MLP, P, thr)
    g.switch_interval(4)
    g.switch_dilation(1)
    rx, ry, _, _ = g.switch_classify(dilate=False)
    seen = []
    for k in range(5):
        dt = g.step()
        fx, fy = g.switch_flags()
        if k == 0:
            assert same(fx, dil_masks(rx, 1)) and same(fy, dil_masks(ry, 1))
        seen.append(fx)
        o.set_flags(fx, fy)
        assert o.delta_t() == dt
        o.step(dt)
        assert same(g.get_state(), o.get_state())
    assert all(same(seen[k], seen[0]) for k in range(4))
    assert not same(seen[4], seen[0])  # reclassified
    g.close()


def test_classification_interval_reuses_dilated_flags():
    g, o = pair(5, 40, 34, seed=19)
    P, thr = W.load_switch()
    g.switch_set(W.SWITCH_MLP, P, thr)
    g.switch_interval(2)
    rx, ry, _, _ = g.switch_classify(dilate=False)
    seen = []
    for k in range(4):
        dt = g.step()
        fx, fy = g.switch_flags()
        if k == 0:
            assert same(fx, dil_masks(rx, 2)) and same(fy, dil_masks(ry, 2))
        seen.append((fx, fy))
        o.set_flags(fx, fy)
        assert o.delta_t() == dt
        o.step(dt)
        assert same(g.get_state(), o.get_state())
    assert same(seen[1][0], seen[0][0]) and same(seen[3][0], seen[2][0])  # reused for a second step
    assert not same(seen[2][0], seen[1][0])  # reclassified
    g.close()
