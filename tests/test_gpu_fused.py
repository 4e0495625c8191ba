"""FP_CONTRACT = 1: the sweeps built with FMA contraction.

Bar (the tolerance of the drop-in contract): relative max-norm per conserved
variable <= 1e-11 against both reference builds (the makefile's FMA-contracted
one and the one without contraction) on the golden trajectories, the 800^2
four-quadrant case after 200 steps and the single-stage rhs. RTI is compared
through its rhs only: its trajectories amplify any rounding difference (the
two reference builds already differ by 1e-4 after a few steps). Within the
mode the library's own paths agree bit for bit (batch vs single solve, device
loop vs host loop), as they do without contraction.
"""
import math

import numpy as np
import pytest

import wenofd as W
from conftest import relmax, same

pytestmark = pytest.mark.gpu

TOL = 1e-11


def solver(tc, sc, rk, cfl, nx, ny, **extra):
    print(f"[gpu] fused solver tc={tc} scheme={sc} rk={rk} cfl={cfl} {nx}x{ny} {extra}", flush=True)
    keys = dict(TESTCASE=tc, SCHEME=sc, RKMETH=rk, CFL="%.17g" % cfl, ELEMNUMX=nx, ELEMNUMY=ny, FP_CONTRACT=1)
    keys.update(extra)
    return W.Solver(**keys)


@pytest.mark.parametrize("tc", [0, 1, 2, 3])
@pytest.mark.parametrize("scheme", [0, 1, 2, 3])
def test_fused_rhs_and_dt(golden, tc, scheme):
    g = golden("rhs_cases")
    U = g[f"tc{tc}_state"]
    nx, ny = U.shape[:2]
    with solver(tc, scheme, 2, 0.6, nx, ny) as s:
        s.set_state(U)
        s.set_boundary()
        s.get_flux()
        rhs = s.get_rhs()
        dt = s.calculate_delta_t()
    nat, nof = g[f"tc{tc}_s{scheme}_ref_native_rhs"], g[f"tc{tc}_s{scheme}_ref_nofma_rhs"]
    if tc != 3:
        assert relmax(rhs, nat) <= TOL and relmax(rhs, nof) <= TOL
    else:
        # RTI's rhs cancels to ~1e-3 of its terms (hydrostatic balance): the
        # bound is the two reference builds' own distance, widened 10x
        assert relmax(rhs, nat) <= 10 * max(relmax(nat, nof), TOL)
    ref_dt = float(g[f"tc{tc}_s{scheme}_ref_native_dt"])
    assert abs(dt - ref_dt) <= 1e-14 * ref_dt


def _traj_names(g):
    return sorted({k[: -len("_spec")] for k in g.files if k.endswith("_spec")})


def test_fused_trajectories_within_tolerance(golden):
    g = golden("trajectories")
    for name in _traj_names(g):
        if name.startswith("rti"):
            continue
        tc, sc, rk, cfl, nx, ny, ns, fx, te = g[f"{name}_spec"]
        tc, sc, rk, nx, ny, ns, fx = map(int, (tc, sc, rk, nx, ny, ns, fx))
        extra = g[f"{name}_extra"]
        keys = {}
        if te >= 0:
            keys["TEND"] = "%.17g" % te
        if tc == 4 and extra.size == 3:
            keys.update(VORTEX_EPS="%.17g" % extra[0], VORTEX_XC="%.17g" % extra[1], VORTEX_YC="%.17g" % extra[2])
        if tc == 5 and extra.size == 16:
            keys["QUAD_STATES"] = ",".join("%.17g" % v for v in extra)
        with solver(tc, sc, rk, cfl, nx, ny, **keys) as s:
            if ns < 0:
                while abs(s.info().now - te) > 1e-9:
                    s.step()
            else:
                dt0 = 0.0
                for k in range(ns):
                    dt = s.step(dt0 if fx else 0.0)
                    if k == 0:
                        dt0 = dt
            U = s.get_state()
        assert relmax(U, g[f"{name}_ref_native_state"]) <= TOL, name
        assert relmax(U, g[f"{name}_ref_nofma_state"]) <= TOL, name


def test_fused_c2_quadrant_800_200_steps(golden):
    g = golden("c2_quadrant_800")
    dt0 = float(g["ref_native_dt0"])  # the fixed step of the reference's run
    with solver(5, 1, 2, 0.7, 800, 800) as s:
        assert abs(s.calculate_delta_t() - dt0) <= 1e-14 * dt0
        for _ in range(200):
            s.step(dt0)
        U = s.get_state()
    idx = g["sample_index"]
    samp = U[np.ix_(idx, idx)]
    for key in ("ref_native", "ref_nofma"):
        assert relmax(samp, g[f"{key}_sample"]) <= TOL, key
        ref_sum = g[f"{key}_sum"]
        assert np.all(np.abs(U.sum(axis=(0, 1)) - ref_sum) <= TOL * np.abs(ref_sum)), key


def test_fused_vortex_100_to_t2(golden):
    """C1 with contraction: same step count and state within the bar of the
    reference's run to t = 2."""
    g = golden("vortex_100")
    with solver(4, 0, 2, 0.05, 100, 100) as s:
        n = s.advance()
        U = s.get_state()
    assert n == int(g["ref_native_nsteps"]) == int(g["ref_nofma_nsteps"])
    assert relmax(U, g["ref_nofma_state"]) <= TOL


@pytest.mark.parametrize("scheme", [0, 1])
def test_fused_smooth_orders(golden, scheme):
    """The SMOOTH error norms (README.md:64-74) and observed orders: the
    contraction changes the norms only at the rounding floor."""
    res = {}
    for n in (100, 140, 200):
        with solver(0, scheme, 2, 0.05, n, n) as s:
            s.advance()
            res[n] = s.accuracy()
    ref = golden("run_accuracy")[f"smooth_s{scheme}_n100_cfl0.05"]
    assert np.all(np.abs(res[100] - ref[1:]) <= 1e-12), (res[100], ref[1:])
    ns = sorted(res)
    for a, b in zip(ns[:-1], ns[1:]):
        order = math.log(res[a][1] / res[b][1]) / math.log(b / a)
        assert 4.8 < order < 5.4, (a, b, order)


def test_fused_differs_from_exact_at_the_rounding_floor():
    """The mode reaches the kernels: the two builds give different bits, and
    the difference stays at the rounding level."""
    base = dict(TESTCASE=4, SCHEME=1, RKMETH=2, CFL="0.6", ELEMNUMX=64, ELEMNUMY=64)
    with W.Solver(**base) as a, W.Solver(**base, FP_CONTRACT=1) as b:
        a.advance(10)
        b.advance(10)
        Ua, Ub = a.get_state(), b.get_state()
    assert not same(Ua, Ub)
    assert relmax(Ub, Ua) <= 1e-13


def test_fused_batch_equals_single_solves():
    cfgs = [c + "FP_CONTRACT = 1\n" for c in W.vortex_batch_configs(n=6, seed=5, cfl=0.5, sizes=range(20, 101, 8))]
    ref = []
    for c in cfgs:
        with W.Solver(c) as s:
            s.set_device_loop(0)
            s.advance(20)
            ref.append(s.get_state())
    with W.Batch(cfgs) as b:
        done, active = b.advance(20)
        assert done == 20 and active == len(cfgs)
        for k, U in enumerate(ref):
            assert same(b.get_state(k), U), k


@pytest.mark.parametrize("tc,rk", [(4, 2), (5, 1), (3, 2)])
def test_fused_device_loop_equals_host_loop(tc, rk):
    extra = {"QUAD_SEED": 4} if tc == 5 else {}
    keys = dict(TESTCASE=tc, SCHEME=1, RKMETH=rk, CFL="0.5", ELEMNUMX=26, ELEMNUMY=22, FP_CONTRACT=1, **extra)
    with W.Solver(**keys) as a, W.Solver(**keys) as b:
        a.set_device_loop(0)
        assert a.advance(33) == b.advance(33) == 33
        assert same(a.get_state(), b.get_state()) and a.info().now == b.info().now


@pytest.mark.parametrize("mode,scheme", [("weno", 1), ("linear", 1), ("mlp", 1), ("mlp", 0), ("mlp", 2), ("mlp", 3)])
def test_fused_switched_steps_match_oracle_with_the_same_flags(mode, scheme):
    """The learned switch's sweeps in the contracted mode: each switched step,
    started from the GPU's own state, agrees with the oracle's hybrid step with
    the same flags (all WENO, all upwind-5, or the trained network's; flagged
    faces of every WENO scheme) to the rounding floor."""
    import oracle as O

    base = dict(TESTCASE=5, SCHEME=scheme, RKMETH=2, CFL="0.5", ELEMNUMX=34, ELEMNUMY=29, QUAD_SEED=9, TEND=100,
                FP_CONTRACT=1)
    print(f"[gpu] fused switched steps, flags {mode}, scheme {scheme}", flush=True)
    o = O.Solver(5, scheme, 2, 0.5, 34, 29, t_end=100.0, quad=O.seeded_quadrant(9))
    with W.Solver(**base) as g:
        if mode == "weno":
            g.switch_set(W.SWITCH_WENO)
        elif mode == "linear":
            g.switch_set(W.SWITCH_LINEAR)
        else:
            P, thr = W.load_switch()
            g.switch_set(W.SWITCH_MLP, P, thr)
        for _ in range(3):
            o.set_state(g.get_state())
            dt = g.step()
            fx, fy = g.switch_flags()
            if mode == "mlp":
                # some fields flagged, and some not: both the correction and the cheap flux alone run
                assert (fx != 0).any() and (fx != 15).any() and (fy != 0).any() and (fy != 15).any()
            o.set_flags(fx, fy)
            o.step(dt)
            assert relmax(g.get_state(), o.get_state()) <= 1e-13
