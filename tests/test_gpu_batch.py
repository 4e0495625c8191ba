"""Batched solves (wenofd_batch_*) against the single-solve path and the oracle.

Bar: every solve of a batch equals the same config run alone through
wenofd_advance's host-driven loop bit for bit (IEEE ==), whatever the tile
shape, the graph length and the other members of the batch; the single-solve
path itself is pinned bit for bit to the oracle and the reference in
test_gpu_parity.py. The single solver's device loop (the batch machinery over
one solve) equals its host loop bit for bit.
"""
import numpy as np
import pytest

import oracle as O
import wenofd as W
from conftest import same

pytestmark = pytest.mark.gpu


def cfg(tc, sc, rk, cfl, nx, ny, **extra):
    keys = dict(TESTCASE=tc, SCHEME=sc, RKMETH=rk, CFL="%.17g" % cfl, ELEMNUMX=nx, ELEMNUMY=ny)
    keys.update(extra)
    return W.config_text(**keys)


def single_states(configs, steps):
    """The single-solve path, one host-driven step at a time (the plain kernels)."""
    out = []
    for c in configs:
        with W.Solver(c) as s:
            s.set_device_loop(0)
            s.advance(steps)
            out.append((s.get_state(), s.info()))
    return out


def check_batch(b, ref, label):
    for k, (U, inf) in enumerate(ref):
        got = b.get_state(k)
        bi = b.info(k)
        assert same(got, U), f"{label}: solve {k} ({inf.nx}x{inf.ny}) differs from the single-solve path"
        assert bi.now == inf.now and bi.steps == inf.steps, f"{label}: solve {k} time/steps"


def mixed_vortex(n=6, seed=3, cfl=0.5):
    return W.vortex_batch_configs(n=n, seed=seed, cfl=cfl, sizes=range(20, 101, 8))


def test_vortex_batch_equals_single_solves():
    cfgs = mixed_vortex()
    print(f"[gpu] batch of {len(cfgs)} vortices, 20 steps", flush=True)
    ref = single_states(cfgs, 20)
    with W.Batch(cfgs) as b:
        done, active = b.advance(20)
        assert done == 20 and active == len(cfgs)
        check_batch(b, ref, "graph 16")


@pytest.mark.parametrize("tx", [64, 128, 256])
@pytest.mark.parametrize("ty", [16, 32, 64])
def test_tile_shapes_are_bitwise_identical(tx, ty):
    cfgs = mixed_vortex(n=4, seed=11)
    print(f"[gpu] tiles {tx}x{ty}", flush=True)
    ref = single_states(cfgs, 9)
    with W.Batch(cfgs) as b:
        b.set_tiles(tx, ty)
        assert b.get_tiles() == (tx, ty)
        b.set_graph_steps(4)
        b.advance(9)  # two graph replays + one plain step
        check_batch(b, ref, f"tiles {tx}x{ty}")
    with W.Solver(cfgs[0]) as s:  # the single-solve path under the same shape
        s.set_tiles(tx, ty)
        s.set_device_loop(0)
        s.advance(9)
        assert same(s.get_state(), ref[0][0])


def test_graph_replays_equal_plain_launches():
    cfgs = mixed_vortex(n=5, seed=5)
    print("[gpu] graph vs plain launches", flush=True)
    with W.Batch(cfgs) as a, W.Batch(cfgs) as b:
        a.set_graph_steps(0)
        b.set_graph_steps(3)
        for _ in range(4):  # replays with host work and synchronisation in between
            a.advance(6)
            b.advance(6)
            for k in range(len(cfgs)):
                assert same(a.get_state(k), b.get_state(k))
                assert a.info(k).now == b.info(k).now


@pytest.mark.parametrize("rk,scheme", [(0, 0), (1, 2), (2, 3)])
def test_mixed_test_cases_in_one_batch(rk, scheme):
    """Every test case and boundary type side by side, uneven grids."""
    cfgs = [cfg(0, scheme, rk, 0.4, 17, 13), cfg(1, scheme, rk, 0.4, 23, 9), cfg(2, scheme, rk, 0.4, 19, 21),
            cfg(3, scheme, rk, 0.4, 9, 33), cfg(4, scheme, rk, 0.4, 31, 27),
            cfg(5, scheme, rk, 0.4, 26, 26, QUAD_SEED=17)]
    print(f"[gpu] mixed cases rk={rk} scheme={scheme}", flush=True)
    ref = single_states(cfgs, 7)
    with W.Batch(cfgs) as b:
        b.set_graph_steps(2)
        b.advance(7)
        check_batch(b, ref, f"rk{rk} s{scheme}")


def test_batch_matches_oracle():
    cfgs = [cfg(4, 1, 2, 0.5, 24, 20), cfg(5, 1, 2, 0.5, 16, 28, QUAD_SEED=3)]
    print("[gpu] batch vs oracle", flush=True)
    with W.Batch(cfgs) as b:
        b.advance(5)
        for k, (tc, nx, ny, q) in enumerate([(4, 24, 20, None), (5, 16, 28, 3)]):
            quad = O.seeded_quadrant(q) if q is not None else None
            s = O.Solver(tc, 1, 2, 0.5, nx, ny, quad=quad)
            for _ in range(5):
                s.step(s.delta_t())
            assert same(b.get_state(k), s.get_state())
            assert b.info(k).now == s.now()


def test_solves_stop_at_their_own_end_time():
    ends = [0.05, 0.11, 0.2, 0.083]
    cfgs = [cfg(4, 1, 2, 0.5, 20 + 6 * k, 24, TEND="%.17g" % t) for k, t in enumerate(ends)]
    print("[gpu] run to t_end", flush=True)
    ref = []
    for c in cfgs:
        with W.Solver(c) as s:
            s.set_device_loop(0)
            s.advance(0)
            ref.append((s.get_state(), s.info()))
    with W.Batch(cfgs) as b:
        b.set_graph_steps(8)
        done, active = b.advance(0)
        assert active == 0
        assert done >= max(inf.steps for _, inf in ref)
        check_batch(b, ref, "to t_end")
        for k, t in enumerate(ends):
            assert abs(b.info(k).now - t) <= 1e-9
        assert b.advance(5) == (0, 0)  # nothing left to run


def test_initialize_ave_restarts_every_solve():
    cfgs = mixed_vortex(n=3, seed=2)
    with W.Batch(cfgs) as b:
        first = [b.get_state(k) for k in range(3)]
        b.advance(4)
        b.initialize_ave()
        assert all(same(b.get_state(k), first[k]) for k in range(3))
        assert all(b.info(k).now == 0.0 and b.info(k).steps == 0 for k in range(3))
        b.advance(4)
        ref = single_states(cfgs, 4)
        check_batch(b, ref, "after restart")


def test_autotune_keeps_the_state_and_picks_a_valid_shape():
    cfgs = mixed_vortex(n=6, seed=9)
    print("[gpu] autotune", flush=True)
    ref = single_states(cfgs, 6)
    with W.Batch(cfgs) as b:
        b.advance(3)
        x, y, ms = b.autotune(steps=4)
        assert x in (64, 128, 256) and y in (16, 32, 64) and ms > 0
        assert b.get_tiles() == (x, y)
        b.advance(3)
        check_batch(b, ref, "after autotune")


def test_nan_step_stops_only_that_solve():
    good = cfg(4, 1, 2, 0.5, 20, 20)
    # a quadrant with negative pressure: NaN sound speed in the first dt
    bad = cfg(5, 1, 2, 0.5, 16, 16, QUAD_STATES="1 0 0 1  1 0 0 1  1 0 0 -1  1 0 0 1")
    with W.Batch([good, bad]) as b:
        with pytest.raises(W.WenofdError) as ei:
            b.advance(3)
        assert ei.value.code == W.E_NAN_DT and "solve(s) 1" in str(ei.value)
        assert b.info(0).steps == 3 and b.info(1).steps == 0
        ref = single_states([good], 3)
        assert same(b.get_state(0), ref[0][0])


@pytest.mark.parametrize("tc,rk,steps", [(4, 2, 37), (5, 0, 21), (5, 1, 18), (3, 2, 16), (0, 2, 40)])
def test_single_solver_device_loop_equals_host_loop(tc, rk, steps):
    """wenofd_advance on the device loop (graph replays, dt on the device) gives
    the host-driven loop's state, time and step count bit for bit, including a
    stop at t_end inside a graph replay and a later call that continues."""
    extra = {"QUAD_SEED": 4} if tc == 5 else {}
    c = cfg(tc, 1, rk, 0.5, 26, 22, **extra)
    print(f"[gpu] device loop tc={tc} rk={rk}", flush=True)
    with W.Solver(c) as a, W.Solver(c) as b:
        a.set_device_loop(0)
        assert a.advance(steps) == b.advance(steps) == steps
        assert same(a.get_state(), b.get_state()) and a.info().now == b.info().now
        a.step()  # a host step in between: the device loop picks up its partials
        b.step()
        assert a.advance(5) == b.advance(5) == 5
        assert same(a.get_state(), b.get_state())
        assert a.info().steps == b.info().steps == steps + 6
    t_end = 0.0375
    c2 = cfg(tc, 1, rk, 0.5, 26, 22, TEND="%.17g" % t_end, **extra)
    with W.Solver(c2) as a, W.Solver(c2) as b:
        a.set_device_loop(0)
        na, nb = a.advance(0), b.advance(0)
        assert na == nb and same(a.get_state(), b.get_state())
        assert abs(b.info().now - t_end) <= 1e-9 and b.advance(3) == 0


def test_device_loop_reports_a_nan_step():
    bad = cfg(5, 1, 2, 0.5, 16, 16, QUAD_STATES="1 0 0 1  1 0 0 1  1 0 0 -1  1 0 0 1")
    with W.Solver(bad) as s:
        with pytest.raises(W.WenofdError) as ei:
            s.advance(5)
        assert ei.value.code == W.E_NAN_DT and s.info().steps == 0
