"""GPU tests of the benchmark workloads at their own sizes.

- The batched workload: 64 seeded isentropic-vortex solves with N drawn from
  {100, 120, ..., 400}, WENO-Z RK3 CFL 0.05, advanced as one batch on the
  device loop. Every solve equals the same config advanced alone through the
  host-driven single-solve loop bit for bit, and the smallest and the largest
  solve equal the oracle bit for bit.
- The bandwidth-scale four-quadrant workload: the seeded (QUAD_SEED = 0)
  4096^2 four-quadrant problem with WENO-Z and WENO-ZP, RK3, CFL 0.6. After 20
  device-loop steps one more step equals the oracle's step from the same state
  bit for bit without contraction, and to the rounding floor (1e-13 relative
  max-norm) with it.

Reference: src/WENOFD.cpp:352-396 (RunRK3), :398-452 (calculateDeltaT).
"""
import re

import numpy as np
import pytest

import oracle as O
import wenofd as W
from conftest import relmax, same

pytestmark = pytest.mark.gpu


def _key(text, key):
    m = re.search(r"^%s = (.+)$" % key, text, re.M)
    return m.group(1) if m else None


def test_batch_of_64_vortex_solves_n100_to_400():
    cfgs = W.vortex_batch_configs(n=64, seed=0)
    sizes = [int(_key(c, "ELEMNUMX")) for c in cfgs]
    print(f"[gpu] batch: 64 vortex solves N {min(sizes)}..{max(sizes)}, 5 steps", flush=True)
    steps = 5
    with W.Batch(cfgs) as b:
        done, active = b.advance(steps)
        assert done == steps and active == 64
        got = [b.get_state(k) for k in range(64)]
        infos = [b.info(k) for k in range(64)]
    for k, c in enumerate(cfgs):
        with W.Solver(c) as s:
            s.set_device_loop(0)
            assert s.advance(steps) == steps
            U = s.get_state()
            inf = s.info()
        assert same(got[k], U), f"solve {k} ({sizes[k]}^2) differs from its single-solve host loop"
        assert infos[k].now == inf.now and infos[k].steps == inf.steps, k
    for k in (int(np.argmin(sizes)), int(np.argmax(sizes))):
        c, n = cfgs[k], sizes[k]
        vort = [float(_key(c, v)) for v in ("VORTEX_EPS", "VORTEX_XC", "VORTEX_YC")]
        print(f"[gpu] batch solve {k}: {n}^2 against the oracle", flush=True)
        o = O.Solver(4, 1, 2, 0.05, n, n, vortex=vort)
        for _ in range(steps):
            o.step(o.delta_t())
        assert same(got[k], o.get_state()), f"solve {k} ({n}^2) differs from the oracle"


@pytest.mark.parametrize("scheme,fp", [(1, 0), (1, 1), (2, 0), (2, 1)])
def test_seeded_quadrant_4096_step_matches_oracle(scheme, fp):
    n = 4096
    print(f"[gpu] seeded quadrant {n}^2 scheme {scheme} FP_CONTRACT={fp}: 20 steps + 1 against the oracle", flush=True)
    keys = dict(TESTCASE=5, SCHEME=scheme, RKMETH=2, CFL="0.6", ELEMNUMX=n, ELEMNUMY=n, QUAD_SEED=0, TEND=100,
                FP_CONTRACT=fp)
    with W.Solver(**keys) as s:
        assert s.advance(20) == 20
        o = O.Solver(5, scheme, 2, 0.6, n, n, t_end=100.0, quad=O.seeded_quadrant(0))
        o.set_state(s.get_state())
        dt = s.step()
        if fp == 0:
            assert o.delta_t() == dt
        else:  # the contracted CFL step multiplies by 1/dx: a few ulps apart at most
            assert abs(o.delta_t() - dt) <= 1e-13 * dt
        o.step(dt)
        U = s.get_state()
    V = o.get_state()
    if fp == 0:
        assert same(U, V)
    else:
        assert relmax(U, V) <= 1e-13
