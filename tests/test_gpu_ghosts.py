"""Ghost points written by the x sweep.

With RK2 / RK3 on grids of at least 5 points a side, each stage's x sweep also
writes the ghost points of its output (the ghost fill's map inverted), and
only the first stage of a host-driven step or of a device-loop launch runs the
ghost fill (src/WENOFD.cpp:850-929, :998-1159 map the ghosts). After any
number of steps the state's ghost layer must therefore be exactly what a
fresh ghost fill of the same interior gives, for every boundary kind: periodic
(TESTCASE 0), symmetric (1, 5), slip sides with SPECIAL top and bottom (3).
"""
import numpy as np
import pytest

import wenofd as W
from conftest import same

pytestmark = pytest.mark.gpu


def _interior_ghosts(nx, ny):
    """Every point but the corner ghosts (never read, returned as zero)."""
    m = np.ones((nx + 8, ny + 8), bool)
    m[:4, :4] = m[:4, -4:] = m[-4:, :4] = m[-4:, -4:] = False
    return m


@pytest.mark.parametrize("fp", [0, 1])
@pytest.mark.parametrize("graph_steps", [0, 16])
@pytest.mark.parametrize("tc,rk", [(0, 2), (1, 2), (3, 2), (5, 1)])
def test_steps_leave_the_ghost_layer_filled(tc, rk, graph_steps, fp):
    nx, ny = 23, 19
    extra = {"QUAD_SEED": 7} if tc == 5 else {}
    print(f"[gpu] ghosts tc={tc} rk={rk} graph_steps={graph_steps} fp={fp} {nx}x{ny}", flush=True)
    with W.Solver(TESTCASE=tc, SCHEME=1, RKMETH=rk, CFL="0.5", ELEMNUMX=nx, ELEMNUMY=ny, FP_CONTRACT=fp,
                  **extra) as s:
        s.set_device_loop(graph_steps)
        m = _interior_ghosts(nx, ny)
        for n in (3, 17):  # a partial replay, then a whole one plus a remainder
            assert s.advance(n) == n
            G = s.get_state(W.LAYOUT_REF_GHOSTED)
            s.set_boundary()
            F = s.get_state(W.LAYOUT_REF_GHOSTED)
            assert same(G[m], F[m]), (tc, n)


@pytest.mark.parametrize("nx,ny,rk", [(5, 5, 2), (5, 7, 1), (7, 5, 2), (9, 5, 1)])
def test_ghost_layer_at_the_smallest_grids(nx, ny, rk):
    """The smallest grids the library accepts (5 points a side: every ghost
    then copies an interior point), RK2 and RK3, every boundary kind. (RK1
    writes no ghosts: its next step fills them before they are read.)"""
    for tc in (0, 1, 3):
        print(f"[gpu] ghosts small tc={tc} {nx}x{ny} rk={rk}", flush=True)
        with W.Solver(TESTCASE=tc, SCHEME=1, RKMETH=rk, CFL="0.3", ELEMNUMX=nx, ELEMNUMY=ny) as s:
            m = _interior_ghosts(nx, ny)
            for graph_steps in (0, 16):
                s.set_device_loop(graph_steps)
                s.advance(18)
                G = s.get_state(W.LAYOUT_REF_GHOSTED)
                s.set_boundary()
                assert same(G[m], s.get_state(W.LAYOUT_REF_GHOSTED)[m]), (tc, graph_steps)


def test_batch_ghost_layers_equal_the_single_solves():
    """The batch's stages write the same ghost points as the single solves'."""
    cfgs = [dict(TESTCASE=tc, SCHEME=1, RKMETH=2, CFL="0.5", ELEMNUMX=nx, ELEMNUMY=ny, **extra)
            for tc, nx, ny, extra in ((0, 21, 17, {}), (1, 30, 12, {}), (3, 9, 36, {}), (5, 26, 22, {"QUAD_SEED": 3}))]
    ref = []
    for c in cfgs:
        with W.Solver(**c) as s:
            s.set_device_loop(0)
            s.advance(20)
            ref.append(s.get_state(W.LAYOUT_REF_GHOSTED))
    with W.Batch(cfgs) as b:
        done, active = b.advance(20)
        assert done == 20 and active == len(cfgs)
        for k, G in enumerate(ref):
            nx, ny = b.shapes[k]
            m = _interior_ghosts(nx, ny)
            assert same(b.get_state(k, W.LAYOUT_REF_GHOSTED)[m], G[m]), k
