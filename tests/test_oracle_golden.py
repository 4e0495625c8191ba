"""The CPU oracle against the golden vectors written by the compiled reference.

Every function of the oracle must equal the reference's no-contraction build
(ref_nofma) bit for bit, and stay within rounding of the reference's own
makefile build (ref_native, which contracts FMAs). See tests/golden/make_golden.py.
"""
import math

import numpy as np
import pytest

import oracle as O
from conftest import relmax, same


@pytest.mark.parametrize("scheme", [0, 1, 2, 3])
def test_weno_vectors(golden, scheme):
    g = golden("weno_vectors")
    lam = math.pow(float(g["deltaX"]), 2.0 / 3.0)
    out = O.weno5(scheme, g["inputs"], lam)
    assert same(out, g[f"ref_nofma_s{scheme}"])
    nat = g[f"ref_native_s{scheme}"]
    assert np.max(np.abs(out - nat) / np.maximum(np.abs(nat), 1e-300)) < 1e-12


@pytest.mark.parametrize("tag", ["g14", "g53"])
def test_physics_vectors(golden, tag):
    g = golden("physics_vectors")
    gamma = float(g[f"{tag}_gamma"])
    ref = g[f"{tag}_ref_nofma"]
    for k, U in enumerate(g[f"{tag}_states"]):
        Lx, Rx = O.eigen(0, U, gamma)
        Ly, Ry = O.eigen(1, U, gamma)
        Fmx, Fpx = O.split(0, U, gamma)
        Fmy, Fpy = O.split(1, U, gamma)
        mine = np.concatenate([Lx.ravel(), Rx.ravel(), Ly.ravel(), Ry.ravel(), Fmx, Fpx, Fmy, Fpy,
                               [O.max_eigen(U, 1, 0, gamma), O.max_eigen(U, 0, 1, gamma)],
                               O.phys_flux(0, U, gamma), O.phys_flux(1, U, gamma)])
        assert same(mine, ref[k]), k


@pytest.mark.parametrize("tc", [0, 1, 2, 3])
@pytest.mark.parametrize("scheme", [0, 1, 2, 3])
def test_rhs_ghost_dt(golden, tc, scheme):
    g = golden("rhs_cases")
    U = g[f"tc{tc}_state"]
    nx, ny = U.shape[:2]
    s = O.Solver(tc, scheme, 2, 0.6, nx, ny)
    s.set_state(U)
    rhs = s.assemble_rhs()
    assert same(rhs, g[f"tc{tc}_s{scheme}_ref_nofma_rhs"])
    assert relmax(rhs, g[f"tc{tc}_s{scheme}_ref_native_rhs"]) < 1e-12
    assert s.delta_t() == float(g[f"tc{tc}_s{scheme}_ref_nofma_dt"])
    # ghost fill incl. corners (the reference's fill order)
    assert same(s.get_full(), g[f"tc{tc}_ghosted"])


def _traj_names(g):
    return sorted({k[: -len("_spec")] for k in g.files if k.endswith("_spec")})


def test_trajectories(golden):
    g = golden("trajectories")
    names = _traj_names(g)
    assert len(names) >= 10
    for name in names:
        tc, sc, rk, cfl, nx, ny, ns, fx, te = g[f"{name}_spec"]
        tc, sc, rk, nx, ny, ns, fx = map(int, (tc, sc, rk, nx, ny, ns, fx))
        extra = g[f"{name}_extra"]
        quad = extra.reshape(4, 4) if (tc == 5 and extra.size == 16) else None
        vort = extra if (tc == 4 and extra.size == 3) else None
        s = O.Solver(tc, sc, rk, cfl, nx, ny, t_end=te, quad=quad, vortex=vort)
        dts = []
        dt0 = None
        k = 0
        while (k < ns) if ns >= 0 else abs(s.now() - (te if te >= 0 else 0)) > 1e-9:
            dt = dt0 if (fx and k > 0) else s.delta_t()
            if k == 0:
                dt0 = dt
            s.step(dt)
            dts.append(dt)
            k += 1
        U = s.get_state()
        assert same(np.array(dts), g[f"{name}_ref_nofma_dts"]), name
        assert same(U, g[f"{name}_ref_nofma_state"]), name
        # Against the FMA-contracted build: 1e-11 except for RTI, where one rhs
        # evaluation at the initial state amplifies 1-ulp input perturbations to
        # ~3e-7 relative (test_rti_rhs_is_ill_conditioned), so any two correct
        # roundings of the reference differ at that level.
        tol = 1e-11 if not name.startswith("rti") else 1e-3
        assert relmax(U, g[f"{name}_ref_native_state"]) < tol, name


def test_reference_is_rank_count_independent(golden):
    # the fixture generator ran the same quadrant case on 1 and 4 MPI ranks
    assert bool(golden("trajectories")["quadrant_mpi4_equal"])


def test_seeded_quadrant_generator(golden):
    g = golden("trajectories")
    assert same(O.seeded_quadrant(3), g["seeded_quadrant_3"])
    assert same(O.seeded_quadrant(0), g["seeded_quadrant_0"])
    q = O.seeded_quadrant(12345)
    assert np.all((q[:, 0] >= 0.1) & (q[:, 0] < 2.0))
    assert np.all((q[:, 3] >= 0.05) & (q[:, 3] < 2.0))
    assert np.all(np.abs(q[:, 1:3]) <= 1.5)


@pytest.mark.parametrize("key", ["smooth_s0_n16_cfl0.5", "smooth_s1_n16_cfl0.5", "smooth_s1_n12_cfl0.5",
                                 "smooth_s0_n20_cfl0.3"])
def test_reference_run_accuracy(golden, key):
    """The reference's own run() (time loop to t_end with dt clipping) on SMOOTH."""
    ref = golden("run_accuracy")[key]
    sc = int(key.split("_s")[1][0])
    n = int(key.split("_n")[1].split("_")[0])
    cfl = float(key.split("cfl")[1])
    s = O.Solver(0, sc, 2, cfl, n, n)
    assert s.run() > 0
    assert abs(s.now() - 1.0) <= 1e-9
    acc = s.accuracy()
    assert int(ref[0]) == n
    # the CSV prints 16 significant digits of the makefile build (FMA-contracted)
    assert np.allclose(acc, ref[1:], rtol=1e-11, atol=0)


def test_rti_rhs_is_ill_conditioned(golden):
    """The floor for RTI parity: the rhs at the RTI initial state moves by
    ~1e-7 (relative) under random 1-ulp perturbations of the input, in the
    oracle and in the reference alike (the two reference builds differ by as
    much on the same input, rhs_cases fixture tc3)."""
    g = golden("rhs_cases")
    nat = g["tc3_s1_ref_native_rhs"]
    nof = g["tc3_s1_ref_nofma_rhs"]
    s = O.Solver(3, 1, 1, 0.6, 12, 36)
    U0 = s.get_state()
    r0 = s.assemble_rhs()
    rng = np.random.default_rng(0)
    s.set_state(U0 * (1 + rng.choice([-1.0, 1.0], U0.shape) * 2.0 ** -52))
    r1 = s.assemble_rhs()
    sens = np.abs(r1 - r0).max() / np.abs(r0).max()
    assert sens > 1e-9  # a well-conditioned rhs would move by ~1e-15
    assert relmax(nof, nat) < 1e-3
