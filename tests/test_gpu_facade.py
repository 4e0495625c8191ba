"""The C++ facade (host/CWENOFD.hpp): a compiled program (host/facade_check.cpp)
calls every method the way the reference's own code calls its CWENOFD
(include/WENOFD.hpp:28-124): the constructor from an EulerEquation pointer,
initializeSolver, initializeAve, exchangeGhostCellsValue, setBoundary,
assembleRHS, assembleBoundaryTerm, assembleSourceTerm, calculateDeltaT,
RunRK1/2/3, useWENO, run and the three outputs. Its results are checked here
against the oracle bit for bit."""
import math
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import ROOT, same

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "fd-weno-eulerequation_amd", "bin", "facade_check")


def _bin(d, name, nx, ny):
    return np.fromfile(os.path.join(d, name), dtype="<f8").reshape(nx, ny, 4)


def test_every_facade_method(tmp_path):
    print("[gpu] facade_check: every CWENOFD method", flush=True)
    r = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, timeout=120)
    rep = {}
    with open(tmp_path / "facade.txt") as f:
        for line in f:
            k, *v = line.split()
            rep.setdefault(k, []).append(v)
    assert r.returncode == 0, (r.stderr, rep.get("error"))
    assert rep["ok"] == [["1"]]

    # RTI 12 x 40, WENO-ZP, RK2
    o = O.Solver(3, 2, 1, 0.6, 12, 40)
    rhs = o.assemble_rhs()
    assert same(_bin(tmp_path, "rti_rhs.bin", 12, 40), rhs)
    assert same(_bin(tmp_path, "rti_boundary_source.bin", 12, 40), rhs)
    # the boundary term alone lacks exactly the gravity source (src/WENOFD.cpp:812-813)
    U0 = o.get_state()
    bnd = _bin(tmp_path, "rti_boundary.bin", 12, 40)
    assert same(bnd[..., :2], rhs[..., :2])
    assert same(bnd[..., 3] + U0[..., 2], rhs[..., 3])
    assert same(bnd[..., 2] + U0[..., 0], rhs[..., 2])
    dt = o.delta_t()
    assert float(rep["rti_dt"][0][0]) == dt
    o.step(dt)
    assert same(_bin(tmp_path, "rti_rk2.bin", 12, 40), o.get_state())
    assert rep["rti_equation_kept"] == [["1"]]
    dxdy, _, _ = o.grid()
    line = rep["rti_deltaX"][0]
    assert float(line[0]) == dxdy[0] and float(line[2]) == dxdy[1]
    assert line[4:6] == ["12", "40"] and line[7] == "4"

    # Riemann2 20 x 18, WENO-Z: RK1, then RK3, each with calculateDeltaT
    o = O.Solver(2, 1, 2, 0.5, 20, 18)
    d1 = o.delta_t()
    o.step(d1, 0)
    d3 = o.delta_t()
    o.step(d3, 2)
    assert [float(v) for v in rep["r2_dts"][0]] == [d1, d3]
    assert same(_bin(tmp_path, "r2_rk1_rk3.bin", 20, 18), o.get_state())
    assert same(_bin(tmp_path, "r2_initial.bin", 20, 18), O.Solver(2, 1, 2, 0.5, 20, 18).get_state())

    # useWENO (src/WENOFD.cpp:823-848) of each scheme
    v5 = np.array([[0.3, -1.2, 2.5, 0.7, 0.05]])
    for sc, dx, val in rep["useWENO"]:
        lam = math.pow(float(dx), 2.0 / 3.0)
        assert float(val) == O.weno5(int(sc), v5, lam)[0], sc

    # SMOOTH 16 x 16, WENO-JS, RK3, CFL 0.05: run() to t = 1
    o = O.Solver(0, 0, 2, 0.05, 16, 16)
    o.run()
    assert float(rep["smooth_now"][0][0]) == o.now()
    assert same(_bin(tmp_path, "smooth_final.bin", 16, 16), o.get_state())
    outs = [p for p in os.listdir(tmp_path) if p.endswith("_SMOOTH_WENOJS_16x16")]
    assert len(outs) == 1
    names = set(os.listdir(tmp_path / outs[0]))
    assert {"solution_initial.plt", "solution_final.plt", "accuracy_final.csv", "error_final.plt",
            "solution_again.plt", "error_again.plt", "accuracy_again.csv"} <= names
