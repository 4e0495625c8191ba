"""run() and the output files (src/WENOFD.cpp:189-299, 1161-1382), through the
C ABI and through the drop-in program."""
import glob
import os
import subprocess

import numpy as np
import pytest

import wenofd as W
from conftest import ROOT

pytestmark = pytest.mark.gpu


def _read_accuracy(path):
    with open(path) as f:
        lines = f.read().split("\n")
    assert lines[0] == "elemNum, Linf-norm (rho), L1-norm (rho), L2-norm (rho)"
    return np.array([float(v) for v in lines[1].split(",")])


def test_run_writes_reference_files(tmp_path, golden):
    print("[gpu] run() SMOOTH 16x16 JS RK3 CFL 0.5", flush=True)
    ref = golden("run_accuracy")["smooth_s0_n16_cfl0.5"]
    with W.Solver(TESTCASE=0, SCHEME=0, RKMETH=2, CFL=0.5, ELEMNUMX=16, ELEMNUMY=16,
                  OUTPUTDIR=str(tmp_path) + "/") as s:
        s.run()
        d = s.output_dir()
        assert abs(s.info().now - 1.0) <= 1e-9
    assert os.path.basename(d.rstrip("/")).endswith("_SMOOTH_WENOJS_16x16")
    names = sorted(os.listdir(d))
    assert names == ["accuracy_final.csv", "error_final.plt", "solution_final.plt", "solution_initial.plt"]
    acc = _read_accuracy(os.path.join(d, "accuracy_final.csv"))
    assert acc[0] == 16
    assert np.allclose(acc[1:], ref[1:], rtol=1e-11, atol=0)
    with open(os.path.join(d, "solution_final.plt")) as f:
        lines = f.read().splitlines()
    assert lines[0] == "TITLE=DGSolution"
    assert lines[1] == "VARIABLES=X Y rho u v pre c E "
    assert lines[2] == "ZONE T=TA, I=16, J=16, DATAPACKING=POINT"
    assert len(lines) == 3 + 16 * 16
    x, y, rho = (float(v) for v in lines[3].split()[:3])
    assert x == 0 and y == 0 and abs(rho - 1.0) < 0.01
    with open(os.path.join(d, "error_final.plt")) as f:
        assert f.readline().strip() == "TITLE=DGSolution"


def test_dropin_program_end_to_end(tmp_path, golden):
    exe = os.path.join(ROOT, "fd-weno-eulerequation_amd", "bin", "wenofd_amd")
    print("[gpu] drop-in program SMOOTH 16x16 Z RK3 CFL 0.5", flush=True)
    os.makedirs(tmp_path / "input")
    (tmp_path / "input" / "config.cfg").write_text(
        "# 0:SMOOTH 1:riemmnn1 2:riemmnn2 3:RTI\nTESTCASE = 0\nELEMNUMX = 16\nELEMNUMY = 16\n"
        "# 0:RK1 1:RK2 2:RK3\nRKMETH = 2\nCFL = 0.5\n# 0:WENOJS 1:WENOZ 2:WENOZP 3:WENOZPI \nSCHEME = 1\n")
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Iteration count: 1," in r.stdout
    csv = glob.glob(str(tmp_path / "output" / "*_SMOOTH_WENOZ_16x16" / "accuracy_final.csv"))
    assert len(csv) == 1
    ref = golden("run_accuracy")["smooth_s1_n16_cfl0.5"]
    assert np.allclose(_read_accuracy(csv[0])[1:], ref[1:], rtol=1e-11, atol=0)


@pytest.mark.parametrize("scheme", [0, 1])
def test_run_files_equal_the_references(tmp_path, golden, scheme):
    """CWENOFD::run on SMOOTH 16^2 (RK3, CFL 0.5) writes the same four files as
    the reference's own run() (src/WENOFD.cpp:1161-1382): the Tecplot solution
    (initial and final) and error files and the accuracy CSV, byte for byte
    against the reference built without contraction (whose state the GPU path
    equals bit for bit), and the solution files also against its makefile
    build."""
    print(f"[gpu] run() SMOOTH 16x16 scheme {scheme} RK3 CFL 0.5: files against the reference's", flush=True)
    with W.Solver(TESTCASE=0, SCHEME=scheme, RKMETH=2, CFL=0.5, ELEMNUMX=16, ELEMNUMY=16,
                  OUTPUTDIR=str(tmp_path) + "/") as s:
        s.run()
        d = s.output_dir()
    ref = os.path.join(ROOT, "tests", "golden", "run_files")
    for build, names in (("ref_nofma", ("solution_initial.plt", "solution_final.plt", "error_final.plt",
                                        "accuracy_final.csv")),
                         ("ref_native", ("solution_initial.plt", "solution_final.plt"))):
        for name in names:
            with open(os.path.join(d, name), "rb") as f:
                got = f.read()
            with open(os.path.join(ref, f"{build}_smooth_s{scheme}_n16_cfl0.5", name), "rb") as f:
                want = f.read()
            assert got == want, (build, name)
