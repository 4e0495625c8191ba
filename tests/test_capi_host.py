"""Host-side checks of libwenofd_amd.so that need no GPU: the library loads,
exports every entry point include/wenofd_amd.h declares, and parses config
text with the reference's rules and error behaviour."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import wenofd as W
from conftest import ROOT


def header_functions():
    text = open(os.path.join(ROOT, "include", "wenofd_amd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(wenofd_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    L = W.lib()
    declared = header_functions()
    assert len(declared) >= 25
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(W.EXPORTS) == declared
    assert L.wenofd_version() >= 100


def test_parse_config_reference_file_format():
    # input/config.cfg of the reference, with CR-LF line ends and a stray line
    text = ("# 0:SMOOTH 1:riemmnn1 2:riemmnn2 3:RTI\r\nTESTCASE = 3\r\nELEMNUMX = 200\r\nELEMNUMY = 800\r\n"
            "# 0:RK1 1:RK2 2:RK3\r\nRKMETH = 1\r\nCFL = 0.7\r\n\r\nnot a key line\r\n"
            "# 0:WENOJS 1:WENOZ 2:WENOZP 3:WENOZPI \r\nSCHEME = 0\r\n")
    inf = W.parse_config(text)
    assert (inf.testcase, inf.nx, inf.ny, inf.rkmeth, inf.scheme) == (3, 200, 800, 1, 0)
    assert inf.cfl == 0.7
    # RTI: [0, 0.25] x [0, 1], gamma 5/3, t_end 1.95 (utils/equations.hpp:605-614)
    assert (inf.xL, inf.xR, inf.yL, inf.yR, inf.t_end) == (0.0, 0.25, 0.0, 1.0, 1.95)
    assert inf.gamma == 5.0 / 3.0
    assert inf.dx == 0.25 / 199 and inf.dy == 1.0 / 799
    assert inf.zp_lambda == (0.25 / 199) ** (2.0 / 3.0)


def test_parse_config_values_use_numeric_prefix():
    # std::stoi / std::stod read a numeric prefix and ignore the rest
    inf = W.parse_config(W.config_text(TESTCASE="2 # riemann", SCHEME="1", RKMETH="2", CFL="0.6x",
                                       ELEMNUMX="64", ELEMNUMY="32"))
    assert (inf.testcase, inf.cfl, inf.nx, inf.ny) == (2, 0.6, 64, 32)


@pytest.mark.parametrize("key,val,code", [
    ("TESTCASE", "7", W.E_BAD_ENUM), ("SCHEME", "4", W.E_BAD_ENUM), ("RKMETH", "3", W.E_BAD_ENUM),
    ("SCHEME", "-1", W.E_BAD_ENUM), ("CFL", "abc", W.E_CONFIG), ("ELEMNUMX", "3", W.E_CONFIG),
    ("ELEMNUMY", "", W.E_CONFIG), ("FP_CONTRACT", "2", W.E_BAD_ENUM), ("FP_CONTRACT", "x", W.E_CONFIG),
])
def test_parse_config_errors(key, val, code):
    keys = dict(TESTCASE="0", SCHEME="1", RKMETH="2", CFL="0.5", ELEMNUMX="32", ELEMNUMY="32")
    keys[key] = val
    with pytest.raises(W.WenofdError) as e:
        W.parse_config(W.config_text(**keys))
    assert e.value.code == code


@pytest.mark.parametrize("missing", ["TESTCASE", "SCHEME", "RKMETH", "CFL", "ELEMNUMX", "ELEMNUMY"])
def test_parse_config_missing_key(missing):
    keys = dict(TESTCASE="0", SCHEME="1", RKMETH="2", CFL="0.5", ELEMNUMX="32", ELEMNUMY="32")
    del keys[missing]
    with pytest.raises(W.WenofdError) as e:
        W.parse_config(W.config_text(**keys))
    assert e.value.code == W.E_CONFIG
    assert missing in str(e.value)


def test_extension_keys():
    inf = W.parse_config(W.config_text(TESTCASE="4", SCHEME="1", RKMETH="2", CFL="0.05", ELEMNUMX="100",
                                       ELEMNUMY="100", TEND="0.5"))
    assert (inf.xL, inf.xR, inf.t_end, inf.gamma) == (0.0, 10.0, 0.5, 1.4)
    with pytest.raises(W.WenofdError):
        W.parse_config(W.config_text(TESTCASE="5", SCHEME="1", RKMETH="2", CFL="0.5", ELEMNUMX="32",
                                     ELEMNUMY="32", QUAD_STATES="1,2,3"))


def test_create_errors_before_touching_a_device():
    # config errors are reported before any device work (so also without a GPU)
    h = ctypes.c_void_p()
    rc = W.lib().wenofd_create(b"TESTCASE = 9\nSCHEME = 0\nRKMETH = 2\nCFL = 0.5\nELEMNUMX = 8\nELEMNUMY = 8\n",
                               0, ctypes.byref(h))
    assert rc == W.E_BAD_ENUM and not h.value
    rc = W.lib().wenofd_create_from_file(b"/nonexistent/config.cfg", 0, ctypes.byref(h))
    assert rc == W.E_CONFIG


def test_batch_config_errors_before_touching_a_device():
    """A batch shares SCHEME, RKMETH and FP_CONTRACT; a bad member config names its index."""
    a = W.config_text(TESTCASE=4, SCHEME=1, RKMETH=2, CFL=0.5, ELEMNUMX=16, ELEMNUMY=16)
    b = W.config_text(TESTCASE=4, SCHEME=0, RKMETH=2, CFL=0.5, ELEMNUMX=16, ELEMNUMY=16)
    bad = W.config_text(TESTCASE=4, SCHEME=1, RKMETH=2, CFL=0.5, ELEMNUMX=3, ELEMNUMY=16)
    fused = a + "FP_CONTRACT = 1\n"
    h = ctypes.c_void_p()
    for texts, code, text in (([a, b], W.E_CONFIG, "share SCHEME"), ([a, bad], W.E_CONFIG, "solve 1"),
                              ([a, fused], W.E_CONFIG, "FP_CONTRACT")):
        arr = (ctypes.c_char_p * len(texts))(*[t.encode() for t in texts])
        rc = W.lib().wenofd_batch_create(arr, len(texts), 0, ctypes.byref(h))
        assert rc == code and not h.value
        assert text in W.lib().wenofd_last_error().decode()
    assert W.lib().wenofd_batch_create(None, 0, 0, ctypes.byref(h)) == W.E_ARG


def test_vortex_batch_configs_cover_the_size_range():
    cfgs = W.vortex_batch_configs(n=64, seed=7)
    infos = [W.parse_config(c) for c in cfgs]
    sizes = {i.nx for i in infos}
    assert len(cfgs) == 64 and sizes <= set(range(100, 401, 20)) and len(sizes) > 8
    assert all(i.nx == i.ny and i.testcase == 4 and i.scheme == 1 and i.rkmeth == 2 for i in infos)
    assert cfgs == W.vortex_batch_configs(n=64, seed=7)  # seeded


def test_no_silent_fallback_without_gpu():
    """Without a usable GPU the library refuses to run rather than computing on the CPU."""
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    with pytest.raises(W.WenofdError) as e:
        W.Solver(TESTCASE=0, SCHEME=1, RKMETH=2, CFL=0.5, ELEMNUMX=16, ELEMNUMY=16)
    assert e.value.code in (W.E_HIP, W.E_ARG)
    with pytest.raises(W.WenofdError):
        W.probe_weno5(1, np.zeros((4, 5)))


def test_dropin_binary_reports_missing_config(tmp_path):
    exe = os.path.join(ROOT, "fd-weno-eulerequation_amd", "bin", "wenofd_amd")
    if not os.path.exists(exe):
        pytest.skip("drop-in binary not built")
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 1
    assert "Failed to open file./input/config.cfg" in r.stderr
