"""CPU safety net, run under the address and undefined-behaviour sanitizers
(host code only; no GPU needed):

- bin/wenofd_index_check brute-forces the kernels' own index maps
  (csrc/wenofd_index.hpp: ghost fill and its inverse, XCD block order, x and y
  tiles, the y sweep's LDS ring schedule) over grid sizes 5..300, every
  boundary kind and every tile shape, with bounds and bijection checks;
- oracle/bin/oracle_san_check drives every entry point of the CPU oracle over
  every test case, scheme and RK method.

Reference: src/WENOFD.cpp:998-1159 (ghost maps), utils/array.h:273-303 (the
reference's own bounds asserts).
"""
import os
import subprocess

from conftest import ROOT

PKG = os.path.join(ROOT, "fd-weno-eulerequation_amd")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")


def _built(exe, make_dir, target):
    if not os.path.exists(exe):
        subprocess.run(["make", "-C", make_dir, target], check=True, capture_output=True)
    return exe


def test_kernel_index_maps_brute_force():
    exe = _built(os.path.join(PKG, "bin", "wenofd_index_check"), PKG, "bin/wenofd_index_check")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=ENV)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "checks passed" in r.stdout


def test_oracle_under_sanitizers():
    d = os.path.join(ROOT, "oracle")
    exe = _built(os.path.join(d, "bin", "oracle_san_check"), d, "bin/oracle_san_check")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=ENV)
    assert r.returncode == 0, r.stderr[-4000:]
    assert r.stdout.strip().endswith("ok")
