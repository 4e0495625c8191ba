"""Writes the golden vectors under tests/golden/ from the compiled reference.

Runs only in a container that has the reference checkout (its sources are
compiled by oracle/build_ref.sh into oracle/_ref/; nothing of the reference is
stored here, only inputs and outputs). Two builds are driven:

  ref_nofma   -O3 -ffp-contract=off   — the IEEE operation order; the oracle and
                                        the GPU path are held to it bit for bit
  ref_native  -O3 -march=native       — the reference makefile's flags (GCC
                                        contracts FMAs); the 1e-11 parity target
                                        is measured against it

Usage:  python tests/golden/make_golden.py [--skip-big]
"""
import argparse
import glob
import math
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref")
MPIEXEC = "/opt/conda/bin/mpiexec"
BUILDS = ("ref_nofma", "ref_native")


def harness(build, args, cwd, nproc=1):
    cmd = [os.path.join(REF, build)] + [str(a) for a in args]
    if nproc > 1:
        cmd = [MPIEXEC, "-n", str(nproc)] + cmd
    subprocess.run(cmd, check=True, cwd=cwd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)


def f17(x):
    return "%.17g" % x


def random_state(rng, nx, ny, gamma):
    """A valid, non-trivial state: smooth waves plus jumps, all positive rho, p."""
    x = np.linspace(0, 1, nx)[:, None]
    y = np.linspace(0, 1, ny)[None, :]
    rho = 1 + 0.3 * np.sin(6 * x + 2 * y) + 0.2 * (x > 0.5) + 0.05 * rng.uniform(size=(nx, ny))
    u = 0.5 * np.cos(3 * y) + 0.3 * (y > 0.3) + 0.05 * rng.uniform(size=(nx, ny))
    v = 0.4 * np.sin(5 * x * y) - 0.2 * (x < 0.2)
    p = 1 + 0.2 * np.cos(4 * x) + 0.5 * (x + y > 1) + 0.02 * rng.uniform(size=(nx, ny))
    E = p / (gamma - 1) + 0.5 * rho * (u * u + v * v)
    return np.ascontiguousarray(np.stack([rho, rho * u, rho * v, E], -1))


def weno_inputs(rng):
    n = 1024
    v = rng.uniform(-1, 1, (n, 5))
    v[0:128] = 1.0 + np.cumsum(rng.normal(size=(128, 5)) * 1e-3, axis=1)  # smooth
    v[128:192] = rng.uniform(-3, 3, (64, 1))  # constant
    t = np.arange(5)[None, :]
    v[192:256] = rng.uniform(-2, 2, (64, 1)) + rng.uniform(-1, 1, (64, 1)) * t  # linear
    v[256:384] = np.where(t < rng.integers(1, 5, (128, 1)), rng.uniform(0, 2, (128, 1)), rng.uniform(-2, 0, (128, 1)))  # steps
    v[384:448] = 0.5 + 1e-13 * rng.normal(size=(64, 5))  # tiny smoothness indicators
    v[448:512] = 1e3 * rng.uniform(-1, 1, (64, 5))
    v[512:576] = 1e-7 * rng.uniform(-1, 1, (64, 5))
    v[576:640] = np.sin(0.3 * t + rng.uniform(0, 6, (64, 1)))  # smooth extrema
    return np.ascontiguousarray(v)


def physics_inputs(rng, n, gamma):
    rho = rng.uniform(0.1, 2.0, n)
    u = rng.uniform(-2, 2, n)
    v = rng.uniform(-2, 2, n)
    p = rng.uniform(0.05, 3.0, n)
    u[:4] = 0.0
    v[4:8] = 0.0
    E = p / (gamma - 1) + 0.5 * rho * (u * u + v * v)
    return np.ascontiguousarray(np.stack([rho, rho * u, rho * v, E], -1))


def gen_weno(tmp, out):
    rng = np.random.default_rng(20240601)
    v = weno_inputs(rng)
    dx = 1.0 / 127.0
    v.tofile(os.path.join(tmp, "w.bin"))
    res = {"inputs": v, "deltaX": np.array(dx)}
    for b in BUILDS:
        for sc in range(4):
            harness(b, ["weno", sc, f17(dx), "w.bin", "wo.bin"], tmp)
            res[f"{b}_s{sc}"] = np.fromfile(os.path.join(tmp, "wo.bin"))
    np.savez_compressed(os.path.join(out, "weno_vectors.npz"), **res)


def gen_physics(tmp, out):
    rng = np.random.default_rng(20240602)
    res = {}
    for tag, gamma in (("g14", 1.4), ("g53", 5.0 / 3.0)):
        U = physics_inputs(rng, 64, gamma)
        U.tofile(os.path.join(tmp, "p.bin"))
        res[f"{tag}_gamma"] = np.array(gamma)
        res[f"{tag}_states"] = U
        for b in BUILDS:
            harness(b, ["phys", f17(gamma), "p.bin", "po.bin"], tmp)
            res[f"{tag}_{b}"] = np.fromfile(os.path.join(tmp, "po.bin")).reshape(64, 90)
    np.savez_compressed(os.path.join(out, "physics_vectors.npz"), **res)


RHS_CASES = [  # testcase, nx, ny, gamma
    (0, 20, 16, 1.4),
    (1, 16, 20, 1.4),
    (2, 20, 16, 1.4),
    (3, 12, 30, 5.0 / 3.0),
]


def gen_rhs(tmp, out):
    rng = np.random.default_rng(20240603)
    res = {}
    for tc, nx, ny, gamma in RHS_CASES:
        U = random_state(rng, nx, ny, gamma)
        U.tofile(os.path.join(tmp, "s.bin"))
        key = f"tc{tc}"
        res[f"{key}_state"] = U
        for sc in range(4):
            for b in BUILDS:
                harness(b, ["rhs", tc, sc, nx, ny, "s.bin", "r.bin", "g.bin"], tmp)
                r = np.fromfile(os.path.join(tmp, "r.bin"))
                res[f"{key}_s{sc}_{b}_rhs"] = r[:-1].reshape(nx, ny, 4)
                res[f"{key}_s{sc}_{b}_dt"] = np.array(r[-1])
                if sc == 0 and b == "ref_nofma":
                    res[f"{key}_ghosted"] = np.fromfile(os.path.join(tmp, "g.bin")).reshape(nx + 8, ny + 8, 4)
    np.savez_compressed(os.path.join(out, "rhs_cases.npz"), **res)


# name: testcase, scheme, rk, cfl, nx, ny, nsteps, fixed_dt, t_end, extra
TRAJ_CASES = {
    "smooth_z_rk3": (0, 1, 2, 0.6, 24, 24, 10, 0, -1, []),
    "riemann2_js_rk2": (2, 0, 1, 0.7, 32, 32, 12, 0, -1, []),
    "rti_zpi_rk1": (3, 3, 0, 0.5, 12, 40, 6, 0, -1, []),
    "riemann1_zp_rk3": (1, 2, 2, 0.6, 24, 24, 8, 0, -1, []),
    "quadrant_z_rk3_fixed": (5, 1, 2, 0.7, 40, 40, 40, 1, -1, []),
    "vortex_js_rk3": (4, 0, 2, 0.05, 30, 30, 10, 0, -1, []),
    "vortex_z_rk3_param": (4, 1, 2, 0.6, 30, 30, 10, 0, -1, [4.0, 4.5, 5.5]),
    "riemann2_js_rk3_tend": (2, 0, 2, 0.7, 16, 16, -1, 0, 0.02, []),
    "rti_z_rk2": (3, 1, 1, 0.6, 12, 36, 15, 0, -1, []),
}


def seeded_quadrant(seed):
    """splitmix64 -> 53-bit uniforms; per quadrant (UR, UL, LL, LR): rho, u, v, p."""
    mask = (1 << 64) - 1
    s = seed & mask
    out = np.empty((4, 4))

    def nxt():
        nonlocal s
        s = (s + 0x9E3779B97F4A7C15) & mask
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        z = z ^ (z >> 31)
        return (z >> 11) * (1.0 / 9007199254740992.0)

    for q in range(4):
        out[q, 0] = 0.1 + 1.9 * nxt()
        out[q, 1] = -1.5 + 3.0 * nxt()
        out[q, 2] = -1.5 + 3.0 * nxt()
        out[q, 3] = 0.05 + 1.95 * nxt()
    return out


def traj(tmp, b, spec, nproc=1):
    tc, sc, rk, cfl, nx, ny, ns, fx, te, extra = spec
    harness(b, ["traj", tc, sc, rk, f17(cfl), nx, ny, ns, fx, f17(te), "t.bin"] + [f17(e) for e in extra], tmp, nproc)
    r = np.fromfile(os.path.join(tmp, "t.bin"))
    return r[: nx * ny * 4].reshape(nx, ny, 4), r[nx * ny * 4:]


def gen_traj(tmp, out):
    res = {}
    specs = dict(TRAJ_CASES)
    specs["seeded3_zp_rk3"] = (5, 2, 2, 0.6, 32, 32, 15, 0, 0.05, list(seeded_quadrant(3).reshape(16)))
    for name, spec in specs.items():
        res[f"{name}_spec"] = np.array(spec[:9], dtype=np.float64)
        res[f"{name}_extra"] = np.array(spec[9], dtype=np.float64)
        for b in BUILDS:
            U, dts = traj(tmp, b, spec)
            res[f"{name}_{b}_state"] = U
            res[f"{name}_{b}_dts"] = dts
    # one rank vs four ranks: the reference's result does not depend on the slab count
    U1, _ = traj(tmp, "ref_nofma", specs["quadrant_z_rk3_fixed"], 1)
    U4, _ = traj(tmp, "ref_nofma", specs["quadrant_z_rk3_fixed"], 4)
    res["quadrant_mpi4_equal"] = np.array(np.array_equal(U1, U4))
    res["seeded_quadrant_3"] = seeded_quadrant(3)
    res["seeded_quadrant_0"] = seeded_quadrant(0)
    np.savez_compressed(os.path.join(out, "trajectories.npz"), **res)


def gen_run(tmp, out, big):
    """The reference's own CWENOFD::run on SMOOTH: accuracy_final.csv values."""
    res = {}
    runs = [(0, 16, 0.5), (1, 16, 0.5), (1, 12, 0.5), (0, 20, 0.3)]
    if big:
        runs += [(0, 100, 0.05), (1, 100, 0.05)]
    for sc, n, cfl in runs:
        d = tempfile.mkdtemp(dir=tmp)
        harness("ref_native", ["run", 0, sc, 2, f17(cfl), n, n], d)
        csv = glob.glob(os.path.join(d, "output", "*", "accuracy_final.csv"))[0]
        with open(csv) as f:
            line = f.read().splitlines()[1]
        vals = [float(x) for x in line.split(",")]
        res[f"smooth_s{sc}_n{n}_cfl{cfl}"] = np.array(vals)  # N, Linf, L1, L2
    np.savez_compressed(os.path.join(out, "run_accuracy.npz"), **res)


def gen_run_files(tmp, out):
    """The reference's own CWENOFD::run on SMOOTH 16^2 (WENO-JS and WENO-Z, RK3,
    CFL 0.5): its Tecplot solution (initial, final) and error files and its
    accuracy CSV, kept verbatim as the expected output files, from the build
    without contraction (the one the GPU path equals bit for bit) and from the
    makefile build."""
    for b in BUILDS:
        for sc in (0, 1):
            d = tempfile.mkdtemp(dir=tmp)
            harness(b, ["run", 0, sc, 2, f17(0.5), 16, 16], d)
            src = glob.glob(os.path.join(d, "output", "*_SMOOTH_*_16x16"))[0]
            dst = os.path.join(out, "run_files", f"{b}_smooth_s{sc}_n16_cfl0.5")
            os.makedirs(dst, exist_ok=True)
            for name in ("solution_initial.plt", "solution_final.plt", "error_final.plt", "accuracy_final.csv"):
                with open(os.path.join(src, name), "rb") as f:
                    data = f.read()
                with open(os.path.join(dst, name), "wb") as f:
                    f.write(data)


def gen_c2(tmp, out):
    """C2: four-quadrant 800^2, WENO-Z, RK3, step-0 CFL-0.7 dt, 200 steps."""
    spec = (5, 1, 2, 0.7, 800, 800, 200, 1, -1, [])
    res = {}
    full = {}
    for b in BUILDS:
        U, dts = traj(tmp, b, spec, nproc=8)
        full[b] = U
        res[f"{b}_dt0"] = np.array(dts[0])
        res[f"{b}_sum"] = U.sum(axis=(0, 1))
        res[f"{b}_absmax"] = np.abs(U).max(axis=(0, 1))
    idx = np.unique(np.concatenate([np.arange(0, 800, 9), [799]]))
    res["sample_index"] = idx
    for b in BUILDS:
        res[f"{b}_sample"] = np.ascontiguousarray(full[b][np.ix_(idx, idx)])
    a, c = full["ref_nofma"], full["ref_native"]
    res["native_vs_nofma_relmax"] = np.abs(a - c).max(axis=(0, 1)) / np.abs(c).max(axis=(0, 1))
    np.savez_compressed(os.path.join(out, "c2_quadrant_800.npz"), **res)


def gen_vortex(tmp, out):
    """C1: isentropic vortex 100^2, WENO-JS, RK3, CFL 0.05, t_end 2.0 (reference loop)."""
    spec = (4, 0, 2, 0.05, 100, 100, -1, 0, 2.0, [])
    res = {}
    for b in BUILDS:
        U, dts = traj(tmp, b, spec, nproc=1)  # one rank: per-rank coordinates round differently (WENOFD.cpp:76)
        res[f"{b}_nsteps"] = np.array(len(dts))
        res[f"{b}_dt_last"] = np.array(dts[-1])
        if b == "ref_nofma":
            res["ref_nofma_state"] = U
            keep = U
        else:
            res["native_vs_nofma_relmax"] = np.abs(U - keep).max(axis=(0, 1)) / np.abs(U).max(axis=(0, 1))
    np.savez_compressed(os.path.join(out, "vortex_100.npz"), **res)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--skip-big", action="store_true", help="skip the 800^2, vortex and N=100 runs")
    ap.add_argument("--only", default="")
    args = ap.parse_args()
    if not all(os.path.exists(os.path.join(REF, b)) for b in BUILDS):
        subprocess.run([os.path.join(ROOT, "oracle", "build_ref.sh")], check=True)
    tmp = tempfile.mkdtemp(prefix="wenofd_golden_")
    jobs = {"weno": lambda: gen_weno(tmp, HERE), "physics": lambda: gen_physics(tmp, HERE),
            "rhs": lambda: gen_rhs(tmp, HERE), "traj": lambda: gen_traj(tmp, HERE),
            "run": lambda: gen_run(tmp, HERE, not args.skip_big), "run_files": lambda: gen_run_files(tmp, HERE)}
    if not args.skip_big:
        jobs["c2"] = lambda: gen_c2(tmp, HERE)
        jobs["vortex"] = lambda: gen_vortex(tmp, HERE)
    for name, fn in jobs.items():
        if args.only and name not in args.only.split(","):
            continue
        print("generating", name, flush=True)
        fn()


if __name__ == "__main__":
    sys.exit(main())
