"""Headline benchmark: WENO-Z RK3 cell-updates/s, FP64, 2048^2, one MI355X.

Workload (one step = one iteration of the reference's run() loop,
src/WENOFD.cpp:213-239): calculateDeltaT, then RunRK3 = 3 x (ghost fill, x-face
fluxes, y-face fluxes, fused rhs + stage update), on the isentropic vortex on
[0,10]^2 (periodic), 2048 x 2048 points, WENO-Z, CFL 0.6. The state is resident
in HBM before the timed region; the time step is folded and applied on the
device (the reference loop's calculateDeltaT + now += dt), steps are replayed
from a HIP graph.

FP64 arithmetic mode (--fp): "fused" (the default, config key FP_CONTRACT = 1)
runs the sweeps built with FMA contraction, like the reference's own
optimised CPU build, and division by the reciprocal; "exact" the build
without either, bit-identical to the reference compiled without contraction. Both meet the drop-in bar (relative
max-norm <= 1e-11 against the reference), and the line reports both rates.

Prints one JSON line (see the fields below). --gpus > 1 is refused: this
solver runs one whole domain on one GPU.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "fd-weno-eulerequation_amd")
sys.path.insert(0, PKG)

METRIC = "WENO-Z RK3 cell-updates/s (FP64, 2048²) + learned-hybrid speedup at equal L1"
SCHEMES = {"JS": 0, "Z": 1, "ZP": 2, "ZPI": 3}
HBM_PEAK_GBS = 8000.0       # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6     # MI355X FP64 vector peak
STEP_BYTES_PER_CELL_UPDATE = 256  # RK3: Un, U1, U2 read and U1, U2, Un+1 written once, 4 doubles each


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def flop_counts():
    exe = os.path.join(PKG, "bin", "wenofd_flops")
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    return json.loads(out)


def kernel_algorithmic(kind, counts, scheme, nx, ny, rk_first=False):
    """Algorithmic FLOPs and HBM bytes of one launch (per-unit figures x units)."""
    sname = {0: "js", 1: "z", 2: "zp", 3: "zpi"}[scheme]
    if kind == "flux_x_rk":
        faces, points, cells = ny * (nx + 1), ny * (nx + 6), nx * ny
        flops = (faces * counts[f"face_x_{sname}"]["flops"] + points * counts["split_x"]["flops"]
                 + cells * 4 * counts["cell_var_stage"]["flops"])
        # U stencil points, y-face fluxes, Un, updated U
        nbytes = points * 32 + nx * (ny + 1) * 32 + cells * 32 + cells * 32
    elif kind == "flux_y":
        faces, points = nx * (ny + 1), nx * (ny + 6)
        flops = faces * counts[f"face_y_{sname}"]["flops"] + points * counts["split_y"]["flops"]
        nbytes = points * 32 + faces * 32
    elif kind == "delta_t":  # after the first step: folds the partial minima of the last stage
        flops = 0
        nbytes = 0
    else:  # ghost_fill
        flops = 0
        nbytes = 8 * (nx + ny) * 64
    return flops, nbytes


FP_MODES = {"exact": 0, "fused": 1}


def _profile_line(path, tag):
    path = os.path.join(ROOT, "profiles", path)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(tag + ":"):
            return dict(kv.split("=") for kv in line.split(":", 1)[1].split())
    return None


def measured_traffic(kernel, n, fp):
    """HBM bytes per launch from the committed rocprofv3 counter summary
    (FETCH_SIZE + WRITE_SIZE, calibrated), or None when not measured for n."""
    if n != 2048:
        return None
    vals = _profile_line("hbm_traffic_vortex2048_wenoz_rk3.md", "traffic_2048" + ("_fused" if fp == "fused" else ""))
    return int(vals[kernel]) if vals and kernel in vals else None


def measured_valu(kernel, n, fp):
    """VALU wave-instructions per launch from the committed counter summary
    (SQ_INSTS_VALU), or None when not measured for n."""
    if n != 2048:
        return None
    vals = _profile_line("valu_vortex2048_wenoz_rk3.md", "valu_2048" + ("_fused" if fp == "fused" else ""))
    return int(vals[kernel]) if vals and kernel in vals else None


# FP64 VALU issue capacity: 256 CUs x 4 SIMDs, one wave64 FP64 instruction per 4 cycles, 2.4 GHz.
VALU_PEAK_WAVE_INSTR = 256 * 4 * 2.4e9 / 4


def valu_issue(kernel, n, avg_us, fp):
    """Measured VALU wave-instructions per launch over this run's duration,
    against the chip's FP64 issue capacity."""
    w = measured_valu(kernel, n, fp)
    if w is None:
        return None
    achieved = w / (avg_us * 1e-6)
    return {"wave_instructions_per_launch": w, "achieved": achieved, "peak": VALU_PEAK_WAVE_INSTR,
            "unit": "wave-instructions/s", "frac": achieved / VALU_PEAK_WAVE_INSTR}


REF_EXE = os.path.join(ROOT, "oracle", "_ref", "ref_native")
CPU_SHARE = 16  # host cores per GPU on the MI355X boxes
MPIEXEC = "/opt/conda/bin/mpiexec"


def _host_cpu():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _ref_bench(cmd, workdir):
    """Runs the reference build's bench mode; returns its stepping seconds."""
    out = subprocess.run(cmd, cwd=workdir, capture_output=True, text=True, timeout=300, check=True).stdout
    for line in out.splitlines():
        if line.startswith("ranks "):
            return float(line.split()[5])
    raise RuntimeError("no timing line from the reference build")


def _cpu_facts():
    """CPU model, nproc and the CPU share of one GPU on this host."""
    nproc = os.cpu_count() or 1
    return {"host": _host_cpu(), "nproc": nproc, "cpu_share": min(nproc, CPU_SHARE)}


def cpu_baseline(nx, ny, scheme, cfl, steps, mpi_ranks):
    """The reference's own CPU path (its sources compiled with the makefile's
    -O3 -march=native into oracle/_ref/ref_native, driven through its class
    surface) timed on this host over a bounded sample of the same workload:
    serially, and under mpiexec at each rank count of mpi_ranks (capped at the
    host's CPU share of one GPU); the serial CPU oracle when that build is
    absent."""
    sname = [k for k, v in SCHEMES.items() if v == scheme][0]
    facts = _cpu_facts()
    if os.path.exists(REF_EXE):
        args = ["bench", "4", str(scheme), "2", "%.17g" % cfl, str(nx), str(ny), str(steps)]
        try:
            with tempfile.TemporaryDirectory() as wd:  # the reference writes ./output/ at start-up
                el = _ref_bench([REF_EXE] + args, wd)
                rec = {"value": nx * ny * steps / el, "unit": "cell-updates/s", "cores": 1, "kind": "reference",
                       "sample": f"{steps} steps (calculateDeltaT + RunRK3) of the same {nx}x{ny} vortex, WENO-"
                                 f"{sname}, the reference's CWENOFD built with its makefile flags, one MPI rank, "
                                 f"{el:.1f} s"}
                rec.update(facts)
                runs = []
                for ranks in sorted({min(r, facts["cpu_share"]) for r in mpi_ranks if r > 1}):
                    if not os.path.exists(MPIEXEC):
                        break
                    log(f"cpu baseline: {steps} reference steps on {ranks} MPI ranks")
                    el_m = _ref_bench([MPIEXEC, "-n", str(ranks), REF_EXE] + args, wd)
                    runs.append({"ranks": ranks, "cores": ranks, "value": nx * ny * steps / el_m, "seconds": el_m})
                if runs:
                    rec["mpi"] = runs
                return rec
        except (subprocess.SubprocessError, RuntimeError, OSError, ValueError, IndexError) as e:
            log(f"cpu baseline: reference build failed ({e}); timing the oracle instead")
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    s = O.Solver(4, scheme, 2, cfl, nx, ny, t_end=100.0)
    t0 = time.perf_counter()
    for _ in range(steps):
        s.step(s.delta_t())
    el = time.perf_counter() - t0
    rec = {"value": nx * ny * steps / el, "unit": "cell-updates/s", "cores": 1, "kind": "port",
           "sample": f"{steps} steps (calculateDeltaT + RunRK3) of the same {nx}x{ny} vortex, WENO-"
                     f"{sname}, serial C++ oracle, {el:.1f} s"}
    rec.update(facts)
    return rec


def copy_calibration(nbytes):
    """A plain device copy of nbytes (torch's copy kernel: nbytes read, nbytes
    written), for scaling the HBM counters of a profiled run: FETCH_SIZE and
    WRITE_SIZE of this kernel against its known bytes."""
    import torch

    n = nbytes // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 8, "launches": 5}), flush=True)


def batched(W, steps, seq_steps, fp, seed=0):
    """64 seeded vortex solves (N in {100, 120, ..., 400}, strength ~ U[3, 6],
    centre ~ U[3, 7]^2), WENO-Z RK3 CFL 0.05, advanced as one batch (one
    launch per phase, graph-captured steps, autotuned tiles) against the same
    64 solves run one after another through the single-solve path."""
    cfgs = [c + f"FP_CONTRACT = {FP_MODES[fp]}\n" for c in W.vortex_batch_configs(n=64, seed=seed, cfl=0.05)]
    cells = sum(W.parse_config(c).nx * W.parse_config(c).ny for c in cfgs)
    log(f"batch: 64 vortex solves, {cells} cells; autotune")
    b = W.Batch(cfgs)
    x, y, _ = b.autotune(steps=8)
    b.advance(16)
    b.synchronize()
    t0 = time.perf_counter()
    done, active = b.advance(steps)
    b.synchronize()
    el = time.perf_counter() - t0
    assert done == steps and active == 64
    b.close()
    log(f"batch: {el * 1e3 / steps:.3f} ms/step, tiles {x}x{y}; sequential {seq_steps} steps each")
    value = cells * steps / el
    out = {"workload": "64 isentropic-vortex solves, N seeded-uniform in {100,120,...,400}, strength ~U[3,6], "
                       "centre ~U[3,7]^2, WENO-Z RK3 CFL 0.05, per-solve dt on the device",
           "cells": cells, "steps": steps, "value": value, "unit": "cell-updates/s",
           "ms_per_step": el * 1e3 / steps, "tiles": [x, y], "graph_steps": 16}
    if seq_steps <= 0:
        return out
    solvers = [W.Solver(c) for c in cfgs]
    for s in solvers:
        s.advance(2)
        s.synchronize()
    t0 = time.perf_counter()
    for s in solvers:
        s.advance(seq_steps)
    for s in solvers:
        s.synchronize()
    el_seq = time.perf_counter() - t0
    for s in solvers:
        s.close()
    seq = cells * seq_steps / el_seq
    out.update({"sequential_value": seq, "sequential_steps": seq_steps, "speedup_vs_sequential": value / seq})
    return out


HYBRID_CONFIGS = ((8, 1), (16, 1), (16, 2))  # (classification interval, dilation radius)


def hybrid(W, n, seed, t_end, truth, fp):
    """Learned switch against plain WENO-Z on an unseen seeded four-quadrant
    problem (seed outside the training seeds 0..255), RK3 CFL 0.6 to t_end,
    each run timed whole, for each (classification interval, dilation radius)
    of HYBRID_CONFIGS. With truth, all are also compared with plain WENO-Z on
    the nested (2n-1)^2 grid sampled at even indices; the reported speedup is
    the best among the configurations whose L1 error is within 5 % of the
    plain solver's."""
    P, thr = W.load_switch()
    cfg = W.config_text(TESTCASE=5, SCHEME=1, RKMETH=2, CFL=0.6, ELEMNUMX=n, ELEMNUMY=n, QUAD_SEED=seed,
                        TEND="%.17g" % t_end, FP_CONTRACT=FP_MODES[fp])

    def run(interval):
        s = W.Solver(cfg)
        if interval:
            s.switch_set(W.SWITCH_MLP, P, thr)
            s.switch_interval(interval[0])
            s.switch_dilation(interval[1])
        s.synchronize()
        t0 = time.perf_counter()
        steps = s.advance(0)
        s.synchronize()
        el = time.perf_counter() - t0
        r = {"steps": steps, "seconds": el, "ms_per_step": el * 1e3 / steps, "value": n * n * steps / el,
             "rho": s.get_state()[..., 0]}
        if interval:
            fx, fy = s.switch_flags()
            m = np.concatenate([fx.ravel(), fy.ravel()])
            r["flag_rate_final_dilated"] = float((m != 0).mean())  # faces with any field flagged
            r["field_rate_final_dilated"] = float(np.unpackbits(m[:, None], axis=1)[:, 4:].mean())
            r["fallbacks"] = s.switch_fallbacks()
        s.close()
        return r

    log(f"hybrid comparison: plain {n}x{n} seed {seed} to t={t_end}")
    plain = run(None)
    rt = None
    if truth:
        m = 2 * n - 1
        log(f"hybrid comparison: truth {m}x{m}")
        cfg_t = W.config_text(TESTCASE=5, SCHEME=1, RKMETH=2, CFL=0.6, ELEMNUMX=m, ELEMNUMY=m, QUAD_SEED=seed,
                              TEND="%.17g" % t_end, FP_CONTRACT=FP_MODES[fp])
        with W.Solver(cfg_t) as s:
            s.advance(0)
            rt = s.get_state()[::2, ::2, 0]
    out = {"workload": f"unseen seeded four-quadrant problem (seed {seed}) {n}x{n}, WENO-Z RK3 CFL 0.6, t_end {t_end}",
           "threshold": thr, "plain_value": plain["value"], "plain_ms_per_step": plain["ms_per_step"],
           "configs": []}
    if rt is not None:
        ep = np.abs(plain["rho"] - rt)
        out["truth"] = f"plain WENO-Z {2 * n - 1}x{2 * n - 1} at even indices"
        out["plain_error_vs_truth"] = {"L1": float(ep.mean()), "Linf": float(ep.max())}
    for interval, radius in HYBRID_CONFIGS:
        log(f"hybrid comparison: hybrid {n}x{n}, classification every {interval} steps, dilation {radius}")
        h = run((interval, radius))
        d = np.abs(h["rho"] - plain["rho"])
        c = {"interval": interval, "dilation": radius, "value": h["value"], "ms_per_step": h["ms_per_step"],
             "speedup": h["value"] / plain["value"], "steps": h["steps"],
             "flag_rate_final_dilated": h["flag_rate_final_dilated"],
             "field_rate_final_dilated": h["field_rate_final_dilated"], "fallbacks": h["fallbacks"],
             "rho_delta_vs_plain": {"L1": float(d.mean()), "Linf": float(d.max())}}
        if rt is not None:
            eh = np.abs(h["rho"] - rt)
            c["error_vs_truth"] = {"L1": float(eh.mean()), "Linf": float(eh.max())}
            c["L1_ratio"] = float(eh.mean() / out["plain_error_vs_truth"]["L1"])
        out["configs"].append(c)
    ok = [c for c in out["configs"] if c.get("L1_ratio", 1.0) <= 1.05]
    best = max(ok or out["configs"], key=lambda c: c["speedup"])
    out["speedup"] = best["speedup"]
    out["speedup_config"] = {"interval": best["interval"], "dilation": best["dilation"]}
    return out


def hybrid_breakdown(W, n, seed, t_end, fp, interval, radius):
    """Per-kernel time per step (HIP events on the solver's stream, summed
    over the whole run) of plain WENO-Z and of the learned switch with one
    (interval, radius), on the hybrid comparison's problem, plus the switch's
    flag-rate trajectory (faces with any field flagged, and fields, every 256
    steps). The events separate the launches, so the sums run a little above
    the un-instrumented step time."""
    P, thr = W.load_switch()
    keys = dict(TESTCASE=5, SCHEME=1, RKMETH=2, CFL=0.6, ELEMNUMX=n, ELEMNUMY=n, QUAD_SEED=seed,
                TEND="%.17g" % t_end, FP_CONTRACT=FP_MODES[fp])
    out = {"workload": f"seeded four-quadrant problem (seed {seed}) {n}x{n}, WENO-Z RK3 CFL 0.6, t_end {t_end}",
           "interval": interval, "dilation": radius}
    with W.Solver(**keys) as s:
        s.set_timing(True)
        steps = s.advance(0)
        s.synchronize()
        kt = s.get_timing()
        out["plain"] = {"steps": steps, "us_per_step": {k: v[0] / steps * 1e3 for k, v in kt.items() if v[1]}}
    with W.Solver(**keys) as s:
        s.switch_set(W.SWITCH_MLP, P, thr)
        s.switch_interval(interval)
        s.switch_dilation(radius)
        s.set_timing(True)
        total, traj = 0, []
        while True:
            k = s.advance(256)
            total += k
            fx, fy = s.switch_flags()
            m = np.concatenate([fx.ravel(), fy.ravel()])
            traj.append([total, float((m != 0).mean()), float(np.unpackbits(m[:, None], axis=1)[:, 4:].mean())])
            if k < 256:
                break
        s.synchronize()
        kt = s.get_timing()
        out["hybrid"] = {"steps": total, "us_per_step": {k: v[0] / total * 1e3 for k, v in kt.items() if v[1]},
                         "launches": {k: v[1] for k, v in kt.items() if v[1]},
                         "fallbacks": s.switch_fallbacks(), "flag_trajectory": traj}
    for side in ("plain", "hybrid"):
        out[side]["us_per_step_total"] = sum(out[side]["us_per_step"].values())
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--scheme", default="Z", choices=list(SCHEMES))
    ap.add_argument("--cfl", type=float, default=0.6)
    ap.add_argument("--case", default="vortex", choices=["vortex", "quadrant"],
                    help="isentropic vortex (the headline) or a seeded four-quadrant problem")
    ap.add_argument("--quad-seed", type=int, default=0)
    ap.add_argument("--fp", default="fused", choices=list(FP_MODES),
                    help="FP64 arithmetic of the sweeps: fused (FMA contraction) or exact (none)")
    ap.add_argument("--other-steps", type=int, default=100,
                    help="steps timed in the other FP mode for comparison (0: skip)")
    ap.add_argument("--cpu-steps", type=int, default=2, help="reference CPU steps for cpu_baseline (0: skip)")
    ap.add_argument("--cpu-ranks", default="8,16",
                    help="MPI rank counts of the reference's parallel CPU runs (capped at the CPU share of one GPU)")
    ap.add_argument("--graph-steps", type=int, default=16,
                    help="steps per graph replay of the device loops (0: every step launched from the host)")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="only run a plain device copy of this many bytes (HBM counter calibration) and exit")
    ap.add_argument("--batch-steps", type=int, default=200, help="steps of the 64-solve batch (0: skip)")
    ap.add_argument("--seq-steps", type=int, default=20, help="steps per solve of the sequential comparison")
    ap.add_argument("--hybrid-size", type=int, default=2048, help="grid of the learned-switch comparison (0: skip)")
    ap.add_argument("--hybrid-seed", type=int, default=1000)
    ap.add_argument("--hybrid-tend", type=float, default=0.25)
    ap.add_argument("--no-hybrid-truth", dest="hybrid_truth", action="store_false",
                    help="skip the comparison with plain WENO-Z on the nested (2N-1)^2 grid")
    ap.add_argument("--hybrid-breakdown", default="",
                    help="INTERVAL:RADIUS: only print the per-kernel time per step of plain WENO-Z and of the "
                         "learned switch over the hybrid comparison's whole run, with the flag-rate trajectory")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this solver runs one whole domain on one GPU; --gpus must be 1", file=sys.stderr)
        return 2

    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
        return 0

    import wenofd as W

    if args.hybrid_breakdown:
        iv, rad = (int(v) for v in args.hybrid_breakdown.split(":"))
        print(json.dumps(hybrid_breakdown(W, args.hybrid_size, args.hybrid_seed, args.hybrid_tend, args.fp, iv, rad)),
              flush=True)
        return 0

    n, sc = args.size, SCHEMES[args.scheme]
    if args.graph_steps != 16:
        W.DEFAULT_GRAPH_STEPS = args.graph_steps

    def make_solver(fp):
        keys = dict(SCHEME=sc, RKMETH=2, CFL=args.cfl, ELEMNUMX=n, ELEMNUMY=n, TEND=1e9, FP_CONTRACT=FP_MODES[fp])
        if args.case == "vortex":
            return W.Solver(TESTCASE=4, **keys)
        return W.Solver(TESTCASE=5, QUAD_SEED=args.quad_seed, **keys)

    def timed_run(s, warmup, steps):
        s.advance(warmup)
        s.synchronize()
        t0 = time.perf_counter()
        done = s.advance(steps)  # the device loop: dt folded and applied on the GPU, graph replays
        s.synchronize()
        el = time.perf_counter() - t0
        assert done == steps
        return el

    modes = {}
    other = [m for m in FP_MODES if m != args.fp][0]
    if args.other_steps > 0:
        log(f"{args.case} {n}x{n} WENO-{args.scheme} RK3 CFL {args.cfl}, {other} FP64: {args.other_steps} steps")
        with make_solver(other) as so:
            el_o = timed_run(so, args.warmup, args.other_steps)
        modes[other] = {"value": n * n * args.other_steps / el_o, "ms_per_step": el_o * 1e3 / args.other_steps,
                        "steps": args.other_steps}
    log(f"{args.case} {n}x{n} WENO-{args.scheme} RK3 CFL {args.cfl}, {args.fp} FP64: warmup {args.warmup}, "
        f"timed {args.steps} steps")
    s = make_solver(args.fp)
    el = timed_run(s, args.warmup, args.steps)
    modes[args.fp] = {"value": n * n * args.steps / el, "ms_per_step": el * 1e3 / args.steps, "steps": args.steps}
    # per-kernel durations: HIP events around each launch of host-driven steps
    # (the same tile functions the device loop runs), outside the timed region
    ksteps = min(args.steps, 50)
    s.set_timing(True)
    s.advance(ksteps)
    s.synchronize()
    kt = s.get_timing()
    s.set_timing(False)
    log(f"{el * 1e3 / args.steps:.3f} ms/step; kernels over {ksteps} host-driven steps: " +
        ", ".join(f"{k} {v[0] / max(v[1], 1) * 1e3:.1f} us x{v[1]}" for k, v in kt.items()))

    counts = flop_counts()
    kernels = {}
    for k, (ms, cnt) in kt.items():
        if cnt == 0:
            continue
        flops, nbytes = kernel_algorithmic(k, counts, sc, n, n)
        avg_s = ms / cnt / 1e3
        kernels[k] = {"avg_us": avg_s * 1e6, "launches": cnt, "flops_per_launch": flops,
                      "bytes_per_launch": nbytes, "tflops": flops / avg_s / 1e12, "gbs": nbytes / avg_s / 1e9}
    dom = max(kernels, key=lambda k: kt[k][0])
    dk = kernels[dom]
    value = n * n * args.steps / el
    step_flops = sum(kernels[k]["flops_per_launch"] * kernels[k]["launches"] for k in kernels) / ksteps
    rec = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (initial condition generated on the host)",
        "config": {"workload": (f"isentropic vortex {n}x{n}" if args.case == "vortex" else
                                f"seeded four-quadrant problem (seed {args.quad_seed}) {n}x{n}") +
                               f", WENO-{args.scheme}, RK3, CFL {args.cfl}, dt from calculateDeltaT every step",
                   "nx": n, "ny": n, "scheme": args.scheme, "rk": "RK3", "cfl": args.cfl, "case": args.case,
                   "fp": args.fp + (" (FP_CONTRACT = 1: FMA contraction, as the reference's -O3 -march=native "
                                    "build, division by the reciprocal, WENO weights over a common "
                                    "denominator)" if args.fp == "fused" else
                                    " (FP_CONTRACT = 0: no contraction)")},
        "fp_modes": modes,
        "roofline": {
            "kernel": dom,
            # the binding roof is the FP64 vector ALU: no kernel of the plain
            # path issues a matrix instruction, and at ~24 FLOP per byte of
            # compulsory traffic the step sits above the HBM ridge point
            "bound": "fp64-valu",
            "compute": "FP64 VALU, 78.6 TFLOP/s on MI355X (FMA counted as 2 FLOP)",
            "achieved": dk["tflops"],
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": dk["tflops"] / FP64_PEAK_TFLOPS,
            "traffic": measured_traffic(dom, n, args.fp) if (args.scheme == "Z" and args.case == "vortex") else None,
            # step level: the compulsory 256 B per cell-update (RK3: read Un,
            # U1, U2 and write U1, U2, Un+1 once each) over the step time
            "hbm": {"achieved": STEP_BYTES_PER_CELL_UPDATE * n * n / (el / args.steps) / 1e9, "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "bytes_per_cell_update": STEP_BYTES_PER_CELL_UPDATE,
                    "frac": STEP_BYTES_PER_CELL_UPDATE * n * n / (el / args.steps) / 1e9 / HBM_PEAK_GBS},
            # the dominant kernel's own design bytes (its reads and writes as
            # written, per launch) over its duration
            "kernel_design_bytes": {"bytes_per_launch": dk["bytes_per_launch"], "achieved": dk["gbs"],
                                    "unit": "GB/s", "frac": dk["gbs"] / HBM_PEAK_GBS},
            "valu_issue": (valu_issue(dom, n, dk["avg_us"], args.fp)
                           if (args.scheme == "Z" and args.case == "vortex") else None),
            "step_fp64_tflops": step_flops / (el / args.steps) / 1e12,
            "kernels": kernels,
        },
    }
    if args.batch_steps > 0:
        rec["batched"] = batched(W, args.batch_steps, args.seq_steps, args.fp)
    if args.hybrid_size > 0:
        rec["hybrid"] = hybrid(W, args.hybrid_size, args.hybrid_seed, args.hybrid_tend, args.hybrid_truth, args.fp)
    if args.cpu_steps > 0:
        ranks = [int(r) for r in args.cpu_ranks.split(",") if r]
        log(f"cpu baseline: {args.cpu_steps} reference steps at {n}x{n}, serial and on {ranks} MPI ranks")
        rec["cpu_baseline"] = cpu_baseline(n, n, sc, args.cfl, args.cpu_steps, ranks)
    else:
        rec["cpu_baseline"] = None
    print(json.dumps(rec), flush=True)
    s.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
