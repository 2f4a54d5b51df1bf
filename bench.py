"""Benchmark: SIMPLE-TS time steps per second on one MI355X.

Workload: supersonic flow past a confined square in a 2-D channel (channel
15H x H, square of side H/4 at x = 4H on the axis), Ma 2.43 / Kn 0.00283 /
Re 1415, uniform staggered grid 1920 x 384, first-order upwind, every time
step iterated to the 1e-8 successive-difference tolerance of the reference
(loop 2), two fused p/T Jacobi sweeps per outer iteration, FP64.  Synthetic
input: the case is generated in the reference's input format
(EnteredData.txt + solve_polyhedrons_b.txt), uniform initial field.  The time
step keeps the Courant number of the 240 x 48 case at dt = 0.01 H/u_in, i.e.
dt = 0.01 (240/1920) H/u_in: with dt = 0.01 H/u_in on this grid the reference
solver itself diverges in the first step.

A "step" is one converged time step of the GPU program (sts_gpu, the C++
mirror of the reference's main()).  The first W steps of the run are not
timed; the next K are.  value = Nx*Ny*K / timed seconds / 1e6 (Mcell.steps/s).

The CPU baseline is the reference solver itself (oracle/_ref/sts_ref, serial,
one pinned core) on the first steps of the same case: the time of step 2
is the difference of a 2-step and a 1-step run (setup excluded), i.e. step 2.

roofline: the fused p/T Jacobi sweep kernel, 144 algorithmic bytes per cell
(16 FP64 arrays read, 2 written) per launch over all Nx*Ny cells, divided by
its average launch duration from HIP events recorded around every launch in
the timed region, against the 8 TB/s HBM peak.
"""
import argparse
import importlib
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mcell·steps/s to 1e-8 residual, surrogate+FP32 vs plain FP64; % HBM peak"
NX, NY = 1920, 384
RE, KN = 1415.0, 0.00283
DT_OVER_TIN = 0.01 * 240 / NX
HBM_PEAK_GBS = 8000.0
SWEEP_BYTES_PER_CELL = 144  # FP64: Fx Fy DTx DTy T_old p_old u v ScT1 rho apx apy bpx bpy p_pr T_pr + T p


def workload_case(dst, steps):
    cases = importlib.import_module("simple-ts_amd.cases")
    return cases.confined_square(dst, NX, NY, RE, KN, steps, dt_over_tin=DT_OVER_TIN)


def run_gpu(steps, warmup, kernel="sweep", args=()):
    pkg = importlib.import_module("simple-ts_amd")
    if not os.access(pkg.DRIVER, os.X_OK):
        raise SystemExit(f"{pkg.DRIVER} is not built; run __graft_entry__.build()")
    work = tempfile.mkdtemp(prefix="sts_bench_")
    try:
        d = workload_case(os.path.join(work, "gpu"), warmup + steps)
        timing = os.path.join(work, "timing.json")
        t0 = time.time()
        subprocess.run([pkg.DRIVER, d, "--warmup", str(warmup), "--time-kernel", kernel, "--timing", timing,
                        *args], check=True)
        wall = time.time() - t0
        with open(timing) as f:
            t = json.load(f)
        t["process_seconds"] = wall
        return t
    finally:
        shutil.rmtree(work, ignore_errors=True)


def run_cpu_baseline():
    """The reference solver on one pinned core: step 2 of the same case."""
    ref = os.path.join(ROOT, "oracle", "_ref", "sts_ref")
    if not os.access(ref, os.X_OK):
        return None
    work = tempfile.mkdtemp(prefix="sts_cpu_")
    try:
        times = {}
        for n in (1, 2):
            d = workload_case(os.path.join(work, f"n{n}"), n)
            cmd = [ref]
            if shutil.which("taskset"):
                cmd = ["taskset", "-c", "0"] + cmd
            t0 = time.perf_counter()
            subprocess.run(cmd, cwd=d, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
            times[n] = time.perf_counter() - t0
        dt = times[2] - times[1]
        return {"value": NX * NY * 1 / dt / 1e6, "unit": "Mcell·steps/s", "cores": 1, "kind": "reference",
                "sample": f"reference serial build, step 2 of the same {NX}x{NY} case from a uniform start "
                          f"(2-step run {times[2]:.2f} s minus 1-step run {times[1]:.2f} s, setup excluded), "
                          f"1 of {os.cpu_count()} host cores"}
    finally:
        shutil.rmtree(work, ignore_errors=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=490)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--mixed-args", default="", help="extra sts_gpu arguments of the mixed run (space separated)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the workload does not shard "
              f"across GPUs in this build)", file=sys.stderr)
        sys.exit(2)
    t = run_gpu(args.steps, args.warmup)
    tm = run_gpu(args.steps, args.warmup, "sweep32", ["--mixed"] + args.mixed_args.split())
    print(json.dumps({"mixed_timing": tm, "fp64_timing": t}), file=sys.stderr)
    na = NX * NY
    value = na * t["timed_steps"] / t["timed_seconds"] / 1e6
    ms_step = 1e3 * t["timed_seconds"] / max(1, t["timed_steps"])
    roof = None
    if t.get("kernel_launches"):
        avg_s = t["kernel_ms"] / t["kernel_launches"] / 1e3
        achieved = SWEEP_BYTES_PER_CELL * na / avg_s / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "k_sweep (fused p/T Jacobi sweep)",
                "avg_launch_us": avg_s * 1e6, "launches": t["kernel_launches"]}
    cpu = None if args.no_cpu else run_cpu_baseline()
    line = {
        "metric": METRIC, "value": value, "unit": "Mcell·steps/s", "n_gpus": 1, "steps": t["timed_steps"],
        "warmup": t["warmup_steps"], "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: confined-square channel generated in the reference's input format, uniform start",
        "config": {"workload": f"supersonic confined square, Ma 2.43 Kn {KN} Re {RE:g}, {NX}x{NY}, "
                               f"dt {DT_OVER_TIN:g} H/u_in, each step converged to 1e-8, FP64", "Nx": NX, "Ny": NY,
                   "outer_iterations_per_step": t["timed_outer_iterations"] / max(1, t["timed_steps"]),
                   "path": "plain FP64"},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
