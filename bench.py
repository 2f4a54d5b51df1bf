"""Benchmark: SIMPLE-TS time steps per second on one MI355X.

Workload: supersonic flow past a confined square in a 2-D channel (channel
15H x H, square of side H/4 at x = 4H on the axis), Ma 2.43 / Kn 0.00283 /
Re 1415, uniform staggered grid 1920 x 384, first-order upwind, every time
step iterated to the 1e-8 successive-difference tolerance of the reference
(loop 2), two fused p/T Jacobi sweeps per outer iteration.  Synthetic input:
the case is generated in the reference's input format (EnteredData.txt +
solve_polyhedrons_b.txt), uniform initial field.  The time step keeps the
Courant number of the 240 x 48 case at dt = 0.01 H/u_in, i.e. dt = 0.01
(240/1920) H/u_in: with dt = 0.01 H/u_in on this grid the reference solver
itself diverges in the first step.

Every path runs the same W + K steps; the first W are not timed, and every
step ends on the reference's loop-2 test (1e-8):
  fp64             the plain FP64 path (sts_gpu, the C++ mirror of the
                   reference's main()), bit-identical to the reference;
  plain            the same FP64 time step (sts_time_step) through the Python
                   binding that runs every path below;
  predicted        FP64 loop 2 started from the degree-2 polynomial in time
                   through the last three converged steps (every cell, the
                   derived fields rebuilt), guarded: an attempt that has not
                   met the test within the recent iteration count of steps
                   from the previous state is restarted from that state
                   (sts_time_step_mixed, extrapolate = 2, guard = 64);
  anderson         FP64 loop 2 from the previous step's state with Anderson
                   acceleration of the outer iterations (history depth 2);
  accelerated      the guarded predicted start of degree 3 with Anderson
                   acceleration (depth 1), all FP64;
  two_phase        an FP32 loop-2 phase until it stagnates, then the FP64
                   phase, from the previous step's state;
  mixed            the guarded degree-3 predicted start, 6 FP32 outer
                   iterations with Anderson acceleration, then the FP64
                   accelerated phase whose FP64 residuals, equation of state
                   and test decide the step;
  surrogate        the method the metric names (surrogate warm start + FP32
                   with FP64 defect correction): the mixed path with the
                   trained corrector network refining the predicted starts
                   (sts_set_corrector, kept only when it lowers the first
                   sweep's residual; after a rejection the network sits out
                   1, 2, 4, ... predicted steps, at most 64), when its
                   parameters are present (simple-ts_amd/unet_corr_b32.npz,
                   written by train_unet.py --corrector);
  surrogate_accelerated  the accelerated path with the same corrector.
value = Nx*Ny*K / timed seconds / 1e6 (Mcell.steps/s) of surrogate (of
accelerated when the network's parameters are absent); every path's figure,
its speed-up over the faster of fp64 and plain, and every field's relative
L2 error against the fp64 path after the W + K steps are reported beside it.

roofline (plain FP64 program, timed steps only): the kernel with the largest
share of the outer iteration's time, the coefficient assembly of the loops
(k_assemble_inl: u and v pseudo-velocities and the energy source), at the
reference-formula bytes of A2 + A3 + A4, 128 + 128 + 40 = 296 B per cell
(12 + 12 + 4 arrays read, 4 + 4 + 1 written, 8 B each, shared arrays counted
once per equation) over all Nx*Ny cells, divided by its average launch
duration from HIP events around every launch in the timed region that did
work (launches that the device-side loop gates out past its stop are not
counted), against the 8 TB/s HBM peak.  Beside it, frac_measured: the bytes
one launch really moves (traffic) over the same duration.  traffic comes from
two rocprofv3 counter passes over a 3-step run of the same case (FETCH_SIZE,
then WRITE_SIZE, median KiB per working launch), each scaled by what the
counter reports for the 5-array copy of the time-step prologue (k_copy5, 5 x
8 x Nx*Ny bytes read and written).  The same figures are given for the fused
p/T sweep pass (k_sweep2: two sweeps that read their 16 arrays once and write
p and T once, 144 B per cell per launch) and for the Jacobi sweeps together
(assembly + loop 3 of one outer iteration).  At 1920 x 384 the state (about
212 MB) fits the 256 MB last-level cache, so the measured bytes are memory
traffic that may be served by that cache rather than by HBM.

The CPU baseline is the reference solver itself (oracle/_ref/sts_ref, serial,
one pinned core) inside the GPU's timed window: the GPU program, bit-identical
to the reference, writes the state after the W warm-up steps; the reference
restarts from those files (ToReadSolvedDataFromFile) for 1 and for 1 + n
steps, and the difference of the two runs times steps W+2 .. W+1+n with setup
excluded.  Outer iterations per step are reported for both sides.
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

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mcell·steps/s to 1e-8 residual, surrogate+FP32 vs plain FP64; % HBM peak"
NX, NY = 1920, 384
RE, KN = 1415.0, 0.00283
DT_OVER_TIN = 0.01 * 240 / NX
HBM_PEAK_GBS = 8000.0
# FP64 reference-formula bytes per cell: assembly A2 + A3 + A4; the fused
# pass of two p/T sweeps reads Fx Fy DTx DTy T_old p_old u v ScT1 rho apx apy
# bpx bpy p_pr T_pr once and writes T p once
ASSEMBLE_BYTES_PER_CELL = 128 + 128 + 40
SWEEP2_BYTES_PER_CELL = 144
CPU_STEPS = 2  # steps of the CPU sample (after one restart step)
WEIGHTS = os.path.join(ROOT, "simple-ts_amd", "unet_corr_b32.npz")
TOL32 = (1e-7, 1e-7, 1e-7)  # FP32-phase tolerances
DEGREE = 2        # predicted start: polynomial degree in time
GUARD = 64        # predicted start: restart from x^n past the budget, pauses up to this many steps
N_I32_MIXED = 6   # mixed: FP32 outer iterations (accelerated) before the FP64 phase
MIX_FP32_ABOVE = 17  # mixed: the FP32 phase runs after steps of more than this many outer iterations
AA_DEPTH = 1      # accelerated: Anderson history depth (with the predicted start)
AA_DEPTH_FULL = 2  # Anderson depth without a predicted start and in the batch (from-rest transients)
AA_DEGREE = 3     # accelerated: predicted-start degree
CORR_PAUSE = 64   # corrector back-off: at most this many predicted steps sat out after a rejection
_ACC = dict(extrapolate=AA_DEGREE, guard=GUARD, anderson=AA_DEPTH)
_MIX = dict(_ACC, n32=N_I32_MIXED, anderson32=AA_DEPTH, fp32_above=MIX_FP32_ABOVE)
PATHS = {  # sts_time_step_mixed settings of each path (n32: FP32 cap; stall: FP32 stagnation window, ratio)
    "plain": None,  # sts_time_step through the same binding as the other paths
    "predicted": dict(extrapolate=DEGREE, guard=GUARD),
    "anderson": dict(anderson=AA_DEPTH_FULL),
    "accelerated": _ACC,
    "two_phase": dict(n32=5000, stall=(6, 0.95)),
    "mixed": _MIX,
    "surrogate": dict(_MIX, net=True, corr_pause=CORR_PAUSE),
    "surrogate_accelerated": dict(_ACC, net=True, corr_pause=CORR_PAUSE),
}
HEADLINE = "surrogate"  # the metric's method; "accelerated" when the network is absent
FALLBACK = "accelerated"


def workload_case(dst, steps, nx=NX, ny=NY, cfl=1.0):
    """The supersonic square on an nx x ny grid at cfl times the Courant
    number of the 240 x 48 case (dt = 0.01 cfl (240/nx) H/u_in)."""
    cases = importlib.import_module("simple-ts_amd.cases")
    return cases.confined_square(dst, nx, ny, RE, KN, steps, dt_over_tin=0.01 * cfl * 240 / nx)


def read_fields(d):
    return {f: np.fromfile(os.path.join(d, f + ".bin"), dtype=np.float64) for f in ("u", "v", "p", "Temper")}


def run_driver(steps, warmup, kernel="assemble_inl,sweep"):
    """The FP64 program; returns its timing record and final fields."""
    pkg = importlib.import_module("simple-ts_amd")
    if not os.access(pkg.DRIVER, os.X_OK):
        raise SystemExit(f"{pkg.DRIVER} is not built; run __graft_entry__.build()")
    work = tempfile.mkdtemp(prefix="sts_bench_")
    try:
        d = workload_case(os.path.join(work, "gpu"), warmup + steps)
        timing = os.path.join(work, "timing.json")
        print(f"[bench] fp64 program: {warmup} + {steps} steps", file=sys.stderr, flush=True)
        subprocess.run([pkg.DRIVER, d, "--warmup", str(warmup), "--time-kernel", kernel, "--timing", timing],
                       check=True, stdout=subprocess.DEVNULL)
        with open(timing) as f:
            t = json.load(f)
        return t, read_fields(d)
    finally:
        shutil.rmtree(work, ignore_errors=True)


def run_python_path(steps, warmup, name, net=None, nx=NX, ny=NY, cfl=1.0):
    """One of PATHS through the Python binding of the C ABI; net: the
    corrector for the paths that use one."""
    spec = dict(PATHS[name] or {})
    n_i32 = spec.pop("n32", 0)
    stall_w, stall_r = spec.pop("stall", (0, 0.95))
    spec.pop("net", None)
    pause = spec.pop("corr_pause", 0)
    ws = importlib.import_module("simple-ts_amd.warmstart")
    work = tempfile.mkdtemp(prefix="sts_bench_")
    try:
        run = ws.Run(workload_case(os.path.join(work, "c"), warmup + steps, nx, ny, cfl),
                     precision=32 if n_i32 else 64)
        if net is not None:
            run.ctx.set_corrector(net, ws.corr_norm(run.case))
            run.ctx.set_corrector_pause(pause)
        n32 = n64 = 0
        t0 = None
        print(f"[bench] {name} path: {warmup} + {steps} steps", file=sys.stderr, flush=True)
        for s in range(warmup + steps):
            if s == warmup:
                t0 = time.perf_counter()
            # steps hand back no residual log, so they return without waiting
            # for their last launches; the last warm-up step and the last timed
            # step wait, so the timed region holds exactly the timed steps
            wait = s in (warmup - 1, warmup + steps - 1)
            if PATHS[name] is None:
                a, b = 0, run.ctx.time_step(*run.loop, wait=wait)[0]
            else:
                a, b, _, _ = run.ctx.time_step_mixed(*run.loop, n_i32, TOL32, stall_w, stall_r, **spec, wait=wait)
            if s >= warmup:
                n32, n64 = n32 + a, n64 + b
        dt = time.perf_counter() - t0
        stats = run.ctx.predict_stats()
        corr = (*run.ctx.corrector_stats(), run.ctx.corrector_skipped()) if net is not None else None
        # the FP64 mass-imbalance check of the final state (S:10621-10646 over the fluid cells)
        mass = run.ctx.continuity()[2]
        run.ctx.jump_overwrite()  # as the program does before it writes the fields
        f = {k: run.ctx.get(k) for k in ("u", "v", "p", "Temper")}
        run.close()
        return dict(seconds=dt, fp32_it=n32, fp64_it=n64, stats=stats, corr=corr, mass=mass.tolist()), f
    finally:
        shutil.rmtree(work, ignore_errors=True)


KERNELS = {"assemble_inl": "k_assemble_inl<double", "sweep": "k_sweep2<double", "copy": "k_copy5"}


def counter_pass(counter, case_dir, out_dir):
    """Median KiB per working launch of each kernel of KERNELS for one counter."""
    import csv
    import statistics
    pkg = importlib.import_module("simple-ts_amd")
    prof = shutil.which("rocprofv3")
    if not prof:
        return None
    d = os.path.join(out_dir, counter)
    cmd = ["timeout", "-s", "KILL", "120", prof, "--pmc", counter, "--kernel-include-regex",
           "k_assemble_inl|k_sweep2|k_copy5", "--output-format", "csv", "-d", d, "-o", "c", "--", pkg.DRIVER,
           case_dir]
    print(f"[bench] counter pass {counter}", file=sys.stderr, flush=True)
    env = dict(os.environ, TMPDIR=out_dir)
    r = subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, cwd=out_dir, env=env)
    path = os.path.join(d, "c_counter_collection.csv")
    if r.returncode != 0 or not os.path.exists(path):
        return None
    by = {k: [] for k in KERNELS}
    for row in csv.DictReader(open(path)):
        for k, pat in KERNELS.items():
            if pat in row["Kernel_Name"]:
                by[k].append(float(row["Counter_Value"]))
    if not all(by.values()):
        return None
    out = {}
    for k, vals in by.items():
        top = max(vals)
        live = [v for v in vals if v > 0.05 * top]  # launches past the loop-2 stop return at once
        out[k] = statistics.median(live)
    return out


def measure_traffic():
    """Memory bytes per launch of k_assemble_inl and k_sweep2: FETCH_SIZE +
    WRITE_SIZE, each scaled by the same counter's reading for k_copy5 (known
    bytes)."""
    work = tempfile.mkdtemp(prefix="sts_pmc_")
    try:
        d = workload_case(os.path.join(work, "c"), 3)
        copy_bytes = 5 * 8 * NX * NY
        total, factors = {"assemble_inl": 0.0, "sweep": 0.0}, {}
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            r = counter_pass(counter, d, work)
            if r is None:
                return None, None
            f = copy_bytes / (r["copy"] * 1024.0)
            factors[counter] = f
            for k in total:
                total[k] += r[k] * 1024.0 * f
        return total, factors
    finally:
        shutil.rmtree(work, ignore_errors=True)


def load_net(nx=NX, ny=NY):
    """The trained corrector for an nx x ny grid, or (None, None)."""
    if not os.path.exists(WEIGHTS):
        return None, None
    pkg = importlib.import_module("simple-ts_amd")
    z = np.load(WEIGHTS, allow_pickle=False)
    net = pkg.UNet(int(z["cin"]), int(z["cout"]), int(z["base"]), 1, ny, nx)
    net.set_params(z["params"])
    return net, {"samples": int(z["samples"]), "train_grid": z["grid"].tolist(),
                 "holdout_loss": float(z["holdout_loss"][-1])}


def _set_entries(case_dir, **values):
    """Rewrites entries of a case's EnteredData.txt in place."""
    cases = importlib.import_module("simple-ts_amd.cases")
    names = cases.entry_names()
    path = os.path.join(case_dir, "EnteredData.txt")
    lines = open(path).read().split("\n")
    for k, name in enumerate(names):
        if name in values:
            lines[k] = cases.fmt(values[name]) + "\t" + lines[k].split("\t", 1)[1]
    open(path, "w").write("\n".join(lines))


def run_cpu_baseline(warmup, n=CPU_STEPS):
    """The reference solver on one pinned core inside the GPU's timed window:
    restarted from the GPU program's state after `warmup` steps, the
    difference of a (1 + n)-step and a 1-step run times steps warmup+2 ..
    warmup+1+n."""
    pkg = importlib.import_module("simple-ts_amd")
    ref = os.path.join(ROOT, "oracle", "_ref", "sts_ref")
    if not os.access(ref, os.X_OK) or not os.access(pkg.DRIVER, os.X_OK):
        return None
    work = tempfile.mkdtemp(prefix="sts_cpu_")
    try:
        base = workload_case(os.path.join(work, "w"), warmup)
        _set_entries(base, Nt_save_solved_data=warmup)
        subprocess.run([pkg.DRIVER, base], check=True, stdout=subprocess.DEVNULL)
        times, iters = {}, {}
        for steps in (1, 1 + n):
            d = os.path.join(work, f"r{steps}")
            shutil.copytree(base, d)
            # the program's saves rewrite max_time_for_solve with its elapsed hours,
            # which would end the restarted run's time loop at once
            _set_entries(d, ToReadSolvedDataFromFile=1, Nt=steps, Nt_save_temp_data=1, max_time_for_solve=1000.0)
            for f in ("Tmp.txt", "Iterations_in_loops.txt"):
                if os.path.exists(os.path.join(d, f)):
                    os.remove(os.path.join(d, f))
            cmd = [ref]
            if shutil.which("taskset"):
                cmd = ["taskset", "-c", "0"] + cmd
            print(f"[bench] cpu reference: restart after step {warmup}, {steps} step(s)", file=sys.stderr,
                  flush=True)
            t0 = time.perf_counter()
            subprocess.run(cmd, cwd=d, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
            times[steps] = time.perf_counter() - t0
            # a restarted run appends its rows without the header line
            rows = open(os.path.join(d, "Iterations_in_loops.txt")).read().split("\n")
            iters[steps] = [int(r.split()[2]) for r in rows if r.strip() and not r.startswith("it")]
        dt = times[1 + n] - times[1]
        its = iters[1 + n][1:]
        if len(its) != n:
            raise RuntimeError(f"the restarted reference ran {len(iters[1 + n])} steps, not {1 + n}")
        return {"value": NX * NY * n / dt / 1e6, "unit": "Mcell·steps/s", "cores": 1, "kind": "reference",
                "outer_iterations_per_step": sum(its) / len(its), "steps": [warmup + 2, warmup + 1 + n],
                "sample": f"reference serial build restarted from the GPU program's (bit-identical) state after step "
                          f"{warmup} of the same {NX}x{NY} case: steps {warmup + 2}-{warmup + 1 + n} "
                          f"({'+'.join(map(str, its))} outer iterations), the {1 + n}-step run {times[1 + n]:.2f} s "
                          f"minus the 1-step run {times[1]:.2f} s (setup excluded), 1 of {os.cpu_count()} host "
                          f"cores"}
    finally:
        shutil.rmtree(work, ignore_errors=True)


def run_batch_workload(n, steps, warmup, mc):
    """The batched-launch workload: n independent cases of uneven size
    (cases.batch_cases) advanced together by sts_batch_time_step after the
    tile autotuner, against the same cases stepped one after another; mc is
    None (plain FP64) or the mixed control tuple of Batch.time_step."""
    pkg = importlib.import_module("simple-ts_amd")
    ws = importlib.import_module("simple-ts_amd.warmstart")
    cases = importlib.import_module("simple-ts_amd.cases")
    work = tempfile.mkdtemp(prefix="sts_batch_")
    mixed = mc is not None
    try:
        res = {}
        for mode in ("batched", "one_by_one"):
            lst = cases.batch_cases(os.path.join(work, mode), n, warmup + steps)
            runs = [ws.Run(d, precision=32 if mixed and mc[0] > 0 else 64) for d, _ in lst]
            cells = sum(r.case.Nx * r.case.Ny for r in runs)
            loop = runs[0].loop
            print(f"[bench] {mode}: {n} cases, {cells} cells, {warmup} + {steps} steps", file=sys.stderr, flush=True)
            B = None
            if mode == "batched":
                B = pkg.Batch([r.ctx for r in runs])
                cv, sv, ms = B.autotune(3)
                res["tiles"] = {"cell_ms": list(ms[:4]), "sweep_ms": list(ms[4:]), "chosen": [cv, sv]}

                def step():
                    B.time_step(*loop, mixed=mc)
            else:
                def step():
                    for r in runs:
                        if mixed:
                            r.ctx.time_step_mixed(*r.loop, *mc[:4], extrapolate=mc[4], fp32_above=mc[5],
                                                  guard=mc[6], anderson=mc[7] if len(mc) > 7 else 0)
                        else:
                            r.ctx.time_step(*r.loop)
            t0 = None
            for s in range(warmup + steps):
                if s == warmup:
                    t0 = time.perf_counter()
                step()
            dt = time.perf_counter() - t0
            res[mode] = cells * steps / dt / 1e6
            res["cells"] = cells
            if B is not None:
                B.close()
            for r in runs:
                r.close()
        res["cases"] = [desc for _, desc in cases.batch_cases(os.path.join(work, "list"), n, 1)]
        return res
    finally:
        shutil.rmtree(work, ignore_errors=True)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def roofline(t, na, measured):
    """Roofline figures of the plain FP64 program's timed steps from its
    per-kernel HIP-event totals; measured = (traffic per launch, counter
    scales) or (None, None)."""
    ks = t.get("kernels", {})
    if not ks.get("assemble_inl", {}).get("launches") or not ks.get("sweep", {}).get("launches"):
        return None
    traffic, factors = measured if measured else (None, None)
    traffic = traffic or {}

    def entry(name, bytes_per_cell, seconds, moved):
        achieved = bytes_per_cell * na / seconds / 1e9
        e = {"achieved": achieved, "frac": achieved / HBM_PEAK_GBS, "bytes_per_launch": bytes_per_cell * na,
             "avg_launch_us": seconds * 1e6, "traffic": moved}
        if moved:
            e["frac_measured"] = moved / seconds / 1e9 / HBM_PEAK_GBS
        return e

    a = ks["assemble_inl"]
    w = ks["sweep"]
    ta = a["ms"] / a["launches"] / 1e3
    tw = w["ms"] / w["launches"] / 1e3
    asm = entry("assemble_inl", ASSEMBLE_BYTES_PER_CELL, ta, traffic.get("assemble_inl"))
    swp = entry("sweep", SWEEP2_BYTES_PER_CELL, tw, traffic.get("sweep"))
    both = None if not traffic else traffic["assemble_inl"] + traffic["sweep"]
    agg = entry("jacobi", ASSEMBLE_BYTES_PER_CELL + SWEEP2_BYTES_PER_CELL, ta + tw, both)
    roof = {"bound": "hbm", "achieved": asm["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": asm["frac"], "traffic": asm["traffic"], "frac_measured": asm.get("frac_measured"),
            "kernel": "k_assemble_inl (FP64 coefficient assembly of the loops: u, v pseudo-velocities and energy "
                      "source, the largest share of the outer iteration's time)",
            "bytes_per_launch": asm["bytes_per_launch"], "avg_launch_us": asm["avg_launch_us"],
            "launches": a["launches"], "counter_scale": factors,
            "k_sweep2": dict(swp, kernel="fused pass of two p/T Jacobi sweeps, distinct arrays read and written "
                                         "once", launches=w["launches"]),
            "jacobi_sweeps": dict(agg, kernel="k_assemble_inl + k_sweep2 per outer iteration (A2 + A3 + A4 + "
                                              "loop 3 at N_I_p_c = 2)"),
            "cache_note": "the 1920x384 FP64 state (about 212 MB) fits the 256 MB last-level cache: traffic is "
                          "counted memory traffic, part of which that cache may serve instead of HBM"}
    return roof


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=490)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--fp64-only", action="store_true", help="run only the FP64 program")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 counter passes")
    ap.add_argument("--write-case", metavar="DIR", help="only write the workload case (warmup + steps steps) to DIR")
    ap.add_argument("--workload", default="square", choices=["square", "batch", "fine", "unet"],
                    help="square: the 1920x384 case (default); batch: uneven cases in shared launches; fine: "
                         "the same case on 3840x768, mixed and accelerated paths against FP64 on that grid; "
                         "unet: inference of the corrector network at 1920x384")
    ap.add_argument("--cases", type=int, default=64)
    args = ap.parse_args()
    if args.write_case:
        print(workload_case(args.write_case, args.warmup + args.steps))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the workload does not shard "
              f"across GPUs in this build)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "fine":
        # at the full Courant number the plain FP64 solver diverges on this grid
        # (a non-finite residual within the first 500 steps); half of it converges
        nx, ny, cfl = 2 * NX, 2 * NY, 0.5
        paths, ref = {}, None
        for name in ("plain", "two_phase", "mixed", "accelerated"):
            try:
                r, f = run_python_path(args.steps, args.warmup, name, None, nx, ny, cfl)
            except Exception as e:  # a path that fails is reported, the others still run
                paths[name] = {"error": str(e)}
                continue
            if ref is None and name == "plain":
                ref = f
            if ref is None:
                paths[name] = {"error": "no FP64 reference fields (the plain path failed)"}
                continue
            paths[name] = {"value": nx * ny * args.steps / r["seconds"] / 1e6,
                           "ms_per_step": 1e3 * r["seconds"] / args.steps,
                           "fp32_outer_iterations_per_step": r["fp32_it"] / args.steps,
                           "fp64_outer_iterations_per_step": r["fp64_it"] / args.steps,
                           "field_rel_l2_vs_fp64": {k: rel_l2(f[k], ref[k]) for k in f},
                           "field_max_abs_vs_fp64": {k: float(np.max(np.abs(f[k] - ref[k]))) for k in f},
                           "mass_imbalance_max_l2_l1": r["mass"]}
        for name in paths:
            if "value" in paths[name] and "value" in paths.get("plain", {}):
                paths[name]["speedup_vs_fp64"] = paths[name]["value"] / paths["plain"]["value"]
        r = paths["mixed"] if "value" in paths.get("mixed", {}) else {"value": None, "ms_per_step": None}
        print(json.dumps({
            "metric": "Mcell·steps/s to 1e-8 residual, mixed precision on the finer grid", "value": r["value"],
            "unit": "Mcell·steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32+f64", "data": "synthetic: confined-square channel in the reference's input format",
            "config": {"workload": f"supersonic confined square, {nx}x{ny}, dt {0.01 * cfl * 240 / nx:g} H/u_in; mixed: "
                                   f"predicted start, {N_I32_MIXED} accelerated FP32 outer iterations after steps "
                                   f"of more than {MIX_FP32_ABOVE} outer iterations, then the accelerated FP64 "
                                   f"phase on the reference's test; errors against the plain FP64 "
                                   f"path on the same grid after warmup + steps steps",
                       "Nx": nx, "Ny": ny},
            "paths": paths}))
        return
    if args.workload == "unet":
        # the corrector's network (9 inputs, 4 outputs, base 32) at the benchmark grid; the
        # time includes the host transfers of sts_unet_infer (a profiler's kernel trace
        # gives the device part)
        pkg = importlib.import_module("simple-ts_amd")
        ws = importlib.import_module("simple-ts_amd.warmstart")
        net = pkg.UNet(ws.CORR_CIN, ws.CORR_COUT, 32, 1, NY, NX)
        net.set_params(ws.he_init(ws.CORR_CIN, ws.CORR_COUT, 32, 0))
        x = np.random.default_rng(0).standard_normal((1, NY, NX, ws.CORR_CIN)).astype(np.float32)
        for _ in range(args.warmup):
            net.infer(x)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            y = net.infer(x)
        dt = (time.perf_counter() - t0) / args.steps
        net.close()
        print(json.dumps({
            "metric": "U-Net inference at 1920x384 (host transfers included)", "value": 1e3 * dt, "unit": "ms",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": False,
            "dtype": "bf16", "data": "synthetic input, He-initialised weights",
            "config": {"workload": "corrector U-Net, 9 inputs, 4 outputs, base 32, depth 4", "Nx": NX, "Ny": NY},
            "finite": bool(np.all(np.isfinite(y)))}))
        return
    if args.workload == "batch":
        out = {"fp64": run_batch_workload(args.cases, args.steps, args.warmup, None),
               "mixed": run_batch_workload(args.cases, args.steps, args.warmup,
                                           (N_I32_MIXED, TOL32, 0, 0.95, AA_DEGREE, 0, GUARD, AA_DEPTH_FULL)),
               "accelerated": run_batch_workload(args.cases, args.steps, args.warmup,
                                                 (0, TOL32, 0, 0.95, AA_DEGREE, 0, GUARD, AA_DEPTH_FULL))}
        r = out["accelerated"]
        print(json.dumps({
            "metric": "Mcell·steps/s to 1e-8 residual, batched uneven cases", "value": r["batched"],
            "unit": "Mcell·steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic: seeded confined-square and Rayleigh-Benard cases in the reference's input format",
            "config": {"workload": f"{args.cases} independent cases, grids 256x64 .. 1024x256, shared launches "
                                   f"with autotuned tiles, guarded degree-{AA_DEGREE} predicted starts and Anderson "
                                   f"acceleration (depth {AA_DEPTH_FULL}) of the FP64 outer iterations",
                       "cells": r["cells"],
                       "speedup_vs_fp64_batched": r["batched"] / out["fp64"]["batched"]},
            "paths": out}))
        return
    na = NX * NY
    t, f64 = run_driver(args.steps, args.warmup)
    fp64_value = na * t["timed_steps"] / t["timed_seconds"] / 1e6
    paths = {"fp64": {"value": fp64_value, "ms_per_step": 1e3 * t["timed_seconds"] / t["timed_steps"],
                      "outer_iterations_per_step": t["timed_outer_iterations"] / t["timed_steps"]}}
    best, best_name = fp64_value, "fp64"
    if not args.fp64_only:
        for name, spec in PATHS.items():
            net = info = None
            if spec and spec.get("net"):
                net, info = load_net()
                if net is None:
                    continue
            r, f = run_python_path(args.steps, args.warmup, name, net)
            v = na * args.steps / r["seconds"] / 1e6
            paths[name] = {"value": v, "ms_per_step": 1e3 * r["seconds"] / args.steps,
                           "fp32_outer_iterations_per_step": r["fp32_it"] / args.steps,
                           "fp64_outer_iterations_per_step": r["fp64_it"] / args.steps,
                           "field_rel_l2_vs_fp64": {k: rel_l2(f[k], f64[k]) for k in f},
                           "mass_imbalance_max_l2_l1": r["mass"]}
            if spec and spec.get("guard"):
                s = r["stats"]
                paths[name]["guard"] = {"predicted_steps": s[0], "restarted": s[1],
                                        "abandoned_fp32_iterations": s[2], "abandoned_fp64_iterations": s[3],
                                        "steps_from_previous_state": s[4]}
            if net is not None:
                paths[name]["corrector"] = {"kept": r["corr"][0], "rejected": r["corr"][1], "skipped": r["corr"][2],
                                            "max_pause": CORR_PAUSE}
                paths[name]["network"] = info
                net.close()
            if name == HEADLINE or (name == FALLBACK and best_name != HEADLINE):
                best, best_name = v, name
    # speed-ups against the faster of the two plain FP64 measurements (the
    # program, and sts_time_step through the binding the other paths use)
    base = max(fp64_value, paths["plain"]["value"]) if "plain" in paths else fp64_value
    for name in paths:
        paths[name]["speedup_vs_fp64"] = paths[name]["value"] / base
    roof = roofline(t, na, None if args.no_traffic else measure_traffic())
    cpu = None if args.no_cpu else run_cpu_baseline(t["warmup_steps"])
    if cpu is not None:
        # the GPU's FP64 program over its whole timed window, per outer iteration, beside the CPU's
        gpu_it = t["timed_outer_iterations"] / t["timed_steps"]
        cpu["gpu_fp64_outer_iterations_per_step"] = gpu_it
        cpu["gpu_fp64_vs_cpu_per_outer_iteration"] = (
            (fp64_value / gpu_it) / (cpu["value"] / cpu["outer_iterations_per_step"]))
    path_desc = {"fp64": "plain FP64",
                 "accelerated": f"start predicted by degree-{AA_DEGREE} extrapolation in time (every cell, derived "
                                f"fields rebuilt), guarded, + Anderson acceleration of the FP64 outer iterations "
                                f"(history depth {AA_DEPTH}, FP32 differences), every step ending on the "
                                f"reference's test",
                 "mixed": f"start predicted by degree-{AA_DEGREE} extrapolation in time, guarded, "
                          f"{N_I32_MIXED} FP32 outer iterations with Anderson acceleration after steps of more "
                          f"than {MIX_FP32_ABOVE} outer iterations, then FP64 outer iterations with Anderson "
                          f"acceleration whose FP64 test ends the step",
                 "surrogate": "the mixed path with the trained corrector network refining the predicted starts "
                              f"(after a rejected refinement it sits out 1, 2, 4, ... predicted steps, at most "
                              f"{CORR_PAUSE})"}
    line = {
        "metric": METRIC, "value": best, "unit": "Mcell·steps/s", "n_gpus": 1, "steps": t["timed_steps"],
        "warmup": t["warmup_steps"], "ms_per_step": paths[best_name]["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": {"fp64": "f64", "accelerated": "f64", "surrogate": "bf16+f32+f64"}.get(best_name, "f32+f64"),
        "data": "synthetic: confined-square channel generated in the reference's input format, uniform start",
        "config": {"workload": f"supersonic confined square, Ma 2.43 Kn {KN} Re {RE:g}, {NX}x{NY}, "
                               f"dt {DT_OVER_TIN:g} H/u_in, each step converged to 1e-8", "Nx": NX, "Ny": NY,
                   "path": path_desc[best_name], "speedup_vs_fp64": paths[best_name]["speedup_vs_fp64"]},
        "paths": paths, "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
