"""Benchmark: cell-RK-stage updates/s of the device path on one MI355X, in
the FP32-mixed precision form with the FP64 path timed beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--config batch|ramp|channel|cylinder|airfoil]
                    [--precision both|fp32|fp64] [--no-autotune]

A "step" is one SSP-RK3 iteration (3 residual evaluations + stage updates)
over every active cell, with all inputs resident in HBM.  Prints one JSON
line; `value` = (sum over instances of cells * 3 * steps) / wall seconds of
the timed run of the FP32-mixed path (the library call returns after its
stream has drained).

--config batch (default): 256 independent channel/ramp instances (about 11M
  cells, Euler, second order, SQUEEZE, LOCAL dt) in one batch, run to a 1e-6
  density-residual drop or K steps, converged instances retired.  Beside the
  FP32-mixed batch: the FP64 batch, and 256 sequential FP64 runs of the plain
  single-instance path (each instance timed alone over a few steps, priced at
  the steps its FP64 run took); per-instance FP32-mixed errors against FP64.
  Then, untimed for `value`, the same instances at first order run to the drop
  in both precisions (at most --conv-steps steps, stall retirement, the
  FP32-mixed run's stalled instances finished in FP64): the errors at
  convergence, which the second-order SQUEEZE batch does not reach.
--config ramp|channel|cylinder|airfoil: one mesh (the 512x256 ramp, the
  64x32 bump channel, the 1024x512 NS O-grid, the 7.8M-cell Delaunay-dual
  airfoil), FP32-mixed and FP64, K steps each (the ramp: to a 1e-8 density
  residual drop, at most K = 20000), with the density residual level reached,
  the FP64 sensitivity floor and (wall cases) wall pressure and force
  coefficients, including the mean lift and drag over the second half of the
  run.

Before the timed runs a launch-shape autotuner times a few steps of each
candidate (cells per block, XCD block order) per precision and keeps the
fastest; every candidate gives the same bits.  --no-autotune uses
--block-size / --xcd-remap (default: the library's automatic choice).

`roofline` prices the dominant kernel's algorithmic bytes per launch against
its average launch time measured with HIP events on the library's stream;
`traffic` is measured with rocprofv3 counters in child runs.  `cpu_baseline`
is the reference solver built from its own sources (or, where that build is
absent, the C oracle) timed on one host core over a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cfd-2d-ns-fv_amd"))

import numpy as np  # noqa: E402

from fv2d import inputs  # noqa: E402
from fv2d import workloads as W  # noqa: E402
from fv2d.lib import (FP32_MIXED, FP64, RETIRED_DROP, RETIRED_NONFINITE,  # noqa: E402
                      RETIRED_STALL, Batch, Solver, finish_stalled)

METRIC = "cell-RK-stage updates/s, FP32-mixed vs FP64 at stated error; % HBM peak"
HBM_PEAK_GBS = 8000.0
PRECISIONS = {"fp64": FP64, "fp32": FP32_MIXED}


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(case, nc: int, nf: int, s: int = 8, g: int = 8, i: int = 4):
    """Compulsory HBM bytes per launch of each kernel (minimal-traffic model:
    every array read or written once per pass, neighbour re-reads free,
    derived geometry recomputed from centroids, face centres, normals,
    lengths, volumes and LSQ inverses).  s = state bytes, g = geometry bytes,
    i = index bytes."""
    if case.poly == 0:
        # one pass: Q (4s) + Q_n (8s/3) + dt (s) + write Q (4s) + sigma (s/3) + V (g)
        b = nc * (12 * s + g) + nf * (2 * i + 2 * g)
        return {"pass_b": b}
    a = nc * (24 * s + 5 * g) + nf * (2 * i + 2 * g)
    if case.limiter != "SQUEEZE":
        a = nc * (12 * s + 5 * g) + nf * (2 * i + 2 * g)
    b = nc * (20 * s + 3 * g) + nf * (2 * i + 4 * g)
    return {"pass_a": a, "pass_b": b}


def max_faces(grid) -> int:
    deg = np.bincount(grid.face_c1, minlength=grid.num_cells)
    deg += np.bincount(grid.face_c2[grid.face_c2 >= 0], minlength=grid.num_cells)
    return int(deg.max())


def cpu_baseline(samples, target_s: float = 15.0, all_cores: bool = False):
    """Time the reference solver (oracle/_ref/ref_harness) or the C oracle on
    one core over a bounded number of steps of each (grid, case) sample.  The
    reference stores at most 6 faces per cell (readmesh.cpp:38), so meshes
    with more go to the C oracle."""
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if any(max_faces(g) > 6 for g, _ in samples):
        harness = ""
    per = target_s / len(samples)
    work = 0.0
    secs = 0.0
    parts = []
    runs = []
    for grid, case in samples:
        nc = grid.num_cells
        if harness and os.path.exists(harness):
            d = tempfile.mkdtemp()
            grid.write(os.path.join(d, "grid.in"))
            c = inputs.Case(**{**case.__dict__, "grid_file": "grid.in"})
            c.write(os.path.join(d, "input.in"))

            def run(n):
                r = subprocess.run(["taskset", "-c", "0", harness, "input.in", "out.bin", str(n),
                                    "--dump", "0", "--time"], cwd=d, capture_output=True,
                                   text=True, check=True)
                return float(r.stdout.strip().splitlines()[-1].split()[1])
            kind = "reference"
        else:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle as O
            o = O.Oracle(grid, case)

            def run(n):
                t0 = time.perf_counter()
                o.step(n)
                return time.perf_counter() - t0
            kind = "port"
        t1 = run(2)
        n = max(2, min(100000, int(per / max(t1 / 2, 1e-6))))
        t = run(n)
        work += nc * 3 * n
        secs += t
        parts.append(f"{nc} cells x {n} steps")
        if kind == "reference":
            runs.append((d, nc, n, t))
    what = ("reference solver built from its own sources" if kind == "reference"
            else "C oracle (serial restatement)")
    out = {"value": work / secs, "unit": "cell-stage/s", "cores": 1, "kind": kind,
           "sample": f"{what}, {'; '.join(parts)} of the same workload ({secs:.2f} s of step loop)"}
    if all_cores and runs:
        out["all_cores"] = cpu_all_cores(harness, runs)
    import shutil
    for d, _, _, _ in runs:
        shutil.rmtree(d, ignore_errors=True)
    return out


def cpu_all_cores(harness, runs, max_procs: int = 16):
    """The same samples as concurrent single-core reference processes, one
    per core of the host's share (independent instances need no more): P
    processes, process j pinned to core j and running sample j mod len(runs)
    for that sample's step count; value = all their cell-stages over the
    group's wall time."""
    import shutil
    cores = sorted(os.sched_getaffinity(0))[:max_procs]
    procs, work = [], 0.0
    dirs = []
    for j, core in enumerate(cores):
        d, nc, n, _ = runs[j % len(runs)]
        dj = tempfile.mkdtemp()
        for f in ("grid.in", "input.in"):
            shutil.copy(os.path.join(d, f), dj)
        dirs.append(dj)
        work += nc * 3 * n
        procs.append((core, dj, n))
    t0 = time.perf_counter()
    running = [subprocess.Popen(["taskset", "-c", str(core), harness, "input.in", "out.bin",
                                 str(n), "--dump", "0", "--time"], cwd=dj,
                                stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
               for core, dj, n in procs]
    rcs = [p.wait() for p in running]
    wall = time.perf_counter() - t0
    for dj in dirs:
        shutil.rmtree(dj, ignore_errors=True)
    if any(rcs):
        return {"error": f"exit codes {rcs}"}
    return {"value": work / wall, "unit": "cell-stage/s", "cores": len(procs),
            "kind": "reference",
            "how": f"{len(procs)} concurrent single-core reference processes (the samples in "
                   f"turn), {wall:.2f} s wall"}


def copy_calibration(mbytes: int) -> None:
    """One device-to-device copy of a known size through the HIP runtime
    (a plain copy kernel), so the profiler's FETCH_SIZE / WRITE_SIZE per
    dispatch can be scaled against bytes that are known exactly."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mbytes << 20
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(n)) == 0
    assert hip.hipMemset(src, 1, ctypes.c_size_t(n)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(dst, src, ctypes.c_size_t(n), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_calibration_bytes": n, "copies": 3}))


def stream_probe(mbytes: int) -> None:
    """Achievable bandwidth on this GPU, beside the 8 TB/s peak: the HIP
    runtime's device-to-device copy of `mbytes` MiB (read + write bytes /
    time, 20 timed copies)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mbytes << 20
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(n)) == 0
    assert hip.hipMemset(src, 1, ctypes.c_size_t(n)) == 0
    assert hip.hipEventCreate(ctypes.byref(e0)) == 0 and hip.hipEventCreate(ctypes.byref(e1)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(dst, src, ctypes.c_size_t(n), 3) == 0   # hipMemcpyDeviceToDevice
    reps = 20
    assert hip.hipEventRecord(e0, None) == 0
    for _ in range(reps):
        assert hip.hipMemcpyAsync(dst, src, ctypes.c_size_t(n), 3, None) == 0
    assert hip.hipEventRecord(e1, None) == 0 and hip.hipEventSynchronize(e1) == 0
    ms = ctypes.c_float()
    assert hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1) == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    gbs = 2.0 * n * reps / (ms.value * 1e-3) / 1e9
    out = {"stream_probe_mib": mbytes, "copies": reps, "GB_per_s": round(gbs, 1)}
    print(json.dumps(out))


def _pmc_mean_kib(counter: str, child_args, match, timeout: int = 300):
    """Mean value (KiB) of `counter` per dispatch of the kernels whose name
    contains `match` (or satisfies it, if callable), from one
    `rocprofv3 --pmc` run of this script."""
    import csv
    import shutil
    if not shutil.which("rocprofv3"):
        return None
    d = tempfile.mkdtemp(dir="/tmp")
    cmd = ["rocprofv3", "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "run", "--",
           sys.executable, os.path.abspath(__file__)] + child_args
    env = dict(os.environ, TMPDIR="/tmp")
    try:   # the child's log lines pass through to this process's stderr
        subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE, text=True,
                       timeout=timeout, check=True)
    except (subprocess.SubprocessError, OSError):
        return None
    vals = []
    for root, _, files in os.walk(d):
        for fn in files:
            if fn.endswith("counter_collection.csv"):
                for r in csv.DictReader(open(os.path.join(root, fn))):
                    name = r["Kernel_Name"]
                    hit = match(name) if callable(match) else match in name
                    if hit and r["Counter_Name"] == counter:
                        vals.append(float(r["Counter_Value"]))
    return sum(vals) / len(vals) if vals else None


def measure_traffic(kernel: str, child_args, tchar: str):
    """HBM-side bytes per launch of `kernel` (rocprofv3 FETCH_SIZE and
    WRITE_SIZE), each scaled by what the same counter reads for a
    device-to-device copy of known size.  `tchar` is the mangled state type
    ("d" double, "f" float) that tells the two instantiations apart."""
    copy_mb = 256
    tname = {"f": "float", "d": "double"}[tchar]

    def match(name):   # demangled (k_pass_b_lds<float, 2, false>) or mangled (IfLi2...)
        return ("k_" + kernel + "_") in name and (f"<{tname}" in name or f"I{tchar}L" in name)
    log(f"traffic: FETCH_SIZE / WRITE_SIZE of {kernel} ({tname})")
    fetch = _pmc_mean_kib("FETCH_SIZE", child_args, match)
    write = _pmc_mean_kib("WRITE_SIZE", child_args, match)
    cf = _pmc_mean_kib("FETCH_SIZE", ["--copy-calibration", str(copy_mb)], "copyBuffer")
    cw = _pmc_mean_kib("WRITE_SIZE", ["--copy-calibration", str(copy_mb)], "copyBuffer")
    if None in (fetch, write, cf, cw) or cf <= 0 or cw <= 0:
        return None, None
    sf = (copy_mb << 20) / (cf * 1024.0)
    sw = (copy_mb << 20) / (cw * 1024.0)
    traffic = fetch * 1024.0 * sf + write * 1024.0 * sw
    detail = {"fetch_kib": round(fetch, 1), "write_kib": round(write, 1),
              "fetch_scale": round(sf, 4), "write_scale": round(sw, 4),
              "calibration": f"{copy_mb} MiB device-to-device copy"}
    return int(traffic), detail


def errors_vs_fp64(q32, q64, gamma):
    """Relative L2 and Linf errors of density and pressure."""
    def p(q):
        return (q[3] - 0.5 * (q[1] ** 2 + q[2] ** 2) / q[0]) * (gamma - 1)

    def l2(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    def linf(a, b):
        return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))

    return {"rho_l2_rel": l2(q32[0], q64[0]), "rho_linf_rel": linf(q32[0], q64[0]),
            "p_l2_rel": l2(p(q32), p(q64)), "p_linf_rel": linf(p(q32), p(q64))}


PERTURBATION = 1e-7


def perturbed_freestream(grid, case, seed: int = 7, amplitude: float = PERTURBATION):
    """Q_inf * (1 + 1e-7 u), u uniform in [-1, 1]: an FP32-rounding-sized
    perturbation of the freestream start."""
    qinf = np.array([case.rho, case.rho * case.u, case.rho * case.v,
                     case.P / (case.gamma - 1) + 0.5 * case.rho * (case.u ** 2 + case.v ** 2)])
    u = np.random.default_rng(seed).uniform(-1.0, 1.0, (4, grid.num_cells))
    return qinf[:, None] * (1.0 + amplitude * u)


def roofline(per_launch, algo, head, args, child):
    """Roofline object of the dominant kernel (longest average launch)."""
    dom = max(per_launch, key=per_launch.get)
    achieved = algo[dom] / (per_launch[dom] * 1e-6) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": f"{dom} ({head})", "algorithmic_bytes_per_launch": int(algo[dom]),
            "avg_launch_us": round(per_launch[dom], 2),
            "all_kernels_avg_us": {k: round(v, 2) for k, v in per_launch.items()}}
    if not args.no_traffic:
        traffic, detail = measure_traffic(dom, child, "f" if head == "fp32" else "d")
        roof["traffic"] = traffic
        if detail:
            roof["traffic_detail"] = detail
    return roof


# ---------------------------------------------------------------------------
# single mesh
# ---------------------------------------------------------------------------
def run_single(prec: str, args, grid, case, opts, walls=None):
    """Step 1, then the rest of the warm-up, then the timed steps in chunks of
    args.chunk: the clock runs only around the stepping calls.  Between
    chunks, the density residual is checked against its value after step 1
    (stop at args.drop_single when set), and with `walls` (the mesh's host
    geometry for the wall integral; True: take it from this run's solver)
    the lift and drag coefficients are sampled.  Then each kernel kind is
    timed per launch."""
    from fv2d import analysis
    nc = grid.num_cells
    s = Solver(grid, case, kernel_variant=args.variant, precision=PRECISIONS[prec], **opts)
    s.init_freestream()
    if walls is True:
        walls = analysis.wall_geometry(grid, case, s)
    s.step(1)
    r0 = float(s.residual_l2()[0])
    s.step(max(0, args.warmup - 1))
    done, wall, samples, reached = 0, 0.0, [], False
    while done < args.steps and not reached:
        n = min(args.chunk, args.steps - done)
        t0 = time.perf_counter()
        s.step(n)
        wall += time.perf_counter() - t0
        done += n
        r = float(s.residual_l2()[0])
        reached = bool(args.drop_single) and r <= args.drop_single * r0
        if walls:
            q, g = s.state_and_grad()
            _, _, cl, cd = analysis.wall_coefficients(grid, case, q, g, geo=walls)
            samples.append((args.warmup + done, cl, cd))
    conv = {"steps": args.warmup + done, "res0_rho": r0, "res_end_rho": r,
            "drop": r / r0 if r0 > 0 else None, "drop_target": args.drop_single or None,
            "reached": reached}
    per_launch = {k: v for k, v in s.time_kernels(args.kernel_reps).items() if v > 0}
    return {"solver": s, "wall": wall, "value": nc * 3 * done / wall, "steps": done,
            "per_launch": per_launch, "convergence": conv, "forces": samples,
            "walls": walls or None}


def wake_kick(grid, case, solver, amplitude: float = 0.1):
    """Freestream plus a transverse-velocity blob in the near wake,
    amplitude * U * exp(-((x - 1)^2 + y^2) / 0.25), which breaks the mirror
    symmetry so that shedding starts within a few shedding periods."""
    x = solver.geometry("cell_centerx")
    y = solver.geometry("cell_centery")
    U = float(np.hypot(case.u, case.v))
    v = case.v + amplitude * U * np.exp(-((x - 1.0) ** 2 + y ** 2) / 0.25)
    rho = np.full_like(x, case.rho)
    E = case.P / (case.gamma - 1) + 0.5 * rho * (case.u ** 2 + v ** 2)
    return np.stack([rho, rho * case.u, rho * v, E])


def strouhal_run(prec: str, args, grid, case, walls):
    """The cylinder with the uniform (time-accurate) step from a kicked
    start: lift and drag sampled every args.strouhal_every steps, the
    shedding frequency from the lift history (fv2d.analysis.strouhal)."""
    from fv2d import analysis
    c = inputs.Case(**{**case.__dict__, "dt_type": "UNIFORM", "CFL": args.strouhal_cfl})
    s = Solver(grid, c, precision=PRECISIONS[prec])
    s.set_state(wake_kick(grid, c, s))
    t, cl, cd = [], [], []
    done, t0 = 0, time.perf_counter()
    while done < args.strouhal_steps:
        n = min(args.strouhal_every, args.strouhal_steps - done)
        s.step(n)
        done += n
        q, g = s.state_and_grad()
        _, _, l, d = analysis.wall_coefficients(grid, c, q, g, geo=walls)
        t.append(s.time())
        cl.append(l)
        cd.append(d)
        if len(t) % 50 == 0:
            log(f"strouhal {prec}: {done} steps, t = {t[-1]:.4g} s, CL = {l:.4f}")
    U = float(np.hypot(case.u, case.v))
    st = analysis.strouhal(t, cl, case.length, U)
    half = len(cd) // 2
    s.close()
    return {**st, "cd_mean": float(np.mean(cd[half:])), "steps": done,
            "time_s": t[-1] if t else 0.0, "dt_s": t[-1] / done if done else None,
            "flow_times": t[-1] * U / case.length if t else 0.0,
            "wall_s": round(time.perf_counter() - t0, 1), "sample_every": args.strouhal_every,
            "cfl": args.strouhal_cfl}


def max_aspect_ratio(grid, area):
    """Largest ratio of a cell's longest to its shortest face."""
    lo = np.full(grid.num_cells, np.inf)
    hi = np.zeros(grid.num_cells)
    for cc in (grid.face_c1, grid.face_c2):
        m = cc >= 0
        np.minimum.at(lo, cc[m], area[m])
        np.maximum.at(hi, cc[m], area[m])
    return float((hi / lo).max())


def mean_coefficients(samples):
    """Mean lift and drag over the second half of the samples."""
    half = samples[len(samples) // 2:]
    return (float(np.mean([c[1] for c in half])), float(np.mean([c[2] for c in half])),
            len(half), half[0][0] if half else None)


def bench_single(args):
    log(f"building the {args.config} mesh")
    if args.config == "airfoil":
        name, grid, case = W.airfoil(log=log, cache=args.grid_npz)
    else:
        name, grid, case = W.SINGLE[args.config]()
    log(f"{grid.num_cells} cells, {grid.num_faces} faces")
    nc, nf = grid.num_cells, grid.num_faces
    precs = ["fp32", "fp64"] if args.precision == "both" else [args.precision]
    opts = {"block_size": args.block_size, "xcd_remap": args.xcd_remap,
            "accumulate": args.accumulate}
    tuned = {}
    if args.autotune and args.variant in (0, 3, 4):
        from fv2d import autotune
        for prec in precs:
            best, table = autotune.tune_single(
                grid, case, PRECISIONS[prec], log=log, kernel_variant=args.variant,
                extra=autotune.RCM + (autotune.ATOMIC if prec == "fp32" else []))
            tuned[prec] = {"choice": best,
                           "ms_per_step": [{**o, "ms": round(ms, 3)} for o, ms in table]}
    walls = case.BC1 in ("SLIP_WALL", "NO_SLIP_WALL")   # the geometry, from the first run
    runs = {}
    for prec in precs:
        runs[prec] = run_single(prec, args, grid, case, tuned[prec]["choice"] if tuned else opts,
                                walls)
        walls = runs[prec]["walls"]
        c = runs[prec]["convergence"]
        log(f"{prec} {runs[prec]['value']:.4g} cell-stage/s, {c['steps']} steps, "
            f"density residual drop {c['drop']:.3g}")
    head = "fp32" if "fp32" in runs else "fp64"
    r = runs[head]
    algo = algorithmic_bytes(case, nc, nf, s=4 if head == "fp32" else 8)
    chosen = r["solver"].launch_options()
    child = ["--no-cpu", "--no-traffic", "--no-autotune", "--steps", "5", "--warmup", "2",
             "--kernel-reps", "5", "--config", args.config,
             "--variant", str(chosen["kernel_variant"]),
             "--block-size", str(chosen["block_size"]), "--xcd-remap", str(chosen["xcd_remap"]),
             "--accumulate", str(chosen["accumulate"]), "--precision", head]
    tmp = None
    if args.config == "airfoil" and not args.no_traffic:
        # the counter runs load this mesh instead of generating it again
        tmp = tempfile.mkdtemp(dir="/tmp")
        grid.save_npz(os.path.join(tmp, "grid.npz"))
        child += ["--grid-npz", os.path.join(tmp, "grid.npz")]
    roof = roofline(r["per_launch"], algo, head, args, child)
    if tmp:
        import shutil
        shutil.rmtree(tmp, ignore_errors=True)
    q = r["solver"].state()
    cfg = {"workload": name, "cells": nc, "faces": nf,
           "precision": "FP32-mixed (FP32 perturbation state)" if head == "fp32" else "FP64 plain",
           "launch_options": chosen, "autotune": tuned or None,
           "state_finite": bool(np.isfinite(q).all())}
    out = {"metric": METRIC, "value": r["value"], "unit": "cell-stage/s", "n_gpus": 1,
           "steps": r["steps"], "warmup": args.warmup,
           "ms_per_step": r["wall"] * 1e3 / r["steps"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32" if head == "fp32" else "f64",
           "data": "synthetic", "config": cfg, "roofline": roof,
           "convergence": {p: runs[p]["convergence"] for p in runs}}
    if walls:
        out["mean_coefficients"] = {}
        for p in runs:
            cl, cd, n, first = mean_coefficients(runs[p]["forces"])
            out["mean_coefficients"][p] = {"cl": cl, "cd": cd, "samples": n,
                                           "from_step": first, "every": args.chunk}
        if "fp64" in runs and head == "fp32":
            m32, m64 = out["mean_coefficients"]["fp32"], out["mean_coefficients"]["fp64"]
            out["mean_coefficients"]["cd_rel_err"] = abs(m32["cd"] - m64["cd"]) / abs(m64["cd"])
    if args.config == "cylinder":
        out["strouhal"] = None
        out["strouhal_why"] = ("not defined for this scheme: its GLOBAL time step is "
                               "2*CFL*V_i/max(sum), proportional to each cell's volume "
                               "(ssp_rk3.cpp:38-42), so cells advance through different "
                               "physical times; see strouhal_uniform_dt")
        if args.strouhal_steps:
            out["strouhal_uniform_dt"] = {
                "how": ("time_step_type 3 (one dt = 2 CFL / max(Sigma/V), time-accurate) at "
                        f"CFL {args.strouhal_cfl}, from freestream plus a transverse-velocity "
                        "blob in the wake; St from upward zero crossings of CL over the last "
                        "half, FFT peak beside it; mean CD over the last half")}
            for p in runs:
                log(f"strouhal run, {p}")
                out["strouhal_uniform_dt"][p] = strouhal_run(p, args, grid, case, walls)
        if walls:
            out["config"]["max_cell_aspect_ratio"] = max_aspect_ratio(grid, walls["face_area"])
    if "fp64" in runs and head == "fp32":
        r64 = runs["fp64"]
        out["fp64"] = {"value": r64["value"], "ms_per_step": r64["wall"] * 1e3 / r64["steps"],
                       "avg_launch_us": {k: round(v, 2) for k, v in r64["per_launch"].items()}}
        out["speedup_vs_fp64"] = round(r["value"] / r64["value"], 4)
        q64 = r64["solver"].state()
        err = errors_vs_fp64(q, q64, case.gamma)
        out["error_vs_fp64"] = {k: float(f"{v:.3e}") for k, v in err.items()}
        out["error_vs_fp64"]["after_steps"] = {p: runs[p]["convergence"]["steps"] for p in runs}
        # the FP64 scheme's own response to an FP32-rounding-sized perturbation
        log("FP64 sensitivity floor run")
        fl = Solver(grid, case)
        fl.set_state(perturbed_freestream(grid, case, seed=args.floor_seed,
                                          amplitude=args.floor_amplitude))
        fl.step(runs["fp64"]["convergence"]["steps"])
        floor = errors_vs_fp64(fl.state(), q64, case.gamma)
        out["fp64_sensitivity_floor"] = {k: float(f"{v:.3e}") for k, v in floor.items()}
        out["fp64_sensitivity_floor"]["how"] = (f"FP64 from Q_inf(1+{args.floor_amplitude:g} u), "
                                                f"u in [-1,1] (seed {args.floor_seed}), against "
                                                "FP64 from Q_inf, same steps")
        if case.BC1 in ("SLIP_WALL", "NO_SLIP_WALL"):
            # wall pressure and force coefficients (fv2d.analysis), FP32-mixed vs FP64,
            # and the same comparison for the perturbed FP64 run
            from fv2d import analysis
            log("wall coefficients")
            _, g32 = r["solver"].state_and_grad()
            _, g64 = r64["solver"].state_and_grad()
            _, gfl = fl.state_and_grad()
            cmp = analysis.compare_walls(grid, case, q, g32, q64, g64, walls)
            wfl = analysis.compare_walls(grid, case, fl.state(), gfl, q64, g64, walls)
            out["wall_vs_fp64"] = {k: float(f"{v:.4e}") for k, v in cmp.items()}
            out["wall_vs_fp64"]["floor_cd_rel_err"] = float(f"{wfl['cd_rel_err']:.3e}")
            out["wall_vs_fp64"]["floor_cp_linf_err"] = float(f"{wfl['cp_linf_err']:.3e}")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline([(grid, case)])
    return out


# ---------------------------------------------------------------------------
# batch
# ---------------------------------------------------------------------------
def run_batch(prec: str, args, cases, nc, opts):
    """Warm up, restart, then time one run to the residual drop (or
    args.steps steps) with retirement; then time each kernel kind per launch
    over all instances."""
    b = Batch(cases, precision=PRECISIONS[prec], **opts)
    if args.stall_checks:
        b.set_stall(args.stall_checks, args.stall_factor)
    b.init()
    if args.warmup:
        # the warm-up captures the step graph, which the restart keeps
        b.step(args.warmup)
        b.init()
    t0 = time.perf_counter()
    steps, r0, r1 = b.run_to_drop(args.drop, args.steps, args.check_every)
    wall = time.perf_counter() - t0
    why = b.retired()
    work = float((nc * 3 * steps).sum())
    b.set_active(np.ones(len(cases), dtype=np.uint8))
    per_launch = {k: v for k, v in b.time_kernels(args.kernel_reps).items() if v > 0}
    return {"batch": b, "wall": wall, "value": work / wall, "steps": steps, "r0": r0, "r1": r1,
            "why": why, "per_launch": per_launch, "work": work}


def sequential_fp64(cases, steps, reps: int = 20):
    """The sequential FP64 runs of the plain single-instance path: each
    instance's per-step time measured alone (freestream start, 3 untimed
    steps, then `reps` timed steps), priced at the steps its FP64 run took."""
    total_t = 0.0
    total_work = 0.0
    for (grid, case), n in zip(cases, steps):
        s = Solver(grid, case)
        s.init_freestream()
        s.step(3)
        t0 = time.perf_counter()
        s.step(reps)
        per_step = (time.perf_counter() - t0) / reps
        s.close()
        total_t += per_step * int(n)
        total_work += grid.num_cells * 3 * int(n)
    return total_work / total_t, total_t


def _conv_summary(e):
    e = np.asarray(e, dtype=np.float64)
    return {"instances": int(e.size),
            "rho_l2_rel_median": float(f"{np.median(e):.3e}") if e.size else None,
            "rho_l2_rel_max": float(f"{e.max():.3e}") if e.size else None,
            "instances_rho_l2_le_1e-5": int((e <= 1e-5).sum())}


def batch_errors(args, cases, r, r64, fp64_opts):
    """The FP32-mixed batch's errors against FP64: after the timed run, at
    convergence, against an FP64 run converged further (--reference-drop),
    and with the stalled instances finished in FP64 (--finish-fp64)."""
    out = {}
    why32 = r["why"]
    errs, bad32, bad64, conv = [], 0, 0, []
    # instances converged in both precisions: FP64 retired by the residual
    # drop, FP32-mixed by the drop or, with stall retirement, by a stalled
    # residual (its state's rounding stops the residual above the drop)
    conv_why = (RETIRED_DROP, RETIRED_STALL) if args.stall_checks else (RETIRED_DROP,)
    reached = np.isin(why32, conv_why) & (r64["why"] == RETIRED_DROP)
    q64s = [r64["batch"].state(i) for i in range(len(cases))]
    for i, (g, c) in enumerate(cases):
        q32 = r["batch"].state(i)
        f32, f64 = bool(np.isfinite(q32).all()), bool(np.isfinite(q64s[i]).all())
        bad32 += not f32
        bad64 += not f64
        if f32 and f64:
            errs.append(errors_vs_fp64(q32, q64s[i], c.gamma))
            if reached[i]:
                conv.append(errs[-1]["rho_l2_rel"])
    rho = np.array([e["rho_l2_rel"] for e in errs])
    pl2 = np.array([e["p_l2_rel"] for e in errs])
    # the FP64 scheme's own response to an FP32-rounding-sized perturbation
    log("FP64 sensitivity floor batch")
    fb = Batch(cases, **fp64_opts)
    if args.stall_checks:
        fb.set_stall(args.stall_checks, args.stall_factor)
    fb.init([perturbed_freestream(g, c, seed=args.floor_seed + 4 + i,
                                  amplitude=args.floor_amplitude)
             for i, (g, c) in enumerate(cases)])
    fb.run_to_drop(args.drop, args.steps, args.check_every)
    fl_rho = []
    for i, (g, c) in enumerate(cases):
        qf = fb.state(i)
        if np.isfinite(q64s[i]).all() and np.isfinite(qf).all():
            fl_rho.append(errors_vs_fp64(qf, q64s[i], c.gamma)["rho_l2_rel"])
    fb.close()
    fl_rho = np.array(fl_rho)
    out["fp64_sensitivity_floor"] = {
        "rho_l2_rel_median": float(f"{np.median(fl_rho):.3e}"),
        "rho_l2_rel_max": float(f"{fl_rho.max():.3e}"),
        "how": f"FP64 batch from Q_inf(1+{args.floor_amplitude:g} u) against the FP64 "
               "batch from Q_inf"}
    out["error_vs_fp64"] = {
        "rho_l2_rel_median": float(f"{np.median(rho):.3e}"),
        "rho_l2_rel_max": float(f"{rho.max():.3e}"),
        "p_l2_rel_median": float(f"{np.median(pl2):.3e}"),
        "p_l2_rel_max": float(f"{pl2.max():.3e}"),
        "instances_rho_l2_le_1e-5": int((rho <= 1e-5).sum()),
        "instances_compared": len(errs), "nonfinite_fp32": bad32, "nonfinite_fp64": bad64,
        "compared": "each instance's final FP32-mixed state vs its final FP64 state"}
    out["error_at_convergence"] = {
        **_conv_summary(conv),
        "compared": f"instances whose FP64 density residual fell to {args.drop:g} of "
                    "its first-step value and whose FP32-mixed residual did too"
                    + (" or stalled" if args.stall_checks else "")
                    + "; each at its own final step"}

    # the stalled instances continued in FP64 from their FP32-mixed states
    fin = None
    if args.finish_fp64:
        log("FP64 finish of the stalled FP32-mixed instances")
        pb = Batch(cases, **fp64_opts)
        if args.stall_checks:
            pb.set_stall(args.stall_checks, args.stall_factor)
        t0 = time.perf_counter()
        # the timed run's retirement codes (time_kernels re-activated the batch)
        finished, fsteps, _, fwhy = finish_stalled(r["batch"], pb, r["r0"], args.drop,
                                                   args.steps, args.check_every,
                                                   which=why32 == RETIRED_STALL)
        fwall = time.perf_counter() - t0
        nc = np.array([g.num_cells for g, _ in cases], dtype=np.float64)
        fwork = float((nc * 3 * fsteps).sum())
        value = (r["work"] + fwork) / (r["wall"] + fwall)
        why_fin = np.where(finished, fwhy, why32)
        fin = {"batch": pb, "finished": finished, "why": why_fin}
        conv_f = []
        reached_f = (why_fin == RETIRED_DROP) & (r64["why"] == RETIRED_DROP)
        for i, (g, c) in enumerate(cases):
            if reached_f[i]:
                q = pb.state(i) if finished[i] else r["batch"].state(i)
                conv_f.append(errors_vs_fp64(q, q64s[i], c.gamma)["rho_l2_rel"])
        out["with_fp64_finish"] = {
            "value": value, "unit": "cell-stage/s",
            "speedup_vs_fp64_batched": round(value / r64["value"], 4),
            "wall_s": round(r["wall"] + fwall, 3), "finish_wall_s": round(fwall, 3),
            "instances_finished": int(finished.sum()),
            "finish_steps_median_max": ([int(np.median(fsteps[finished])),
                                         int(fsteps[finished].max())]
                                        if finished.any() else [0, 0]),
            "finish_cell_stage_share": round(fwork / (r["work"] + fwork), 4),
            "retired": {"drop": int((why_fin == RETIRED_DROP).sum()),
                        "stall": int((why_fin == RETIRED_STALL).sum()),
                        "nonfinite": int((why_fin == RETIRED_NONFINITE).sum())},
            "error_at_convergence": {
                **_conv_summary(conv_f),
                "compared": f"instances whose FP64 density residual fell to {args.drop:g} of "
                            "its first-step value and whose FP32-mixed run, with the stalled "
                            "instances continued in FP64 from their FP32-mixed states, did "
                            "too; each at its own final step"},
            "how": "FP32-mixed batch with stall retirement, then an FP64 batch continuing "
                   "the stalled instances to the same drop of the same reference norms "
                   "(fv2d.lib.finish_stalled); value = all cell-stages of both runs over "
                   "both runs' wall time (the finish's state hand-over included)"}

    if args.reference_drop:
        # two states that both meet a residual drop still differ along the
        # slowly converging modes by about drop / (their decay rate), so
        # compare the converged FP32-mixed states with an FP64 reference
        # converged much further (untimed)
        log(f"FP64 reference batch to a {args.reference_drop:g} drop")
        rb = Batch(cases, **fp64_opts)
        if args.stall_checks:
            rb.set_stall(args.stall_checks, args.stall_factor)
        rb.init()
        rsteps, _, _ = rb.run_to_drop(args.reference_drop, args.reference_steps,
                                      args.check_every)
        rwhy = rb.retired()
        ok64 = np.isin(rwhy, (RETIRED_DROP, RETIRED_STALL))
        qref = [rb.state(i) if ok64[i] else None for i in range(len(cases))]
        rb.close()

        def versus_reference(why, state):
            ok = np.isin(why, conv_why) & ok64
            err, wref = [], []
            for i, (g, c) in enumerate(cases):
                if ok[i]:
                    err.append(errors_vs_fp64(state(i), qref[i], c.gamma)["rho_l2_rel"])
                    wref.append(rwhy[i])
            err, wref = np.array(err), np.array(wref)
            return {**_conv_summary(err),
                    "converged": int(np.isin(why, conv_why).sum()),
                    "where_fp64_reference_reached_the_drop": _conv_summary(
                        err[wref == RETIRED_DROP] if err.size else err),
                    "where_fp64_reference_stalled": _conv_summary(
                        err[wref == RETIRED_STALL] if err.size else err)}

        out["error_at_convergence_vs_converged_fp64"] = {
            **versus_reference(why32, r["batch"].state),
            "fp64_reference_converged": int(ok64.sum()),
            "fp64_reference_retired": {
                "drop": int((rwhy == RETIRED_DROP).sum()),
                "stall": int((rwhy == RETIRED_STALL).sum())},
            "fp64_reference_steps_median_max": [int(np.median(rsteps)), int(rsteps.max())],
            "compared": "each FP32-mixed state at its retirement (drop"
                        + (" or stall" if args.stall_checks else "")
                        + f") against FP64 run to a {args.reference_drop:g} drop"
                        + (" or its stall" if args.stall_checks else "")}
        # the stopping rule's own share: the FP64 batch at the drop against the
        # same reference
        out["error_at_convergence_vs_converged_fp64"]["fp64_batch_at_the_drop"] = \
            versus_reference(r64["why"], r64["batch"].state)
        if fin is not None:
            pb = fin["batch"]
            out["with_fp64_finish"]["error_at_convergence_vs_converged_fp64"] = {
                **versus_reference(fin["why"], lambda i: (pb.state(i) if fin["finished"][i]
                                                          else r["batch"].state(i))),
                "compared": "each final state (FP32-mixed, or FP64 where the FP32-mixed run "
                            "stalled) against the same FP64 reference"}
    if fin is not None:
        fin["batch"].close()
    return out


def convergence_leg(args, cases, nc, opts32, opts64):
    """The configured instances at first order, run to the residual drop in
    both precisions with stall retirement, the FP32-mixed run's stalled
    instances finished in FP64: errors at convergence, which the second-order
    SQUEEZE batch does not reach (its lagged limiter stalls the residual in
    both precisions).  Runs after the measurement; its figures are reported
    beside it, not in `value`."""
    cases1 = [(g, inputs.Case(**{**c.__dict__, "poly": 0})) for g, c in cases]
    a = argparse.Namespace(**vars(args))
    a.steps, a.check_every, a.warmup, a.kernel_reps = args.conv_steps, 200, 0, 1
    a.stall_checks, a.stall_factor = 20, 0.5
    a.finish_fp64, a.reference_drop = True, 0.0
    log(f"convergence leg: first order, to a {a.drop:g} drop or {a.steps} steps")
    r32 = run_batch("fp32", a, cases1, nc, opts32)
    r64 = run_batch("fp64", a, cases1, nc, opts64)
    e = batch_errors(a, cases1, r32, r64, opts64)
    r32["batch"].close()
    r64["batch"].close()
    fin = e["with_fp64_finish"]
    return {
        "workload": "the same 256 instances at first order (poly 0), Euler, Rusanov, LOCAL dt",
        "why": "the second-order SQUEEZE instances stall above the drop in both precisions",
        "drop": a.drop, "max_steps": a.steps, "check_every": a.check_every,
        "stall_retirement": {"checks": a.stall_checks, "factor": a.stall_factor},
        "fp32_mixed": {"value": r32["value"], "wall_s": round(r32["wall"], 3),
                       "retired": {"drop": int((r32["why"] == RETIRED_DROP).sum()),
                                   "stall": int((r32["why"] == RETIRED_STALL).sum())}},
        "fp64_batched": {"value": r64["value"], "wall_s": round(r64["wall"], 3),
                         "retired": {"drop": int((r64["why"] == RETIRED_DROP).sum()),
                                     "stall": int((r64["why"] == RETIRED_STALL).sum())}},
        "speedup_vs_fp64_batched": round(r32["value"] / r64["value"], 4),
        "error_at_convergence": e["error_at_convergence"],
        "with_fp64_finish": {k: fin[k] for k in ("value", "speedup_vs_fp64_batched",
                                                 "instances_finished", "retired",
                                                 "error_at_convergence")},
        "fp64_sensitivity_floor": e["fp64_sensitivity_floor"]}


def bench_batch(args):
    name, plan = W.batch(args.instances)
    cases = [(i.grid, i.case) for i in plan]
    if args.order or args.limiter:
        # a variant of the batch (the configured workload is o2 SQUEEZE)
        over = {}
        if args.order:
            over["poly"] = args.order - 1
        if args.limiter:
            over["limiter"] = args.limiter
        cases = [(g, inputs.Case(**{**c.__dict__, **over})) for g, c in cases]
        name += " -- variant: " + ", ".join(f"{k}={v}" for k, v in over.items())
    nc = W.cells_of(plan)
    nf = np.array([i.grid.num_faces for i in plan], dtype=np.int64)
    log(f"{name}: {len(cases)} instances, {int(nc.sum())} cells")
    precs = ["fp32", "fp64"] if args.precision == "both" else [args.precision]
    opts = {"block_size": args.block_size, "xcd_remap": args.xcd_remap,
            "accumulate": args.accumulate, "kernel_variant": args.variant}
    tuned = {}
    if args.autotune:
        from fv2d import autotune
        for prec in precs:
            best, table = autotune.tune_batch(
                cases, PRECISIONS[prec], log=log,
                extra=autotune.RCM + (autotune.ATOMIC if prec == "fp32" else []))
            tuned[prec] = {"choice": best,
                           "ms_per_step": [{**o, "ms": round(ms, 3)} for o, ms in table]}
    runs = {}
    for prec in precs:
        runs[prec] = run_batch(prec, args, cases, nc, tuned[prec]["choice"] if tuned else opts)
        r = runs[prec]
        log(f"{prec} batch {r['value']:.4g} cell-stage/s, {r['wall']:.2f} s, "
            f"retired by drop {int((r['why'] == RETIRED_DROP).sum())}, "
            f"by stall {int((r['why'] == RETIRED_STALL).sum())}")
    head = "fp32" if "fp32" in runs else "fp64"
    r = runs[head]
    algo = algorithmic_bytes(cases[0][1], int(nc.sum()), int(nf.sum()),
                             s=4 if head == "fp32" else 8)
    chosen = r["batch"].launch_options()
    child = ["--no-cpu", "--no-traffic", "--no-seq", "--no-autotune", "--steps", "4",
             "--warmup", "0", "--kernel-reps", "3", "--config", "batch",
             "--instances", str(args.instances), "--block-size", str(chosen["block_size"]),
             "--xcd-remap", str(chosen["xcd_remap"]), "--accumulate", str(chosen["accumulate"]),
             "--variant", str(chosen["kernel_variant"]),
             "--precision", head, "--order", str(args.order)] + \
        (["--limiter", args.limiter] if args.limiter else [])
    roof = roofline(r["per_launch"], algo, head, args, child)
    steps = r["steps"]
    cfg = {"workload": name, "instances": len(cases), "cells": int(nc.sum()),
           "faces": int(nf.sum()),
           "precision": "FP32-mixed (FP32 perturbation state)" if head == "fp32" else "FP64 plain",
           "drop": args.drop, "max_steps": args.steps, "check_every": args.check_every,
           "launch_options": chosen, "autotune": tuned or None,
           "graph": "two-step hipGraph replay",
           "steps_min_median_max": [int(steps.min()), int(np.median(steps)), int(steps.max())],
           "stall_retirement": ({"checks": args.stall_checks, "factor": args.stall_factor}
                                if args.stall_checks else None),
           "retired": {p: {"drop": int((runs[p]["why"] == RETIRED_DROP).sum()),
                           "stall": int((runs[p]["why"] == RETIRED_STALL).sum()),
                           "nonfinite": int((runs[p]["why"] == RETIRED_NONFINITE).sum())}
                       for p in runs}}
    out = {"metric": METRIC, "value": r["value"], "unit": "cell-stage/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": r["wall"] * 1e3 / max(1, int(steps.max())), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32" if head == "fp32" else "f64",
           "data": "synthetic", "config": cfg, "roofline": roof}
    if "fp64" in runs:
        r64 = runs["fp64"]
        out["fp64_batched"] = {"value": r64["value"], "wall_s": round(r64["wall"], 3),
                               "avg_launch_us": {k: round(v, 2) for k, v in r64["per_launch"].items()}}
        if not args.no_seq:
            log("sequential FP64 runs of the plain path")
            seq, seq_t = sequential_fp64(cases, r64["steps"])
            out["fp64_sequential"] = {"value": seq, "modelled_wall_s": round(seq_t, 3),
                                      "how": "each instance alone through fvk_ctx (20 timed "
                                             "steps after 3), times its FP64 step count"}
            if head == "fp32":
                out["speedup_vs_fp64_sequential"] = round(r["value"] / seq, 4)
        if head == "fp32":
            out["speedup_vs_fp64_batched"] = round(r["value"] / r64["value"], 4)
            fp64_opts = tuned["fp64"]["choice"] if tuned else opts
            out.update(batch_errors(args, cases, r, r64, fp64_opts))
            if args.conv_steps and not (args.order or args.limiter):
                out["convergence_first_order"] = convergence_leg(
                    args, cases, nc, tuned["fp32"]["choice"] if tuned else opts, fp64_opts)
    if not args.no_cpu:
        # a bounded sample: the smallest, the median and the largest instance
        order = np.argsort(nc)
        pick = [order[0], order[len(order) // 2], order[-1]]
        out["cpu_baseline"] = cpu_baseline([cases[i] for i in pick], all_cores=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (single mesh; default 2000) or the step cap of the "
                         "run to the residual drop (batch; default 1000)")
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--config", default="batch", choices=["batch", *W.SINGLE])
    ap.add_argument("--precision", choices=["both", "fp32", "fp64"], default="both",
                    help="both: FP32-mixed is the measured value, FP64 is timed beside it and "
                         "is the oracle of the stated error")
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--drop", type=float, default=1e-6)
    ap.add_argument("--order", type=int, default=0, choices=[0, 1, 2],
                    help="batch: a variant at this order (0: the configured order 2)")
    ap.add_argument("--limiter", default="", choices=["", "NONE", "SQUEEZE"],
                    help="batch: a variant with this limiter (default: the configured SQUEEZE)")
    ap.add_argument("--check-every", type=int, default=100)
    ap.add_argument("--stall-checks", type=int, default=0,
                    help="batch: also retire an instance after this many checks in a row "
                         "without its density residual falling below --stall-factor times "
                         "its lowest counted value (0: off)")
    ap.add_argument("--stall-factor", type=float, default=0.5)
    ap.add_argument("--floor-amplitude", type=float, default=PERTURBATION,
                    help="relative size of the FP64 sensitivity-floor run's start perturbation")
    ap.add_argument("--floor-seed", type=int, default=7,
                    help="seed of the floor run's perturbation (batch: instance i uses seed + 4 + i)")
    ap.add_argument("--finish-fp64", action="store_true",
                    help="batch: continue the instances the FP32-mixed run retired by a stall "
                         "in FP64 from their states, to the same drop (timed; reported "
                         "beside the FP32-mixed run)")
    ap.add_argument("--reference-drop", type=float, default=0.0,
                    help="batch: also compare the converged FP32-mixed states with an FP64 "
                         "batch run to this residual drop (untimed; 0: off)")
    ap.add_argument("--reference-steps", type=int, default=40000,
                    help="batch: step cap of the --reference-drop run")
    ap.add_argument("--conv-steps", type=int, default=20000,
                    help="batch: step cap of the convergence leg (the instances at first "
                         "order, run to the drop in both precisions with stall retirement "
                         "and the FP64 finish); 0: skip it")
    ap.add_argument("--chunk", type=int, default=None,
                    help="single mesh: steps between residual checks and force samples")
    ap.add_argument("--drop-single", type=float, default=None,
                    help="single mesh: stop at this density-residual drop (default: ramp "
                         "1e-8, else none)")
    ap.add_argument("--no-seq", action="store_true", help="batch: skip the sequential FP64 runs")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel-reps", type=int, default=None,
                    help="back-to-back launches per kernel for the per-launch timing")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 FETCH_SIZE / WRITE_SIZE passes")
    ap.add_argument("--variant", type=int, default=0,
                    help="0 default, 1 cell-parallel, 2 slot-parallel (single mesh), 3 LDS-staged, "
                         "4 LDS-staged on a persistent grid")
    ap.add_argument("--block-size", type=int, default=0,
                    help="cells per mesh block (0: the library's automatic choice)")
    ap.add_argument("--xcd-remap", type=int, default=-1,
                    help="XCD block order: -1 automatic, 0 off, 1 contiguous ranges, "
                         "G>1 runs of G blocks")
    ap.add_argument("--accumulate", type=int, default=0,
                    help="with --no-autotune: 0 owner gather (bit-exact), 1 atomic scatter")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="skip the launch-shape autotuner (use --block-size/--xcd-remap)")
    ap.add_argument("--strouhal-steps", type=int, default=0,
                    help="cylinder: also run this many steps with the uniform (time-accurate) "
                         "time step and report the shedding frequency")
    ap.add_argument("--strouhal-cfl", type=float, default=0.1,
                    help="cylinder: CFL of the Strouhal run (the uniform step takes the "
                         "convective spectral sums only; 0.1 keeps the wall cells' viscous "
                         "terms stable, 0.25 diverges)")
    ap.add_argument("--strouhal-every", type=int, default=2000,
                    help="cylinder: steps between lift samples of the Strouhal run")
    ap.add_argument("--grid-npz", default=None,
                    help="airfoil: load the generated mesh from this file (Grid.save_npz) "
                         "instead of generating it")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                    help="only run device-to-device copies of MB MiB (counter calibration)")
    ap.add_argument("--stream-probe", type=int, default=0, metavar="MB",
                    help="only time device-to-device copies of MB MiB (achievable bandwidth)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.stream_probe:
        stream_probe(args.stream_probe)
        return
    if args.gpus != 1:
        print(f"bench.py: the solver runs on one GPU only; --gpus {args.gpus} is not supported "
              "(instances or mesh partitions would shard across GPUs)",
              file=sys.stderr)
        sys.exit(2)
    batch = args.config == "batch"
    if args.steps is None:
        args.steps = {"batch": 1000, "airfoil": 500, "cylinder": 20000,
                      "ramp": 20000}.get(args.config, 2000)
    if args.chunk is None:
        args.chunk = {"airfoil": 500, "channel": 500}.get(args.config, 1000)
    if args.drop_single is None:
        args.drop_single = 1e-8 if args.config == "ramp" else 0.0
    if args.warmup is None:
        args.warmup = 10 if batch else 50
    if args.kernel_reps is None:
        args.kernel_reps = 20 if batch else 200
    out = bench_batch(args) if batch else bench_single(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
