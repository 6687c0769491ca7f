"""Benchmark of the Hermite-UGKS time step on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c5|c3|c2]
                [--no-cpu] [--no-fp64]
python bench.py --copy-probe                  (calibration copy for counters)
python bench.py --rocprof-summary DIR [DIR..] (summarise rocprofv3 CSV output)

A step is one full reference time step (timestep, interpolation,
evolution, update) over the whole phase space of the workload.  The
default workload is the large 2D Rayleigh-Taylor case (512 x 512 cells,
40 x 40 Gauss-Hermite nodes, Nm = 12, about 419M phase-space points), the
largest single-GPU configuration of the benchmark set.  The headline value
is the mixed-precision path's phase-space updates per second
(cells * nv^2 per step); the plain FP64 path is timed on the same
workload and the mixed-path error against it is reported per field.

Prints one JSON line (rank 0).
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
sys.path[:0] = [os.path.join(ROOT, "hermite-ugks_amd"), os.path.join(ROOT, "tests")]

from hugks import capi, cases  # noqa: E402

METRIC = ("phase-space updates/s (cells\u00b7nv\u00b2) mixed vs FP64 at stated err; "
          "% HBM roofline")
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, GB/s
COPY_PROBE_BYTES = 1 << 30     # bytes read (and written) by one probe copy
FP64_PEAK_TFLOPS = 78.6

WORKLOADS = {
    # large 2D Rayleigh-Taylor: 512x512, GH40, Kn=1e-3, Nm=12
    "c5": dict(nx=512, ny=512, xlength=0.5, ylength=1.0, nv=40, vmax=5.0, Tref=1.0,
               quad_type="gauss_hermite", force_type="hermite_symmetric", Nm=12, cfl=0.5,
               maxtime=0.14, case_id=6, knudsen=1e-3, phi=1.0, TwL=1.0, TwR=1.0),
    # 2D Rayleigh-Taylor 120x120, GH28, Kn=1e-3, Nm=10
    "c3": dict(nx=120, ny=120, xlength=0.5, ylength=1.0, nv=28, vmax=5.0, Tref=1.0,
               quad_type="gauss_hermite", force_type="hermite_symmetric", Nm=10, cfl=0.5,
               maxtime=0.14, case_id=6, knudsen=1e-3, phi=1.0, TwL=1.0, TwR=1.0),
    # 2D force-driven Poiseuille 4x64, GH24, Kn=0.1, phi=0.05, Nm=8
    "c2": dict(nx=4, ny=64, xlength=0.1, ylength=1.0, nv=24, vmax=5.0, Tref=1.0,
               quad_type="gauss_hermite", force_type="hermite_symmetric", Nm=8, cfl=0.8,
               maxtime=200.0, case_id=5, knudsen=0.1, phi=0.05, TwL=1.0, TwR=1.0),
}
VARIANT = {"c5": "cartesian", "c3": "cartesian", "c2": "reference"}
NAMES = {"c5": "rt2d_512x512_gh40_nm12", "c3": "rt2d_120x120_gh28_nm10",
         "c2": "poiseuille2d_4x64_gh24_nm8"}


def algorithmic_bytes(case, kernel):
    """Algorithmic HBM bytes of one launch: each array the kernel must
    touch counted once (N = cells*nv^2, s = bytes per stored element)."""
    d = case.desc
    s = 8 if d.precision == capi.FP64 else 4
    nn = d.nv * d.nv
    N = d.nx * d.ny * nn
    nvf = (d.nx + 1) * d.ny * nn
    nhf = d.nx * (d.ny + 1) * nn
    if kernel == "slope2d":          # read h, b; write 4 slope arrays
        return 6 * s * N
    if kernel == "flux_vface":       # read h, b and 4 slopes; write 2 face arrays
        return 6 * s * N + 2 * s * nvf
    if kernel == "flux_hface":
        return 6 * s * N + 2 * s * nhf
    if kernel == "update2d":         # read h, b, 4 face arrays; write h, b
        return 4 * s * N + 2 * s * (nvf + nhf)
    # interior-face pipeline: both orientations per launch
    xper, yper = d.bc[0] == capi.BC_PERIODIC, d.bc[2] == capi.BC_PERIODIC
    nint = (d.nx + 1 if xper else d.nx - 1) * d.ny + d.nx * (d.ny + 1 if yper else d.ny - 1)
    if kernel == "flux_sum":         # read 6 upwind arrays per face node; 37 sums per face
        return 6 * s * nint * nn + 8 * 37 * nint
    if kernel == "flux_node":        # read 6 upwind arrays, write 2 face arrays per face node
        return 8 * s * nint * nn
    return None


def run_gpu(case, warmup, steps, timing=True):
    import torch
    s = capi.Solver(case)
    if warmup > 0:               # advance(0) would run to the loop exit
        s.advance(warmup)
    torch.cuda.synchronize()
    s.set_timing(timing)
    s.reset_kernel_times()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    info = s.advance(steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    if info.steps != steps:
        raise RuntimeError(f"the timed region ran {info.steps} steps, not {steps}")
    kt = s.kernel_times()
    w = s.macro(with_q=False)
    s.close()
    return dict(seconds=t1 - t0, info=info, kernels=kt, w=w)


def cpu_baseline(wl, params, budget_rows):
    """Reference Fortran solver (oracle/_ref) on a bounded sample of the
    workload: the full row width and velocity grid, budget_rows rows,
    timed loop of a few steps.  Falls back to the C restatement."""
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness2d")
    d = WORKLOADS[wl]
    rows = min(budget_rows, d["ny"])
    steps = 1
    if os.path.exists(exe):
        with tempfile.TemporaryDirectory() as td:
            nml = os.path.join(td, "bench.namelist")
            with open(nml, "w") as f:
                f.write("&params\n")
                for k, v in params.items():
                    f.write(f"  {k} = {v!r}\n" if isinstance(v, str) else f"  {k} = {v}\n")
                f.write("/\n")
            env = dict(os.environ)
            r = subprocess.run([exe, nml, "-", str(steps), "1" if VARIANT[wl] == "cartesian" else "0",
                                "1", str(rows)], capture_output=True, text=True, cwd=td, env=env,
                               timeout=600)
            line = [ln for ln in r.stdout.splitlines() if "HARNESS_TIMING" in ln]
            if r.returncode == 0 and line:
                loop_s = float(line[0].split("loop_s=")[1].split()[0])
                n = d["nx"] * rows * d["nv"] ** 2 * steps
                return {"value": n / loop_s, "unit": "phase-space updates/s", "cores": 8,
                        "kind": "reference",
                        "sample": f"{d['nx']}x{rows} rows of the workload (full nv={d['nv']}, "
                                  f"Nm={d['Nm']}), {steps} step, reference Fortran built with "
                                  f"amdflang -O2 -fopenmp; 8 OpenMP threads (hard-coded in the "
                                  f"reference), update() serial; loop {loop_s:.1f} s"}
    # C restatement (single thread)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py
    p = dict(params, ny=max(2, rows // 4))
    case = cases.build_2d(p, capi.FP64, rt_variant=VARIANT[wl], seed=1)
    o = oracle_py.Oracle(case)
    t0 = time.perf_counter()
    o.step(1)
    dt = time.perf_counter() - t0
    n = case.ncell * case.nn
    return {"value": n / dt, "unit": "phase-space updates/s", "cores": 1, "kind": "port",
            "sample": f"{p['nx']}x{p['ny']} cells of the workload, 1 step, C restatement"}


def copy_probe():
    """A plain device copy of COPY_PROBE_BYTES, three times: the counter
    runs of rocprofv3 over this give the FETCH_SIZE / WRITE_SIZE reading of
    a kernel whose HBM traffic is known, which calibrates the same counters
    read over the solver's kernels."""
    import torch
    n = COPY_PROBE_BYTES // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": COPY_PROBE_BYTES, "copies": 3}))


def _csv_rows(path):
    import csv
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def rocprof_summary(dirs):
    """Per-kernel summary of rocprofv3 CSV output directories (kernel
    trace: calls and average duration; counter collection: average counter
    value per dispatch).  Prints JSON."""
    import glob
    out = {}
    for d in dirs:
        for path in glob.glob(os.path.join(d, "**", "*kernel_trace.csv"), recursive=True):
            for r in _csv_rows(path):
                k = r.get("Kernel_Name", "?")
                e = out.setdefault(k, {})
                dur = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-6
                e["calls"] = e.get("calls", 0) + 1
                e["total_ms"] = e.get("total_ms", 0.0) + dur
        for path in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            acc = {}
            for r in _csv_rows(path):
                key = (r.get("Kernel_Name", "?"), r["Counter_Name"])
                v = acc.setdefault(key, {})
                disp = r.get("Dispatch_Id", r.get("Correlation_Id", "0"))
                v[disp] = v.get(disp, 0.0) + float(r["Counter_Value"])
            for (k, cn), per in acc.items():
                e = out.setdefault(k, {})
                e[cn + "_avg"] = sum(per.values()) / len(per)
                e[cn + "_dispatches"] = len(per)
    for e in out.values():
        if "calls" in e:
            e["avg_ms"] = e["total_ms"] / e["calls"]
    print(json.dumps(out, indent=1, sort_keys=True))


def traffic_from_profiles(workload, kernel):
    """HBM bytes per launch of kernel from the committed counter summary
    profiles/traffic_<workload>.json (FETCH_SIZE + WRITE_SIZE, KiB per
    dispatch, each scaled by its reading over a copy of known size);
    None when no summary exists."""
    path = os.path.join(ROOT, "profiles", f"traffic_{NAMES[workload]}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    k = t.get("kernels", {}).get(kernel)
    if not k:
        return None
    return k.get("hbm_bytes_per_launch")


def c4_params(seed=7, n=64):
    """The batched sweep: n force-driven channel cases (case 5), nx = 4,
    Nm = 8, GH grids; per case Kn log-uniform on [0.01, 1], phi uniform on
    [0.01, 0.2], ny uniform integer on [32, 128], nv from {16, 20, 24, 28,
    32}; every second case is a Fourier-type channel with TwR uniform on
    [1.1, 1.5] (TwL = 1).  Drawn with numpy's PCG64 from a fixed seed."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for k in range(n):
        kn = float(10 ** rng.uniform(-2.0, 0.0))
        phi = float(rng.uniform(0.01, 0.2))
        ny = int(rng.integers(32, 129))
        nv = int(rng.choice([16, 20, 24, 28, 32]))
        twr = float(rng.uniform(1.1, 1.5)) if k % 2 else 1.0
        out.append(dict(nx=4, ny=ny, xlength=0.1, ylength=1.0, nv=nv, vmax=5.0, Tref=1.0,
                        quad_type="gauss_hermite", force_type="hermite_symmetric", Nm=8,
                        cfl=0.8, maxtime=200.0, case_id=5, knudsen=kn, phi=phi, TwL=1.0,
                        TwR=twr))
    return out


def bench_c4(a):
    """Aggregate updates/s of the batched sweep: all cases advanced by one
    graph replay per step (grouped launches for the small cases, tuned
    schedules for the others), against the same cases stepped one after the
    other, and the mixed path against FP64 per case."""
    import torch
    from metrics import field_errors
    plist = c4_params()
    mixed = [cases.build_2d(p, capi.MIXED_FP32) for p in plist]
    N = sum(c.ncell * c.nn for c in mixed)

    def run_batch(cs):
        b = capi.Batch(cs, autotune=True)
        b.advance(a.warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        infos = b.advance(a.steps)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        if any(i.steps != a.steps for i in infos):
            raise RuntimeError("a batch case stopped inside the timed region")
        ws = [s.macro(with_q=False) for s in b.solvers]
        scheds = [s.schedule() for s in b.solvers]
        run_batch.grouped = b.grouped_cases
        b.close()
        return t1 - t0, ws, scheds

    tm, wm, scheds = run_batch(mixed)
    grouped = run_batch.grouped
    value = N * a.steps / tm
    # the same cases one after the other (eager launches), for comparison;
    # per-kernel times from the HIP-event hooks give the roofline
    kt_all = {}
    t_seq = 0.0
    for c, sc in zip(mixed, scheds):
        s = capi.Solver(c)
        s.set_schedule(*sc)
        if a.warmup > 0:
            s.advance(a.warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.advance(a.steps)
        torch.cuda.synchronize()
        t_seq += time.perf_counter() - t0
        s.set_timing(True)
        s.reset_kernel_times()
        s.advance(1)
        for k, (ms, calls) in s.kernel_times().items():
            ab = algorithmic_bytes(c, k)
            e = kt_all.setdefault(k, [0.0, 0, 0])
            e[0] += ms
            e[1] += calls
            e[2] += ab or 0
        s.close()
    with_bytes = [k for k in kt_all if kt_all[k][2]]
    dom = max(with_bytes, key=lambda k: kt_all[k][0]) if with_bytes else None
    roof = None
    if dom:
        gbs = kt_all[dom][2] / (kt_all[dom][0] * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dom,
                "note": "sum over the cases of one step's algorithmic bytes / summed launch time"}
    out = {}
    if not a.no_fp64:
        f64 = [c.with_precision(capi.FP64) for c in mixed]
        tf, wf, _ = run_batch(f64)
        errs = np.array([field_errors(x, y) for x, y in zip(wm, wf)])
        errs_th = np.array([field_errors(x, y, thermal=True) for x, y in zip(wm, wf)])
        # the plain FP64 path: every case on its own, one after the other,
        # with the library's default schedule
        t_p64 = 0.0
        for c in f64:
            s = capi.Solver(c)
            if a.warmup > 0:
                s.advance(a.warmup)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            s.advance(a.steps)
            torch.cuda.synchronize()
            t_p64 += time.perf_counter() - t0
            s.close()
        out["fp64"] = {"value": N * a.steps / tf, "ms_per_step": 1e3 * tf / a.steps,
                       "plain_sequential_value": N * a.steps / t_p64}
        out["batched_mixed_over_plain_fp64"] = value * t_p64 / (N * a.steps)
        out["mixed_vs_fp64"] = {
            "speedup": value / out["fp64"]["value"],
            "max_rel_err_conserved": errs.max(axis=0).tolist(),
            "max_rel_err_conserved_thermal_momentum_scale": errs_th.max(axis=0).tolist(),
            "steps_compared": a.warmup + a.steps,
            "note": "channel flows started from rest: momentum is small for the first "
                    "hundreds of steps, so errors relative to the momentum magnitude are "
                    "large; the second figure scales momentum by sqrt(rho * rho E)"}
    from collections import Counter
    line = {
        "metric": METRIC, "value": value, "unit": "phase-space updates/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1e3 * tm / a.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 storage/flux, f64 moments", "data": "synthetic (seeded sweep)",
        "config": {"workload": "batched_channel_sweep_64", "cases": len(plist),
                   "phase_space_points": N, "seed": 7, "parallelism": "single GPU"},
        "roofline": roof, "cpu_baseline": None,
        "sequential_value": N * a.steps / t_seq,
        "batch_over_sequential": t_seq / tm,
        "schedules": {f"flux{f}_update{u}": n for (f, u), n in Counter(scheds).items()},
        "grouped_cases": grouped,
    }
    line.update(out)
    print(json.dumps(line))


def main():
    if "--copy-probe" in sys.argv:
        return copy_probe()
    if "--rocprof-summary" in sys.argv:
        return rocprof_summary(sys.argv[sys.argv.index("--rocprof-summary") + 1:])
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c5", choices=sorted(WORKLOADS) + ["c4"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--cpu-rows", type=int, default=16)
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {a.gpus} is not supported "
              f"(multi-GPU slab decomposition is a design note, not an implementation)",
              file=sys.stderr)
        sys.exit(2)
    wl = a.workload
    if wl == "c4":
        return bench_c4(a)
    params = WORKLOADS[wl]
    mixed = cases.build_2d(params, capi.MIXED_FP32, rt_variant=VARIANT[wl], seed=1)
    N = mixed.ncell * mixed.nn
    rm = run_gpu(mixed, a.warmup, a.steps)
    value = N * a.steps / rm["seconds"]
    out = {}
    if not a.no_fp64:
        fp64 = mixed.with_precision(capi.FP64)
        rf = run_gpu(fp64, a.warmup, a.steps)
        from metrics import field_errors, prim_errors
        gamma = mixed.desc.gamma
        pe = prim_errors(rm["w"], rf["w"], gamma)
        fe = field_errors(rm["w"], rf["w"])
        area = mixed.desc.dx * mixed.desc.dy
        w0 = np.zeros_like(rf["w"])
        out["fp64"] = {
            "value": N * a.steps / rf["seconds"], "ms_per_step": 1e3 * rf["seconds"] / a.steps,
            "kernels_ms_per_launch": {k: v[0] / v[1] for k, v in rf["kernels"].items()},
        }
        out["mixed_vs_fp64"] = {
            "speedup": value / out["fp64"]["value"],
            "rel_err": {"rho": pe[0], "U": pe[1], "V": pe[2], "T": pe[3]},
            "rel_err_conserved": {"rho": fe[0], "rhoU": fe[1], "rhoV": fe[2], "E": fe[3]},
            "mass_rel_diff": float(abs(rm["w"][:, 0].sum() - rf["w"][:, 0].sum()) /
                                   rf["w"][:, 0].sum()),
            "energy_rel_diff": float(abs(rm["w"][:, 3].sum() - rf["w"][:, 3].sum()) /
                                     rf["w"][:, 3].sum()),
            "steps_compared": a.warmup + a.steps,
        }
        del w0, area
    # dominant kernel of the headline (mixed) path
    kt = rm["kernels"]
    dom = max(kt, key=lambda k: kt[k][0])
    avg_ms = kt[dom][0] / kt[dom][1]
    ab = algorithmic_bytes(mixed, dom)
    roof = None
    if ab is not None:
        gbs = ab / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(wl, dom),
                "kernel": dom,
                "avg_launch_ms": avg_ms, "algorithmic_bytes_per_launch": ab}
    cpu = None
    if not a.no_cpu:
        try:
            cpu = cpu_baseline(wl, params, a.cpu_rows)
        except Exception as e:  # reported, not fatal
            cpu = {"error": str(e)}
    ms = 1e3 * rm["seconds"] / a.steps
    line = {
        "metric": METRIC,
        "value": value, "unit": "phase-space updates/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 storage/flux, f64 moments",
        "data": "synthetic (seeded initial condition)",
        "config": {"workload": NAMES[wl], "cells": [params["nx"], params["ny"]],
                   "nv": params["nv"], "Nm": params["Nm"], "knudsen": params["knudsen"],
                   "phase_space_points": N, "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernels_ms_per_launch": {k: v[0] / v[1] for k, v in kt.items()},
    }
    line.update(out)
    if cpu and "value" in cpu:
        line["gpu_over_cpu"] = value / cpu["value"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
