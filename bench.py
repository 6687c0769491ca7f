"""Benchmark of the Hermite-UGKS time step on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c5|c3|c2|c4]
                [--no-cpu] [--no-fp64] [--no-bf16x3]
python bench.py --copy-probe                  (calibration copy for counters)
python bench.py --rocprof-summary DIR [DIR..] (summarise rocprofv3 CSV output)
python bench.py --traffic-summary WORKLOAD FETCH WRITE COPY_FETCH COPY_WRITE TRACE
                                              (profiles/traffic_<workload>.json)

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
FP64_PEAK_TFLOPS = 78.6       # MI355X FP64 vector and matrix
FP32_MFMA_PEAK_TFLOPS = 157.3  # MI355X dense FP32 matrix

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


SCHED_NAMES = {"flux": {capi.FLUX_FUSED: "fused", capi.FLUX_SPLIT: "split",
                        capi.FLUX_STENCIL: "stencil", capi.FLUX_CELLSUM: "cellsum"},
               "update": {capi.UPDATE_LDS: "lds", capi.UPDATE_MFMA: "mfma",
                          capi.UPDATE_MFMA_NODE: "mfma-node",
                          capi.UPDATE_MFMA_FACE: "mfma-face"}}


def algorithmic_bytes(case, kernel, update=None):
    """Algorithmic HBM bytes of one launch: each array the kernel must
    touch counted once (N = cells*nv^2, s = bytes per stored element).
    update: the case's update kernel (the face-forming one reads per-cell
    divergence arrays instead of the face arrays)."""
    d = case.desc
    s = 8 if d.precision == capi.FP64 else 4
    nn = d.nv * d.nv
    N = d.nx * d.ny * nn
    nvf = (d.nx + 1) * d.ny * nn
    nhf = d.nx * (d.ny + 1) * nn
    if kernel == "slope2d":          # read h, b; write 4 slope arrays
        return 6 * s * N
    if kernel == "slope_sum":        # read h, b; write 4 slope arrays and 4 bins of 37 sums per cell
        return 6 * s * N + 8 * 4 * 37 * d.nx * d.ny
    if kernel == "flux_vface":       # read h, b and 4 slopes; write 2 face arrays
        return 6 * s * N + 2 * s * nvf
    if kernel == "flux_hface":
        return 6 * s * N + 2 * s * nhf
    if kernel == "update2d" and update == capi.UPDATE_MFMA_FACE:
        return 6 * s * N             # read h, b and 2 divergence arrays; write h, b
    if kernel == "update2d":         # read h, b, 4 face arrays; write h, b
        return 4 * s * N + 2 * s * (nvf + nhf)
    if kernel == "flux_div":         # read h, b and 4 slopes; write 2 divergence arrays
        return 8 * s * N
    # interior-face pipeline: both orientations per launch
    xper, yper = d.bc[0] == capi.BC_PERIODIC, d.bc[2] == capi.BC_PERIODIC
    nint = (d.nx + 1 if xper else d.nx - 1) * d.ny + d.nx * (d.ny + 1 if yper else d.ny - 1)
    if kernel == "flux_sum":         # read 6 upwind arrays per face node; 37 sums per face
        return 6 * s * nint * nn + 8 * 37 * nint
    if kernel == "flux_node":        # read 6 upwind arrays, write 2 face arrays per face node
        return 8 * s * nint * nn
    return None


def run_gpu(case, warmup, steps, kernel_steps=3):
    """Times `steps` steps after `warmup` untimed ones, with the per-launch
    timing events off; then `kernel_steps` more steps with them on give the
    per-kernel device times (HIP events on the library's stream)."""
    import torch
    s = capi.Solver(case)
    w0 = s.macro(with_q=False)
    if warmup > 0:               # advance(0) would run to the loop exit
        s.advance(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    info = s.advance(steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    if info.steps != steps:
        raise RuntimeError(f"the timed region ran {info.steps} steps, not {steps}")
    w, q = s.macro()
    kt = {}
    if kernel_steps > 0:
        s.set_timing(True)
        s.reset_kernel_times()
        s.advance(kernel_steps)
        kt = s.kernel_times()
    sched = s.schedule()
    s.close()
    return dict(seconds=t1 - t0, info=info, kernels=kt, w=w, q=q, w0=w0, schedule=sched)


def host_info():
    """nproc and the CPU model of the host the baseline runs on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def hermite_flops(case):
    """Algorithmic flops of the separable Hermite force per step (h and b),
    12 (Nm+1) nv (nv+Nm+1) per cell; the reconstruction diagnostic of the
    reference is excluded (update() discards it)."""
    d = case.desc
    return 12.0 * (d.nm + 1) * d.nv * (d.nv + d.nm + 1) * d.nx * d.ny


def drift(w, w0):
    """Relative change of total mass and total energy from t = 0."""
    m0, e0 = w0[:, 0].sum(), w0[:, -1].sum()
    return {"mass": float(abs(w[:, 0].sum() - m0) / abs(m0)),
            "energy": float(abs(w[:, -1].sum() - e0) / abs(e0))}


def ref_harness_run(params, steps, variant, rows=0, timeout=600):
    """Loop seconds of the reference Fortran solver (oracle/_ref/ref_harness2d)
    on one &params set; None when the harness is not built."""
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness2d")
    if not os.path.exists(exe):
        return None
    with tempfile.TemporaryDirectory() as td:
        nml = os.path.join(td, "bench.namelist")
        with open(nml, "w") as f:
            f.write("&params\n")
            for k, v in params.items():
                f.write(f"  {k} = {v!r}\n" if isinstance(v, str) else f"  {k} = {v}\n")
            f.write("/\n")
        args = [exe, nml, "-", str(steps), "1" if variant == "cartesian" else "0", "1"]
        if rows:
            args.append(str(rows))
        r = subprocess.run(args, capture_output=True, text=True, cwd=td, timeout=timeout)
        line = [ln for ln in r.stdout.splitlines() if "HARNESS_TIMING" in ln]
        if r.returncode != 0 or not line:
            raise RuntimeError(f"reference harness failed: {r.stderr[-500:]}")
        return float(line[0].split("loop_s=")[1].split()[0])


def cpu_baseline(wl, params, budget_rows):
    """Reference Fortran solver (oracle/_ref) on a bounded sample of the
    workload: the full row width and velocity grid, budget_rows rows, one
    timed step.  Falls back to the C restatement."""
    d = WORKLOADS[wl]
    rows = min(budget_rows, d["ny"])
    steps = 1
    loop_s = ref_harness_run(params, steps, VARIANT[wl], rows)
    if loop_s is not None:
        n = d["nx"] * rows * d["nv"] ** 2 * steps
        return dict({"value": n / loop_s, "unit": "phase-space updates/s", "cores": 8,
                     "kind": "reference",
                     "sample": f"{d['nx']}x{rows} rows of the workload (full nv={d['nv']}, "
                               f"Nm={d['Nm']}), {steps} step, reference Fortran built with "
                               f"amdflang -O2 -fopenmp; 8 OpenMP threads (hard-coded in the "
                               f"reference) for timestep/interpolation/evolution, update() "
                               f"serial; loop {loop_s:.1f} s"}, **host_info())
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
    return dict({"value": n / dt, "unit": "phase-space updates/s", "cores": 1, "kind": "port",
                 "sample": f"{p['nx']}x{p['ny']} cells of the workload, 1 step, C restatement"},
                **host_info())


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


# rocprofv3 kernel names of the 2D step -> the library's timer names
TIMER_OF = [("k_update2d_mx", "update2d"), ("k_update2d_mfma", "update2d"),
            ("k_update2d<", "update2d"), ("k_fsum", "flux_sum"), ("k_fnode", "flux_node"),
            ("k_fface", "flux_face"), ("k_slope2d", "slope2d"), ("k_flux_wall", "flux_wall"),
            ("k_fdiv", "flux_div"),
            ("k_cell_sum", "slope_sum"),
            ("k_cell_pre", "cell_pre"), ("k_finalize", "finalize"),
            ("k_part_sum", "finalize_partials"), ("k_rate_part", "timestep_rates"),
            ("k_timestep", "timestep")]


def traffic_summary(workload, dirs):
    """profiles/traffic_<workload>.json content from rocprofv3 CSV output:
    dirs = FETCH_SIZE run, WRITE_SIZE run, the copy probe's FETCH_SIZE and
    WRITE_SIZE runs, and a kernel trace of the same command.  Each counter
    is scaled by what it reads over the 1 GiB copy (bytes read and written
    known), then set against the kernel's algorithmic bytes."""
    import contextlib
    import io

    def summ(d):
        buf = io.StringIO()
        with contextlib.redirect_stdout(buf):
            rocprof_summary([d])
        return json.loads(buf.getvalue())
    fetch, write, cf, cw, trace = (summ(d) for d in dirs)
    copy = "__amd_rocclr_copyBuffer"
    big = lambda s, key: max(v[key] for k, v in s.items() if k.startswith(copy) and key in v)
    gib = float(1 << 30)
    fk, wk = big(cf, "FETCH_SIZE_avg"), big(cw, "WRITE_SIZE_avg")
    cal = {"copy_bytes_read": gib, "copy_bytes_written": gib, "FETCH_SIZE_kib": fk,
           "WRITE_SIZE_kib": wk, "fetch_scale": gib / (fk * 1024), "write_scale": gib / (wk * 1024)}
    case = cases.build_2d(WORKLOADS[workload], capi.MIXED_FP32, rt_variant=VARIANT[workload],
                          seed=1)
    kern = {}
    for name, v in fetch.items():
        timer = next((t for pat, t in TIMER_OF if pat in name), None)
        if timer is None or "FETCH_SIZE_avg" not in v or name not in write:
            continue
        fb = v["FETCH_SIZE_avg"] * 1024 * cal["fetch_scale"]
        wb = write[name]["WRITE_SIZE_avg"] * 1024 * cal["write_scale"]
        ab = algorithmic_bytes(case, timer,
                               capi.UPDATE_MFMA_FACE if "k_update2d_mxf" in name else None)
        ms = trace.get(name, {}).get("avg_ms")
        kern[timer] = {"rocprof_name": name, "FETCH_SIZE_kib": v["FETCH_SIZE_avg"],
                       "WRITE_SIZE_kib": write[name]["WRITE_SIZE_avg"],
                       "dispatches": v["FETCH_SIZE_dispatches"],
                       "hbm_bytes_per_launch": fb + wb, "fetch_bytes": fb, "write_bytes": wb,
                       "algorithmic_bytes_per_launch": ab,
                       "traffic_over_algorithmic": (fb + wb) / ab if ab else None,
                       "avg_ms_trace": ms,
                       "traffic_GBs": (fb + wb) / (ms * 1e-3) / 1e9 if ms else None,
                       "algorithmic_GBs": ab / (ms * 1e-3) / 1e9 if ms and ab else None}
    print(json.dumps({"workload": NAMES[workload], "precision": "mixed", "calibration": cal,
                      "kernels": kern}, indent=1, sort_keys=True))


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
    """The batched sweep's cases (hugks.cases.sweep_c4)."""
    return cases.sweep_c4(seed, n)


def bench_c4(a):
    """Aggregate updates/s of the batched sweep: all cases advanced by one
    graph replay per step (grouped launches), mixed against the same batch
    in FP64, per-case errors after the channel flows have developed, the
    grouped kernels timed live for the roofline, and the reference solver
    timed on a sample of the cases."""
    import torch
    from metrics import (field_errors, prim_errors, prim_errors_thermal, q_error,
                         q_error_thermal)
    plist = c4_params()
    mixed = [cases.build_2d(p, capi.MIXED_FP32) for p in plist]
    N = sum(c.ncell * c.nn for c in mixed)

    def run_batch(cs, develop=0):
        b = capi.Batch(cs, autotune=True)
        w0 = [s.macro(with_q=False) for s in b.solvers]
        b.advance(a.warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        infos = b.advance(a.steps)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        if any(i.steps != a.steps for i in infos):
            raise RuntimeError("a batch case stopped inside the timed region")
        # grouped kernels, timed on eager launches of the same step
        b.set_timing(True)
        b.reset_kernel_times()
        b.advance(2)
        kt = b.kernel_times()
        b.set_timing(False)
        done = a.warmup + a.steps + 2
        if develop > done:
            b.advance(develop - done)
        out = [s.macro() for s in b.solvers]
        scheds = [s.schedule() for s in b.solvers]
        waves = [s.update_waves() for s in b.solvers]
        grouped = b.grouped_cases
        b.close()
        return dict(seconds=t1 - t0, wq=out, w0=w0, kt=kt, scheds=scheds, waves=waves,
                    grouped=grouped)

    develop = 2000     # the channel flows develop: compare there, not near rest
    rm = run_batch(mixed, develop)
    value = N * a.steps / rm["seconds"]
    # roofline of the dominant grouped kernel: algorithmic bytes of one
    # launch (summed over the cases it steps) over its average duration
    kmap = {"kg_slope": "slope2d", "kg_fsum": "flux_sum", "kg_fnode": "flux_node",
            "kg_update": "update2d", "kg_cell_sum": "slope_sum", "kg_fdiv": "flux_div"}
    roof = None
    kt = rm["kt"]
    timed = [k for k in kt if k in kmap]
    if timed:
        dom = max(timed, key=lambda k: kt[k][0])
        # kg_update runs one launch per compiled instance per step: the bytes
        # of one step over the time of all of its launches in one step
        steps_timed = kt["kg_timestep"][1] if "kg_timestep" in kt else 2
        ms_step = kt[dom][0] / steps_timed
        ab = sum(algorithmic_bytes(c, kmap[dom], sc[1]) for c, sc in zip(mixed, rm["scheds"]))
        gbs = ab / (ms_step * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dom,
                "avg_ms_per_step": ms_step, "algorithmic_bytes_per_step": ab,
                "note": "grouped launch over all cases; HIP events on eager launches"}
    out = {"kernels_ms_per_step": {k: v[0] / kt.get("kg_timestep", (0, 2))[1]
                                   for k, v in kt.items()}}
    if not a.no_fp64:
        f64 = [c.with_precision(capi.FP64) for c in mixed]
        rf = run_batch(f64, develop)
        errs, perrs, pterrs, qerrs, qterrs = [], [], [], [], []
        for k, c in enumerate(mixed):
            (wm, qm), (wf, qf) = rm["wq"][k], rf["wq"][k]
            g = c.desc.gamma
            errs.append(field_errors(wm, wf))
            perrs.append(prim_errors(wm, wf, g))
            pterrs.append(prim_errors_thermal(wm, wf, g))
            qerrs.append(q_error(qm, qf))
            qterrs.append(q_error_thermal(qm, qf, wf, g))
        errs, perrs, pterrs = np.array(errs), np.array(perrs), np.array(pterrs)
        out["fp64"] = {"value": N * a.steps / rf["seconds"],
                       "ms_per_step": 1e3 * rf["seconds"] / a.steps}
        out["mixed_vs_fp64"] = {
            "speedup": value / out["fp64"]["value"],
            "max_rel_err": {"rho": perrs[:, 0].max(), "U": perrs[:, 1].max(),
                            "V": perrs[:, 2].max(), "T": perrs[:, 3].max(),
                            "q": float(max(qerrs))},
            "max_rel_err_thermal_scale": {"U": pterrs[:, 1].max(), "V": pterrs[:, 2].max(),
                                          "q": float(max(qterrs))},
            "cases_within_1e-5_flow_scale": int(np.sum(np.all(perrs <= 1e-5, axis=1))),
            "max_rel_err_conserved": {"rho": errs[:, 0].max(), "rhoU": errs[:, 1].max(),
                                      "rhoV": errs[:, 2].max(), "E": errs[:, 3].max()},
            "max_drift_from_t0": {
                "mixed": {q: max(drift(rm["wq"][k][0], rm["w0"][k])[q] for k in range(len(mixed)))
                          for q in ("mass", "energy")},
                "fp64": {q: max(drift(rf["wq"][k][0], rf["w0"][k])[q] for k in range(len(mixed)))
                         for q in ("mass", "energy")}},
            "steps_compared": develop,
            "note": "per case after the flows developed.  max_rel_err: velocities against "
                    "the largest flow speed of the case, q against its largest |q|; "
                    "thermal scale: velocities against sqrt(max T), q against "
                    "max rho (max T)^1.5.  The slowest channels (Mach 0.01) carry ~1e-6 of "
                    "the thermal speed, the resolution of FP32 storage of f"}
    cpu = None
    if not a.no_cpu:
        # the reference solver on every eighth case, 3 steps each
        sample = list(range(0, len(plist), 8))
        n_pts = t_ref = 0.0
        for k in sample:
            ls = ref_harness_run(plist[k], 3, "reference")
            if ls is None:
                break
            n_pts += mixed[k].ncell * mixed[k].nn * 3
            t_ref += ls
        if t_ref > 0:
            cpu = dict({"value": n_pts / t_ref, "unit": "phase-space updates/s", "cores": 8,
                        "kind": "reference",
                        "sample": f"cases {sample} of the sweep, 3 steps each, each case as "
                                  f"its own run of the reference Fortran (amdflang -O2 "
                                  f"-fopenmp; 8 OpenMP threads hard-coded, update() serial); "
                                  f"summed loop {t_ref:.1f} s"}, **host_info())
    from collections import Counter
    line = {
        "metric": METRIC, "value": value, "unit": "phase-space updates/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1e3 * rm["seconds"] / a.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 storage/flux, f64 moments", "data": "synthetic (seeded sweep)",
        "config": {"workload": "batched_channel_sweep_64", "cases": len(plist),
                   "phase_space_points": N, "seed": 7, "parallelism": "single GPU"},
        "roofline": roof, "cpu_baseline": cpu,
        "schedules": {f"flux{f}_update{u}": n for (f, u), n in Counter(rm["scheds"]).items()},
        "grouped_cases": rm["grouped"],
        # the autotuner's wavefronts-per-cell tile of the MFMA node update, per case
        "update_waves": {f"{w} per cell": n for w, n in sorted(Counter(rm["waves"]).items())},
    }
    line.update(out)
    if cpu and "value" in cpu:
        line["gpu_over_cpu"] = value / cpu["value"]
    print(json.dumps(line))


def main():
    if "--copy-probe" in sys.argv:
        return copy_probe()
    if "--rocprof-summary" in sys.argv:
        return rocprof_summary(sys.argv[sys.argv.index("--rocprof-summary") + 1:])
    if "--traffic-summary" in sys.argv:
        i = sys.argv.index("--traffic-summary")
        return traffic_summary(sys.argv[i + 1], sys.argv[i + 2:i + 7])
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c5", choices=sorted(WORKLOADS) + ["c4"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-bf16x3", action="store_true")
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
        from metrics import field_errors, prim_errors, q_error
        gamma = mixed.desc.gamma
        pe = prim_errors(rm["w"], rf["w"], gamma)
        fe = field_errors(rm["w"], rf["w"])
        out["fp64"] = {
            "value": N * a.steps / rf["seconds"], "ms_per_step": 1e3 * rf["seconds"] / a.steps,
            "kernels_ms_per_launch": {k: v[0] / v[1] for k, v in rf["kernels"].items()},
        }
        out["mixed_vs_fp64"] = {
            "speedup": value / out["fp64"]["value"],
            "rel_err": {"rho": pe[0], "U": pe[1], "V": pe[2], "T": pe[3],
                        "q": q_error(rm["q"], rf["q"])},
            "rel_err_conserved": {"rho": fe[0], "rhoU": fe[1], "rhoV": fe[2], "E": fe[3]},
            "drift_from_t0": {"mixed": drift(rm["w"], rm["w0"]),
                              "fp64": drift(rf["w"], rf["w0"])},
            "steps_compared": a.warmup + a.steps,
            "note": "velocities against the largest speed, q against the largest |q|; "
                    "energy is not conserved under gravity (work of the body force)",
        }
    # the bf16x3 split-product Hermite contraction: the same step with the
    # update's tiles as bf16 MFMA products of operands split in three
    if not a.no_bf16x3:
        rb = run_gpu(mixed.with_precision(capi.MIXED_BF16X3), a.warmup, a.steps)
        from metrics import prim_errors
        b3 = {"value": N * a.steps / rb["seconds"], "ms_per_step": 1e3 * rb["seconds"] / a.steps,
              "update2d_ms": (rb["kernels"]["update2d"][0] / rb["kernels"]["update2d"][1]
                              if "update2d" in rb["kernels"] else None)}
        pb = prim_errors(rb["w"], rm["w"], mixed.desc.gamma)
        b3["rel_err_vs_fp32_tiles"] = {"rho": pb[0], "U": pb[1], "V": pb[2], "T": pb[3]}
        if not a.no_fp64:
            pf = prim_errors(rb["w"], rf["w"], mixed.desc.gamma)
            b3["rel_err_vs_fp64"] = {"rho": pf[0], "U": pf[1], "V": pf[2], "T": pf[3]}
        out["bf16x3"] = b3
    # dominant kernel of the headline (mixed) path
    kt = rm["kernels"]
    dom = max(kt, key=lambda k: kt[k][0])
    avg_ms = kt[dom][0] / kt[dom][1]
    ab = algorithmic_bytes(mixed, dom, rm["schedule"][1])
    roof = None
    if ab is not None:
        gbs = ab / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(wl, dom),
                "kernel": dom,
                "avg_launch_ms": avg_ms, "algorithmic_bytes_per_launch": ab}
    # every step kernel with algorithmic bytes: GB/s on them and on the
    # measured traffic of the committed counter summary, as fractions of peak
    per_kernel = {}
    for k, (tot, n) in kt.items():
        kb = algorithmic_bytes(mixed, k, rm["schedule"][1])
        if kb is None or n == 0:
            continue
        kms = tot / n
        g = kb / (kms * 1e-3) / 1e9
        tr = traffic_from_profiles(wl, k)
        per_kernel[k] = {"ms": round(kms, 4), "algorithmic_GBs": round(g, 1),
                         "frac": round(g / HBM_PEAK_GBS, 4),
                         "traffic_GBs": round(tr / (kms * 1e-3) / 1e9, 1) if tr else None,
                         "traffic_frac": round(tr / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                         if tr else None}
    # the Hermite force on the matrix cores: its algorithmic flops over the
    # update kernel's time (the bound of that kernel is HBM; this is the
    # MFMA utilisation the force term reaches inside it)
    mfma = None
    if "update2d" in kt:
        upd_ms = kt["update2d"][0] / kt["update2d"][1]
        tf = hermite_flops(mixed) / (upd_ms * 1e-3) / 1e12
        mfma = {"kernel": "update2d", "flops_per_launch": hermite_flops(mixed),
                "achieved": round(tf, 2), "peak": FP32_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(tf / FP32_MFMA_PEAK_TFLOPS, 4)}
        if not a.no_fp64 and "update2d" in rf["kernels"]:
            k64 = rf["kernels"]["update2d"]
            tf64 = hermite_flops(mixed) / (k64[0] / k64[1] * 1e-3) / 1e12
            mfma["fp64"] = {"achieved": round(tf64, 2), "peak": FP64_PEAK_TFLOPS,
                            "frac": round(tf64 / FP64_PEAK_TFLOPS, 4)}
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
        "schedule": {"flux": SCHED_NAMES["flux"].get(rm["schedule"][0]),
                     "update": SCHED_NAMES["update"].get(rm["schedule"][1])},
        "roofline": roof,
        "kernels_roofline": per_kernel,
        "mfma": mfma,
        "cpu_baseline": cpu,
        "kernels_ms_per_launch": {k: v[0] / v[1] for k, v in kt.items()},
    }
    line.update(out)
    if cpu and "value" in cpu:
        line["gpu_over_cpu"] = value / cpu["value"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
