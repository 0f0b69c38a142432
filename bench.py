"""Benchmark of the MI355X VL MHD step: 3D decaying MHD turbulence, 256^3,
periodic, HLLD + PLM + CT.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--no-cpu]
                    [--no-tune] [--no-fp32c] [--accuracy-steps 100]

Prints one JSON line.  ``value`` is the FP64 plain path (cell-updates/s);
the FP32-compensated path is timed on the same initial state and reported
beside it with its speed-up and its per-variable relative L1 error against
the FP64 path after ``--accuracy-steps`` steps.  A step is one full Van-Leer
step (CFL dt, predictor, corrector, ghost fill) over the whole grid; inputs
are resident in HBM before the timed region starts.  The CPU baseline is the
oracle (a C restatement of the reference's algorithm with its OpenMP
structure) on the host cores, over a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/s at 256^3 HLLD+CT, FP32-comp vs FP64 on 1 MI355X; %HBM roofline"
HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes per unit of each kernel (4- or 8-byte words):
#   flux3_*:  per cell of the union of the face ranges, read 9 prim + 3
#             face-normal B; plus per face written, 8 flux components
#   flux_*:   per face, read 9 prim + 1 face-normal B, write 8 flux components
#   emf_*:    per edge triple, read 6 flux components + 6 prim (v, B), write 3
#   update_*: per cell, FP64: read 9 prim + 18 lower-face fluxes + 3 EMFs +
#             3 faces, write 9 prim + 3 faces (45 words of 8 bytes);
#             FP32C predictor: read 9 prim + 6 hi words + 18 fluxes + 3 EMFs
#             (4 bytes) + 3 faces (8 bytes), write 9 prim + 3 faces (4 bytes);
#             FP32C corrector: read 6 hi + 6 lo + 18 fluxes + 3 EMFs (4 bytes)
#             + 3 faces (8 bytes), write 6 hi + 6 lo + 9 prim (4) + 3 faces (8)
KERNEL_BYTES = {
    "fp64": {"flux": 18 * 8, "emf": 15 * 8, "update_pred": 45 * 8, "update_corr": 45 * 8},
    "fp64fma": {"flux": 18 * 8, "emf": 15 * 8, "update_pred": 45 * 8, "update_corr": 45 * 8},
    "fp32c": {"flux": 18 * 4, "emf": 15 * 4, "update_pred": 48 * 4 + 3 * 8,
              "update_corr": 54 * 4 + 6 * 8},
}
# HBM bytes per launch of the dominant kernels at 256^3 (FETCH_SIZE x 2.0 +
# WRITE_SIZE x 1.0, scaled by a plain device copy) and the fraction of the
# sampled SIMDs' cycles spent issuing VALU work (SQ_ACTIVE_INST_VALU x 4 /
# (GRBM_GUI_ACTIVE x SIMDs); near 1 a launch is bound by arithmetic issue),
# from the rocprofv3 --pmc runs summarised in the named profile, per work
# shape of the kernel's class (0 flat, n > 1 slab columns of n rows; the
# FP32C tile shape runs flat); the bench itself runs no counters.  FETCH_SIZE
# counts reads that the MALL serves too, which the flat launches (workgroups
# in runs dealt round-robin over the XCDs) lean on, so for them it bounds the
# HBM bytes from above.
PROFILE = "profiles/vl_step_kernels_turb256.md"
PROFILED = {
    ("fp64fma", "flux3_plm", 0): (8.298e9, 0.64),
    ("fp64fma", "flux3_plm", 8): (5.995e9, 0.58),
    ("fp64", "flux3_plm", 0): (8.298e9, 0.76),
    ("fp64", "flux3_plm", 8): (5.992e9, 0.71),
    ("fp32c", "update_corr", 0): (5.326e9, 0.38),
    ("fp32c", "update_corr", 16): (4.774e9, 0.32),
    ("fp32c", "flux3_plm", 0): (3.926e9, 0.93),
    ("fp32c", "flux3_plm", 8): (2.944e9, 0.85),
}


def profiled(prec, kernel, tuned):
    """(traffic, valu_busy) profiled for kernel at the work shape its class
    ran with, or (None, None)."""
    cls = 0 if kernel.startswith("flux") else 1 if kernel.startswith("emf") else 2
    shape = tuned[cls]
    if prec == "fp32c" and shape == 1:
        shape = 0
    return PROFILED.get((prec, kernel, shape), (None, None))


# Algorithmic bytes per cell-update of a whole VL step: 5 state volumes of 9
# words (FP64 360 B; FP32C 252 B with hi + lo on the stage-2 read of S_n and
# the write of S_{n+1}).
STEP_BYTES = {"fp64": 360.0, "fp64fma": 360.0, "fp32c": 252.0}


def kernel_units(name, n, ng=3):
    """Units (faces/edges/cells) one launch of a kernel processes at n^3."""
    nt = n + 2 * ng
    order = 1 if name.endswith("plm") or name.endswith("corr") else 0
    lo, hi = order + 1, nt - 2 - order
    m = hi - lo + 1
    if name.startswith("flux3"):
        return (m + 2) ** 3
    if name.startswith("flux"):
        return (m + 1) * (m + 2) * (m + 2)
    if name.startswith("emf"):
        return (m + 1) ** 3
    return nt ** 3


def kernel_bytes(prec, name, n):
    word = 4 if prec == "fp32c" else 8
    if name.startswith("flux3"):
        faces = 3 * kernel_units(name.replace("flux3", "flux"), n)
        return kernel_units(name, n) * 12 * word + faces * 8 * word
    table = KERNEL_BYTES[prec]
    key = name if name in table else name.split("_")[0]
    return table[key] * kernel_units(name, n)


def lscpu():
    """Model, sockets, cores per socket and threads per core of the host."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, check=True).stdout
    except (OSError, subprocess.CalledProcessError):
        return info
    keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
            "Thread(s) per core": "threads_per_core", "CPU(s)": "cpus"}
    for line in out.splitlines():
        k, _, v = line.partition(":")
        if k.strip() in keys:
            v = v.strip()
            info[keys[k.strip()]] = int(v) if v.isdigit() else v
    return info


def cpu_run(n, steps, build, threads):
    """One child process timing the oracle (oracle/cpu_bench.py) with
    ``threads`` OpenMP threads pinned one per core."""
    import subprocess
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close",
               OMP_PLACES="cores")
    r = subprocess.run([sys.executable, os.path.join(ROOT, "oracle", "cpu_bench.py"), "--n",
                        str(n), "--steps", str(steps), "--build", build], env=env,
                       capture_output=True, text=True, timeout=900)
    if r.returncode != 0:
        return {"build": build, "n": n, "threads": threads, "error": r.stderr[-400:]}
    return json.loads(r.stdout.strip().splitlines()[-1])


def cpu_baseline(n):
    """The oracle (a C restatement of the reference's algorithm with its
    OpenMP sweeps) on the host: the benchmark's own workload (n^3, a few
    steps) with one pinned thread per physical core of this process's CPU
    share, in the parity build and in the reference's fast-math compiler
    preset, plus one core on a 64^3 sample."""
    cpu = lscpu()
    physical = int(cpu.get("sockets", 1)) * int(cpu.get("cores_per_socket", os.cpu_count() or 1))
    share = int(os.environ.get("OMP_NUM_THREADS") or physical)
    threads = max(1, min(physical, share))
    runs = [cpu_run(n, 2, "parity", threads), cpu_run(n, 2, "reference-flags", threads),
            cpu_run(64, 3, "parity", 1)]
    main_run = runs[0]
    return {"value": main_run.get("value"), "unit": "cell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle (C restatement, OpenMP sweeps as in the reference) on the {n}^3 "
                      f"turbulence workload, 2 timed steps after 1 untimed, {threads} threads "
                      "pinned one per core (OMP_PROC_BIND=close, OMP_PLACES=cores; the host's "
                      f"{physical} physical cores, capped at this process's CPU share "
                      f"OMP_NUM_THREADS={share}), parity build",
            "cpu": cpu, "runs": runs}


def copy_calibration(mib):
    """A plain device-to-device copy of known size, five times, so that the
    FETCH_SIZE/WRITE_SIZE counters can be scaled to bytes."""
    import torch
    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": 8 * n, "launches": 5}), flush=True)


def run_path(prec, n, prim, bfld, args):
    """Time one precision on the shared initial state; returns timings and the
    state after ``accuracy_steps`` steps (counting warm-up and timed steps)."""
    from bonzo_amd._lib import Context
    from bonzo_amd.autotune import autotune
    g = Context(n, n, n, rsolver="hlld", reconstr="linear", cour=0.4, precision=prec)
    g.upload(prim, bfld)
    g.init_state()
    tuned, table = (0, 0, 0), {}
    if args.shapes:
        tuned = tuple(int(x) for x in args.shapes.split(","))
        g.set_tuning(*tuned)
    elif not args.no_tune:
        # tuning steps advance the state: restart from the initial condition
        tuned, table = autotune(g)
        g.upload(prim, bfld)
        g.init_state()
    g.run_steps(args.warmup)
    g.synchronize()
    # timed region: K steps with the device-resident dt, bracketed by stream
    # syncs; every launch is bracketed by HIP events on the library's stream
    g.set_kernel_timing(True)
    g.synchronize()
    t0 = time.perf_counter()
    g.run_steps(args.steps)
    g.synchronize()
    el = time.perf_counter() - t0
    g.set_kernel_timing(False)
    kt = g.kernel_timing()
    done = args.warmup + args.steps
    rest = max(0, args.accuracy_steps - done)
    g.run_steps(rest)
    d = g.diag()
    state = g.download() if args.accuracy_steps > 0 else None
    g.close()
    value = n ** 3 * args.steps / el
    dname, (dtot, dl) = max(kt.items(), key=lambda kv: kv[1][0])
    davg = dtot / dl
    b = kernel_bytes(prec, dname, n)
    ach = b / (davg * 1e-3) / 1e9
    traffic, busy = profiled(prec, dname, tuned) if n == 256 else (None, None)
    return {
        "value": value, "ms_per_step": 1e3 * el / args.steps, "tuned": tuned, "table": table,
        "kernels": {k: v[0] / v[1] for k, v in sorted(kt.items())},
        "roofline": {"bound": "hbm", "kernel": dname, "achieved": ach, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                     "traffic": traffic, "traffic_source": PROFILE if traffic else None,
                     "valu_busy": busy,
                     "avg_launch_ms": davg, "algorithmic_bytes_per_launch": b},
        "step_GBs": STEP_BYTES[prec] * value / 1e9,
        "divb_max": d["divb_max"], "hlld_guards": d["hlld_guards"],
        "steps_total": max(done, args.accuracy_steps), "state": state,
    }


def run_batch_workload(cases, reps, use_graph=True, precision="fp64", keep_state=False,
                       tune=True, blocks=None):
    """Time a batch of boxes from their initial states to their t_max:
    ``reps`` timed repetitions after one warm-up (which also captures the
    graph); cell-updates = sum over boxes of active cells x steps taken.
    ``blocks`` (flux, emf, update workgroup caps) are applied as given, else
    ``tune`` picks them with the batch autotuner."""
    from bonzo_amd._lib import Batch
    from bonzo_amd.autotune import autotune_batch
    b = Batch([dict(nx=s.nx, ny=s.ny, nz=s.nz, bc=bc, lmin=s.lmin, lmax=s.lmax, tmax=tm)
               for (s, _, _), bc, tm in cases], rsolver="hlld", reconstr="linear", cour=0.4,
              precision=precision)

    def reset():
        for q, ((s, prim, bfld), _, _) in enumerate(cases):
            b.upload(q, prim, bfld)
        b.init_state()

    table = {}
    if blocks is not None:
        tuned = tuple(blocks)
        b.set_tuning(*tuned)
    elif tune:
        tuned, table = autotune_batch(b, reset)
    else:
        tuned = (0, 0, 0)
    reset()
    b.advance(max_steps=2, use_graph=use_graph)
    walls = []
    for _ in range(reps):
        reset()
        b.synchronize()
        t0 = time.perf_counter()
        b.advance(use_graph=use_graph)  # returns after the device has finished
        walls.append(time.perf_counter() - t0)
    t, steps = b.times()
    state = [b.download(q) for q in range(len(cases))] if keep_state else None
    b.close()
    cells = [s.nx * s.ny * s.nz for (s, _, _), _, _ in cases]
    cu = float(sum(c * int(n) for c, n in zip(cells, steps)))
    med = sorted(walls)[len(walls) // 2]
    out = {"value": cu / med, "unit": "cell-updates/s", "wall_s_median": med,
           "wall_s": walls, "cell_updates": cu, "boxes": len(cases), "cells": sum(cells),
           "steps_min": int(steps.min()), "steps_max": int(steps.max()),
           "all_at_tmax": bool(all(t[q] == cases[q][2] for q in range(len(cases)))),
           "graph": use_graph, "precision": precision,
           "tuned_blocks": dict(zip(("flux", "emf", "update"), tuned)),
           "autotune_ms": {k: {str(c): round(v * 1e3, 3) for c, v in t.items()}
                           for k, t in table.items()}}
    return out, state


def batch_rel_l1(cases, a, b):
    """Per-variable relative L1 of batch states ``a`` against ``b`` over the
    active cells of every box together."""
    import numpy as np
    names = ["rho", "vx", "vy", "vz", "p", "psc", "bxc", "byc", "bzc", "bxf", "byf", "bzf"]
    num = np.zeros(12)
    den = np.zeros(12)
    for q, ((s, _, _), _, _) in enumerate(cases):
        sk, sj, si = s.active()
        for v in range(12):
            x = a[q][v // 9][v % 9][sk, sj, si] if v < 9 else a[q][1][v - 9][sk, sj, si]
            y = b[q][v // 9][v % 9][sk, sj, si] if v < 9 else b[q][1][v - 9][sk, sj, si]
            num[v] += float(np.sum(np.abs(x - y)))
            den[v] += float(np.sum(np.abs(y)))
    return {names[v]: num[v] / den[v] for v in range(12) if den[v] > 0}


def batch_benches(reps, fp32c=True, use_graph=True):
    """The batched sweeps; ``use_graph=False`` launches every step eagerly
    (the same kernels, for kernel-trace profiling)."""
    from bonzo_amd import problems
    out = {}
    sweeps = [("tubes", [(c, (1,) * 6, 0.1) for c in problems.tube_sweep(512, seed=4004)], reps,
               "512 seeded 1D MHD Riemann tubes, lengths 64-4096 cells on [0,1], outflow, "
               "VL + PLM + HLLD, CFL 0.4, each to t=0.1"),
              ("blast2d", [(c, (0,) * 6, 0.2) for c in problems.blast_sweep(64, seed=5005)],
               max(1, reps // 2 + 1),
               "64 periodic 2D MHD blast boxes, sides 64-512 on [0,1]^2, VL + PLM + HLLD + CT, "
               "CFL 0.4, each to t=0.2")]
    for name, cases, nrep, desc in sweeps:
        r64, s64 = run_batch_workload(cases, nrep, use_graph=use_graph, keep_state=fp32c)
        r64["workload"] = desc + ", FP64"
        out[name] = r64
        if name == "tubes" and use_graph:
            tb = tuple(r64["tuned_blocks"].values())
            out["tubes_no_graph"], _ = run_batch_workload(cases, 1, use_graph=False, blocks=tb)
        if fp32c:
            r32, s32 = run_batch_workload(cases, nrep, use_graph=use_graph, precision="fp32c",
                                          keep_state=True)
            r32["workload"] = desc + ", FP32-compensated"
            r32["speedup_vs_fp64"] = r32["value"] / r64["value"]
            l1 = batch_rel_l1(cases, s32, s64)
            r32["rel_l1_vs_fp64"] = l1
            r32["rel_l1_max"] = max(l1.values())
            out[name + "_fp32c"] = r32
            del s32
        del s64, cases
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the work-shape autotuner")
    ap.add_argument("--shapes", default="", metavar="F,E,U",
                    help="fixed work shapes of the flux, EMF and update kernels (no autotuning)")
    ap.add_argument("--no-fp32c", action="store_true", help="time the FP64 path only")
    ap.add_argument("--fp64-ieee-only", action="store_true",
                    help="time only the IEEE FP64 path (no FMA mode); it is then the value")
    ap.add_argument("--accuracy-steps", type=int, default=100,
                    help="steps after which FP32C is compared with FP64 (0: skip)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-box workloads")
    ap.add_argument("--batch-reps", type=int, default=5,
                    help="timed repetitions of the batched tube sweep (blast boxes: half)")
    ap.add_argument("--batch-only", action="store_true",
                    help="only run the batched-box workloads")
    ap.add_argument("--batch-eager", action="store_true",
                    help="batched workloads without the captured graph (for kernel traces)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB mebibytes (rocprof counter scale)")
    ap.add_argument("--only", choices=("fp64", "fp64fma", "fp32c"), default="",
                    help="time one path only and print its figures (for profiler runs: the "
                         "IEEE and FMA units hold kernels of the same names)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the path is measured on one GPU's share "
              "(multi-GPU sharding is designed, not built)", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from bonzo_amd import problems

    if args.batch_only:
        print(json.dumps({"batched": batch_benches(args.batch_reps, not args.no_fp32c,
                                                   not args.batch_eager)}),
              flush=True)
        return
    n = args.n
    s, prim, bfld = problems.turbulence(n)
    if args.only:
        r = run_path(args.only, n, prim, bfld, args)
        print(json.dumps({"path": args.only, "n": n, "value": r["value"], "unit": "cell-updates/s",
                          "ms_per_step": r["ms_per_step"], "roofline": r["roofline"],
                          "kernels_ms_per_launch": r["kernels"], "divb_max": r["divb_max"],
                          "shapes": dict(zip(("flux", "emf", "update"), r["tuned"]))}),
              flush=True)
        return
    # FP64, bit for bit the oracle (IEEE, no contraction)
    rie = run_path("fp64", n, prim, bfld, args)
    # FP64 with multiply-adds fused (within 1e-12 of the oracle): the FP64 value
    rfm = None if args.fp64_ieee_only else run_path("fp64fma", n, prim, bfld, args)
    r32 = None if args.no_fp32c else run_path("fp32c", n, prim, bfld, args)
    del prim, bfld
    r64 = rfm if rfm is not None else rie
    prec64 = "fp64fma" if rfm is not None else "fp64"

    def rel_l1(a, b):
        sk, sj, si = s.active()
        names = ["rho", "vx", "vy", "vz", "p", "psc", "bxc", "byc", "bzc"]
        l1 = {}
        for v in range(9):
            x = a[0][v][sk, sj, si]
            y = b[0][v][sk, sj, si]
            den = float(np.sum(np.abs(y)))
            if den > 0:
                l1[names[v]] = float(np.sum(np.abs(x - y)) / den)
        for c, nm in enumerate(("bxf", "byf", "bzf")):
            x = a[1][c][sk, sj, si]
            y = b[1][c][sk, sj, si]
            l1[nm] = float(np.sum(np.abs(x - y)) / np.sum(np.abs(y)))
        return l1

    out = {
        "metric": METRIC,
        "value": r64["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"3D decaying MHD turbulence {n}^3 periodic, VL + PLM + HLLD + CT, "
                               "FP64 plain path, CFL 0.4, seeded k^-2 fields "
                               "(rho=1, p=0.6, Bx=0.2, dB_rms=0.05, Mach 0.5)",
                   "n": n, "precision": prec64, "rsolver": "hlld", "reconstr": "linear",
                   "arithmetic": ("FP64, a*b+c fused into FMA (within 1e-12 relative L1 of the "
                                  "IEEE path and the oracle)" if rfm is not None else
                                  "FP64, IEEE without contraction (bit for bit the oracle)"),
                   "shapes": dict(zip(("flux", "emf", "update"), r64["tuned"]))},
        "roofline": r64["roofline"],
        "step_roofline": {"bytes_per_cell_update": STEP_BYTES["fp64"],
                          "achieved_GBs": r64["step_GBs"],
                          "frac_of_8TBs": r64["step_GBs"] / HBM_PEAK_GBS},
        "kernels_ms_per_launch": r64["kernels"],
        "divb_max": r64["divb_max"],
        "autotune_ms": {str(c): {k: round(v, 3) for k, v in t.items()}
                        for c, t in r64["table"].items()},
    }
    if rfm is not None:
        ieee = {"value": rie["value"], "ms_per_step": rie["ms_per_step"],
                "dtype": "f64 (IEEE, no contraction: bit for bit the oracle)",
                "roofline": rie["roofline"], "kernels_ms_per_launch": rie["kernels"],
                "shapes": dict(zip(("flux", "emf", "update"), rie["tuned"])),
                "divb_max": rie["divb_max"]}
        if rie["state"] is not None and rfm["state"] is not None:
            l1 = rel_l1(rfm["state"], rie["state"])
            out["fma_rel_l1_vs_ieee"] = l1
            out["fma_rel_l1_max"] = max(l1.values())
            out["accuracy_steps"] = rie["steps_total"]
        out["fp64_ieee"] = ieee
    if r32 is not None:
        fp32c = {"value": r32["value"], "ms_per_step": r32["ms_per_step"],
                 "speedup_vs_fp64": r32["value"] / r64["value"],
                 "speedup_vs_fp64_ieee": r32["value"] / rie["value"],
                 "dtype": "f32 (hi+lo state; FP64 update, pressure recovery, CT increments, dt)",
                 "roofline": r32["roofline"],
                 "step_roofline": {"bytes_per_cell_update": STEP_BYTES["fp32c"],
                                   "achieved_GBs": r32["step_GBs"],
                                   "frac_of_8TBs": r32["step_GBs"] / HBM_PEAK_GBS},
                 "shapes": dict(zip(("flux", "emf", "update"), r32["tuned"])),
                 "kernels_ms_per_launch": r32["kernels"], "divb_max": r32["divb_max"],
                 "hlld_guards": r32["hlld_guards"]}
        if rie["state"] is not None and r32["state"] is not None:
            l1 = rel_l1(r32["state"], rie["state"])
            fp32c["rel_l1_vs_fp64"] = l1
            fp32c["rel_l1_max"] = max(l1.values())
            fp32c["accuracy_steps"] = rie["steps_total"]
        out["fp32c"] = fp32c
    if not args.no_batch:
        out["batched"] = batch_benches(args.batch_reps, not args.no_fp32c)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
