"""Benchmark of the MI355X VL MHD step: 3D decaying MHD turbulence, 256^3,
periodic, HLLD + PLM + CT, FP64 (cell-updates/s on one GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--no-cpu]

Prints one JSON line.  A step is one full Van-Leer step (predictor +
corrector + ghost fill + CFL dt) over the whole grid; inputs are resident in
HBM before the timed region starts.  The CPU baseline is the oracle (a C
restatement of the reference's algorithm with its OpenMP structure) on the
host cores, over a bounded sample of the same workload.
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

# Algorithmic HBM bytes per unit of each kernel, FP64 (8-byte words):
#   flux_*: per face, read 9 prim + 1 face-normal B, write 8 flux components
#   emf_*:  per edge triple, read 6 flux components + 6 prim (v, B), write 3
#   update_*: per cell, read 9 prim + 18 lower-face fluxes + 3 EMFs + 3 faces,
#             write 9 prim + 3 faces
KERNEL_BYTES = {"flux": 18 * 8, "emf": 15 * 8, "update": 45 * 8}
STEP_BYTES_FP64 = 360.0  # 5 state volumes of 9 words per cell-update


def face_units(name, n, ng=3):
    """Units (faces/edges/cells) one launch of a kernel processes at n^3."""
    nt = n + 2 * ng
    order = 1 if name.endswith("plm") or name.endswith("corr") else 0
    lo, hi = order + 1, nt - 2 - order
    m = hi - lo + 1
    if name.startswith("flux"):
        return (m + 1) * (m + 2) * (m + 2)
    if name.startswith("emf"):
        return (m + 1) ** 3
    return nt ** 3


def cpu_baseline(budget_s=15.0):
    import numpy as np  # noqa: F401

    from bonzo_amd import problems
    from oracle.oracle import OracleSim
    n = 64
    threads = min(16, os.cpu_count() or 1)
    s, prim, bfld = problems.turbulence(n)
    o = OracleSim(n, n, n, nthreads=threads)
    o.prim[:] = prim
    o.bfld[:] = bfld
    o.init()
    t0 = time.perf_counter()
    o.advance_steps(1)
    one = time.perf_counter() - t0
    k = max(1, min(50, int(budget_s / max(one, 1e-3))))
    t0 = time.perf_counter()
    o.advance_steps(k)
    el = time.perf_counter() - t0
    o.close()
    return {"value": n ** 3 * k / el, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"oracle (C restatement, OpenMP sweeps as in the reference) on {n}^3 "
                      f"turbulence, {k} steps after 1 warm-up step, {el:.1f} s"}


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the chunk autotuner")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB mebibytes (rocprof counter scale)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the path is measured on one GPU's share "
              "(multi-GPU sharding is designed, not built)", file=sys.stderr)
        sys.exit(2)

    from bonzo_amd import problems
    from bonzo_amd._lib import Context
    from bonzo_amd.autotune import autotune

    n = args.n
    s, prim, bfld = problems.turbulence(n)
    g = Context(n, n, n, rsolver="hlld", reconstr="linear", cour=0.4)
    g.upload(prim, bfld)
    g.init_state()
    tuned, table = (0, 0, 0), {}
    if not args.no_tune:
        # tuning steps advance the state: restart from the initial condition
        tuned, table = autotune(g)
        g.upload(prim, bfld)
        g.init_state()
        g.set_time(0.0, 0)
    del prim, bfld
    g.run_steps(args.warmup)
    g.synchronize()

    # timed region: K steps, device-resident dt, bracketed by stream syncs;
    # every launch is bracketed by HIP events on the library's stream
    g.set_kernel_timing(True)
    g.synchronize()
    t0 = time.perf_counter()
    g.run_steps(args.steps)
    g.synchronize()
    el = time.perf_counter() - t0
    g.set_kernel_timing(False)
    kt = g.kernel_timing()
    d = g.diag()
    g.close()

    ms = 1e3 * el / args.steps
    value = n ** 3 * args.steps / el
    dom = max(kt.items(), key=lambda kv: kv[1][0])
    dname, (dtot, dlaunch) = dom
    davg_ms = dtot / dlaunch
    kind = dname.split("_")[0]
    bytes_per_launch = KERNEL_BYTES[kind] * face_units(dname, n)
    achieved = bytes_per_launch / (davg_ms * 1e-3) / 1e9
    step_achieved = STEP_BYTES_FP64 * value / 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"3D decaying MHD turbulence {n}^3 periodic, VL + PLM + HLLD + CT, "
                               "FP64 plain path, CFL 0.4, seeded k^-2 fields",
                   "n": n, "precision": "fp64", "rsolver": "hlld", "reconstr": "linear",
                   "chunks": {"flux": tuned[0], "emf": tuned[1], "update": tuned[2]}},
        "roofline": {"bound": "hbm", "kernel": dname, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "avg_launch_ms": davg_ms, "algorithmic_bytes_per_launch": bytes_per_launch},
        "step_roofline": {"bytes_per_cell_update": STEP_BYTES_FP64, "achieved_GBs": step_achieved,
                          "frac_of_8TBs": step_achieved / HBM_PEAK_GBS},
        "kernels_ms_per_launch": {k: v[0] / v[1] for k, v in sorted(kt.items())},
        "divb_max": d["divb_max"],
        "autotune_ms": {str(c): {k: round(v, 3) for k, v in t.items()} for c, t in table.items()},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
