#!/usr/bin/env python3
"""Benchmark of the D3Q19 channel-flow step on one MI355X.

One "step" is the reference's per-step update (DNS-solver/src/
dns-const-q-xycut.c:487-518: encalc + wshearstr + xcomputations) over the
whole grid: the fused collide-stream kernel plus the per-step log
reductions, with the state resident in HBM.

Default workload: the Re_tau = 180 smooth-wall channel, grid
768 x 384 x 192 (xdv streamwise, ydv spanwise, zdv wall-normal; 56.6 M
fluid nodes), constant body force, plain FP64 populations (the value line)
and FP32-shifted populations (reported beside it).

Prints ONE JSON line (rank 0).  Multi-GPU runs are not part of this build.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "dns-solvers_amd"))

import dnsgpu  # noqa: E402

PEAK_HBM_GBS = 8000.0

# Workloads (sizes named by what they are).  u_tau is lowered from 0.05 to
# 0.0054 lattice units to keep the Mach number of the centreline near 0.17;
# grid, Re_tau and h are the configuration's.
WORKLOADS = {
    "retau180_768x384x192": dict(X=768, Y=384, Z=192, retau=180.0, utau=0.0054),
    "retau395_1536x768x384": dict(X=1536, Y=768, Z=384, retau=395.0, utau=0.0054),
    # riblets on both walls: triangular, spacing 16 cells, height 8 cells
    "retau395_riblets_1536x768x384": dict(X=1536, Y=768, Z=384, retau=395.0, utau=0.0054,
                                          riblets=(16, 8)),
    "laminar_64x32x32": dict(X=64, Y=32, Z=32, retau=None, utau=None),
}


def case_params(w):
    """omega, force, ub of a workload."""
    if w["retau"] is None:
        nu, F = 0.01, 1e-6
        ub = F * w["Z"] ** 2 / (12 * nu)
    else:
        h = w["Z"] / 2.0
        nu = w["utau"] * h / w["retau"]
        F = w["utau"] ** 2 / h
        ub = 15.7 * w["utau"]
    omega = 1.0 / (3.0 * nu + 0.5)
    return omega, F, ub, nu


def riblet_bands(X, Y, spacing, height):
    """Solid rows of triangular riblets per column (the rule of the driver's
    voxeliser, dns-solvers_amd/driver/dnscase.c)."""
    import numpy as np
    yc = np.arange(Y) + 0.5
    H = height * (1.0 - np.abs(2.0 * (np.mod(yc, spacing) / spacing) - 1.0))
    hb = np.where(H > 0.5, np.ceil(H - 0.5), 0).astype(np.int16)
    return np.ascontiguousarray(np.broadcast_to(hb[None, :], (X, Y)))


def run_precision(wname, prec, steps, warmup):
    w = WORKLOADS[wname]
    omega, F, ub, nu = case_params(w)
    X, Y, Z = w["X"], w["Y"], w["Z"]
    nodes = X * Y * Z
    ch = dnsgpu.Channel(X, Y, Z, Rebs=ub * Z / nu, ub=ub, CFL=0.1, TSTR=0, forcing=dnsgpu.CONST_F,
                        force=F, precision=prec, omega=omega)
    try:
        if w.get("riblets"):
            ch.set_solid_bands(*([riblet_bands(X, Y, *w["riblets"])] * 2))
        if w["retau"] is None:
            ch.init_field(dnsgpu.INIT_POISEUILLE, uc=F * Z * Z / (8 * nu), amp=0.05, seed=1)
        else:
            ch.init_field(dnsgpu.INIT_TURBULENT, uc=w["utau"], amp=0.1, seed=2, nmodes=4)
        blocks, _ = ch.autotune(trial_steps=5)
        ch.step(warmup, cntr0=0, logs=True)
        t0 = time.perf_counter()
        logs = ch.step(steps, cntr0=warmup, logs=True)  # returns after the stream drained
        t1 = time.perf_counter()
        wall = t1 - t0
        # kernel-only timing of the same step sequence with HIP events on the
        # library's stream (the stream the kernels run on)
        ms_k, ms_w = ch.time_steps(steps, stats=True)
        ms_k_nostats, _ = ch.time_steps(steps, stats=False)
        ok = all(v == v for v in logs["stress"])
    finally:
        ch.close()
    esize = 8 if prec == dnsgpu.FP64 else 4
    bytes_per_node = 2 * 19 * esize
    avg_kernel_s = ms_k / 1e3 / steps
    achieved = bytes_per_node * nodes / avg_kernel_s / 1e9
    return {
        "mlups": nodes * steps / wall / 1e6,
        "ms_per_step": wall / steps * 1e3,
        "kernel_ms": avg_kernel_s * 1e3,
        "kernel_ms_nostats": ms_k_nostats / steps,
        "events_wall_ms_per_step": ms_w / steps,
        "achieved_gbs": achieved,
        "bytes_per_node": bytes_per_node,
        "nodes": nodes,
        "finite": ok,
        "step_blocks": blocks,
        "stress_last": float(logs["stress"][-1]),
    }


def cpu_baseline(wname, budget_s=12.0):
    """The oracle (CPU restatement of the reference loop, one core) on a
    bounded sample of the workload: the same y/z extents with 4 x-planes."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import numpy as np
    import pyoracle
    w = WORKLOADS[wname]
    omega, F, ub, nu = case_params(w)
    X, Y, Z = 4, w["Y"], w["Z"]
    o = pyoracle.Oracle(X, Y, Z, omega, ub, 0.1, forcing=1, F=F)
    ft = np.zeros((X, Y, Z + 2, 19))
    wts = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    ft[:, :, 1:Z + 1, :] = wts
    f = ft.copy()
    n = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        o.run(ft, f, 1, cntr0=n)
        n += 1
    dt = time.perf_counter() - t0
    return {"value": X * Y * Z * n / dt / 1e6, "unit": "MLUPS", "cores": 1, "kind": "port",
            "sample": f"oracle/lbm_oracle.c, {X}x{Y}x{Z} slab of the workload, {n} steps, FP64, 1 thread"}


def measured_traffic(wname):
    """HBM bytes per launch of the step kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of this workload (calibrated against a
    copy of known size), or None when no summary exists."""
    path = os.path.join(HERE, "profiles", f"traffic_{wname}.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="retau180_768x384x192", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fp64-only", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs the whole channel on one GPU; multi-GPU runs are not part of it "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)
    dnsgpu.load()
    r64 = run_precision(args.workload, dnsgpu.FP64, args.steps, args.warmup)
    print(f"# fp64 done: {r64['mlups']:.0f} MLUPS", file=sys.stderr, flush=True)
    r32 = None
    if not args.fp64_only:
        r32 = run_precision(args.workload, dnsgpu.FP32, args.steps, args.warmup)
        print(f"# fp32 done: {r32['mlups']:.0f} MLUPS", file=sys.stderr, flush=True)
    cpu = None if args.no_cpu else cpu_baseline(args.workload)
    w = WORKLOADS[args.workload]
    tr = measured_traffic(args.workload)
    t64 = tr["fp64_stats"]["traffic_bytes_per_launch"] if tr else None
    t32 = tr["fp32_stats"]["traffic_bytes_per_launch"] if tr else None
    out = {
        "metric": "MLUPS + % HBM roofline (FP32 vs FP64 D3Q19); ensemble solves/h + p99 latency",
        "value": round(r64["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r64["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "grid_xdv_ydv_zdv": [w["X"], w["Y"], w["Z"]],
                   "fluid_nodes": r64["nodes"], "forcing": "constant body force",
                   "texture": ("riblets spacing %d height %d" % w["riblets"]) if w.get("riblets") else "smooth",
                   "per_step": "fused collide-stream + per-step log reductions (encalc/wshearstr)",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(r64["achieved_gbs"], 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(r64["achieved_gbs"] / PEAK_HBM_GBS, 4), "traffic": t64,
                     "kernel": "step_kernel<double,false,true>", "kernel_ms": round(r64["kernel_ms"], 4),
                     "bytes_per_node": r64["bytes_per_node"],
                     "traffic_source": f"profiles/traffic_{args.workload}.json" if tr else None},
        "fp64": r64,
    }
    if r32:
        out["fp32_shifted"] = dict(r32, speedup_vs_fp64=round(r32["mlups"] / r64["mlups"], 3),
                                   roofline={"bound": "hbm", "achieved": round(r32["achieved_gbs"], 1),
                                             "peak": PEAK_HBM_GBS, "unit": "GB/s",
                                             "frac": round(r32["achieved_gbs"] / PEAK_HBM_GBS, 4), "traffic": t32,
                                             "kernel": "step_kernel<float,false,true>"})
    if cpu:
        out["cpu_baseline"] = cpu
    print(json.dumps(out))


if __name__ == "__main__":
    main()
