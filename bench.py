#!/usr/bin/env python3
"""Benchmark of the D3Q19 channel-flow step on one MI355X.

One "step" is the reference's per-step update (DNS-solver/src/
dns-const-q-xycut.c:487-518: encalc + wshearstr + xcomputations) over the
whole grid: the fused collide-stream kernel plus the per-step log
reductions, with the state resident in HBM.

Default workload: the Re_tau = 395 channel with triangular riblets
(spacing 16, height 8 cells) on both walls, grid 1536 x 768 x 384 (xdv
streamwise, ydv spanwise, zdv wall-normal; 453 M fluid nodes), constant
body force.  The value line is FP32-shifted storage with FP32 collision;
the plain FP64 path (the speed-up's reference) and FP16-shifted storage are
reported beside it, and so is the serving figure: the 64-solve FP32
texture sweep, all requests at t = 0, against the same solves run one after
another on the plain FP64 path.

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
            ch.init_field(dnsgpu.INIT_TURBULENT, uc=w["utau"], amp=0.1, seed=2, nmodes=8)
        blocks, _ = ch.autotune(trial_steps=5)
        _, waves, rows = ch.step_shape(rows=True)
        pieces = ch.step_pieces()
        placement = ch.pair_placement()
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
    esize = {dnsgpu.FP64: 8, dnsgpu.FP32: 4, dnsgpu.FP16: 2}[prec]
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
        "step_waves_per_block": waves,
        "step_rows_per_lane": rows,
        "step_z_pieces": pieces,
        "pair_placement": {"destinations_tried": placement[0], "step_gain": round(placement[1], 4)},
        "stress_last": float(logs["stress"][-1]),
    }


def cpu_baseline(wname, budget_s=12.0, w=None):
    """The oracle (CPU restatement of the reference loop, one core) on a
    bounded sample of the workload: the same y/z extents with 4 x-planes."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import numpy as np
    import pyoracle
    w = w or WORKLOADS[wname]
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


# The reference solver itself as the CPU baseline: oracle/build_ref.sh builds
# it from its own sources (where they lie, in the build container) for a
# 4-plane slab of a workload, twice, with two values of ITER (the step count
# is a compile-time macro, definitions.h:52); the binaries travel with the
# tree.  The per-step cost is the difference of the two runs' wall times
# over the extra steps, so start-up, the restart read and the step-0 outputs
# cancel.  The reference has solids only in its wall rows, so a riblet
# workload is timed on its smooth channel (the same grid and physics).
REF_ITERS = {"retau180_768x384x192": (10, 110), "retau395_1536x768x384": (10, 60)}
REF_BASE = {"retau395_riblets_1536x768x384": "retau395_1536x768x384"}
# the reference's own parallel mode beside it: x-slabs over REF_MPI_RANKS
# ranks (the host cores a one-GPU share of the machine has), two planes each
REF_MPI_RANKS = 16
REF_MPI_PLANES = 2


def _ref_case(wname, ranks=1):
    w = WORKLOADS[wname]
    omega, F, ub, nu = case_params(w)
    # the reference's omega = 1/((3/Rebs) ub zdv + 1/2) (dns-const-q-xycut.c:216-217) equals 1/(3 nu + 1/2)
    X = 4 if ranks == 1 else REF_MPI_PLANES * ranks
    return dict(X=X, Y=w["Y"], Z=w["Z"], Rebs=ub * w["Z"] / nu, ub=ub, CFL=0.1)


def _ref_exe(wname, iters, ranks=1):
    tag = "" if ranks == 1 else f"_p{ranks}"
    return os.path.join(HERE, "oracle", "_ref", f"bench_{wname}{tag}_n{iters}", "run-dns")


def build_reference_baseline():
    """Build the reference binaries of the CPU baseline (only where the
    reference sources are present; a build step, not a measurement)."""
    import subprocess
    for wname, iters_pair in REF_ITERS.items():
        for ranks in (1, REF_MPI_RANKS):
            c = _ref_case(wname, ranks)
            tag = "" if ranks == 1 else f"_p{ranks}"
            for iters in iters_pair:
                subprocess.run([os.path.join(HERE, "oracle", "build_ref.sh"), f"bench_{wname}{tag}_n{iters}",
                                str(c["X"]), str(c["Y"]), str(c["Z"]), repr(c["Rebs"]), repr(c["ub"]), repr(c["CFL"]),
                                "1", "2", str(iters)], check=True, capture_output=True,
                               env=dict(os.environ, REF_XDIM=str(ranks)))


def reference_baseline(wname, ranks=1):
    """Per-step rate of the reference solver (its own constant-flow-rate
    forcing, from a rest state): one MPI rank on one core on the 4-plane
    slab, or `ranks` ranks splitting a 2-planes-per-rank slab in x (its
    parallel mode, throughput only: each rank exchanges its x faces with its
    neighbours every step); None when its binaries or mpirun are absent."""
    import shutil
    import subprocess
    import tempfile

    import numpy as np
    mpirun = os.environ.get("MPIRUN", "/opt/conda/bin/mpirun")
    wname = REF_BASE.get(wname, wname)
    iters_pair = REF_ITERS.get(wname)
    exes = [_ref_exe(wname, n, ranks) for n in iters_pair] if iters_pair else []
    if not exes or not os.path.exists(mpirun) or not all(os.path.exists(e) for e in exes):
        return None
    c = _ref_case(wname, ranks)
    X, Y, Z = c["X"], c["Y"], c["Z"]
    ft = np.zeros((X, Y, Z + 2, 19))
    ft[:, :, 1:Z + 1, :] = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    secs = []
    tmp = tempfile.mkdtemp(prefix="dnsref_")
    try:
        for exe, iters in zip(exes, iters_pair):
            wd = os.path.join(tmp, str(iters))
            os.makedirs(wd)
            # restart pair at TSTR = 1 (dns-const-q-xycut.c:405-435): ftemp || f, then the force
            with open(os.path.join(wd, "lbf.0001"), "wb") as fh:
                fh.write(ft.tobytes())
                fh.write(ft.tobytes())
            with open(os.path.join(wd, "force.0001"), "wb") as fh:
                fh.write(np.float64(0.0).tobytes())
            t0 = time.perf_counter()
            p = subprocess.run([mpirun, "-np", str(ranks), exe, "-p"], cwd=wd, capture_output=True, timeout=600)
            secs.append(time.perf_counter() - t0)
            shutil.rmtree(wd, ignore_errors=True)  # the restart and the step-0 outputs
            if p.returncode != 0:
                print(f"# reference baseline run failed (rc {p.returncode})", file=sys.stderr, flush=True)
                return None
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    per_step = (secs[1] - secs[0]) / (iters_pair[1] - iters_pair[0])
    if per_step <= 0:
        return None
    return {"value": X * Y * Z / per_step / 1e6, "unit": "MLUPS", "cores": ranks, "kind": "reference",
            "sample": f"the reference solver (DNS-solver/src, built from its sources, {ranks} MPI rank(s)"
                      f"{', x-split' if ranks > 1 else ''}) on a {X}x{Y}x{Z} slab of the {wname} grid (smooth walls), "
                      f"constant flow rate, from rest: ({iters_pair[1]} - {iters_pair[0]}) steps between two runs of "
                      f"{secs[0]:.2f} s and {secs[1]:.2f} s"}


def measured_traffic(wname):
    """HBM bytes per launch of the step kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of this workload (calibrated against a
    copy of known size), or None when no summary exists."""
    path = os.path.join(HERE, "profiles", f"traffic_{wname}.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh)


SWEEP = "texture_sweep_64"
PNAME = {dnsgpu.FP64: "fp64", dnsgpu.FP32: "fp32", dnsgpu.FP16: "fp16"}

# Drag reduction of a textured channel against the smooth one at the same
# constant body force, FP64 and FP32-shifted from the same seeded field:
# time averages over [avg_from, steps).
DRAG = {
    "drag_retau180_grooves_768x384x192": dict(base="retau180_768x384x192", texture=("grooves", 32, 16),
                                              steps=20000, avg_from=10000),
    "drag_retau395_riblets_1536x768x384": dict(base="retau395_1536x768x384", texture=("riblets", 16, 8),
                                               steps=20000, avg_from=10000),
}


def stripes(X, Y, period, slip):
    """Streamwise-invariant slip stripes on a wall (code -2 slip, -1 no-slip)."""
    import numpy as np
    row = np.where(np.arange(Y) % period < slip, -2, -1).astype(np.int8)
    return np.ascontiguousarray(np.broadcast_to(row[None, :], (X, Y)))


def drag_run(name, prec, textured, steps, avg_from, batch=500):
    """One run of a drag workload: the smooth (textured=False) or textured
    channel at the workload's constant force from the seeded turbulent
    field, averaged over [avg_from, steps): the time averages, the mean
    profile, and the batch means (batch steps each) of the bulk velocity
    and the wall stress."""
    import numpy as np
    d = DRAG[name]
    w = WORKLOADS[d["base"]]
    omega, F, ub, nu = case_params(w)
    X, Y, Z = w["X"], w["Y"], w["Z"]
    ch = dnsgpu.Channel(X, Y, Z, Rebs=ub * Z / nu, ub=ub, CFL=0.1, TSTR=0, forcing=dnsgpu.CONST_F,
                        force=F, precision=prec, omega=omega)
    tag = f"{PNAME[prec]} {'textured' if textured else 'smooth'}"
    try:
        kind, a, b = d["texture"]
        if textured and kind == "grooves":
            m = stripes(X, Y, a, b)
            ch.set_walls(m, m)
        elif textured:
            ch.set_solid_bands(*([riblet_bands(X, Y, a, b)] * 2))
        ch.init_field(dnsgpu.INIT_TURBULENT, uc=w["utau"], amp=0.1, seed=2, nmodes=8)
        ch.autotune(trial_steps=3)
        t0 = time.perf_counter()
        done = 0
        while done < avg_from:
            k = min(2000, avg_from - done)
            ch.step(k, cntr0=done, logs=False)
            done += k
            print(f"# {tag}: {done} steps, {time.perf_counter() - t0:.0f} s", file=sys.stderr, flush=True)
        ch.tavg_reset()
        ub_b, st_b = [], []
        while done < steps:
            k = min(batch, steps - done)
            logs = ch.step(k, cntr0=done, logs=True)
            ch.tavg_sample()
            ub_b.append(float(np.mean(logs["flowrate"])) / (Y * Z))
            st_b.append(float(np.mean(logs["stress"])))
            done += k
            if done % 2000 == 0 or done == steps:
                print(f"# {tag}: {done} steps, {time.perf_counter() - t0:.0f} s", file=sys.stderr, flush=True)
        wall = time.perf_counter() - t0
        tv = ch.tavg(profiles=True)
    finally:
        ch.close()
    tv["profiles"] = tv["profiles"].tolist()
    return {"tavg": tv, "wall_s": wall, "Ub_batches": ub_b, "stress_batches": st_b, "batch_steps": batch,
            "steps": steps, "avg_from": avg_from}


def _batch_stats(x):
    import numpy as np
    x = np.asarray(x, dtype=float)
    n = len(x)
    return float(x.mean()), (float(x.std(ddof=1) / np.sqrt(n)) if n > 1 else float("nan")), n


def run_drag(name, args):
    """Drag reduction of a textured channel against the smooth one at the
    same constant body force, FP64 / FP32 / FP16 storage from the same
    seeded field, averaged over [avg_from, steps).  With --dr-dir each
    finished run is kept as a file there and runs already there are read
    back, so the six runs can be made in several processes; --dr-runs
    selects the runs of this process."""
    import numpy as np
    d = DRAG[name]
    w = WORKLOADS[d["base"]]
    omega, F, ub, nu = case_params(w)
    X, Y, Z = w["X"], w["Y"], w["Z"]
    steps = args.dr_steps or d["steps"]
    avg_from = min(args.dr_avg_from or d["avg_from"], steps - 500)
    want = None
    if args.dr_runs:
        want = {(p, t) for p, t in (r.split(":") for r in args.dr_runs.split(","))}
    runs = {}
    for prec in (dnsgpu.FP64, dnsgpu.FP32, dnsgpu.FP16):
        for textured in (False, True):
            key = (PNAME[prec], "textured" if textured else "smooth")
            path = os.path.join(args.dr_dir, f"{name}.{key[0]}.{key[1]}.json") if args.dr_dir else None
            if path and os.path.exists(path):
                with open(path) as fh:
                    r = json.load(fh)
                if r["steps"] == steps and r["avg_from"] == avg_from:
                    runs[(prec, textured)] = r
                    continue
            if want is not None and key not in want:
                continue
            r = drag_run(name, prec, textured, steps, avg_from)
            runs[(prec, textured)] = r
            if path:
                os.makedirs(args.dr_dir, exist_ok=True)
                with open(path, "w") as fh:
                    json.dump(r, fh)
            print(f"# {key[0]} {key[1]}: Ub={r['tavg']['Ub']:.6e} {r['wall_s']:.1f} s", file=sys.stderr, flush=True)

    h = Z / 2.0
    utau_cfg = np.sqrt(F * h)
    out = {
        "metric": "MLUPS + % HBM roofline (FP32 vs FP64 D3Q19); ensemble solves/h + p99 latency",
        "unit": "MLUPS", "n_gpus": 1, "steps": steps, "warmup": 0, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": name, "grid_xdv_ydv_zdv": [X, Y, Z], "texture": "%s %d %d" % d["texture"],
                   "forcing": "constant body force", "averaging_window": [avg_from, steps],
                   "window_eddy_turnovers": round((steps - avg_from) * utau_cfg / h, 3),
                   "run_eddy_turnovers": round(steps * utau_cfg / h, 3),
                   "drag_reduction": "1 - (Ub_smooth / Ub_textured)^2 at constant force",
                   "sampling_error": "standard error of 500-step batch means of Ub (batches of a developing "
                                     "flow are correlated: a lower bound)",
                   "parallelism": "single GPU"},
        "runs_present": sorted(f"{PNAME[p]}_{'textured' if t else 'smooth'}" for (p, t) in runs),
    }
    if (dnsgpu.FP64, True) in runs:
        r64 = runs[(dnsgpu.FP64, True)]
        out["value"] = round(X * Y * Z * steps / r64["wall_s"] / 1e6, 1)
        out["ms_per_step"] = round(r64["wall_s"] / steps * 1e3, 4)
    out["Ub"] = {f"{PNAME[p]}_{'textured' if t else 'smooth'}": r["tavg"]["Ub"] for (p, t), r in runs.items()}
    out["wall_s"] = {f"{PNAME[p]}_{'textured' if t else 'smooth'}": round(r["wall_s"], 2) for (p, t), r in runs.items()}

    def dr_of(prec):
        if (prec, False) not in runs or (prec, True) not in runs:
            return None
        u0, s0, _ = _batch_stats(runs[(prec, False)]["Ub_batches"])
        u1, s1, n = _batch_stats(runs[(prec, True)]["Ub_batches"])
        dr = dnsgpu.drag_reduction(runs[(prec, False)]["tavg"], runs[(prec, True)]["tavg"], dnsgpu.CONST_F)["DR_Ub"]
        err = 200.0 * (u0 / u1) ** 2 * np.sqrt((s0 / u0) ** 2 + (s1 / u1) ** 2)
        return dr, err, n

    base = dr_of(dnsgpu.FP64)
    t64 = runs.get((dnsgpu.FP64, True), {}).get("tavg")
    for prec in (dnsgpu.FP64, dnsgpu.FP32, dnsgpu.FP16):
        r = dr_of(prec)
        if r is None:
            continue
        key = "fp64" if prec == dnsgpu.FP64 else PNAME[prec] + "_shifted"
        t = runs[(prec, True)]["tavg"]
        ent = {"drag_reduction_percent": round(r[0], 4), "sampling_error_points": round(r[1], 4),
               "batches": r[2], "u_tau_wall_stress": float(np.sqrt(abs(t["stress"]))),
               "slip_velocity": 0.5 * (t["u_slip_bottom"] + t["u_slip_top"])}
        if prec != dnsgpu.FP64 and base is not None and t64 is not None:
            u64 = np.asarray(t64["profiles"][0])
            ut64 = np.sqrt(abs(t64["stress"]))
            slip64 = 0.5 * (t64["u_slip_bottom"] + t64["u_slip_top"])
            ent.update({
                "drag_reduction_error_points": round(abs(r[0] - base[0]), 4),
                "error_over_sampling_error": round(abs(r[0] - base[0]) / base[1], 3) if base[1] > 0 else None,
                "mean_profile_max_rel_error": float(np.abs(np.asarray(t["profiles"][0]) - u64).max()
                                                    / np.abs(u64).max()),
                "u_tau_wall_stress_rel_error": float(abs(np.sqrt(abs(t["stress"])) - ut64) / ut64),
                "slip_velocity_rel_error": (abs(ent["slip_velocity"] - slip64) / abs(slip64)) if slip64 else None,
            })
        out[key] = ent
    return out


def sweep_measure(prec, burst, sequential=True, kernel=True):
    """The serving ensemble texture sweep: 64 seeded solves (driver/
    dnsserve.c: grids, Re_tau, grooves or riblets, Poisson arrivals of mean
    gap 2 s, or all at t = 0 when burst) served through one ensemble until
    the running u_tau settles (0.5% over 2000 steps after a 2000-step
    transient, cap 10000); then, with sequential, the same solves for the
    same step counts one after another on the plain FP64 single-case path;
    with kernel, the multi-case step kernel timed with all 64 resident."""
    import dnsserve
    reqs = dnsserve.generate(64, seed=5, mean_gap=2.0)
    cfg = dnsserve.defaults(realtime=0 if burst else 1, verbose=1, precision=prec)
    res, summ = dnsserve.sweep(reqs, cfg)
    print(f"# sweep done: {summ['wall']:.1f} s, {summ['mlups']:.0f} MLUPS", file=sys.stderr, flush=True)
    steps = [r["steps"] if r["status"] >= 0 else 0 for r in res]
    seq = fifo = None
    if sequential:
        secs = dnsserve.sequential(reqs, cfg, dnsgpu.FP64, steps)
        seq = sum(secs)
        print(f"# sequential fp64 done: {seq:.1f} s", file=sys.stderr, flush=True)
        # the same requests served one at a time in arrival order by the FP64
        # single-case path (a queue alone), from the measured per-solve times
        t, ttr = 0.0, []
        for r, sec in zip(reqs, secs):
            arr = r.arrival if not burst else 0.0
            t = max(t, arr) + sec
            ttr.append(t - arr)
        ttr.sort()
        k = lambda p: ttr[max(0, min(len(ttr) - 1, int(-(-p * len(ttr) // 100)) - 1))]
        fifo = {"wall_s": round(t, 3), "p50_s": round(k(50), 3), "p99_s": round(k(99), 3)}
    nodes = sum(r.xdv * r.ydv * r.zdv for r in reqs)
    bpn = 152 if prec == dnsgpu.FP32 else 76
    status = {str(k): sum(1 for r in res if r["status"] == k) for k in (1, 0, -1, -2)}
    out = {
        "solves": 64, "seed": 5, "storage": PNAME[prec] + "-shifted",
        "arrivals": "all at t=0" if burst else "Poisson, mean gap 2 s",
        "solves_per_hour": round(summ["solves_per_hour"], 1),
        "p50_s": round(summ["p50"], 3),
        "p99_s": round(summ["p99"], 3),
        "mean_time_to_result_s": round(summ["mean_ttr"], 3),
        "wall_s": round(summ["wall"], 3),
        "busy_s": round(summ["busy"], 3),
        "aggregate_mlups": round(summ["mlups"], 1),
        "aggregate_mlups_busy": round(summ["mlups_busy"], 1),
        "max_resident": summ["max_resident"],
        "peak_device_bytes": summ["peak_bytes"],
        "pair_arena_s": round(summ["arena_s"], 3),
        "pair_arena_bytes_per_half": summ["arena_bytes"],
        "status_counts": status,
        "sequential_fp64_s": None if seq is None else round(seq, 3),
        "speedup_vs_sequential_fp64": None if seq is None else round(seq / summ["wall"], 3),
        "sequential_fp64_queue": fifo,
        "grids": "{128,256,512} x {64,128,256} x {64,96,128} (xdv x ydv x zdv)",
        "re_tau": "U[90,180]", "textures": "grooves (period 8-64, gas fraction 0.3-0.9) or riblets "
        "(spacing 8-32, height 4-12)", "convergence": "running u_tau within 0.5% over 2000 steps "
        "after a 2000-step transient, cap 10000",
        "total_steps": int(sum(steps)),
    }
    if kernel:
        pk = pool_kernel_measure(prec)
        out["ms_per_step_all_resident"] = pk["ms_per_step"]
        out["roofline"] = pk["roofline"]
    out["per_solve"] = res
    return out


def pool_kernel_measure(prec, n=10):
    """The serving sweep's dominant kernel: the multi-case step with all 64
    seeded solves resident, their population pairs from the pair arenas as
    in the sweep, autotuned, then n steps timed with HIP events on the
    ensemble's stream (step kernel alone, and the whole step with the record
    reduction and the finishers)."""
    import dnsserve
    reqs = dnsserve.generate(64, seed=5, mean_gap=2.0)
    cfg = dnsserve.defaults(realtime=0, verbose=0, precision=prec)
    nodes = sum(r.xdv * r.ydv * r.zdv for r in reqs)
    bpn = 152 if prec == dnsgpu.FP32 else 76
    esize = 4 if prec == dnsgpu.FP32 else 2
    dnsgpu.reserve_pair_arenas(sum(r.xdv * r.ydv * r.zdv * 19 * esize for r in reqs))
    hs = [dnsserve.Handle(r, cfg) for r in reqs]
    ens = dnsgpu.Ensemble(precision=prec)
    for h in hs:
        ens.add(h)
    blocks, _ = ens.autotune(3)
    ens.step(2)
    tk, tw = ens.time_kernels(n)
    ens.close()
    for h in hs:
        h.close()
    dnsgpu.release_pair_arenas()
    achieved = nodes * bpn / (tk / n / 1e3) / 1e9
    tr = measured_traffic(SWEEP)
    traffic = tr[PNAME[prec]]["traffic_bytes_per_launch"] if tr and PNAME[prec] in tr else None
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
            "traffic_source": f"profiles/traffic_{SWEEP}.json" if traffic else None,
            "kernel": "ens_step_kernel<%s>" % ("float" if prec == dnsgpu.FP32 else "_Float16"),
            "kernel_ms": round(tk / n, 4), "bytes_per_node": bpn, "nodes": nodes, "grid_blocks": blocks,
            "frac_of_copy_ceiling": round(achieved / CEILING_GBS, 4), "copy_ceiling": CEILING_GBS}
    if traffic:
        roof["traffic_gbs"] = round(traffic / (tk / n / 1e3) / 1e9, 1)
    return {"ms_per_step": round(tw / n, 4), "roofline": roof}


def run_pool_only(args):
    """The multi-case step kernel of the 64-solve sweep with every solve
    resident (pool_kernel_measure), as its own line."""
    prec = dnsgpu.FP16 if args.sweep_fp16 else dnsgpu.FP32
    pk = pool_kernel_measure(prec, n=args.steps)
    r = pk["roofline"]
    return {"metric": "multi-case step kernel, 64 solves resident", "value": round(r["nodes"] / r["kernel_ms"] / 1e3, 1),
            "unit": "MLUPS (step kernel)", "n_gpus": 1, "steps": args.steps, "ms_per_step": pk["ms_per_step"],
            "higher_is_better": True, "dtype": "f32" if prec == dnsgpu.FP32 else "f32 (f16 storage)",
            "data": "synthetic", "config": {"workload": SWEEP, "storage": PNAME[prec] + "-shifted"}, "roofline": r}


def run_ens_traffic_summary(args):
    """HBM bytes per dispatch of the multi-case step kernel and the record
    reduction of the 64-solve pool, from rocprofv3 --pmc runs of bench.py
    --pool-only (FETCH_SIZE, WRITE_SIZE) for FP32 and FP16 storage, scaled by
    what the same counters read for the plain copy of exactly 4 GiB (copybw
    4 2).  Argument: fp32 fetch, fp32 write, copy fetch, copy write, fp16
    fetch, fp16 write CSVs.  The step kernel's grid is that of its last
    dispatch (the timed steps, after the autotuner's trials)."""
    import csv
    import statistics
    import dnsserve
    paths = args.traffic_summary.split(",")
    f32f, f32w, cf, cw = paths[:4]
    copy_kib = 4 * 1024 * 1024

    def cal(path):
        vals = [float(r["Counter_Value"]) for r in csv.DictReader(open(path)) if "copy_f4" in r["Kernel_Name"]]
        return copy_kib / statistics.median(vals)

    scale = {"fetch": cal(cf), "write": cal(cw)}
    reqs = dnsserve.generate(64, seed=5, mean_gap=2.0)
    nodes = sum(r.xdv * r.ydv * r.zdv for r in reqs)
    out = {"workload": SWEEP, "nodes": nodes, "fetch_scale": round(scale["fetch"], 4),
           "write_scale": round(scale["write"], 4)}
    pairs = [("fp32", f32f, f32w)] + ([("fp16", paths[4], paths[5])] if len(paths) >= 6 else [])
    for pname, fpath, wpath in pairs:
        per = {}
        for kind, path in (("fetch", fpath), ("write", wpath)):
            sums, name, grid = {}, {}, {}
            for r in csv.DictReader(open(path)):
                kn = r["Kernel_Name"]
                if "ens_step_kernel" not in kn and "ens_reduce_kernel" not in kn:
                    continue
                d = int(r["Dispatch_Id"])
                sums[d] = sums.get(d, 0.0) + float(r["Counter_Value"])
                name[d] = kn
                grid[d] = int(r["Grid_Size"]) // int(r["Workgroup_Size"])
            step = sorted(d for d in sums if "ens_step_kernel" in name[d])
            red = [sums[d] for d in sums if "ens_reduce_kernel" in name[d]]
            g = grid[step[-1]]
            vals = [sums[d] for d in step if grid[d] == g]
            kname = name[step[-1]].split("(")[0].replace("void ", "")
            per[kind] = (g, statistics.median(vals), statistics.median(red) if red else 0.0, kname, len(vals))
        g, f, fr, kname, nd = per["fetch"]
        rb = f * 1024 * scale["fetch"]
        wb = per["write"][1] * 1024 * scale["write"]
        out[pname] = {"kernel": kname, "grid_blocks": g, "dispatches": nd,
                      "read_B_per_node": round(rb / nodes, 3), "write_B_per_node": round(wb / nodes, 3),
                      "traffic_bytes_per_launch": round(rb + wb),
                      "reduce_read_B_per_node": round(fr * 1024 * scale["fetch"] / nodes, 3)}
    return out


def run_sweep(args):
    """The texture sweep as its own line (value = solves/h)."""
    prec = dnsgpu.FP16 if args.sweep_fp16 else dnsgpu.FP32
    m = sweep_measure(prec, args.burst, sequential=not args.no_sequential)
    out = {
        "metric": "MLUPS + % HBM roofline (FP32 vs FP64 D3Q19); ensemble solves/h + p99 latency",
        "value": m["solves_per_hour"],
        "unit": "solves/h",
        "n_gpus": 1,
        "steps": m["total_steps"],
        "warmup": 0,
        "ms_per_step": m.get("ms_per_step_all_resident"),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec == dnsgpu.FP32 else "f32 (f16 storage)",
        "data": "synthetic",
        "config": {"workload": SWEEP, "parallelism": "single GPU, one ensemble"},
        "roofline": m.get("roofline"),
        "serving": m,
    }
    if not args.no_cpu:
        r0 = dnsserve_first_request()
        w = dict(X=r0.xdv, Y=r0.ydv, Z=r0.zdv, retau=None, utau=None)
        out["cpu_baseline"] = cpu_baseline(SWEEP, w=w)
    return out


def dnsserve_first_request():
    import dnsserve
    return dnsserve.generate(64, seed=5, mean_gap=2.0)[0]


def run_z_sweep(args):
    """Step-kernel bandwidth against the wall-normal size Z at about the
    Re_tau 180 node count (Y = 384, X = 56623104 // (384 Z)), smooth walls,
    each precision autotuned, HIP-event kernel times over --steps logged
    steps: how the 64-row chunking and the direction-segment length
    (Z * sizeof(T) bytes) bear on each storage precision."""
    zs = [int(z) for z in args.z_list.split(",")]
    precs = [{"fp64": dnsgpu.FP64, "fp32": dnsgpu.FP32, "fp16": dnsgpu.FP16}[p] for p in args.z_prec.split(",")]
    out = []
    for Z in zs:
        Y = 384
        X = max(2, 56623104 // (Y * Z))
        for prec in precs:
            ch = dnsgpu.Channel(X, Y, Z, Rebs=3000.0, ub=0.1, CFL=0.1, forcing=dnsgpu.CONST_Q, precision=prec)
            try:
                ch.init_field(dnsgpu.INIT_TURBULENT, uc=0.006, amp=0.1, seed=1, nmodes=2)
                ch.step(2, logs=False)
                blocks, _ = ch.autotune(trial_steps=3)
                ms_k, _ = ch.time_steps(args.steps, stats=True)
            finally:
                ch.close()
            esize = {dnsgpu.FP64: 8, dnsgpu.FP32: 4, dnsgpu.FP16: 2}[prec]
            k = ms_k / args.steps
            gbs = 2 * 19 * esize * X * Y * Z / (k / 1e3) / 1e9
            out.append({"Z": Z, "X": X, "Y": Y, "precision": PNAME[prec], "blocks": blocks,
                        "kernel_us": round(k * 1e3, 1), "GBs": round(gbs), "frac": round(gbs / PEAK_HBM_GBS, 4)})
            print(f"# Z={Z} {PNAME[prec]}: {k * 1e3:.0f} us, {gbs:.0f} GB/s", file=sys.stderr, flush=True)
    return {"metric": "step-kernel GB/s against Z", "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
            "data": "synthetic", "z_sweep": out}


def run_rows_sweep(args):
    """FP16 storage, one against two rows per lane (rows z and z + 1 in one
    4-byte access per direction, the autotuner's cells-per-thread axis):
    step-kernel time (statistics variant, HIP events, --steps launches) at
    every waves-per-block x grid shape of the autotuner's search, each shape
    timed twice in alternation, on the Re_tau 395 riblet grid and on smooth
    grids of about 56.6 M nodes with zdv = 192, 96, 160, 128 (the last
    chunk full, a quarter full, half full, full)."""
    cases = [("1536x768x384 riblets s16 h8", 1536, 768, 384, (16, 8)), ("768x384x192 smooth", 768, 384, 192, None),
             ("1536x384x96 smooth", 1536, 384, 96, None), ("920x384x160 smooth", 920, 384, 160, None),
             ("1152x384x128 smooth", 1152, 384, 128, None)]
    out = []
    for name, X, Y, Z, rib in cases:
        ch = dnsgpu.Channel(X, Y, Z, Rebs=3000.0, ub=0.1, CFL=0.1, forcing=dnsgpu.CONST_F, force=1e-6,
                            precision=dnsgpu.FP16)
        try:
            if rib:
                hb = riblet_bands(X, Y, *rib)
                ch.set_solid_bands(hb, hb)
            ch.init_field(dnsgpu.INIT_TURBULENT, uc=0.0054, amp=0.1, seed=2, nmodes=8)
            ch.step(2, logs=False)
            times = {}
            for _ in range(2):
                for rows in (1, 2):
                    for W in (2, 3, 4):
                        for wpc in (8, 12, 16, 24):
                            ch.set_step_shape((256 * wpc + W - 1) // W, W, rows)
                            shape = ch.step_shape(rows=True)
                            ms, _ = ch.time_steps(args.steps, stats=True)
                            times.setdefault(shape, []).append(ms / args.steps)
        finally:
            ch.close()
        best = {}
        for (b, w, r), t in times.items():
            if r not in best or min(t) < best[r][0]:
                best[r] = (min(t), w, b)
        ent = {"grid": name, "nodes": X * Y * Z}
        for r in (1, 2):
            tm, w, b = best[r]
            ent[f"rows{r}"] = {"kernel_ms": round(tm, 4), "waves": w, "blocks": b,
                               "GBs": round(76 * X * Y * Z / (tm / 1e3) / 1e9)}
        ent["two_over_one_speed"] = round(best[1][0] / best[2][0], 3)
        out.append(ent)
        print(f"# {name}: one row {best[1][0]:.3f} ms, two rows {best[2][0]:.3f} ms", file=sys.stderr, flush=True)
    return {"metric": "FP16 step kernel, one against two rows per lane", "unit": "ms", "n_gpus": 1,
            "steps": args.steps, "data": "synthetic", "rows_sweep": out}


def run_fp16_drift(args):
    """Error of the reduced storages against FP64 over the run length, by
    channel height: the x,y-uniform (laminar) channel at the workloads'
    lattice viscosity nu = u_tau h / Re_tau with Re_tau = zdv (u_tau 0.0054,
    so nu = 0.0027 at every zdv), u_b = 15.7 u_tau, on a 4 x 4 x zdv grid,
    started from the parabola of that u_b.  Constant force (the workloads'
    forcing) at the parabola's force F = 12 nu u_b / zdv^2: the FP64 state
    stays on the steady parabola and the storage's error shows as a drift
    of the bulk velocity u_b; constant flow rate (encalc.c:121): as a drift
    of the force, a difference of two numbers near u_b (per step, and its
    mean over the steps since the previous mark).  Per-step changes
    below half a storage ulp of g are lost, and the force per step falls as
    1/zdv^2.  Logged at steps 10^2 .. --drift-steps."""
    import numpy as np
    zs = [int(z) for z in args.drift_z.split(",")]
    marks = [m for m in (100, 300, 1000, 3000, 10000, 30000, 100000, 300000, 1000000) if m <= args.drift_steps]
    precs = (dnsgpu.FP64, dnsgpu.FP32, dnsgpu.FP16)
    out = []
    for Z in zs:
        utau = 0.0054
        nu = utau * (Z / 2.0) / float(Z)
        ub = 15.7 * utau
        F = 12.0 * nu * ub / float(Z * Z)
        omega = 1.0 / (3.0 * nu + 0.5)
        for forcing, fname in ((dnsgpu.CONST_F, "constant force"), (dnsgpu.CONST_Q, "constant flow rate")):
            traj = {}
            for prec in precs:
                ch = dnsgpu.Channel(4, 4, Z, Rebs=ub * Z / nu, ub=ub, CFL=0.1, TSTR=0, forcing=forcing,
                                    force=F if forcing == dnsgpu.CONST_F else 0.0, precision=prec, omega=omega)
                rows = []
                try:
                    ch.init_field(dnsgpu.INIT_POISEUILLE, uc=1.5 * ub, amp=0.0, seed=1)
                    done = 0
                    for m in marks:
                        lg = ch.step(m - done, cntr0=done, logs=True)
                        done = m
                        prof = ch.statistics()[0]  # UAVG
                        # the last step's force and the mean over the steps since the previous mark
                        rows.append((float(lg["force"][-1]), float(lg["flowrate"][-1]) / (4 * Z), prof,
                                     float(np.mean(lg["force"]))))
                finally:
                    ch.close()
                traj[prec] = rows
            for i, m in enumerate(marks):
                f64, u64, p64, fm64 = traj[dnsgpu.FP64][i]
                ent = {"Z": Z, "forcing": fname, "steps": m, "force_fp64": f64, "Ub_fp64": u64,
                       "force_mean_fp64": fm64, "mean_over": [marks[i - 1] if i else 0, m]}
                for prec in (dnsgpu.FP32, dnsgpu.FP16):
                    f, u, pr, fm = traj[prec][i]
                    ent[PNAME[prec]] = {"Ub_rel": abs(u - u64) / abs(u64),
                                        "force_rel": abs(f - f64) / abs(f64) if f64 else None,
                                        "force_mean_rel": abs(fm - fm64) / abs(fm64) if fm64 else None,
                                        "profile_max_rel": float(np.max(np.abs(pr - p64)) / np.max(np.abs(p64)))}
                out.append(ent)
            last = out[-1]
            key = "Ub_rel" if forcing == dnsgpu.CONST_F else "force_mean_rel"
            print(f"# Z={Z} {fname}: after {marks[-1]} steps {key} fp32 {last['fp32'][key]:.2e} fp16 "
                  f"{last['fp16'][key]:.2e}, profile fp32 {last['fp32']['profile_max_rel']:.2e} fp16 "
                  f"{last['fp16']['profile_max_rel']:.2e}", file=sys.stderr, flush=True)
    return {"metric": "relative error against FP64 of the bulk velocity (constant force), of the force "
            "(constant flow rate) and of the mean profile", "unit": "1", "n_gpus": 1, "data": "synthetic",
            "grid": "4 x 4 x zdv, laminar", "drift": out}


def run_placement_sweep(args):
    """Step-kernel time of fresh handles of one workload with the autotuner's
    population pair placement on and off, alternating (DNSGPU_PAIR_PLACEMENT),
    each handle autotuned and timed on its own buffers."""
    w = WORKLOADS[args.workload]
    omega, F, ub, nu = case_params(w)
    X, Y, Z = w["X"], w["Y"], w["Z"]
    pmap = {"fp64": dnsgpu.FP64, "fp32": dnsgpu.FP32, "fp16": dnsgpu.FP16}
    esize = {"fp64": 8, "fp32": 4, "fp16": 2}
    out = {"workload": args.workload, "handles": []}
    for pname in args.precisions.split(","):
        for i in range(args.placement_handles):
            on = i % 2 == 0
            os.environ["DNSGPU_PAIR_PLACEMENT"] = "1" if on else "0"
            ch = dnsgpu.Channel(X, Y, Z, Rebs=ub * Z / nu, ub=ub, CFL=0.1, TSTR=0, forcing=dnsgpu.CONST_F,
                                force=F, precision=pmap[pname], omega=omega)
            try:
                if w.get("riblets"):
                    ch.set_solid_bands(*([riblet_bands(X, Y, *w["riblets"])] * 2))
                ch.init_field(dnsgpu.INIT_TURBULENT, uc=w["utau"], amp=0.1, seed=2, nmodes=8)
                t0 = time.perf_counter()
                ch.autotune(trial_steps=5)
                t_tune = time.perf_counter() - t0
                tried, gain = ch.pair_placement()
                shape = ch.step_shape(rows=True)
                ch.step(3, logs=False)
                ms = min(ch.time_steps(10, stats=True)[0], ch.time_steps(10, stats=True)[0]) / 10
            finally:
                ch.close()
            rec = {"precision": pname, "placement": on, "destinations_tried": tried, "step_gain": round(gain, 4),
                   "shape": list(shape), "autotune_s": round(t_tune, 2), "kernel_ms": round(ms, 3),
                   "gbs": round(2 * 19 * esize[pname] * X * Y * Z / (ms / 1e3) / 1e9, 1)}
            out["handles"].append(rec)
            print("# " + json.dumps(rec), file=sys.stderr, flush=True)
    os.environ.pop("DNSGPU_PAIR_PLACEMENT", None)
    return out


def _step_variant(name):
    """(precision, FIRST, STATS, rows per lane, z-pieces) of a step_kernel
    dispatch name as rocprofv3 writes it (demangled, or mangled for the FP16
    instances), or None for any other kernel.  z-pieces is True for the
    instance that splits columns into z-pieces (its launch's grid then holds
    one group of blocks per piece)."""
    import re
    m = re.match(r"void lbm::step_kernel<(\w+), (\w+), (\w+), (\w+), (\d)(?:, (\w+))?>", name)
    if m:
        t, first, stats, _, rows, pieces = m.groups()
        return ({"double": "fp64", "float": "fp32", "_Float16": "fp16"}[t], first == "true", stats == "true",
                int(rows), pieces == "true")
    m = re.match(r"_ZN3lbmL11step_kernelI(d|f|DF16_)Lb(\d)ELb(\d)ELb(\d)ELi(\d)E(?:Lb(\d)E)?", name)
    if m:
        t, first, stats, _, rows, pieces = m.groups()
        return ({"d": "fp64", "f": "fp32", "DF16_": "fp16"}[t], first == "1", stats == "1", int(rows),
                pieces == "1")
    return None


def run_traffic_summary(args):
    """HBM bytes per step-kernel dispatch from two rocprofv3 --pmc runs of
    bench.py (FETCH_SIZE, then WRITE_SIZE; both with the same
    DNSGPU_TUNE_CACHE, so that they run the same launch shapes) and the same
    counters over the plain copy of exactly 4 GiB (copybw 4 2): FETCH_SIZE and
    WRITE_SIZE are scaled by what they read for that copy.  Grouped by
    kernel variant and launch shape; the shape of a variant's last dispatch
    is the autotuner's choice (the logged steps).  Prints the traffic record
    bench.py reads (profiles/traffic_<workload>.json)."""
    import csv
    import statistics
    fetch_csv, write_csv, cal_fetch_csv, cal_write_csv = args.traffic_summary.split(",")
    copy_kib = 4 * 1024 * 1024

    def cal(path):
        vals = [float(r["Counter_Value"]) for r in csv.DictReader(open(path)) if "copy_f4" in r["Kernel_Name"]]
        return copy_kib / statistics.median(vals)

    scale = {"fetch": cal(cal_fetch_csv), "write": cal(cal_write_csv)}
    w = WORKLOADS[args.workload]
    nodes = w["X"] * w["Y"] * w["Z"]
    sums, last = {}, {}
    for kind, path in (("fetch", fetch_csv), ("write", write_csv)):
        for r in sorted(csv.DictReader(open(path)), key=lambda r: int(r["Dispatch_Id"])):
            v = _step_variant(r["Kernel_Name"])
            if v is None:
                continue
            pname, first, stats, rows, pieces = v
            key = f"{pname}_{'first_' if first else ''}{'stats' if stats else 'nostats'}"
            wg = int(r["Workgroup_Size"])
            shape = f"{int(r['Grid_Size']) // wg}x{wg}" + ("x2rows" if rows == 2 else "") + \
                ("xpieces" if pieces else "")
            ent = sums.setdefault(key, {}).setdefault(shape, {"fetch": [], "write": []})
            ent[kind].append(float(r["Counter_Value"]))
            last[(key, kind)] = shape  # the logged steps come last
    out = {"workload": args.workload, "nodes": nodes, "fetch_scale": round(scale["fetch"], 4),
           "write_scale": round(scale["write"], 4)}
    for key in sorted(sums):
        by = {}
        for shape, ent in sums[key].items():
            if not ent["fetch"] or not ent["write"]:
                continue
            f = statistics.mean(ent["fetch"])
            wr = statistics.mean(ent["write"])
            rb = f * 1024 * scale["fetch"]
            wb = wr * 1024 * scale["write"]
            by[shape] = {"fetch_kib": round(f), "write_kib": round(wr), "read_B_per_node": round(rb / nodes, 3),
                         "write_B_per_node": round(wb / nodes, 3), "traffic_bytes_per_launch": round(rb + wb),
                         "dispatches": min(len(ent["fetch"]), len(ent["write"]))}
        if not by:
            continue
        chosen = last.get((key, "fetch"))
        if chosen not in by:
            chosen = max(by, key=lambda k: by[k]["dispatches"])
        out[key] = dict(chosen_shape=chosen, **{k: v for k, v in by[chosen].items() if k != "dispatches"},
                        by_shape=by)
    return out


# The read+write ceiling of this device (a plain copy between buffers in
# different halves of device memory, profiles/copybw_memory_halves.md)
CEILING_GBS = 5550.0


def roofline_of(r, traffic, kernel):
    out = {"bound": "hbm", "achieved": round(r["achieved_gbs"], 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
           "frac": round(r["achieved_gbs"] / PEAK_HBM_GBS, 4), "traffic": traffic, "kernel": kernel,
           "kernel_ms": round(r["kernel_ms"], 4), "bytes_per_node": r["bytes_per_node"],
           # the same kernel against the device's measured read+write ceiling,
           # and the rate of the bytes it moves by the HBM counters
           "frac_of_copy_ceiling": round(r["achieved_gbs"] / CEILING_GBS, 4), "copy_ceiling": CEILING_GBS,
           "copy_ceiling_source": "profiles/copybw_memory_halves.md"}
    if traffic:
        out["traffic_gbs"] = round(traffic / (r["kernel_ms"] / 1e3) / 1e9, 1)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="retau395_riblets_1536x768x384",
                    choices=sorted(WORKLOADS) + [SWEEP] + sorted(DRAG))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serving", action="store_true", help="single-case workloads: skip the serving sweep")
    ap.add_argument("--precisions", default="fp32,fp64,fp16",
                    help="single-case workloads: storage precisions to run (fp32 is the value line)")
    ap.add_argument("--burst", action="store_true", help=f"{SWEEP}: every request at t = 0")
    ap.add_argument("--no-sequential", action="store_true", help=f"{SWEEP}: skip the sequential FP64 solves")
    ap.add_argument("--sweep-fp16", action="store_true", help=f"{SWEEP}: FP16-shifted storage")
    ap.add_argument("--pool-only", action="store_true",
                    help=f"{SWEEP}: only the multi-case step kernel with all 64 solves resident (--steps timed)")
    ap.add_argument("--dr-steps", type=int, default=0, help="drag workloads: steps per run")
    ap.add_argument("--dr-avg-from", type=int, default=0, help="drag workloads: first averaged step")
    ap.add_argument("--dr-runs", default="", help="drag workloads: runs of this process, e.g. fp64:smooth,fp64:textured")
    ap.add_argument("--dr-dir", default="", help="drag workloads: directory keeping each finished run")
    ap.add_argument("--z-sweep", action="store_true", help="step-kernel GB/s against Z (see run_z_sweep)")
    ap.add_argument("--rows-sweep", action="store_true", help="FP16 one against two rows per lane "
                    "(see run_rows_sweep)")
    ap.add_argument("--fp16-drift", action="store_true", help="reduced-storage error against FP64 over steps, "
                    "by zdv (see run_fp16_drift)")
    ap.add_argument("--drift-z", default="64,128,192,256,384,512")
    ap.add_argument("--drift-steps", type=int, default=100000)
    ap.add_argument("--placement-sweep", action="store_true", help="step time of fresh handles with the "
                    "population pair placement on and off (see run_placement_sweep)")
    ap.add_argument("--placement-handles", type=int, default=6)
    ap.add_argument("--traffic-summary", default="", help="FETCH.csv,WRITE.csv,COPY_FETCH.csv,COPY_WRITE.csv: "
                    "the traffic record from rocprofv3 counter runs (see run_traffic_summary)")
    ap.add_argument("--z-list", default="64,96,128,192,384")
    ap.add_argument("--z-prec", default="fp64,fp32,fp16")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs the whole channel on one GPU; multi-GPU runs are not part of it "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)
    if args.traffic_summary:
        summ = run_ens_traffic_summary(args) if args.workload == SWEEP else run_traffic_summary(args)
        print(json.dumps(summ, indent=1))
        return
    # the CPU baselines run before this process touches the GPU (the
    # reference runs as a child process)
    cpu = cpu_ref = cpu_mpi = None
    if not args.no_cpu and args.workload in WORKLOADS and not (args.z_sweep or args.fp16_drift or args.rows_sweep
                                                               or args.placement_sweep):
        cpu_ref = reference_baseline(args.workload)
        print("# cpu reference (1 rank) done", file=sys.stderr, flush=True)
        cpu_mpi = reference_baseline(args.workload, ranks=REF_MPI_RANKS)
        print(f"# cpu reference ({REF_MPI_RANKS} ranks) done", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.workload, budget_s=8.0)
        print("# cpu oracle done", file=sys.stderr, flush=True)
    dnsgpu.load()
    if args.z_sweep:
        print(json.dumps(run_z_sweep(args)))
        return
    if args.fp16_drift:
        print(json.dumps(run_fp16_drift(args)))
        return
    if args.rows_sweep:
        print(json.dumps(run_rows_sweep(args)))
        return
    if args.placement_sweep:
        print(json.dumps(run_placement_sweep(args)))
        return
    if args.workload == SWEEP:
        if args.pool_only:
            print(json.dumps(run_pool_only(args)))
        else:
            print(json.dumps(run_sweep(args)))
        return
    if args.workload in DRAG:
        print(json.dumps(run_drag(args.workload, args)))
        return
    precs = [{"fp64": dnsgpu.FP64, "fp32": dnsgpu.FP32, "fp16": dnsgpu.FP16}[p] for p in args.precisions.split(",")]
    if dnsgpu.FP32 not in precs:
        precs = [dnsgpu.FP32] + precs
    runs = {}
    for prec in precs:
        runs[prec] = run_precision(args.workload, prec, args.steps, args.warmup)
        print(f"# {PNAME[prec]} done: {runs[prec]['mlups']:.0f} MLUPS", file=sys.stderr, flush=True)
    w = WORKLOADS[args.workload]
    tr = measured_traffic(args.workload) or {}

    def traffic_of(p):
        # the counters of the launch shape (blocks x threads) this run's
        # autotuner chose, when that shape was profiled, else those of the
        # profiled run's own choice
        ent = tr.get(f"{PNAME[p]}_stats") or {}
        shape = f"{runs[p]['step_blocks']}x{64 * runs[p]['step_waves_per_block']}"
        if runs[p]["step_rows_per_lane"] == 2:
            shape += "x2rows"
        if runs[p].get("step_z_pieces", 1) > 1:
            shape += "xpieces"
        by = (ent.get("by_shape") or {}).get(shape)
        return (by or ent).get("traffic_bytes_per_launch"), shape if by else ent.get("chosen_shape")

    traffic_shape = {p: traffic_of(p) for p in runs}
    traffic = {p: t for p, (t, _) in traffic_shape.items()}
    tname = {dnsgpu.FP64: "double", dnsgpu.FP32: "float", dnsgpu.FP16: "_Float16"}
    bands = "true" if w.get("riblets") else "false"
    kname = {p: f"step_kernel<{tname[p]},false,true,{bands}" +
             (",2>" if runs[p]["step_rows_per_lane"] == 2 else ",1,true>" if runs[p].get("step_z_pieces", 1) > 1
              else ">")
             for p in runs}
    r32 = runs[dnsgpu.FP32]
    r64 = runs.get(dnsgpu.FP64)
    out = {
        "metric": "MLUPS + % HBM roofline (FP32 vs FP64 D3Q19); ensemble solves/h + p99 latency",
        "value": round(r32["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r32["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload, "grid_xdv_ydv_zdv": [w["X"], w["Y"], w["Z"]],
                   "fluid_nodes_incl_solid": r32["nodes"], "forcing": "constant body force",
                   "texture": ("riblets spacing %d height %d, both walls" % w["riblets"]) if w.get("riblets")
                   else "smooth",
                   "storage": "fp32-shifted populations (f - w), fp32 collision, fp64 sums",
                   "initial_field": "log-law mean + curl of 8x8 Fourier modes, rms 10% of the centreline, seed 2",
                   "per_step": "fused collide-stream + per-step log reductions (encalc/wshearstr)",
                   "parallelism": "single GPU"},
        "roofline": dict(roofline_of(r32, traffic[dnsgpu.FP32], kname[dnsgpu.FP32]),
                         traffic_source=f"profiles/traffic_{args.workload}.json" if traffic[dnsgpu.FP32] else None,
                         traffic_shape=traffic_shape[dnsgpu.FP32][1]),
        "speedup_fp32_vs_fp64": round(r32["mlups"] / r64["mlups"], 3) if r64 else None,
        "fp32_shifted": r32,
    }
    if r64:
        out["fp64"] = dict(r64, roofline=roofline_of(r64, traffic[dnsgpu.FP64], kname[dnsgpu.FP64]))
    if dnsgpu.FP16 in runs:
        r16 = runs[dnsgpu.FP16]
        out["fp16_shifted"] = dict(r16, speedup_vs_fp64=round(r16["mlups"] / r64["mlups"], 3) if r64 else None,
                                   roofline=roofline_of(r16, traffic[dnsgpu.FP16], kname[dnsgpu.FP16]))
    # the reference itself when its binaries are here, else the oracle (a
    # port of the same loop, bitwise equal to it); both are reported
    if cpu_ref:
        out["cpu_baseline"] = cpu_ref
        if cpu:
            out["cpu_baseline_port"] = cpu
    elif cpu:
        out["cpu_baseline"] = cpu
    if cpu_mpi:
        out["cpu_baseline_mpi"] = cpu_mpi
    if not args.no_serving:
        # the serving figure: the FP32 sweep with every request at t = 0
        # against the same solves one after another on the plain FP64 path,
        # and the roofline of its dominant kernel (the pool, every solve resident)
        m = sweep_measure(dnsgpu.FP32, burst=True, sequential=True, kernel=True)
        m.pop("per_solve")
        out["serving"] = m
    print(json.dumps(out))


if __name__ == "__main__":
    main()
