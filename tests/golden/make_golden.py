#!/usr/bin/env python3
"""Write the golden fixtures: outputs of the reference solver itself.

Builds DNS-solver from its own sources (oracle/build_ref.sh), starts it with
`-p` from a seeded restart pair written in its checkpoint layout, and keeps
its outputs as small fixtures in tests/golden/*.npz:
  * the seven per-step logs, as parsed numbers and as raw bytes
    (energy.txt ... p-grad.txt, encalc.c:125-138, wshearstr.c:133-154),
  * every turb-fld.000.NNN written at cadence (statistics.c:137-147),
  * the checkpoint lbf after N steps (dns-const-q-xycut.c:491-503): in full
    for the small grids, as a strided sample plus SHA-256 for the large one,
  * vel/den of the last cadence output (vel_print.c), full or sampled.

The initial state is the library's seeded POISEUILLE generator
(tests/fields.device_poiseuille is its bit-exact replica), so the GPU side
regenerates it on the device without these files carrying it.

Run in a container that has the reference sources and MPICH:
    python tests/golden/make_golden.py
"""
import hashlib
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import fields  # noqa: E402
import refrun  # noqa: E402

# name: grid, physics, cadence, steps, walls, initial field, storage
CASES = {
    # 8x8x8 smooth walls, constant Q, utime = 16 steps, 160 steps
    "smooth8": dict(X=8, Y=8, Z=8, Rebs=60.0, ub=0.05, CFL=0.25, TSTR=100, N=160, slip="",
                    uc=0.06, amp=0.05, seed=7, full=True),
    # same with the reference's slip posts (PSLIP, slip_posts_init.c)
    "posts8": dict(X=8, Y=8, Z=8, Rebs=60.0, ub=0.05, CFL=0.25, TSTR=100, N=160, slip="PSLIP",
                   uc=0.06, amp=0.05, seed=7, full=True),
    # posts on a taller grid: wall-normal extent above one 64-row wave segment
    "posts8x8x70": dict(X=8, Y=8, Z=70, Rebs=200.0, ub=0.05, CFL=3.5, TSTR=0, N=100, slip="PSLIP",
                        uc=0.06, amp=0.05, seed=11, full=False),
    # the laminar 64x32x32 channel: nu = 0.01, Poiseuille bulk velocity of
    # F = 1e-6 (ub = F H^2 / (12 nu)), utime = 50 steps, 500 steps, constant Q
    "laminar64x32x32": dict(X=64, Y=32, Z=32, Rebs=0.008533333333333333 * 32 / 0.01,
                            ub=0.008533333333333333, CFL=0.32, TSTR=0, N=500, slip="",
                            uc=0.0128, amp=0.05, seed=1, full=False),
}

SAMPLE_STRIDE = 997


def utime_of(c):
    return int(0.5 * c["Z"] / c["CFL"])


def run_case(name, c, workroot):
    X, Y, Z = c["X"], c["Y"], c["Z"]
    ut = utime_of(c)
    assert c["N"] % (10 * ut) == 0, "N must land on a checkpoint"
    exe = refrun.build("golden_" + name, X, Y, Z, c["Rebs"], c["ub"], c["CFL"], c["TSTR"], c["TSTR"] + 1,
                       c["N"], c["slip"])
    f0 = fields.device_poiseuille(X, Y, Z, c["uc"], c["amp"], c["seed"])
    wd = refrun.run(exe, os.path.join(workroot, name), f0, f0, c["TSTR"])
    out = {k: np.array(v) for k, v in c.items() if k != "slip"}
    out["slip"] = np.array(c["slip"] == "PSLIP")
    out["utime"] = np.array(ut)
    for key, fn in refrun.LOG_FILES.items():
        raw = open(os.path.join(wd, fn), "rb").read()
        out["log_raw_" + key] = np.frombuffer(raw, dtype=np.uint8)
        d = np.loadtxt(os.path.join(wd, fn), ndmin=2)
        out["log_" + key] = d[:, 1]
        out["log_t"] = d[:, 0]
    ncad = c["N"] // ut + 1
    tf = []
    for m in range(ncad):
        p = os.path.join(wd, "turb-fld.000.%.3d" % (c["TSTR"] + m))
        tf.append(np.fromfile(p, dtype=np.float64).reshape(9, Z + 2))
    out["turbfld"] = np.array(tf)
    lbf = os.path.join(wd, "lbf.%.4d" % (c["TSTR"] + c["N"] // ut))
    raw = open(lbf, "rb").read()
    out["lbf_sha256"] = np.array(hashlib.sha256(raw).hexdigest())
    ft, f = refrun.read_lbf(lbf, X, Y, Z)
    last = c["TSTR"] + c["N"] // ut
    vel = np.fromfile(os.path.join(wd, "vel.%.4d" % last), dtype=np.float64)
    den = np.fromfile(os.path.join(wd, "den.%.4d" % last), dtype=np.float64)
    out["vel_sha256"] = np.array(hashlib.sha256(vel.tobytes()).hexdigest())
    out["den_sha256"] = np.array(hashlib.sha256(den.tobytes()).hexdigest())
    if c["full"]:
        out["lbf_ftemp"] = ft
        out["lbf_f"] = f
        out["vel"] = vel
        out["den"] = den
    else:
        idx = np.arange(0, ft.size, SAMPLE_STRIDE)
        out["sample_idx"] = idx
        out["lbf_ftemp_sample"] = ft.reshape(-1)[idx]
        out["lbf_f_sample"] = f.reshape(-1)[idx]
        vidx = np.arange(0, vel.size, SAMPLE_STRIDE)
        out["vel_sample_idx"] = vidx
        out["vel_sample"] = vel[vidx]
        out["den_sample"] = den[np.arange(0, den.size, SAMPLE_STRIDE)]
    for fn in ("2dresults-exact.dat", "2dresults-lb.dat", "uavg-temp.dat"):
        raw = open(os.path.join(wd, fn), "rb").read()
        out["raw_" + fn.replace("-", "_").replace(".", "_")] = np.frombuffer(raw, dtype=np.uint8)
    return out


def main():
    if not refrun.available():
        sys.exit("reference sources or mpirun not available")
    # the reference's run directories (restart pair, logs, fields) are
    # scratch: only the fixtures under tests/golden are kept
    with tempfile.TemporaryDirectory(prefix="dns_golden_") as work:
        for name, c in CASES.items():
            out = run_case(name, c, work)
            path = os.path.join(HERE, name + ".npz")
            np.savez_compressed(path, **out)
            print(name, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
