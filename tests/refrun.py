"""Drive the reference solver binary built from source by oracle/build_ref.sh.

Test infrastructure: writes a restart pair lbf.TSTR / force.TSTR in the
reference's checkpoint layout (DNS-solver/src/dns-const-q-xycut.c:405-435),
runs ``mpirun -np 1 run-dns -p`` (one rank: the only layout that integrates
the global periodic problem) and reads its outputs back.
"""
import os
import shutil
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SRC = os.environ.get("DNS_REFERENCE_SRC", "/root/reference/DNS-solver/src")
MPIRUN = os.environ.get("MPIRUN", "/opt/conda/bin/mpirun")

LOG_FILES = {
    "K": "energy.txt", "flowrate": "flowrate.txt", "massrate": "massrate.txt",
    "stup": "str-up.txt", "stdown": "str-do.txt", "stress": "w-str.txt", "pgrad": "p-grad.txt",
}


def available():
    return os.path.isdir(REF_SRC) and os.path.exists(MPIRUN)


def build(name, xdv, ydv, zdv, Rebs, ub, CFL, TSTR, TEND, ITER, slip=""):
    exe = os.path.join(REPO, "oracle", "_ref", name, "run-dns")
    args = [os.path.join(REPO, "oracle", "build_ref.sh"), name, str(xdv), str(ydv), str(zdv),
            repr(float(Rebs)), repr(float(ub)), repr(float(CFL)), str(TSTR), str(TEND), str(ITER)]
    if slip:
        args.append(slip)
    defs = os.path.join(REPO, "oracle", "_ref", name, "definitions.h")
    if not (os.path.exists(exe) and os.path.exists(defs) and _defs_match(defs, args)):
        subprocess.run(args, check=True, capture_output=True)
    return exe


def _defs_match(defs, args):
    txt = open(defs).read()
    name, xdv, ydv, zdv, Rebs, ub, CFL, TSTR, TEND, ITER = args[1:11]
    want = {"xdv": xdv, "ydv": ydv, "zdv": zdv, "Rebs": Rebs, "ub": ub, "CFL": CFL,
            "TSTR": TSTR, "TEND": TEND, "ITER": ITER}
    for k, v in want.items():
        if f"#define {k} {v}\n" not in txt:
            return False
    slip = len(args) > 11
    return ("\n#define PSLIP" in txt) == slip


def write_restart(workdir, TSTR, ftemp, f, force=0.0):
    os.makedirs(workdir, exist_ok=True)
    with open(os.path.join(workdir, "lbf.%.4d" % TSTR), "wb") as fh:
        fh.write(np.ascontiguousarray(ftemp, dtype=np.float64).tobytes())
        fh.write(np.ascontiguousarray(f, dtype=np.float64).tobytes())
    with open(os.path.join(workdir, "force.%.4d" % TSTR), "wb") as fh:
        fh.write(np.float64(force).tobytes())


def run(exe, workdir, ftemp, f, TSTR, force=0.0, timeout=600):
    if os.path.isdir(workdir):
        shutil.rmtree(workdir)
    write_restart(workdir, TSTR, ftemp, f, force)
    p = subprocess.run([MPIRUN, "-np", "1", exe, "-p"], cwd=workdir, capture_output=True,
                       text=True, timeout=timeout)
    if p.returncode != 0:
        raise RuntimeError(f"reference run failed rc={p.returncode}\n{p.stdout}\n{p.stderr}")
    return workdir


def read_lbf(path, X, Y, Z):
    a = np.fromfile(path, dtype=np.float64)
    n = X * Y * (Z + 2) * 19
    assert a.size == 2 * n, (a.size, n)
    return a[:n].reshape(X, Y, Z + 2, 19), a[n:].reshape(X, Y, Z + 2, 19)


def read_logs(workdir):
    out = {}
    for k, fn in LOG_FILES.items():
        d = np.loadtxt(os.path.join(workdir, fn), ndmin=2)
        out["t"] = d[:, 0]
        out[k] = d[:, 1]
    return out
