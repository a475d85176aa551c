"""The reference's own main (DNS-solver/src/dns-const-q-xycut.c, with its
encalc, wshearstr, statistics, vel_print and output_twoD, compiled from its
sources) linked against the GPU step through xcomputations' own signature
(dns-solvers_amd/driver/xcomputations_gpu.c, oracle/build_ref.sh with
REF_SHIM=1).  Started from the golden runs' restart pair, every output it
writes is byte for byte the reference's own run's (tests/golden): the
seven logs, turb-fld at each cadence point, the checkpoint, vel/den,
uavg-temp.dat and 2dresults.  Constant flow rate: the force each step uses
is the reference encalc's own V.force.

The binaries are built in the container that holds the reference sources
(build() of the package entry point); the test is skipped where they were
not built."""
import hashlib
import os

import numpy as np
import pytest

import golden_util as gu
import refrun

pytestmark = pytest.mark.gpu
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("name", gu.CASES)
def test_reference_main_on_the_gpu_step_writes_the_reference_outputs(tmp_path, name):
    exe = os.path.join(REPO, "oracle", "_ref", "shim_" + name, "run-dns-gpu")
    if not os.path.exists(exe):
        pytest.skip("run-dns-gpu not built (reference sources absent where build() ran)")
    if not os.path.exists(refrun.MPIRUN):
        pytest.skip("mpirun absent")
    g = gu.load(name)
    p = gu.params(g)
    TSTR, N, ut, Z = p["TSTR"], p["N"], p["utime"], p["Z"]
    f0 = gu.initial_state(p)
    print(f"case: reference main + GPU step, {name}", flush=True)
    wd = refrun.run(exe, str(tmp_path / name), f0, f0, TSTR, timeout=300)
    for key, fn in refrun.LOG_FILES.items():
        assert open(os.path.join(wd, fn), "rb").read() == g["log_raw_" + key].tobytes(), fn
    for m in range(N // ut + 1):
        st = np.fromfile(os.path.join(wd, "turb-fld.000.%.3d" % (TSTR + m)), dtype=np.float64)
        assert np.array_equal(st.reshape(9, Z + 2), g["turbfld"][m]), m
    last = TSTR + N // ut
    raw = open(os.path.join(wd, "lbf.%.4d" % last), "rb").read()
    assert hashlib.sha256(raw).hexdigest() == str(g["lbf_sha256"])
    for fn, key in (("vel", "vel_sha256"), ("den", "den_sha256")):
        data = np.fromfile(os.path.join(wd, "%s.%.4d" % (fn, last)), dtype=np.float64)
        assert hashlib.sha256(data.tobytes()).hexdigest() == str(g[key]), fn
    for fn in ("2dresults-exact.dat", "2dresults-lb.dat", "uavg-temp.dat"):
        assert open(os.path.join(wd, fn), "rb").read() == g["raw_" + fn.replace("-", "_").replace(".", "_")].tobytes()
