"""The C host driver (dns-solvers_amd/bin/dnsrun) end to end on the GPU:
restart from a seeded lbf pair exactly as the reference is started in the
golden runs, then every output file against the reference's own outputs.
Bars as in test_gpu_golden.py (constant-Q: 1e-12 on populations, 1e-10 on
statistics and logs); 2dresults-exact.dat and the file set byte-exact."""
import os
import subprocess

import numpy as np
import pytest

import golden_util as gu

pytestmark = pytest.mark.gpu
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DNSRUN = os.path.join(REPO, "dns-solvers_amd", "bin", "dnsrun")


def write_case(path, p, extra=""):
    txt = (f"Rebs = {p['Rebs']!r}\nub = {p['ub']!r}\nxdv = {p['X']}\nydv = {p['Y']}\nzdv = {p['Z']}\n"
           f"TSTR = {p['TSTR']}\nTEND = {p['TSTR'] + 1}\nCFL = {p['CFL']!r}\nITER = {p['N']}\n"
           f"PSLIP = {1 if p['slip'] else 0}\n" + extra)
    open(path, "w").write(txt)


def run(args, cwd, timeout=120):
    print("case: dnsrun", args, flush=True)
    return subprocess.run([DNSRUN] + args, cwd=cwd, capture_output=True, text=True, timeout=timeout)


@pytest.mark.parametrize("name", ["posts8", "laminar64x32x32"])
def test_driver_restart_matches_reference_outputs(tmp_path, name):
    g = gu.load(name)
    p = gu.params(g)
    X, Y, Z, TSTR, N, ut = p["X"], p["Y"], p["Z"], p["TSTR"], p["N"], p["utime"]
    f0 = gu.initial_state(p)
    with open(tmp_path / ("lbf.%.4d" % TSTR), "wb") as fh:
        fh.write(f0.tobytes())
        fh.write(f0.tobytes())
    with open(tmp_path / ("force.%.4d" % TSTR), "wb") as fh:
        fh.write(np.float64(0.0).tobytes())
    write_case(tmp_path / "dns.case", p)
    r = run(["-p"], tmp_path)
    assert r.returncode == 0, r.stderr
    # logs: N+1 rows each
    for k, fn in (("K", "energy.txt"), ("flowrate", "flowrate.txt"), ("massrate", "massrate.txt"),
                  ("stup", "str-up.txt"), ("stdown", "str-do.txt"), ("stress", "w-str.txt"),
                  ("pgrad", "p-grad.txt")):
        d = np.loadtxt(tmp_path / fn, ndmin=2)
        assert d.shape[0] == N + 1
        assert np.array_equal(d[:, 0], g["log_t"])
        ref = g["log_" + k]
        # both sides printed: one unit in the last printed decimal may flip
        tol = 1e-10 * np.abs(ref).max() + (1.01e-12 if k != "K" else 1.01e-14)
        assert np.abs(d[:, 1] - ref).max() <= tol, k
    # turb-fld at every cadence point
    for m in range(N // ut + 1):
        st = np.fromfile(tmp_path / ("turb-fld.000.%.3d" % (TSTR + m)), dtype=np.float64).reshape(9, Z + 2)
        ref = g["turbfld"][m]
        for q in range(9):
            assert np.abs(st[q] - ref[q]).max() <= 1e-10 * np.abs(ref[q]).max() + 1e-300
    # checkpoint after N steps
    lbf = np.fromfile(tmp_path / ("lbf.%.4d" % (TSTR + N // ut)), dtype=np.float64)
    n = X * Y * (Z + 2) * 19
    assert lbf.size == 2 * n
    ft, f = lbf[:n], lbf[n:]
    if "lbf_ftemp" in g:
        ref = g["lbf_ftemp"].reshape(-1)
        assert np.abs(ft - ref).max() <= 1e-12 * np.abs(ref).max()
    else:
        idx = g["sample_idx"]
        assert np.abs(ft[idx] - g["lbf_ftemp_sample"]).max() <= 1e-12 * np.abs(g["lbf_ftemp_sample"]).max()
        assert np.abs(f[idx] - g["lbf_f_sample"]).max() <= 1e-12 * np.abs(g["lbf_ftemp_sample"]).max()
    assert os.path.exists(tmp_path / ("force.%.4d" % (TSTR + N // ut)))
    vel = np.fromfile(tmp_path / ("vel.%.4d" % (TSTR + N // ut)), dtype=np.float64)
    if "vel" in g:
        assert np.abs(vel - g["vel"]).max() <= 1e-10 * np.abs(g["vel"]).max()
    else:
        vi = g["vel_sample_idx"]
        assert np.abs(vel[vi] - g["vel_sample"]).max() <= 1e-10 * np.abs(g["vel_sample"]).max()
    assert open(tmp_path / "2dresults-exact.dat", "rb").read() == g["raw_2dresults_exact_dat"].tobytes()
    lb = np.loadtxt(tmp_path / "2dresults-lb.dat")
    rl = np.loadtxt(__import__("io").BytesIO(g["raw_2dresults_lb_dat"].tobytes()))
    assert np.abs(lb - rl).max() <= 1e-10 * np.abs(rl).max() + 1e-12
    assert "Total computational time is:" in open(tmp_path / "report.code").read()


def test_driver_rest_start_and_constant_force(tmp_path):
    p = dict(Rebs=60.0, ub=0.05, X=16, Y=8, Z=10, TSTR=0, CFL=0.25, N=40, slip=False)
    write_case(tmp_path / "dns.case", p, "forcing = constf\nforce = 1e-5\nRSLIP = 1\ntavg_start = 10\ntavg_every = 5\n")
    r = run(["-s"], tmp_path)
    assert r.returncode == 0, r.stderr
    pg = np.loadtxt(tmp_path / "p-grad.txt")
    assert np.allclose(pg[:, 1], 1e-5 * 10 / 2)
    assert "slip option: RSLIP" in open(tmp_path / "report.code").read()
    # time averages over steps 10..40 (31 logged steps, 7 profile samples)
    summ = dict(l.split() for l in open(tmp_path / "tavg-summary.txt"))
    assert float(summ["steps"]) == 31 and float(summ["profile_samples"]) == 7
    assert abs(float(summ["u_tau"]) - np.sqrt(1e-5 * 10 / 2)) < 1e-12
    ws = np.loadtxt(tmp_path / "w-str.txt")[10:, 1]
    assert abs(float(summ["stress"]) - ws.mean()) < 1e-10
    prof = list(tmp_path.glob("tavg-fld.000.*"))
    assert len(prof) == 1 and prof[0].stat().st_size == 9 * 12 * 8


def test_driver_riblets(tmp_path):
    p = dict(Rebs=60.0, ub=0.05, X=8, Y=32, Z=24, TSTR=0, CFL=0.25, N=48, slip=False)
    write_case(tmp_path / "dns.case", p, "texture = riblets\nriblet_spacing = 8\nriblet_height = 4\n"
                                         "init = poiseuille\ninit_uc = 0.06\ninit_amp = 0.05\n")
    r = run(["-s"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "riblets" in open(tmp_path / "report.code").read()
    den = np.fromfile(tmp_path / "den.0000", dtype=np.float64).reshape(8, 32, 26)
    # solid riblet nodes are written as zero density, fluid ones near 1
    assert (den[:, 4, 1:4] == 0).all() and abs(den[:, 0, 1] - 1).max() < 0.1


@pytest.mark.parametrize("precision,tol", [("fp32", 1e-5), ("fp16", 1e-3)])
def test_driver_reduced_storage_runs(tmp_path, precision, tol):
    """The driver with `precision = fp32 | fp16` writes the same file set as
    the FP64 run of the case, with logs and the final checkpoint within the
    storage's bar of the FP64 ones (constant force, riblets)."""
    p = dict(Rebs=60.0, ub=0.05, X=8, Y=32, Z=24, TSTR=0, CFL=0.25, N=48, slip=False)
    extra = ("forcing = constf\nforce = 1e-5\ntexture = riblets\nriblet_spacing = 8\nriblet_height = 4\n"
             "init = poiseuille\ninit_uc = 0.06\ninit_amp = 0.05\n")
    runs = {}
    for prec in ("fp64", precision):
        d = tmp_path / prec
        d.mkdir()
        write_case(d / "dns.case", p, extra + f"precision = {prec}\n")
        r = run(["-s"], d)
        assert r.returncode == 0, r.stderr
        runs[prec] = d
    a, b = runs["fp64"], runs[precision]
    assert sorted(x.name for x in a.iterdir() if x.name != "dns.case") == \
        sorted(x.name for x in b.iterdir() if x.name != "dns.case")
    for log in ("w-str.txt", "flowrate.txt", "energy.txt"):
        la, lb = np.loadtxt(a / log), np.loadtxt(b / log)
        assert np.array_equal(la[:, 0], lb[:, 0]), log
        assert np.abs(lb[:, 1] - la[:, 1]).max() <= tol * np.abs(la[:, 1]).max(), log
    last = sorted(x.name for x in a.iterdir() if x.name.startswith("lbf."))[-1]
    fa, fb = np.fromfile(a / last), np.fromfile(b / last)
    assert np.abs(fb - fa).max() <= tol * np.abs(fa).max(), last


def test_driver_rejects_truncated_checkpoint(tmp_path):
    p = dict(Rebs=60.0, ub=0.05, X=8, Y=8, Z=8, TSTR=3, CFL=0.25, N=16, slip=False)
    write_case(tmp_path / "dns.case", p)
    with open(tmp_path / "lbf.0003", "wb") as fh:
        fh.write(np.zeros(100).tobytes())
    r = run(["-p"], tmp_path)
    assert r.returncode != 0 and "expected" in r.stderr


def test_driver_aborts_on_nan(tmp_path):
    # a NaN population reaches the wall stress within two steps
    # (wshearstr.c:126-130); the kernels only compute with it
    p = dict(Rebs=60.0, ub=0.05, X=8, Y=8, Z=8, TSTR=3, CFL=0.25, N=16, slip=False)
    write_case(tmp_path / "dns.case", p)
    f0 = np.zeros((8, 8, 10, 19))
    f0[:, :, 1:9] = gu.fields.W
    f0[3, 4, 1, 5] = np.nan
    with open(tmp_path / "lbf.0003", "wb") as fh:
        fh.write(f0.tobytes())
        fh.write(f0.tobytes())
    r = run(["-p"], tmp_path)
    assert r.returncode != 0
    assert "Eroor: NAN, found in wall shear stress" in r.stderr
