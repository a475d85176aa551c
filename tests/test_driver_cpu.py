"""Host-side parts of the C driver (libdnsio.so: case reader, wall-texture
masks, file writers) against the reference's own files in the golden
fixtures.  No GPU needed."""
import ctypes
import os

import numpy as np
import pytest

import fields
import golden_util as gu
import pyoracle

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(REPO, "dns-solvers_amd", "lib", "libdnsio.so")


class Row(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("t", "K", "flowrate", "massrate", "stup", "stdown", "stress", "pgrad", "force", "rho")]


class RunCase(ctypes.Structure):
    _fields_ = [("Rebs", ctypes.c_double), ("ub", ctypes.c_double), ("CFL", ctypes.c_double),
                ("xdv", ctypes.c_int), ("ydv", ctypes.c_int), ("zdv", ctypes.c_int),
                ("TSTR", ctypes.c_int), ("TEND", ctypes.c_int), ("ITER", ctypes.c_long),
                ("forcing", ctypes.c_int), ("force", ctypes.c_double), ("precision", ctypes.c_int),
                ("omega", ctypes.c_double), ("texture", ctypes.c_int), ("riblet_spacing", ctypes.c_double),
                ("riblet_height", ctypes.c_double), ("riblet_offset", ctypes.c_double), ("walls", ctypes.c_int),
                ("period_x", ctypes.c_int), ("period_y", ctypes.c_int), ("slip_x", ctypes.c_int),
                ("slip_y", ctypes.c_int), ("offset_x", ctypes.c_int), ("offset_y", ctypes.c_int),
                ("init", ctypes.c_int), ("init_uc", ctypes.c_double), ("init_amp", ctypes.c_double),
                ("init_seed", ctypes.c_uint64), ("init_modes", ctypes.c_int), ("block_x", ctypes.c_int),
                ("block_y", ctypes.c_int), ("block_height", ctypes.c_int), ("tavg_start", ctypes.c_long),
                ("tavg_every", ctypes.c_int), ("outdir", ctypes.c_char * 512)]


@pytest.fixture(scope="module")
def lib():
    L = ctypes.CDLL(LIB)
    L.dnsio_format_row.argtypes = [ctypes.c_int, ctypes.POINTER(Row), ctypes.c_char_p, ctypes.c_int]
    L.dns_case_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(RunCase), ctypes.c_char_p, ctypes.c_int]
    L.dns_case_defaults.argtypes = [ctypes.POINTER(RunCase)]
    L.dns_case_masks.argtypes = [ctypes.POINTER(RunCase), ctypes.c_void_p, ctypes.c_void_p]
    L.dns_case_bands.argtypes = [ctypes.POINTER(RunCase), ctypes.c_void_p, ctypes.c_void_p]
    L.dnsio_write_turbfld.argtypes = [ctypes.c_char_p, ctypes.c_long, ctypes.c_void_p, ctypes.c_int]
    L.dnsio_write_twoD.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
    return L


def oracle_logs(name):
    g = gu.load(name)
    p = gu.params(g)
    X, Y, Z = p["X"], p["Y"], p["Z"]
    b, t = gu.wall_mask(p)
    o = pyoracle.Oracle(X, Y, Z, pyoracle.omega(p["Rebs"], p["ub"], Z), p["ub"], p["CFL"], TM=p["TSTR"],
                        mask=pyoracle.wall_mask(X, Y, Z, b, t))
    ft = gu.initial_state(p)
    f = ft.copy()
    logs = o.run(ft, f, p["N"])
    last = o.logs_of(ft, p["N"])
    rows = {k: np.append(logs[k], last[k]) for k in logs}
    o.run(ft, f, 1, cntr0=p["N"])  # output_twoD sees the state after up+1 steps
    return g, p, o, rows, ft


@pytest.mark.parametrize("name", ["posts8", "smooth8"])
def test_log_rows_byte_identical_to_reference(lib, name):
    g, p, o, rows, _ = oracle_logs(name)
    keys = ("K", "flowrate", "massrate", "stup", "stdown", "stress", "pgrad")
    buf = ctypes.create_string_buffer(128)
    for w, k in enumerate(keys):
        out = bytearray()
        for i in range(len(rows["t"])):
            r = Row(*[rows[f][i] for f in ("t", "K", "flowrate", "massrate", "stup", "stdown", "stress",
                                            "pgrad", "force", "rho")])
            n = lib.dnsio_format_row(w, ctypes.byref(r), buf, 128)
            out += buf.raw[:n]
        assert bytes(out) == g["log_raw_" + k].tobytes(), k


def test_uavg_and_twod_files_byte_identical(lib, tmp_path):
    g, p, o, rows, ft_end = oracle_logs("posts8")
    zd = p["Z"] + 2
    stats = np.ascontiguousarray(g["turbfld"][-1])
    d = str(tmp_path).encode()
    assert lib.dnsio_write_turbfld(d, 110, stats.ctypes.data, zd) == 0
    assert open(tmp_path / "turb-fld.000.110", "rb").read() == stats.tobytes()
    assert open(tmp_path / "uavg-temp.dat", "rb").read() == g["raw_uavg_temp_dat"].tobytes()
    ex, lb = o.twoD(ft_end)
    lb = np.ascontiguousarray(lb)
    assert lib.dnsio_write_twoD(d, lb.ctypes.data, zd, p["ub"]) == 0
    assert open(tmp_path / "2dresults-exact.dat", "rb").read() == g["raw_2dresults_exact_dat"].tobytes()
    assert open(tmp_path / "2dresults-lb.dat", "rb").read() == g["raw_2dresults_lb_dat"].tobytes()


def test_case_file_and_reference_textures(lib, tmp_path):
    cf = tmp_path / "dns.case"
    cf.write_text("# posts case\nRebs = 60.0\nub = 0.05\nxdv = 12\nydv = 8\nzdv = 10\nTSTR = 5\nTEND = 6\n"
                  "CFL = 0.25\nITER = 40\nPSLIP = 1\nforcing = constf\nforce = 1e-6\nprecision = fp32\n")
    c = RunCase()
    err = ctypes.create_string_buffer(512)
    assert lib.dns_case_read(str(cf).encode(), ctypes.byref(c), err, 512) == -1  # defaults not applied
    lib.dns_case_defaults(ctypes.byref(c))
    assert lib.dns_case_read(str(cf).encode(), ctypes.byref(c), err, 512) == 0, err.value
    assert (c.xdv, c.ydv, c.zdv, c.TSTR, c.ITER, c.texture, c.forcing, c.precision) == (12, 8, 10, 5, 40, 1, 1, 1)
    for name, code in (("fp64", 0), ("fp32", 1), ("fp16", 2)):
        cf.write_text(f"precision = {name}\ntavg_start = 100\ntavg_every = 20\n")
        c2 = RunCase()
        lib.dns_case_defaults(ctypes.byref(c2))
        assert lib.dns_case_read(str(cf).encode(), ctypes.byref(c2), err, 512) == 0, err.value
        assert (c2.precision, c2.tavg_start, c2.tavg_every) == (code, 100, 20)
    cf.write_text("precision = fp8\n")
    assert lib.dns_case_read(str(cf).encode(), ctypes.byref(c2), err, 512) == -1
    b = np.zeros((12, 8), np.int8)
    t = np.zeros((12, 8), np.int8)
    lib.dns_case_masks(ctypes.byref(c), b.ctypes.data, t.ctypes.data)
    # PSLIP defaults are the pattern of slip_posts_init.c (pinned by the
    # golden posts runs through fields.posts_mask)
    assert np.array_equal(b, fields.posts_mask(12, 8)) and np.array_equal(t, b)
    c.texture = 2  # RSLIP: slip_ridges_init.c stripes j in {1,2} mod 4
    lib.dns_case_masks(ctypes.byref(c), b.ctypes.data, t.ctypes.data)
    assert np.array_equal(b, fields.stripes_mask(12, 8, 4, 2, offset=1))
    bad = tmp_path / "bad.case"
    bad.write_text("xdv = 8\nnonsense = 3\n")
    lib.dns_case_defaults(ctypes.byref(c))
    assert lib.dns_case_read(str(bad).encode(), ctypes.byref(c), err, 512) == -1
    assert b"nonsense" in err.value


@pytest.mark.parametrize("spacing,height,offset", [(16, 8, 0.0), (8, 4, 0.0), (12, 5, 3.0)])
def test_riblet_voxeliser_matches_test_replica(lib, spacing, height, offset):
    c = RunCase()
    lib.dns_case_defaults(ctypes.byref(c))
    c.xdv, c.ydv, c.zdv = 4, 48, 40
    c.texture = 3
    c.riblet_spacing, c.riblet_height, c.riblet_offset = spacing, height, offset
    b = np.zeros((4, 48), np.int16)
    t = np.zeros((4, 48), np.int16)
    lib.dns_case_bands(ctypes.byref(c), b.ctypes.data, t.ctypes.data)
    ref = fields.riblet_bands(4, 48, spacing, height, offset)
    assert np.array_equal(b, ref) and np.array_equal(t, ref)
    # cell centres sample the triangle: the tallest column is within one cell of the tip
    assert height - 1 <= b.max() <= height


def test_block_voxeliser_matches_test_replica(lib):
    c = RunCase()
    lib.dns_case_defaults(ctypes.byref(c))
    c.xdv, c.ydv, c.zdv = 12, 16, 40
    c.texture = 4
    c.period_x, c.period_y, c.offset_x, c.offset_y = 6, 8, 1, 3
    c.block_x, c.block_y, c.block_height = 2, 3, 5
    c.walls = 1  # bottom wall only
    b = np.zeros((12, 16), np.int16)
    t = np.zeros((12, 16), np.int16)
    lib.dns_case_bands(ctypes.byref(c), b.ctypes.data, t.ctypes.data)
    assert np.array_equal(b, fields.block_bands(12, 16, 6, 8, 2, 3, 5, 1, 3))
    assert not t.any()
    # footprint area: 2 x 3 cells in each of the 2 x 2 periods
    assert (b == 5).sum() == 2 * 3 * 2 * 2 and set(np.unique(b)) == {0, 5}


@pytest.mark.parametrize("zdv,forcing,warned", [(300, "constq", True), (256, "constq", False),
                                                (300, "constf", False)])
def test_driver_warns_for_fp16_constant_flow_rate_in_tall_channels(tmp_path, zdv, forcing, warned):
    """FP16 storage with a constant flow rate above DNSGPU_FP16_CONST_Q_ZMAX
    prints a warning before the device is touched (here the run then stops:
    no GPU in this container)."""
    import subprocess
    exe = os.path.join(REPO, "dns-solvers_amd", "bin", "dnsrun")
    (tmp_path / "dns.case").write_text(f"Rebs = 3000.0\nub = 0.08\nxdv = 4\nydv = 4\nzdv = {zdv}\nTSTR = 0\n"
                                       f"TEND = 1\nCFL = 1.0\nITER = 2\nprecision = fp16\nforcing = {forcing}\n"
                                       "force = 1e-8\n")
    env = dict(os.environ, HIP_VISIBLE_DEVICES="")
    r = subprocess.run([exe, "-s", "dns.case"], cwd=tmp_path, capture_output=True, text=True, timeout=60,
                       env=env)
    assert ("warning: FP16 storage with a constant flow rate" in r.stderr) == warned, r.stderr
