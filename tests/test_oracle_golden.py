"""The oracle reproduces the reference's own outputs (golden fixtures written
by the reference binary, tests/golden/make_golden.py) bit for bit:
checkpoint state after N steps, the seven per-step log files byte for byte,
every turb-fld statistics file and the vel/den fields.  Runs without the
reference sources or a GPU."""
import numpy as np
import pytest

import golden_util as gu
import pyoracle


def run_oracle(g):
    p = gu.params(g)
    X, Y, Z = p["X"], p["Y"], p["Z"]
    bot, top = gu.wall_mask(p)
    o = pyoracle.Oracle(X, Y, Z, pyoracle.omega(p["Rebs"], p["ub"], Z), p["ub"], p["CFL"], TM=p["TSTR"],
                        mask=pyoracle.wall_mask(X, Y, Z, bot, top))
    ft = gu.initial_state(p)
    f = ft.copy()
    logs = {k: [] for k in ("t",) + gu.LOG_KEYS}
    stats = []
    moments = None
    for ss in range(p["N"] + 1):
        row = o.logs_of(ft, ss)
        for k in logs:
            logs[k].append(row[k])
        if ss % p["utime"] == 0:
            stats.append(o.statistics(ft))
        if ss == p["N"]:
            moments = o.moments(ft)
            break
        o.run(ft, f, 1, cntr0=ss)
    return p, o, ft, f, {k: np.array(v) for k, v in logs.items()}, np.array(stats), moments


@pytest.fixture(scope="module", params=gu.CASES)
def case(request):
    g = gu.load(request.param)
    return request.param, g, run_oracle(g)


def test_state_after_n_steps_bitwise(case):
    name, g, (p, o, ft, f, logs, stats, mom) = case
    assert gu.lbf_sha(ft, f) == str(g["lbf_sha256"]), name
    if "lbf_ftemp" in g:
        assert np.array_equal(ft, g["lbf_ftemp"]) and np.array_equal(f, g["lbf_f"])
    else:
        idx = g["sample_idx"]
        assert np.array_equal(ft.reshape(-1)[idx], g["lbf_ftemp_sample"])


def test_log_files_byte_identical(case):
    name, g, (p, o, ft, f, logs, stats, mom) = case
    for k in gu.LOG_KEYS:
        assert gu.log_text(k, logs["t"], logs[k]) == g["log_raw_" + k].tobytes(), (name, k)


def test_turbfld_bitwise(case):
    name, g, (p, o, ft, f, logs, stats, mom) = case
    assert stats.shape == g["turbfld"].shape
    assert np.array_equal(stats, g["turbfld"]), name


def test_vel_den_bitwise(case):
    import hashlib
    name, g, (p, o, ft, f, logs, stats, (vel, den)) = case
    assert hashlib.sha256(vel.tobytes()).hexdigest() == str(g["vel_sha256"])
    assert hashlib.sha256(den.tobytes()).hexdigest() == str(g["den_sha256"])
