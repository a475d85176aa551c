"""The oracle against the reference solver compiled from its own sources in
this container (oracle/build_ref.sh): bitwise checkpoint state and
byte-identical log files on grids and seeds the golden fixtures do not
cover.  Skipped where the reference sources or MPICH are absent (a
machine that only receives this repository); tests/test_oracle_golden.py
covers that case from fixtures."""
import os

import numpy as np
import pytest

import fields
import golden_util as gu
import pyoracle
import refrun

pytestmark = pytest.mark.skipif(not refrun.available(), reason="reference sources / mpirun not present")

CASES = [
    # X, Y, Z, Rebs, ub, CFL (utime = 0.5 Z / CFL), N, slip, seed
    (12, 8, 9, 90.0, 0.04, 0.5625, 80, "", 21),
    (8, 12, 6, 45.0, 0.07, 0.375, 80, "PSLIP", 22),
    (16, 8, 11, 150.0, 0.03, 0.55, 100, "PSLIP", 23),
]


@pytest.mark.parametrize("X,Y,Z,Rebs,ub,CFL,N,slip,seed", CASES)
def test_oracle_bitwise_equal_to_reference(tmp_path, X, Y, Z, Rebs, ub, CFL, N, slip, seed):
    TSTR = 7
    ut = int(0.5 * Z / CFL)
    assert N % (10 * ut) == 0
    exe = refrun.build(f"t_{X}x{Y}x{Z}_{slip or 'smooth'}", X, Y, Z, Rebs, ub, CFL, TSTR, TSTR + 1, N, slip)
    f0 = fields.poiseuille_field(X, Y, Z, 2 * ub, noise=0.08, seed=seed)
    f1 = fields.poiseuille_field(X, Y, Z, 2 * ub, noise=0.08, seed=seed + 100)
    wd = refrun.run(exe, str(tmp_path / "run"), f0, f1, TSTR)
    rft, rf = refrun.read_lbf(os.path.join(wd, "lbf.%.4d" % (TSTR + N // ut)), X, Y, Z)
    m = fields.posts_mask(X, Y) if slip else None
    o = pyoracle.Oracle(X, Y, Z, pyoracle.omega(Rebs, ub, Z), ub, CFL, TM=TSTR,
                        mask=pyoracle.wall_mask(X, Y, Z, m, m))
    ft, f = f0.copy(), f1.copy()
    logs = o.run(ft, f, N)
    assert np.array_equal(ft, rft)
    assert np.array_equal(f, rf)
    last = o.logs_of(ft, N)
    for k in gu.LOG_KEYS:
        txt = gu.log_text(k, np.append(logs["t"], last["t"]), np.append(logs[k], last[k]))
        assert txt == open(os.path.join(wd, refrun.LOG_FILES[k]), "rb").read(), k
    # the checkpoint written at step 0 is the input pair, byte for byte
    r0 = refrun.read_lbf(os.path.join(wd, "lbf.%.4d" % TSTR), X, Y, Z)
    assert np.array_equal(r0[0], f0) and np.array_equal(r0[1], f1)
