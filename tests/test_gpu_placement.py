"""Population pair placement by dnsgpu_autotune (include/dnsgpu.h,
dnsgpu_pair_placement): the autotuner may replace the buffer that does not
hold the state by a fresh allocation in the other half of device memory.
Only where the bytes live changes, so a placed handle gives bitwise the
populations and logs of an unplaced one at the same launch shape.

* The Re_tau 180 grid 768x384x192 in FP64 (8.6 GB per buffer, above the
  2 GiB threshold): autotune with placement, then 6 logged steps from the
  seeded field, against a handle with DNSGPU_PAIR_PLACEMENT=0 at the same
  shape; two x-slabs of both checkpoint halves and every log row bitwise.
* A small case (below the threshold): the search does not run.
"""
import os
import sys

import numpy as np
import pytest

import dnsgpu

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402  (the workload definitions)

pytestmark = pytest.mark.gpu

C2 = "retau180_768x384x192"


def run(wname, placement, shape=None, steps=6):
    w = bench.WORKLOADS[wname]
    omega, F, ub, nu = bench.case_params(w)
    X, Y, Z = w["X"], w["Y"], w["Z"]
    os.environ["DNSGPU_PAIR_PLACEMENT"] = "1" if placement else "0"
    try:
        with dnsgpu.Channel(X, Y, Z, Rebs=ub * Z / nu, ub=ub, CFL=0.1, forcing=dnsgpu.CONST_Q,
                            omega=omega) as ch:
            ch.init_field(dnsgpu.INIT_TURBULENT, uc=w["utau"], amp=0.1, seed=2, nmodes=8)
            if shape is None:
                ch.autotune(trial_steps=3)
            else:
                ch.set_step_shape(*shape)
            got_shape = ch.step_shape(rows=True)
            pp = ch.pair_placement()
            logs = ch.step(steps)
            slabs = ch.save_slab(0, 2) + ch.save_slab(X // 2, 2)
    finally:
        os.environ.pop("DNSGPU_PAIR_PLACEMENT", None)
    return got_shape, pp, logs, slabs


def test_placed_pair_gives_the_same_state():
    print(f"case: pair placement {C2} FP64, autotuned with placement", flush=True)
    shape, pp, logs, slabs = run(C2, True)
    print(f"placement: {pp[0]} destination buffers tried, step gain {pp[1]:+.4f}; shape {shape}", flush=True)
    assert pp[0] >= 1  # 8.6 GB buffers on a device with room for a candidate
    assert pp[1] == 0.0 or pp[1] >= 0.015
    print(f"case: pair placement {C2} FP64, placement off, shape {shape}", flush=True)
    shape0, pp0, logs0, slabs0 = run(C2, False, shape=shape)
    assert pp0 == (0, 0.0)
    for a, b in zip(slabs, slabs0):
        assert np.array_equal(a, b)
    for k in logs:
        assert np.array_equal(logs[k], logs0[k]), k


def test_small_case_is_not_placed():
    print("case: pair placement, 32x16x24 FP32 (below the threshold)", flush=True)
    os.environ["DNSGPU_PAIR_PLACEMENT"] = "1"
    try:
        with dnsgpu.Channel(32, 16, 24, Rebs=100.0, ub=0.05, CFL=0.1, precision=dnsgpu.FP32) as ch:
            ch.init_rest()
            ch.autotune(trial_steps=2)
            assert ch.pair_placement() == (0, 0.0)
    finally:
        os.environ.pop("DNSGPU_PAIR_PLACEMENT", None)
