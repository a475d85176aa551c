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
* Pair arenas (dnsgpu_reserve_pair_arenas): three small members of an
  ensemble, FP32 and FP16, created while the arenas are reserved (each
  member's two buffers ranges of blocks in different memory halves) and
  stepped 37 times, against the same members on plain allocations: states
  and log rows bitwise; a second reservation while the members hold ranges
  grows each half until the request is free; after the release every
  block is freed.
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


def _members(precision):
    """Three small members of different grids and textures (seeded
    Poiseuille fields)."""
    import fields
    out = []
    for (X, Y, Z), tex, seed in (((16, 24, 40), "riblets", 3), ((32, 16, 24), "smooth", 5), ((8, 32, 70), "posts", 7)):
        ch = dnsgpu.Channel(X, Y, Z, Rebs=60.0, ub=0.05, CFL=0.25, TSTR=5, precision=precision)
        if tex == "riblets":
            hb = fields.riblet_bands(X, Y, 8, 4)
            ch.set_solid_bands(hb, hb)
        elif tex == "posts":
            m = fields.posts_mask(X, Y)
            ch.set_walls(m, m)
        ch.init_field(dnsgpu.INIT_POISEUILLE, uc=0.06, amp=0.05, seed=seed)
        out.append(ch)
    return out


def _ensemble_run(precision, arenas):
    if arenas:
        got = dnsgpu.reserve_pair_arenas(1 << 30)
        assert got >= 1 << 30
        b0, b1, used = dnsgpu.pair_arena_stats(0)
        assert b0 >= 1 and b1 >= 1 and used == 0
    try:
        ms = _members(precision)
        if arenas:
            # every member's two buffers are ranges of the arenas
            used = dnsgpu.pair_arena_stats(0)[2]
            assert used >= sum(2 * m.X * m.Y * m.Z * 19 * 4 for m in ms)
            # a second reservation counts the bytes not held by the members:
            # 8 GiB free per half needs a second block in each
            got = dnsgpu.reserve_pair_arenas(8 << 30)
            b0, b1, used2 = dnsgpu.pair_arena_stats(0)
            assert got >= 8 << 30 and b0 >= 2 and b1 >= 2 and used2 == used
        ens = dnsgpu.Ensemble(precision=precision)
        for m in ms:
            ens.add(m)
        ens.step(17)
        ens.step(20)
        rows = [ens.rows(m, 20) for m in ms]
        states = [m.save_aos() for m in ms]
        ens.close()
        for m in ms:
            m.close()
    finally:
        if arenas:
            dnsgpu.release_pair_arenas()
    return rows, states


@pytest.mark.parametrize("precision", [dnsgpu.FP32, dnsgpu.FP16])
def test_arena_pairs_give_the_same_ensemble_state(precision):
    print(f"case: pair arenas, three members, precision {precision}, with arenas", flush=True)
    rows_a, states_a = _ensemble_run(precision, True)
    assert dnsgpu.pair_arena_stats(0) == (0, 0, 0)  # released: every block freed
    print(f"case: pair arenas, three members, precision {precision}, plain allocations", flush=True)
    rows_p, states_p = _ensemble_run(precision, False)
    for sa, sp in zip(states_a, states_p):
        assert np.array_equal(sa[0], sp[0]) and np.array_equal(sa[1], sp[1])
    for ra, rp in zip(rows_a, rows_p):
        for k in dnsgpu.LOG_FIELDS:
            assert np.array_equal(ra[k], rp[k]), k
