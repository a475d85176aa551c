"""Refused calls on live handles: each returns DNSGPU_EARG with a message and
leaves the handle's state as it was (the handle then steps bitwise like one
that never received the call)."""
import ctypes

import numpy as np
import pytest

import dnsgpu

pytestmark = pytest.mark.gpu

EARG = -3
X, Y, Z = 8, 6, 12


def seeded(precision=dnsgpu.FP64):
    ch = dnsgpu.Channel(X, Y, Z, Rebs=60.0, ub=0.05, CFL=0.25, TSTR=100, forcing=dnsgpu.CONST_F, force=2e-6,
                        precision=precision)
    ch.init_field(dnsgpu.INIT_POISEUILLE, uc=0.06, amp=0.05, seed=3)
    return ch


def refused(fn, *args, match=None):
    with pytest.raises(dnsgpu.DnsGpuError) as ei:
        fn(*args)
    assert ei.value.code == EARG
    if match:
        assert match in str(ei.value)


@pytest.mark.parametrize("precision", [dnsgpu.FP64, dnsgpu.FP16])
def test_refused_calls_leave_the_state_unchanged(precision):
    print("case: refused calls, precision", precision, flush=True)
    a, b = seeded(precision), seeded(precision)
    L = dnsgpu.load()
    slip = np.full((X, Y), -1, np.int8)
    slip[:, ::2] = -2
    bad = np.full((X, Y), -1, np.int8)
    bad[3, 2] = 5
    # a valid bottom wall with an invalid top wall: neither is applied
    refused(a.set_walls, slip, bad, match="wall code 5")
    refused(a.set_solid_bands, np.full((X, Y), 5, np.int16), np.full((X, Y), 6, np.int16),
            match="fewer than 2 fluid rows")
    refused(a.set_solid_bands, np.full((X, Y), -1, np.int16), None, match="fewer than 2 fluid rows")
    refused(a.init_field, 7, match="unknown init kind")
    refused(a.step, -1, match="nsteps < 0")
    buf = np.zeros((X + 1, Y, Z + 2, 19))
    vel = np.zeros((X + 1, Y, Z + 2, 3))
    pb = buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    pv = vel.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    for name, args in (("dnsgpu_load_aos", (0, X + 1, pb, pb)), ("dnsgpu_load_aos", (-1, 2, pb, pb)),
                       ("dnsgpu_save_aos", (X - 1, 2, pb, pb)), ("dnsgpu_save_aos", (0, 0, pb, pb)),
                       ("dnsgpu_fields", (X, 1, pv, pb))):
        assert getattr(L, name)(a.h, *args) == EARG, (name, args)
        assert "slab" in L.dnsgpu_last_error().decode()
    refused(a.time_steps, 0)
    refused(a.autotune, 0)
    la, lb = a.step(4), b.step(4)
    for k in la:
        assert np.array_equal(la[k], lb[k]), k
    fa, fb = a.save_aos(), b.save_aos()
    assert np.array_equal(fa[0], fb[0]) and np.array_equal(fa[1], fb[1])
    # bands go on before the first step only
    refused(a.set_solid_bands, np.ones((X, Y), np.int16), None, match="before the first step")
    a.close()
    b.close()


def test_bands_and_slip_walls_exclude_each_other():
    print("case: bands vs slip walls", flush=True)
    slip = np.full((X, Y), -1, np.int8)
    slip[::2, :] = -2
    bands = np.zeros((X, Y), np.int16)
    bands[:, ::3] = 2
    with dnsgpu.Channel(X, Y, Z, precision=dnsgpu.FP32) as c:
        c.set_walls(slip, None)
        refused(c.set_solid_bands, bands, None, match="no-slip walls")
        # the refused bands left none behind: slip walls may change again
        c.set_walls(None, slip)
    with dnsgpu.Channel(X, Y, Z, precision=dnsgpu.FP32) as c:
        c.set_solid_bands(bands, bands)
        refused(c.set_walls, None, slip, match="cannot be combined")
        c.set_walls(None, None)  # no-slip walls remain allowed
        c.step(2)


def test_a_handle_belongs_to_one_ensemble():
    print("case: one owner per handle", flush=True)
    ch = seeded(dnsgpu.FP32)
    other = seeded(dnsgpu.FP32)
    ea, eb = dnsgpu.Ensemble(precision=dnsgpu.FP32), dnsgpu.Ensemble(precision=dnsgpu.FP32)
    ea.add(ch)
    refused(ea.add, ch, match="already a member")
    refused(eb.add, ch, match="another ensemble")
    refused(eb.remove, other, match="not a member")
    assert eb.count() == 0 and eb.bytes_used() == 0
    ea.step(3)
    ea.remove(ch)
    eb.add(ch)  # free again once removed
    eb.step(2)
    for x in (ea, eb, ch, other):
        x.close()
