"""Argument checks of the C-ABI that return before any HIP call, so they run
without a GPU: bad grids and precisions are refused with DNSGPU_EARG and a
message, null handles are refused by every entry point, and the size queries
answer 0 for what cannot be built.  The reference has no return codes (it
aborts, wshearstr.c:126-130, or fails in main, dns-const-q-xycut.c:240-356);
the library reports instead, and these are the cases it must report."""
import ctypes

import pytest

import dnsgpu

EARG = -3


def lib():
    return dnsgpu.load()


def case(xdv=8, ydv=8, zdv=16, precision=dnsgpu.FP64):
    return dnsgpu.DnsCase(xdv=xdv, ydv=ydv, zdv=zdv, Rebs=100.0, ub=0.05, CFL=0.1, TSTR=1, forcing=0,
                          force=0.0, precision=precision, omega=0.0)


@pytest.mark.parametrize("grid,why", [
    ((8, 8, 3), "4<=zdv"),            # fewer fluid rows than the wall rule needs
    ((8, 8, 1025), "4<=zdv"),         # more planes than the finisher's block covers
    ((1, 8, 16), "xdv,ydv>=2"),
    ((8, 1, 16), "xdv,ydv>=2"),
    ((0, 0, 0), "xdv,ydv>=2"),
    ((2048, 2048, 512), "2^31"),      # 32-bit node indices
])
def test_create_refuses_unsupported_grids(grid, why):
    L = lib()
    c = case(*grid)
    h = ctypes.c_void_p(1234)
    assert L.dnsgpu_create(ctypes.byref(c), ctypes.byref(h)) == EARG
    assert h.value is None  # the handle is cleared on failure
    assert why in L.dnsgpu_last_error().decode()
    assert L.dnsgpu_footprint(ctypes.byref(c)) == 0
    assert L.dnsens_footprint(ctypes.byref(c)) == 0


@pytest.mark.parametrize("field,value,why", [
    ("Rebs", -50.0, "omega"),          # tau < 1/2
    ("Rebs", 0.0, "omega"),            # tau infinite
    ("omega", 2.0, "omega"),           # the BGK stability limit
    ("omega", -1.0, "negative"),
    ("CFL", 0.0, "CFL"),
    ("forcing", 2, "forcing 2"),
    ("ub", float("nan"), "not finite"),
    ("force", float("inf"), "not finite"),
])
def test_create_refuses_unphysical_parameters(field, value, why):
    L = lib()
    c = case()
    setattr(c, field, value)
    h = ctypes.c_void_p()
    assert L.dnsgpu_create(ctypes.byref(c), ctypes.byref(h)) == EARG
    assert why in L.dnsgpu_last_error().decode()
    assert L.dnsgpu_footprint(ctypes.byref(c)) == 0


def test_create_refuses_unknown_precision():
    L = lib()
    c = case(precision=7)
    h = ctypes.c_void_p()
    assert L.dnsgpu_create(ctypes.byref(c), ctypes.byref(h)) == EARG
    assert "precision 7" in L.dnsgpu_last_error().decode()
    e = ctypes.c_void_p()
    assert L.dnsens_create(7, 0, 0, ctypes.byref(e)) == EARG
    assert e.value is None


def test_null_arguments():
    L = lib()
    c = case()
    assert L.dnsgpu_create(None, None) == EARG
    assert L.dnsgpu_create(ctypes.byref(c), None) == EARG
    assert L.dnsens_create(dnsgpu.FP64, 0, 0, None) == EARG
    assert L.dnsgpu_footprint(None) == 0
    assert L.dnsens_footprint(None) == 0
    assert L.dnsgpu_device_bytes(None) == 0
    assert L.dnsens_member_bytes(None) == 0
    assert L.dnsens_count(None) == 0
    assert L.dnsens_bytes_used(None) == 0
    assert L.dnsens_budget(None) == 0
    L.dnsgpu_destroy(None)
    L.dnsens_destroy(None)


def test_every_handle_call_refuses_a_null_handle():
    L = lib()
    d = (ctypes.c_double * 64)()
    dp = ctypes.cast(d, ctypes.POINTER(ctypes.c_double))
    rows = (dnsgpu.DnsStepLog * 4)()
    init = dnsgpu.DnsInit()
    tavg = dnsgpu.DnsTavg()
    info = dnsgpu.DnsEnsInfo()
    n = ctypes.c_int()
    calls = {
        "dnsgpu_set_walls": (None, None, None),
        "dnsgpu_set_solid_bands": (None, None, None),
        "dnsgpu_init_rest": (None,),
        "dnsgpu_init_field": (None, ctypes.byref(init)),
        "dnsgpu_load_aos": (None, 0, 1, dp, dp),
        "dnsgpu_commit_load": (None,),
        "dnsgpu_save_aos": (None, 0, 1, dp, dp),
        "dnsgpu_get_force": (None, dp),
        "dnsgpu_set_force": (None, 1e-6),
        "dnsgpu_step": (None, 0, 1, rows),
        "dnsgpu_statistics": (None, dp),
        "dnsgpu_fields": (None, 0, 1, dp, dp),
        "dnsgpu_time_steps": (None, 1, 1, dp, dp),
        "dnsgpu_autotune": (None, 1, ctypes.byref(n), dp),
        "dnsgpu_step_shape": (None, ctypes.byref(n), ctypes.byref(n), ctypes.byref(n)),
        "dnsgpu_set_step_shape": (None, 1, 1, 1),
        "dnsgpu_fields_noalloc": (None, 0, 1, dp, dp),
        "dnsgpu_tavg_reset": (None,),
        "dnsgpu_tavg_sample": (None,),
        "dnsgpu_tavg_get": (None, ctypes.byref(tavg), dp),
        "dnsens_add": (None, None),
        "dnsens_remove": (None, None),
        "dnsens_step": (None, 1),
        "dnsens_info": (None, None, ctypes.byref(info)),
        "dnsens_rows": (None, None, rows, 4),
        "dnsens_autotune": (None, 1, ctypes.byref(n), dp),
        "dnsens_time_steps": (None, 1, dp),
        "dnsens_time_kernels": (None, 1, dp, dp),
        "dnsens_grid": (None, ctypes.byref(n), ctypes.byref(n)),
        "dnsens_bind_thread": (None,),
    }
    for name, args in calls.items():
        assert getattr(L, name)(*args) == EARG, name
        assert L.dnsgpu_last_error().decode(), name
