"""The configured workloads on the GPU, against the oracle and through
size-independent properties.

* Initial field: the library's seeded turbulent field (log-law mean plus
  the curl of 8x8 Fourier modes at 10% rms of the centreline velocity,
  seed 2) on the Re_tau 180 grid 768x384x192 and the Re_tau 395 grid
  1536x768x384 is bitwise the replica in tests/fields.py at the x-planes
  compared.
* x-slabs of the workloads: 4 x-planes of the seeded field, periodic in x,
  taken as a restart state by both the library and the oracle, as the
  reference's restart path takes a state
  (DNS-solver/src/dns-const-q-xycut.c:399-454): smooth walls (Re_tau 180),
  grooves of period 32 with gas fraction 0.5 on both walls (Re_tau 180),
  riblets of spacing 16 and height 8 on both walls (Re_tau 395), under the
  workloads' constant body force.  FP64: populations bitwise equal to the
  oracle; FP32 and FP16 storage against the FP64 path within POP_BAR.
* Full grids, FP64 and FP32, constant flow rate, a few logged steps: a
  finite state, the mean density conserved (MASS_BAR, relative), and the
  force that the fused budget of step n hands to step n+1 equal to the one
  the moments of state n+1 give (rho ub - j_x, encalc.c:121) within
  BUDGET_BAR ub.
"""
import os
import sys

import numpy as np
import pytest

import dnsgpu
import fields
import pyoracle

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402  (the workload definitions)

pytestmark = pytest.mark.gpu

C2 = "retau180_768x384x192"
C4 = "retau395_riblets_1536x768x384"
SLAB = 4
# FP32 / FP16 storage against FP64 after the slab runs (absolute, f ~ 1/3)
POP_BAR = {dnsgpu.FP32: 1e-5, dnsgpu.FP16: 1e-4}
# relative change of the mean density over the full-grid runs: FP64 sums
# differ only in order; FP32 storage rounds every collision (the mean of the
# rounding over the grid, a few 1e-11 per step from this initial field)
MASS_BAR = {dnsgpu.FP64: 1e-13, dnsgpu.FP32: 1e-9}
# budget force against the moments' force, in units of ub: FP64 sums differ
# only in order; with FP32 storage both sides are FP32 moments (each node's
# j_x is a single-precision sum of its 19 values, rounded alike at the many
# similar nodes of a smooth plane, so the rounding does not average out:
# 2e-9 ub measured at 768x384x192)
BUDGET_BAR = {dnsgpu.FP64: 1e-13, dnsgpu.FP32: 1e-8}

_fields_cache = {}


def workload(wname):
    w = bench.WORKLOADS[wname]
    omega, F, ub, nu = bench.case_params(w)
    return w, omega, F, ub, nu


def seeded_slab(wname):
    """The FP64 seeded turbulent field of the full workload grid (the
    library's generator), x-planes [0, SLAB), checked against the replica."""
    if wname in _fields_cache:
        return _fields_cache[wname]
    w, omega, F, ub, nu = workload(wname)
    X, Y, Z = w["X"], w["Y"], w["Z"]
    print(f"case: seeded field {wname} FP64, planes 0..{SLAB - 1}", flush=True)
    with dnsgpu.Channel(X, Y, Z, Rebs=ub * Z / nu, ub=ub, CFL=0.1, forcing=dnsgpu.CONST_F, force=F,
                        omega=omega) as ch:
        ch.init_field(dnsgpu.INIT_TURBULENT, uc=w["utau"], amp=0.1, seed=2, nmodes=8)
        ft, f = ch.save_slab(0, SLAB)
    ref = fields.device_turbulent(X, Y, Z, omega, w["utau"], 0.1, 2, 8, xs=np.arange(SLAB))
    _fields_cache[wname] = (ft, f, ref)
    return ft, f, ref


@pytest.mark.parametrize("wname", [C2, C4])
def test_seeded_field_of_the_workload_is_the_replica(wname):
    ft, f, ref = seeded_slab(wname)
    assert np.array_equal(ft, ref)
    assert np.array_equal(f, ref)


def slab_run(wname, texture, precision, steps):
    """A SLAB-plane periodic channel from the workload's seeded field: the
    library's state after `steps` steps (ftemp), its log rows, and the
    oracle's case for the same slab."""
    w, omega, F, ub, nu = workload(wname)
    _, Y, Z = w["X"], w["Y"], w["Z"]
    f0, _, _ = seeded_slab(wname)
    f0 = f0.copy()
    X = SLAB
    bot = top = None
    hb = None
    if texture == "grooves":
        bot = top = fields.stripes_mask(X, Y, 32, 16)
        mask = pyoracle.wall_mask(X, Y, Z, bot, top)
    elif texture == "riblets":
        hb = fields.riblet_bands(X, Y, 16, 8)
        mask = fields.band_mask(X, Y, Z, hb, hb)
        f0[mask != 0] = 0.0
    else:
        mask = pyoracle.wall_mask(X, Y, Z)
    print(f"case: slab {X}x{Y}x{Z} of {wname}, {texture}, precision {precision}, {steps} steps", flush=True)
    with dnsgpu.Channel(X, Y, Z, Rebs=ub * Z / nu, ub=ub, CFL=0.1, TSTR=0, forcing=dnsgpu.CONST_F, force=F,
                        precision=precision, omega=omega) as ch:
        if bot is not None:
            ch.set_walls(bot, top)
        ch.load_aos(f0, f0)
        if hb is not None:
            ch.set_solid_bands(hb, hb)
        logs = ch.step(steps)
        ft, _ = ch.save_aos()
    o = pyoracle.Oracle(X, Y, Z, omega, ub, 0.1, TM=0, forcing=1, F=F, mask=mask)
    return f0, ft, logs, o


@pytest.mark.parametrize("wname,texture,steps", [(C2, "smooth", 20), (C2, "grooves", 20), (C4, "riblets", 10)])
def test_workload_slab_fp64_bitwise_to_the_oracle(wname, texture, steps):
    f0, ft, logs, o = slab_run(wname, texture, dnsgpu.FP64, steps)
    oft, of = f0.copy(), f0.copy()
    ologs = o.run(oft, of, steps)
    assert np.array_equal(ft, oft), float(np.abs(ft - oft).max())
    for k in ("K", "flowrate", "massrate", "stup", "stdown", "stress", "rho"):
        assert np.allclose(logs[k], ologs[k], rtol=1e-10, atol=1e-300), k


@pytest.mark.parametrize("wname,texture,steps", [(C2, "grooves", 20), (C4, "riblets", 10)])
@pytest.mark.parametrize("precision", [dnsgpu.FP32, dnsgpu.FP16])
def test_workload_slab_reduced_storage_against_fp64(wname, texture, steps, precision):
    _, a, la, _ = slab_run(wname, texture, dnsgpu.FP64, steps)
    _, b, lb, _ = slab_run(wname, texture, precision, steps)
    err = float(np.abs(a - b).max())
    print(f"storage {precision} vs FP64: populations {err:.3e}", flush=True)
    assert err <= POP_BAR[precision]
    assert np.abs(la["flowrate"] - lb["flowrate"]).max() <= 1e-4 * np.abs(la["flowrate"]).max()


@pytest.mark.parametrize("wname", [C2, C4])
@pytest.mark.parametrize("precision", [dnsgpu.FP64, dnsgpu.FP32])
def test_full_grid_budget_and_mass(wname, precision):
    w, omega, F, ub, nu = workload(wname)
    X, Y, Z = w["X"], w["Y"], w["Z"]
    print(f"case: full grid {wname} constant Q, precision {precision}", flush=True)
    with dnsgpu.Channel(X, Y, Z, Rebs=ub * Z / nu, ub=ub, CFL=0.1, forcing=dnsgpu.CONST_Q,
                        precision=precision, omega=omega) as ch:
        if w.get("riblets"):
            hb = bench.riblet_bands(X, Y, *w["riblets"])
            ch.set_solid_bands(hb, hb)
        ch.init_field(dnsgpu.INIT_TURBULENT, uc=w["utau"], amp=0.1, seed=2, nmodes=8)
        logs = ch.step(4)
    for k in ("K", "flowrate", "massrate", "stress", "force", "rho"):
        assert np.all(np.isfinite(logs[k])), k
    rho = logs["rho"]
    direct = logs["rho"][1:] * ub - logs["massrate"][1:] / (Y * Z)
    print(f"mean density change {np.abs(rho - rho[0]).max() / rho[0]:.3e}, budget - moments "
          f"{np.abs(logs['force'][1:] - direct).max() / ub:.3e} ub", flush=True)
    assert np.abs(rho - rho[0]).max() <= MASS_BAR[precision] * rho[0], rho
    # the force of step n+1 from the fused budget of step n, against the
    # moments of state n+1: rho_mean ub - j_x_mean (massrate = j_x_mean Y Z)
    assert np.abs(logs["force"][1:] - direct).max() <= BUDGET_BAR[precision] * ub, (logs["force"], direct)
