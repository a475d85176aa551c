"""Time averages on the GPU (FP64 accumulators): against the oracle's
per-step logs and statistics, the slip velocity against the oracle's
fields, and the FP32-shifted path's error on u_tau, the mean profile and
the drag reduction against FP64 (the bars of the FP32 claim)."""
import numpy as np
import pytest

import dnsgpu
import fields
import pyoracle

pytestmark = pytest.mark.gpu


def oracle_tavg(o, ft, f, nsteps, every):
    """Oracle replay: per-step rows and profile samples every `every` steps,
    combined with the library's formulas (<u^2> = plane variance + mean^2)."""
    rows = []
    acc = np.zeros((9, o.Z + 2))
    slip = np.zeros(4)
    ns = 0
    mb = o.mask[:, :, 0]
    mt = o.mask[:, :, o.Z + 1]
    for s in range(nsteps):
        if s % every == 0:
            st = o.statistics(ft)
            u, v, w, r = st[0], st[1], st[2], st[7]
            acc[0:3] += st[0:3]
            acc[3] += r
            acc[4] += st[3] ** 2 + u * u
            acc[5] += st[4] ** 2 + v * v
            acc[6] += st[5] ** 2 + w * w
            acc[7] += st[6] + u * w
            acc[8] += st[8] ** 2 + r * r
            vel, _ = o.moments(ft)
            slip += [vel[:, :, 1, 0][mb == -2].sum(), vel[:, :, 1, 0][mb == -1].sum(),
                     vel[:, :, o.Z, 0][mt == -2].sum(), vel[:, :, o.Z, 0][mt == -1].sum()]
            ns += 1
        rows.append(o.run(ft, f, 1, cntr0=s))
    logs = {k: np.concatenate([r[k] for r in rows]) for k in rows[0]}
    return logs, acc / ns, slip / ns, ns


def test_time_averages_match_oracle():
    X, Y, Z = 12, 8, 10
    m = fields.posts_mask(X, Y)
    ch = dnsgpu.Channel(X, Y, Z, Rebs=60.0, ub=0.05, CFL=0.25, forcing=dnsgpu.CONST_Q)
    ch.set_walls(m, m)
    ch.init_field(dnsgpu.INIT_POISEUILLE, uc=0.06, amp=0.05, seed=3)
    f0 = fields.device_poiseuille(X, Y, Z, 0.06, 0.05, 3)
    o = pyoracle.Oracle(X, Y, Z, pyoracle.omega(60.0, 0.05, Z), 0.05, 0.25, mask=pyoracle.wall_mask(X, Y, Z, m, m))
    n, every = 60, 5
    ch.tavg_reset()
    for s in range(0, n, every):
        ch.tavg_sample()
        ch.step(every, cntr0=s)
    t = ch.tavg(profiles=True)
    ft, f = f0.copy(), f0.copy()
    logs, acc, slip, ns = oracle_tavg(o, ft, f, n, every)
    assert t["nsteps"] == n and t["nsamples"] == ns
    for k, ok in (("pgrad", "pgrad"), ("stress", "stress"), ("K", "K")):
        assert abs(t[k] - logs[ok].mean()) <= 1e-10 * abs(logs[ok]).max(), k
    assert abs(t["Ub"] - (logs["flowrate"] / (Y * Z)).mean()) <= 1e-12
    assert abs(t["utau"] - np.sqrt(logs["pgrad"].mean())) <= 1e-10 * t["utau"]
    prof = t["profiles"]
    exp = np.zeros_like(prof)
    exp[0:3] = acc[0:3]
    exp[7] = acc[3]
    exp[3] = np.sqrt(np.maximum(acc[4] - acc[0] ** 2, 0))
    exp[4] = np.sqrt(np.maximum(acc[5] - acc[1] ** 2, 0))
    exp[5] = np.sqrt(np.maximum(acc[6] - acc[2] ** 2, 0))
    exp[6] = acc[7] - acc[0] * acc[2]
    exp[8] = np.sqrt(np.maximum(acc[8] - acc[3] ** 2, 0))
    for q in range(9):
        assert np.abs(prof[q] - exp[q]).max() <= 1e-9 * max(np.abs(exp[q]).max(), 1e-30), q
    cols = [(m == -2).sum(), (m == -1).sum()] * 2
    for i, key in enumerate(("u_slip_bottom", "u_noslip_bottom", "u_slip_top", "u_noslip_top")):
        assert abs(t[key] - slip[i] / cols[i]) <= 1e-12, key


@pytest.mark.parametrize("forcing", [dnsgpu.CONST_Q, dnsgpu.CONST_F])
def test_fp32_drag_reduction_and_profile_error(forcing):
    # super-hydrophobic grooves (period 8, gas fraction 0.5) against smooth
    # walls: FP32-shifted must reproduce the FP64 drag reduction within 0.2
    # percentage points and the mean profile within 0.5%
    X, Y, Z = 4, 32, 32
    nu = 0.1
    omega = 1.0 / (3 * nu + 0.5)
    F = 1e-6
    ub = F * Z * Z / (12 * nu)
    res = {}
    for prec in (dnsgpu.FP64, dnsgpu.FP32):
        for tex in ("smooth", "grooves"):
            ch = dnsgpu.Channel(X, Y, Z, Rebs=ub * Z / nu, ub=ub, CFL=0.5, forcing=forcing, force=F, precision=prec,
                                omega=omega)
            if tex == "grooves":
                g = fields.stripes_mask(X, Y, 8, 4)
                ch.set_walls(g, g)
            ch.init_rest()
            ch.step(12000, logs=False)
            ch.tavg_reset()
            for s in range(4):
                ch.tavg_sample()
                ch.step(500)
            res[(prec, tex)] = ch.tavg(profiles=True)
            ch.close()
    dr64 = dnsgpu.drag_reduction(res[(dnsgpu.FP64, "smooth")], res[(dnsgpu.FP64, "grooves")], forcing)
    dr32 = dnsgpu.drag_reduction(res[(dnsgpu.FP32, "smooth")], res[(dnsgpu.FP32, "grooves")], forcing)
    print("DR fp64", dr64, "fp32", dr32)
    for k in dr64:
        assert dr64[k] > 5.0  # the grooves reduce drag
        assert abs(dr32[k] - dr64[k]) <= 0.2, k
    for tex in ("smooth", "grooves"):
        p64 = res[(dnsgpu.FP64, tex)]["profiles"][0]
        p32 = res[(dnsgpu.FP32, tex)]["profiles"][0]
        assert np.abs(p32 - p64).max() <= 0.005 * np.abs(p64).max(), tex
        assert abs(res[(dnsgpu.FP32, tex)]["utau"] - res[(dnsgpu.FP64, tex)]["utau"]) <= 0.005 * res[(dnsgpu.FP64, tex)]["utau"]
    # slip velocity on the grooves is positive and below the bulk velocity
    g64 = res[(dnsgpu.FP64, "grooves")]
    assert 0 < g64["u_slip_bottom"] < g64["Ub"]
    assert abs(res[(dnsgpu.FP32, "grooves")]["u_slip_bottom"] - g64["u_slip_bottom"]) <= 0.005 * g64["u_slip_bottom"]
