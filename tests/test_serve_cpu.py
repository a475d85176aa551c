"""The sweep generator and the request -> case mapping of the serving
layer (host code only; no GPU)."""
import math

import numpy as np

import dnsgpu
import dnsserve


def test_generator_is_seeded_and_in_range():
    a = dnsserve.generate(64, seed=5)
    b = dnsserve.generate(64, seed=5)
    c = dnsserve.generate(64, seed=6)
    da = [dnsserve.as_dict(r) for r in a]
    assert da == [dnsserve.as_dict(r) for r in b]
    assert da != [dnsserve.as_dict(r) for r in c]
    assert da[0]["arrival"] == 0.0
    arr = [r["arrival"] for r in da]
    assert all(x <= y for x, y in zip(arr, arr[1:]))
    textures = set()
    for r in da:
        assert r["xdv"] in (128, 256, 512) and r["ydv"] in (64, 128, 256) and r["zdv"] in (64, 96, 128)
        assert 90.0 <= r["retau"] < 180.0
        textures.add(r["texture"])
        if r["texture"] == dnsserve.GROOVES:
            assert 8 <= r["period"] <= 64 and 0.3 <= r["param"] < 0.9
        else:
            assert r["texture"] == dnsserve.RIBLETS
            assert 8 <= r["period"] <= 32 and r["param"] in range(4, 13)
    assert textures == {dnsserve.GROOVES, dnsserve.RIBLETS}


def test_arrivals_are_a_poisson_stream():
    reqs = dnsserve.generate(4000, seed=11, mean_gap=2.0)
    gaps = np.diff([r.arrival for r in reqs])
    # exponential gaps: mean 2, coefficient of variation 1
    assert abs(gaps.mean() - 2.0) < 0.1
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.06


def test_case_mapping_hits_the_target_re_tau():
    cfg = dnsserve.defaults()
    for r in dnsserve.generate(16, seed=5):
        c, i = dnsserve.case_of(r, cfg)
        assert (c.xdv, c.ydv, c.zdv) == (r.xdv, r.ydv, r.zdv)
        assert c.forcing == dnsgpu.CONST_Q and c.precision == dnsgpu.FP32 and c.ub == cfg.ub
        reb = (r.retau / 0.09) ** (1 / 0.88)  # Dean's correlation
        assert math.isclose(c.Rebs, reb, rel_tol=1e-14)
        nu = c.ub * c.zdv / c.Rebs  # tau = 3 ub zdv / Rebs + 1/2
        # the initial friction velocity gives the target Re_tau on half the height
        assert math.isclose(i.uc * (c.zdv / 2) / nu, r.retau, rel_tol=1e-12)
        assert i.kind == dnsgpu.INIT_TURBULENT and i.nmodes == 8 and i.amp == 0.1 and i.seed == r.seed
