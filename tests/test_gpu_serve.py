"""The continuous-batching service: every solve's result is the same when it
is served alone and in a mix that arrives over time (bitwise: the member's
state does not depend on the others, and its convergence decisions use its
own statistics only)."""
import pytest

import dnsgpu
import dnsserve

pytestmark = pytest.mark.gpu

# grid Reynolds number Re_b / zdv at most about 70 (resolved enough for BGK
# to stay bounded on these coarse grids)
SMALL = [
    dict(xdv=32, ydv=16, zdv=24, retau=55.0, texture=dnsserve.GROOVES, period=8, param=0.5, seed=1),
    dict(xdv=16, ydv=32, zdv=40, retau=90.0, texture=dnsserve.RIBLETS, period=16, param=4.0, seed=2),
    dict(xdv=24, ydv=24, zdv=70, retau=120.0, texture=dnsserve.GROOVES, period=12, param=0.75, seed=3),
    dict(xdv=16, ydv=16, zdv=32, retau=70.0, texture=dnsserve.SMOOTH, period=0, param=0.0, seed=4),
    dict(xdv=40, ydv=16, zdv=24, retau=60.0, texture=dnsserve.RIBLETS, period=8, param=5.0, seed=5),
]


def cfg(**kw):
    base = dict(transient=300, window=400, check=100, cap=1500, tol=0.01, realtime=0, init_modes=4)
    base.update(kw)
    return dnsserve.defaults(**base)


KEYS = ("steps", "status", "utau", "Re_tau", "Cf_w", "Cf_p", "Ub", "stress", "u_slip")


def test_each_result_is_the_same_alone_and_in_the_mix():
    print("case: serve small mix, burst then realtime", flush=True)
    mix, summ = dnsserve.sweep(dnsserve.requests(SMALL), cfg())
    assert summ["completed"] == len(SMALL) and summ["max_resident"] == len(SMALL), [r["status"] for r in mix]
    for r in mix:
        assert r["status"] in (0, 1) and r["steps"] >= 301 and r["utau"] > 0
    for i, item in enumerate(SMALL):
        alone, _ = dnsserve.sweep(dnsserve.requests([item]), cfg())
        for k in KEYS:
            assert alone[0][k] == mix[i][k], (i, k)
    # arrivals over time: requests join while others run
    timed = [dict(item, arrival=0.02 * i) for i, item in enumerate(SMALL)]
    late, s2 = dnsserve.sweep(dnsserve.requests(timed), cfg(realtime=1))
    assert s2["completed"] == len(SMALL)
    for i in range(len(SMALL)):
        for k in KEYS:
            assert late[i][k] == mix[i][k], (i, k)
        assert late[i]["finished"] >= late[i]["admitted"] >= late[i]["arrival"]
    assert s2["p99"] >= s2["p50"] > 0


def test_admission_waits_for_memory():
    print("case: serve with a budget of one solve", flush=True)
    # three solves of one grid: a budget of 1.5 footprints holds one at a time
    items = [dict(SMALL[0], seed=11), dict(SMALL[0], texture=dnsserve.RIBLETS, period=8, param=4.0, seed=12),
             dict(SMALL[0], param=0.25, seed=13)]
    need = dnsgpu.member_footprint(32, 16, 24, dnsgpu.FP32)
    res, summ = dnsserve.sweep(dnsserve.requests(items), cfg(budget=need + need // 2))
    assert summ["max_resident"] == 1 and summ["completed"] == 3
    # served one at a time, in arrival order
    assert res[0]["finished"] <= res[1]["admitted"] and res[1]["finished"] <= res[2]["admitted"]
    ref, _ = dnsserve.sweep(dnsserve.requests(items), cfg())
    for a, b in zip(res, ref):
        for k in KEYS:
            assert a[k] == b[k], k


def test_sequential_single_case_path_runs_the_same_steps():
    print("case: sequential single-case path", flush=True)
    items = SMALL[:2]
    res, _ = dnsserve.sweep(dnsserve.requests(items), cfg())
    secs = dnsserve.sequential(dnsserve.requests(items), cfg(), dnsgpu.FP64, [r["steps"] for r in res])
    assert all(s > 0 for s in secs)


def test_dnsserve_command_line(tmp_path):
    """The request-loop command: a short burst of the seeded sweep (three
    requests, reduced transient/window/cap) prints one JSON object per
    solve and a summary line."""
    import json
    import os
    import subprocess
    print("case: dnsserve command", flush=True)
    exe = os.path.join(os.path.dirname(dnsgpu.LIB_PATH), "..", "bin", "dnsserve")
    r = subprocess.run([exe, "-b", "-n", "3", "-c", "400", "-t", "100", "-w", "100", "-k", "100", "-p", "fp16"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = [json.loads(x) for x in r.stdout.splitlines()]
    solves, summary = lines[:-1], lines[-1]
    assert len(solves) == 3 and summary["summary"] and summary["completed"] == 3
    for s in solves:
        assert s["status"] in (0, 1) and 201 <= s["steps"] <= 401 and s["u_tau"] > 0
    assert "admit" in r.stderr and "done" in r.stderr


def test_texture_sweep_of_64_requests_alone_and_in_the_mix():
    """The configured sweep: 64 seeded requests (grids {128,256,512} x
    {64,128,256} x {64,96,128}, Re_tau in [90,180], grooves or riblets),
    FP32-shifted storage, all arriving at once and admitted as they are
    prepared (the first solves run while the later ones are created), with
    a short transient, window and cap; a sample of the solves (the first
    groove, the first riblet and the largest grid) served alone gives
    bitwise the same results as in the mix of 64."""
    print("case: serve the 64-request sweep, FP32, cap 400", flush=True)
    reqs = dnsserve.generate(64, seed=5, mean_gap=2.0)
    c = cfg(transient=200, window=200, check=100, cap=400, precision=dnsgpu.FP32, init_modes=8)
    mix, summ = dnsserve.sweep(reqs, c)
    assert summ["completed"] == 64 and summ["max_resident"] >= 8, ([r["status"] for r in mix], summ)
    for r in mix:
        assert r["status"] in (0, 1) and 201 <= r["steps"] <= 401 and r["utau"] > 0, r
    tex = [r.texture for r in reqs]
    sizes = [r.xdv * r.ydv * r.zdv for r in reqs]
    sample = sorted({tex.index(dnsserve.GROOVES), tex.index(dnsserve.RIBLETS), sizes.index(max(sizes))})
    for i in sample:
        one = dnsserve.requests([{k: getattr(reqs[i], k) for k, _ in dnsserve.Request._fields_}])
        alone, _ = dnsserve.sweep(one, c)
        for k in KEYS:
            assert alone[0][k] == mix[i][k], (i, k)


def test_request_larger_than_the_budget_is_refused():
    """A request whose footprint exceeds the whole budget is refused (status
    -2) when it comes up; the requests before and after it are served and
    give the results they give alone."""
    print("case: serve with one request above the budget", flush=True)
    items = [SMALL[0], SMALL[2], SMALL[3]]
    fp = [dnsgpu.member_footprint(r["xdv"], r["ydv"], r["zdv"], dnsgpu.FP32) for r in items]
    budget = fp[0] + fp[2]
    assert fp[1] > budget
    res, summ = dnsserve.sweep(dnsserve.requests(items), cfg(budget=budget))
    assert [r["status"] for r in res][1] == -2 and summ["completed"] == 2
    for i in (0, 2):
        assert res[i]["status"] in (0, 1)
        alone, _ = dnsserve.sweep(dnsserve.requests([items[i]]), cfg())
        for k in KEYS:
            assert alone[0][k] == res[i][k], (i, k)
