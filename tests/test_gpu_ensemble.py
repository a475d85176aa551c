"""The ensemble (many resident cases advanced by one multi-case kernel per
step) against the single-case path, and a member alone against the same
member in a mix.

Bars:
  * constant-F FP64: populations bitwise equal to the single-case path
    after the same steps (the force is a constant, so no reduction feeds
    the state); log rows within 1e-10 relative (plane sums in another
    order).
  * constant-Q FP64: populations within 1e-12 relative of the single-case
    path; log rows within 1e-10.
  * alone == in the mix: bitwise (populations, log rows, time averages) in
    FP64, FP32 and FP16 storage, for members admitted and retired at
    different steps.
  * graph replays: bitwise equal to the same launches made directly.
"""
import numpy as np
import pytest

import dnsgpu
import fields
import pyoracle

pytestmark = pytest.mark.gpu

REBS, UB = 60.0, 0.05


def make(shape, texture, forcing=dnsgpu.CONST_Q, F=0.0, precision=dnsgpu.FP64, seed=7, omega=0.0):
    X, Y, Z = shape
    ch = dnsgpu.Channel(X, Y, Z, Rebs=REBS, ub=UB, CFL=0.25, TSTR=5, forcing=forcing, force=F,
                        precision=precision, omega=omega)
    if texture == "posts":
        m = fields.posts_mask(X, Y)
        ch.set_walls(m, m)
    elif texture == "stripes":
        ch.set_walls(fields.stripes_mask(X, Y, 6, 3), fields.stripes_mask(X, Y, 4, 1, offset=2))
    elif texture == "riblets":
        hb = fields.riblet_bands(X, Y, 8, 4)
        ch.set_solid_bands(hb, hb)
    elif texture == "blocks":
        ch.set_solid_bands(fields.block_bands(X, Y, 4, 4, 2, 2, 3), fields.block_bands(X, Y, 6, 4, 3, 1, 2, 1))
    ch.init_field(dnsgpu.INIT_POISEUILLE, uc=0.06, amp=0.05, seed=seed)
    return ch


def oracle_of(ch, shape, texture, forcing, F):
    """The oracle case of a member built by make()."""
    X, Y, Z = shape
    if texture == "posts":
        m = fields.posts_mask(X, Y)
        mask = pyoracle.wall_mask(X, Y, Z, m, m)
    elif texture == "stripes":
        mask = pyoracle.wall_mask(X, Y, Z, fields.stripes_mask(X, Y, 6, 3), fields.stripes_mask(X, Y, 4, 1, offset=2))
    elif texture == "riblets":
        hb = fields.riblet_bands(X, Y, 8, 4)
        mask = fields.band_mask(X, Y, Z, hb, hb)
    elif texture == "blocks":
        mask = fields.band_mask(X, Y, Z, fields.block_bands(X, Y, 4, 4, 2, 2, 3),
                                fields.block_bands(X, Y, 6, 4, 3, 1, 2, 1))
    else:
        mask = pyoracle.wall_mask(X, Y, Z)
    o = pyoracle.Oracle(X, Y, Z, pyoracle.omega(REBS, UB, Z), UB, 0.25, TM=5,
                        forcing=1 if forcing == dnsgpu.CONST_F else 0, F=F, mask=mask)
    return o, mask


CASES = [((16, 8, 10), "smooth"), ((8, 16, 24), "riblets"), ((16, 12, 70), "stripes"), ((8, 8, 130), "posts"),
         ((12, 16, 20), "blocks")]


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("precision", [dnsgpu.FP64, dnsgpu.FP32, dnsgpu.FP16])
@pytest.mark.parametrize("forcing", [dnsgpu.CONST_F, dnsgpu.CONST_Q])
def test_members_match_single_case_path(forcing, precision):
    """Every member, flat walls, slip textures and solid bands (riblets,
    blocks), in each storage precision: the same state as the single-case
    path of the same precision (bitwise under constant force; under constant
    flow rate the force is a global sum taken in another order: 1e-12 in
    FP64, 1e-6 absolute in FP32 and FP16 storage)."""
    print(f"case: ensemble vs single, forcing {forcing}, precision {precision}", flush=True)
    n = 40
    F = 2e-5 if forcing == dnsgpu.CONST_F else 0.0
    ens = dnsgpu.Ensemble(precision=precision)
    members = [make(s, t, forcing, F, precision=precision) for s, t in CASES]
    for ch in members:
        ens.add(ch)
    ens.step(n)
    for (shape, tex), ch in zip(CASES, members):
        ref = make(shape, tex, forcing, F, precision=precision)
        rl = ref.step(n + 1, cntr0=0)  # dnsens_add takes the first step
        el = ens.rows(ch, n)
        a = ch.save_aos()[0]
        b = ref.save_aos()[0]
        if forcing == dnsgpu.CONST_F:
            assert np.array_equal(a, b), (shape, tex, rel(a, b))
        elif precision == dnsgpu.FP64:
            assert rel(a, b) <= 1e-12, (shape, tex, rel(a, b))
        else:
            assert np.abs(a - b).max() <= 1e-6, (shape, tex, rel(a, b))
        for k in ("K", "flowrate", "massrate", "stup", "stdown", "stress", "pgrad", "force", "rho"):
            ref_k = rl[k][1:]
            assert np.abs(el[k] - ref_k).max() <= 1e-10 * np.abs(ref_k).max() + 1e-300, (shape, tex, k)
        assert np.array_equal(el["t"], rl["t"][1:])
        # time accumulators: every logged step, the add's first one included
        ta, tb = ch.tavg(), ref.tavg()
        assert ta["nsteps"] == tb["nsteps"] == n + 1
        for k in ("pgrad", "stress", "Ub", "K", "rho"):
            assert abs(ta[k] - tb[k]) <= 1e-10 * abs(tb[k]) + 1e-300, k
        ref.close()
    info = ens.info(members[0])
    assert info["finite"] == 1 and info["steps"] == n + 1
    ens.close()


@pytest.mark.parametrize("forcing", [dnsgpu.CONST_F, dnsgpu.CONST_Q])
def test_members_match_the_oracle(forcing):
    """Each FP64 member's state after its steps in the mix against the
    oracle's run of the reference loop (DNS-solver/src/dns-const-q-xycut.c:
    487-518) from the same seeded field: bitwise under constant force, within
    1e-12 relative under constant flow rate; log rows within 1e-10."""
    print(f"case: ensemble members vs oracle, forcing {forcing}", flush=True)
    n = 30
    F = 2e-5 if forcing == dnsgpu.CONST_F else 0.0
    ens = dnsgpu.Ensemble(precision=dnsgpu.FP64)
    members = [make(s, t, forcing, F) for s, t in CASES]
    for ch in members:
        ens.add(ch)
    ens.step(n)
    for (shape, tex), ch in zip(CASES, members):
        X, Y, Z = shape
        o, mask = oracle_of(ch, shape, tex, forcing, F)
        f0 = fields.device_poiseuille(X, Y, Z, 0.06, 0.05, 7)
        f0[mask != 0] = 0.0
        oft, of = f0.copy(), f0.copy()
        ol = o.run(oft, of, n + 1)  # the add's first step and n more
        ft = ch.save_aos()[0]
        if forcing == dnsgpu.CONST_F:
            assert np.array_equal(ft, oft), (shape, tex, rel(ft, oft))
        else:
            assert rel(ft, oft) <= 1e-12, (shape, tex, rel(ft, oft))
        el = ens.rows(ch, n)
        for k in ("K", "flowrate", "massrate", "stup", "stdown", "stress", "pgrad", "rho"):
            ref_k = ol[k][1:]
            assert np.abs(el[k] - ref_k).max() <= 1e-10 * np.abs(ref_k).max() + 1e-300, (shape, tex, k)
    for ch in members:
        ch.close()
    ens.close()


def test_tall_member_matches_single_case_path():
    """A member taller than 292 rows makes the step kernel read the plane
    shifts from global memory instead of LDS: the same bitwise results
    against the single-case path (constant F).  A case above 1024 rows has
    no member footprint (neither path takes it)."""
    print("case: tall member", flush=True)
    n, F = 12, 2e-5
    ens = dnsgpu.Ensemble(precision=dnsgpu.FP64)
    shapes = [(8, 8, 300), (16, 8, 10)]
    members = [make(s, "smooth", dnsgpu.CONST_F, F) for s in shapes]
    for ch in members:
        ens.add(ch)
    ens.step(n)
    for shape, ch in zip(shapes, members):
        ref = make(shape, "smooth", dnsgpu.CONST_F, F)
        ref.step(n + 1, cntr0=0)
        assert np.array_equal(ch.save_aos()[0], ref.save_aos()[0]), shape
        ref.close()
    assert dnsgpu.member_footprint(8, 8, 1025, dnsgpu.FP64) == 0
    for ch in members:
        ch.close()
    ens.close()


@pytest.mark.parametrize("precision", [dnsgpu.FP64, dnsgpu.FP32, dnsgpu.FP16])
def test_alone_equals_in_mix(precision):
    """A member's result does not depend on the other members: the same
    solve run alone and in a changing mix (admissions and retirements at
    other steps) is bitwise identical."""
    print(f"case: alone vs mix, precision {precision}", flush=True)
    probe_shape, probe_tex = (12, 24, 40), "riblets"
    alone = dnsgpu.Ensemble(precision=precision)
    p1 = make(probe_shape, probe_tex, precision=precision, seed=11)
    alone.add(p1)
    for _ in range(3):
        alone.step(30)
    p1_rows = alone.rows(p1, 30)

    mix = dnsgpu.Ensemble(precision=precision)
    a = make((8, 16, 24), "riblets", precision=precision, seed=3)
    mix.add(a)
    mix.step(17)  # an odd chunk: the resident member's buffers swap
    p2 = make(probe_shape, probe_tex, precision=precision, seed=11)
    b = make((8, 8, 130), "posts", precision=precision, seed=5)
    mix.add(p2)
    mix.add(b)
    mix.step(30)
    mix.remove(a)
    c = make((16, 8, 10), "smooth", precision=precision, seed=9)
    mix.add(c)
    mix.step(30)
    mix.step(30)
    p2_rows = mix.rows(p2, 30)

    assert np.array_equal(p1.save_aos()[0], p2.save_aos()[0])
    for k in dnsgpu.LOG_FIELDS:
        assert np.array_equal(p1_rows[k], p2_rows[k]), k
    t1, t2 = p1.tavg(), p2.tavg()
    assert t1 == t2
    for ch in (p1, p2, a, b, c):
        ch.close()
    alone.close()
    mix.close()


def test_graph_replays_match_direct_launches():
    print("case: graph vs direct", flush=True)
    shapes = [((16, 8, 10), "posts"), ((8, 16, 24), "riblets")]
    g = dnsgpu.Ensemble(precision=dnsgpu.FP32)
    d = dnsgpu.Ensemble(precision=dnsgpu.FP32, flags=dnsgpu.ENS_NO_GRAPH)
    gm = [make(s, t, precision=dnsgpu.FP32) for s, t in shapes]
    dm = [make(s, t, precision=dnsgpu.FP32) for s, t in shapes]
    for ch in gm:
        g.add(ch)
    for ch in dm:
        d.add(ch)
    for n in (10, 10, 7, 10, 10):
        g.step(n)
        d.step(n)
        for x, y in zip(gm, dm):
            # host-side work between replays: a statistics pass on each
            x.tavg_sample()
            y.tavg_sample()
            assert np.array_equal(x.save_aos()[0], y.save_aos()[0])
            assert np.array_equal(g.rows(x, n)["stress"], d.rows(y, n)["stress"])
    for x, y in zip(gm, dm):
        assert x.tavg(profiles=True)["profiles"].tolist() == y.tavg(profiles=True)["profiles"].tolist()
    g.close()
    d.close()


def test_admission_by_footprint_and_member_rules():
    print("case: admission", flush=True)
    chs = [make((16, 8, 10), "smooth", precision=dnsgpu.FP32, seed=s) for s in (1, 2, 3)]
    need = dnsgpu.Ensemble.member_bytes(chs[0])
    # the handle plus its tile records: 128 columns in tiles of 8, 4*10+4 doubles each
    assert need == chs[0].device_bytes() + 16 * 44 * 8
    assert dnsgpu.footprint(16, 8, 10, dnsgpu.FP32) == chs[0].device_bytes()
    assert dnsgpu.member_footprint(16, 8, 10, dnsgpu.FP32) == need
    # FP16 storage: the same tiles of 8 columns
    assert dnsgpu.member_footprint(16, 8, 10, dnsgpu.FP16) == dnsgpu.footprint(16, 8, 10, dnsgpu.FP16) + 16 * 44 * 8
    ens = dnsgpu.Ensemble(precision=dnsgpu.FP32, budget=2 * need + need // 2)
    ens.add(chs[0])
    ens.add(chs[1])
    with pytest.raises(dnsgpu.DnsGpuError) as ei:
        ens.add(chs[2])
    assert ei.value.code == -4
    assert ens.count() == 2 and ens.bytes_used() == 2 * need
    # a member's state is the ensemble's to advance
    with pytest.raises(dnsgpu.DnsGpuError):
        chs[0].step(1)
    with pytest.raises(dnsgpu.DnsGpuError):
        chs[0].init_rest()
    ens.step(5)
    ens.remove(chs[0])
    ens.add(chs[2])
    ens.step(5)
    assert ens.count() == 2
    chs[0].step(2)  # a former member steps on its own again
    # precision mismatch is refused
    other = make((16, 8, 10), "smooth", precision=dnsgpu.FP64)
    with pytest.raises(dnsgpu.DnsGpuError):
        ens.add(other)
    # destroying a member handle removes it
    chs[1].close()
    assert ens.count() == 1
    ens.step(3)
    for ch in chs + [other]:
        ch.close()
    ens.close()


def test_info_is_current_from_admission():
    """dnsens_info right after dnsens_add reports the member's counter and
    force as they are then: a fresh state has taken its first step (counter
    1), a state that stepped alone keeps its counter; after a dnsens_step
    the counters advance by the steps taken."""
    print("case: info at admission", flush=True)
    ens = dnsgpu.Ensemble(precision=dnsgpu.FP64)
    fresh = make((16, 8, 10), "smooth")
    stepped = make((16, 8, 10), "smooth", seed=3)
    stepped.step(6)
    g = stepped.force()
    ens.add(fresh)
    ens.add(stepped)
    a, b = ens.info(fresh), ens.info(stepped)
    assert a["steps"] == 1 and b["steps"] == 6
    assert a["finite"] == 1 and b["finite"] == 1
    assert b["force"] == g and a["force"] != 0.0
    ens.step(4)
    assert ens.info(fresh)["steps"] == 5 and ens.info(stepped)["steps"] == 10
    for ch in (fresh, stepped):
        ch.close()
    ens.close()


def test_nan_member_is_flagged_and_others_unaffected():
    print("case: NaN member", flush=True)
    ens = dnsgpu.Ensemble(precision=dnsgpu.FP64)
    good = make((16, 8, 10), "smooth", dnsgpu.CONST_F, 1e-5)
    ref = make((16, 8, 10), "smooth", dnsgpu.CONST_F, 1e-5)
    # tau -> 0.5 with a strong force: the state overflows within a few hundred steps
    bad = make((16, 8, 10), "smooth", dnsgpu.CONST_F, 0.05, omega=1.9999)
    ens.add(good)
    ens.add(bad)
    ens.step(400)
    assert ens.info(bad)["finite"] == 0
    assert ens.info(good)["finite"] == 1
    ref.step(401)
    assert np.array_equal(good.save_aos()[0], ref.save_aos()[0])
    for ch in (good, ref, bad):
        ch.close()
    ens.close()
