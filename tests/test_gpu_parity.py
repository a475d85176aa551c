"""Parity of the HIP library against the oracle (CPU restatement, pinned
bitwise to the reference binary) on seeded inputs.

Bars (stated per test):
  * constant-F FP64: populations bitwise equal to the oracle after N steps
    (same arithmetic in the same order; no reduction feeds the state).
  * constant-Q FP64: populations within 1e-12 relative (the force is a
    global sum; the GPU sums in a fixed tree order, the reference
    sequentially); log rows within 1e-10 relative.
  * FP32-shifted and FP16-shifted storage: error against FP64 stated per
    quantity.
"""
import numpy as np
import pytest

import dnsgpu
import fields
import pyoracle

pytestmark = pytest.mark.gpu


def make_pair(X, Y, Z, forcing, F=0.0, mask2d=None, precision=dnsgpu.FP64, Rebs=60.0, ub=0.05, CFL=0.25,
              TSTR=100):
    ch = dnsgpu.Channel(X, Y, Z, Rebs=Rebs, ub=ub, CFL=CFL, TSTR=TSTR, forcing=forcing, force=F,
                        precision=precision)
    bot = top = None
    if mask2d is not None:
        bot, top = mask2d
        ch.set_walls(bot, top)
    om = pyoracle.omega(Rebs, ub, Z)
    o = pyoracle.Oracle(X, Y, Z, om, ub, CFL, TM=TSTR, forcing=1 if forcing == dnsgpu.CONST_F else 0, F=F,
                        mask=pyoracle.wall_mask(X, Y, Z, bot, top))
    return ch, o


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def seeded(ch, uc=0.06, amp=0.05, seed=7):
    ch.init_field(dnsgpu.INIT_POISEUILLE, uc=uc, amp=amp, seed=seed)
    return fields.device_poiseuille(ch.X, ch.Y, ch.Z, uc, amp, seed)


def test_init_field_matches_replica_bitwise():
    print("case: init POISEUILLE 12x8x70", flush=True)
    ch = dnsgpu.Channel(12, 8, 70)
    f0 = seeded(ch)
    ft, f = ch.save_aos()
    assert np.array_equal(ft, f0)
    assert np.array_equal(f, f0)


def test_save_load_roundtrip_bitwise():
    X, Y, Z = 8, 6, 12
    rng = np.random.default_rng(3)
    a = rng.uniform(0.01, 0.4, size=(X, Y, Z + 2, 19))
    b = rng.uniform(0.01, 0.4, size=(X, Y, Z + 2, 19))
    a[:, :, 0] = a[:, :, -1] = 0
    b[:, :, 0] = b[:, :, -1] = 0
    ch = dnsgpu.Channel(X, Y, Z)
    ch.load_aos(a, b, slab=3)
    ft, f = ch.save_aos(slab=5)
    assert np.array_equal(ft, a) and np.array_equal(f, b)


@pytest.mark.parametrize("shape,walls", [
    ((16, 8, 10), "smooth"),
    ((16, 8, 10), "posts"),
    ((16, 12, 70), "stripes"),
    ((8, 8, 130), "posts"),
    ((2, 2, 4), "smooth"),      # the smallest grid the library takes
    ((3, 5, 65), "stripes"),    # odd x/y extents, a one-row last chunk
    ((4, 4, 512), "posts"),     # the largest zdv (8 chunks, 3 waves per block)
])
def test_const_force_fp64_bitwise(shape, walls):
    X, Y, Z = shape
    print(f"case: const-F FP64 {shape} {walls}", flush=True)
    mask = None
    if walls == "posts":
        m = fields.posts_mask(X, Y)
        mask = (m, m)
    elif walls == "stripes":
        mask = (fields.stripes_mask(X, Y, 6, 3), fields.stripes_mask(X, Y, 4, 1, offset=2))
    ch, o = make_pair(X, Y, Z, dnsgpu.CONST_F, F=2e-5, mask2d=mask)
    f0 = seeded(ch)
    n = 60
    logs = ch.step(n, cntr0=0)
    ft, f = ch.save_aos()
    oft, of = f0.copy(), f0.copy()
    ologs = o.run(oft, of, n)
    assert np.array_equal(ft, oft), rel(ft, oft)
    assert np.array_equal(f[:, :, 1:Z + 1], of[:, :, 1:Z + 1])
    for k in ("K", "flowrate", "massrate", "stup", "stdown", "stress", "pgrad", "rho"):
        assert np.allclose(logs[k], ologs[k], rtol=1e-10, atol=1e-300), (k, logs[k][:3], ologs[k][:3])
    assert np.array_equal(logs["t"], ologs["t"])


@pytest.mark.parametrize("walls", ["smooth", "posts"])
def test_const_q_fp64_within_1e12(walls):
    X, Y, Z = 16, 8, 10
    print(f"case: const-Q FP64 {walls}", flush=True)
    mask = None
    if walls == "posts":
        m = fields.posts_mask(X, Y)
        mask = (m, m)
    ch, o = make_pair(X, Y, Z, dnsgpu.CONST_Q, mask2d=mask)
    f0 = seeded(ch)
    n = 200
    logs = ch.step(n)
    ft, f = ch.save_aos()
    oft, of = f0.copy(), f0.copy()
    ologs = o.run(oft, of, n)
    assert rel(ft, oft) <= 1e-12
    for k in ("K", "flowrate", "massrate", "stup", "stdown", "stress", "pgrad", "force"):
        scale = np.abs(ologs[k]).max()
        assert np.abs(logs[k] - ologs[k]).max() <= 1e-10 * scale, k


def test_statistics_and_fields_match_oracle():
    X, Y, Z = 16, 8, 10
    print("case: statistics/fields", flush=True)
    m = fields.posts_mask(X, Y)
    ch, o = make_pair(X, Y, Z, dnsgpu.CONST_Q, mask2d=(m, m))
    seeded(ch)
    ch.step(37, logs=False)
    ft, _ = ch.save_aos()
    st = ch.statistics()
    ost = o.statistics(ft)
    assert np.allclose(st, ost, rtol=1e-10, atol=1e-15)
    vel, den = ch.fields()
    ovel, oden = o.moments(ft)
    assert np.array_equal(vel, ovel) and np.array_equal(den, oden)


def test_restart_equals_uninterrupted():
    X, Y, Z = 16, 8, 10
    print("case: restart", flush=True)
    ch, _ = make_pair(X, Y, Z, dnsgpu.CONST_F, F=1e-5)
    seeded(ch)
    ch.step(20, logs=False)
    a, b = ch.save_aos()
    ch.step(30, logs=False)
    end1 = ch.save_aos()[0]
    ch2, _ = make_pair(X, Y, Z, dnsgpu.CONST_F, F=1e-5)
    ch2.load_aos(a, b)
    ch2.step(30, logs=False)
    end2 = ch2.save_aos()[0]
    assert np.array_equal(end1, end2)


@pytest.mark.parametrize("precision", [dnsgpu.FP32, dnsgpu.FP16])
def test_restart_equals_uninterrupted_reduced_storage(precision):
    """A checkpoint of an FP32- or FP16-shifted state (written as FP64
    f = w + g, the reference's lbf layout) reloads to the same stored values:
    the restarted run continues bitwise as the uninterrupted one (constant F,
    riblets included).  The stored values convert exactly to FP64 and back:
    w + g is exact in FP64 for an FP16 g·2^16 (11 significant bits) and for
    an FP32 g that is not smaller than about 2^-29 w, as everywhere in this
    flow."""
    X, Y, Z = 16, 16, 24
    print(f"case: restart, precision {precision}", flush=True)
    hb = fields.riblet_bands(X, Y, 8, 4)

    def make():
        ch = dnsgpu.Channel(X, Y, Z, Rebs=60.0, ub=0.05, CFL=0.25, TSTR=100, forcing=dnsgpu.CONST_F, force=1e-5,
                            precision=precision)
        ch.set_solid_bands(hb, hb)
        return ch

    ch = make()
    ch.init_field(dnsgpu.INIT_POISEUILLE, uc=0.06, amp=0.05, seed=3)
    ch.step(20, logs=False)
    a, b = ch.save_aos()
    ch.step(30, logs=False)
    end1 = ch.save_aos()[0]
    ch2 = make()
    ch2.load_aos(a, b)
    ch2.step(30, logs=False)
    end2 = ch2.save_aos()[0]
    assert np.array_equal(end1, end2), rel(end2, end1)
    ch.close()
    ch2.close()


def test_fp32_shifted_against_fp64():
    X, Y, Z = 16, 8, 24
    print("case: fp32 vs fp64", flush=True)
    m = fields.stripes_mask(X, Y, 4, 2, offset=1)
    ch64, _ = make_pair(X, Y, Z, dnsgpu.CONST_Q, mask2d=(m, m))
    ch32, _ = make_pair(X, Y, Z, dnsgpu.CONST_Q, mask2d=(m, m), precision=dnsgpu.FP32)
    seeded(ch64)
    seeded(ch32)
    l64 = ch64.step(300)
    l32 = ch32.step(300)
    s64 = ch64.statistics()
    s32 = ch32.statistics()
    # mean profile within 1e-4 relative, populations within 1e-5 of f
    assert np.abs(s32[0] - s64[0]).max() <= 1e-4 * np.abs(s64[0]).max()
    f64 = ch64.save_aos()[0]
    f32 = ch32.save_aos()[0]
    assert np.abs(f32 - f64).max() <= 1e-5
    assert np.abs(l32["stress"] - l64["stress"]).max() <= 1e-3 * np.abs(l64["stress"]).max()


def test_fp16_shifted_against_fp64():
    """FP16 storage of g = f - w (scaled by 2^16), FP32 collision: the
    stored populations carry 11 significant bits of g, so the state drifts
    from FP64 at about 1e-5 of f after 300 steps (measured 1.7e-5)."""
    X, Y, Z = 16, 8, 24
    print("case: fp16 vs fp64", flush=True)
    m = fields.stripes_mask(X, Y, 4, 2, offset=1)
    ch64, _ = make_pair(X, Y, Z, dnsgpu.CONST_Q, mask2d=(m, m))
    ch16, _ = make_pair(X, Y, Z, dnsgpu.CONST_Q, mask2d=(m, m), precision=dnsgpu.FP16)
    seeded(ch64)
    seeded(ch16)
    # the initial field itself is rounded once to FP16
    f0 = ch16.save_aos()[0]
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    g0 = fields.device_poiseuille(X, Y, Z, 0.06, 0.05, 7)[:, :, 1:Z + 1] - w
    assert np.abs(f0[:, :, 1:Z + 1] - w - g0).max() <= np.abs(g0).max() * 2.0 ** -11
    l64 = ch64.step(300)
    l16 = ch16.step(300)
    s64 = ch64.statistics()
    s16 = ch16.statistics()
    assert np.abs(s16[0] - s64[0]).max() <= 3e-4 * np.abs(s64[0]).max()
    assert np.abs(ch16.save_aos()[0] - ch64.save_aos()[0]).max() <= 1e-4
    assert np.abs(l16["stress"] - l64["stress"]).max() <= 2e-3 * np.abs(l64["stress"]).max()
    assert np.abs(l16["force"] - l64["force"]).max() <= 3e-4 * np.abs(l64["force"]).max()


def make_riblet_pair(X, Y, Z, forcing, spacing, height, F=0.0, precision=dnsgpu.FP64, Rebs=60.0, ub=0.05,
                     blocks=None):
    ch = dnsgpu.Channel(X, Y, Z, Rebs=Rebs, ub=ub, CFL=0.25, TSTR=5, forcing=forcing, force=F,
                        precision=precision)
    hb = fields.riblet_bands(X, Y, spacing, height) if blocks is None else fields.block_bands(X, Y, *blocks)
    ch.set_solid_bands(hb, hb)
    s = fields.band_mask(X, Y, Z, hb, hb)
    o = pyoracle.Oracle(X, Y, Z, pyoracle.omega(Rebs, ub, Z), ub, 0.25, TM=5,
                        forcing=1 if forcing == dnsgpu.CONST_F else 0, F=F, mask=s)
    return ch, o, s


@pytest.mark.parametrize("shape,spacing,height", [((8, 16, 24), 8, 4), ((4, 24, 140), 12, 6), ((4, 16, 512), 8, 5)])
def test_riblets_const_force_fp64_bitwise(shape, spacing, height):
    X, Y, Z = shape
    print(f"case: riblets const-F {shape} s={spacing} h={height}", flush=True)
    ch, o, s = make_riblet_pair(X, Y, Z, dnsgpu.CONST_F, spacing, height, F=2e-5)
    f0 = seeded(ch)
    f0[s != 0] = 0.0
    n = 40
    logs = ch.step(n)
    ft, f = ch.save_aos()
    oft, of = f0.copy(), f0.copy()
    ologs = o.run(oft, of, n)
    assert np.array_equal(ft, oft), rel(ft, oft)
    for k in ("K", "flowrate", "massrate", "stup", "stdown", "stress", "rho"):
        assert np.allclose(logs[k], ologs[k], rtol=1e-10, atol=1e-300), k


def test_riblets_const_q_fp64_and_outputs():
    X, Y, Z = 8, 16, 24
    print("case: riblets const-Q", flush=True)
    ch, o, s = make_riblet_pair(X, Y, Z, dnsgpu.CONST_Q, 8, 4)
    f0 = seeded(ch)
    f0[s != 0] = 0.0
    n = 150
    logs = ch.step(n)
    ft, _ = ch.save_aos()
    oft, of = f0.copy(), f0.copy()
    ologs = o.run(oft, of, n)
    assert rel(ft, oft) <= 1e-12
    for k in ("K", "flowrate", "massrate", "stress", "pgrad", "force", "rho"):
        assert np.abs(logs[k] - ologs[k]).max() <= 1e-10 * np.abs(ologs[k]).max(), k
    st = ch.statistics()
    assert np.allclose(st, o.statistics(ft), rtol=1e-10, atol=1e-15)
    vel, den = ch.fields()
    ovel, oden = o.moments(ft)
    assert np.array_equal(vel, ovel) and np.array_equal(den, oden)


# bars of the reduced storages against FP64 after 300 constant-Q steps:
# populations (absolute, f ~ 1/3) and the force
REDUCED_BARS = {dnsgpu.FP32: (1e-5, 1e-3), dnsgpu.FP16: (1e-4, 3e-3)}


@pytest.mark.parametrize("precision,texture",
                         [(p, t) for t in ("riblets", "blocks") for p in (dnsgpu.FP32, dnsgpu.FP16)]
                         + [(dnsgpu.FP32, "riblets_tall")])
def test_solid_bands_reduced_storage_against_fp64(texture, precision):
    X, Y, Z = {"riblets": (8, 16, 32), "blocks": (12, 16, 32), "riblets_tall": (4, 16, 512)}[texture]
    print(f"case: {texture} precision {precision} vs fp64", flush=True)
    kw = dict(blocks=(4, 4, 2, 2, 3)) if texture == "blocks" else {}
    sp, h = {"riblets": (8, 4), "blocks": (0, 0), "riblets_tall": (8, 5)}[texture]
    a, _, s = make_riblet_pair(X, Y, Z, dnsgpu.CONST_Q, sp, h, **kw)
    b, _, _ = make_riblet_pair(X, Y, Z, dnsgpu.CONST_Q, sp, h, precision=precision, **kw)
    seeded(a)
    seeded(b)
    la = a.step(300)
    lb = b.step(300)
    fa = a.save_aos()[0]
    fb = b.save_aos()[0]
    pop_bar, force_bar = REDUCED_BARS[precision]
    assert np.abs(fa - fb).max() <= pop_bar
    assert np.abs(la["force"] - lb["force"]).max() <= force_bar * np.abs(la["force"]).max()
    # solid nodes stay zero in the checkpoint layout
    assert np.all(fb[s != 0] == 0.0)


def test_fp16_tall_channel_bands_add_no_error():
    """FP16 storage at the largest zdv (512): the error against FP64 grows with
    the step count in the bulk of the channel, with or without riblets (2.0e-4
    in the populations after 300 steps here, at k = 490 far from any band):
    per-step changes below half an FP16 ulp of g 2^16 are lost while the tall
    channel's flow evolves slowly.  The band path must add nothing to it."""
    X, Y, Z = 4, 16, 512
    print("case: FP16 tall channel, riblets vs smooth", flush=True)

    def run(precision, bands):
        ch = dnsgpu.Channel(X, Y, Z, Rebs=60.0, ub=0.05, CFL=0.25, TSTR=5, forcing=dnsgpu.CONST_Q,
                            precision=precision)
        if bands:
            hb = fields.riblet_bands(X, Y, 8, 5)
            ch.set_solid_bands(hb, hb)
        seeded(ch)
        force = ch.step(300)["force"]
        f = ch.save_aos()[0]
        ch.close()
        return f, force

    err = {}
    for bands in (True, False):
        fa, ga = run(dnsgpu.FP64, bands)
        fb, gb = run(dnsgpu.FP16, bands)
        err[bands] = (float(np.abs(fa - fb).max()), float(np.abs(ga - gb).max() / np.abs(ga).max()))
    print("FP16 vs FP64 (population, relative force):", err, flush=True)
    assert err[False][0] <= 5e-4 and err[True][0] <= 5e-4, err
    assert err[True][0] <= 1.2 * err[False][0] + 1e-5, err
    assert err[False][1] <= 2e-2 and err[True][1] <= 2e-2, err


@pytest.mark.parametrize("blocks", [(4, 4, 2, 2, 3), (6, 8, 3, 2, 5)])
def test_blocks_const_force_fp64_bitwise(blocks):
    """Solid posts with side walls (height fields from the wall): lateral
    faces between columns of different height take the same generalised
    bounce-back as the riblet flanks."""
    X, Y, Z = 12, 16, 20
    print(f"case: blocks const-F {blocks}", flush=True)
    ch, o, s = make_riblet_pair(X, Y, Z, dnsgpu.CONST_F, 0, 0, F=2e-5, blocks=blocks)
    f0 = seeded(ch)
    f0[s != 0] = 0.0
    n = 40
    logs = ch.step(n)
    ft, _ = ch.save_aos()
    oft, of = f0.copy(), f0.copy()
    ologs = o.run(oft, of, n)
    assert np.array_equal(ft, oft), rel(ft, oft)
    for k in ("K", "flowrate", "massrate", "stress", "rho"):
        assert np.allclose(logs[k], ologs[k], rtol=1e-10, atol=1e-300), k


def _steady_flow_rate(X, Y, Z, F, hb=None, tol=1e-11, cap=200000):
    ch = dnsgpu.Channel(X, Y, Z, Rebs=100.0, ub=0.01, CFL=0.25, forcing=dnsgpu.CONST_F, force=F, omega=1.0)
    if hb is not None:
        ch.set_solid_bands(hb, hb)
    ch.init_rest()
    q_prev, steps = None, 0
    while steps < cap:
        q = ch.step(1000, cntr0=steps)["flowrate"][-1]
        steps += 1000
        if q_prev is not None and abs(q - q_prev) < tol * abs(q):
            break
        q_prev = q
    ch.close()
    return q


def test_grid_convergence_of_flow_over_rectangular_ribs():
    """Steady Stokes flow through a channel with streamwise rectangular ribs
    (width H/4, height H/8, one per spanwise period H) on both walls: the
    flow-rate ratio to the smooth channel converges under refinement
    (H = 16, 32, 64; measured 0.71946, 0.70980, 0.70578, 0.70416 at H = 128,
    observed order about 1.3: the rib corners are singular)."""
    print("case: grid convergence, rectangular ribs", flush=True)
    R = []
    for n in (1, 2, 4):
        Z = Y = 16 * n
        F = 3e-7 * (4.0 / n) ** 2
        hb = fields.block_bands(2, Y, 1, Y, 1, 4 * n, 2 * n)
        R.append(_steady_flow_rate(2, Y, Z, F, hb) / _steady_flow_rate(2, Y, Z, F))
    e1, e2 = R[0] - R[1], R[1] - R[2]
    assert e1 > 0 and e2 > 0, R
    order = np.log2(e1 / e2)
    assert 0.9 <= order <= 2.5, (R, order)
    assert abs(R[2] - 0.7058) < 2e-3, R


# relative change of the total x-momentum over 1000 steps with slip walls and
# no force (momentum is a collision and streaming invariant there): FP64
# sums in the reference's order; FP32 storage: the collision's coefficient
# sums are exact, so only unbiased rounding is left (single-precision
# weights fl(1/18), fl(1/36) would gain 1.5e-8 of the momentum per step,
# 1.5e-5 here); FP16 storage: its rounding of each stored population
MOMENTUM_BARS = {dnsgpu.FP64: 1e-12, dnsgpu.FP32: 1e-7, dnsgpu.FP16: 3e-5}


@pytest.mark.parametrize("precision", [dnsgpu.FP64, dnsgpu.FP32, dnsgpu.FP16])
def test_momentum_conserved_with_slip_walls(precision):
    X, Y, Z = 16, 16, 32
    print(f"case: slip walls, no force, precision {precision}", flush=True)
    # nu = 0.005: the seeded noise lives through the run, so nodes differ
    ch = dnsgpu.Channel(X, Y, Z, ub=0.05, CFL=0.25, forcing=dnsgpu.CONST_F, force=0.0, precision=precision,
                        omega=1.0 / (3 * 0.005 + 0.5))
    slip = np.full((X, Y), -2, dtype=np.int8)
    ch.set_walls(slip, slip)
    ch.init_field(dnsgpu.INIT_POISEUILLE, uc=0.08, amp=0.1, seed=4)
    j = ch.step(1000)["massrate"]
    drift = float(np.abs(j - j[0]).max() / abs(j[0]))
    print(f"relative momentum drift {drift:.3e}", flush=True)
    assert drift <= MOMENTUM_BARS[precision]


@pytest.mark.parametrize("precision", [dnsgpu.FP64, dnsgpu.FP32, dnsgpu.FP16])
def test_statistics_over_the_fluid_of_banded_planes(precision):
    """A velocity uniform over the fluid of a channel with riblets on one
    wall and blocks on the other: no fluctuation in any plane (the rms,
    <u'w'>, the density rms and K vanish), the means over the fluid are U and
    1, and the FP64 statistics and log row equal the oracle's."""
    import test_oracle_invariants as inv
    X, Y, Z, U = 6, 16, 20, 0.04
    print(f"case: uniform banded state, precision {precision}", flush=True)
    hb = fields.riblet_bands(X, Y, 8, 4)
    ht = fields.block_bands(X, Y, 3, 4, 1, 2, 3)
    f, s = inv.uniform_banded_state(X, Y, Z, hb, ht, U)
    with dnsgpu.Channel(X, Y, Z, Rebs=60.0, ub=0.05, CFL=0.25, forcing=dnsgpu.CONST_F, force=0.0,
                        precision=precision) as ch:
        ch.load_aos(f, f)
        ch.set_solid_bands(hb, ht)
        st = ch.statistics()
        row = ch.step(1)
    fl = slice(1, Z + 1)
    assert np.abs(st[3:7, fl]).max() <= 1e-14 * U
    assert np.abs(st[8, fl]).max() <= 1e-13
    # the stored populations of U: exact in FP64, rounded to FP32 / FP16
    tu, tr = {dnsgpu.FP64: (1e-12, 1e-14), dnsgpu.FP32: (1e-6, 1e-7), dnsgpu.FP16: (1e-3, 1e-4)}[precision]
    assert np.allclose(st[0, fl], U, rtol=tu) and np.allclose(st[7, fl], 1.0, rtol=tr)
    assert abs(row["K"][0]) <= 1e-28
    if precision == dnsgpu.FP64:
        o = pyoracle.Oracle(X, Y, Z, pyoracle.omega(60.0, 0.05, Z), 0.05, 0.25, forcing=1, F=0.0, mask=s)
        assert np.allclose(st, o.statistics(f), rtol=1e-12, atol=1e-15)  # rho rms: rounding of sums of ones
        ol = o.logs_of(f)
        # the wall stresses of a uniform velocity are rounding (1e-19) on both sides
        for k in ("flowrate", "massrate", "stup", "stdown", "rho"):
            assert np.isclose(row[k][0], ol[k], rtol=1e-12, atol=1e-17), k
