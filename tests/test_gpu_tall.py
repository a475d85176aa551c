"""Columns taller than 512 rows (up to zdv = 1024): the step kernel splits
each column into z-pieces of at most 512 rows, one group of blocks per
piece, and the moments and statistics kernels take the piece from their
grid's y index.  The same parity bars as the shorter columns, against the
oracle (pinned bitwise to the reference):
  * constant F, FP64: populations bitwise equal after N steps, for two
    launch shapes (the state does not depend on the grid); log rows within
    1e-10 relative;
  * constant Q, FP64: populations within 1e-12 relative, log rows within
    1e-10; the first step's force comes from the moments kernel; two-pass
    statistics within 1e-10 and the fields bitwise;
  * riblets (solid bands, read-side bounce-back), constant F: bitwise;
  * FP32-shifted storage with riblets, constant Q, against FP64: the bars
    of the shorter columns (1e-5 in the populations, 1e-3 of the force).
The step grid is a multiple of the piece count; two rows per lane (FP16)
takes one piece only and falls back to one row; an ensemble member above 512
rows runs in z-pieces too (one tile and record per piece), bitwise the
single-case path under constant force and the same alone and in a mix.

Columns of 6 or 8 chunks (321-384, 449-512 rows) may also run as two
pieces (dnsgpu_set_step_pieces; the default for FP64 and FP16 storage): the
same state bitwise as one piece and as the oracle, under constant force.
"""
import numpy as np
import pytest

import dnsgpu
import fields
import pyoracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def seeded(ch, uc=0.06, amp=0.05, seed=7):
    ch.init_field(dnsgpu.INIT_POISEUILLE, uc=uc, amp=amp, seed=seed)
    return fields.device_poiseuille(ch.X, ch.Y, ch.Z, uc, amp, seed)


def walls_of(X, Y, walls):
    if walls == "posts":
        m = fields.posts_mask(X, Y)
        return m, m
    if walls == "stripes":
        return fields.stripes_mask(X, Y, 6, 3), fields.stripes_mask(X, Y, 4, 1, offset=2)
    return None, None


def make(X, Y, Z, forcing, F=0.0, walls="smooth", precision=dnsgpu.FP64, Rebs=60.0, ub=0.05, CFL=0.25,
         TSTR=100):
    ch = dnsgpu.Channel(X, Y, Z, Rebs=Rebs, ub=ub, CFL=CFL, TSTR=TSTR, forcing=forcing, force=F,
                        precision=precision)
    bot, top = walls_of(X, Y, walls)
    if bot is not None:
        ch.set_walls(bot, top)
    o = pyoracle.Oracle(X, Y, Z, pyoracle.omega(Rebs, ub, Z), ub, CFL, TM=TSTR,
                        forcing=1 if forcing == dnsgpu.CONST_F else 0, F=F,
                        mask=pyoracle.wall_mask(X, Y, Z, bot, top))
    return ch, o


LOG_KEYS = ("K", "flowrate", "massrate", "stup", "stdown", "stress", "pgrad", "rho")


@pytest.mark.parametrize("shape,walls", [
    ((4, 4, 640), "posts"),      # 10 chunks: two pieces of 5
    ((3, 4, 1024), "stripes"),   # the tallest column: two pieces of 8
    ((2, 3, 577), "smooth"),     # a one-row last chunk in the second piece
])
def test_tall_const_force_fp64_bitwise(shape, walls):
    X, Y, Z = shape
    n = 60
    ostate = None
    for blocks in (0, 6):  # the library's default grid, then 6 blocks of one wave
        print(f"case: tall const-F FP64 {shape} {walls}, grid {blocks or 'default'}", flush=True)
        ch, o = make(X, Y, Z, dnsgpu.CONST_F, F=2e-5, walls=walls)
        f0 = seeded(ch)
        if blocks:
            ch.set_step_shape(blocks, 1)
        b, w = ch.step_shape()
        assert b % 2 == 0, (b, w)  # a multiple of the two pieces
        logs = ch.step(n, cntr0=0)
        ft, f = ch.save_aos()
        ch.close()
        if ostate is None:
            oft, of = f0.copy(), f0.copy()
            ologs = o.run(oft, of, n)
            ostate = (oft, of, ologs)
        oft, of, ologs = ostate
        assert np.array_equal(ft, oft), rel(ft, oft)
        assert np.array_equal(f[:, :, 1:Z + 1], of[:, :, 1:Z + 1])
        for k in LOG_KEYS:
            assert np.allclose(logs[k], ologs[k], rtol=1e-10, atol=1e-300), (k, logs[k][:3], ologs[k][:3])
        assert np.array_equal(logs["t"], ologs["t"])


def test_tall_const_q_fp64_statistics_and_fields():
    X, Y, Z = 4, 4, 1024
    print(f"case: tall const-Q FP64 {X}x{Y}x{Z} posts", flush=True)
    ch, o = make(X, Y, Z, dnsgpu.CONST_Q, walls="posts")
    f0 = seeded(ch)
    n = 100
    logs = ch.step(n)
    ft, _ = ch.save_aos()
    oft, of = f0.copy(), f0.copy()
    ologs = o.run(oft, of, n)
    assert rel(ft, oft) <= 1e-12
    for k in LOG_KEYS + ("force",):
        scale = np.abs(ologs[k]).max()
        assert np.abs(logs[k] - ologs[k]).max() <= 1e-10 * scale, k
    st = ch.statistics()
    assert np.allclose(st, o.statistics(ft), rtol=1e-10, atol=1e-15)
    vel, den = ch.fields()
    ovel, oden = o.moments(ft)
    assert np.array_equal(vel, ovel) and np.array_equal(den, oden)
    ch.close()


def make_riblets(X, Y, Z, forcing, F=0.0, precision=dnsgpu.FP64):
    ch = dnsgpu.Channel(X, Y, Z, Rebs=60.0, ub=0.05, CFL=0.25, TSTR=5, forcing=forcing, force=F,
                        precision=precision)
    hb = fields.riblet_bands(X, Y, 8, 5)
    ch.set_solid_bands(hb, hb)
    s = fields.band_mask(X, Y, Z, hb, hb)
    o = pyoracle.Oracle(X, Y, Z, pyoracle.omega(60.0, 0.05, Z), 0.05, 0.25, TM=5,
                        forcing=1 if forcing == dnsgpu.CONST_F else 0, F=F, mask=s)
    return ch, o, s


def test_tall_riblets_const_force_fp64_bitwise():
    X, Y, Z = 4, 16, 1024
    print(f"case: tall riblets const-F FP64 {X}x{Y}x{Z} s=8 h=5", flush=True)
    ch, o, s = make_riblets(X, Y, Z, dnsgpu.CONST_F, F=2e-5)
    f0 = seeded(ch)
    f0[s != 0] = 0.0
    n = 40
    logs = ch.step(n)
    ft, _ = ch.save_aos()
    oft, of = f0.copy(), f0.copy()
    ologs = o.run(oft, of, n)
    assert np.array_equal(ft, oft), rel(ft, oft)
    for k in ("K", "flowrate", "massrate", "stup", "stdown", "stress", "rho"):
        assert np.allclose(logs[k], ologs[k], rtol=1e-10, atol=1e-300), k
    ch.close()


def test_tall_riblets_fp32_against_fp64():
    X, Y, Z = 4, 16, 1024
    print(f"case: tall riblets const-Q FP32 vs FP64 {X}x{Y}x{Z}", flush=True)
    a, _, s = make_riblets(X, Y, Z, dnsgpu.CONST_Q)
    b, _, _ = make_riblets(X, Y, Z, dnsgpu.CONST_Q, precision=dnsgpu.FP32)
    seeded(a)
    seeded(b)
    la = a.step(300)
    lb = b.step(300)
    fa = a.save_aos()[0]
    fb = b.save_aos()[0]
    assert np.abs(fa - fb).max() <= 1e-5
    assert np.abs(la["force"] - lb["force"]).max() <= 1e-3 * np.abs(la["force"]).max()
    assert np.all(fb[s != 0] == 0.0)
    a.close()
    b.close()


def test_tall_launch_shapes():
    """The grid stays a multiple of the piece count for any request and the
    autotuner keeps it so; two rows per lane (FP16) and more waves than the
    LDS of a piece allows are refused; every shape gives the same state
    (constant F)."""
    X, Y, Z = 6, 4, 1000
    print(f"case: tall launch shapes FP16 {X}x{Y}x{Z}", flush=True)
    ref = None
    for shape in ((7, 3), (1, 1), (5, 2), None):
        ch = dnsgpu.Channel(X, Y, Z, Rebs=60.0, ub=0.05, CFL=0.25, forcing=dnsgpu.CONST_F, force=2e-5,
                            precision=dnsgpu.FP16)
        seeded(ch)
        if shape is None:
            ch.autotune(trial_steps=2)
            with pytest.raises(dnsgpu.DnsGpuError, match="rows per lane"):
                ch.set_step_shape(8, 2, 2)
            with pytest.raises(dnsgpu.DnsGpuError, match="waves"):
                ch.set_step_shape(8, 4, 1)  # 4 x 16 KiB of plane sums + the block's globals > 64 KiB
        else:
            ch.set_step_shape(*shape)
        b, w, r = ch.step_shape(rows=True)
        print(f"requested {shape}: {b} blocks x {w} waves, {r} rows per lane", flush=True)
        assert b >= 2 and b % 2 == 0 and r == 1
        if shape is not None:
            assert b == max(2, shape[0] // 2 * 2) and w == shape[1]
        ch.step(20, logs=False)
        ft = ch.save_aos()[0]
        ch.close()
        if ref is None:
            ref = ft
        assert np.array_equal(ft, ref)


@pytest.mark.parametrize("precision", [dnsgpu.FP64, dnsgpu.FP32])
def test_tall_member_in_an_ensemble(precision):
    """A member of 640 rows (two z-pieces: one tile and record per piece)
    beside a short one: bitwise the single-case path's state under constant
    force, and the same state alone and in the mix."""
    print(f"case: ensemble with a 640-row member, precision {precision}", flush=True)
    n, F = 12, 2e-5
    shapes = [(4, 6, 640), (8, 8, 40)]

    def member(shape):
        X, Y, Z = shape
        ch = dnsgpu.Channel(X, Y, Z, Rebs=60.0, ub=0.05, CFL=0.25, TSTR=5, forcing=dnsgpu.CONST_F, force=F,
                            precision=precision)
        m = fields.posts_mask(X, Y)
        ch.set_walls(m, m)
        ch.init_field(dnsgpu.INIT_POISEUILLE, uc=0.06, amp=0.05, seed=7)
        return ch

    assert dnsgpu.member_footprint(4, 6, 640, precision) > dnsgpu.footprint(4, 6, 640, precision)
    mix = dnsgpu.Ensemble(precision=precision)
    members = [member(s) for s in shapes]
    for ch in members:
        mix.add(ch)
    mix.step(n)
    alone = dnsgpu.Ensemble(precision=precision)
    solo = member(shapes[0])
    alone.add(solo)
    alone.step(n)
    for shape, ch in zip(shapes, members):
        ref = member(shape)
        ref.step(n + 1, cntr0=0)
        assert np.array_equal(ch.save_aos()[0], ref.save_aos()[0]), shape
        ref.close()
    assert np.array_equal(solo.save_aos()[0], members[0].save_aos()[0])
    ra, rm = alone.rows(solo, n), mix.rows(members[0], n)
    for k in dnsgpu.LOG_FIELDS:
        assert np.array_equal(ra[k], rm[k]), k
    mix.close()
    alone.close()
    for ch in members + [solo]:
        ch.close()


@pytest.mark.parametrize("Z,precision,default", [(384, dnsgpu.FP64, 2), (512, dnsgpu.FP32, 1),
                                                 (384, dnsgpu.FP16, 2)])
def test_two_pieces_of_a_short_column(Z, precision, default):
    X, Y = 4, 4
    n = 40
    print(f"case: z-pieces {X}x{Y}x{Z} precision {precision}, one and two pieces", flush=True)
    out = {}
    for pieces in (1, 2):
        ch, o = make(X, Y, Z, dnsgpu.CONST_F, F=2e-5, walls="posts", precision=precision)
        f0 = seeded(ch)
        assert ch.step_pieces() == default
        ch.set_step_pieces(pieces)
        b, w = ch.step_shape()
        assert ch.step_pieces() == pieces and b % pieces == 0, (b, w)
        logs = ch.step(n, cntr0=0)
        out[pieces] = (ch.save_aos()[0], logs)
        ch.close()
    assert np.array_equal(out[1][0], out[2][0])
    for k in LOG_KEYS:
        assert np.allclose(out[1][1][k], out[2][1][k], rtol=1e-10, atol=1e-300), k
    if precision == dnsgpu.FP64:
        oft, of = f0.copy(), f0.copy()
        o.run(oft, of, n)
        assert np.array_equal(out[2][0], oft)


def test_piece_counts_and_rows_per_lane():
    print("case: z-piece options", flush=True)
    for Z, ok in ((320, False), (384, True), (448, False), (512, True), (256, False)):
        ch = dnsgpu.Channel(4, 4, Z, Rebs=60.0, ub=0.05, CFL=0.25, precision=dnsgpu.FP32)
        ch.init_rest()
        if ok:
            ch.set_step_pieces(2)
            assert ch.step_pieces() == 2
        else:
            with pytest.raises(dnsgpu.DnsGpuError, match="z-pieces"):
                ch.set_step_pieces(2)
            assert ch.step_pieces() == 1
        with pytest.raises(dnsgpu.DnsGpuError, match="z-pieces"):
            ch.set_step_pieces(3)
        ch.close()
    # FP16 two rows per lane take one piece
    ch = dnsgpu.Channel(8, 8, 384, Rebs=60.0, ub=0.05, CFL=0.25, precision=dnsgpu.FP16)
    ch.init_rest()
    assert ch.step_pieces() == 2
    ch.set_step_shape(8, 2, 2)
    assert ch.step_shape(rows=True)[2] == 2 and ch.step_pieces() == 1
    with pytest.raises(dnsgpu.DnsGpuError, match="one z-piece"):
        ch.set_step_pieces(2)
    ch.set_step_shape(8, 2, 1)
    ch.set_step_pieces(2)
    assert ch.step_pieces() == 2 and ch.step_shape()[0] % 2 == 0
    # the autotuner keeps a valid combination
    ch.autotune(trial_steps=2)
    b, w, r = ch.step_shape(rows=True)
    p = ch.step_pieces()
    print(f"autotuned: {b} blocks x {w} waves, {r} rows per lane, {p} pieces", flush=True)
    assert p in (1, 2) and b % p == 0 and (r == 1 or p == 1)
    ch.close()
