"""The FP16 step with two rows per lane (one 4-byte access per direction
holding rows z and z + 1) against the one-row step.

Bars:
  * constant force: populations bitwise equal after the same steps (each
    node's arithmetic is the one-row form's; bands bounce on the read side,
    which gives every fluid node the write side's values); the logs taken
    from the plane sums (wall stress, flow rate, density) bitwise; K and
    the mass rate (per-lane sums) within 1e-12 relative (another summation
    order).
  * switching between the forms mid-run, and a two-row state joining the
    ensemble (whose FP16 step bounces on the write side), give the same
    populations as the one-row run throughout.
"""
import numpy as np
import pytest

import dnsgpu
import fields

pytestmark = pytest.mark.gpu

REBS, UB, F = 60.0, 0.05, 2e-5


def make(shape, texture, seed=7, forcing=dnsgpu.CONST_F):
    X, Y, Z = shape
    ch = dnsgpu.Channel(X, Y, Z, Rebs=REBS, ub=UB, CFL=0.25, TSTR=5, forcing=forcing,
                        force=F if forcing == dnsgpu.CONST_F else 0.0, precision=dnsgpu.FP16)
    if texture == "posts":
        m = fields.posts_mask(X, Y)
        ch.set_walls(m, m)
    elif texture == "stripes":
        ch.set_walls(fields.stripes_mask(X, Y, 6, 3), fields.stripes_mask(X, Y, 4, 1, offset=2))
    elif texture == "riblets":
        hb = fields.riblet_bands(X, Y, 8, 4)
        ch.set_solid_bands(hb, hb)
    elif texture == "tall_riblets":
        hb = fields.riblet_bands(X, Y, 8, 5)
        ch.set_solid_bands(hb, hb)
    elif texture == "blocks":
        ch.set_solid_bands(fields.block_bands(X, Y, 4, 4, 2, 2, 3), fields.block_bands(X, Y, 6, 4, 3, 1, 2, 1))
    ch.init_field(dnsgpu.INIT_POISEUILLE, uc=0.06, amp=0.05, seed=seed)
    return ch


def two_rows(ch):
    ch.set_step_pieces(1)  # two rows per lane take one z-piece
    blocks, waves = ch.step_shape()
    ch.set_step_shape(blocks, waves, 2)
    assert ch.step_shape(rows=True)[2] == 2


def one_row(ch):
    blocks, waves = ch.step_shape()
    ch.set_step_shape(blocks, waves, 1)


CASES = [
    ((16, 8, 10), "smooth"),
    ((2, 2, 4), "smooth"),          # the smallest grid
    ((16, 8, 10), "posts"),
    ((16, 12, 70), "stripes"),      # one 128-row chunk, 35 lanes
    ((8, 8, 130), "posts"),         # a one-lane second chunk
    ((8, 16, 24), "riblets"),
    ((4, 16, 256), "riblets"),      # two full chunks, shifts in LDS
    ((4, 16, 258), "riblets"),      # shifts read from global memory
    ((4, 16, 512), "tall_riblets"),  # the largest zdv
    ((12, 16, 20), "blocks"),
]


def compare_logs(l1, l2):
    for k in ("stress", "stup", "stdown", "pgrad", "force", "t", "flowrate", "rho"):
        assert np.array_equal(l1[k], l2[k]), k
    for k in ("K", "massrate"):
        assert np.allclose(l1[k], l2[k], rtol=1e-12, atol=1e-300), (k, np.abs(l1[k] - l2[k]).max())


@pytest.mark.parametrize("shape,texture", CASES)
def test_two_rows_per_lane_equals_one(shape, texture):
    print(f"case: fp16 two rows per lane {shape} {texture}", flush=True)
    a = make(shape, texture)
    b = make(shape, texture)
    two_rows(b)
    n = 40
    la = a.step(n, cntr0=0)  # the first step (local gather) is one-row in both
    lb = b.step(n, cntr0=0)
    fa, ga = a.save_aos()
    fb, gb = b.save_aos()
    assert np.array_equal(fa, fb) and np.array_equal(ga, gb)
    compare_logs(la, lb)
    s1, s2 = a.statistics(), b.statistics()
    assert np.array_equal(s1, s2)
    a.close()
    b.close()


@pytest.mark.parametrize("texture", ["riblets", "blocks"])
def test_switching_forms_mid_run(texture):
    """two rows, one row (a write-side step of a read-side state), two rows
    again: the one-row run's populations after every segment."""
    shape = (12, 16, 20) if texture == "blocks" else (8, 16, 24)
    print(f"case: fp16 rows 2 -> 1 -> 2, {texture}", flush=True)
    a = make(shape, texture)
    b = make(shape, texture)
    k = 12
    a.step(1, cntr0=0)
    b.step(1, cntr0=0)
    for seg, form in enumerate((two_rows, one_row, two_rows)):
        form(b)
        a.step(k, cntr0=1 + seg * k, logs=False)
        b.step(k, cntr0=1 + seg * k, logs=False)
        assert np.array_equal(a.save_aos()[0], b.save_aos()[0]), seg
    a.close()
    b.close()


def test_two_row_state_joins_the_ensemble():
    """A member whose state comes from two-row steps (read side) is
    advanced by the ensemble's FP16 kernel (write side) as the one-row
    single-case path advances it."""
    print("case: fp16 two-row state joins an ensemble", flush=True)
    shapes = [((8, 16, 24), "riblets"), ((12, 16, 20), "blocks"), ((16, 12, 70), "stripes")]
    ens = dnsgpu.Ensemble(precision=dnsgpu.FP16)
    members, refs = [], []
    for shape, tex in shapes:
        b = make(shape, tex)
        two_rows(b)
        b.step(9, cntr0=0, logs=False)
        ens.add(b)
        members.append(b)
        refs.append(make(shape, tex))
    ens.step(10)
    for (shape, tex), b, a in zip(shapes, members, refs):
        a.step(19, cntr0=0, logs=False)
        assert np.array_equal(a.save_aos()[0], b.save_aos()[0]), (shape, tex)
        a.close()
    ens.close()
    for b in members:
        b.close()


def test_shape_setter_rejects_two_rows_where_it_does_not_apply():
    print("case: two rows refused for FP32 and odd zdv", flush=True)
    ch = dnsgpu.Channel(8, 8, 24, precision=dnsgpu.FP32)
    blocks, waves = ch.step_shape()
    with pytest.raises(RuntimeError):
        ch.set_step_shape(blocks, waves, 2)
    ch.close()
    ch = dnsgpu.Channel(8, 8, 25, precision=dnsgpu.FP16)
    with pytest.raises(RuntimeError):
        ch.set_step_shape(blocks, waves, 2)
    ch.close()
