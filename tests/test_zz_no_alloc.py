"""dnsgpu_fields_noalloc against dnsgpu_fields: the same velocity and density
fields bit for bit (FP64, FP32 and FP16 storage, flat walls and solid bands),
on two calls in a row, the second with fewer x-planes, and on a second
handle of another grid sharing the device's static staging; a request
beyond the staging is refused with a clear error."""
import numpy as np
import pytest

import dnsgpu
import fields

pytestmark = pytest.mark.gpu

MAX_DOUBLES = 8 << 20  # DNSGPU_FIELDS_NOALLOC_MAX_DOUBLES


def channel(X, Y, Z, precision, bands):
    ch = dnsgpu.Channel(X, Y, Z, Rebs=60.0, ub=0.05, CFL=0.25, forcing=dnsgpu.CONST_F, force=2e-5,
                        precision=precision)
    if bands:
        hb = fields.riblet_bands(X, Y, 8, 4)
        ch.set_solid_bands(hb, hb)
    else:
        m = fields.posts_mask(X, Y)
        ch.set_walls(m, m)
    ch.init_field(dnsgpu.INIT_POISEUILLE, uc=0.06, amp=0.05, seed=5)
    ch.step(7)
    return ch


@pytest.mark.parametrize("precision", [dnsgpu.FP64, dnsgpu.FP32, dnsgpu.FP16])
@pytest.mark.parametrize("bands", [False, True])
def test_noalloc_fields_equal_the_allocating_entry_point(precision, bands):
    print(f"case: fields_noalloc precision {precision} bands {bands}", flush=True)
    with channel(12, 16, 20, precision, bands) as ch:
        for x0, nx in ((0, 12), (3, 5)):  # the second call smaller than the first
            v0, d0 = ch.fields_slab(x0, nx)
            v1, d1 = ch.fields_slab(x0, nx, noalloc=True)
            assert np.array_equal(v0, v1) and np.array_equal(d0, d1), (x0, nx)
            assert np.all(np.isfinite(v1)) and np.abs(d1[:, :, 1:-1]).max() > 0.5
        # another handle, another grid, the same static staging
        with channel(8, 24, 36, precision, bands) as other:
            v0, d0 = other.fields_slab(1, 6)
            v1, d1 = other.fields_slab(1, 6, noalloc=True)
            assert np.array_equal(v0, v1) and np.array_equal(d0, d1)
        # the first handle again after the other one used the staging
        v0, d0 = ch.fields_slab(2, 9)
        v1, d1 = ch.fields_slab(2, 9, noalloc=True)
        assert np.array_equal(v0, v1) and np.array_equal(d0, d1)


def test_noalloc_refuses_more_than_the_staging():
    print("case: fields_noalloc beyond the staging", flush=True)
    X, Y, Z = 40, 256, 254  # 40 planes need 40*256*256*4 = 10.5 M doubles
    per_plane = Y * (Z + 2) * 4
    fit = MAX_DOUBLES // per_plane
    with dnsgpu.Channel(X, Y, Z) as ch:
        with pytest.raises(dnsgpu.DnsGpuError, match="static array"):
            ch.fields_slab(0, fit + 1, noalloc=True)
        v1, d1 = ch.fields_slab(0, fit, noalloc=True)
        v0, d0 = ch.fields_slab(0, fit)
        assert np.array_equal(v0, v1) and np.array_equal(d0, d1)
