"""The seeded turbulent initial field (DNSGPU_INIT_TURBULENT) through its
replica in tests/fields.py, on the CPU: the properties that fix it, so that
the GPU test that compares the library with the replica bit for bit pins the
library's field too.

The reference starts from rest (DNS-solver/src/init.c:3-20) or from a
restart file; this generator seeds the turbulent configurations with a
log-law mean plus divergence-free Fourier modes.  What the mathematics fixes:
  * the rms of |u'| over the grid is amp times the log law's centreline
    velocity (the modes are orthogonal on the periodic planes),
  * every plane mean of u' vanishes, so the plane mean of u_x is the log
    law u_tau (ln(y+)/kappa + B) (u_tau y+ in the viscous sublayer),
  * u' vanishes at the walls: u'_x, u'_y like the distance, u'_z like its
    square (the sin^2 envelope of the vector potential),
  * the discrete divergence is the truncation error of a curl: O(h^2), a
    quarter when the grid is refined twice in every direction,
  * the populations are the equilibrium of that velocity at rho = 1.
"""
import math

import numpy as np
import pytest

import fields

OMEGA = 1.0 / (3.0 * 0.0054 * 96 / 180.0 + 0.5)  # the Re_tau 180 workload (u_tau 0.0054, h 96)
UTAU, AMP = 0.0054, 0.1


def fluct(X, Y, Z, nmodes, seed=2, amp=AMP, omega=OMEGA):
    t = fields.turbulent_tables(X, Y, Z, omega, UTAU, amp, seed, nmodes)
    u0, u1, u2 = fields.turbulent_velocity(X, Y, Z, omega, UTAU, amp, seed, nmodes, tables=t)
    return t, np.stack([u0 - t["U"][None, None, :], u1, u2])


@pytest.mark.parametrize("shape,nmodes", [((24, 20, 16), 3), ((32, 48, 40), 8), ((6, 5, 12), 8)])
def test_rms_is_amp_times_the_centreline_velocity(shape, nmodes):
    X, Y, Z = shape
    t, up = fluct(X, Y, Z, nmodes)
    rms = math.sqrt(float((up ** 2).sum(0).mean()))
    assert abs(rms - AMP * t["ucl"]) <= 1e-12 * AMP * t["ucl"]
    # the modes are clamped to the grid's resolvable ones (6x5: |i| <= 2, |j| <= 2)
    mx, my = min(nmodes, (X - 1) // 2), min(nmodes, (Y - 1) // 2)
    assert len(t["modes"]) == mx * (2 * my + 1) + my


def test_plane_means_and_log_law():
    X, Y, Z = 32, 48, 40
    t, up = fluct(X, Y, Z, 8)
    assert np.abs(up.mean(axis=(1, 2))).max() <= 1e-15 * t["ucl"]
    nu = (1. / 3.) * (1. / OMEGA - 0.5)
    y = np.arange(Z) + 0.5
    yw = np.minimum(y, Z - y)
    yp = yw * UTAU / nu
    law = UTAU * np.where(yp < 11.0, yp, np.log(yp) / 0.41 + 5.2)
    assert np.allclose(t["U"], law, rtol=1e-14, atol=0)
    # centreline velocity of the law at y = h
    assert abs(t["ucl"] - UTAU * (math.log(0.5 * Z * UTAU / nu) / 0.41 + 5.2)) <= 1e-15


def test_fluctuation_vanishes_at_the_walls():
    X, Y, Z = 32, 48, 128
    _, up = fluct(X, Y, Z, 8)
    for side in (slice(0, 3), slice(Z - 1, Z - 4, -1)):
        r = [np.sqrt((up[c][:, :, side] ** 2).mean(axis=(0, 1))) for c in range(3)]
        # rows at distance 1/2, 3/2, 5/2 from the wall
        for c in (0, 1):
            assert abs(r[c][1] / r[c][0] - 3.0) < 0.05, (c, r[c])
        assert abs(r[2][1] / r[2][0] - 9.0) < 0.3, r[2]
        assert abs(r[2][2] / r[2][0] - 25.0) < 1.5, r[2]


def divergence(up):
    """Central differences, periodic in x and y, interior rows in z."""
    ux, uy, uz = up
    d = ((np.roll(ux, -1, 0) - np.roll(ux, 1, 0)) + (np.roll(uy, -1, 1) - np.roll(uy, 1, 1))) / 2.0
    d = d[:, :, 1:-1] + (uz[:, :, 2:] - uz[:, :, :-2]) / 2.0
    dudx = np.abs((np.roll(ux, -1, 0) - np.roll(ux, 1, 0)) / 2.0).max()
    return np.abs(d).max() / dudx


def test_divergence_is_second_order():
    # the same modes on two grids, one refined twice in x, y and z: the
    # relative discrete divergence falls by four
    coarse = divergence(fluct(24, 24, 32, 3)[1])
    fine = divergence(fluct(48, 48, 64, 3)[1])
    assert coarse < 0.1
    assert 3.6 < coarse / fine < 4.4, (coarse, fine)
    # a planted sign error in one term of the curl is not divergence-free
    t, up = fluct(24, 24, 32, 3)
    bad = up.copy()
    bad[0] = -bad[0]
    assert divergence(bad) > 10 * coarse


def test_populations_are_the_equilibrium_of_the_velocity():
    X, Y, Z = 6, 8, 10
    f = fields.device_turbulent(X, Y, Z, OMEGA, UTAU, AMP, 3, 4)
    u0, u1, u2 = fields.turbulent_velocity(X, Y, Z, OMEGA, UTAU, AMP, 3, 4)
    g = f[:, :, 1:Z + 1, :]
    rho = g.sum(-1)
    u = (g @ fields.E) / rho[..., None]
    assert np.abs(rho - 1.0).max() <= 1e-15
    assert np.allclose(u[..., 0], u0, rtol=0, atol=1e-16)
    assert np.allclose(u[..., 1], u1, rtol=0, atol=1e-16)
    assert np.allclose(u[..., 2], u2, rtol=0, atol=1e-16)
    assert np.all(f[:, :, 0] == 0) and np.all(f[:, :, Z + 1] == 0)


def test_seed_changes_the_field_and_not_its_rms():
    a = fluct(24, 20, 16, 3, seed=2)
    b = fluct(24, 20, 16, 3, seed=9)
    assert np.abs(a[1] - b[1]).max() > 0.1 * AMP * a[0]["ucl"]
    ra = math.sqrt(float((a[1] ** 2).sum(0).mean()))
    rb = math.sqrt(float((b[1] ** 2).sum(0).mean()))
    assert abs(ra - rb) <= 1e-12 * ra
