"""Checks on the oracle that do not depend on the reference binary: closed
forms, conservation laws and symmetries of the D3Q19 BGK channel
(DNS-solver/src/xcomputations.c), chosen so that a dropped term, a wrong
sign, a wrong lattice table or a wrong index fails at least one of them."""
import numpy as np
import pytest

import fields
import pyoracle

W = fields.W
E = fields.E


def mk(X, Y, Z, omega=1.0, forcing=1, F=0.0, bot=None, top=None, ub=0.05):
    return pyoracle.Oracle(X, Y, Z, omega, ub, 0.25, forcing=forcing, F=F,
                           mask=pyoracle.wall_mask(X, Y, Z, bot, top))


def rest(X, Y, Z):
    f = np.zeros((X, Y, Z + 2, 19))
    f[:, :, 1:Z + 1] = W
    return f


def test_lattice_tables():
    # weights sum to 1, isotropic second moment 1/3, zero odd moments
    assert abs(W.sum() - 1) < 1e-15
    assert np.allclose((W[:, None, None] * E[:, :, None] * E[:, None, :]).sum(0), np.eye(3) / 3)
    assert np.allclose((W[:, None] * E).sum(0), 0)


def test_rest_state_is_a_fixed_point():
    o = mk(6, 5, 7, omega=1.3, F=0.0)
    ft = rest(6, 5, 7)
    f = ft.copy()
    o.run(ft, f, 10)
    # the 19 weights sum to 1 only to rounding, so the fixed point holds to it
    assert np.abs(ft - rest(6, 5, 7)).max() < 1e-16


def test_force_adds_exactly_G_of_momentum_per_node_per_step():
    # one collision on a uniform state: the x-momentum rises by G per node,
    # mass is unchanged (sum_a Fi = 0, sum_a e_x Fi = G)
    X, Y, Z = 4, 3, 5
    o = mk(X, Y, Z, omega=1.1, F=3e-4)
    f0 = fields.poiseuille_field(X, Y, Z, 0.05, noise=0.1, seed=2)
    ft = f0.copy()
    f = np.zeros_like(ft)
    o.collide_stream(ft, f, 3e-4)  # ft now holds the post-collision values
    fl = slice(1, Z + 1)
    dm = ft[:, :, fl].sum(-1) - f0[:, :, fl].sum(-1)
    dj = (ft[:, :, fl] @ E[:, 0]) - (f0[:, :, fl] @ E[:, 0])
    assert np.abs(dm).max() < 1e-15
    assert np.abs(dj - 3e-4).max() < 1e-15


def test_mass_conserved_with_slip_and_noslip_walls():
    X, Y, Z = 8, 8, 9
    m = fields.posts_mask(X, Y)
    o = mk(X, Y, Z, omega=1.7, F=1e-5, bot=m, top=fields.stripes_mask(X, Y, 4, 1))
    ft = fields.poiseuille_field(X, Y, Z, 0.05, noise=0.1, seed=5)
    f = ft.copy()
    m0 = ft.sum()
    o.run(ft, f, 50)
    assert abs(ft.sum() - m0) < 1e-13 * m0


@pytest.mark.parametrize("walls", ["noslip", "posts", "stripes", "random"])
def test_streaming_writes_every_slot_exactly_once(walls):
    X, Y, Z = 8, 8, 5
    rng = np.random.default_rng(4)
    if walls == "noslip":
        b = t = None
    elif walls == "posts":
        b = t = fields.posts_mask(X, Y)
    elif walls == "stripes":
        b, t = fields.stripes_mask(X, Y, 4, 2), fields.stripes_mask(X, Y, 8, 3, offset=5)
    else:
        b = np.where(rng.random((X, Y)) < 0.5, -2, -1).astype(np.int8)
        t = np.where(rng.random((X, Y)) < 0.5, -2, -1).astype(np.int8)
    cnt = mk(X, Y, Z, bot=b, top=t).coverage()
    assert (cnt[:, :, 1:Z + 1] == 1).all()
    assert (cnt[:, :, [0, Z + 1]] == 0).all()


def test_periodic_translation_equivariance_bitwise():
    X, Y, Z = 8, 6, 7
    m = fields.posts_mask(X, Y)
    f0 = fields.poiseuille_field(X, Y, Z, 0.05, noise=0.1, seed=9)
    a, fa = f0.copy(), f0.copy()
    mk(X, Y, Z, omega=1.6, F=2e-5, bot=m, top=m).run(a, fa, 20)
    sh = (3, 2)
    b = np.roll(f0, sh, axis=(0, 1)).copy()
    fb = b.copy()
    ms = np.roll(m, sh, axis=(0, 1)).copy()
    mk(X, Y, Z, omega=1.6, F=2e-5, bot=ms, top=ms).run(b, fb, 20)
    assert np.array_equal(np.roll(a, sh, axis=(0, 1)), b)


def test_wall_normal_mirror_symmetry():
    # mirror z -> Z+1-z and direction a -> mirror[a]: a symmetric field with
    # symmetric walls stays symmetric (to rounding: the sums run in a order)
    X, Y, Z = 6, 6, 8
    mir = [0, 1, 2, 3, 4, 6, 5, 7, 8, 9, 10, 12, 11, 14, 13, 16, 15, 18, 17]
    f0 = fields.poiseuille_field(X, Y, Z, 0.05, noise=0.1, seed=3)
    sym = 0.5 * (f0 + f0[:, :, ::-1][..., mir])
    m = fields.posts_mask(X, Y)
    a, fa = sym.copy(), sym.copy()
    mk(X, Y, Z, omega=1.5, F=2e-5, bot=m, top=m).run(a, fa, 30)
    assert np.abs(a - a[:, :, ::-1][..., mir]).max() < 1e-14


def test_all_slip_walls_give_plug_flow_without_wall_stress():
    # free-slip (specular) walls exert no stress: under a body force every
    # node gains the same momentum and the profile stays flat
    X, Y, Z = 4, 4, 6
    s = np.full((X, Y), -2, dtype=np.int8)
    o = mk(X, Y, Z, omega=1.2, F=1e-5, bot=s, top=s)
    ft = rest(X, Y, Z)
    f = ft.copy()
    o.run(ft, f, 40)
    ux = (ft[:, :, 1:Z + 1] @ E[:, 0]) / ft[:, :, 1:Z + 1].sum(-1)
    assert np.abs(ux - ux.mean()).max() < 1e-14
    assert abs(ux.mean() - 40 * 1e-5) < 1e-12


def test_laminar_poiseuille_profile():
    # steady body-force channel, halfway walls at y = 0 and H (y = k - 1/2):
    # the hydrodynamic velocity is u(y) = F/(2 nu) y (H - y), exact for
    # halfway bounce-back at tau = 1/2 + sqrt(3/16).  The reference reports
    # u = j/rho without the half-force shift of its force term
    # (xcomputations.c:323-336), so the printed velocity is u - F/2: a
    # wrong weight, sign or the Guo prefactor breaks this at 1e-12.
    X, Y, Z = 2, 2, 12
    tau = 0.5 + np.sqrt(3 / 16)
    nu = (tau - 0.5) / 3
    F = 1e-6
    o = mk(X, Y, Z, omega=1 / tau, F=F)
    ft = rest(X, Y, Z)
    f = ft.copy()
    o.run(ft, f, 6000)
    ux = ((ft[0, 0, 1:Z + 1] @ E[:, 0]) / ft[0, 0, 1:Z + 1].sum(-1))
    y = np.arange(1, Z + 1) - 0.5
    exact = F / (2 * nu) * y * (Z - y) - F / 2
    assert np.abs(ux - exact).max() < 1e-9 * exact.max()


def test_constant_q_feedback():
    # encalc.c:121: every step's force is the momentum deficit of the state
    # it acts on, G_n = rho_n ub - jx_n; at the steady state the force
    # balances the wall friction (laminar: G = 12 nu ub / H^2 to the
    # discretisation error) and the mean wall stress equals F h.
    X, Y, Z = 4, 4, 10
    ub = 0.02
    omega = 1.5
    nu = (1 / omega - 0.5) / 3
    o = pyoracle.Oracle(X, Y, Z, omega, ub, 0.25, forcing=0)
    ft = rest(X, Y, Z)
    f = ft.copy()
    for n in range(3000):
        rho = ft[:, :, 1:Z + 1].sum() / (X * Y * Z)
        j = (ft[:, :, 1:Z + 1] @ E[:, 0]).sum() / (X * Y * Z)
        logs = o.run(ft, f, 1, cntr0=n)
        assert abs(logs["force"][0] - (rho * ub - j)) < 1e-15
    G = logs["force"][0]
    assert abs(G - 12 * nu * ub / Z ** 2) < 0.02 * G
    assert abs(logs["stress"][0] - logs["pgrad"][0]) < 0.02 * abs(logs["pgrad"][0])


def test_slip_changes_the_answer():
    # the slip parity cases are not trivially equal to the no-slip ones
    X, Y, Z = 8, 8, 6
    f0 = fields.poiseuille_field(X, Y, Z, 0.05, noise=0.05, seed=1)
    a, fa = f0.copy(), f0.copy()
    b, fb = f0.copy(), f0.copy()
    mk(X, Y, Z, F=1e-5).run(a, fa, 5)
    m = fields.posts_mask(X, Y)
    mk(X, Y, Z, F=1e-5, bot=m, top=m).run(b, fb, 5)
    assert np.abs(a - b).max() > 1e-6


@pytest.mark.parametrize("spacing,height", [(8, 4), (6, 3), (16, 8)])
def test_riblet_streaming_exactly_once_and_mass(spacing, height):
    # solids inside the fluid rows: every fluid slot is written exactly once
    # by the generalised rule (bounce-back whenever the target is solid) and
    # no population is lost or created
    X, Y, Z = 4, 16, 20
    hb = fields.riblet_bands(X, Y, spacing, height)
    s = fields.band_mask(X, Y, Z, hb, hb)
    o = pyoracle.Oracle(X, Y, Z, 1.6, 0.05, 0.25, forcing=1, F=1e-5, mask=s)
    cnt = o.coverage()
    fluid = s == 0
    assert (cnt[fluid] == 1).all()
    assert (cnt[~fluid] == 0).all()
    ft = fields.poiseuille_field(X, Y, Z, 0.05, noise=0.1, seed=8, solid=~fluid)
    f = ft.copy()
    m0 = ft.sum()
    o.run(ft, f, 30)
    assert abs(ft.sum() - m0) < 1e-13 * m0
    assert (ft[~fluid] == 0).all()


def test_riblets_increase_drag_over_flat_wall_at_fixed_flow():
    # same bulk flow (constant Q), riblet walls need more force than flat
    # walls once laminar flow has settled (the riblets narrow the channel)
    X, Y, Z = 2, 16, 16
    o_flat = pyoracle.Oracle(X, Y, Z, 1.5, 0.02, 0.25, forcing=0)
    hb = fields.riblet_bands(X, Y, 8, 3)
    o_rib = pyoracle.Oracle(X, Y, Z, 1.5, 0.02, 0.25, forcing=0, mask=fields.band_mask(X, Y, Z, hb, hb))
    f0 = rest(X, Y, Z)
    a, fa = f0.copy(), f0.copy()
    b, fb = f0.copy(), f0.copy()
    b[fields.band_mask(X, Y, Z, hb, hb) != 0] = 0
    fb[:] = b
    la = o_flat.run(a, fa, 2000)
    lb = o_rib.run(b, fb, 2000)
    assert lb["force"][-1] > 1.2 * la["force"][-1]


@pytest.mark.parametrize("blocks", [(4, 4, 2, 2, 3), (8, 8, 3, 5, 6)])
def test_block_streaming_exactly_once_and_mass(blocks):
    # posts with side walls (height fields from both walls): every fluid
    # slot written exactly once, mass conserved, solid slots untouched
    X, Y, Z = 8, 16, 20
    hb = fields.block_bands(X, Y, *blocks)
    ht = fields.block_bands(X, Y, *blocks, offset_x=1, offset_y=2)
    s = fields.band_mask(X, Y, Z, hb, ht)
    o = pyoracle.Oracle(X, Y, Z, 1.6, 0.05, 0.25, forcing=1, F=1e-5, mask=s)
    cnt = o.coverage()
    fluid = s == 0
    assert (cnt[fluid] == 1).all()
    assert (cnt[~fluid] == 0).all()
    ft = fields.poiseuille_field(X, Y, Z, 0.05, noise=0.1, seed=9, solid=~fluid)
    f = ft.copy()
    m0 = ft.sum()
    o.run(ft, f, 30)
    assert abs(ft.sum() - m0) < 1e-13 * m0
    assert (ft[~fluid] == 0).all()


def uniform_banded_state(X, Y, Z, hb, ht, U=0.04):
    """feq(1, (U, 0, 0)) at every fluid node, zeros at wall rows and band nodes."""
    s = fields.band_mask(X, Y, Z, hb, ht)
    u = np.zeros((X, Y, Z + 2, 3))
    u[..., 0] = U
    f = fields.feq(np.ones((X, Y, Z + 2)), u)
    f[s != 0] = 0.0
    return np.ascontiguousarray(f), s


def test_statistics_over_the_fluid_of_banded_planes():
    # a velocity uniform over the fluid has no fluctuation in any plane, band
    # planes included: the rms, <u'w'>, the density rms and K vanish (a plane
    # mean over the whole plane with solids as zero would give an rms of
    # sqrt(phi) (1 - phi) U at fluid fraction phi), the means over the fluid
    # are U and 1, and the flow rate counts the fluid only
    X, Y, Z, U = 6, 16, 20, 0.04
    hb = fields.riblet_bands(X, Y, 8, 4)
    ht = fields.block_bands(X, Y, 3, 4, 1, 2, 3)
    f, s = uniform_banded_state(X, Y, Z, hb, ht, U)
    o = pyoracle.Oracle(X, Y, Z, 1.2, 0.05, 0.25, forcing=1, F=0.0, mask=s)
    st = o.statistics(f)
    fl = slice(1, Z + 1)
    assert np.abs(st[3:7, fl]).max() <= 1e-15 * U
    assert np.abs(st[8, fl]).max() <= 1e-14
    assert np.allclose(st[0, fl], U, rtol=1e-14) and np.allclose(st[7, fl], 1.0, rtol=1e-14)
    row = o.logs_of(f)
    assert abs(row["K"]) <= 1e-30
    fluid = (s[:, :, fl] == 0).sum()
    assert abs(row["flowrate"] - U * fluid / X) <= 1e-13 * U * fluid / X
    # the same state with flat walls: the reference's plane statistics
    f2, s2 = uniform_banded_state(X, Y, Z, np.zeros_like(hb), np.zeros_like(ht), U)
    o2 = pyoracle.Oracle(X, Y, Z, 1.2, 0.05, 0.25, forcing=1, F=0.0, mask=s2)
    assert abs(o2.logs_of(f2)["flowrate"] - U * Y * Z) <= 1e-13 * U * Y * Z
