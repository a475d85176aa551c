"""Seeded test fields in the reference's AoS layout [X][Y][Z+2][19].

Test data only: an equilibrium distribution (the feq of
DNS-solver/src/xcomputations.c:330-331 at rho=1) of a laminar channel
profile plus seeded uniform noise in all three velocity components.
"""
import numpy as np

W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
E = np.array([
    [0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
    [1, 1, 0], [-1, 1, 0], [-1, -1, 0], [1, -1, 0], [1, 0, 1], [1, 0, -1], [-1, 0, -1],
    [-1, 0, 1], [0, 1, 1], [0, 1, -1], [0, -1, -1], [0, -1, 1]], dtype=np.float64)


def feq(rho, u):
    """Equilibrium populations for rho[...], u[..., 3] -> [..., 19]."""
    eu = u @ E.T
    uu = (u * u).sum(-1)[..., None]
    return W * rho[..., None] * (1 + 3 * eu + 4.5 * eu * eu - 1.5 * uu)


def poiseuille_field(X, Y, Z, uc, noise=0.05, seed=1, solid=None):
    """feq of u_x = 4 uc y(H-y)/H^2 (y = k - 0.5, H = Z) plus uniform noise of
    +-noise*uc in all components; wall rows (and `solid` nodes) are zero."""
    rng = np.random.default_rng(seed)
    k = np.arange(Z + 2, dtype=np.float64)
    y = k - 0.5
    ux = 4 * uc * y * (Z - y) / (Z * Z)
    u = np.zeros((X, Y, Z + 2, 3))
    u[..., 0] = ux[None, None, :]
    u += noise * uc * rng.uniform(-1, 1, size=u.shape)
    rho = np.ones((X, Y, Z + 2))
    f = feq(rho, u)
    f[:, :, 0, :] = 0
    f[:, :, Z + 1, :] = 0
    if solid is not None:
        f[solid] = 0
    return np.ascontiguousarray(f)


# --------------------------------------------------------------------------
# Bit-exact replica of the library's DNSGPU_INIT_POISEUILLE generator
# (dns-solvers_amd/csrc/lbm_io_kernels.h, init_kernel): counter-based
# splitmix64 noise and the feq expression order of xcomputations.c:331.
# Every step is an IEEE operation numpy performs identically.

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x):
    x = (x + np.uint64(0x9E3779B97F4A7C15))
    x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


def _urand(seed, ctr):
    with np.errstate(over="ignore"):
        h = _splitmix64(np.uint64(seed) ^ _splitmix64(ctr.astype(np.uint64)))
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0) * 2.0 - 1.0


def _edot(a, u0, u1, u2):
    comps = [(int(E[a, 0]), u0), (int(E[a, 1]), u1), (int(E[a, 2]), u2)]
    s = None
    for e, c in comps:
        if e == 0:
            continue
        if s is None:
            s = c if e > 0 else -c
        else:
            s = (s + c) if e > 0 else (s - c)
    return np.zeros_like(u0) if s is None else s


def feq_ref_order(u0, u1, u2):
    """feq(rho=1, u) in the reference's association order -> [..., 19]."""
    v = u0 * u0 + u1 * u1 + u2 * u2
    out = np.empty(u0.shape + (19,))
    for a in range(19):
        u = _edot(a, u0, u1, u2)
        if a == 0:
            out[..., a] = (W[a] * 1.0) * (1. - 1.5 * v)
        else:
            out[..., a] = (W[a] * 1.0) * (((1. + 3. * u) + (4.5 * u) * u) - 1.5 * v)
    return out


def device_poiseuille(X, Y, Z, uc, amp, seed):
    """AoS [X][Y][Z+2][19] of the library's POISEUILLE initial state (FP64)."""
    n = np.arange(X * Y * Z, dtype=np.uint64).reshape(X, Y, Z)
    z = (np.arange(Z, dtype=np.float64) + 0.5)[None, None, :]
    H = float(Z)
    u0 = 4.0 * uc * z * (H - z) / (H * H)
    u0 = np.broadcast_to(u0, (X, Y, Z)).copy()
    with np.errstate(over="ignore"):
        c4 = n * np.uint64(4)
    u0 = u0 + amp * uc * _urand(seed, c4)
    u1 = amp * uc * _urand(seed, c4 + np.uint64(1))
    u2 = amp * uc * _urand(seed, c4 + np.uint64(2))
    f = np.zeros((X, Y, Z + 2, 19))
    f[:, :, 1:Z + 1, :] = feq_ref_order(u0, u1, u2)
    return f


def posts_mask(X, Y):
    """Wall codes of slip_posts_init.c:11-41 on the periodic domain: 2x2
    slip squares with period 4 in x and y, starting at index 1."""
    i = np.arange(X)[:, None] % 4
    j = np.arange(Y)[None, :] % 4
    m = np.full((X, Y), -1, dtype=np.int8)
    m[((i == 1) | (i == 2)) & ((j == 1) | (j == 2))] = -2
    return m


def stripes_mask(X, Y, period, slip_width, offset=0, along="x"):
    """Streamwise-invariant slip stripes (grooves): slip where
    (j - offset) mod period < slip_width (j spanwise)."""
    j = (np.arange(Y) - offset) % period
    row = np.where(j < slip_width, -2, -1).astype(np.int8)
    return np.ascontiguousarray(np.broadcast_to(row[None, :], (X, Y)))


def riblet_bands(X, Y, spacing, height, offset=0.0):
    """Solid rows of streamwise-invariant triangular riblets (spanwise
    period `spacing`, tip height `height` cells): cell (y, z) is solid when
    its centre z + 1/2 lies below the triangle height at y + 1/2.  Same rule
    as dns-solvers_amd/driver/dnscase.c (dns_case_bands)."""
    yc = np.arange(Y) + 0.5 - offset
    phase = np.mod(yc, spacing) / spacing
    H = height * (1.0 - np.abs(2.0 * phase - 1.0))
    hb = np.where(H > 0.5, np.ceil(H - 0.5), 0).astype(np.int16)
    return np.ascontiguousarray(np.broadcast_to(hb[None, :], (X, Y)))


def block_bands(X, Y, period_x, period_y, block_x, block_y, height, offset_x=0, offset_y=0):
    """Solid posts with side walls (blocks): `height` solid rows over the
    footprint (x - offset_x) mod period_x < block_x and (y - offset_y) mod
    period_y < block_y.  Same rule as dns_case_bands (texture blocks)."""
    x = (np.arange(X) - offset_x) % period_x < block_x
    y = (np.arange(Y) - offset_y) % period_y < block_y
    return np.ascontiguousarray(np.where(x[:, None] & y[None, :], height, 0).astype(np.int16))


def band_mask(X, Y, Z, hb, ht, bottom=None, top=None):
    """Reference node codes [X][Y][Z+2] with solid bands as -1 nodes."""
    s = np.zeros((X, Y, Z + 2), dtype=np.int8)
    s[:, :, 0] = -1 if bottom is None else bottom
    s[:, :, Z + 1] = -1 if top is None else top
    k = np.arange(Z + 2)[None, None, :]
    hb = np.asarray(hb)[:, :, None]
    ht = np.asarray(ht)[:, :, None]
    s[(k >= 1) & (k <= hb)] = -1
    s[(k <= Z) & (k >= Z + 1 - ht)] = -1
    return s


# --------------------------------------------------------------------------
# Replica of the library's DNSGPU_INIT_TURBULENT generator (the table
# builder turbulent_tables in dns-solvers_amd/csrc/dnsgpu.hip and the
# TURBULENT branch of init_kernel in csrc/lbm_io_kernels.h).  The tables are
# built with Python's math module, which calls the C library's cos, sin and
# log on arguments formed in the library's order; the per-node sums are the
# library's IEEE products and sums in its order, so the result is bit-exact.

import math as _math

_PI = 3.141592653589793
_TWO_PI = 6.283185307179586


def turbulent_tables(X, Y, Z, omega, uc, amp, seed, nmodes):
    """The generator's modes, x/y/z tables and scale: a dict with 'modes'
    (list of (i, j, sc, a1, a2, p1, p2, kx, ky)), 'cx', 'sx' [mx+1][X], 'by'
    [nm][4][Y], 'U', 'env', 'denv' [Z], 'nrm', 'ucl', 'nu'."""
    kinv, Bplus = 1.0 / 0.41, 5.2
    nu = (1. / 3.) * (1. / omega - 0.5)
    hh = 0.5 * float(Z)
    mx, my = min(nmodes, (X - 1) // 2), min(nmodes, (Y - 1) // 2)
    modes = []
    for i in range(mx + 1):
        for j in range(-my, my + 1):
            if i == 0 and j <= 0:
                continue
            m = np.uint64((i * 64 + (j + 32)) & 0xFFFF)
            ctr = np.array([m * np.uint64(8) + np.uint64(k) for k in range(4)], dtype=np.uint64)
            r = _urand(seed, ctr)
            a1, a2 = float(r[0]), float(r[1])
            p1, p2 = _PI * float(r[2]), _PI * float(r[3])
            modes.append((i, j, 1.0 / float(i * i + j * j), a1, a2, p1, p2,
                          (_TWO_PI * float(i)) / float(X), (_TWO_PI * float(j)) / float(Y)))
    cx = np.array([[_math.cos((_TWO_PI * float((i * x) % X)) / float(X)) for x in range(X)] for i in range(mx + 1)])
    sx = np.array([[_math.sin((_TWO_PI * float((i * x) % X)) / float(X)) for x in range(X)] for i in range(mx + 1)])
    by = np.zeros((len(modes), 4, Y))
    for m, md in enumerate(modes):
        j = md[1]
        for q, p in enumerate((md[5], md[6])):
            for y in range(Y):
                n = ((j * y) % Y + Y) % Y
                arg = (_TWO_PI * float(n)) / float(Y) + p
                by[m, 2 * q, y] = _math.cos(arg)
                by[m, 2 * q + 1, y] = _math.sin(arg)
    U = np.zeros(Z)
    env = np.zeros(Z)
    denv = np.zeros(Z)
    for z in range(Z):
        yy = float(z) + 0.5
        yw = yy if yy < hh else float(Z) - yy
        yplus = yw * uc / nu
        uplus = yplus if yplus < 11.0 else (_math.log(yplus) * kinv + Bplus)
        U[z] = uc * uplus
        sn = _math.sin(_PI * (float(z) + 0.5) / float(Z))
        cs = _math.cos(_PI * (float(z) + 0.5) / float(Z))
        env[z] = sn * sn
        denv[z] = (((2.0 * sn) * cs) * _PI) / float(Z)
    S = 0.0
    for z in range(Z):
        d2, e2 = float(denv[z]) * float(denv[z]), float(env[z]) * float(env[z])
        for md in modes:
            _, _, sc, a1, a2, p1, p2, kx, ky = md
            q1, q2 = a1 * ky, a2 * kx
            term = (sc * sc) * (d2 * ((a1 * a1 + a2 * a2) * 0.5) +
                                e2 * (((q1 * q1 + q2 * q2) - 2.0 * (q1 * q2) * _math.cos(p1 - p2)) * 0.5))
            S = S + term
    ucl = uc * (_math.log(hh * uc / nu) * kinv + Bplus)
    nrm = amp * ucl / _math.sqrt(S / float(Z)) if S > 0.0 else 0.0
    return dict(modes=modes, cx=cx, sx=sx, by=by, U=U, env=env, denv=denv, nrm=nrm, ucl=ucl, nu=nu)


def turbulent_velocity(X, Y, Z, omega, uc, amp, seed, nmodes, xs=None, tables=None):
    """Velocity (u0, u1, u2), each [len(xs)][Y][Z], of the generator at the
    x-planes xs (default all), before the populations are formed."""
    t = tables or turbulent_tables(X, Y, Z, omega, uc, amp, seed, nmodes)
    xs = np.arange(X) if xs is None else np.asarray(xs)
    shp = (len(xs), Y, Z)
    env = np.broadcast_to(t["env"][None, None, :], shp)
    denv = np.broadcast_to(t["denv"][None, None, :], shp)
    ux = np.zeros(shp)
    uy = np.zeros(shp)
    uz = np.zeros(shp)
    for m, (i, j, sc, a1, a2, p1, p2, kx, ky) in enumerate(t["modes"]):
        cA = t["cx"][i][xs][:, None, None]
        sA = t["sx"][i][xs][:, None, None]
        cb1, sb1, cb2, sb2 = (t["by"][m, q][None, :, None] for q in range(4))
        c1 = cA * cb1 - sA * sb1
        s1 = sA * cb1 + cA * sb1
        c2 = cA * cb2 - sA * sb2
        s2 = sA * cb2 + cA * sb2
        ux = ux + sc * (-((denv * a2) * c2))
        uy = uy + sc * ((denv * a1) * c1)
        uz = uz + sc * (env * ((a1 * ky) * s1 - (a2 * kx) * s2))
    u0 = t["U"][None, None, :] + t["nrm"] * ux
    return u0, t["nrm"] * uy, t["nrm"] * uz


def device_turbulent(X, Y, Z, omega, uc, amp, seed, nmodes, xs=None):
    """AoS [len(xs)][Y][Z+2][19] of the library's TURBULENT initial state
    (FP64) at the x-planes xs (default all)."""
    u0, u1, u2 = turbulent_velocity(X, Y, Z, omega, uc, amp, seed, nmodes, xs)
    f = np.zeros(u0.shape[:2] + (Z + 2, 19))
    f[:, :, 1:Z + 1, :] = feq_ref_order(u0, u1, u2)
    return f
