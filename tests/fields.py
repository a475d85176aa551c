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
