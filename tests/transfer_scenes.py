"""Scenes and closed forms for the SH transfer-matrix tests
(RenderEngine::shComputeTMatrix, Renderer/RenderEngine.cpp:410-446, and
__raygen__ComputeTMatrix, Renderer/device_functions/cameraShTmatrix.cu).

Floor scene: a large diffuse cookTorrance floor (ks = 0, normal +z) and
small triangles 1 above it whose vertices are the transfer points.  Each
basis map i is a constant colour col_i except for one texel of the -z face
that is 1e7 times brighter: it takes almost all of the map's sampling
tables, so an environment light sample at the floor lands below the floor
(no contribution) and the environment pdf of every upward direction stays
below the 1e-4 cut-off.  A path from a vertex then contributes only when it
goes down, hits the floor and its BSDF continuation leaves upwards, with
weight pb^2 / (pb^2 + pe^2) ~ 1.  Its expectation is

  T[k][i][c] = col_i[c] * Z_k[c],
  Z_k[c] = integral over the lower hemisphere of Y'_k(w) rho_c(-w_z) dw,
  rho_c(mu_o) = integral over the upper hemisphere of f_c(wi, wo) w_B dwi

with Y'_k = Y_k / (1 + 0.005 l^2 (l+1)^2), f the BSDF times the cosine.
rho depends on wo only through mu_o (isotropic BSDF, constant radiance), so
only the zonal coefficients (m = 0) are non-zero, and Z_k follows from the
Legendre polynomials (numpy), independently of the SH code under test."""
import numpy as np
from numpy.polynomial.legendre import leggauss, legval

FLOOR_MAT = [0.7, 0.5, 0.3, 0.0, 0.0, 0.0, 0.8, 2]
CHIP_MAT = [0.5, 0.5, 0.5, 0.0, 0.0, 0.0, 0.5, 2]


def floor_scene(n_chips: int = 8, half: float = 1e4, mirror: bool = False):
    """vertices, indices, material ids, materials and the chip vertices."""
    fl = np.array([[-half, -half, 0], [half, -half, 0], [half, half, 0], [-half, half, 0]],
                  np.float32)
    verts, tris = [fl], [[0, 1, 2], [0, 2, 3]]  # counter-clockwise from +z
    nv = 4
    for j in range(n_chips):
        x, y = (j % 4) - 1.5, (j // 4) - 0.5
        verts.append(np.array([[x, y, 1.0], [x + 0.004, y, 1.0], [x, y + 0.004, 1.0]],
                              np.float32))
        tris.append([nv, nv + 1, nv + 2])
        nv += 3
    floor = list(FLOOR_MAT)
    if mirror:
        floor[7] = 3
    return (np.concatenate(verts), np.array(tris, np.int32),
            np.array([0, 0] + [1] * n_chips, np.int32), [floor, CHIP_MAT], np.arange(4, nv))


def map_colours(nc: int) -> np.ndarray:
    return np.array([[1.0 + i, 2.0, 0.5 + 0.25 * i] for i in range(nc)])


def floor_basis(nc: int, n: int = 8) -> np.ndarray:
    cols = map_colours(nc)
    basis = np.empty((nc, 6, n, n, 3), np.float32)
    for i in range(nc):
        basis[i] = cols[i]
        basis[i, 4, n // 2 - 1, n // 2 + 1] = cols[i] * 1e7
    return basis


def floor_expectation(orc, degree: int, basis: np.ndarray, n_mu: int = 16, n_in: int = 32,
                      n_phi: int = 64):
    """Expected T[k][i][c] of the floor scene ([nc][nc][3], k then i) by
    Gauss-Legendre quadrature, the BSDF, its pdf and the environment pdf
    taken from the oracle's own pieces (pinned by the direct-light tests).
    Also returns the largest environment pdf met over the upper hemisphere."""
    nc = (degree + 1) ** 2
    n = basis.shape[2]
    up = np.array([0.0, 0.0, 1.0])
    mo, wo = leggauss(n_mu)
    mo, wo = 0.5 * (mo + 1), 0.5 * wo        # mu_o in (0, 1)
    mi, wi = leggauss(n_in)
    mi, wi = 0.5 * (mi + 1), 0.5 * wi        # mu_i in (0, 1)
    phis = (np.arange(n_phi) + 0.5) * 2 * np.pi / n_phi
    dirs = [np.array([np.sqrt(1 - m * m) * np.cos(p), np.sqrt(1 - m * m) * np.sin(p), m])
            for m in mi for p in phis]
    dw = np.repeat(wi, n_phi) * 2 * np.pi / n_phi
    tabs = orc.env_tables(basis[0])
    pe = np.array([orc.env_pdf(tabs, n, d) for d in dirs])
    rho = np.zeros((n_mu, 3))
    for a, m in enumerate(mo):
        vo = np.array([np.sqrt(1 - m * m), 0.0, m])
        for d, w, p in zip(dirs, dw, pe):
            pb = orc.ct_pdf(FLOOR_MAT, up, up, d, vo)
            wb = pb * pb / max(pb * pb + p * p, 1e-8)
            rho[a] += orc.ct_eval(FLOOR_MAT, up, up, d, vo) * wb * w
    z = np.zeros((nc, 3))
    for l in range(degree + 1):
        c = np.zeros(l + 1)
        c[l] = 1.0
        y = np.sqrt((2 * l + 1) / (4 * np.pi)) * legval(-mo, c)   # start direction mu = -mu_o
        filt = 1.0 / (1.0 + 0.005 * l * l * (l + 1) ** 2)
        z[l * l + l] = 2 * np.pi * filt * (wo * y) @ rho
    cols = map_colours(nc)
    return z[:, None, :] * cols[None, :, :], float(pe.max())


def panel_scene():
    """Two facing diffuse panels 0.6 apart (interreflections, paths past the
    roulette depth), a mirror panel, and transfer points on both panels."""
    def quad(c, u, v):
        c, u, v = (np.asarray(x, np.float32) for x in (c, u, v))
        return np.array([c - u - v, c + u - v, c + u + v, c - u + v], np.float32)
    vs = [quad([0, 0, 0], [1, 0, 0], [0, 1, 0]),          # floor, normal +z
          quad([0, 0, 0.6], [0, 1, 0], [1, 0, 0]),        # ceiling, normal -z
          quad([1.2, 0, 0.3], [0, 0, 0.3], [0, 1, 0])]    # mirror wall, normal -x
    tris = []
    for q in range(3):
        b = 4 * q
        tris += [[b, b + 1, b + 2], [b, b + 2, b + 3]]
    grid = []
    for z in (0.0, 0.6):
        for x in (-0.5, 0.0, 0.5):
            for y in (-0.4, 0.3):
                grid.append([x, y, z])
    v = np.concatenate(vs + [np.array(grid, np.float32)])
    mats = [[0.8, 0.6, 0.4, 0.0, 0.0, 0.0, 0.9, 2], [0.6, 0.6, 0.6, 0.04, 0.04, 0.04, 0.5, 0],
            [0.9, 0.9, 0.9, 0.9, 0.9, 0.9, 0.0, 3]]
    mids = np.array([0, 0, 1, 1, 2, 2], np.int32)
    return v, np.array(tris, np.int32), mids, mats, np.arange(12, len(v))


def smooth_basis(nc: int, n: int = 8, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.lognormal(0.0, 0.6, (nc, 6, n, n, 3)).astype(np.float32)
