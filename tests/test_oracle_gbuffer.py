"""CPU checks that pin the oracle's G-buffer producer (orc_gbuffer_cast)
without retyping it: pinhole closed forms on planes, a 45-degree mirror, an
independent numpy intersection (plane hit + edge functions) on random
triangle soups, cube-face lookups at texel centres, translation invariance
and sphere normals.

The oracle restates the first-hit part of the reference's OptiX pass:
__raygen__fovCamera (Renderer/device_functions/camera.cu:82-290),
__closesthit__radiance / __miss__radiance
(Renderer/device_functions/devicePrograms.cu:47-311, 517-564) and
EnvDirecToUV (Renderer/device_functions/cubeMap.cu:21-92)."""
import math

import numpy as np
import pytest

from oracle import binding as orc

FOV = 0.8575560548920328          # RenderEngine::setCamera
T = math.tan(0.5 * FOV)
LOOK_Z = [0, 0, 0, 0, 0, -1, 0, 1, 0]  # from origin along -z, up +y
MAT = [[0.5, 0.25, 0.125, 0.0625, 0.03125, 0.015625, 0.75, 0]]


def _pinhole_dirs(w, h):
    """Unnormalised ray directions of a camera at the origin looking along
    -z with +y up: pixel (ix, iy) maps to ((ix - w/2), (iy - h/2)) * 2 tan(fov/2) / h
    on the z = -1 plane (square pixels, jitter at its mean)."""
    iy, ix = np.mgrid[0:h, 0:w]
    s = 2.0 * T / h
    return np.stack([(ix - w / 2) * s, (iy - h / 2) * s, -np.ones((h, w))], -1).reshape(-1, 3)


def _quad(z, half=100.0):
    v = np.array([[-half, -half, z], [half, -half, z], [half, half, z], [-half, half, z]],
                 np.float32)
    return v, np.array([[0, 1, 2], [0, 2, 3]], np.int32)


def test_plane_closed_form():
    w, h, z = 16, 12, -4.0
    v, ix = _quad(z)
    bounds = [-10, -10, -8, 10, 10, 2]
    g = orc.gbuffer_cast(v, ix, MAT, LOOK_Z, w, h, bounds=bounds)
    d = _pinhole_dirs(w, h)
    hit = d * (z / d[:, 2:3])  # ray meets z = -4
    want = (hit + 0.001 - np.array(bounds[:3])) / (np.array(bounds[3:]) - np.array(bounds[:3]))
    assert np.all(g["env_flag"] == 1)
    assert np.allclose(g["position"][:, :3], want, rtol=0, atol=1e-12)
    # CCW quad seen from +z: geometric normal +z; reflection flips z.
    assert np.allclose(g["normal"][:, :3], [0.5, 0.5, 1.0], atol=1e-15)
    dn = d / np.linalg.norm(d, axis=1, keepdims=True)
    r = dn * [1, 1, -1]
    assert np.allclose(g["reflect"][:, :3], r / 2 + 0.5, atol=1e-12)
    assert np.allclose(g["kd"][:, :3], MAT[0][0:3]) and np.allclose(g["ks"][:, :3], MAT[0][3:6])
    assert np.allclose(g["roughness"], MAT[0][6])
    assert np.all(g["position"][:, 3] == 1) and np.all(g["kd"][:, 3] == 1)
    # The centre pixel's ray is the view direction itself.
    c = (h // 2) * w + w // 2
    assert np.allclose(g["position"][c, :3], (np.array([0, 0, z]) + 0.001 - bounds[:3]) /
                       (np.array(bounds[3:]) - bounds[:3]), atol=1e-15)


def test_miss_encodings():
    w, h = 10, 8
    v, ix = _quad(+5.0)  # behind the camera
    bounds = [-1, -2, -3, 1, 2, 3]
    g = orc.gbuffer_cast(v, ix, MAT, LOOK_Z, w, h, bounds=bounds)
    d = _pinhole_dirs(w, h)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    assert np.all(g["env_flag"] == 0) and np.all(g["prim"] == -1)
    lo, hi = np.array(bounds[:3]), np.array(bounds[3:])
    assert np.allclose(g["position"][:, :3], (-1e5 - lo) / (hi - lo), atol=1e-9)
    assert np.allclose(g["normal"][:, :3], -d / 2 + 0.5, atol=1e-12)
    assert np.allclose(g["reflect"][:, :3], -d / 2 + 0.5, atol=1e-12)
    assert np.all(g["kd"][:, :3] == 0) and np.all(g["ks"][:, :3] == 0)
    assert np.all(g["roughness"] == 1.0)


def _centre_miss_kd(faces, direction):
    """kd of pixel (2, 2) of a 4 x 4 view along `direction` (the pixel whose
    ray is the view direction: ix - cx = 0) with no geometry in view (a tiny
    triangle far behind)."""
    d = np.asarray(direction, np.float64)
    up = [0, 1, 0] if abs(d[1]) < 0.9 * np.linalg.norm(d) else [1, 0, 0]
    v = (-d * 1e3 + np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]])).astype(np.float32)
    cam = [0, 0, 0, *d, *up]
    g = orc.gbuffer_cast(v, [[0, 1, 2]], MAT, cam, 4, 4, bounds=[-1e4] * 3 + [1e4] * 3,
                         env_faces=faces)
    assert g["env_flag"][10] == 0
    return g["kd"][10, :3]


def test_env_face_per_axis():
    # EnvDirecToUV: +x 0, -x 1, +y 2, -y 3, -z 4, +z 5.
    n = 4
    faces = np.zeros((6, n, n, 3), np.float32)
    for f in range(6):
        faces[f] = f + 1
    for f, d in enumerate([(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, -1), (0, 0, 1)]):
        assert np.allclose(_centre_miss_kd(faces, d), f + 1, atol=1e-12)


@pytest.mark.parametrize("face,i,j", [(4, 0, 0), (4, 5, 2), (0, 7, 3), (5, 1, 6), (2, 3, 3),
                                      (3, 6, 0), (1, 2, 7)])
def test_env_texel_centre_exact(face, i, j):
    # A direction through the centre of texel (row i, column j) reads that
    # texel with zero bilinear weight on its neighbours.
    n = 8
    rng = np.random.default_rng(face * 100 + i * 10 + j)
    faces = rng.uniform(0.1, 5.0, (6, n, n, 3)).astype(np.float32)
    uc, vc = 2 * (j + 0.5) / n - 1, 2 * (i + 0.5) / n - 1
    d = {0: (1, vc, uc), 1: (-1, vc, -uc), 2: (uc, 1, vc), 3: (uc, -1, -vc), 4: (uc, vc, -1),
         5: (-uc, vc, 1)}[face]
    assert np.allclose(_centre_miss_kd(faces, d), faces[face, i, j], rtol=1e-12)


def test_mirror_45_degrees():
    # A mirror at z = -4 tilted 45 degrees about y sends the centre ray
    # (0, 0, -1) to +x; a target wall at x = 3 takes the first hit.
    s = math.sqrt(0.5)
    n = np.array([s, 0, s])
    e1 = np.array([s, 0, -s])  # in the mirror plane
    e2 = np.array([0, 1, 0])
    c0 = np.array([0, 0, -4.0])
    mirror = [c0 - 5 * e1 - 5 * e2, c0 + 5 * e1 - 5 * e2, c0 + 5 * e2]
    wall = [[3, -10, -20], [3, 10, -20], [3, 0, 10]]
    v = np.array(mirror + wall, np.float32)
    ix = np.array([[0, 1, 2], [3, 4, 5]], np.int32)
    mats = [[0, 0, 0, 0, 0, 0, 0, 3], [0.9, 0.8, 0.7, 0.2, 0.3, 0.4, 0.35, 0]]
    bounds = [-10, -10, -10, 10, 10, 10]
    w, h = 10, 8
    c = (h // 2) * w + w // 2  # the pixel whose ray is the view direction
    lo, hi = np.array(bounds[:3]), np.array(bounds[3:])
    g2 = orc.gbuffer_cast(v, ix, mats, LOOK_Z, w, h, material_ids=[0, 1], max_bounce=2,
                          bounds=bounds)
    # Mirror hit p1 = (0, 0, -4) + 0.001; the reflected ray along +x meets
    # x = 3 at y = 0.001, z = -3.999, plus the offset again.
    want = np.array([3.001, 0.002, -3.998])
    assert g2["env_flag"][c] == 1 and g2["prim"][c] == 1
    assert np.allclose(g2["position"][c, :3], (want - lo) / (hi - lo), atol=1e-7)
    assert np.allclose(g2["kd"][c, :3], mats[1][0:3]) and g2["roughness"][c] == pytest.approx(0.35)
    # Wall normal: cross((0, 20, 0), (0, 10, 30)) = (600, 0, 0) -> +x.
    assert np.allclose(g2["normal"][c, :3], [1.0, 0.5, 0.5], atol=1e-12)
    # The reflected ray (+x) bounces back to -x off the wall.
    assert np.allclose(g2["reflect"][c, :3], [0.0, 0.5, 0.5], atol=1e-12)
    # max_bounce 1 (wavelet mode): the path ends on the mirror, no first hit.
    g1 = orc.gbuffer_cast(v, ix, mats, LOOK_Z, w, h, material_ids=[0, 1], max_bounce=1,
                          bounds=bounds)
    assert g1["env_flag"][c] == 0 and g1["prim"][c] == -1
    assert np.allclose(g1["position"][c, :3], -lo / (hi - lo))
    assert np.all(g1["normal"][c, :3] == 0.5) and np.all(g1["kd"][c, :3] == 0)
    assert g1["roughness"][c] == 0.0
    # Mirror then miss: kd is 0 even with an environment (reference MIS weight 0).
    faces = np.ones((6, 2, 2, 3), np.float32)
    gm = orc.gbuffer_cast(np.array(mirror + [[50, 50, 50], [51, 50, 50], [50, 51, 50]],
                                   np.float32), ix, mats, LOOK_Z, w, h, material_ids=[0, 1],
                          max_bounce=3, bounds=bounds, env_faces=faces)
    assert gm["env_flag"][c] == 0 and np.all(gm["kd"][c, :3] == 0)
    assert np.allclose(gm["normal"][c, :3], [0.0, 0.5, 0.5], atol=1e-12)  # -(+x) / 2 + 0.5


def _numpy_closest(v, ix, d):
    """Closest hit per ray from the origin, t > 0.01, by plane intersection
    and edge-function containment (a different formulation from the
    oracle's Moller-Trumbore).  Returns (prim, t), prim -1 for none."""
    A, B, C = v[ix[:, 0]].astype(np.float64), v[ix[:, 1]].astype(np.float64), \
        v[ix[:, 2]].astype(np.float64)
    n = np.cross(B - A, C - A)                      # [T, 3]
    dn = d @ n.T                                    # [R, T]
    with np.errstate(divide="ignore", invalid="ignore"):
        t = np.einsum("tk,tk->t", n, A)[None, :] / dn
    q = t[..., None] * d[:, None, :]                # [R, T, 3]
    inside = np.ones(t.shape, bool)
    for P0, P1 in ((A, B), (B, C), (C, A)):
        e = np.cross(P1 - P0, q - P0[None])          # [R, T, 3]
        inside &= np.einsum("rtk,tk->rt", e, n) >= 0
    ok = inside & (t > 0.01) & np.isfinite(t)
    tt = np.where(ok, t, np.inf)
    prim = np.argmin(tt, axis=1)
    best = tt[np.arange(len(d)), prim]
    return np.where(np.isfinite(best), prim, -1), best


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_soup_matches_numpy(seed):
    rng = np.random.default_rng(seed)
    nt, w, h = 60, 24, 18
    centres = rng.uniform([-3, -2, -9], [3, 2, -2], (nt, 3))
    v = (centres[:, None, :] + rng.normal(0, 0.8, (nt, 3, 3))).reshape(-1, 3).astype(np.float32)
    ix = np.arange(3 * nt, dtype=np.int32).reshape(nt, 3)
    g = orc.gbuffer_cast(v, ix, MAT, LOOK_Z, w, h, bounds=[-10, -10, -12, 10, 10, 2])
    d = _pinhole_dirs(w, h)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    prim, t = _numpy_closest(v, ix, d)
    safe = g["margin"] > 1e-6
    assert safe.mean() > 0.9
    assert np.array_equal(g["prim"][safe], prim[safe])
    hit = safe & (prim >= 0)
    assert hit.sum() > 50
    p = d[hit] * t[hit, None] + 0.001
    lo, hi = np.array([-10, -10, -12]), np.array([10, 10, 2])
    assert np.allclose(g["position"][hit, :3], (p - lo) / (hi - lo), atol=1e-10)


def test_translation_invariance():
    rng = np.random.default_rng(7)
    nt = 40
    v = (rng.uniform([-3, -2, -9], [3, 2, -2], (nt, 1, 3)) +
         rng.normal(0, 0.7, (nt, 3, 3))).reshape(-1, 3)
    ix = np.arange(3 * nt, dtype=np.int32).reshape(nt, 3)
    bounds = np.array([-10, -10, -12, 10, 10, 2], np.float64)
    off = np.array([8.0, -4.0, 16.0])
    g0 = orc.gbuffer_cast(v.astype(np.float32), ix, MAT, LOOK_Z, 20, 16, bounds=bounds)
    cam = np.array(LOOK_Z, np.float64)
    cam[0:3] += off
    cam[3:6] += off
    g1 = orc.gbuffer_cast((v + off).astype(np.float32), ix, MAT, cam, 20, 16,
                          bounds=bounds + np.concatenate([off, off]))
    safe = (g0["margin"] > 1e-5) & (g1["margin"] > 1e-5)
    assert np.array_equal(g0["prim"][safe], g1["prim"][safe])
    hit = safe & (g0["env_flag"] == 1)
    assert hit.sum() > 20
    # Hits move with the scene; misses sit at the fixed -1e5 and do not.
    assert np.allclose(g0["position"][hit], g1["position"][hit], atol=2e-5)
    for k in ("normal", "reflect"):
        assert np.allclose(g0[k][safe], g1[k][safe], atol=2e-5), k


def _uv_sphere(r, nu, nv, centre):
    th = np.linspace(0, np.pi, nv + 1)
    ph = np.linspace(0, 2 * np.pi, nu, endpoint=False)
    T_, P_ = np.meshgrid(th, ph, indexing="ij")
    n = np.stack([np.sin(T_) * np.cos(P_), np.cos(T_), np.sin(T_) * np.sin(P_)], -1).reshape(-1, 3)
    idx = []
    for i in range(nv):
        for j in range(nu):
            a, b = i * nu + j, i * nu + (j + 1) % nu
            c, d = a + nu, b + nu
            idx += [[a, c, b], [b, c, d]]
    return (centre + r * n).astype(np.float32), n.astype(np.float32), np.array(idx, np.int32)


def test_sphere_vertex_normals_are_radial():
    centre = np.array([0.0, 0.0, -6.0])
    v, n, ix = _uv_sphere(2.0, 48, 24, centre)
    g = orc.gbuffer_cast(v, ix, MAT, LOOK_Z, 32, 24, normals=n, bounds=[-3, -3, -9, 3, 3, -3])
    hit = (g["env_flag"] == 1) & (g["margin"] > 1e-6)
    assert hit.sum() > 100
    lo, hi = np.array([-3, -3, -9]), np.array([3, 3, -3])
    p = g["position"][hit, :3] * (hi - lo) + lo - 0.001
    radial = (p - centre) / np.linalg.norm(p - centre, axis=1, keepdims=True)
    nn = g["normal"][hit, :3] * 2 - 1
    assert np.allclose(np.linalg.norm(nn, axis=1), 1, atol=1e-12)
    assert np.abs(nn - radial).max() < 0.02
    # Without vertex normals the geometric normal is perpendicular to its triangle.
    g2 = orc.gbuffer_cast(v, ix, MAT, LOOK_Z, 32, 24, bounds=[-3, -3, -9, 3, 3, -3])
    h2 = np.flatnonzero((g2["env_flag"] == 1) & (g2["margin"] > 1e-6))
    tri = ix[g2["prim"][h2]]
    ng = g2["normal"][h2, :3] * 2 - 1
    for k in range(3):
        e = v[tri[:, (k + 1) % 3]].astype(np.float64) - v[tri[:, k]]
        assert np.abs(np.einsum("ij,ij->i", ng, e)).max() < 1e-9
