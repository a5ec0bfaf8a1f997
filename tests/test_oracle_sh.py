"""CPU checks that pin the oracle's SH-PRT pieces: pbrt's real SH basis
(Renderer/device_functions/shaderUtils.h:230-337) against closed forms and
orthonormality, and the render (RenderEngine::shRender,
Renderer/RenderEngine.cpp:459-486; devicePrograms.cu:312-501) against the
expectation it must converge to for a single-entry transfer matrix."""
import math

import numpy as np
import pytest

from oracle import binding as orc


def _grid_dirs(m):
    g = (np.arange(m) + 0.5) / m
    U, V = np.meshgrid(g, g)
    a, b = 2 * U - 1, 2 * V - 1
    J = 4.0 / (1 + a * a + b * b) ** 1.5 / (m * m)
    one = np.ones_like(a)
    dirs = []
    for d in ((one, a, b), (-one, a, b), (a, one, b), (a, -one, b), (a, b, one), (a, b, -one)):
        x = np.stack(d, -1).reshape(-1, 3)
        dirs.append(x / np.linalg.norm(x, axis=1, keepdims=True))
    return np.concatenate(dirs), np.tile(J.ravel(), 6)


def test_sh_closed_forms_bands_0_to_2():
    rng = np.random.default_rng(0)
    c0, c1 = 0.5 / math.sqrt(math.pi), math.sqrt(3 / (4 * math.pi))
    h = 0.5 * math.sqrt(15 / math.pi)
    for _ in range(10):
        w = rng.normal(size=3)
        w /= np.linalg.norm(w)
        x, y, z = w
        want = [c0, -c1 * y, c1 * z, -c1 * x,           # Condon-Shortley phase on odd m
                h * x * y, -h * y * z, 0.25 * math.sqrt(5 / math.pi) * (3 * z * z - 1),
                -h * x * z, 0.25 * math.sqrt(15 / math.pi) * (x * x - y * y)]
        assert np.allclose(orc.sh_eval(w)[:9], want, atol=1e-12)
    # Poles: sin(phi), cos(phi) taken as 0, 1.
    assert np.allclose(orc.sh_eval([0, 0, 1])[[1, 3]], 0.0)


def test_sh_orthonormal():
    dirs, wts = _grid_dirs(48)
    Y = np.array([orc.sh_eval(d) for d in dirs])
    G = (Y * wts[:, None]).T @ Y
    assert np.abs(G - np.eye(81)).max() < 2e-3


def test_sh_render_single_entry_transfer():
    """T_v[k][j] = 1 only at (k, j) = (0, 3) and L[3] = c: every vertex's
    light response is c in coefficient 0, so a vertex's colour is
    c * Y_00 * integral(f cos) (cosine sampling of a diffuse BSDF is exact).
    A transposed T would give 0."""
    s = 0.01
    v = np.array([[-s, 0, -s], [s, 0, -s], [s, 0, s], [-s, 0, s]], np.float32)
    ix = np.array([[0, 2, 1], [0, 3, 2]], np.int32)
    n = np.tile([0.0, 1.0, 0.0], (4, 1)).astype(np.float32)
    mat = [0.7, 0.5, 0.3, 0.0, 0.0, 0.0, 0.4, 2]
    deg, nc = 2, 9
    T = np.zeros((4, nc, nc, 3), np.float32)
    T[:, 0, 3, :] = 1.0
    L = np.zeros((nc, 3), np.float32)
    L[3] = [2.0, 1.0, 0.5]
    cam = [0.0, 3.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, -1.0]
    g = orc.sh_render(v, ix, n, [mat], cam, 2, 2, T, L, deg, samples=40000, seed=3)
    assert g["prim"][3] >= 0
    view = np.array([0.0, 1.0, 0.0])
    dirs, wts = _grid_dirs(48)
    up = np.array([0.0, 1.0, 0.0])
    albedo = sum(orc.ct_eval(mat, up, up, d, view) * w for d, w in zip(dirs, wts) if d[1] > 0)
    want = np.array([2.0, 1.0, 0.5]) * 0.5 / math.sqrt(math.pi) * albedo
    assert np.allclose(g["color"][3, :3], want, rtol=0.02), (g["color"][3, :3], want)
    T2 = np.ascontiguousarray(np.swapaxes(T, 1, 2))  # transposed: R_3 = L[0] = 0
    g2 = orc.sh_render(v, ix, n, [mat], cam, 2, 2, T2, L, deg, samples=64, seed=3)
    assert np.all(g2["color"][3, :3] == 0)
