"""Pins the oracle's yaw rotation of a cubemap (orc_rotate_cubemap) with
closed forms of the face conventions of Renderer/device_functions/cubeMap.cu:
  - yaw 0 and four quarter turns are the identity;
  - a quarter turn of a face-constant map permutes faces as the axes say:
    output +X looks at source +Z, -X at -Z, +Z at -X, -Z at +X, +-Y stay;
  - two quarter turns equal a half turn (nearest sampling of texel centres);
  - the +Y face turns in place: out[r][c] = src[c][N - 1 - r].
"""
import math

import numpy as np
import pytest


def _maps(n, seed=0):
    return np.random.default_rng(seed).lognormal(0, 1, (6, n, n, 3)).astype(np.float32)


@pytest.mark.parametrize("n", [2, 8, 32])
def test_identity_and_full_turn(orc, n):
    m = _maps(n, n)
    assert np.array_equal(orc.rotate_cubemap(m, 0.0), m)
    r = m
    for _ in range(4):
        r = orc.rotate_cubemap(r, math.pi / 2)
    assert np.array_equal(r, m)


def test_quarter_turn_face_permutation(orc):
    n = 16
    m = np.zeros((6, n, n, 3), np.float32)
    for f in range(6):
        m[f] = f + 1
    r = orc.rotate_cubemap(m, math.pi / 2)
    src_face = [5, 4, 2, 3, 0, 1]   # +X <- +Z, -X <- -Z, +Z <- -X, -Z <- +X
    for f in range(6):
        assert np.all(r[f] == src_face[f] + 1), f


def test_two_quarters_make_a_half(orc):
    m = _maps(16, 3)
    q = orc.rotate_cubemap(orc.rotate_cubemap(m, math.pi / 2), math.pi / 2)
    assert np.array_equal(q, orc.rotate_cubemap(m, math.pi))


def test_top_face_turns_in_place(orc):
    n = 8
    m = _maps(n, 5)
    r = orc.rotate_cubemap(m, math.pi / 2)
    for row in range(n):
        for col in range(n):
            assert np.array_equal(r[2, row, col], m[2, col, n - 1 - row])


def test_small_yaw_moves_energy_continuously(orc):
    # A small turn keeps most texels (nearest sampling) and never invents values.
    m = _maps(32, 7)
    r = orc.rotate_cubemap(m, 0.01)
    vals = set(map(tuple, m.reshape(-1, 3)))
    assert all(tuple(v) in vals for v in r.reshape(-1, 3))
    assert np.mean(np.all(r == m, axis=-1)) > 0.9
