"""Pins the oracle's hash-grid encoding (and the product's host level table)
against the committed level table, an FP64 evaluation of the level formula,
brute-force trilinear interpolation and hand-computed spatial hashes."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden" / "hashgrid_levels.json"


def test_level_table_matches_fixture_and_oracle(nw, orc):
    fx = json.loads(GOLDEN.read_text())
    g = fx["grid"]
    prod = nw.hashgrid_levels(nw.HashgridDesc(g["n_levels"], g["n_features_per_level"],
                                              g["log2_hashmap_size"], g["base_resolution"],
                                              g["per_level_scale"]))
    ora = orc.hashgrid_levels(g["n_levels"], g["log2_hashmap_size"], g["base_resolution"],
                              g["per_level_scale"])
    scale = np.array([float.fromhex(h) for h in fx["scale_f32_hex"]], np.float32)
    for tab in (prod, ora):
        np.testing.assert_array_equal(tab["scale"], scale)
        for key in ("resolution", "dense", "size", "offset"):
            np.testing.assert_array_equal(tab[key], np.array(fx[key], np.uint32), err_msg=key)
    assert fx["offset"][-1] == 6513496


def test_level_table_against_fp64_formula(nw):
    """scale_l = b^l N_min - 1 to FP32 accuracy; res = ceil(scale) + 1 except
    where FP32 rounding lands exactly on an integer (documented)."""
    lv = nw.hashgrid_levels(nw.HashgridDesc(16, 2, 19, 16, 1.5))
    T = 1 << 19
    for l in range(16):
        exact = 1.5 ** l * 16 - 1
        assert abs(float(lv["scale"][l]) - exact) <= 1e-6 * exact  # exp2f/log2f rounding
        r = int(lv["resolution"][l])
        assert r == math.ceil(float(lv["scale"][l])) + 1
        assert r in (math.ceil(exact) + 1, math.ceil(exact))  # FP32 may round onto the integer
        dense = r ** 3 <= T
        assert bool(lv["dense"][l]) == dense
        assert lv["size"][l] == (min(-(-r ** 3 // 8) * 8, T) if dense else T)


def _scene(orc, levels, log2T, table, F=2):
    lv = orc.hashgrid_levels(levels, log2T, 16, 1.5)
    wf = np.zeros((levels * F, 6 * 4 * 4), np.float32)
    layers = np.zeros(64 * 32 + 2 * 64 * 64 + 3 * 64, np.float32)
    return orc.Scene(lv, F, table, wf, None, layers), lv


def test_dense_level_is_trilinear_interpolation(orc):
    """With one dense level (16^3 <= 2^13), the encoding is the trilinear
    interpolation of the grid values at pos = x * 15 + 0.5."""
    rng = np.random.default_rng(0)
    sc, lv = _scene(orc, 1, 13, rng.standard_normal((16 ** 3 // 8 * 8 + 8, 2)).astype(np.float32))
    assert lv["dense"][0] == 1 and lv["resolution"][0] == 16
    table = sc.keep["table"]
    for _ in range(200):
        x = rng.random(3).astype(np.float32)
        _, feat = sc.hash_encode(x)
        # FP32 grid position fmaf(x, 15, 0.5): the FP64 product is exact.
        p = (x.astype(np.float64) * 15.0 + 0.5).astype(np.float32).astype(np.float64)
        g = np.floor(p).astype(int)
        f = p - g
        ref = np.zeros(2)
        for c in range(8):
            b = [(c >> d) & 1 for d in range(3)]
            w = np.prod([f[d] if b[d] else 1 - f[d] for d in range(3)])
            idx = ((g[0] + b[0]) + (g[1] + b[1]) * 16 + (g[2] + b[2]) * 256) % lv["size"][0]
            ref += w * table[idx]
        np.testing.assert_allclose(feat, ref, rtol=1e-12, atol=1e-14)


def test_vertex_positions_are_one_hot(orc):
    """A position on a grid vertex (frac = 0 in every dimension) returns that
    vertex's entry exactly."""
    rng = np.random.default_rng(1)
    # A one-level table with scale 16: x = (n - 0.5) / 16 is exact in FP32 and
    # lands on vertex n.
    lv = {"scale": np.array([16.0], np.float32), "resolution": np.array([17], np.uint32),
          "dense": np.array([1], np.uint32), "size": np.array([4920], np.uint32),
          "offset": np.array([0, 4920], np.uint32)}
    table = rng.standard_normal((4920, 2)).astype(np.float32)
    sc = orc.Scene(lv, 2, table, np.zeros((2, 96), np.float32), None,
                   np.zeros(64 * 32 + 2 * 64 * 64 + 3 * 64, np.float32))
    for (i, j, k) in [(1, 1, 1), (3, 7, 11), (16, 1, 9), (8, 8, 8)]:
        x = np.array([(i - 0.5) / 16, (j - 0.5) / 16, (k - 0.5) / 16], np.float32)
        idx, feat = sc.hash_encode(x)
        np.testing.assert_array_equal(feat, table[i + j * 17 + k * 289].astype(np.float64))
        assert idx[0, 0] == i + j * 17 + k * 289


def test_hashed_level_indices_match_published_hash(orc):
    """Hashed level: entry = (x ^ y * 2654435761 ^ z * 805459861) mod 2^32 mod T."""
    rng = np.random.default_rng(2)
    T = 1 << 10
    sc, lv = _scene(orc, 8, 10, rng.standard_normal((8 * T, 2)).astype(np.float32))
    assert lv["dense"][5] == 0
    for _ in range(50):
        x = rng.random(3).astype(np.float32)
        idx, _ = sc.hash_encode(x)
        for l in range(8):
            s = np.float32(lv["scale"][l])
            g = [int(np.floor(np.float32(np.float64(v) * np.float64(s) + 0.5))) for v in x]
            for c in range(8):
                cx, cy, cz = (g[d] + ((c >> d) & 1) for d in range(3))
                if lv["dense"][l]:
                    r = int(lv["resolution"][l])
                    want = (cx + cy * r + cz * r * r) % int(lv["size"][l])
                else:
                    h = (cx ^ ((cy * 2654435761) & 0xFFFFFFFF) ^ ((cz * 805459861) & 0xFFFFFFFF))
                    want = (h & 0xFFFFFFFF) % int(lv["size"][l])
                assert idx[l, c] == want, (l, c)


def _corner_map(orc, sc, lv, level, positions):
    """(grid coordinate of corner c -> entry index) seen at `positions`,
    corner c taking offset (c >> d) & 1 along dimension d."""
    seen = {}
    scale = np.float32(lv["scale"][level])
    for x in positions:
        idx, _ = sc.hash_encode(x)
        p = (x.astype(np.float64) * float(scale) + 0.5).astype(np.float32)
        g = np.floor(p).astype(int)
        for c in range(8):
            key = tuple(int(g[d] + ((c >> d) & 1)) for d in range(3))
            v = int(idx[level, c])
            assert seen.setdefault(key, v) == v, (level, key)  # a function of the corner
    return seen


def test_corner_indices_lie_in_their_level_and_dense_levels_are_injective(orc):
    """Properties the level layout fixes whatever the hash: every corner
    index lies in [0, size) of its level, and a dense level (res^3 <= T)
    gives distinct grid vertices of [0, res)^3 distinct entries."""
    levels, log2T = 8, 15
    rng = np.random.default_rng(3)
    sc, lv = _scene(orc, levels, log2T, np.zeros((1 << 15) * levels, np.float32).reshape(-1, 2))
    pos = rng.random((3000, 3)).astype(np.float32)
    pos[:8] = [[0, 0, 0], [1, 1, 1], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [0.5, 0.5, 0.5],
               [1e-7, 1 - 1e-7, 0.25]]
    for x in pos[:400]:
        idx, _ = sc.hash_encode(x)
        assert np.all(idx < lv["size"][:, None]), x
    n_dense = 0
    for level in range(levels):
        if not lv["dense"][level]:
            continue
        n_dense += 1
        res = int(lv["resolution"][level])
        seen = _corner_map(orc, sc, lv, level, pos)
        inside = {k: v for k, v in seen.items() if all(0 <= c < res for c in k)}
        assert len(inside) > min(2000, res ** 3 // 2)
        assert len(set(inside.values())) == len(inside), level  # injective
        assert max(inside.values()) < lv["size"][level]
    assert 0 < n_dense < levels  # both kinds of level are exercised
