"""Pins the oracle's non-standard Haar transform and area-weighted top-K
against closed forms, Parseval, an inverse round trip and a textbook numpy
recursion, and the selection against a numpy lexicographic sort."""
import numpy as np
import pytest


def haar_textbook(x):
    """Non-standard 2-D Haar, FP64, written from the textbook recursion: at
    each level, one 1-D orthonormal step over the rows of the LL block
    (averages left, details right), then one over its columns (averages on
    the low rows).  Row 0 is the first row."""
    a = np.array(x, np.float64)
    n = a.shape[0]
    s = n
    r2 = np.sqrt(2.0)
    while s >= 2:
        blk = a[:s, :s].copy()
        lo = (blk[:, 0::2] + blk[:, 1::2]) / r2
        hi = (blk[:, 0::2] - blk[:, 1::2]) / r2
        blk = np.concatenate([lo, hi], axis=1)
        lo = (blk[0::2, :] + blk[1::2, :]) / r2
        hi = (blk[0::2, :] - blk[1::2, :]) / r2
        a[:s, :s] = np.concatenate([lo, hi], axis=0)
        s //= 2
    return a


def haar_inverse(c):
    a = np.array(c, np.float64)
    n = a.shape[0]
    s = 2
    r2 = np.sqrt(2.0)
    while s <= n:
        blk = a[:s, :s].copy()
        h = s // 2
        lo, hi = blk[:h, :], blk[h:, :]
        rows = np.empty_like(blk)
        rows[0::2, :] = (lo + hi) / r2
        rows[1::2, :] = (lo - hi) / r2
        lo, hi = rows[:, :h], rows[:, h:]
        out = np.empty_like(blk)
        out[:, 0::2] = (lo + hi) / r2
        out[:, 1::2] = (lo - hi) / r2
        a[:s, :s] = out
        s *= 2
    return a


@pytest.mark.parametrize("n", [2, 4, 16, 64])
def test_haar_matches_textbook_and_inverts(orc, n):
    rng = np.random.default_rng(n)
    x = rng.lognormal(0, 1.5, (n, n)).astype(np.float32)
    got = orc.haar_face(x).astype(np.float64)
    ref = haar_textbook(x)
    np.testing.assert_allclose(got, ref, rtol=1e-6, atol=1e-6 * np.abs(ref).max())
    np.testing.assert_allclose(haar_inverse(got), x, rtol=1e-5, atol=1e-5 * x.max())
    # Parseval (orthonormal basis)
    assert abs(np.sum(got ** 2) / np.sum(x.astype(np.float64) ** 2) - 1) < 1e-6


def test_haar_constant_face_is_dc_only(orc):
    n, c = 32, 3.25
    got = orc.haar_face(np.full((n, n), c, np.float32))
    assert got[0, 0] == c * n
    got[0, 0] = 0
    assert np.all(got == 0)


@pytest.mark.parametrize("pos", [(0, 0), (5, 9), (31, 31), (16, 0)])
def test_haar_delta_has_3log2n_plus_1_terms(orc, pos):
    n = 32
    x = np.zeros((n, n), np.float32)
    x[pos] = 1.0
    got = orc.haar_face(x)
    assert np.count_nonzero(got) == 3 * 5 + 1
    assert got[0, 0] == pytest.approx(1.0 / n)


def test_envmap_haar_layout(orc):
    rng = np.random.default_rng(0)
    faces = rng.random((6, 8, 8, 3)).astype(np.float32)
    coef = orc.envmap_haar(faces)
    for f in range(6):
        for c in range(3):
            np.testing.assert_array_equal(coef[f, :, :, c], orc.haar_face(faces[f, :, :, c]))


def _scores_numpy(coef):
    n = coef.shape[1]
    r, g, b = (coef[..., i].astype(np.float32) for i in range(3))
    lum = (r * np.float32(0.212671) + g * np.float32(0.715160)) + b * np.float32(0.072169)
    rows, cols = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    m = np.maximum(rows, cols)
    j = np.where(m == 0, 0, np.floor(np.log2(np.maximum(m, 1))).astype(int))
    s = np.ldexp(np.abs(lum), -2 * j[None]).astype(np.float32)
    s[~(s >= np.finfo(np.float32).tiny)] = 0
    return s.reshape(-1)


@pytest.mark.parametrize("k", [1, 7, 64, 300, 6 * 16 * 16])
def test_topk_matches_lexicographic_sort(orc, k):
    n = 16
    rng = np.random.default_rng(k)
    faces = rng.lognormal(0, 1.5, (6, n, n, 3)).astype(np.float32)
    faces[1, 4:7, 4:7] = 1e3
    coef = orc.envmap_haar(faces)
    k = min(k, 6 * n * n)
    idx, st = orc.topk(coef, k)
    s = _scores_numpy(coef)
    for t in range(0, 6 * n * n, 97):
        assert orc.score(coef, t) == s[t]
    order = np.lexsort((np.arange(len(s)), -s.astype(np.float64)))
    np.testing.assert_array_equal(idx, order[:k])
    np.testing.assert_array_equal(st, coef.reshape(-1, 3)[order[:k]])


def test_topk_area_weights_and_ties(orc):
    n = 8
    coef = np.zeros((6, n, n, 3), np.float32)
    coef[0, 0, 0] = 1.0      # DC, area 1
    coef[0, 0, 1] = 1.0      # j = 0 detail, area 1 (tie with DC -> lower index first)
    coef[0, 2, 3] = 4.0      # j = 1 detail: 4 * 1/4 = 1 (tie)
    coef[0, 5, 0] = 15.9     # j = 2 detail: 15.9 / 16 < 1
    coef[3, 0, 0] = -1.0     # |lum| counts, sign kept in the strength
    idx, st = orc.topk(coef, 5)
    first = [0, 1, 2 * n + 3, 3 * n * n]  # four ties at score 1, by index
    np.testing.assert_array_equal(idx[:4], first)
    assert idx[4] == 5 * n
    np.testing.assert_array_equal(st[3], [-1, -1, -1])
