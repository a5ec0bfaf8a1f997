"""CPU checks that pin the oracle's guided a-trous denoiser (orc_denoise)
without retyping it: closed forms on tiny images, an independent scipy
normalised convolution for unbounded sigmas, invariants of the filter.

The denoiser replaces the reference's closed learned model behind
OptixBackend::denoise (Renderer/OptixBackend.cpp:841-906); its images are
not reproducible, so these checks pin the published filter the build
implements (edge-avoiding a-trous, Dammertz et al. 2010)."""
import math

import numpy as np
import pytest
from scipy import ndimage

from oracle import binding as orc

H3 = np.array([1, 4, 6, 4, 1], np.float64) / 16


def _img(rng, h, w, scale=1.0):
    c = np.empty((h, w, 4), np.float32)
    c[..., :3] = rng.lognormal(0.0, 1.0, (h, w, 3)) * scale
    c[..., 3] = rng.uniform(0, 1, (h, w))
    return c


def _unit(rng, h, w):
    v = rng.normal(size=(h, w, 3))
    v /= np.linalg.norm(v, axis=-1, keepdims=True)
    out = np.zeros((h, w, 4), np.float32)
    out[..., :3] = v
    return out


def _gray(vals):
    c = np.zeros((1, len(vals), 4), np.float32)
    c[0, :, :3] = np.asarray(vals, np.float32)[:, None]
    c[0, :, 3] = 1.0
    return c


def test_two_pixel_closed_form():
    # 1 x 2 image, constant guides, grey colours a and b: one pass mixes
    # the pair with tap weights 6/16 (centre) and 4/16 (neighbour) x
    # exp(-kappa^2 3 (b - a)^2 / sc^2), kappa = 1 / sqrt(a b); a second pass
    # (step 2) has no in-image taps and changes nothing.
    a, b, sc = 0.5, 2.0, 1.3
    c = _gray([a, b])
    g = np.zeros_like(c)
    kappa = 2.0 ** (-(math.log2(a) + math.log2(b)) / 2)
    e = math.exp(-kappa ** 2 * 3 * (b - a) ** 2 / sc ** 2)
    want0 = (6 * a + 4 * e * b) / (6 + 4 * e)
    want1 = (6 * b + 4 * e * a) / (6 + 4 * e)
    for it in (1, 2, 3):
        out = orc.denoise(c, g, g, iterations=it, sigma_color=sc)
        assert np.allclose(out[0, 0, :3], want0, rtol=1e-12)
        assert np.allclose(out[0, 1, :3], want1, rtol=1e-12)
        assert np.all(out[0, :, 3] == 1.0)


def test_three_pixel_closed_form_second_pass_scale():
    # 1 x 3 image [a, a, b]: pass 0 (step 1) and pass 1 (step 2, colour
    # sigma halved: distance weight 4^1) by hand.
    a, b, sc = 1.0, 3.0, 2.0
    c = _gray([a, a, b])
    g = np.zeros_like(c)
    kappa = 2.0 ** (-(2 * math.log2(a) + math.log2(b)) / 3)
    k2 = kappa ** 2

    def e(d, i):
        return math.exp(-k2 * 4 ** i * 3 * d * d / sc ** 2)

    e1 = e(b - a, 0)
    p0 = (6 * a + 4 * a + 1 * e1 * b) / (6 + 4 + e1)
    p1 = (4 * a + 6 * a + 4 * e1 * b) / (4 + 6 + 4 * e1)
    p2 = (1 * e1 * a + 4 * e1 * a + 6 * b) / (e1 + 4 * e1 + 6)
    out1 = orc.denoise(c, g, g, iterations=1, sigma_color=sc)
    assert np.allclose(out1[0, :, 0], [p0, p1, p2], rtol=1e-12)
    # Pass 1: step 2, pixel 0 <-> pixel 2; pixel 1 alone.
    e2 = e(p2 - p0, 1)
    q0 = (6 * p0 + 4 * e2 * p2) / (6 + 4 * e2)
    q2 = (6 * p2 + 4 * e2 * p0) / (6 + 4 * e2)
    out2 = orc.denoise(c, g, g, iterations=2, sigma_color=sc)
    assert np.allclose(out2[0, :, 0], [q0, p1, q2], rtol=1e-12)


def _atrous_linear(img, iterations):
    """Normalised B3-spline a-trous smoothing with taps outside the image
    skipped: per pass, correlate(img) / correlate(ones) with the dilated
    5 x 5 kernel and zero padding."""
    cur = img.astype(np.float64)
    ones = np.ones(img.shape[:2])
    for i in range(iterations):
        s = 1 << i
        k1 = np.zeros(4 * s + 1)
        k1[::s] = H3
        k2 = np.outer(k1, k1)
        den = ndimage.correlate(ones, k2, mode="constant", cval=0.0)
        cur = np.stack([ndimage.correlate(cur[..., ch], k2, mode="constant", cval=0.0) / den
                        for ch in range(cur.shape[-1])], axis=-1)
    return cur


@pytest.mark.parametrize("h,w,it", [(17, 23, 3), (9, 40, 5), (1, 12, 2), (33, 1, 4)])
def test_unbounded_sigmas_equal_linear_atrous(h, w, it):
    rng = np.random.default_rng(h * 100 + w)
    c = _img(rng, h, w)
    n = _unit(rng, h, w)
    a = _img(rng, h, w)
    out = orc.denoise(c, n, a, iterations=it, sigma_color=1e30, sigma_normal=1e30,
                      sigma_albedo=1e30)
    want = _atrous_linear(c[..., :3], it)
    assert np.allclose(out[..., :3], want, rtol=1e-12, atol=0)
    assert np.array_equal(out[..., 3], c[..., 3].astype(np.float64))


def test_constant_image_and_blend_one_unchanged():
    rng = np.random.default_rng(3)
    h, w = 13, 19
    c = np.empty((h, w, 4), np.float32)
    c[...] = [3.0, 2.0, 0.25, 0.5]
    n, a = _unit(rng, h, w), _img(rng, h, w)
    out = orc.denoise(c, n, a)
    assert np.allclose(out, c, rtol=1e-13, atol=0)
    c2 = _img(rng, h, w)
    out = orc.denoise(c2, n, a, blend=1.0)
    assert np.array_equal(out, c2.astype(np.float64))
    out = orc.denoise(c2, n, a, blend=0.25)
    full = orc.denoise(c2, n, a, blend=0.0)
    assert np.allclose(out[..., :3], 0.25 * c2[..., :3] + 0.75 * full[..., :3], rtol=1e-13)


def test_distinct_normals_with_tiny_sigma_keep_input():
    rng = np.random.default_rng(4)
    h, w = 11, 14
    c, a = _img(rng, h, w), np.zeros((h, w, 4), np.float32)
    n = _unit(rng, h, w)
    out = orc.denoise(c, n, a, sigma_normal=1e-6)
    assert np.allclose(out[..., :3], c[..., :3], rtol=1e-14, atol=0)


def test_albedo_edge_stops_flow():
    # Left half albedo 0, right half 1, sigma_albedo tiny: the left half's
    # output does not depend on the right half's colours at all.
    rng = np.random.default_rng(5)
    h, w = 16, 24
    c, n = _img(rng, h, w), np.zeros((h, w, 4), np.float32)
    a = np.zeros((h, w, 4), np.float32)
    a[:, w // 2:, :3] = 1.0
    c2 = c.copy()
    c2[:, w // 2:, :3] *= 7.0
    # Unbounded colour sigma: the HDR normalisation (which sees the right
    # half) does not enter, so only the albedo guide gates the flow.
    o1 = orc.denoise(c, n, a, sigma_color=1e30, sigma_albedo=0.01)
    o2 = orc.denoise(c2, n, a, sigma_color=1e30, sigma_albedo=0.01)
    assert np.array_equal(o1[:, : w // 2], o2[:, : w // 2])
    assert not np.array_equal(o1[:, w // 2:], o2[:, w // 2:])


def test_power_of_two_scaling_and_transpose():
    rng = np.random.default_rng(6)
    h, w = 12, 17
    c, n, a = _img(rng, h, w), _unit(rng, h, w), _img(rng, h, w)
    out = orc.denoise(c, n, a)
    c8 = c.copy()
    c8[..., :3] *= 8.0
    out8 = orc.denoise(c8, n, a)
    assert np.allclose(out8[..., :3], 8.0 * out[..., :3], rtol=1e-11)
    t = lambda x: np.ascontiguousarray(np.transpose(x, (1, 0, 2)))
    outt = orc.denoise(t(c), t(n), t(a))
    assert np.allclose(outt, t(out), rtol=1e-12)
