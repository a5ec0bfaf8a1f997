"""GPU tests of nwprt_denoise (the guided a-trous replacement for
OptixBackend::denoise, Renderer/OptixBackend.cpp:841-906) against the FP64
oracle (orc_denoise).

Tolerance: the passes run in FP32 with one hardware exp2 per tap, the HDR
normalisation as a compensated FP32 sum of hardware log2 values.  Per pixel and channel,
|gpu - oracle| <= 1e-4 * (|oracle| + 1e-3 * max|oracle|)."""
import numpy as np
import pytest

from neural_wavelet_prt_demo_amd import DeviceBuffer

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _inputs(rng, h, w):
    c = np.empty((h, w, 4), np.float32)
    c[..., :3] = rng.lognormal(0.0, 1.5, (h, w, 3))
    c[..., 3] = rng.uniform(0, 1, (h, w))
    # Normals: smooth field with a hard crease and some per-pixel noise.
    yy, xx = np.mgrid[0:h, 0:w]
    n = np.stack([np.sin(xx / 7.0), np.cos(yy / 5.0), np.where(xx > w // 2, 1.0, -1.0)], -1)
    n += rng.normal(0, 0.05, n.shape)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    nn = np.zeros((h, w, 4), np.float32)
    nn[..., :3] = n
    # Albedo: a few flat regions with noise.
    a = np.zeros((h, w, 4), np.float32)
    a[..., :3] = ((xx // 9 + yy // 7) % 3)[..., None] * 0.3 + rng.uniform(0, 0.02, (h, w, 3))
    return c, nn, a


def _run(nw, ctx, c, n, a, blend=0.0, **kw):
    h, w = c.shape[:2]
    bufs = [DeviceBuffer.from_array(x) for x in (c, n, a)]
    out = DeviceBuffer(c.nbytes)
    p = nw.Context.denoise_default_params()
    for k, v in kw.items():
        setattr(p, k, v)
    ctx.denoise(bufs[0], bufs[1], bufs[2], out, w, h, blend, p)
    ctx.sync()
    return out.download(np.float32, (h, w, 4))


def _check(got, want):
    want = want.astype(np.float64)
    scale = 1e-3 * np.abs(want[..., :3]).max()
    err = np.abs(got[..., :3] - want[..., :3]) / (np.abs(want[..., :3]) + scale)
    assert err.max() <= TOL, err.max()


@pytest.mark.parametrize("h,w,it,blend", [(29, 37, 5, 0.0), (48, 64, 1, 0.0), (48, 64, 8, 0.3),
                                          (1, 1, 5, 0.0), (1, 50, 3, 0.0), (50, 1, 4, 0.5),
                                          (17, 16, 2, 0.0)])
def test_denoise_matches_oracle(nw, orc, ctx, h, w, it, blend):
    rng = np.random.default_rng(h * 1000 + w + it)
    c, n, a = _inputs(rng, h, w)
    got = _run(nw, ctx, c, n, a, blend, iterations=it)
    want = orc.denoise(c, n, a, iterations=it, blend=blend)
    _check(got, want)
    assert np.array_equal(got[..., 3], c[..., 3])  # alpha copied


def test_denoise_custom_sigmas_and_sizes_in_a_row(nw, orc, ctx):
    # A larger image after a smaller one (work buffers grow), then smaller again.
    for (h, w, sc, sn, sa) in [(20, 30, 0.5, 0.2, 0.05), (130, 97, 2.0, 0.6, 0.3),
                               (9, 11, 1.0, 0.3, 0.1)]:
        rng = np.random.default_rng(h + w)
        c, n, a = _inputs(rng, h, w)
        got = _run(nw, ctx, c, n, a, iterations=5, sigma_color=sc, sigma_normal=sn,
                   sigma_albedo=sa)
        want = orc.denoise(c, n, a, iterations=5, sigma_color=sc, sigma_normal=sn,
                           sigma_albedo=sa)
        _check(got, want)


def test_denoise_blend_one_and_constant_image(nw, ctx):
    rng = np.random.default_rng(9)
    c, n, a = _inputs(rng, 24, 40)
    got = _run(nw, ctx, c, n, a, blend=1.0)
    assert np.array_equal(got, c)
    k = np.empty_like(c)
    k[...] = [2.5, 0.75, 0.125, 0.5]
    got = _run(nw, ctx, k, n, a)
    assert np.allclose(got, k, rtol=1e-6, atol=0)


def test_denoise_errors(nw, ctx):
    h, w = 8, 8
    bufs = [DeviceBuffer(h * w * 16) for _ in range(4)]
    c, n, a, o = bufs
    bad = []
    p = nw.Context.denoise_default_params()
    for field, value in (("iterations", 0), ("iterations", 9), ("sigma_color", 0.0),
                         ("sigma_normal", -1.0), ("sigma_albedo", float("inf"))):
        q = nw.Context.denoise_default_params()
        setattr(q, field, value)
        bad.append(lambda q=q: ctx.denoise(c, n, a, o, w, h, 0.0, q))
    bad += [lambda: ctx.denoise(c, n, a, o, w, h, 1.5, p),
            lambda: ctx.denoise(c, n, a, o, 0, h, 0.0, p),
            lambda: ctx.denoise(c, n, a, c, w, h, 0.0, p),         # output = colour
            lambda: ctx.denoise(c, n, a, n.ptr + 64, w, h, 0.0, p),  # overlaps normal
            lambda: ctx.denoise(c, n, a, 0, w, h, 0.0, p),
            lambda: ctx.denoise(c.ptr + 4, n, a, o, w, h, 0.0, p)]   # misaligned
    for f in bad:
        with pytest.raises(nw.NwprtError) as e:
            f()
        assert e.value.status == nw.ERR_INVALID_ARGUMENT
