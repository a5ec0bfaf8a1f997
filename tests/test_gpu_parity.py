"""GPU parity of the HIP path against the CPU oracle, through the C ABI.

Bars: bit-exact for Haar coefficients, top-K index sets (and strengths) and
hash-grid entry indices; radiance within the tolerances stated per test
against the FP64 oracle.
"""
import numpy as np
import pytest

from cases import make_case
from conftest import max_rel_err, psnr

pytestmark = pytest.mark.gpu

# Tolerances (radiance vs the FP64 oracle, per pixel and channel, relative
# to max(|ref|, 1e-3 * max|ref|)).
TOL_FP32 = 1e-4
TOL_HALF = {"bf16": 2e-2, "fp16": 5e-3}
MIN_PSNR_DB = 50.0


def _gpu_relight(nw, ctx, case, precision, env=None, scene=None, tonemap=None):
    if scene is None:
        scene = ctx.scene_create(case.desc, case.params)
    if env is None:
        env = ctx.envmap_from_faces(case.faces, case.k)
    frame = nw.DeviceFrame(case.width, case.height, case.gbuf)
    ctx.relight(scene, env, frame, nw.PRECISIONS[precision],
                nw.TONEMAP_FULL if tonemap is None else tonemap)
    ctx.sync()
    return scene, env, frame, frame.results()


@pytest.mark.parametrize("n", [2, 4, 32, 64, 128])
def test_haar_and_topk_bit_exact(nw, orc, ctx, n):
    faces = nw.synth_cubemap(7 + n, n) if n >= 4 else np.random.default_rng(n).lognormal(
        0, 1.5, (6, n, n, 3)).astype(np.float32)
    k = min(300, 6 * n * n)
    env = ctx.envmap_from_faces(faces, k)
    coef = ctx.envmap_coefficients(env)
    ref = orc.envmap_haar(faces)
    assert np.array_equal(coef.view(np.uint32), ref.view(np.uint32))
    idx, st = ctx.envmap_selection(env)
    ridx, rst = orc.topk(ref, k)
    assert np.array_equal(idx, ridx)
    assert np.array_equal(st.view(np.uint32), rst.view(np.uint32))
    ctx.envmap_destroy(env)


@pytest.mark.parametrize("n,n_maps", [(2, 3), (8, 5), (32, 2), (64, 3), (128, 7)])
def test_haar_many_maps_bit_exact(nw, orc, ctx, n, n_maps):
    """nwprt_envmap_haar over several maps at once (tiles of 32 x 32 texels,
    then the levels above them) equals the oracle's whole-face transform of
    each map, bit for bit."""
    from neural_wavelet_prt_demo_amd import DeviceBuffer
    rng = np.random.default_rng(n * 10 + n_maps)
    faces = rng.lognormal(0, 1.5, (n_maps, 6, n, n, 3)).astype(np.float32)
    faces[0, 2, n // 2, n // 3] = 1e3  # a sun texel
    dfaces = DeviceBuffer.from_array(faces)
    dcoef = DeviceBuffer(faces.nbytes)
    ctx.envmap_haar(dfaces, n, n_maps, dcoef)
    ctx.sync()
    got = dcoef.download(np.float32, faces.shape)
    for m in range(n_maps):
        assert np.array_equal(got[m].view(np.uint32), orc.envmap_haar(faces[m]).view(np.uint32)), m
    with pytest.raises(nw.NwprtError) as e:
        ctx.envmap_haar(dfaces, n, n_maps, dfaces)   # in place
    assert e.value.status == nw.ERR_INVALID_ARGUMENT


def test_topk_ties_break_by_index(nw, orc, ctx):
    n = 16
    faces = np.ones((6, n, n, 3), np.float32)  # all DCs equal, all details 0
    faces[2, 3, 5] = 4.0                          # one bump on face 2
    for k in (1, 6, 7, 40, 6 * n * n):
        env = ctx.envmap_from_faces(faces, min(k, 1024))
        idx, _ = ctx.envmap_selection(env)
        ridx, _ = orc.topk(orc.envmap_haar(faces), min(k, 1024))
        assert np.array_equal(idx, ridx), k
        ctx.envmap_destroy(env)


def test_topk_large_tied_bucket(nw, orc, ctx):
    """A per-pixel checkerboard gives every finest diagonal detail the same
    score: 6 x 32^2 = 6144 equal keys at n = 64, more than the selection
    keeps in LDS, so the kernel resolves them with passes over all keys."""
    n = 64
    yy, xx = np.mgrid[0:n, 0:n]
    board = np.where((xx + yy) % 2 == 0, 1.0, -1.0).astype(np.float32)
    faces = np.repeat(board[None, :, :, None], 6, axis=0).repeat(3, axis=3) * 5.0 + 10.0
    faces = np.ascontiguousarray(faces, np.float32)
    for k in (300, 1024):
        env = ctx.envmap_from_faces(faces, k)
        idx, st = ctx.envmap_selection(env)
        ridx, rst = orc.topk(orc.envmap_haar(faces), k)
        assert np.array_equal(idx, ridx), k
        assert np.array_equal(st.view(np.uint32), rst.view(np.uint32)), k
        ctx.envmap_destroy(env)


def test_hash_indices_bit_exact(nw, orc, ctx):
    case = make_case(nw, orc, width=64, height=32, hash_dtype=nw.DTYPE_F16)
    scene = ctx.scene_create(case.desc, case.params)
    pos = case.gbuf["position"].copy()
    special = np.array([[0, 0, 0, 1], [1, 1, 1, 1], [0.5, 0.5, 0.5, 1], [1, 0, 0.25, 1],
                        [1 / 15, 2 / 15, 3 / 15, 1], [0.999999, 1e-7, 0.5, 1]], np.float32)
    pos[:len(special)] = special
    gidx, gfeat = ctx.debug_hash_encode(scene, pos)
    for i in range(0, len(pos), 7):
        ridx, rfeat = case.oscene.hash_encode(pos[i, :3])
        assert np.array_equal(gidx[i], ridx), i
        np.testing.assert_allclose(gfeat[i], rfeat, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("precision", ["fp32", "bf16", "fp16"])
def test_relight_config0_vs_oracle(nw, orc, ctx, precision):
    """128 x 128, N = 32, K = 64 (the CPU-runnable configuration)."""
    case = make_case(nw, orc)
    scene, env, frame, res = _gpu_relight(nw, ctx, case, precision)
    idx, st = ctx.envmap_selection(env)
    ind, wd = case.oscene.relight(idx, st, case.gbuf)
    got = res["indirect"][:, :3]
    err = max_rel_err(ind, got)
    p = psnr(ind, got)
    print(f"{precision}: max rel err {err:.3e}, PSNR {p:.1f} dB")
    assert np.all(res["indirect"][:, 3] == 1.0) and np.all(res["indirect_with_direct"][:, 3] == 1.0)
    np.testing.assert_allclose(res["indirect_with_direct"][:, :3],
                               got + case.gbuf["direct"][:, :3], rtol=1e-6, atol=1e-6)
    if precision == "fp32":
        assert err <= TOL_FP32
    else:
        assert err <= TOL_HALF[precision]
        assert p >= MIN_PSNR_DB
    # RGBA8 is the tone map of the with-direct output, byte for byte.
    tm = orc.tonemap(res["indirect_with_direct"][:, :3])
    assert np.array_equal(tm, res["rgba8"])


@pytest.mark.parametrize("precision", ["fp32", "bf16", "fp16"])
def test_transport_amplified_vs_oracle(nw, orc, ctx, precision):
    """Per-pair transport with unit-scale features, so every layer matters."""
    case = make_case(nw, orc, width=32, height=16, k=40, hash_std=1.0, feature_std=1.0, seed=3)
    scene = ctx.scene_create(case.desc, case.params)
    env = ctx.envmap_from_faces(case.faces, case.k)
    frame = nw.DeviceFrame(case.width, case.height, case.gbuf)
    begin, count = 100, 200
    t_gpu = ctx.debug_transport(scene, env, frame, begin, count, nw.PRECISIONS[precision])
    idx, st = ctx.envmap_selection(env)
    _, _, t_ref = case.oscene.relight(idx, st, case.gbuf, pixels=np.arange(begin, begin + count),
                                      transport=True)
    spread = float(np.std(t_ref))
    err = np.abs(t_gpu - t_ref)
    print(f"{precision}: T spread {spread:.3f}, max |dT| {err.max():.3e}")
    assert spread > 0.1  # the network is not near-constant in this case
    tol = {"fp32": 1e-4, "bf16": 0.05, "fp16": 0.01}[precision]
    assert err.max() <= tol * max(1.0, float(np.abs(t_ref).max()))


def test_background_pixels(nw, orc, ctx):
    case = make_case(nw, orc, width=40, height=30, background=0.4)
    for precision in ("fp32", "bf16"):
        _, _, _, res = _gpu_relight(nw, ctx, case, precision)
        bg = case.gbuf["env_flag"] == 0
        assert bg.any() and (~bg).any()
        assert np.all(res["indirect"][bg, :3] == 0.0)
        assert np.array_equal(res["indirect_with_direct"][bg, :3], case.gbuf["direct"][bg, :3])
        assert np.all(res["indirect"][~bg, :3] > 0.0)


@pytest.mark.parametrize("shape", [(1, 1), (37, 13), (65, 3), (300, 1)])
@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_ragged_frames(nw, orc, ctx, shape, precision):
    w, h = shape
    case = make_case(nw, orc, width=w, height=h, k=17, face_size=8, seed=w * h)
    _, env, _, res = _gpu_relight(nw, ctx, case, precision)
    idx, st = ctx.envmap_selection(env)
    ind, _ = case.oscene.relight(idx, st, case.gbuf)
    tol = TOL_FP32 if precision == "fp32" else TOL_HALF[precision]
    assert max_rel_err(ind, res["indirect"][:, :3]) <= tol


@pytest.mark.parametrize("precision,hash_f16,shape,background",
                         [("bf16", False, (96, 40), 0.0), ("bf16", True, (97, 41), 0.3),
                          ("fp16", False, (33, 7), 0.2), ("fp16", True, (64, 64), 0.0)])
def test_split_passes_match_relight(nw, orc, ctx, precision, hash_f16, shape, background):
    """nwprt_relight encodes inside the fused kernel (each wave its own
    pixels); p1 + p2 run the separate planar encode.  Both give the same
    outputs bit for bit."""
    w, h = shape
    case = make_case(nw, orc, width=w, height=h, background=background,
                     hash_dtype=nw.DTYPE_F16 if hash_f16 else nw.DTYPE_F32)
    prec = nw.PRECISIONS[precision]
    scene, env, frame, res = _gpu_relight(nw, ctx, case, precision)
    frame2 = nw.DeviceFrame(case.width, case.height, case.gbuf)
    ctx.TCNNPass_p1(scene, 1, frame2, prec)
    ctx.TCNNPass_p2(scene, env, 1, frame2, prec)
    ctx.sync()
    res2 = frame2.results()
    for key in res:
        a, b = res[key], res2[key]
        bad = ~((a == b) | (np.isnan(a.astype(np.float64)) & np.isnan(b.astype(np.float64))))
        where = np.argwhere(bad)[:5].tolist()
        assert not bad.any(), (key, int(bad.sum()), where,
                               [(a[tuple(i)], b[tuple(i)]) for i in where],
                               bool(np.isfinite(a.astype(np.float64)).all()))
    # p2 on a slot encoded for another precision is a call-order error.
    with pytest.raises(nw.NwprtError) as e:
        ctx.TCNNPass_p2(scene, env, 1, frame2, nw.PREC_FP32)
    assert e.value.status == nw.ERR_STATE


def test_selection_path_matches_faces_path(nw, orc, ctx):
    case = make_case(nw, orc, width=64, height=64)
    scene, env, _, res = _gpu_relight(nw, ctx, case, "bf16")
    idx, st = orc.topk(orc.envmap_haar(case.faces), case.k)
    env2 = ctx.envmap_from_selection(idx, st, case.desc.face_size)
    _, _, _, res2 = _gpu_relight(nw, ctx, case, "bf16", env=env2, scene=scene)
    for key in res:
        assert np.array_equal(res[key], res2[key]), key


def test_light_scaling_is_exact(nw, orc, ctx):
    """Doubling every light coefficient doubles each clamped term exactly in
    FP32, so the compensated sum doubles bit for bit (any size)."""
    case = make_case(nw, orc, width=256, height=64, k=96)
    scene = ctx.scene_create(case.desc, case.params)
    idx, st = orc.topk(orc.envmap_haar(case.faces), case.k)
    e1 = ctx.envmap_from_selection(idx, st, case.desc.face_size)
    e2 = ctx.envmap_from_selection(idx, st * 2, case.desc.face_size)
    for precision in ("fp32", "bf16", "fp16"):
        _, _, _, r1 = _gpu_relight(nw, ctx, case, precision, env=e1, scene=scene)
        _, _, _, r2 = _gpu_relight(nw, ctx, case, precision, env=e2, scene=scene)
        assert np.array_equal(r1["indirect"][:, :3] * 2, r2["indirect"][:, :3]), precision


def test_tonemap_entry_point(nw, orc, ctx):
    rng = np.random.default_rng(5)
    n = 4096
    src = np.zeros((n, 4), np.float32)
    src[:, :3] = rng.lognormal(-1, 1.5, (n, 3))
    src[:8, :3] = [[0, 0, 0], [1, 1, 1], [1e6, 0, 0], [0, 1e-7, 2], [0.5, 0.25, 0.125],
                   [0.0031308, 0.04, 0.9], [10, 10, 10], [1e-30, 0, 0]]
    ds, dd = nw.DeviceBuffer.from_array(src), nw.DeviceBuffer(n * 4)
    ctx.computeFinalPixelColors(ds, dd, 64, 64)
    ctx.sync()
    got = dd.download(np.uint32, (n,))
    ref = orc.tonemap(src[:, :3])
    assert np.array_equal(got, ref)
    assert got[0] == 0xFF000000 and got[1] == 0xFFFFFFFF


def test_tonemap_thresholds_exact(nw, orc, ctx):
    """Every byte boundary, the float below it, NaN, +-inf and negatives:
    the device byte equals the oracle's."""
    t = orc.tonemap_thresholds()
    below = np.nextafter(t, np.float32(-np.inf))
    above = np.nextafter(t, np.float32(np.inf))
    special = np.array([np.nan, np.inf, -np.inf, -1.0, -0.0, 0.0, 1e-38, 3.4e38], np.float32)
    vals = np.concatenate([t, below, above, special])
    vals = np.concatenate([vals, np.zeros((-len(vals)) % 3, np.float32)])
    n = len(vals) // 3
    src = np.ones((n, 4), np.float32)
    src[:, :3] = vals.reshape(n, 3)
    ds, dd = nw.DeviceBuffer.from_array(src), nw.DeviceBuffer(n * 4)
    ctx.computeFinalPixelColors(ds, dd, n, 1)
    ctx.sync()
    got = dd.download(np.uint32, (n,))
    assert np.array_equal(got, orc.tonemap(src[:, :3]))


def test_error_paths(nw, orc, ctx):
    case = make_case(nw, orc, width=8, height=8, k=8, face_size=8)
    scene = ctx.scene_create(case.desc, case.params)
    env = ctx.envmap_from_faces(case.faces, 8)
    frame = nw.DeviceFrame(8, 8, case.gbuf)
    with pytest.raises(nw.NwprtError) as e:
        ctx.relight(scene + 1000, env, frame)
    assert e.value.status == nw.ERR_NOT_FOUND
    with pytest.raises(nw.NwprtError) as e:  # no such precision
        ctx.relight(scene, env, frame, 7)
    assert e.value.status == nw.ERR_INVALID_ARGUMENT
    with pytest.raises(nw.NwprtError) as e:
        ctx.envmap_from_faces(case.faces, 6 * 8 * 8 + 1)
    assert e.value.status == nw.ERR_INVALID_ARGUMENT
    with pytest.raises(nw.NwprtError) as e:
        ctx.envmap_from_selection(np.array([6 * 64], np.int32), np.ones((1, 3), np.float32), 8)
    assert e.value.status == nw.ERR_INVALID_ARGUMENT
    other = ctx.envmap_from_faces(nw.synth_cubemap(1, 16), 8)
    with pytest.raises(nw.NwprtError) as e:  # face size differs from the scene's table
        ctx.relight(scene, other, frame)
    assert e.value.status == nw.ERR_INVALID_ARGUMENT
    with pytest.raises(nw.NwprtError) as e:  # slot 0 never encoded
        ctx.TCNNPass_p2(scene, env, 0, frame)
    assert e.value.status == nw.ERR_STATE


def test_config1_fp16_1080p(nw, orc, ctx):
    """1920 x 1080, N = 64, K = 128, FP16 hash grid and MLP: bit-exact hash
    indices and top-K set, radiance on a pixel sample within tolerance."""
    case = make_case(nw, orc, width=1920, height=1080, face_size=64, k=128,
                     hash_dtype=nw.DTYPE_F16, seed=1)
    scene, env, frame, res = _gpu_relight(nw, ctx, case, "fp16")
    idx, st = ctx.envmap_selection(env)
    ridx, rst = orc.topk(orc.envmap_haar(case.faces), case.k)
    assert np.array_equal(idx, ridx)
    rng = np.random.default_rng(11)
    pix = rng.choice(case.width * case.height, 1500, replace=False)
    gidx, _ = ctx.debug_hash_encode(scene, case.gbuf["position"][pix])
    for j in range(0, len(pix), 25):
        ridx_h, _ = case.oscene.hash_encode(case.gbuf["position"][pix[j], :3])
        assert np.array_equal(gidx[j], ridx_h)
    ind, _ = case.oscene.relight(idx, st, case.gbuf, pixels=pix)
    got = res["indirect"][pix, :3]
    err, p = max_rel_err(ind, got), psnr(ind, got)
    print(f"config1 fp16: max rel err {err:.3e}, PSNR {p:.1f} dB")
    assert err <= TOL_HALF["fp16"] and p >= MIN_PSNR_DB


def test_every_tile_variant_matches_and_autotune_records(nw, orc, ctx):
    """All fused-kernel tile variants compute the same outputs bit for bit
    (same per-lane operation order), and the autotuner's choice is used."""
    case = make_case(nw, orc, width=200, height=37, k=45)
    scene = ctx.scene_create(case.desc, case.params)
    env = ctx.envmap_from_faces(case.faces, case.k)
    ref = None
    for v in range(len(nw.Context.relight_variants())):
        ctx.set_relight_variant(v)
        _, _, _, res = _gpu_relight(nw, ctx, case, "bf16", env=env, scene=scene)
        if ref is None:
            ref = res
        for key in ref:
            assert np.array_equal(ref[key], res[key]), (v, key)
    ctx.set_relight_variant(-1)
    frame = nw.DeviceFrame(case.width, case.height, case.gbuf)
    best, ms = ctx.autotune_relight(scene, env, frame, nw.PREC_BF16)
    assert 0 <= best < len(nw.Context.relight_variants()) and ms > 0
    assert ctx.selected_relight_variant(nw.PREC_BF16, case.k, case.width * case.height) == best
