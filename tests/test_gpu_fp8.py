"""GPU tests of the FP8 form (hidden layers FP8-E4M3 with per-layer
power-of-two scales from calibrated activation maxima, first layer BF16,
FP32 accumulation) against the FP64 oracle, through the C ABI.

FP8-E4M3 keeps 3 mantissa bits (relative rounding step 2^-4), so the per-pair
transport carries a few percent of error; the per-pixel K-sum averages much
of it out.  The bars below are stated per test.
"""
import numpy as np
import pytest

from cases import make_case
from conftest import max_rel_err, psnr

pytestmark = pytest.mark.gpu

# Radiance vs the FP64 oracle (per pixel and channel, relative to
# max(|ref|, 1e-3 max|ref|)) and PSNR on linear radiance.
TOL_FP8 = 0.1
MIN_PSNR_FP8_DB = 50.0


def _relight(nw, ctx, case, precision, scene=None):
    scene = ctx.scene_create(case.desc, case.params) if scene is None else scene
    env = ctx.envmap_from_faces(case.faces, case.k)
    frame = nw.DeviceFrame(case.width, case.height, case.gbuf)
    ctx.relight(scene, env, frame, nw.PRECISIONS[precision])
    ctx.sync()
    return scene, env, frame.results()


def test_fp8_config0_vs_oracle(nw, orc, ctx):
    """128 x 128, N = 32, K = 64: FP8 next to BF16 on the same frame."""
    case = make_case(nw, orc)
    scene, env, res = _relight(nw, ctx, case, "fp8")
    idx, st = ctx.envmap_selection(env)
    ref, _ = case.oscene.relight(idx, st, case.gbuf)
    got = res["indirect"][:, :3]
    err, p = max_rel_err(ref, got), psnr(ref, got)
    _, _, res16 = _relight(nw, ctx, case, "bf16", scene)
    err16, p16 = max_rel_err(ref, res16["indirect"][:, :3]), psnr(ref, res16["indirect"][:, :3])
    print(f"fp8: max rel err {err:.3e}, PSNR {p:.1f} dB; bf16: {err16:.3e}, {p16:.1f} dB")
    assert err <= TOL_FP8 and p >= MIN_PSNR_FP8_DB
    np.testing.assert_allclose(res["indirect_with_direct"][:, :3], got + case.gbuf["direct"][:, :3],
                               rtol=1e-6, atol=1e-6)


def test_fp8_transport_amplified_vs_oracle(nw, orc, ctx):
    """Per-pair transport with unit-scale features, so every layer matters."""
    case = make_case(nw, orc, width=32, height=16, k=40, hash_std=1.0, feature_std=1.0, seed=3)
    scene = ctx.scene_create(case.desc, case.params)
    env = ctx.envmap_from_faces(case.faces, case.k)
    frame = nw.DeviceFrame(case.width, case.height, case.gbuf)
    begin, count = 100, 200
    t8 = ctx.debug_transport(scene, env, frame, begin, count, nw.PREC_FP8)
    idx, st = ctx.envmap_selection(env)
    _, _, t_ref = case.oscene.relight(idx, st, case.gbuf, pixels=np.arange(begin, begin + count),
                                      transport=True)
    scale = max(1.0, float(np.abs(t_ref).max()))
    err = np.abs(t8 - t_ref) / scale
    print(f"fp8 transport: spread {np.std(t_ref):.3f}, max |dT|/scale {err.max():.3e}, "
          f"mean {err.mean():.3e}")
    assert np.std(t_ref) > 0.1
    assert err.max() <= 0.25 and err.mean() <= 0.03


def test_fp8_scale_equivariance_is_exact(nw, orc, ctx):
    """Per-layer amax scaling with power-of-two scales: multiplying the hash
    table by 2^s multiplies every activation by 2^s, the calibrated maxima by
    2^s and the chosen exponents by -s, so with an identity output the
    radiance is multiplied by exactly 2^s, bit for bit."""
    case = make_case(nw, orc, width=64, height=24, k=32, softplus=False, seed=11)
    _, _, res = _relight(nw, ctx, case, "fp8")
    base = res["indirect"][:, :3]
    assert np.count_nonzero(base) > base.size // 4
    for s in (6, -5):
        params = dict(case.params)
        params["hash_table"] = (case.params["hash_table"] * np.float32(2.0 ** s)).astype(np.float32)
        _, _, res_s = _relight(nw, ctx, case, "fp8", ctx.scene_create(case.desc, params))
        assert np.array_equal(res_s["indirect"][:, :3], base * np.float32(2.0 ** s)), s


def test_fp8_background_and_ragged(nw, orc, ctx):
    for (w, h), bgf in (((1, 1), 0.0), ((37, 13), 0.3), ((300, 1), 0.0), ((65, 3), 0.5)):
        case = make_case(nw, orc, width=w, height=h, k=17, face_size=8, seed=w * h,
                         background=bgf)
        _, env, res = _relight(nw, ctx, case, "fp8")
        idx, st = ctx.envmap_selection(env)
        ref, _ = case.oscene.relight(idx, st, case.gbuf)
        got = res["indirect"][:, :3]
        bg = case.gbuf["env_flag"] == 0
        assert np.all(got[bg] == 0.0)
        assert max_rel_err(ref, got) <= TOL_FP8, (w, h)


def test_fp8_fused_encode_equals_p1_p2(nw, orc, ctx):
    """The one-call FP8 relight encodes inside the kernel (and encodes its
    calibration sample itself); p1 + p2 read the standalone encode's
    features.  Both give the same outputs bit for bit, FP16 and FP32 tables,
    ragged frames with background pixels."""
    for hd, (w, h), bg in ((nw.DTYPE_F16, (97, 41), 0.2), (nw.DTYPE_F32, (64, 64), 0.0),
                           (nw.DTYPE_F16, (1, 1), 0.0)):
        case = make_case(nw, orc, width=w, height=h, k=37, face_size=16, hash_dtype=hd,
                         background=bg, seed=w + h)
        scene, env, fused = _relight(nw, ctx, case, "fp8")
        frame = nw.DeviceFrame(w, h, case.gbuf)
        ctx.TCNNPass_p1(scene, 0, frame, nw.PREC_FP8)
        ctx.TCNNPass_p2(scene, env, 0, frame, nw.PREC_FP8, nw.TONEMAP_FULL)
        ctx.sync()
        split = frame.results()
        for key in fused:
            assert np.array_equal(fused[key], split[key]), (w, h, key)


def test_fp8_variants_agree(nw, orc, ctx):
    """The FP8 kernel forms (plain, software-pipelined, groups in sequence
    at three waves per SIMD) run the same per-lane arithmetic in the same
    order: identical outputs, also with aux inputs and an odd K."""
    for aux in (False, True):
        case = make_case(nw, orc, width=150, height=37, k=45, aux=aux, seed=21)
        scene = ctx.scene_create(case.desc, case.params)
        env = ctx.envmap_from_faces(case.faces, case.k)
        ref = None
        try:
            for v in range(len(nw.Context.relight_fp8_variants())):
                ctx.set_relight_variant(v)
                frame = nw.DeviceFrame(case.width, case.height, case.gbuf)
                ctx.relight(scene, env, frame, nw.PREC_FP8)
                ctx.sync()
                res = frame.results()
                if ref is None:
                    ref = res
                for key in ref:
                    assert np.array_equal(ref[key], res[key]), (aux, v, key)
        finally:
            ctx.set_relight_variant(-1)


def test_fp8_errors_and_autotune(nw, orc, ctx):
    case = make_case(nw, orc, width=64, height=32, k=24, seed=5)
    scene = ctx.scene_create(case.desc, case.params)
    env = ctx.envmap_from_faces(case.faces, case.k)
    frame = nw.DeviceFrame(case.width, case.height, case.gbuf)
    names = nw.Context.relight_fp8_variants()
    assert len(names) >= 1
    best, ms = ctx.autotune_relight(scene, env, frame, nw.PREC_FP8)
    assert 0 <= best < len(names) and ms > 0
    assert ctx.selected_relight_variant(nw.PREC_FP8, case.k, case.width * case.height) == best
