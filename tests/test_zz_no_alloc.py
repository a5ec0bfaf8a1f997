"""GPU test of nwprt_relight_noalloc, the allocation-free relight: through
the C ABI, its outputs equal nwprt_relight's bit for bit (same kernels, the
library's per-device scratch array as work memory) on two calls in a row,
the second with a smaller frame and K, for BF16, FP16 and FP8; the forms
that keep per-pixel buffers are refused."""
import numpy as np
import pytest

from cases import make_case

pytestmark = pytest.mark.gpu

SIZES = [(203, 117, 300, 32, 0.1), (40, 9, 17, 8, 0.0)]


@pytest.mark.parametrize("precision", ["bf16", "fp16", "fp8"])
def test_noalloc_equals_relight(nw, orc, ctx, precision):
    prec = nw.PRECISIONS[precision]
    for i, (w, h, k, n, bg) in enumerate(SIZES):
        case = make_case(nw, orc, width=w, height=h, k=k, face_size=n, background=bg,
                         hash_dtype=nw.DTYPE_F16, seed=60 + i)
        scene = ctx.scene_create(case.desc, case.params)
        env = ctx.envmap_from_faces(case.faces, k)
        ref = nw.DeviceFrame(w, h, case.gbuf)
        ctx.relight(scene, env, ref, prec, nw.TONEMAP_FULL)
        ctx.sync()
        want = ref.results()
        got_frame = nw.DeviceFrame(w, h, case.gbuf)
        ctx.relight_noalloc(scene, env, got_frame, prec, nw.TONEMAP_FULL)
        ctx.sync()
        got = got_frame.results()
        for key in want:
            assert np.array_equal(got[key], want[key]), (precision, i, key)
        assert np.count_nonzero(got["indirect"][:, :3]) > 0
        print(f"noalloc {precision} call {i}: {w}x{h} K={k} equal to nwprt_relight")


def test_noalloc_refuses_forms_with_buffers(nw, orc, ctx):
    case = make_case(nw, orc, width=16, height=8, k=8, face_size=8, aux=True, seed=3)
    scene = ctx.scene_create(case.desc, case.params)
    env = ctx.envmap_from_faces(case.faces, 8)
    frame = nw.DeviceFrame(16, 8, case.gbuf)
    with pytest.raises(nw.NwprtError) as e:      # aux inputs: per-pixel buffers
        ctx.relight_noalloc(scene, env, frame, nw.PREC_BF16)
    assert e.value.status == nw.ERR_UNSUPPORTED
    plain = make_case(nw, orc, width=16, height=8, k=8, face_size=8, seed=4)
    scene2 = ctx.scene_create(plain.desc, plain.params)
    env2 = ctx.envmap_from_faces(plain.faces, 8)
    with pytest.raises(nw.NwprtError) as e:      # the plain FP32 form
        ctx.relight_noalloc(scene2, env2, nw.DeviceFrame(16, 8, plain.gbuf), nw.PREC_FP32)
    assert e.value.status == nw.ERR_UNSUPPORTED
