"""GPU test of the two-slot frame pipeline (Renderer/OutputVisualizer.cpp:71-123):
p1 and the map update of frame f + 1 run on a second stream while p2 of
frame f runs, ordered by events.  Every frame's outputs must equal, bit for
bit, a single nwprt_relight of the same G-buffer and cubemap."""
import numpy as np
import pytest

from cases import make_case

pytestmark = pytest.mark.gpu


def _ref(nw, ctx, scene, env, frame, prec):
    ref = nw.DeviceFrame(frame.width, frame.height)
    ref.bufs = frame.bufs  # same G-buffer, own outputs
    ctx.relight(scene, env, ref, prec)
    ctx.sync()
    return ref.results()


@pytest.mark.parametrize("precision", ["bf16", "fp16"])
def test_two_slot_pipeline_matches_single_relights(nw, orc, ctx, precision):
    from neural_wavelet_prt_demo_amd import DeviceBuffer
    from neural_wavelet_prt_demo_amd.pipeline import FramePipeline
    w, h, k, n = 96, 80, 40, 8
    prec = nw.PRECISIONS[precision]
    base = make_case(nw, orc, width=8, height=8, k=k, face_size=n, seed=70)
    scene = ctx.scene_create(base.desc, base.params)
    # Four distinct frames (G-buffer + cubemap), cycled through the pipeline.
    cases = [make_case(nw, orc, width=w, height=h, k=k, face_size=n, seed=80 + i,
                       background=0.1) for i in range(4)]
    faces = [DeviceBuffer.from_array(c.faces) for c in cases]
    envs = [ctx.envmap_from_faces(cases[i].faces, k) for i in range(2)]
    check_env = ctx.envmap_from_faces(cases[0].faces, k)
    frames = [nw.DeviceFrame(w, h, c.gbuf) for c in cases]
    pipe = FramePipeline(ctx, scene, envs, prec)
    for f in range(7):
        i = f % 4
        pipe.submit(frames[i], faces[i])
    pipe.drain()
    # Frames 3..6 were the last to write frames[3], [0], [1], [2].
    for i in range(4):
        ctx.envmap_update_faces(check_env, faces[i])
        want = _ref(nw, ctx, scene, check_env, frames[i], prec)
        got = frames[i].results()
        for key in ("indirect", "indirect_with_direct", "rgba8"):
            assert np.array_equal(got[key], want[key]), (precision, i, key)


def test_pipeline_needs_two_maps(nw):
    from neural_wavelet_prt_demo_amd.pipeline import FramePipeline
    with pytest.raises(ValueError):
        FramePipeline(None, 1, [1], nw.PREC_BF16)
