"""GPU tests of the batch form: uneven relight jobs (own scene, map, K, frame
size) in one captured graph with one persistent kernel.  Each job's outputs
must equal, bit for bit, a single nwprt_relight with the same precision and
tile variant -- on every replay of the graph, after an in-place update of a
job's environment map, and after batch autotuning."""
import numpy as np
import pytest

from cases import make_case

pytestmark = pytest.mark.gpu

# (width, height, K, face size, scene): uneven pixel counts and K, two scenes.
JOBS = [(40, 30, 17, 8, 0), (64, 64, 64, 8, 1), (1, 1, 5, 8, 0), (100, 7, 130, 8, 1),
        (33, 17, 300, 8, 0), (128, 96, 32, 8, 1)]


def _setup(nw, orc, ctx, precision="bf16"):
    scenes = [make_case(nw, orc, width=8, height=8, k=8, face_size=8, seed=100 + s)
              for s in (0, 1)]
    sids = [ctx.scene_create(c.desc, c.params) for c in scenes]
    jobs, frames, envs, cases = [], [], [], []
    for j, (w, h, k, n, s) in enumerate(JOBS):
        case = make_case(nw, orc, width=w, height=h, k=k, face_size=n, seed=j, background=0.1)
        env = ctx.envmap_from_faces(case.faces, k)
        frame = nw.DeviceFrame(w, h, case.gbuf)
        jobs.append((sids[s], env, frame))
        frames.append(frame)
        envs.append(env)
        cases.append(case)
    return sids, jobs, frames, envs, cases


def _single(nw, ctx, job, precision, variant):
    scene, env, frame = job
    ref = nw.DeviceFrame(frame.width, frame.height)
    ref.bufs = frame.bufs  # same G-buffer, own outputs
    ctx.set_relight_variant(variant)
    ctx.relight(scene, env, ref, nw.PRECISIONS[precision])
    ctx.sync()
    ctx.set_relight_variant(-1)
    return ref.results()


def _assert_same(got, want, what):
    for key in ("indirect", "indirect_with_direct", "rgba8"):
        assert np.array_equal(got[key], want[key]), (what, key)


@pytest.mark.parametrize("precision", ["bf16", "fp16"])
def test_batch_matches_single_relights(nw, orc, ctx, precision):
    sids, jobs, frames, envs, cases = _setup(nw, orc, ctx)
    bid = ctx.batch_create(jobs, nw.PRECISIONS[precision])
    n_items, n_pairs, variant = ctx.batch_info(bid)
    assert n_items == sum((w * h + 63) // 64 for w, h, *_ in JOBS)
    assert n_pairs == sum(w * h * k for w, h, k, *_ in JOBS)
    want = [_single(nw, ctx, job, precision, variant) for job in jobs]
    extra = nw.DeviceFrame(8, 8, nw.synth_gbuffer(1, 8, 8, 0.0))
    for rep in range(3):  # graph replays, with eager work in between
        for f in frames:
            for b in f.out.values():
                b.zero()
        ctx.batch_launch(bid)
        ctx.sync()
        for j, f in enumerate(frames):
            _assert_same(f.results(), want[j], (rep, j))
        ctx.relight(jobs[0][0], jobs[0][1], extra)
        ctx.sync()
    ctx.batch_destroy(bid)


def test_batch_follows_envmap_update_and_autotune(nw, orc, ctx):
    sids, jobs, frames, envs, cases = _setup(nw, orc, ctx)
    bid = ctx.batch_create(jobs, nw.PREC_BF16)
    ctx.batch_launch(bid)
    ctx.sync()
    # New faces for job 3's map, same K: the next launch sees the new selection.
    ctx.envmap_update_faces(envs[3], nw.synth_cubemap(77, 8))
    ctx.batch_launch(bid)
    ctx.sync()
    _, _, variant = ctx.batch_info(bid)
    _assert_same(frames[3].results(), _single(nw, ctx, jobs[3], "bf16", variant), "updated")
    best, ms = ctx.batch_autotune(bid)
    assert best in (0, 1, 5) and ms > 0 and ctx.batch_info(bid)[2] == best
    ctx.batch_launch(bid)
    ctx.sync()
    for j, f in enumerate(frames):
        _assert_same(f.results(), _single(nw, ctx, jobs[j], "bf16", best), ("tuned", j))
    # A destroyed map makes the next launch fail cleanly.
    ctx.envmap_destroy(envs[2])
    with pytest.raises(nw.NwprtError) as e:
        ctx.batch_launch(bid)
    assert e.value.status == nw.ERR_NOT_FOUND
    ctx.batch_destroy(bid)
    with pytest.raises(nw.NwprtError) as e:
        ctx.batch_launch(bid)
    assert e.value.status == nw.ERR_NOT_FOUND


def test_batch_rejects_unsupported(nw, orc, ctx):
    sids, jobs, frames, envs, cases = _setup(nw, orc, ctx)
    for prec in (nw.PREC_FP32, nw.PREC_FP8):
        with pytest.raises(nw.NwprtError) as e:
            ctx.batch_create(jobs, prec)
        assert e.value.status == nw.ERR_UNSUPPORTED
    aux = make_case(nw, orc, width=8, height=8, k=8, face_size=8, aux=True, seed=3)
    aux_scene = ctx.scene_create(aux.desc, aux.params)
    with pytest.raises(nw.NwprtError) as e:
        ctx.batch_create([(aux_scene, envs[0], frames[0])], nw.PREC_BF16)
    assert e.value.status == nw.ERR_UNSUPPORTED
    lin = make_case(nw, orc, width=8, height=8, k=8, face_size=8, softplus=False, seed=4)
    lin_scene = ctx.scene_create(lin.desc, lin.params)
    with pytest.raises(nw.NwprtError) as e:
        ctx.batch_create([jobs[0], (lin_scene, envs[1], frames[1])], nw.PREC_BF16)
    assert e.value.status == nw.ERR_UNSUPPORTED
    with pytest.raises(nw.NwprtError):
        ctx.batch_create([], nw.PREC_BF16)
