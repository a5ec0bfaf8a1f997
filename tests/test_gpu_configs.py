"""GPU parity at the full sizes of the precision study, the batched-jobs and
the served-stream workloads, against the FP64 oracle, through the C ABI.

- Precision study: 1920 x 1080, 6 x 128 x 128 cubemap, K = 300, FP16 hash
  grid; the network in FP32, BF16 and FP8.  Bit-exact top-K selection and
  hash-grid entry indices; radiance on a pixel sample (random pixels plus
  the brightest ones) within the stated tolerance and at PSNR >= 50 dB for
  every form; FP8 and BF16 also agree with each other over the whole frame.
- Batched jobs: the 32 uneven jobs of workloads.batch_jobs() in one graph;
  every job equals its single relight bit for bit and a pixel sample per
  job matches the oracle.
- Served stream: 4 resident scenes, 1080p and 4K requests at K = 300 with
  yaw-rotated and fresh 6 x 128 x 128 maps; every request equals itself
  alone bit for bit, and a pixel sample per request matches the oracle.
Each test prints its per-form error and PSNR.
"""
import numpy as np
import pytest

from cases import make_case
from conftest import max_rel_err, psnr
from neural_wavelet_prt_demo_amd import DeviceBuffer

pytestmark = pytest.mark.gpu

# max relative error per pixel and channel (floor 1e-3 x max) vs the FP64 oracle
TOL = {"fp32": 1e-4, "fp16": 5e-3, "bf16": 2e-2, "fp8": 0.1}
MIN_PSNR_DB = 50.0
N_SAMPLE = 3000


def _sample(rng, P, n, extra=()):
    pix = rng.choice(P, min(n, P), replace=False)
    return np.unique(np.concatenate([pix, np.asarray(extra, np.int64)]))


@pytest.fixture(scope="module")
def c2(nw, orc, ctx):
    case = make_case(nw, orc, width=1920, height=1080, face_size=128, k=300,
                     hash_dtype=nw.DTYPE_F16, seed=2)
    scene = ctx.scene_create(case.desc, case.params)
    env = ctx.envmap_from_faces(case.faces, case.k)
    frame = nw.DeviceFrame(case.width, case.height, case.gbuf)
    out = {}
    for prec in ("fp32", "bf16", "fp8"):
        ctx.relight(scene, env, frame, nw.PRECISIONS[prec], nw.TONEMAP_FULL)
        ctx.sync()
        out[prec] = frame.results()
    yield case, scene, env, out
    ctx.envmap_destroy(env)
    ctx.scene_destroy(scene)


def test_precision_study_selection_and_hash_indices(nw, orc, ctx, c2):
    case, scene, env, _ = c2
    idx, st = ctx.envmap_selection(env)
    ridx, rst = orc.topk(orc.envmap_haar(case.faces), case.k)
    assert np.array_equal(idx, ridx)
    assert np.array_equal(st.view(np.uint32), rst.view(np.uint32))
    pix = _sample(np.random.default_rng(21), case.width * case.height, 1200)
    gidx, _ = ctx.debug_hash_encode(scene, case.gbuf["position"][pix])
    for j, p in enumerate(pix):
        ridx_h, _ = case.oscene.hash_encode(case.gbuf["position"][p, :3])
        assert np.array_equal(gidx[j], ridx_h), p


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp8"])
def test_precision_study_radiance(nw, orc, ctx, c2, prec):
    case, scene, env, out = c2
    got_all = out[prec]["indirect"][:, :3]
    lum = got_all.sum(1)
    brightest = np.argsort(lum)[-200:]
    pix = _sample(np.random.default_rng(22), case.width * case.height, N_SAMPLE, brightest)
    idx, st = ctx.envmap_selection(env)
    ref, wd = case.oscene.relight(idx, st, case.gbuf, pixels=pix)
    got = got_all[pix]
    err, p = max_rel_err(ref, got), psnr(ref, got)
    print(f"precision study {prec}: {len(pix)} pixels, max rel err {err:.3e}, PSNR {p:.1f} dB")
    assert err <= TOL[prec] and p >= MIN_PSNR_DB
    # RGBA8 of the sampled pixels: the tone map of the with-direct output.
    assert np.array_equal(out[prec]["rgba8"][pix], orc.tonemap(out[prec]["indirect_with_direct"][pix, :3]))


def test_precision_study_fp8_whole_frame(nw, ctx, c2):
    """No pixel of the frame is clipped or lost in FP8: every pixel agrees
    with the BF16 frame (itself within 2e-2 of the oracle on its sample)."""
    _, _, _, out = c2
    a, b = out["bf16"]["indirect"][:, :3], out["fp8"]["indirect"][:, :3]
    assert np.all(np.isfinite(b))
    err, p = max_rel_err(a, b), psnr(a, b)
    print(f"fp8 vs bf16, whole frame: max rel err {err:.3e}, PSNR {p:.1f} dB")
    assert err <= 0.1 and p >= MIN_PSNR_DB


# ------------------------------------------------------------ batched jobs --

@pytest.mark.parametrize("prec", ["bf16", "fp8"])
def test_batched_jobs_workload(nw, orc, ctx, prec):
    from neural_wavelet_prt_demo_amd import workloads
    specs = workloads.batch_jobs()
    assert len(specs) == 32
    desc = nw.config_scene_desc(64, nw.DTYPE_F16)
    params = [nw.synth_scene(desc, s) for s in workloads.BATCH_SCENE_SEEDS]
    scenes = [ctx.scene_create(desc, p) for p in params]
    lv = orc.hashgrid_levels()
    oscenes = [orc.Scene.from_synth(lv, desc, p) for p in params]
    jobs, frames, gbufs, faces_all = [], [], [], []
    for sp in specs:
        faces = nw.synth_cubemap(sp.seed, 64)
        gb = nw.synth_gbuffer(sp.seed, sp.width, sp.height, 0.0)
        env = ctx.envmap_from_faces(faces, sp.k)
        frame = nw.DeviceFrame(sp.width, sp.height, gb)
        jobs.append((scenes[sp.scene], env, frame))
        frames.append(frame)
        gbufs.append(gb)
        faces_all.append(faces)
    bid = ctx.batch_create(jobs, nw.PRECISIONS[prec])
    _, n_pairs, variant = ctx.batch_info(bid)
    assert n_pairs == sum(sp.width * sp.height * sp.k for sp in specs)
    ctx.batch_launch(bid)
    ctx.sync()
    got = [f.results() for f in frames]
    rng = np.random.default_rng(31)
    worst = (0.0, np.inf)
    ctx.set_relight_variant(variant)
    try:
        for j, (sp, (scene, env, frame)) in enumerate(zip(specs, jobs)):
            single = nw.DeviceFrame(sp.width, sp.height)
            single.bufs = frame.bufs
            ctx.relight(scene, env, single, nw.PRECISIONS[prec], nw.TONEMAP_FULL)
            ctx.sync()
            want = single.results()
            for key in ("indirect", "indirect_with_direct", "rgba8"):
                assert np.array_equal(got[j][key], want[key]), (j, key)
            idx, st = ctx.envmap_selection(env)
            ridx, _ = orc.topk(orc.envmap_haar(faces_all[j]), sp.k)
            assert np.array_equal(idx, ridx), j
            pix = _sample(rng, sp.width * sp.height, 300)
            ref, _ = oscenes[sp.scene].relight(idx, st, gbufs[j], pixels=pix)
            e, p = max_rel_err(ref, got[j]["indirect"][pix, :3]), psnr(ref, got[j]["indirect"][pix, :3])
            worst = (max(worst[0], e), min(worst[1], p))
            assert e <= TOL[prec] and p >= MIN_PSNR_DB, (j, e, p)
            ctx.envmap_destroy(env)
    finally:
        ctx.set_relight_variant(-1)
    print(f"batched jobs {prec}: 32 jobs, worst max rel err {worst[0]:.3e}, "
          f"worst PSNR {worst[1]:.1f} dB")
    ctx.batch_destroy(bid)
    for s in scenes:
        ctx.scene_destroy(s)


# ---------------------------------------------------------- served stream --

class _Outputs:
    """Output buffers of one request (float4 indirect, float4 with-direct,
    RGBA8)."""

    def __init__(self, nw, w, h):
        self.nw, self.P = nw, w * h
        self.out = [DeviceBuffer(self.P * 16), DeviceBuffer(self.P * 16), DeviceBuffer(self.P * 4)]

    def frame_out(self):
        return self.nw.FrameOut(*(b.ptr for b in self.out))

    def results(self):
        P = self.P
        return {"indirect": self.out[0].download(np.float32, (P, 4)),
                "indirect_with_direct": self.out[1].download(np.float32, (P, 4)),
                "rgba8": self.out[2].download(np.uint32, (P,))}


@pytest.mark.parametrize("prec", ["bf16", "fp8"])
def test_served_stream_workload(nw, orc, ctx, prec):
    """4 resident scenes; 1080p requests and one 4K request, K = 300, with
    rotated resident and fresh host cubemaps; alone == mix, bit for bit, and
    a pixel sample per request against the oracle."""
    from neural_wavelet_prt_demo_amd import workloads
    desc = nw.config_scene_desc(128, nw.DTYPE_F16)
    params = [nw.synth_scene(desc, s) for s in workloads.SERVED_SCENE_SEEDS]
    scenes = [ctx.scene_create(desc, p) for p in params]
    lv = orc.hashgrid_levels()
    oscenes = [orc.Scene.from_synth(lv, desc, p) for p in params]
    bases = [nw.synth_cubemap(s, 128) for s in workloads.SERVED_BASE_SEEDS]
    dbases = [DeviceBuffer.from_array(b) for b in bases]
    fresh = [nw.synth_cubemap(s, 128) for s in workloads.SERVED_FRESH_SEEDS[:2]]
    views = {"1080p": (1920, 1080), "4k": (3840, 2160)}
    gbufs = {v: nw.synth_gbuffer(5000 + i, w, h, 0.0) for i, (v, (w, h)) in enumerate(views.items())}
    # (view, scene, rotated?, yaw or fresh index)
    reqs = [("1080p", 0, True, 0.7), ("1080p", 1, False, 0), ("4k", 2, True, 4.1),
            ("1080p", 3, True, -2.3), ("1080p", 2, False, 1), ("1080p", 1, True, 3.0)]
    dframes = {v: nw.DeviceFrame(w, h, gbufs[v]) for v, (w, h) in views.items()}

    def outs():
        return [_Outputs(nw, *views[v]) for v, *_ in reqs]

    def submit(srv, i, of):
        v, s, rot, arg = reqs[i]
        fo = of.frame_out()
        if rot:
            return ctx.server_submit(srv, scenes[s], dframes[v], faces=dbases[s], yaw=arg,
                                     k=300, face_size=128, out=fo)
        return ctx.server_submit(srv, scenes[s], dframes[v], faces=fresh[arg], k=300,
                                 face_size=128, out=fo)

    prec_id = nw.PRECISIONS[prec]
    alone = outs()
    for i, of in enumerate(alone):
        srv = ctx.server_create(budget_bytes=200 << 30, max_active=8, precision=prec_id)
        submit(srv, i, of)
        done = ctx.server_poll(srv, wait=2)
        assert len(done) == 1 and done[0][1] > 0
        ctx.server_destroy(srv)
    mix = outs()
    srv = ctx.server_create(budget_bytes=200 << 30, max_active=8, precision=prec_id)
    ids = [submit(srv, i, of) for i, of in enumerate(mix)]
    done = ctx.server_poll(srv, wait=2)
    assert sorted(d[0] for d in done) == sorted(ids) and all(d[1] > 0 for d in done)
    st = ctx.server_stats(srv)
    print("served mix stats:", st)
    assert st["peak_active"] > 1 and st["failed"] == 0
    ctx.server_destroy(srv)
    rng = np.random.default_rng(41)
    worst = (0.0, np.inf)
    for i, (v, s, rot, arg) in enumerate(reqs):
        a, b = alone[i].results(), mix[i].results()
        for key in a:
            assert np.array_equal(a[key], b[key]), (i, key)
        faces = orc.rotate_cubemap(bases[s], arg) if rot else fresh[arg]
        idx, strength = orc.topk(orc.envmap_haar(faces), 300)
        w, h = views[v]
        pix = _sample(rng, w * h, 1000)
        ref, _ = oscenes[s].relight(idx, strength, gbufs[v], pixels=pix)
        got = b["indirect"][pix, :3]
        e, p = max_rel_err(ref, got), psnr(ref, got)
        worst = (max(worst[0], e), min(worst[1], p))
        assert e <= TOL[prec] and p >= MIN_PSNR_DB, (i, e, p)
    print(f"served stream {prec}: {len(reqs)} requests (1080p + 4K), worst max rel err "
          f"{worst[0]:.3e}, worst PSNR {worst[1]:.1f} dB")
    for sc in scenes:
        ctx.scene_destroy(sc)
