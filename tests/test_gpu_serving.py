"""GPU tests of the serving path: rotated environment maps (bit-exact
against the oracle's rotation + Haar + top-K) and the request server
(continuous batching): every request's outputs are the same alone and in the
mix, the same as nwprt_relight of its inputs, bit for bit, and within the
precision's tolerance of the FP64 oracle.  Scene 2 has the reference
network's aux inputs."""
import math

import numpy as np
import pytest

from cases import make_case
from conftest import max_rel_err, psnr
from neural_wavelet_prt_demo_amd import DeviceBuffer

pytestmark = pytest.mark.gpu

N = 8
# (width, height, scene, env source, K, yaw): uneven sizes, three scenes,
# resident / rotated / fresh (device and host) maps.
REQUESTS = [(40, 30, 0, "resident", 17, None), (64, 48, 1, "rotated", 64, 0.7),
            (1, 1, 0, "fresh_dev", 5, None), (100, 7, 1, "fresh_host", 130, None),
            (33, 17, 0, "rotated", 300, -2.5), (128, 96, 1, "resident", 32, None),
            (70, 70, 0, "rotated", 200, 3.9), (45, 21, 2, "rotated", 50, 1.3),
            (20, 9, 2, "resident", 24, None)]
TOL = {"bf16": 2e-2, "fp16": 5e-3, "fp8": 0.1}
# FP8 on the aux scene (scene 2): its wavelet-independent aux term dominates
# the hidden activations of this random network, so FP8's quantization
# errors are alike across a pixel's K terms and do not average out (BF16
# loses the same ~30 dB there).
MIN_PSNR_FP8_AUX = 40.0


class _Setup:
    def __init__(self, nw, orc, ctx):
        self.nw, self.ctx, self.orc = nw, ctx, orc
        cs = [make_case(nw, orc, width=8, height=8, k=8, face_size=N, seed=200 + s, aux=s == 2)
              for s in (0, 1, 2)]
        self.oscenes = [c.oscene for c in cs]
        self.scenes = [ctx.scene_create(c.desc, c.params) for c in cs]
        self.base = [nw.synth_cubemap(50 + s, N) for s in (0, 1, 2)]
        self.dbase = [DeviceBuffer.from_array(b) for b in self.base]
        self.reqs = []
        for i, (w, h, s, src, k, yaw) in enumerate(REQUESTS):
            gb = nw.synth_gbuffer(300 + i, w, h, 0.15)
            fresh = nw.synth_cubemap(400 + i, N)
            r = {"w": w, "h": h, "scene": self.scenes[s], "src": src, "k": k, "yaw": yaw,
                 "gbuf": gb, "fresh": fresh, "dfresh": DeviceBuffer.from_array(fresh),
                 "env": ctx.envmap_from_faces(self.base[s], k) if src == "resident" else None,
                 "s": s}
            self.reqs.append(r)

    def frame(self, r):
        return self.nw.DeviceFrame(r["w"], r["h"], r["gbuf"])

    def submit(self, server, r, frame):
        ctx = self.ctx
        if r["src"] == "resident":
            return ctx.server_submit(server, r["scene"], frame, env=r["env"])
        if r["src"] == "rotated":
            return ctx.server_submit(server, r["scene"], frame, faces=self.dbase[r["s"]],
                                     yaw=r["yaw"], k=r["k"], face_size=N)
        faces = r["dfresh"] if r["src"] == "fresh_dev" else r["fresh"]
        return ctx.server_submit(server, r["scene"], frame, faces=faces, k=r["k"], face_size=N)

    def oracle_check(self, r, got, precision):
        """The request's radiance against the FP64 oracle (its own map's
        rotation, Haar and top-K in the oracle)."""
        orc = self.orc
        if r["src"] == "resident":
            idx, st = self.ctx.envmap_selection(r["env"])
        else:
            faces = (orc.rotate_cubemap(self.base[r["s"]], r["yaw"]) if r["src"] == "rotated"
                     else r["fresh"])
            idx, st = orc.topk(orc.envmap_haar(faces), r["k"])
        ref, _ = self.oscenes[r["s"]].relight(idx, st, r["gbuf"])
        out = got["indirect"][:, :3]
        e, p = max_rel_err(ref, out), psnr(ref, out)
        bar = MIN_PSNR_FP8_AUX if precision == "fp8" and r["s"] == 2 else 50.0
        assert e <= TOL[precision] and p >= bar, (r["w"], r["h"], r["s"], r["src"], e, p)
        return e, p

    def reference(self, r, variant, precision="bf16"):
        """nwprt_relight of the same inputs with the same tile variant."""
        ctx = self.ctx
        if r["src"] == "resident":
            env = r["env"]
        elif r["src"] == "rotated":
            env = ctx.envmap_from_faces(self.base[r["s"]], r["k"])
            ctx.envmap_update_rotated(env, self.dbase[r["s"]], r["yaw"])
        else:
            env = ctx.envmap_from_faces(r["fresh"], r["k"])
        f = self.frame(r)
        ctx.set_relight_variant(variant)
        ctx.relight(r["scene"], env, f, self.nw.PRECISIONS[precision])
        ctx.sync()
        ctx.set_relight_variant(-1)
        return f.results()


def _same(a, b, what):
    for key in ("indirect", "indirect_with_direct", "rgba8"):
        assert np.array_equal(a[key], b[key]), (what, key)


@pytest.mark.parametrize("precision", ["bf16", "fp8"])
def test_server_alone_equals_mix(nw, orc, ctx, precision):
    st = _Setup(nw, orc, ctx)
    variant = 1 if precision == "fp8" else 5  # the pipelined form of each
    prec = nw.PRECISIONS[precision]
    refs = [st.reference(r, variant, precision) for r in st.reqs]
    # Alone: one request per server lifetime.
    for i, r in enumerate(st.reqs):
        srv = ctx.server_create(max_active=4, variant=variant, precision=prec)
        f = st.frame(r)
        rid = st.submit(srv, r, f)
        done = ctx.server_poll(srv, wait=2)
        assert [d[0] for d in done] == [rid] and done[0][1] > 0
        _same(f.results(), refs[i], ("alone", i))
        ctx.server_destroy(srv)
    # Mix: all at once, small ticks so every tick holds items of several
    # requests and requests span many ticks.
    srv = ctx.server_create(max_active=16, tick_pairs=64 * 300 * 3, variant=variant,
                            precision=prec)
    frames = [st.frame(r) for r in st.reqs]
    ids = [st.submit(srv, r, f) for r, f in zip(st.reqs, frames)]
    done = ctx.server_poll(srv, wait=2)
    assert sorted(d[0] for d in done) == sorted(ids)
    stats = ctx.server_stats(srv)
    print("mix stats:", stats)
    assert stats["completed"] == len(ids) and stats["queued"] == 0 and stats["active"] == 0
    assert stats["ticks"] > len(ids) and stats["peak_active"] > 1 and stats["bytes_in_use"] == 0
    assert stats["failed"] == 0
    worst = (0.0, np.inf)
    for i, f in enumerate(frames):
        got = f.results()
        _same(got, refs[i], ("mix", i))
        e, p = st.oracle_check(st.reqs[i], got, precision)
        worst = (max(worst[0], e), min(worst[1], p))
    print(f"served {precision} vs oracle: worst max rel err {worst[0]:.3e}, PSNR {worst[1]:.1f} dB")
    ctx.server_destroy(srv)


def test_server_memory_budget_admission(nw, orc, ctx):
    """A budget that fits one request's state at a time: requests are
    admitted one by one as earlier ones complete, and results are unchanged."""
    st = _Setup(nw, orc, ctx)
    variant = 5
    # the largest request's block: features and aux inputs, tiles, strengths,
    # cubemap / coefficients / scores (+ one 1 KB score histogram) / indices,
    # each rounded up to 256 bytes
    big = max(r["w"] * r["h"] * 96 + r["k"] * 4096 + 6 * N * N * 28 + r["k"] * 16 + 8 * 256
              + 1024 for r in st.reqs)
    srv = ctx.server_create(budget_bytes=big, max_active=8, tick_pairs=64 * 300 * 2,
                            variant=variant)
    frames = [st.frame(r) for r in st.reqs]
    ids = []
    for r, f in zip(st.reqs, frames):
        ids.append(st.submit(srv, r, f))
    got = []
    while len(got) < len(ids):
        got += ctx.server_poll(srv, wait=1)
    assert [g[0] for g in got] == ids  # arrival order
    stats = ctx.server_stats(srv)
    print("budget stats:", stats, "budget", big)
    assert stats["peak_bytes"] <= big and stats["peak_active"] < len(ids)
    for i, f in enumerate(frames):
        _same(f.results(), st.reference(st.reqs[i], variant), ("budget", i))
    ctx.server_destroy(srv)


def test_server_errors(nw, orc, ctx):
    st = _Setup(nw, orc, ctx)
    r = st.reqs[1]
    with pytest.raises(nw.NwprtError) as e:
        ctx.server_create(precision=nw.PREC_FP32)
    assert e.value.status == nw.ERR_UNSUPPORTED
    srv = ctx.server_create(budget_bytes=1 << 16, max_active=2)
    with pytest.raises(nw.NwprtError) as e:  # needs more than the whole budget
        st.submit(srv, r, st.frame(r))
    assert e.value.status == nw.ERR_INVALID_ARGUMENT
    ctx.server_destroy(srv)
    srv = ctx.server_create(max_active=2)
    with pytest.raises(nw.NwprtError) as e:  # face size of another scene
        ctx.server_submit(srv, r["scene"], st.frame(r), faces=st.dbase[0], yaw=0.1, k=8,
                          face_size=16)
    assert e.value.status == nw.ERR_INVALID_ARGUMENT
    assert ctx.server_poll(srv, wait=2) == []
    ctx.server_destroy(srv)
    with pytest.raises(nw.NwprtError) as e:
        ctx.server_poll(srv)
    assert e.value.status == nw.ERR_NOT_FOUND


def test_server_holds_scene_and_map(nw, orc, ctx):
    """While a request that reads a scene or a resident map is queued or
    running, destroying them or updating the map in place is refused
    (NWPRT_ERR_STATE); once the request is reported, both are allowed."""
    st = _Setup(nw, orc, ctx)
    r0, r5 = st.reqs[0], st.reqs[5]          # resident maps, scenes 0 and 1
    env0 = r0["env"]
    # one slot: the second request stays queued until the first completes
    srv = ctx.server_create(max_active=1, variant=5)
    f0, f5 = st.frame(r0), st.frame(r5)
    ids = [st.submit(srv, r0, f0), st.submit(srv, r5, f5)]
    for what, call in (("scene 0", lambda: ctx.scene_destroy(r0["scene"])),
                       ("scene 1 (queued)", lambda: ctx.scene_destroy(r5["scene"])),
                       ("map", lambda: ctx.envmap_destroy(env0)),
                       ("map update", lambda: ctx.envmap_update_faces(env0, st.base[0])),
                       ("map rotation", lambda: ctx.envmap_update_rotated(env0, st.dbase[0], 0.2))):
        with pytest.raises(nw.NwprtError) as e:
            call()
        assert e.value.status == nw.ERR_STATE, what
    done = ctx.server_poll(srv, wait=2)
    assert sorted(d[0] for d in done) == sorted(ids) and all(d[1] > 0 for d in done)
    assert ctx.server_stats(srv)["failed"] == 0
    _same(f0.results(), st.reference(r0, 5), "held 0")
    _same(f5.results(), st.reference(r5, 5), "held 5")
    ctx.envmap_update_faces(env0, st.base[1])   # allowed again
    ctx.server_destroy(srv)
    # A server destroyed with requests still queued lets go of them too.
    srv = ctx.server_create(max_active=1, variant=5)
    st.submit(srv, r0, st.frame(r0))
    st.submit(srv, r5, st.frame(r5))
    ctx.server_destroy(srv)
    ctx.envmap_destroy(env0)
    ctx.scene_destroy(r5["scene"])


@pytest.mark.parametrize("n,k,yaw", [(8, 40, 0.3), (64, 300, -2.1), (128, 300, math.pi / 2),
                                     (32, 100, 5.0)])
def test_rotated_envmap_bit_exact(nw, orc, ctx, n, k, yaw):
    faces = nw.synth_cubemap(n + 1, n)
    dfaces = DeviceBuffer.from_array(faces)
    env = ctx.envmap_from_faces(nw.synth_cubemap(999, n), k)
    ctx.envmap_update_rotated(env, dfaces, yaw)
    ctx.sync()
    rot = orc.rotate_cubemap(faces, yaw)
    ref = orc.envmap_haar(rot)
    coef = ctx.envmap_coefficients(env)
    assert np.array_equal(coef.view(np.uint32), ref.view(np.uint32))
    idx, st = ctx.envmap_selection(env)
    ridx, rst = orc.topk(ref, k)
    assert np.array_equal(idx, ridx) and np.array_equal(st.view(np.uint32), rst.view(np.uint32))
    # A second rotation reuses the map's buffers and matches again.
    ctx.envmap_update_rotated(env, dfaces, -yaw)
    ctx.sync()
    ref2 = orc.envmap_haar(orc.rotate_cubemap(faces, -yaw))
    assert np.array_equal(ctx.envmap_coefficients(env).view(np.uint32), ref2.view(np.uint32))


def test_rotation_errors(nw, ctx):
    env = ctx.envmap_from_selection(np.arange(4, dtype=np.int32), np.ones((4, 3), np.float32), 8)
    d = DeviceBuffer.from_array(nw.synth_cubemap(1, 8))
    with pytest.raises(nw.NwprtError) as e:
        ctx.envmap_update_rotated(env, d, 0.5)
    assert e.value.status == nw.ERR_STATE
    env2 = ctx.envmap_from_faces(nw.synth_cubemap(2, 8), 8)
    with pytest.raises(nw.NwprtError) as e:
        ctx.envmap_update_rotated(env2, d, float("nan"))
    assert e.value.status == nw.ERR_INVALID_ARGUMENT
