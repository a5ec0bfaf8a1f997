"""GPU tests of the G-buffer producer (nwprt_mesh_create + nwprt_gbuffer_cast,
the first-hit part of the reference's OptiX pass:
Renderer/device_functions/camera.cu:82-290 and devicePrograms.cu:47-311,
517-564) against the FP64 brute-force oracle (orc_gbuffer_cast).

Bar: on pixels whose oracle margin is >= 1e-4 (no triangle edge, depth tie,
t_min crossing or grazing hit within that distance, so FP32 rounding cannot
change the first hit), env_flag, kd / ks / roughness of hits and the hit or
miss class agree exactly; position, normal and reflect encodings agree
within 2e-4 of max(1, |value|) (FP32 against FP64 on [0, 1] encodings and
on the far-off miss positions; after a mirror
bounce the reflected ray's FP32 direction error grows with the distance,
hence the looser bound there); miss kd (bilinear environment lookup in
FP32) within 1e-5 relative.  At most 2 % of pixels may fall under the
margin (5 % in the dense 110K-triangle scene)."""
import numpy as np
import pytest

from cases import mesh_scene, uv_sphere

pytestmark = pytest.mark.gpu

MARGIN = 1e-4
TOL_ENC = 2e-4


def _download(frame):
    P = frame.width * frame.height
    out = {}
    for k in ("position", "normal", "reflect", "kd", "ks"):
        out[k] = frame.bufs[k].download(np.float32, (P, 4)).astype(np.float64)
    out["roughness"] = frame.bufs["roughness"].download(np.float32, (P,)).astype(np.float64)
    out["env_flag"] = frame.bufs["env_flag"].download(np.uint8, (P,))
    return out


def _cast(nw, ctx, s, w, h, max_bounce=1, faces=None):
    mesh = ctx.mesh_create(s["vertices"], s["indices"], s["materials"], normals=s.get("normals"),
                           material_ids=s.get("material_ids"))
    frame = nw.DeviceFrame(w, h)
    env = None
    if faces is not None:
        env = nw.DeviceBuffer.from_array(np.ascontiguousarray(faces, np.float32))
    ctx.gbuffer_cast(mesh, np.reshape(s["camera"], (3, 3)), frame, max_bounce=max_bounce,
                     bounds=s["bounds"], env_faces=env, env_n=0 if faces is None else faces.shape[1])
    ctx.sync()
    got = _download(frame)
    info = ctx.mesh_info(mesh)
    ctx.mesh_destroy(mesh)
    return got, info


def _oracle(orc, s, w, h, max_bounce=1, faces=None):
    return orc.gbuffer_cast(s["vertices"], s["indices"], s["materials"], s["camera"], w, h,
                            normals=s.get("normals"), material_ids=s.get("material_ids"),
                            max_bounce=max_bounce, bounds=s["bounds"], env_faces=faces)


def _compare(got, want, tol_enc=TOL_ENC, min_safe=0.98):
    safe = want["margin"] >= MARGIN
    assert safe.mean() >= min_safe, safe.mean()
    assert np.array_equal(got["env_flag"][safe], want["env_flag"][safe])
    hit = safe & (want["env_flag"] == 1)
    miss = safe & (want["env_flag"] == 0)
    assert np.array_equal(got["kd"][hit].astype(np.float32), want["kd"][hit].astype(np.float32))
    # ks is the material's own value, or DisneyBSDF's packed specular colour
    # (computed: FP32 against FP64)
    np.testing.assert_allclose(got["ks"][hit], want["ks"][hit], rtol=2e-6, atol=1e-7)
    assert np.array_equal(got["roughness"][safe].astype(np.float32),
                          want["roughness"][safe].astype(np.float32))
    for k in ("position", "normal", "reflect"):
        # Miss positions sit near -1e5 / extent: FP32 keeps them to ~1e-7
        # relative, hence the relative term.
        err = (np.abs(got[k][safe] - want[k][safe]) /
               np.maximum(1.0, np.abs(want[k][safe]))).max() if safe.any() else 0.0
        assert err <= tol_enc, (k, err)
    if miss.any():
        kd_w = want["kd"][miss, :3]
        assert np.all(np.abs(got["kd"][miss, :3] - kd_w) <= 1e-5 * (np.abs(kd_w) + 1e-6))
        assert np.array_equal(got["ks"][miss], want["ks"][miss])
    return hit.sum(), miss.sum()


@pytest.mark.parametrize("seed,w,h,bounce", [(0, 64, 48, 1), (1, 37, 29, 2), (2, 96, 54, 4),
                                              (3, 1, 1, 1), (4, 1, 50, 2), (5, 50, 1, 1)])
def test_cast_matches_oracle(nw, orc, ctx, seed, w, h, bounce):
    s = mesh_scene(seed)
    faces = np.random.default_rng(seed).lognormal(0, 1, (6, 8, 8, 3)).astype(np.float32)
    got, info = _cast(nw, ctx, s, w, h, bounce, faces)
    want = _oracle(orc, s, w, h, bounce, faces)
    _compare(got, want, TOL_ENC if bounce == 1 else 1e-3)
    assert 1 <= info["bvh_depth"] <= 24 and info["triangles"] == len(s["indices"])


def test_cast_geometric_normals_and_single_material(nw, orc, ctx):
    s = mesh_scene(11, mirror=False)
    s.pop("normals")
    s.pop("material_ids")
    got, _ = _cast(nw, ctx, s, 48, 40)
    want = _oracle(orc, s, 48, 40)
    nh, _ = _compare(got, want)
    assert nh > 300


def test_cast_large_mesh_deep_bvh(nw, orc, ctx):
    # ~110K triangles of a finely tessellated sphere plus a dense soup.
    v, n, ix = uv_sphere(1.2, 256, 192, [0.0, 0.0, -1.0])
    rng = np.random.default_rng(5)
    soup = 12000
    c = rng.uniform([-3, -2, -5], [3, 2, 1], (soup, 1, 3))
    t = (c + rng.normal(0, 0.05, (soup, 3, 3))).reshape(-1, 3).astype(np.float32)
    s = {"vertices": np.concatenate([v, t]),
         "indices": np.concatenate([ix, np.arange(3 * soup, dtype=np.int32).reshape(-1, 3) +
                                    len(v)]),
         "materials": np.array([[0.3, 0.4, 0.5, 0.1, 0.1, 0.1, 0.5, 0]], np.float32),
         "camera": [0.0, 0.0, 5.0, 0.0, 0.0, -1.0, 0.0, 1.0, 0.0]}
    s["bounds"] = np.concatenate([s["vertices"].min(0), s["vertices"].max(0)])
    got, info = _cast(nw, ctx, s, 40, 30)
    want = _oracle(orc, s, 40, 30)
    nh, _ = _compare(got, want, min_safe=0.95)  # 12K small triangles: edges everywhere
    assert nh > 500 and info["triangles"] == len(s["indices"]) and info["bvh_depth"] <= 24


@pytest.mark.parametrize("nt", [1, 5, 8, 9, 100])
def test_cast_small_and_degenerate_meshes(nw, orc, ctx, nt):
    # nt triangles sharing one centroid (no SAH split exists: median splits
    # by index), scaled copies of one triangle facing the camera.
    base = np.array([[-1, -1, 0], [1, -1, 0], [0, 1.5, 0]], np.float64)
    v = np.concatenate([base * (0.5 + 0.02 * i) + [0, 0, -3 - 0.01 * i] for i in range(nt)])
    s = {"vertices": v.astype(np.float32),
         "indices": np.arange(3 * nt, dtype=np.int32).reshape(-1, 3),
         "materials": np.array([[0.2, 0.3, 0.4, 0.0, 0.0, 0.0, 0.6, 0]], np.float32),
         "camera": [0.0, 0.0, 2.0, 0.0, 0.0, -1.0, 0.0, 1.0, 0.0],
         "bounds": np.array([-2, -2, -5, 2, 2, 1], np.float32)}
    got, _ = _cast(nw, ctx, s, 24, 18)
    want = _oracle(orc, s, 24, 18)
    _compare(got, want)


def test_cast_null_fields_are_skipped(nw, ctx):
    s = mesh_scene(3)
    mesh = ctx.mesh_create(s["vertices"], s["indices"], s["materials"], normals=s["normals"],
                           material_ids=s["material_ids"])
    frame = nw.DeviceFrame(16, 8)
    sentinel = np.full((16 * 8, 4), 7.0, np.float32)
    for k in ("normal", "kd", "direct"):
        frame.bufs[k].upload(sentinel)
    gb = frame.gbuffer()
    gb.normal = None
    gb.kd = None
    cam = nw.Camera((nw.c_float * 3)(*s["camera"][0:3]), (nw.c_float * 3)(*s["camera"][3:6]),
                    (nw.c_float * 3)(*s["camera"][6:9]))
    b = np.ascontiguousarray(s["bounds"], np.float32)
    params = nw.CastParams(1, b.ctypes.data, None, 0)
    nw._check(nw.lib().nwprt_gbuffer_cast(ctx.h, mesh, nw.ctypes.byref(cam), 16, 8,
                                          nw.ctypes.byref(params), nw.ctypes.byref(gb), None))
    ctx.sync()
    for k in ("normal", "kd", "direct"):  # skipped fields and `direct` keep their contents
        assert np.all(frame.bufs[k].download(np.float32, (16 * 8, 4)) == 7.0), k
    pos = frame.bufs["position"].download(np.float32, (16 * 8, 4))
    assert np.all(pos[:, 3] == 1.0)
    ctx.mesh_destroy(mesh)


def test_mesh_and_cast_errors(nw, ctx):
    s = mesh_scene(0)
    v, ix, m = s["vertices"], s["indices"], s["materials"]
    bad = [lambda: ctx.mesh_create(v, ix + len(v), m),                     # index out of range
           lambda: ctx.mesh_create(v, ix, []),                             # no materials
           lambda: ctx.mesh_create(v, ix, m, material_ids=np.full(len(ix), 9)),
           lambda: ctx.mesh_create(np.where(np.arange(v.size).reshape(v.shape) == 5, np.nan, v),
                                   ix, m)]
    for f in bad:
        with pytest.raises(nw.NwprtError) as e:
            f()
        assert e.value.status == nw.ERR_INVALID_ARGUMENT
    flat = np.array([[0, 0, -2], [1, 0, -2], [0, 1, -2]], np.float32)
    mesh = ctx.mesh_create(flat, [[0, 1, 2]], m[:1])
    frame = nw.DeviceFrame(8, 8)
    cam = [[0, 0, 0], [0, 0, -1], [0, 1, 0]]
    with pytest.raises(nw.NwprtError):        # flat mesh needs explicit bounds
        ctx.gbuffer_cast(mesh, cam, frame)
    with pytest.raises(nw.NwprtError):        # from == at
        ctx.gbuffer_cast(mesh, [[0, 0, 0], [0, 0, 0], [0, 1, 0]], frame, bounds=[-1] * 3 + [1] * 3)
    with pytest.raises(nw.NwprtError):        # up parallel to the view
        ctx.gbuffer_cast(mesh, [[0, 0, 0], [0, 1, 0], [0, 1, 0]], frame, bounds=[-1] * 3 + [1] * 3)
    with pytest.raises(nw.NwprtError):
        ctx.gbuffer_cast(mesh, cam, frame, max_bounce=0, bounds=[-1] * 3 + [1] * 3)
    ctx.gbuffer_cast(mesh, cam, frame, bounds=[-1, -1, -3, 1, 1, 0])
    ctx.sync()
    ctx.mesh_destroy(mesh)
    with pytest.raises(nw.NwprtError) as e:
        ctx.mesh_info(mesh)
    assert e.value.status == nw.ERR_NOT_FOUND


def test_cast_feeds_relight(nw, orc, ctx):
    """The producer's G-buffer drives the relighting path unchanged: relight
    on the cast frame matches the oracle relight of the same (downloaded)
    G-buffer."""
    from cases import make_case
    w, h = 48, 32
    case = make_case(nw, orc, width=w, height=h, k=64, seed=4)
    s = mesh_scene(4)
    mesh = ctx.mesh_create(s["vertices"], s["indices"], s["materials"], normals=s["normals"],
                           material_ids=s["material_ids"])
    frame = nw.DeviceFrame(w, h)
    frame.bufs["direct"].upload(case.gbuf["direct"])
    ctx.gbuffer_cast(mesh, np.reshape(s["camera"], (3, 3)), frame, bounds=s["bounds"])
    scene = ctx.scene_create(case.desc, case.params)
    env = ctx.envmap_from_faces(case.faces, case.k)
    ctx.relight(scene, env, frame, nw.PREC_FP32, nw.TONEMAP_FULL)
    ctx.sync()
    res = frame.results()
    g = _download(frame)
    P = w * h
    gb = {k: g[k].astype(np.float32) for k in ("position", "normal", "reflect", "kd", "ks")}
    gb["roughness"] = g["roughness"].astype(np.float32)
    gb["env_flag"] = g["env_flag"]
    gb["direct"] = case.gbuf["direct"]
    idx, st = ctx.envmap_selection(env)
    ind, _ = case.oscene.relight(idx, st, gb)
    assert 0 < g["env_flag"].sum() < P
    ref = ind
    got = res["indirect"][:, :3].astype(np.float64)
    den = np.maximum(np.abs(ref), 1e-3 * np.abs(ref).max())
    assert np.max(np.abs(got - ref) / den) <= 1e-4
    ctx.scene_destroy(scene)
    ctx.envmap_destroy(env)
    ctx.mesh_destroy(mesh)


# DisneyBSDF parameter sets (specular_transmission, metallic, subsurface,
# specular, specular_tint, anisotropic, sheen, sheen_tint, clearcoat,
# clearcoat_gloss, eta): a plastic with sheen and clearcoat, a tinted metal,
# a rough glass.
DISNEY = [(0.0, 0.0, 0.4, 0.5, 0.2, 0.0, 0.7, 0.5, 0.6, 0.3, 1.5),
          (0.0, 0.9, 0.0, 0.5, 0.5, 0.6, 0.0, 0.5, 0.0, 1.0, 1.45),
          (0.8, 0.0, 0.0, 0.5, 0.0, 0.2, 0.0, 0.5, 0.3, 0.8, 1.5)]


def _direct_scene(seed, mirror=True, kinds=None):
    """kinds: BSDF types for the three non-mirror materials (None keeps the
    scene's roughPlastic x 2 + cookTorrance); DisneyBSDF materials take the
    parameter sets of DISNEY in order."""
    s = mesh_scene(seed, mirror=mirror)
    mats = np.array([nw_material_row(m) for m in s["materials"]], np.float32)
    if kinds is not None:
        mats[:3, 7] = kinds
        for i, k in enumerate(kinds):
            if k == 1:
                mats[i, 8:] = DISNEY[i]
    s["materials"] = mats
    return s


def nw_material_row(m):
    import neural_wavelet_prt_demo_amd as nw
    return nw.material_row(m)


@pytest.mark.parametrize("seed,w,h,samples,bounce,kinds", [(0, 32, 24, 64, 1, (2, 2, 2)),
                                                            (1, 29, 17, 16, 2, (2, 2, 2)),
                                                            (2, 1, 1, 256, 1, (2, 2, 2)),
                                                            (3, 32, 24, 64, 1, None),
                                                            (4, 23, 19, 32, 2, None),
                                                            (5, 32, 24, 48, 1, (4, 0, 4)),
                                                            (6, 32, 24, 64, 1, (1, 1, 1)),
                                                            (7, 29, 17, 32, 2, (1, 0, 2))])
def test_direct_light_replays_oracle(nw, orc, ctx, seed, w, h, samples, bounce, kinds):
    """The direct term draws the same counter-based random numbers as the
    oracle's FP64 replay, so per pixel the two estimates differ only by
    rounding, except where a sample's shadow ray or table bin flips at a
    boundary (rare).  Bar: 98 % of pixels within 1e-3 of
    (|oracle| + 1e-2 max), every pixel within 0.25 max."""
    s = _direct_scene(seed, kinds=kinds)
    rng = np.random.default_rng(seed)
    light = rng.lognormal(0, 1, (6, 8, 8, 3)).astype(np.float32)
    light[2, 3, 4] = 200.0  # a sun on +y
    view = rng.lognormal(0, 1, (6, 4, 4, 3)).astype(np.float32)
    mesh = ctx.mesh_create(s["vertices"], s["indices"], s["materials"], normals=s["normals"],
                           material_ids=s["material_ids"])
    frame = nw.DeviceFrame(w, h)
    dl = nw.DeviceBuffer.from_array(light)
    dv = nw.DeviceBuffer.from_array(view)
    ctx.gbuffer_cast(mesh, np.reshape(s["camera"], (3, 3)), frame, max_bounce=bounce,
                     bounds=s["bounds"], env_faces=dv, env_n=4, light_faces=dl, light_n=8,
                     direct_samples=samples, seed=1234 + seed)
    ctx.sync()
    got = frame.bufs["direct"].download(np.float32, (w * h, 4)).astype(np.float64)
    want = orc.gbuffer_cast(s["vertices"], s["indices"], s["materials"], s["camera"], w, h,
                            normals=s["normals"], material_ids=s["material_ids"],
                            max_bounce=bounce, bounds=s["bounds"], env_faces=view,
                            light_faces=light, samples=samples, seed=1234 + seed)
    ref = want["direct"]
    flags = frame.bufs["env_flag"].download(np.uint8, (w * h,))
    safe = want["margin"] >= MARGIN
    assert np.array_equal(flags[safe], want["env_flag"][safe])
    assert np.all(got[:, 3] == 1.0)
    top = max(np.abs(ref[:, :3]).max(), 1e-6)
    err = np.abs(got[safe, :3] - ref[safe, :3]) / (np.abs(ref[safe, :3]) + 1e-2 * top)
    per_pixel = err.max(axis=1)
    assert np.mean(per_pixel <= 1e-3) >= 0.98, (np.mean(per_pixel <= 1e-3), per_pixel.max())
    assert np.abs(got[safe, :3] - ref[safe, :3]).max() <= 0.25 * top
    if w * h > 1:
        assert (ref[want["env_flag"] == 1, :3] > 0).any()  # lit hits exist
    ctx.mesh_destroy(mesh)


def test_direct_light_errors_and_skips(nw, ctx):
    s = _direct_scene(0, kinds=(1, 2, 2))
    frame = nw.DeviceFrame(8, 8)
    light = nw.DeviceBuffer.from_array(np.ones((6, 4, 4, 3), np.float32))
    for col, bad in ((18, 0.0), (18, -1.5), (9, np.nan)):  # eta <= 0, a NaN parameter
        mats = s["materials"].copy()
        mats[0, col] = bad
        with pytest.raises(nw.NwprtError) as e:
            ctx.mesh_create(s["vertices"], s["indices"], mats, normals=s["normals"],
                            material_ids=s["material_ids"])
        assert e.value.status == nw.ERR_INVALID_ARGUMENT
    mesh = ctx.mesh_create(s["vertices"], s["indices"], s["materials"], normals=s["normals"],
                           material_ids=s["material_ids"])
    for kw in ({"light_n": 0}, {"light_n": 200}, {"light_n": 4, "direct_samples": 0}):
        with pytest.raises(nw.NwprtError) as e:
            ctx.gbuffer_cast(mesh, np.reshape(s["camera"], (3, 3)), frame, bounds=s["bounds"],
                             light_faces=light, **kw)
        assert e.value.status == nw.ERR_INVALID_ARGUMENT
    ctx.mesh_destroy(mesh)


def test_full_size_properties(nw, ctx):
    """1080p over the 1.29 M-triangle bench scene, where a brute-force oracle
    is out of reach: size-independent properties.  Casts are bitwise
    deterministic; hit positions lie in the bounds (plus the 0.001 offset);
    normals and reflections are unit vectors and obey the reflection law
    against the pinhole ray; the direct term is finite, non-negative,
    repeatable for a seed and different for another; the SH render is
    finite and non-negative, on transfer matrices computed for every vertex."""
    import bench
    verts, norms, idx, mids, mats = bench.gbuffer_bench_mesh()
    mesh = ctx.mesh_create(verts, idx, mats, normals=norms, material_ids=mids)
    info = ctx.mesh_info(mesh)
    W, H = 1920, 1080
    cam = ([0.0, 2.5, 7.5], [0.0, -0.8, -2.0], [0.0, 1.0, 0.0])
    faces = nw.synth_cubemap(3, 32)
    df = nw.DeviceBuffer.from_array(faces)
    f1, f2 = nw.DeviceFrame(W, H), nw.DeviceFrame(W, H)
    for f, seed in ((f1, 5), (f2, 5)):
        ctx.gbuffer_cast(mesh, cam, f, env_faces=df, env_n=32, light_faces=df, light_n=32,
                         direct_samples=2, seed=seed)
    ctx.sync()
    P = W * H
    get = lambda f, k, c=4: f.bufs[k].download(np.float32, (P, c) if c > 1 else (P,))
    for k in ("position", "normal", "reflect", "kd", "direct"):
        assert np.array_equal(get(f1, k), get(f2, k)), k
    flag = f1.bufs["env_flag"].download(np.uint8, (P,)).astype(bool)
    assert 0.3 < flag.mean() < 0.9
    lo, hi = np.array(info["lower"]), np.array(info["upper"])
    pos = get(f1, "position")[flag, :3]
    slack = 0.0011 / (hi - lo)
    assert np.all(pos >= -1e-5) and np.all(pos <= 1 + slack + 1e-5)
    n = get(f1, "normal")[flag, :3] * 2 - 1
    r = get(f1, "reflect")[flag, :3] * 2 - 1
    assert np.allclose(np.linalg.norm(n, axis=1), 1, atol=1e-5)
    assert np.allclose(np.linalg.norm(r, axis=1), 1, atol=1e-5)
    # Reflection law: r = d - 2 (d . n) n with d the pixel's pinhole ray.
    iy, ix = np.divmod(np.arange(P)[flag], W)
    fwd = np.array(cam[1]) - np.array(cam[0])
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, cam[2])
    right = right / np.linalg.norm(right) * (W / H)
    up = np.cross(right, fwd)
    up /= np.linalg.norm(up)
    t = np.tan(0.5 * 0.8575560548920328)
    sx, sy = (ix - W / 2) / (0.5 * W / t), (iy - H / 2) / (0.5 * H / t)
    d = fwd + sx[:, None] * right + sy[:, None] * up
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    want = d - 2 * np.sum(d * n, axis=1, keepdims=True) * n
    assert np.abs(want - r).max() < 2e-4
    dl = get(f1, "direct")[:, :3]
    assert np.all(np.isfinite(dl)) and np.all(dl >= 0) and dl[flag].mean() > 0
    ctx.gbuffer_cast(mesh, cam, f2, env_faces=df, env_n=32, light_faces=df, light_n=32,
                     direct_samples=2, seed=6)
    ctx.sync()
    assert not np.array_equal(get(f2, "direct"), get(f1, "direct"))
    # SH-PRT at degree 2 over the same scene: transfer matrices of all
    # vertices from the 9 real-SH basis maps (16 paths per vertex), then the
    # render that reads them.
    nc = 9
    T = ctx.sh_compute_transfer(mesh, nw.sh_basis_faces(2, 16), 16, 2, samples=16, seed=1)
    ctx.sync()
    Th = T.download(np.float32, (len(verts), nc * nc, 3))
    assert np.all(np.isfinite(Th)) and np.count_nonzero(Th) > Th.size // 4
    L = np.zeros((nc, 3), np.float32)
    L[0] = 1.0
    out = nw.DeviceBuffer(P * 16)
    ctx.sh_render(mesh, cam, W, H, T, nw.DeviceBuffer.from_array(L),
                  2, out, brdf_samples=8, seed=1, env_faces=df, env_n=32)
    ctx.sync()
    img = out.download(np.float32, (P, 4))
    assert np.all(np.isfinite(img)) and np.all(img[:, :3] >= 0) and img[flag, :3].mean() > 0
    ctx.mesh_destroy(mesh)


@pytest.mark.parametrize("seed,bounce", [(8, 1), (9, 2)])
def test_cast_disney_first_hit_encodings(nw, orc, ctx, seed, bounce):
    """DisneyBSDF materials at the first hit: kd is the base colour, ks the
    packed specular colour (disneyBSDF.cu:10-25), as the oracle computes it."""
    s = _direct_scene(seed, kinds=(1, 1, 1))
    faces = np.random.default_rng(seed).lognormal(0, 1, (6, 8, 8, 3)).astype(np.float32)
    got, _ = _cast(nw, ctx, s, 48, 40, bounce, faces)
    want = _oracle(orc, s, 48, 40, bounce, faces)
    n_hit, _ = _compare(got, want, TOL_ENC if bounce == 1 else 1e-3)
    assert n_hit > 100


# ---------------------------------------------------------------- textures --
# Material textures (nwprt_texture_create / nwprt_mesh_set_textures; the
# reference's tex2D reads, devicePrograms.cu:66-71, 121-147) against the
# oracle's FP64 sampling.  Texture coordinates are random per vertex over
# [-1.5, 2.5] (so sampling wraps), the textures random bytes of odd sizes.
# Bar for textured fields: within 1e-3 (FP32 bilinear at FP32 barycentrics
# against FP64: on the 16-unit ground quad, whose random texture
# coordinates span 4 wraps, the barycentrics' rounding moves a sample by up
# to a few 1e-4; a wrong texel, channel or orientation errs by ~0.1).

def _texture_set(rng):
    return [rng.integers(0, 256, (6, 8, 4), dtype=np.uint8),     # 8 wide, 6 high
            np.array([[[200, 120, 40, 255]]], np.uint8),          # 1 x 1
            rng.integers(0, 256, (16, 16, 4), dtype=np.uint8),
            rng.integers(0, 256, (5, 3, 4), dtype=np.uint8),
            rng.integers(200, 256, (4, 4, 4), dtype=np.uint8)]   # eta of about 0.58-1


def _textured_scene(nw, seed, kinds):
    s = _direct_scene(seed, kinds=kinds)
    rng = np.random.default_rng(100 + seed)
    f = {n: i for i, n in enumerate(nw.TEX_FIELD_NAMES)}
    tc = rng.uniform(-1.5, 2.5, (len(s["vertices"]), 2)).astype(np.float32)
    mt = -np.ones((len(s["materials"]), nw.TEX_FIELDS), np.int32)
    mt[0, f["kd"]], mt[0, f["roughness"]] = 0, 3
    mt[1, f["ks"]], mt[1, f["kd"]] = 2, 1
    mt[2, f["kd"]], mt[2, f["metallic"]], mt[2, f["specular_tint"]] = 2, 3, 1
    mt[2, f["eta"]], mt[2, f["clearcoat"]], mt[2, f["sheen"]] = 4, 0, 3
    return s, tc, mt, _texture_set(rng)


def _textured_mesh(nw, ctx, s, tc, mt, texs):
    mesh = ctx.mesh_create(s["vertices"], s["indices"], s["materials"], normals=s["normals"],
                           material_ids=s["material_ids"])
    ids = [ctx.texture_create(t) for t in texs]
    ctx.mesh_set_textures(mesh, tc, np.where(mt >= 0, np.take(ids, np.maximum(mt, 0)), -1))
    return mesh, ids


@pytest.mark.parametrize("seed,w,h,kinds", [(20, 64, 48, None), (21, 41, 33, (1, 0, 1)),
                                            (22, 1, 1, (2, 2, 2))])
def test_cast_textured_matches_oracle(nw, orc, ctx, seed, w, h, kinds):
    s, tc, mt, texs = _textured_scene(nw, seed, kinds)
    mesh, ids = _textured_mesh(nw, ctx, s, tc, mt, texs)
    frame = nw.DeviceFrame(w, h)
    ctx.gbuffer_cast(mesh, np.reshape(s["camera"], (3, 3)), frame, bounds=s["bounds"])
    ctx.sync()
    got = _download(frame)
    want = orc.gbuffer_cast(s["vertices"], s["indices"], s["materials"], s["camera"], w, h,
                            normals=s["normals"], material_ids=s["material_ids"],
                            bounds=s["bounds"], texcoords=tc, mat_tex=mt, textures=texs)
    safe = want["margin"] >= MARGIN
    assert safe.mean() >= 0.98
    assert np.array_equal(got["env_flag"][safe], want["env_flag"][safe])
    hit = safe & (want["env_flag"] == 1)
    for k in ("kd", "ks"):
        err = np.abs(got[k][hit, :3] - want[k][hit, :3]).max() if hit.any() else 0.0
        assert err <= 1e-3, (k, err)
    assert np.abs(got["roughness"][hit] - want["roughness"][hit]).max(initial=0.0) <= 1e-3
    for k in ("position", "normal", "reflect"):
        err = (np.abs(got[k][safe] - want[k][safe]) /
               np.maximum(1.0, np.abs(want[k][safe]))).max() if safe.any() else 0.0
        assert err <= TOL_ENC, (k, err)
    if w * h > 100:  # the textures are in use: kd varies across the hits
        assert len(np.unique(np.round(got["kd"][hit, 0], 3))) > 10
    ctx.mesh_destroy(mesh)
    for t in ids:
        ctx.texture_destroy(t)


@pytest.mark.parametrize("seed,w,h,samples,kinds", [(30, 32, 24, 48, None),
                                                    (31, 29, 17, 32, (1, 0, 2))])
def test_direct_light_textured_replays_oracle(nw, orc, ctx, seed, w, h, samples, kinds):
    """The direct term with textured materials, same bar as the untextured
    replay (98 % of pixels within 1e-3 of |oracle| + 1e-2 max)."""
    s, tc, mt, texs = _textured_scene(nw, seed, kinds)
    mesh, ids = _textured_mesh(nw, ctx, s, tc, mt, texs)
    rng = np.random.default_rng(seed)
    light = rng.lognormal(0, 1, (6, 8, 8, 3)).astype(np.float32)
    light[2, 3, 4] = 200.0
    frame = nw.DeviceFrame(w, h)
    dl = nw.DeviceBuffer.from_array(light)
    ctx.gbuffer_cast(mesh, np.reshape(s["camera"], (3, 3)), frame, bounds=s["bounds"],
                     light_faces=dl, light_n=8, direct_samples=samples, seed=77 + seed)
    ctx.sync()
    got = frame.bufs["direct"].download(np.float32, (w * h, 4)).astype(np.float64)
    want = orc.gbuffer_cast(s["vertices"], s["indices"], s["materials"], s["camera"], w, h,
                            normals=s["normals"], material_ids=s["material_ids"],
                            bounds=s["bounds"], light_faces=light, samples=samples,
                            seed=77 + seed, texcoords=tc, mat_tex=mt, textures=texs)
    ref = want["direct"]
    safe = want["margin"] >= MARGIN
    top = max(np.abs(ref[:, :3]).max(), 1e-6)
    err = np.abs(got[safe, :3] - ref[safe, :3]) / (np.abs(ref[safe, :3]) + 1e-2 * top)
    per_pixel = err.max(axis=1)
    print(f"textured direct seed {seed}: within 1e-3 {np.mean(per_pixel <= 1e-3):.4f}, "
          f"worst {per_pixel.max():.2e}")
    assert np.mean(per_pixel <= 1e-3) >= 0.98
    assert np.abs(got[safe, :3] - ref[safe, :3]).max() <= 0.25 * top
    assert (ref[want["env_flag"] == 1, :3] > 0).any()
    ctx.mesh_destroy(mesh)
    for t in ids:
        ctx.texture_destroy(t)


def test_texture_lifetime_and_errors(nw, ctx):
    s, tc, mt, texs = _textured_scene(nw, 40, (2, 2, 2))
    mesh, ids = _textured_mesh(nw, ctx, s, tc, mt, texs)
    with pytest.raises(nw.NwprtError) as e:  # held by the mesh
        ctx.texture_destroy(ids[0])
    assert e.value.status == nw.ERR_STATE
    bad = np.where(mt >= 0, 987654, -1).astype(np.int32)
    with pytest.raises(nw.NwprtError) as e:
        ctx.mesh_set_textures(mesh, tc, bad)
    assert e.value.status == nw.ERR_NOT_FOUND
    with pytest.raises(nw.NwprtError) as e:  # the mesh keeps its textures
        ctx.texture_destroy(ids[0])
    assert e.value.status == nw.ERR_STATE
    frame = nw.DeviceFrame(8, 8)
    with pytest.raises(nw.NwprtError) as e:
        ctx.texture_create(np.zeros((0, 4, 4), np.uint8))
    assert e.value.status == nw.ERR_INVALID_ARGUMENT
    with pytest.raises(nw.NwprtError) as e:
        ctx.mesh_set_textures(mesh, np.full_like(tc, np.nan), np.where(mt >= 0, ids[0], -1))
    assert e.value.status == nw.ERR_INVALID_ARGUMENT
    ctx.mesh_set_textures(mesh, None, None)  # untextured again: the textures are free
    for t in ids:
        ctx.texture_destroy(t)
    ctx.gbuffer_cast(mesh, np.reshape(s["camera"], (3, 3)), frame, bounds=s["bounds"])
    ctx.sync()
    ctx.mesh_destroy(mesh)


def test_sh_render_rejects_textured_meshes(nw, ctx):
    s = mesh_scene(41, mirror=False)
    mesh = ctx.mesh_create(s["vertices"], s["indices"], s["materials"], normals=s["normals"],
                           material_ids=s["material_ids"])
    tid = ctx.texture_create(np.full((2, 2, 4), 128, np.uint8))
    mt = -np.ones((len(s["materials"]), nw.TEX_FIELDS), np.int32)
    mt[0, 0] = tid
    ctx.mesh_set_textures(mesh, np.zeros((len(s["vertices"]), 2), np.float32), mt)
    tm = nw.DeviceBuffer(len(s["vertices"]) * 9 * 3 * 4)
    sh = nw.DeviceBuffer.from_array(np.ones((9, 3), np.float32))
    out = nw.DeviceBuffer(8 * 8 * 16)
    with pytest.raises(nw.NwprtError) as e:
        ctx.sh_render(mesh, np.reshape(s["camera"], (3, 3)), 8, 8, tm, sh, 2, out, brdf_samples=4)
    assert e.value.status == nw.ERR_UNSUPPORTED
    ctx.mesh_destroy(mesh)
    ctx.texture_destroy(tid)
