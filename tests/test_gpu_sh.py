"""GPU tests of the SH-PRT render (nwprt_sh_render: RenderEngine::shRender,
Renderer/RenderEngine.cpp:459-486, and __closesthit__SHradiance,
Renderer/device_functions/devicePrograms.cu:312-501) against the FP64
oracle's replay of the same random stream (orc_sh_render).

Bar: on pixels whose first hit is away from edges and ties (oracle margin
>= 1e-4), |gpu - oracle| <= 1e-3 (|oracle| + 1e-2 max): the two differ
only by FP32 rounding (no shadow rays, so no sample can flip)."""
import numpy as np
import pytest

from cases import mesh_scene

pytestmark = pytest.mark.gpu


def _scene(seed, sphere=(48, 24), soup=40):
    s = mesh_scene(seed, sphere=sphere, soup=soup, mirror=False)
    return s  # roughPlastic x 2 + cookTorrance


@pytest.mark.parametrize("seed,w,h,degree,samples,kinds", [(0, 32, 24, 2, 32, None),
                                                           (1, 29, 17, 0, 16, None),
                                                           (2, 24, 16, 8, 8, None),
                                                           (3, 16, 12, 4, 16, (4, 2, 0)),
                                                           (4, 24, 16, 2, 16, (1, 1, 2))])
def test_sh_render_replays_oracle(nw, orc, ctx, seed, w, h, degree, samples, kinds):
    s = _scene(seed, sphere=(16, 8) if degree == 8 else (48, 24), soup=10 if degree == 8 else 40)
    if kinds is not None:
        # DisneyBSDF rows take a plastic-with-sheen and a tinted-metal set
        disney = [(0.0, 0.0, 0.4, 0.5, 0.2, 0.0, 0.7, 0.5, 0.6, 0.3, 1.5),
                  (0.0, 0.9, 0.0, 0.5, 0.5, 0.6, 0.0, 0.5, 0.0, 1.0, 1.45)]
        mats = np.array([nw.material_row(m) for m in s["materials"]], np.float32)
        mats[:3, 7] = kinds
        for i, k in enumerate(kinds):
            if k == 1:
                mats[i, 8:] = disney[i % 2]
        s["materials"] = mats
    rng = np.random.default_rng(seed)
    nc = (degree + 1) ** 2
    nv = len(s["vertices"])
    T = (rng.normal(0, 0.3, (nv, nc, nc, 3)) + np.eye(nc)[None, :, :, None]).astype(np.float32)
    L = rng.normal(0, 1, (nc, 3)).astype(np.float32)
    L[0] = [3.0, 2.5, 2.0]
    view = rng.lognormal(0, 1, (6, 4, 4, 3)).astype(np.float32)
    mesh = ctx.mesh_create(s["vertices"], s["indices"], s["materials"], normals=s["normals"],
                           material_ids=s["material_ids"])
    dT, dL, dv = (nw.DeviceBuffer.from_array(x) for x in (T, L, view))
    out = nw.DeviceBuffer(w * h * 16)
    ctx.sh_render(mesh, np.reshape(s["camera"], (3, 3)), w, h, dT, dL, degree, out,
                  brdf_samples=samples, seed=77 + seed, env_faces=dv, env_n=4)
    ctx.sync()
    got = out.download(np.float32, (w * h, 4)).astype(np.float64)
    want = orc.sh_render(s["vertices"], s["indices"], s["normals"], s["materials"], s["camera"],
                         w, h, T, L, degree, material_ids=s["material_ids"], samples=samples,
                         seed=77 + seed, env_faces=view)
    safe = want["margin"] >= 1e-4
    assert safe.mean() > 0.95
    ref = want["color"]
    top = np.abs(ref[:, :3]).max()
    err = np.abs(got[safe, :3] - ref[safe, :3]) / (np.abs(ref[safe, :3]) + 1e-2 * top)
    assert err.max() <= 1e-3, err.max()
    assert np.all(got[:, 3] == 1.0)
    assert (want["prim"] >= 0).sum() > 0.2 * w * h
    ctx.mesh_destroy(mesh)


def test_sh_render_errors(nw, ctx):
    s = mesh_scene(0)  # has a mirror
    T = nw.DeviceBuffer(len(s["vertices"]) * 9 * 9 * 12)
    L = nw.DeviceBuffer(9 * 12)
    out = nw.DeviceBuffer(8 * 8 * 16)
    cam = np.reshape(s["camera"], (3, 3))
    mesh = ctx.mesh_create(s["vertices"], s["indices"], s["materials"], normals=s["normals"],
                           material_ids=s["material_ids"])
    with pytest.raises(nw.NwprtError) as e:
        ctx.sh_render(mesh, cam, 8, 8, T, L, 2, out)
    assert e.value.status == nw.ERR_UNSUPPORTED
    ctx.mesh_destroy(mesh)
    s = mesh_scene(0, mirror=False)
    mesh = ctx.mesh_create(s["vertices"], s["indices"], s["materials"])  # no normals
    with pytest.raises(nw.NwprtError) as e:
        ctx.sh_render(mesh, cam, 8, 8, T, L, 2, out)
    assert e.value.status == nw.ERR_INVALID_ARGUMENT
    ctx.mesh_destroy(mesh)
    mesh = ctx.mesh_create(s["vertices"], s["indices"], s["materials"], normals=s["normals"])
    for kw in ({"degree": 9}, {"degree": -1}, {"brdf_samples": 0}):
        args = {"degree": 2, "brdf_samples": 4}
        args.update(kw)
        with pytest.raises(nw.NwprtError) as e:
            ctx.sh_render(mesh, cam, 8, 8, T, L, args["degree"], out,
                          brdf_samples=args["brdf_samples"])
        assert e.value.status == nw.ERR_INVALID_ARGUMENT
    ctx.mesh_destroy(mesh)
