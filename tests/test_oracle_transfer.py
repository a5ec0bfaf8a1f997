"""CPU checks that pin the oracle's SH transfer-matrix computation
(orc_sh_transfer), the FP64 restatement of RenderEngine::shComputeTMatrix
(Renderer/RenderEngine.cpp:410-446) and __raygen__ComputeTMatrix
(Renderer/device_functions/cameraShTmatrix.cu:87-178):

- its mean over vertices and samples against a quadrature of its
  expectation on the floor scene (transfer_scenes), which fixes the sphere
  sampling, the 4 pi / N scale, the SH filter and normalisation, the
  throughput f / pdf, the exit weight, the removal of the first ray's
  environment radiance and the [vertex][k * nc + i][3] layout;
- a mirror floor: every path either leaves on its first ray or after one
  mirror, so all radiance is removed and T is exactly zero;
- Russian roulette is unbiased: T with roulette after depth 3 matches T
  without it within the statistical error, on two facing panels where paths
  run past depth 3.
Plus the host logic of the basis-map folder loader (natural order)."""
import numpy as np
import pytest

from oracle import binding as orc
from transfer_scenes import (floor_basis, floor_expectation, floor_scene, panel_scene,
                             smooth_basis)


def _mean_sem(T):
    return T.mean(0), T.std(0, ddof=1) / np.sqrt(T.shape[0])


def test_oracle_mean_matches_quadrature_on_the_floor():
    deg = 2
    nc = (deg + 1) ** 2
    v, ix, mids, mats, pts = floor_scene()
    basis = floor_basis(nc)
    out = orc.sh_transfer(v, ix, mats, basis, deg, material_ids=mids, samples=16384, seed=7,
                          verts=pts)
    T = out["T"].reshape(len(pts), nc, nc, 3)
    est, sem = _mean_sem(T)
    want, pe_max = floor_expectation(orc, deg, basis)
    assert pe_max < 1e-4  # no light sample above the floor passes the cut-off
    err = np.abs(est - want)
    top = np.abs(want).max()
    assert np.all(err <= 5 * sem + 2e-3 * top), (err / (sem + 1e-12)).max()
    # the zonal rows carry the signal; the others are noise around zero
    zonal = [l * l + l for l in range(deg + 1)]
    assert np.all(np.abs(want[zonal]).max(axis=(1, 2)) > 0.05 * top)
    # T scales with each map's own colour (column i of row k)
    assert np.allclose(est[0, :, 0] / est[0, 0, 0], [1, 2, 3, 4, 5, 6, 7, 8, 9], rtol=1e-9)


def test_oracle_mirror_floor_removes_everything():
    deg = 1
    v, ix, mids, mats, pts = floor_scene(n_chips=4, mirror=True)
    out = orc.sh_transfer(v, ix, mats, floor_basis(4), deg, material_ids=mids, samples=2048,
                          seed=1, verts=pts)
    assert np.all(out["T"] == 0.0)


def test_oracle_roulette_is_unbiased():
    deg = 1
    nc = (deg + 1) ** 2
    v, ix, mids, mats, pts = panel_scene()
    basis = smooth_basis(nc, seed=2)
    kw = dict(material_ids=mids, samples=6000, max_bounce=12, verts=pts)
    a = [orc.sh_transfer(v, ix, mats, basis, deg, rr_depth=3, seed=s, **kw)["T"] for s in range(4)]
    b = [orc.sh_transfer(v, ix, mats, basis, deg, rr_depth=100, seed=s + 10, **kw)["T"]
         for s in range(4)]
    a, b = np.array(a), np.array(b)   # [seed][vertex][row][3]
    ma, sa = _mean_sem(a)
    mb, sb = _mean_sem(b)
    top = np.abs(mb).max()
    assert top > 0
    assert np.all(np.abs(ma - mb) <= 5 * np.sqrt(sa ** 2 + sb ** 2) + 2e-3 * top)


def test_natural_order_of_basis_files(nw, tmp_path):
    n = 4
    names = ["sh10.npy", "sh2.npy", "SH1.npy", "sh0.npy", "notes.txt", "sh3.npy"]
    order = {"sh0.npy": 0, "SH1.npy": 1, "sh2.npy": 2, "sh3.npy": 3, "sh10.npy": 10}
    for name in names:
        if name.endswith(".npy"):
            cross = np.full((3 * n, 4 * n, 3), float(order[name]), np.float32)
            nw.npy_write(str(tmp_path / name), cross)
        else:
            (tmp_path / name).write_text("x")
    b = nw.load_sh_basis_folder(str(tmp_path), 1)
    assert b.shape == (4, 6, n, n, 3)
    assert [float(b[i, 0, 0, 0, 0]) for i in range(4)] == [0.0, 1.0, 2.0, 3.0]
    with pytest.raises(nw.NwprtError):
        nw.load_sh_basis_folder(str(tmp_path), 2)


def test_host_sh_basis_matches_the_oracle(nw):
    rng = np.random.default_rng(0)
    d = rng.normal(size=(64, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d[0] = [0.0, 0.0, 1.0]
    d[1] = [0.0, 0.0, -1.0]
    y = nw.sh_eval_real(d, 8)
    want = np.array([orc.sh_eval(x) for x in d])
    assert np.abs(y - want).max() < 1e-12
    b = nw.sh_basis_faces(2, 4)
    dirs = nw.cube_texel_dirs(4)
    assert b.shape == (9, 6, 4, 4, 3)
    assert np.allclose(b[:, :, :, :, 1], np.moveaxis(nw.sh_eval_real(dirs, 2), -1, 0), atol=1e-6)
    # texel (f, y, x) looks along the direction its cube face puts there
    for f, axis, sign in ((0, 0, 1), (1, 0, -1), (2, 1, 1), (3, 1, -1), (4, 2, -1), (5, 2, 1)):
        assert np.all(np.argmax(np.abs(dirs[f]), -1) == axis)
        assert np.all(np.sign(dirs[f][..., axis]) == sign)
