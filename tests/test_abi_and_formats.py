"""CPU tests of the C ABI surface and the on-disk formats.

The .npy and .hdr readers are held against numpy and against the
reference's own readers (Renderer/npy.h, Renderer/HDRloader.cpp) built
under oracle/_ref from the reference sources; those tests skip where that
checker is not built.
"""
import re
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
REF_BIN = ROOT / "oracle" / "_ref" / "ref_formats"
needs_ref = pytest.mark.skipif(not REF_BIN.exists(), reason="oracle/_ref checker not built")


def test_library_exports_every_declared_symbol(nw):
    header = (ROOT / "include" / "nwprt.h").read_text()
    declared = set(re.findall(r"\b(nwprt_[a-z0-9_]+)\s*\(", header))
    assert len(declared) >= 30
    lib = nw.lib()
    for name in sorted(declared):
        assert hasattr(lib, name), name
        assert name in nw._SIGNATURES, f"{name} missing from the Python binding"
    assert lib.nwprt_abi_version() == 1


def test_status_strings_and_last_error(nw):
    assert nw.lib().nwprt_status_string(nw.ERR_STATE) == b"invalid call order"
    with pytest.raises(nw.NwprtError) as e:
        nw.npy_read("/nonexistent/file.npy")
    assert e.value.status == nw.ERR_IO
    assert b"cannot open" in nw.lib().nwprt_last_error()


# ------------------------------------------------------------------- npy --

@pytest.mark.parametrize("dtype", ["<f4", "<f8", "<f2", "<i4", "<i8", "|u1"])
@pytest.mark.parametrize("order", ["C", "F"])
def test_npy_read_matches_numpy(nw, tmp_path, dtype, order):
    rng = np.random.default_rng(0)
    a = (rng.standard_normal((5, 7, 3)) * 50).astype(np.dtype(dtype))
    a = np.asfortranarray(a) if order == "F" else a
    p = tmp_path / "a.npy"
    np.save(p, a)
    vals, shape, fortran = nw.npy_read(p, np.float32)
    assert shape == (5, 7, 3) and fortran == (order == "F")
    flat = a.ravel(order="K" if order == "C" else "F")  # file order
    np.testing.assert_array_equal(vals, flat.astype(np.float64).astype(np.float32))
    if np.dtype(dtype).kind in "iu":
        ivals, _, _ = nw.npy_read(p, np.int32)
        np.testing.assert_array_equal(ivals, flat.astype(np.int32))


def test_npy_write_roundtrip(nw, tmp_path):
    a = np.arange(24, dtype=np.float32).reshape(2, 3, 4) / 7
    nw.npy_write(tmp_path / "f.npy", a)
    np.testing.assert_array_equal(np.load(tmp_path / "f.npy"), a)
    b = np.arange(-5, 5, dtype=np.int32)
    nw.npy_write(tmp_path / "i.npy", b)
    np.testing.assert_array_equal(np.load(tmp_path / "i.npy"), b)


def test_npy_malformed(nw, tmp_path):
    p = tmp_path / "bad.npy"
    p.write_bytes(b"NOTNUMPY" + b"\0" * 40)
    with pytest.raises(nw.NwprtError) as e:
        nw.npy_read(p)
    assert e.value.status == nw.ERR_FORMAT
    good = tmp_path / "g.npy"
    np.save(good, np.ones(100, np.float32))
    p.write_bytes(good.read_bytes()[:-17])  # truncated payload
    with pytest.raises(nw.NwprtError) as e:
        nw.npy_read(p)
    assert e.value.status == nw.ERR_FORMAT
    np.save(good, np.ones(3, np.float32).astype(">f4"))
    with pytest.raises(nw.NwprtError) as e:
        nw.npy_read(good)
    assert e.value.status == nw.ERR_UNSUPPORTED


def _ref_npy(path, mode):
    out = Path(str(path) + ".bin")
    subprocess.run([str(REF_BIN), mode, str(path), str(out)], check=True)
    b = out.read_bytes()
    nd, fo = struct.unpack("<ii", b[:8])
    shape = struct.unpack(f"<{nd}Q", b[8:8 + 8 * nd])
    data = np.frombuffer(b[8 + 8 * nd:], np.float32 if mode == "npyf" else np.int32)
    return data, tuple(shape), bool(fo)


@needs_ref
@pytest.mark.parametrize("order", ["C", "F"])
def test_npy_matches_reference_reader(nw, tmp_path, order):
    rng = np.random.default_rng(1)
    a = rng.standard_normal((64, 3)).astype(np.float32)  # a waveletstrength-like array
    a = np.asfortranarray(a) if order == "F" else a
    np.save(tmp_path / "s.npy", a)
    ref, rshape, rfo = _ref_npy(tmp_path / "s.npy", "npyf")
    got, shape, fo = nw.npy_read(tmp_path / "s.npy")
    assert shape == rshape and fo == rfo
    np.testing.assert_array_equal(got, ref)
    i = rng.integers(0, 24576, 64).astype(np.int32)  # a waveletcoeffs-like array
    np.save(tmp_path / "c.npy", i)
    ref_i, _, _ = _ref_npy(tmp_path / "c.npy", "npyi")
    got_i, _, _ = nw.npy_read(tmp_path / "c.npy", np.int32)
    np.testing.assert_array_equal(got_i, ref_i)


# ------------------------------------------------------------------- hdr --

def _rgbe_encode(img):
    """Ward's float -> RGBE (test-side, numpy)."""
    v = img.max(axis=-1)
    m, e = np.frexp(v)
    scale = np.where(v > 1e-32, m * 256.0 / np.where(v > 0, v, 1), 0)
    rgb = np.floor(img * scale[..., None]).astype(np.uint8)
    ex = np.where(v > 1e-32, e + 128, 0).astype(np.uint8)
    return np.concatenate([rgb, ex[..., None]], axis=-1)


def _write_hdr(path, rgbe, mode):
    """Writes RGBE scanlines flat, with new-style RLE, or with old-style runs."""
    h, w, _ = rgbe.shape
    out = bytearray(b"#?RADIANCE\n# test\nFORMAT=32-bit_rle_rgbe\n\n")
    out += f"-Y {h} +X {w}\n".encode()
    for y in range(h):
        row = rgbe[y]
        if mode == "flat":
            out += row.tobytes()
        elif mode == "rle":
            out += bytes([2, 2, w >> 8, w & 255])
            for c in range(4):
                comp = row[:, c]
                i = 0
                while i < w:
                    j = i
                    while j < w and j - i < 127 and comp[j] == comp[i]:
                        j += 1
                    if j - i >= 3:
                        out += bytes([128 + (j - i), comp[i]])
                        i = j
                    else:
                        k = i
                        while k < w and k - i < 128 and not (k + 2 < w and comp[k] == comp[k + 1] == comp[k + 2]):
                            k += 1
                        k = max(k, i + 1)
                        out += bytes([k - i]) + comp[i:k].tobytes()
                        i = k
        elif mode == "old":
            x = 0
            while x < w:
                out += row[x].tobytes()
                run = 1
                while x + run < w and run < 255 and np.array_equal(row[x + run], row[x]):
                    run += 1
                if run > 1:
                    out += bytes([1, 1, 1, run - 1])
                x += run
    Path(path).write_bytes(bytes(out))


def _decode_ref_formula(rgbe):
    e = rgbe[..., 3].astype(np.int32) - 128
    return (rgbe[..., :3].astype(np.float32) / np.float32(256.0)) * np.ldexp(
        np.float32(1.0), e)[..., None].astype(np.float32)


@pytest.mark.parametrize("mode", ["flat", "rle", "old"])
def test_hdr_read_decodes_like_reference_formula(nw, tmp_path, mode):
    rng = np.random.default_rng(2)
    img = rng.lognormal(0, 2, (12, 24, 3)).astype(np.float32)
    img[3, 5:15] = img[3, 5]  # runs
    img[0, 0] = 0
    rgbe = _rgbe_encode(img)
    _write_hdr(tmp_path / "x.hdr", rgbe, mode)
    got = nw.hdr_read(tmp_path / "x.hdr")
    np.testing.assert_array_equal(got, _decode_ref_formula(rgbe))


@needs_ref
@pytest.mark.parametrize("mode", ["flat", "rle"])
def test_hdr_matches_reference_loader(nw, tmp_path, mode):
    rng = np.random.default_rng(3)
    img = rng.lognormal(0, 2, (24, 32, 3)).astype(np.float32)
    rgbe = _rgbe_encode(img)
    _write_hdr(tmp_path / "x.hdr", rgbe, mode)
    out = tmp_path / "x.bin"
    subprocess.run([str(REF_BIN), "hdr", str(tmp_path / "x.hdr"), str(out)], check=True)
    b = out.read_bytes()
    w, h = struct.unpack("<ii", b[:8])
    cols = np.frombuffer(b[8:8 + w * h * 12], np.float32).reshape(h, w, 3)
    flipped = np.frombuffer(b[8 + w * h * 12:], np.float32).reshape(h, w, 3)
    got = nw.hdr_read(tmp_path / "x.hdr")
    np.testing.assert_array_equal(got, cols)
    np.testing.assert_array_equal(got[::-1], flipped)  # HDRLoaderResult::get row flip


def test_hdr_write_roundtrip(nw, tmp_path):
    rng = np.random.default_rng(4)
    img = rng.lognormal(0, 2, (9, 17, 3)).astype(np.float32)
    img[0, 0] = 0
    nw.hdr_write(tmp_path / "w.hdr", img)
    back = nw.hdr_read(tmp_path / "w.hdr")
    assert back[0, 0].max() == 0
    rel = np.abs(back - img) / np.maximum(img.max(axis=-1, keepdims=True), 1e-30)
    assert rel.max() < 1.0 / 128


def test_hdr_malformed(nw, tmp_path):
    p = tmp_path / "bad.hdr"
    p.write_bytes(b"#?RADIANCE\n\n-Y 4 +X 4\n" + b"\x01" * 10)
    with pytest.raises(nw.NwprtError) as e:
        nw.hdr_read(p)
    assert e.value.status == nw.ERR_FORMAT
    p.write_bytes(b"P6\n4 4\n255\n")
    with pytest.raises(nw.NwprtError) as e:
        nw.hdr_read(p)
    assert e.value.status == nw.ERR_FORMAT


def test_cross_to_faces_layout(nw):
    """Faces +X,-X,+Y,-Y,-Z,+Z cut from a 4N x 3N cross at (col, row) blocks
    (2,1) (0,1) (1,2) (1,0) (1,1) (3,1) counted from the bottom, rows
    bottom-up (Renderer/EnvironmentMap.cpp:112-126)."""
    n = 5
    cross = np.arange(3 * n * 4 * n * 3, dtype=np.float32).reshape(3 * n, 4 * n, 3)
    faces = nw.cross_to_faces(cross)
    bottom_up = cross[::-1]
    blocks = [(2, 1), (0, 1), (1, 2), (1, 0), (1, 1), (3, 1)]
    for f, (c, r) in enumerate(blocks):
        np.testing.assert_array_equal(faces[f], bottom_up[r * n:(r + 1) * n, c * n:(c + 1) * n])
    with pytest.raises(nw.NwprtError):
        nw.cross_to_faces(np.zeros((9, 10, 3), np.float32))


# -------------------------------------------------------------- generator --

def test_generator_distributions(nw):
    g = nw.synth_gbuffer(0, 128, 96, 0.25)
    P = 128 * 96
    pos = g["position"]
    assert pos.shape == (P, 4) and np.all(pos[:, :3] >= 0) and np.all(pos[:, :3] < 1)
    assert np.all(pos[:, 3] == 1)
    n = g["normal"][:, :3] * 2 - 1
    r = g["reflect"][:, :3] * 2 - 1
    np.testing.assert_allclose(np.linalg.norm(n, axis=1), 1, atol=1e-6)
    np.testing.assert_allclose(np.linalg.norm(r, axis=1), 1, atol=1e-6)
    assert np.all(np.sum(n * r, axis=1) >= -1e-6)  # reflection stays in the hemisphere
    assert abs(np.mean(n[:, 2])) < 0.03  # uniform on the sphere
    assert np.all(g["roughness"] >= 0.05) and np.all(g["roughness"] < 1)
    assert 0.22 < np.mean(g["env_flag"] == 0) < 0.28
    g2 = nw.synth_gbuffer(0, 128, 96, 0.25)
    for k in g:
        assert np.array_equal(g[k], g2[k])
    assert not np.array_equal(nw.synth_gbuffer(1, 128, 96)["position"], pos)


def test_cubemap_generator(nw):
    f = nw.synth_cubemap(3, 64)
    assert f.shape == (6, 64, 64, 3)
    assert np.sum(f == 1e3) == 27  # one 3 x 3 sun, all channels
    rest = np.log(f[f != 1e3])
    assert abs(rest.mean()) < 0.03 and abs(rest.std() - 1.5) < 0.03


def test_scene_generator(nw):
    d = nw.config_scene_desc(32, nw.DTYPE_F16)
    p = nw.synth_scene(d, 0)
    tab = p["hash_table"].view(np.float16).astype(np.float32)
    assert tab.shape == (6513496, 2) and abs(tab.std() - 0.1) < 1e-3
    assert abs(p["wavelet_features"].std() - 0.1) < 2e-3
    w1 = p["layers"][:64 * 32]
    assert abs(w1.std() - np.sqrt(2 / 32)) < 0.01
    assert abs(p["projection"].std() - np.sqrt(1 / 32)) < 0.01


def test_obj_reader(nw, tmp_path):
    """nwprt_obj_read keeps what the reference's loadOBJ keeps
    (Renderer/Model.cpp:73-146): polygons fanned, one vertex per distinct
    (position, normal) corner, cookTorrance materials from the MTL
    (roughness = (1 / Ns)^2), defaults for faces without a material."""
    (tmp_path / "m.mtl").write_text(
        "newmtl red\nKd 0.8 0.1 0.1\nKs 0.2 0.2 0.2\nNs 4\n"
        "newmtl flat\nKd 0.3 0.3 0.3\nNs 0\n")
    (tmp_path / "s.obj").write_text(
        "# test\nmtllib m.mtl\n"
        "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 0 0 1\n"
        "vn 0 0 1\nvn 0 1 0\nvt 0.25 0.75\n"
        "f 1 2 3\n"                      # before any usemtl: default material
        "usemtl red\nf 1//1 2//1 3//1 4//1\n"   # quad -> 2 triangles
        "usemtl flat\nf -5/1/2 -4/1/2 -1/1/2\n")  # negative indices, v/vt/vn
    m = nw.obj_read(tmp_path / "s.obj")
    assert m["indices"].shape == (4, 3)
    # Distinct (position, normal) corners: (1,-),(2,-),(3,-),(1,1),(2,1),(3,1),(4,1),(1,2),(2,2),(5,2)
    assert len(m["vertices"]) == 10 and m["normals"] is not None
    v, ix = m["vertices"], m["indices"]
    np.testing.assert_array_equal(v[ix[1]], [[0, 0, 0], [1, 0, 0], [1, 1, 0]])
    np.testing.assert_array_equal(v[ix[2]], [[0, 0, 0], [1, 1, 0], [0, 1, 0]])
    np.testing.assert_array_equal(v[ix[3]], [[0, 0, 0], [1, 0, 0], [0, 0, 1]])
    np.testing.assert_array_equal(m["normals"][ix[3]], [[0, 1, 0]] * 3)
    np.testing.assert_array_equal(m["normals"][ix[0]], [[0, 0, 0]] * 3)  # corners without vn
    mats, mid = m["materials"], m["material_ids"]
    assert list(mid) == [2, 0, 0, 1]
    np.testing.assert_allclose(mats[0, :8], [0.8, 0.1, 0.1, 0.2, 0.2, 0.2, 1 / 16, 2], rtol=1e-6)
    np.testing.assert_allclose(mats[1, :8], [0.3, 0.3, 0.3, 0, 0, 0, 0, 2])
    np.testing.assert_allclose(mats[2, :8], [0.5, 0.5, 0.5, 0.5, 0.5, 0.5, 1, 2])
    # the Disney parameters and eta keep Material::init's values
    for row in mats:
        np.testing.assert_allclose(row[8:], nw.MATERIAL_TAIL_DEFAULTS, rtol=1e-7)
    (tmp_path / "bad.obj").write_text("v 0 0 0\nf 1 2 3\n")
    with pytest.raises(nw.NwprtError) as e:
        nw.obj_read(tmp_path / "bad.obj")
    assert e.value.status == nw.ERR_FORMAT
    with pytest.raises(nw.NwprtError) as e:
        nw.obj_read(tmp_path / "missing.obj")
    assert e.value.status == nw.ERR_IO


def test_obj_read_textured(nw, tmp_path):
    """nwprt_obj_read_textured: texcoords and map_Kd / map_Ks / map_Ns file
    names (relative to the OBJ's folder; absolute names kept), and the
    reference's vertex order (Renderer/Model.cpp:73-146): per shape ("o" /
    "g"), one mesh per material in ascending id with faces without one first,
    each mesh numbering distinct (position, normal, texcoord) corners by first
    use; a corner with a normal or texcoord fills the mesh's earlier vertices
    that have none (addVertex, Model.cpp:38-70).  Polygons are cut by
    tinyobjloader's ear clipping, so a concave one does not fan from a
    reflex corner."""
    (tmp_path / "t.mtl").write_text(
        "newmtl a\nKd 1 0 0\nmap_Kd -s 1 1 1 wood.png\nmap_Ns /abs/r.png\n"
        "newmtl b\nKd 0 1 0\nmap_Ks spec.tga\n")
    (tmp_path / "t.obj").write_text(
        "mtllib t.mtl\n"
        "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 0.5 0.25 0\n"
        "vt 0 0\nvt 1 0\nvt 1 1\nvt 0 1\nvn 0 0 1\n"
        "o first\n"
        "usemtl b\nf 1/1/1 2/2/1 3/3/1\n"      # shape 0, material b
        "usemtl a\nf 1 3 4/4\n"                 # shape 0, material a: texcoord back-fill
        "usemtl b\nf 3/3/1 4/4/1 1/1/1\n"      # shape 0, material b again
        "o second\n"
        "f 2 3 4 1 5\n")                       # shape 1, material b (kept): concave pentagon
    m = nw.obj_read_textured(tmp_path / "t.obj")
    mats, ix, v, tc = m["materials"], m["indices"], m["vertices"], m["texcoords"]
    assert len(mats) == 2 and m["maps"] == [[str(tmp_path / "wood.png"), "", "/abs/r.png"],
                                           ["", str(tmp_path / "spec.tga"), ""]]
    np.testing.assert_allclose(mats[0, :3], [1, 0, 0])
    # Mesh order: (shape 0, a), (shape 0, b), (shape 1, b).
    assert list(m["material_ids"]) == [0, 1, 1] + [1] * 3
    # Mesh (0, a): corners (1,-,-), (3,-,-), (4,t4,-): the texcoord fills all three.
    np.testing.assert_array_equal(ix[0], [0, 1, 2])
    np.testing.assert_array_equal(tc[:3], [[0, 1]] * 3)
    np.testing.assert_array_equal(m["normals"][:3], np.zeros((3, 3)))  # no vn in this mesh
    # Mesh (0, b): 1/1/1 2/2/1 3/3/1 then 3/3/1 4/4/1 1/1/1 -> vertices 3..6.
    np.testing.assert_array_equal(ix[1:3], [[3, 4, 5], [5, 6, 3]])
    np.testing.assert_array_equal(tc[3:7], [[0, 0], [1, 0], [1, 1], [0, 1]])
    np.testing.assert_array_equal(m["normals"][3:7], [[0, 0, 1]] * 4)
    # Mesh (1, b): the pentagon 2 3 4 1 5 has a reflex corner at 5.  Ear
    # clipping cuts (2,3,4); (2,4,1) holds corner 5, so (4,1,5); then (2,4,5).
    tris = [sorted(map(tuple, v[t])) for t in ix[3:]]
    want = [[(1, 0, 0), (1, 1, 0), (0, 1, 0)], [(0, 1, 0), (0, 0, 0), (0.5, 0.25, 0)],
            [(1, 0, 0), (0, 1, 0), (0.5, 0.25, 0)]]
    assert tris == [sorted(w) for w in want]
    assert len(v) == 7 + 5
    # Total area equals the pentagon's (shoelace), so no triangle overlaps.
    a = sum(0.5 * abs(np.cross(v[t[1]] - v[t[0]], v[t[2]] - v[t[0]])[2]) for t in ix[3:])
    poly = np.array([[1, 0], [1, 1], [0, 1], [0, 0], [0.5, 0.25]])
    x, y = poly[:, 0], poly[:, 1]
    assert a == pytest.approx(0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1))))
    # Texcoord index out of range is a format error (tinyobj only warns).
    (tmp_path / "bad.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1/1 2/1 3/1\n")
    with pytest.raises(nw.NwprtError) as e:
        nw.obj_read_textured(tmp_path / "bad.obj")
    assert e.value.status == nw.ERR_FORMAT


def test_obj_concave_ear_order(nw, tmp_path):
    """Ear clipping on a polygon whose first corner is reflex: the cut moves
    on to the next corner (tiny_obj_loader.h:1463-1566), in the projection on
    the axes the face normal favours (here x-z, a face in the y = 2 plane)."""
    (tmp_path / "c.obj").write_text(
        "v 0 2 0\nv 0.5 2 0.25\nv 1 2 0\nv 1 2 1\nv 0 2 1\nf 1 2 3 4 5\n")
    m = nw.obj_read(tmp_path / "c.obj")
    v, ix = m["vertices"], m["indices"]
    assert len(ix) == 3
    area = sum(0.5 * np.linalg.norm(np.cross(v[t[1]] - v[t[0]], v[t[2]] - v[t[0]])) for t in ix)
    # Unit square minus the notch; a fan from corner 1 would cover 1.125.
    assert area == pytest.approx(1.0 - 0.5 * 0.25)
    # Corner 2 is reflex, so the first ear is (2, 3, 4), then (2, 4, 5), (1, 2, 5).
    np.testing.assert_array_equal(v[ix[0]], [[0.5, 2, 0.25], [1, 2, 0], [1, 2, 1]])
    np.testing.assert_array_equal(v[ix[1]], [[0.5, 2, 0.25], [1, 2, 1], [0, 2, 1]])
    np.testing.assert_array_equal(v[ix[2]], [[0, 2, 0], [0.5, 2, 0.25], [0, 2, 1]])


def test_ply_reader(nw, tmp_path):
    """nwprt_ply_read keeps what loadPly keeps (Renderer/Model.cpp:152-220):
    x y z, optional nx ny nz, face vertex_indices; ASCII and binary
    little-endian, any scalar types, unknown properties and elements
    skipped, polygons fanned, one default cookTorrance material."""
    (tmp_path / "a.ply").write_text(
        "ply\nformat ascii 1.0\ncomment test\nelement vertex 5\nproperty float x\n"
        "property float y\nproperty float z\nproperty float nx\nproperty float ny\n"
        "property float nz\nelement face 2\nproperty list uchar int vertex_indices\nend_header\n"
        "0 0 0 0 0 1\n1 0 0 0 0 1\n1 1 0 0 0 1\n0 1 0 0 0 1\n0 0 1 1 0 0\n"
        "4 0 1 2 3\n3 0 1 4\n")
    m = nw.ply_read(tmp_path / "a.ply")
    assert m["vertices"].shape == (5, 3) and m["normals"].shape == (5, 3)
    np.testing.assert_array_equal(m["indices"], [[0, 1, 2], [0, 2, 3], [0, 1, 4]])
    np.testing.assert_array_equal(m["normals"][4], [1, 0, 0])
    np.testing.assert_allclose(m["materials"], [nw.material_row([0.5, 0.5, 0.5, 0.5, 0.5, 0.5, 1, 2])],
                               rtol=1e-7)
    assert list(m["material_ids"]) == [0, 0, 0]
    # Binary little-endian: doubles, an extra property and an extra element.
    hdr = ("ply\nformat binary_little_endian 1.0\nelement vertex 3\nproperty double x\n"
           "property double y\nproperty double z\nproperty uchar red\n"
           "element edge 1\nproperty int a\nproperty int b\n"
           "element face 1\nproperty list uchar uint vertex_indices\nend_header\n").encode()
    body = b"".join(struct.pack("<dddB", *v) for v in ((0, 0, 0, 7), (2, 0, 0, 8), (0, 3, 0, 9)))
    body += struct.pack("<ii", 0, 1) + struct.pack("<BIII", 3, 2, 1, 0)
    (tmp_path / "b.ply").write_bytes(hdr + body)
    m = nw.ply_read(tmp_path / "b.ply")
    assert m["normals"] is None
    np.testing.assert_array_equal(m["vertices"], [[0, 0, 0], [2, 0, 0], [0, 3, 0]])
    np.testing.assert_array_equal(m["indices"], [[2, 1, 0]])
    for name, text, status in (
            ("be.ply", "ply\nformat binary_big_endian 1.0\nelement vertex 0\nend_header\n",
             nw.ERR_FORMAT),
            ("oob.ply", "ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\n"
                        "property float y\nproperty float z\nelement face 1\n"
                        "property list uchar int vertex_indices\nend_header\n"
                        "0 0 0\n1 0 0\n0 1 0\n3 0 1 5\n", nw.ERR_FORMAT)):
        (tmp_path / name).write_text(text)
        with pytest.raises(nw.NwprtError) as e:
            nw.ply_read(tmp_path / name)
        assert e.value.status == status
    with pytest.raises(nw.NwprtError) as e:
        nw.ply_read(tmp_path / "missing.ply")
    assert e.value.status == nw.ERR_IO


def test_material_defaults_follow_the_reference_constructors(nw):
    """Material::init and the per-type constructors (Renderer/Material.h:79-155)."""
    tail = nw.MATERIAL_TAIL_DEFAULTS
    want = {0: [0.5, 0.5, 0.5, 1, 1, 1, 0.1], 1: [0.5, 0.5, 0.5, 0.25, 0.25, 0.25, 0.5],
            2: [0.5, 0.5, 0.5, 0.5, 0.5, 0.5, 1.0], 3: [0.25] * 6 + [0.5],
            4: [0.5, 0.5, 0.5, 0.25, 0.25, 0.25, 1.0]}
    for t, head in want.items():
        m = nw.material_defaults(t)
        np.testing.assert_allclose(m[:7], head, rtol=1e-7)
        assert m[7] == t
        np.testing.assert_allclose(m[8:], tail, rtol=1e-7)
    assert float(tail[-1]) == float(np.float32(1.49) / np.float32(1.000277))
    with pytest.raises(nw.NwprtError):
        nw.material_defaults(5)
