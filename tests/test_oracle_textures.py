"""CPU checks that pin the oracle's texture sampling (orc_tex_sample, the
reference's tex2D on 8-bit textures: Renderer/textures.cpp:8-47 with wrap
addressing, normalized coordinates, bilinear filtering, normalized-float
reads and the sRGB decode) by what the definitions fix:

- texel centres ((i + 0.5) / w, (j + 0.5) / h) return the texel itself: the
  sRGB decode of each colour byte against published sRGB values and the
  IEC 61966-2-1 formula, alpha as b / 255;
- bilinear filtering reproduces a function linear in x exactly between
  texel centres (the alpha channel carries a byte ramp);
- wrap: (u + 1, v - 2) samples like (u, v), and between the last and the
  first texel the sample mixes the two;
- a constant texture samples to its constant everywhere;
- the committed device table (srgb_table.hpp) equals a fresh run of its
  generator, whose entries are the decode rounded to FP32.
"""
import importlib.util
from pathlib import Path

import numpy as np

from oracle import binding as orc

CSRC = Path(__file__).resolve().parent.parent / "neural_wavelet_prt_demo_amd" / "csrc"


def _srgb(b):
    c = b / 255.0
    return c / 12.92 if c <= 0.04045 else ((c + 0.055) / 1.055) ** 2.4


def test_texel_centres_return_the_decoded_texel():
    rng = np.random.default_rng(0)
    t = rng.integers(0, 256, (5, 7, 4), dtype=np.uint8)
    for j in range(5):
        for i in range(7):
            got = orc.tex_sample(t, (i + 0.5) / 7, (j + 0.5) / 5)
            want = [_srgb(int(t[j, i, c])) for c in range(3)] + [t[j, i, 3] / 255.0]
            np.testing.assert_allclose(got, want, rtol=0, atol=1e-15)
    # Published sRGB decode values (IEC 61966-2-1): 0 -> 0, 255 -> 1,
    # 128 -> 0.21586, 10 -> 10 / 255 / 12.92 (linear segment).
    t = np.array([[[0, 255, 128, 10]]], np.uint8)
    got = orc.tex_sample(t, 0.3, 0.9)
    assert got[0] == 0.0 and abs(got[1] - 1.0) < 1e-15
    assert abs(got[2] - 0.21586050011389926) < 1e-12
    assert abs(got[3] - 10 / 255.0) < 1e-15  # alpha is not decoded


def test_bilinear_reproduces_a_linear_ramp_and_wraps():
    w, h = 9, 4
    t = np.zeros((h, w, 4), np.uint8)
    t[..., 3] = (20 * np.arange(w))[None, :]  # alpha ramp in x: 20 i / 255
    rng = np.random.default_rng(1)
    for _ in range(50):
        x = rng.uniform(0.5, w - 0.5)          # between the first and last centres
        u = x / w
        v = rng.uniform(-3, 3)
        got = orc.tex_sample(t, u, v)[3]
        assert abs(got - 20 * (x - 0.5) / 255.0) < 1e-12
        assert abs(orc.tex_sample(t, u + 1.0, v - 2.0)[3] - got) < 1e-12
    # Across the seam: halfway between the last centre and the first.
    got = orc.tex_sample(t, 1.0, 0.5)[3]
    assert abs(got - 0.5 * (20 * (w - 1) + 0) / 255.0) < 1e-12
    got = orc.tex_sample(t, 0.0, 0.5)[3]
    assert abs(got - 0.5 * (20 * (w - 1)) / 255.0) < 1e-12
    # Constant texture.
    c = np.full((3, 6, 4), [30, 140, 250, 77], np.uint8)
    for u, v in rng.uniform(-5, 5, (20, 2)):
        np.testing.assert_allclose(orc.tex_sample(c, u, v),
                                   [_srgb(30), _srgb(140), _srgb(250), 77 / 255.0],
                                   rtol=0, atol=1e-14)


def test_srgb_table_header_is_current():
    spec = importlib.util.spec_from_file_location("gen_srgb", CSRC / "gen_srgb_table.py")
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    assert (CSRC / "srgb_table.hpp").read_text() == gen.render()
    for b in (0, 1, 10, 11, 128, 200, 255):
        assert float(np.float32(_srgb(b))) == float(np.float32(gen.decode(b)))


def test_textured_cast_reads_the_texture_at_the_hit():
    """A camera facing a textured quad: every pixel's kd is the texture at
    the pixel's texture coordinate, found here in closed form (ray-plane
    intersection, then the quad's affine texture mapping)."""
    # Quad z = -2 spanning x, y in [-2, 2], texcoords u = (x + 2) / 2, v = (y + 2) / 4 + 0.25.
    verts = np.array([[-2, -2, -2], [2, -2, -2], [2, 2, -2], [-2, 2, -2]], np.float32)
    tc = np.array([[(x + 2) / 2, (y + 2) / 4 + 0.25] for x, y, _ in verts], np.float32)
    ix = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    mats = [[0.1, 0.2, 0.3, 0.0, 0.0, 0.0, 0.5, 2]]
    rng = np.random.default_rng(3)
    tex = rng.integers(0, 256, (4, 6, 4), dtype=np.uint8)
    mt = -np.ones((1, 14), np.int32)
    mt[0, 0] = 0
    cam = [0, 0, 0, 0, 0, -1, 0, 1, 0]
    w, h = 12, 9
    out = orc.gbuffer_cast(verts, ix, mats, cam, w, h, texcoords=tc, mat_tex=mt, textures=[tex])
    hits = np.nonzero(out["env_flag"] == 1)[0]
    assert len(hits) == w * h
    for p in hits[:: 7]:
        pos = out["position"][p, :3]  # (p - lower) / extent of the quad's bounds
        x, y = -2 + 4 * pos[0], -2 + 4 * pos[1]
        # the encoded position includes the reference's +0.001 offset
        x, y = x - 0.001, y - 0.001
        want = orc.tex_sample(tex, (x + 2) / 2, (y + 2) / 4 + 0.25)
        np.testing.assert_allclose(out["kd"][p, :3], want[:3], rtol=0, atol=2e-7)
