"""Pins the oracle's network, reduction, aux encodings and tone map against
an independent numpy FP64 implementation and closed forms."""
import math

import numpy as np
import pytest

from cases import make_case


def numpy_relight(case, sel, light, pixels, aux=False):
    """The relight math written directly in numpy (FP64), with the hash-grid
    features taken from the oracle's encoder (pinned separately)."""
    d, prm = case.desc, case.params
    layers = np.asarray(prm["layers"], np.float64)
    H, F = 64, 32
    W1 = layers[:H * F].reshape(H, F)
    W2 = layers[H * F:H * F + H * H].reshape(H, H)
    W3 = layers[H * F + H * H:H * F + 2 * H * H].reshape(H, H)
    W4 = layers[H * F + 2 * H * H:H * F + 2 * H * H + 3 * H].reshape(3, H)
    Waux = layers[-H * 16:].reshape(H, 16) if aux else None
    P = np.asarray(prm["projection"], np.float64) if d.has_projection else np.eye(F)
    wf = np.asarray(prm["wavelet_features"], np.float64)
    g = case.gbuf
    out = np.zeros((len(pixels), 3))
    for q, p in enumerate(pixels):
        if g["env_flag"][p] == 0:
            continue
        _, h = case.oscene.hash_encode(g["position"][p, :3])
        hp = P @ h
        b2 = 0.0
        if aux:
            n, r, kd, ks = (g[k][p, :3].astype(np.float64) for k in ("normal", "reflect", "kd", "ks"))
            rough = math.log(25.0 * float(g["roughness"][p]) + 1.0) / math.log(26.0)
            ax = np.concatenate([n, r, [rough], 2 * kd - 1, 2 * ks - 1, [d.aux_pad_value] * 3])
            b2 = Waux @ ax
        z = hp[None, :] * wf[:, sel].T                       # [K, 32]
        y = np.maximum(z @ W1.T, 0)
        y = np.maximum(y @ W2.T + b2, 0)
        y = np.maximum(y @ W3.T, 0)
        o = y @ W4.T
        t = np.logaddexp(0, o) if d.output_softplus else o
        out[q] = np.maximum(light * t, 0).sum(axis=0)
    return out


@pytest.mark.parametrize("aux", [False, True])
def test_oracle_network_matches_numpy(nw, orc, aux):
    case = make_case(nw, orc, width=6, height=5, face_size=4, k=9, hash_std=1.0,
                     feature_std=1.0, background=0.2, log2_hashmap_size=10, aux=aux, seed=2)
    coef = orc.envmap_haar(case.faces)
    sel, light = orc.topk(coef, case.k)
    light = light.astype(np.float64)
    light[::3] *= -1  # negative terms must clamp per term
    pixels = np.arange(case.width * case.height)
    ind, wd = case.oscene.relight(sel, light.astype(np.float32), case.gbuf)
    ref = numpy_relight(case, sel, light, pixels, aux)
    # aux encodings go through FP32 in the oracle (as in the reference)
    np.testing.assert_allclose(ind, ref, rtol=1e-5 if aux else 1e-12, atol=1e-9)
    np.testing.assert_allclose(wd, ind + case.gbuf["direct"][:, :3], rtol=1e-12)
    assert np.all(ind[case.gbuf["env_flag"] == 0] == 0)


def test_zero_network_gives_ln2_times_clamped_light(nw, orc):
    case = make_case(nw, orc, width=4, height=4, face_size=4, k=6, log2_hashmap_size=10)
    layers = np.zeros_like(case.params["layers"])
    sc = orc.Scene(case.levels, 2, case.oscene.keep["table"], case.params["wavelet_features"],
                   case.params["projection"], layers)
    light = np.array([[1, -2, 3], [0.5, 0.5, -0.5], [2, 0, 1], [-1, -1, -1], [0, 0, 0],
                      [4, 1, 0.25]], np.float32)
    ind, _ = sc.relight(np.arange(6), light, case.gbuf)
    want = math.log(2) * np.maximum(light.astype(np.float64), 0).sum(axis=0)
    np.testing.assert_allclose(ind, np.broadcast_to(want, ind.shape), rtol=1e-14)


def test_aux_encoding_closed_forms(orc):
    g = {"position": np.zeros((3, 4), np.float32),
         "normal": np.array([[0.5, 1, 0, 1]] * 3, np.float32),
         "reflect": np.array([[0, 0.25, 1, 1]] * 3, np.float32),
         "roughness": np.array([0, 1, 0.5], np.float32),
         "kd": np.array([[0, 1, 0.5, 1]] * 3, np.float32),
         "ks": np.array([[1, 0, 0.25, 1]] * 3, np.float32),
         "direct": np.zeros((3, 4), np.float32), "env_flag": np.ones(3, np.uint8)}
    a0 = orc.aux_inputs(g, 0, 1.0)
    np.testing.assert_array_equal(a0[:6], [0.5, 1, 0, 0, 0.25, 1])
    assert a0[6] == 0.0
    np.testing.assert_array_equal(a0[7:13], [-1, 1, 0, 1, -1, -0.5])
    np.testing.assert_array_equal(a0[13:], [1, 1, 1])
    assert abs(orc.aux_inputs(g, 1, 1.0)[6] - 1.0) < 1e-6
    assert abs(orc.aux_inputs(g, 2, 0.0)[6] - math.log(13.5) / math.log(26)) < 1e-6
    assert np.all(orc.aux_inputs(g, 2, 0.0)[13:] == 0)


def test_tonemap_closed_forms(orc):
    tm = orc.tonemap(np.array([[0, 0, 0], [1, 1, 1], [1, 0, 0], [0, 0, 1], [1e9, 2, 3],
                               [-1, -5, 0]], np.float32))
    assert tm[0] == 0xFF000000          # black: 1.055 * 1e-6^(1/2.4) - 0.055 < 0 -> 0
    assert tm[1] == 0xFFFFFFFF          # white clamps to 1 -> 255
    assert tm[2] == 0xFF0000FF          # R in the low byte
    assert tm[3] == 0xFFFF0000          # B in bits 16..23
    assert tm[4] == 0xFFFFFFFF
    assert tm[5] == 0xFF000000          # negative -> NaN power -> clamped to 0
    x = np.float32(0.5)
    want = int(min(1.0, 1.055 * (0.5 + 1e-6) ** (1 / 2.4) - 0.055) * 255.9)
    assert orc.tonemap(np.array([[x, x, x]], np.float32))[0] & 0xFF == want


def _curve64(x):
    return 1.055 * math.pow(float(x) + 1e-6, 1.0 / 2.4) - 0.055


def test_tonemap_thresholds_pinned(orc):
    """Each oracle threshold T_b is the first FP32 value whose FP64 curve
    reaches b / 255.9: the curve at T_b does, at the float just below it
    does not (checked here with Python's own arithmetic)."""
    t = orc.tonemap_thresholds()
    assert np.all(np.diff(t.astype(np.float64)) > 0)
    for b, x in enumerate(t, start=1):
        below = np.nextafter(x, np.float32(-np.inf))
        assert _curve64(x) >= b / 255.9 > _curve64(below), b
    # the byte of each threshold is b, the float below it b - 1
    tm = orc.tonemap(np.stack([t, t, t], 1)) & 0xFF
    assert np.array_equal(tm, np.arange(1, 256))
    below = np.nextafter(t, np.float32(-np.inf))
    tm = orc.tonemap(np.stack([below, below, below], 1)) & 0xFF
    assert np.array_equal(tm, np.arange(0, 255))


def test_tonemap_matches_reference_fp32_formula(orc):
    """Away from the thresholds the byte equals the reference's FP32 formula
    (toneMap.cu:58-70) evaluated in numpy float32; a dropped 1e-6 offset, a
    wrong exponent or a 256 scale would move many bytes."""
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(0, 1.2, 20000), 10 ** rng.uniform(-5, 0.2, 20000)])
    x = x.astype(np.float32)
    f = np.float32
    y = f(1.055) * np.power(x + f(1e-6), f(1.0) / f(2.4)) - f(0.055)
    ref = (np.clip(y, f(0), f(1)) * f(255.9)).astype(np.uint32)
    got = orc.tonemap(np.stack([x, x, x], 1)) & 0xFF
    t = orc.tonemap_thresholds().astype(np.float64)
    near = np.min(np.abs(x[:, None].astype(np.float64) - t[None, :]) /
                  np.maximum(t[None, :], 1e-30), axis=1) < 1e-5
    assert np.array_equal(got[~near], ref[~near])
    assert np.all(np.abs(got[near].astype(int) - ref[near].astype(int)) <= 1)
    assert (~near).sum() > 39000


def test_tonemap_table_header_is_current():
    """The committed device table equals the oracle's thresholds and a fresh
    run of its generator."""
    import importlib.util
    from pathlib import Path
    root = Path(__file__).resolve().parents[1] / "neural_wavelet_prt_demo_amd" / "csrc"
    spec = importlib.util.spec_from_file_location("gen_tm", root / "gen_tonemap_table.py")
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    assert (root / "tonemap_table.hpp").read_text() == gen.render()


def test_tonemap_table_equals_oracle(orc):
    import re
    from pathlib import Path
    hdr = (Path(__file__).resolve().parents[1] / "neural_wavelet_prt_demo_amd" / "csrc" /
           "tonemap_table.hpp").read_text()
    vals = [float.fromhex(v) for v in re.findall(r"^\s+(0x[0-9a-fp.+-]+)f,", hdr, re.M)]
    assert len(vals) == 255
    assert np.array_equal(np.array(vals, np.float32), orc.tonemap_thresholds())


def test_golden_small_relight(nw, orc):
    """Regression pin: the committed reduced relight (generator inputs, Haar,
    top-K, FP64 radiance) is reproduced by this tree's generator and oracle."""
    from pathlib import Path
    import sys
    sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
    import make_golden as mg
    gold = np.load(Path(__file__).resolve().parent / "golden" / "relight_small.npz")
    d, params, gbuf, faces = mg.small_inputs()
    np.testing.assert_array_equal(gbuf["position"], gold["position"])
    np.testing.assert_array_equal(faces, gold["faces"])
    np.testing.assert_array_equal(params["layers"], gold["layers"])
    coef = orc.envmap_haar(faces)
    np.testing.assert_array_equal(coef, gold["coef"])
    idx, st = orc.topk(coef, mg.SMALL["k"])
    np.testing.assert_array_equal(idx, gold["topk_idx"])
    lv = orc.hashgrid_levels(16, mg.SMALL["log2_hashmap_size"], 16, 1.5)
    ind, wd = orc.Scene.from_synth(lv, d, params).relight(idx, st, gbuf, threads=1)
    np.testing.assert_allclose(ind, gold["indirect"], rtol=1e-13)
    np.testing.assert_allclose(wd, gold["with_direct"], rtol=1e-13)
