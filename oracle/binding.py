"""ctypes binding of oracle/liboracle.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg.  The product never imports this module.
"""
from __future__ import annotations

import ctypes
import subprocess
from ctypes import POINTER, Structure, c_double, c_float, c_int, c_uint32, c_void_p
from pathlib import Path
from typing import Dict, Optional

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"


class OrcScene(Structure):
    _fields_ = [("n_levels", c_int), ("n_features", c_int), ("scale", c_void_p),
                ("resolution", c_void_p), ("dense", c_void_p), ("size", c_void_p),
                ("offset", c_void_p), ("table", c_void_p), ("n_coef", c_int),
                ("wfeat", c_void_p), ("proj", c_void_p), ("layers", c_void_p),
                ("feature_dim", c_int), ("hidden", c_int), ("n_hidden_layers", c_int),
                ("softplus", c_int), ("aux", c_int), ("aux_pad", c_float), ("waux", c_void_p)]


class OrcGbuffer(Structure):
    _fields_ = [("position", c_void_p), ("normal", c_void_p), ("reflect", c_void_p),
                ("roughness", c_void_p), ("kd", c_void_p), ("ks", c_void_p),
                ("direct", c_void_p), ("env_flag", c_void_p)]


class OrcMesh(Structure):
    _fields_ = [("vertices", c_void_p), ("normals", c_void_p), ("indices", c_void_p),
                ("n_tris", ctypes.c_int64), ("material_ids", c_void_p), ("materials", c_void_p),
                ("texcoords", c_void_p), ("mat_tex", c_void_p), ("tex_texels", c_void_p),
                ("tex_offset", c_void_p), ("tex_w", c_void_p), ("tex_h", c_void_p)]


TEX_FIELDS = 14


def _with_textures(mesh, keep, texcoords=None, mat_tex=None, textures=None):
    """Fills the texture fields of an OrcMesh.  texcoords [n_vertices][2];
    mat_tex [n_materials][14] indices into `textures` or -1; textures: list
    of uint8 [h][w][4] arrays, row 0 at v = 0.  `keep` holds the arrays."""
    if texcoords is None or mat_tex is None or not textures:
        return mesh
    tc = np.ascontiguousarray(texcoords, np.float32).reshape(-1, 2)
    mt = np.ascontiguousarray(mat_tex, np.int32).reshape(-1, TEX_FIELDS)
    ts = [np.ascontiguousarray(t, np.uint8) for t in textures]
    texels = np.concatenate([t.ravel() for t in ts])
    off = np.cumsum([0] + [t.size for t in ts[:-1]]).astype(np.int64)
    tw = np.array([t.shape[1] for t in ts], np.int32)
    th = np.array([t.shape[0] for t in ts], np.int32)
    keep += [tc, mt, texels, off, tw, th]
    mesh.texcoords, mesh.mat_tex, mesh.tex_texels = _p(tc), _p(mt), _p(texels)
    mesh.tex_offset, mesh.tex_w, mesh.tex_h = _p(off), _p(tw), _p(th)
    return mesh


def tex_sample(texture: np.ndarray, u: float, v: float) -> np.ndarray:
    """orc_tex_sample of one uint8 [h][w][4] texture at (u, v)."""
    keep = []
    mesh = _with_textures(OrcMesh(), keep, np.zeros((1, 2)), -np.ones((1, TEX_FIELDS)), [texture])
    out = np.zeros(4)
    lib().orc_tex_sample(ctypes.byref(mesh), 0, float(u), float(v), _p(out))
    return out


def _gbuf(g):
    keep = {k: np.ascontiguousarray(v) for k, v in g.items()}
    return keep, OrcGbuffer(*[_p(keep.get(k)) for k in ("position", "normal", "reflect",
                                                         "roughness", "kd", "ks", "direct",
                                                         "env_flag")])


_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE), "liboracle.so"], check=True)
    return LIB_PATH


# Material rows: kd[3], ks[3], roughness, type, then the DisneyBSDF
# parameters and eta (ORC_MAT_FLOATS = 19 values); rows of 8 take
# Material::init's values for the rest (Renderer/Material.h:79-97).
MAT_FLOATS = 19
_MAT_TAIL = np.array([0.0, 0.0, 0.0, 0.5, 0.0, 0.0, 0.0, 0.5, 0.0, 1.0,
                      np.float32(1.49) / np.float32(1.000277)], np.float32)


def mat_rows(materials) -> np.ndarray:
    """Material rows (8 or 19 values each) as float32 [n][19]."""
    rows = []
    for m in np.atleast_2d(np.asarray(materials, np.float32)):
        m = m.reshape(-1)
        if m.size == 8:
            m = np.concatenate([m, _MAT_TAIL])
        assert m.size == MAT_FLOATS, m.size
        rows.append(m)
    return np.ascontiguousarray(np.stack(rows), np.float32)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        l = ctypes.CDLL(str(LIB_PATH))
        l.orc_hashgrid_levels.argtypes = [c_int, c_int, c_int, c_float] + [c_void_p] * 5
        l.orc_hash_encode.argtypes = [POINTER(OrcScene), c_void_p, c_void_p, c_void_p]
        l.orc_haar_face.argtypes = [c_void_p, c_int, c_int, c_void_p]
        l.orc_envmap_haar.argtypes = [c_void_p, c_int, c_void_p]
        l.orc_rotate_cubemap.argtypes = [c_void_p, c_int, c_double, c_void_p]
        l.orc_score.argtypes = [c_void_p, c_int, c_int]
        l.orc_score.restype = c_float
        l.orc_topk.argtypes = [c_void_p, c_int, c_int, c_void_p, c_void_p]
        l.orc_relight_pixels.argtypes = [POINTER(OrcScene), c_void_p, c_void_p, c_int,
                                         POINTER(OrcGbuffer), c_void_p, c_int, c_void_p, c_void_p,
                                         c_void_p, c_int]
        l.orc_tonemap.argtypes = [c_float, c_float, c_float]
        l.orc_tonemap_thresholds.argtypes = [c_void_p]
        l.orc_tonemap_thresholds.restype = None
        l.orc_denoise.argtypes = [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_double,
                                  c_double, c_double, c_double, c_void_p]
        l.orc_denoise.restype = None
        l.orc_gbuffer_cast.argtypes = [POINTER(OrcMesh), c_void_p, c_int, c_int, c_int, c_void_p,
                                       c_void_p, c_int] + [c_void_p] * 9
        l.orc_gbuffer_cast.restype = None
        l.orc_gbuffer_cast_direct.argtypes = [POINTER(OrcMesh), c_void_p, c_int, c_int, c_int,
                                              c_void_p, c_void_p, c_int, c_void_p, c_int, c_int,
                                              c_uint32] + [c_void_p] * 10
        l.orc_gbuffer_cast_direct.restype = None
        l.orc_env_table_doubles.argtypes = [c_int]
        l.orc_env_table_doubles.restype = ctypes.c_size_t
        l.orc_env_tables.argtypes = [c_void_p, c_int, c_void_p]
        l.orc_env_tables.restype = None
        l.orc_ct_eval.argtypes = [c_void_p] * 6
        l.orc_ct_eval.restype = None
        l.orc_ct_pdf.argtypes = [c_void_p] * 5
        l.orc_ct_pdf.restype = c_double
        l.orc_env_pdf.argtypes = [c_void_p, c_int, c_void_p]
        l.orc_env_pdf.restype = c_double
        l.orc_sh_eval.argtypes = [c_void_p, c_void_p]
        l.orc_sh_eval.restype = None
        l.orc_sh_render.argtypes = [POINTER(OrcMesh), c_void_p, c_int, c_int, c_void_p, c_void_p,
                                    c_int, c_int, c_uint32, c_void_p, c_int, c_void_p, c_void_p,
                                    c_void_p]
        l.orc_sh_render.restype = None
        l.orc_sh_transfer.argtypes = [POINTER(OrcMesh), c_void_p, c_int, c_int, c_int, c_int,
                                      c_int, c_uint32, c_void_p, ctypes.c_int64, c_void_p,
                                      c_void_p, c_void_p]
        l.orc_sh_transfer.restype = None
        l.orc_tex_sample.argtypes = [POINTER(OrcMesh), c_int, c_double, c_double, c_void_p]
        l.orc_tex_sample.restype = None
        for name in ("orc_bsdf_sample_many", "orc_bsdf_eval_many", "orc_bsdf_pdf_many"):
            getattr(l, name).restype = None
        l.orc_bsdf_sample_many.argtypes = [c_void_p] * 4 + [ctypes.c_int64] + [c_void_p] * 3
        l.orc_bsdf_eval_many.argtypes = [c_void_p] * 4 + [ctypes.c_int64] + [c_void_p] * 2
        l.orc_bsdf_pdf_many.argtypes = [c_void_p] * 4 + [ctypes.c_int64] + [c_void_p] * 2
        l.orc_rng_uniform.argtypes = [c_uint32, c_uint32, c_uint32]
        l.orc_rng_uniform.restype = c_float
        l.orc_tonemap.restype = c_uint32
        l.orc_aux_inputs.argtypes = [POINTER(OrcGbuffer), c_int, c_float, c_void_p]
        l.orc_cpu_relight_fp32.argtypes = [POINTER(OrcScene), c_void_p, c_void_p, c_int,
                                           POINTER(OrcGbuffer), c_int, c_int, c_void_p, c_int]
        l.orc_cpu_relight_fp32.restype = c_double
        _lib = l
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(c_void_p)


def hashgrid_levels(n_levels=16, log2_hashmap_size=19, base_resolution=16, per_level_scale=1.5):
    out = {"scale": np.zeros(n_levels, np.float32), "resolution": np.zeros(n_levels, np.uint32),
           "dense": np.zeros(n_levels, np.uint32), "size": np.zeros(n_levels, np.uint32),
           "offset": np.zeros(n_levels + 1, np.uint32)}
    lib().orc_hashgrid_levels(n_levels, log2_hashmap_size, base_resolution, per_level_scale,
                              _p(out["scale"]), _p(out["resolution"]), _p(out["dense"]),
                              _p(out["size"]), _p(out["offset"]))
    return out


class Scene:
    """Host scene for the oracle: keeps every array alive for the struct."""

    def __init__(self, levels: Dict[str, np.ndarray], n_features: int, table: np.ndarray,
                 wfeat: np.ndarray, proj: Optional[np.ndarray], layers: np.ndarray,
                 softplus: bool = True, hidden: int = 64, n_hidden_layers: int = 3,
                 waux: Optional[np.ndarray] = None, aux_pad: float = 1.0):
        self.keep = dict(levels)
        self.keep["table"] = np.ascontiguousarray(table, np.float32)
        self.keep["wfeat"] = np.ascontiguousarray(wfeat, np.float32)
        self.keep["proj"] = None if proj is None else np.ascontiguousarray(proj, np.float32)
        self.keep["layers"] = np.ascontiguousarray(layers, np.float32)
        self.keep["waux"] = None if waux is None else np.ascontiguousarray(waux, np.float32)
        feature_dim = self.keep["wfeat"].shape[0]
        k = self.keep
        self.s = OrcScene(len(levels["scale"]), n_features, _p(k["scale"]), _p(k["resolution"]),
                          _p(k["dense"]), _p(k["size"]), _p(k["offset"]), _p(k["table"]),
                          k["wfeat"].shape[1], _p(k["wfeat"]), _p(k["proj"]), _p(k["layers"]),
                          feature_dim, hidden, n_hidden_layers, int(softplus),
                          int(waux is not None), aux_pad, _p(k["waux"]))

    @classmethod
    def from_synth(cls, levels, desc, params):
        """From a product SceneDesc + synth_scene() arrays; FP16 tables are
        widened with numpy (exact), not with product code.  With aux inputs
        the last 64 x 16 weights are the second layer's aux columns."""
        tab = params["hash_table"]
        if tab.dtype == np.uint16:
            tab = tab.view(np.float16).astype(np.float32)
        layers = np.asarray(params["layers"], np.float32)
        waux = None
        if desc.aux_dims:
            waux, layers = layers[-64 * 16:].reshape(64, 16), layers[:-64 * 16]
        return cls(levels, desc.grid.n_features_per_level, tab, params["wavelet_features"],
                   params["projection"] if desc.has_projection else None, layers,
                   bool(desc.output_softplus), waux=waux, aux_pad=desc.aux_pad_value)

    def hash_encode(self, pos3: np.ndarray):
        L = self.s.n_levels
        idx = np.zeros((L, 8), np.uint32)
        feat = np.zeros(L * self.s.n_features, np.float64)
        p = np.ascontiguousarray(pos3, np.float32)
        lib().orc_hash_encode(ctypes.byref(self.s), _p(p), _p(idx), _p(feat))
        return idx, feat

    def relight(self, sel, light, gbuf, pixels=None, transport=False, threads=8):
        sel = np.ascontiguousarray(sel, np.int32)
        light = np.ascontiguousarray(light, np.float32)
        pix = None if pixels is None else np.ascontiguousarray(pixels, np.int32)
        n = len(gbuf["position"]) if pix is None else len(pix)
        ind = np.zeros((n, 3), np.float64)
        wd = np.zeros((n, 3), np.float64)
        tr = np.zeros((n, len(sel), 3), np.float64) if transport else None
        keep, g = _gbuf(gbuf)
        lib().orc_relight_pixels(ctypes.byref(self.s), _p(sel), _p(light), len(sel),
                                 ctypes.byref(g), _p(pix), n, _p(ind), _p(wd), _p(tr), threads)
        return (ind, wd, tr) if transport else (ind, wd)

    def cpu_relight_fp32(self, sel, light, gbuf, begin, count, threads):
        sel = np.ascontiguousarray(sel, np.int32)
        light = np.ascontiguousarray(light, np.float32)
        out = np.zeros((count, 3), np.float32)
        keep, g = _gbuf(gbuf)
        secs = lib().orc_cpu_relight_fp32(ctypes.byref(self.s), _p(sel), _p(light), len(sel),
                                          ctypes.byref(g), begin, count, _p(out), threads)
        return out, secs


def aux_inputs(gbuf, p: int, pad: float = 1.0) -> np.ndarray:
    keep, g = _gbuf(gbuf)
    out = np.zeros(16, np.float32)
    lib().orc_aux_inputs(ctypes.byref(g), p, pad, _p(out))
    return out


def envmap_haar(faces: np.ndarray) -> np.ndarray:
    faces = np.ascontiguousarray(faces, np.float32)
    coef = np.zeros_like(faces)
    lib().orc_envmap_haar(_p(faces), faces.shape[1], _p(coef))
    return coef


def rotate_cubemap(faces: np.ndarray, yaw: float) -> np.ndarray:
    faces = np.ascontiguousarray(faces, np.float32)
    out = np.zeros_like(faces)
    lib().orc_rotate_cubemap(_p(faces), faces.shape[1], float(yaw), _p(out))
    return out


def haar_face(plane: np.ndarray) -> np.ndarray:
    plane = np.ascontiguousarray(plane, np.float32)
    out = np.zeros_like(plane)
    lib().orc_haar_face(_p(plane), plane.shape[0], 1, _p(out))
    return out


def score(coef: np.ndarray, t: int) -> float:
    coef = np.ascontiguousarray(coef, np.float32)
    return lib().orc_score(_p(coef), t, coef.shape[1])


def topk(coef: np.ndarray, k: int):
    coef = np.ascontiguousarray(coef, np.float32)
    idx = np.zeros(k, np.int32)
    st = np.zeros((k, 3), np.float32)
    lib().orc_topk(_p(coef), coef.shape[1], k, _p(idx), _p(st))
    return idx, st


def tonemap(rgb: np.ndarray) -> np.ndarray:
    rgb = np.asarray(rgb, np.float32).reshape(-1, rgb.shape[-1])
    return np.array([lib().orc_tonemap(float(r[0]), float(r[1]), float(r[2])) for r in rgb],
                    np.uint32)


def tonemap_thresholds() -> np.ndarray:
    """The oracle's 255 FP32 byte thresholds T_1 .. T_255 of the tone map."""
    out = np.zeros(255, np.float32)
    lib().orc_tonemap_thresholds(_p(out))
    return out


def denoise(color: np.ndarray, normal: np.ndarray, albedo: np.ndarray, iterations: int = 5,
            sigma_color: float = 1.0, sigma_normal: float = 0.3, sigma_albedo: float = 0.1,
            blend: float = 0.0) -> np.ndarray:
    """Guided a-trous denoiser on float4 [h][w] images; returns float64 [h][w][4]."""
    c = np.ascontiguousarray(color, np.float32)
    n = np.ascontiguousarray(normal, np.float32)
    a = np.ascontiguousarray(albedo, np.float32)
    h, w = c.shape[:2]
    assert c.shape == (h, w, 4) and n.shape == c.shape and a.shape == c.shape
    out = np.empty((h, w, 4), np.float64)
    lib().orc_denoise(_p(c), _p(n), _p(a), w, h, iterations, sigma_color, sigma_normal,
                      sigma_albedo, blend, _p(out))
    return out


def gbuffer_cast(vertices: np.ndarray, indices: np.ndarray, materials, camera, width: int,
                 height: int, *, normals: Optional[np.ndarray] = None,
                 material_ids: Optional[np.ndarray] = None, max_bounce: int = 1,
                 bounds=None, env_faces: Optional[np.ndarray] = None,
                 light_faces: Optional[np.ndarray] = None, samples: int = 1,
                 seed: int = 0, texcoords=None, mat_tex=None,
                 textures=None) -> Dict[str, np.ndarray]:
    """First-hit G-buffer by FP64 brute force (textures: see _with_textures).  materials: rows (kd[3], ks[3],
    roughness, type); camera: (from, at, up); bounds: (lower, upper) or the
    vertices' own.  Returns float64 [P][4] position / normal / reflect / kd /
    ks, [P] roughness, uint8 env_flag, int32 prim (-1 none), float64 margin."""
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
    ix = np.ascontiguousarray(indices, np.int32).reshape(-1, 3)
    nr = None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
    mid = None if material_ids is None else np.ascontiguousarray(material_ids, np.int32)
    mats = mat_rows(materials)
    cam = np.ascontiguousarray(camera, np.float32).reshape(9)
    if bounds is None:
        bounds = np.concatenate([v.min(0), v.max(0)])
    b = np.ascontiguousarray(bounds, np.float32).reshape(6)
    env_n = 0
    if env_faces is not None:
        env_faces = np.ascontiguousarray(env_faces, np.float32)
        env_n = env_faces.shape[1]
    keep = []
    mesh = _with_textures(OrcMesh(_p(v), _p(nr), _p(ix), len(ix), _p(mid), _p(mats)), keep,
                          texcoords, mat_tex, textures)
    P = width * height
    out = {k: np.empty((P, 4), np.float64) for k in ("position", "normal", "reflect", "kd", "ks")}
    out["roughness"] = np.empty(P, np.float64)
    out["env_flag"] = np.empty(P, np.uint8)
    out["prim"] = np.empty(P, np.int32)
    out["margin"] = np.empty(P, np.float64)
    keys = ("position", "normal", "reflect", "kd", "ks", "roughness", "env_flag", "prim",
            "margin")
    if light_faces is None:
        lib().orc_gbuffer_cast(ctypes.byref(mesh), _p(cam), width, height, max_bounce, _p(b),
                               _p(env_faces), env_n, *[_p(out[k]) for k in keys])
        return out
    light = np.ascontiguousarray(light_faces, np.float32)
    out["direct"] = np.empty((P, 4), np.float64)
    lib().orc_gbuffer_cast_direct(ctypes.byref(mesh), _p(cam), width, height, max_bounce, _p(b),
                                  _p(env_faces), env_n, _p(light), light.shape[1], samples,
                                  seed & 0xFFFFFFFF, *[_p(out[k]) for k in keys + ("direct",)])
    return out


def env_tables(faces: np.ndarray) -> Dict[str, np.ndarray]:
    """Environment sampling tables of [6][n][n][3] faces, split by field."""
    f = np.ascontiguousarray(faces, np.float32)
    n = f.shape[1]
    t = np.empty(lib().orc_env_table_doubles(n), np.float64)
    lib().orc_env_tables(_p(f), n, _p(t))
    per = n * (n + 1) + n * n + (n + 1) + n
    out = {"face_pmf": t[0:6], "face_cdf": t[6:13], "face_lum": t[13:19], "raw": t}
    faces_ = t[19:].reshape(6, per)
    out["cdf_rows"] = faces_[:, :n * (n + 1)].reshape(6, n, n + 1)
    out["pdf_rows"] = faces_[:, n * (n + 1):n * (n + 1) + n * n].reshape(6, n, n)
    out["cdf_marginals"] = faces_[:, n * (2 * n + 1):n * (2 * n + 1) + n + 1]
    out["pdf_marginals"] = faces_[:, n * (2 * n + 1) + n + 1:]
    return out


def rng_uniform(seed: int, pixel: int, counter: int) -> float:
    return lib().orc_rng_uniform(seed & 0xFFFFFFFF, pixel, counter)


def _d3(x):
    return np.ascontiguousarray(x, np.float64).reshape(3)


def ct_eval(mat, ns, ng, l, v) -> np.ndarray:
    """cookTorrance BSDF value times the cosine, for mat = (kd, ks, roughness, type)."""
    m = mat_rows([mat])[0]
    a = [_d3(x) for x in (ns, ng, l, v)]
    out = np.empty(3, np.float64)
    lib().orc_ct_eval(_p(m), *[_p(x) for x in a], _p(out))
    return out


def ct_pdf(mat, ns, ng, l, v) -> float:
    m = mat_rows([mat])[0]
    a = [_d3(x) for x in (ns, ng, l, v)]
    return lib().orc_ct_pdf(_p(m), *[_p(x) for x in a])


def env_pdf(tables: Dict[str, np.ndarray], n: int, d) -> float:
    dd = _d3(d)
    return lib().orc_env_pdf(_p(tables["raw"]), n, _p(dd))


def sh_eval(w) -> np.ndarray:
    """pbrt's real SH basis (degree 8, 81 values) at unit direction w."""
    ww = _d3(w)
    out = np.empty(81, np.float64)
    lib().orc_sh_eval(_p(ww), _p(out))
    return out


def sh_render(vertices, indices, normals, materials, camera, width: int, height: int,
              transfer, light_sh, degree: int, *, material_ids=None, samples: int = 16,
              seed: int = 0, env_faces=None) -> Dict[str, np.ndarray]:
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
    ix = np.ascontiguousarray(indices, np.int32).reshape(-1, 3)
    nr = np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
    mid = None if material_ids is None else np.ascontiguousarray(material_ids, np.int32)
    mats = mat_rows(materials)
    cam = np.ascontiguousarray(camera, np.float32).reshape(9)
    t = np.ascontiguousarray(transfer, np.float32)
    ls = np.ascontiguousarray(light_sh, np.float32)
    env_n = 0
    if env_faces is not None:
        env_faces = np.ascontiguousarray(env_faces, np.float32)
        env_n = env_faces.shape[1]
    mesh = OrcMesh(_p(v), _p(nr), _p(ix), len(ix), _p(mid), _p(mats))
    P = width * height
    out = {"color": np.empty((P, 4), np.float64), "prim": np.empty(P, np.int32),
           "margin": np.empty(P, np.float64)}
    lib().orc_sh_render(ctypes.byref(mesh), _p(cam), width, height, _p(t), _p(ls), degree,
                        samples, seed & 0xFFFFFFFF, _p(env_faces), env_n, _p(out["color"]),
                        _p(out["prim"]), _p(out["margin"]))
    return out


def sh_transfer(vertices, indices, materials, basis, degree: int, *, normals=None,
                material_ids=None, samples: int = 4096, max_bounce: int = 12, rr_depth: int = 3,
                seed: int = 0, verts=None, texcoords=None, mat_tex=None,
                textures=None) -> Dict[str, np.ndarray]:
    """SH transfer matrices by FP64 replay (orc_sh_transfer; textures: see
    _with_textures).  basis: the
    nc = (degree + 1)^2 basis maps, float32 [nc][6][n][n][3]; verts: vertex
    indices (all by default).  Returns T float64 [len(verts)][nc * nc][3]
    (row k * nc + i), abs_sum (same shape: sums of |terms|) and margin
    float64 [len(verts)][nc]."""
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
    ix = np.ascontiguousarray(indices, np.int32).reshape(-1, 3)
    nr = None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
    mid = None if material_ids is None else np.ascontiguousarray(material_ids, np.int32)
    mats = mat_rows(materials)
    nc = (degree + 1) ** 2
    b = np.ascontiguousarray(basis, np.float32)
    assert b.ndim == 5 and b.shape[0] == nc and b.shape[1] == 6 and b.shape[4] == 3
    n = b.shape[2]
    vs = np.arange(len(v), dtype=np.int64) if verts is None else np.ascontiguousarray(verts, np.int64)
    keep = []
    mesh = _with_textures(OrcMesh(_p(v), _p(nr), _p(ix), len(ix), _p(mid), _p(mats)), keep,
                          texcoords, mat_tex, textures)
    out = {"T": np.empty((len(vs), nc * nc, 3), np.float64),
           "abs_sum": np.empty((len(vs), nc * nc, 3), np.float64),
           "margin": np.empty((len(vs), nc), np.float64)}
    lib().orc_sh_transfer(ctypes.byref(mesh), _p(b), n, degree, samples, max_bounce, rr_depth,
                          seed & 0xFFFFFFFF, _p(vs), len(vs), _p(out["T"]), _p(out["abs_sum"]),
                          _p(out["margin"]))
    return out


def _bsdf_args(mat, ns, ng, v):
    return (mat_rows([mat])[0], _d3(ns), _d3(ng), _d3(v))


def bsdf_sample_many(mat, ns, ng, v, u):
    """The material's sampler on uniforms u [n][3]: (directions [n][3], ok [n])."""
    m, a, b, c = _bsdf_args(mat, ns, ng, v)
    u = np.ascontiguousarray(u, np.float32).reshape(-1, 3)
    out = np.empty((len(u), 3), np.float64)
    ok = np.empty(len(u), np.int32)
    lib().orc_bsdf_sample_many(_p(m), _p(a), _p(b), _p(c), len(u), _p(u), _p(out), _p(ok))
    return out, ok.astype(bool)


def bsdf_eval_many(mat, ns, ng, v, l):
    m, a, b, c = _bsdf_args(mat, ns, ng, v)
    l = np.ascontiguousarray(l, np.float64).reshape(-1, 3)
    out = np.empty((len(l), 3), np.float64)
    lib().orc_bsdf_eval_many(_p(m), _p(a), _p(b), _p(c), len(l), _p(l), _p(out))
    return out


def bsdf_pdf_many(mat, ns, ng, v, l):
    m, a, b, c = _bsdf_args(mat, ns, ng, v)
    l = np.ascontiguousarray(l, np.float64).reshape(-1, 3)
    out = np.empty(len(l), np.float64)
    lib().orc_bsdf_pdf_many(_p(m), _p(a), _p(b), _p(c), len(l), _p(l), _p(out))
    return out
