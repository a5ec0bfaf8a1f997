"""Seeded relight cases shared by the parity tests.

Inputs come from the product's seeded generator (host arrays), so the GPU
path and the oracle read identical bits.
"""
from __future__ import annotations

from dataclasses import dataclass
import json
import math
from typing import Dict, Optional

import numpy as np


@dataclass
class Case:
    desc: object
    params: Dict[str, np.ndarray]
    gbuf: Dict[str, np.ndarray]
    faces: np.ndarray
    width: int
    height: int
    k: int
    oscene: object = None
    levels: Optional[Dict[str, np.ndarray]] = None


def make_case(nw, orc, *, seed=0, width=128, height=128, face_size=32, k=64,
              hash_dtype=None, hash_std=0.1, feature_std=0.1, background=0.0,
              projection=True, softplus=True, levels=16, features=2, log2_hashmap_size=19,
              aux=False) -> Case:
    if hash_dtype is None:
        hash_dtype = nw.DTYPE_F32
    desc = nw.config_scene_desc(face_size, hash_dtype, levels=levels, features=features,
                                log2_hashmap_size=log2_hashmap_size, projection=projection,
                                softplus=softplus, aux=aux)
    params = nw.synth_scene(desc, seed, hash_std, feature_std)
    gbuf = nw.synth_gbuffer(seed, width, height, background)
    faces = nw.synth_cubemap(seed, face_size)
    lv = orc.hashgrid_levels(levels, log2_hashmap_size, 16, 1.5)
    oscene = orc.Scene.from_synth(lv, desc, params)
    return Case(desc, params, gbuf, faces, width, height, k, oscene, lv)


def _f16(a):
    return np.asarray(a, np.float32).astype(np.float16).astype(np.float64)


def _write_model_folder(path, rng, n_face, log2T=14):
    """A weight folder in the reference layout (tcnn.cu:52-74) with random
    parameters: out_prd 32 -> 64 -> 64 (1 hidden layer, linear output),
    cmp_enc Identity over the 13 aux inputs, cmp_net 80 -> 64 -> 64 -> 3
    (padded to 16 rows) with a softplus output."""
    cfg = {"vtx_enc": {"otype": "HashGrid", "n_levels": 16, "n_features_per_level": 2,
                       "log2_hashmap_size": log2T, "base_resolution": 16, "per_level_scale": 1.5},
           "out_prd": {"otype": "FullyFusedMLP", "activation": "ReLU", "output_activation": "None",
                       "n_neurons": 64, "n_hidden_layers": 1},
           "cmp_enc": {"otype": "Identity"},
           "cmp_net": {"otype": "FullyFusedMLP", "activation": "ReLU",
                       "output_activation": "Softplus", "n_neurons": 64, "n_hidden_layers": 2}}
    path.mkdir(parents=True, exist_ok=True)
    (path / "config_comp.json").write_text(json.dumps(cfg))
    import neural_wavelet_prt_demo_amd as nw
    lv = nw.hashgrid_levels(nw.HashgridDesc(16, 2, log2T, 16, 1.5))
    w = {"vtx": rng.normal(0, 0.5, int(lv["offset"][-1]) * 2).astype(np.float32),
         "o1": rng.normal(0, math.sqrt(2 / 32), (64, 32)).astype(np.float32),
         "o2": rng.normal(0, math.sqrt(1 / 64), (64, 64)).astype(np.float32),
         "c1": rng.normal(0, math.sqrt(2 / 80), (64, 80)).astype(np.float32),
         "c2": rng.normal(0, math.sqrt(2 / 64), (64, 64)).astype(np.float32),
         "c3": rng.normal(0, math.sqrt(2 / 64), (16, 64)).astype(np.float32),
         "wve": rng.normal(0, 0.5, (32, 6 * n_face * n_face)).astype(np.float32)}
    np.save(path / "vtx_enc.npy", w["vtx"])
    np.save(path / "out_prd.npy", np.concatenate([w["o1"].ravel(), w["o2"].ravel()]))
    np.save(path / "cmp_net.npy", np.concatenate([w["c1"].ravel(), w["c2"].ravel(), w["c3"].ravel()]))
    np.save(path / "wve_enc.npy", w["wve"])
    return w, lv


def _faces_to_cross(faces):
    n = faces.shape[1]
    cross = np.zeros((3 * n, 4 * n, 3), np.float32)
    bottom_up = cross[::-1]
    for f, (c, r) in enumerate([(2, 1), (0, 1), (1, 2), (1, 0), (1, 1), (3, 1)]):
        bottom_up[r * n:(r + 1) * n, c * n:(c + 1) * n] = faces[f]
    return cross


def uv_sphere(r, nu, nv, centre):
    """UV sphere: vertices, radial unit normals, CCW-outward triangles."""
    th = np.linspace(0, np.pi, nv + 1)
    ph = np.linspace(0, 2 * np.pi, nu, endpoint=False)
    T_, P_ = np.meshgrid(th, ph, indexing="ij")
    n = np.stack([np.sin(T_) * np.cos(P_), np.cos(T_), np.sin(T_) * np.sin(P_)], -1).reshape(-1, 3)
    idx = []
    for i in range(nv):
        for j in range(nu):
            a, b = i * nu + j, i * nu + (j + 1) % nu
            idx += [[a, a + nu, b], [b, a + nu, b + nu]]
    return (np.asarray(centre) + r * n).astype(np.float32), n.astype(np.float32), \
        np.array(idx, np.int32)


def mesh_scene(seed: int, *, sphere=(48, 24), soup=40, mirror=True):
    """A seeded test scene for the G-buffer producer: a sphere with vertex
    normals, a ground quad, a random triangle soup and (optionally) a tilted
    mirror, with four materials.  Returns a dict of host arrays plus camera
    and bounds; the camera sits at +z looking at the origin."""
    rng = np.random.default_rng(seed)
    sv, sn, si = uv_sphere(1.5, sphere[0], sphere[1], [0.3, 0.2, -1.0])
    verts, norms, idx, mids = [sv], [sn], [si], [np.zeros(len(si), np.int32)]
    base = len(sv)
    g = np.array([[-8, -1.6, -9], [8, -1.6, -9], [8, -1.6, 6], [-8, -1.6, 6]], np.float32)
    verts.append(g)
    norms.append(np.tile([0, 1, 0], (4, 1)).astype(np.float32))
    idx.append(np.array([[0, 2, 1], [0, 3, 2]], np.int32) + base)
    mids.append(np.ones(2, np.int32))
    base += 4
    if soup:
        c = rng.uniform([-4, -1.5, -6], [4, 3, 0], (soup, 1, 3))
        t = (c + rng.normal(0, 0.6, (soup, 3, 3))).reshape(-1, 3).astype(np.float32)
        verts.append(t)
        # Per-vertex normals of the soup: each triangle's own face normal.
        fn = np.cross(t[1::3] - t[0::3], t[2::3] - t[0::3])
        fn /= np.linalg.norm(fn, axis=1, keepdims=True)
        norms.append(np.repeat(fn, 3, axis=0).astype(np.float32))
        idx.append(np.arange(3 * soup, dtype=np.int32).reshape(-1, 3) + base)
        mids.append(rng.integers(0, 3, soup).astype(np.int32))
        base += 3 * soup
    if mirror:
        m = np.array([[-6, -1.5, -5], [-1.5, -1.5, -7.5], [-3.5, 4, -6.5]], np.float32)
        verts.append(m)
        fn = np.cross(m[1] - m[0], m[2] - m[0])
        norms.append(np.tile(fn / np.linalg.norm(fn), (3, 1)).astype(np.float32))
        idx.append(np.array([[0, 1, 2]], np.int32) + base)
        mids.append(np.full(1, 3, np.int32))
    mats = np.array([[0.8, 0.3, 0.2, 0.04, 0.04, 0.04, 0.3, 0],
                     [0.5, 0.5, 0.5, 0.0, 0.0, 0.0, 1.0, 0],
                     [0.1, 0.6, 0.9, 0.5, 0.5, 0.5, 0.05, 2],
                     [0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 3]], np.float32)
    v = np.concatenate(verts)
    return {"vertices": v, "normals": np.concatenate(norms), "indices": np.concatenate(idx),
            "material_ids": np.concatenate(mids), "materials": mats,
            "camera": [0.5, 1.0, 6.0, 0.0, 0.0, -1.5, 0.0, 1.0, 0.0],
            "bounds": np.concatenate([v.min(0), v.max(0)]).astype(np.float32)}
