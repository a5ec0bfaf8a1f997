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
