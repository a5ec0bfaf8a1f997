"""Seeded relight cases shared by the parity tests.

Inputs come from the product's seeded generator (host arrays), so the GPU
path and the oracle read identical bits.
"""
from __future__ import annotations

from dataclasses import dataclass
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
              projection=True, softplus=True, levels=16, features=2, log2_hashmap_size=19) -> Case:
    if hash_dtype is None:
        hash_dtype = nw.DTYPE_F32
    desc = nw.config_scene_desc(face_size, hash_dtype, levels=levels, features=features,
                                log2_hashmap_size=log2_hashmap_size, projection=projection,
                                softplus=softplus)
    params = nw.synth_scene(desc, seed, hash_std, feature_std)
    gbuf = nw.synth_gbuffer(seed, width, height, background)
    faces = nw.synth_cubemap(seed, face_size)
    lv = orc.hashgrid_levels(levels, log2_hashmap_size, 16, 1.5)
    oscene = orc.Scene.from_synth(lv, desc, params)
    return Case(desc, params, gbuf, faces, width, height, k, oscene, lv)
