#!/usr/bin/env python3
"""Writes the committed fixtures under tests/golden/.

The reference ships no golden vectors, fixtures or tests for this path, and
its hash-grid / MLP library (tiny-cuda-nn) is an empty submodule, so these
fixtures are produced by this repository's own code and pin it against
regressions:

  hashgrid_levels.json   level table of the 16 x 2, 2^19, N_min 16, b 1.5 grid
                         (product host code; the oracle recomputes it)
  relight_small.npz      a reduced relight (8 x 8 pixels, N = 4, K = 8,
                         2^10-entry hash levels): generator inputs, the oracle's
                         Haar coefficients, top-K and FP64 radiance

Run from the repository root: python tests/golden/make_golden.py
"""
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import neural_wavelet_prt_demo_amd as nw  # noqa: E402
from oracle import binding as orc  # noqa: E402

OUT = Path(__file__).resolve().parent

SMALL = dict(width=8, height=8, face_size=4, k=8, log2_hashmap_size=10, seed=5)


def small_inputs():
    d = nw.config_scene_desc(SMALL["face_size"], nw.DTYPE_F32,
                             log2_hashmap_size=SMALL["log2_hashmap_size"])
    params = nw.synth_scene(d, SMALL["seed"], 0.5, 0.5)
    gbuf = nw.synth_gbuffer(SMALL["seed"], SMALL["width"], SMALL["height"], 0.25)
    faces = nw.synth_cubemap(SMALL["seed"], SMALL["face_size"])
    return d, params, gbuf, faces


def main():
    grid = nw.HashgridDesc(16, 2, 19, 16, 1.5)
    lv = nw.hashgrid_levels(grid)
    (OUT / "hashgrid_levels.json").write_text(json.dumps(
        {"grid": {"n_levels": 16, "n_features_per_level": 2, "log2_hashmap_size": 19,
                  "base_resolution": 16, "per_level_scale": 1.5},
         "scale_f32_hex": [float(s).hex() for s in lv["scale"]],
         "resolution": lv["resolution"].tolist(), "dense": lv["dense"].tolist(),
         "size": lv["size"].tolist(), "offset": lv["offset"].tolist()}, indent=1) + "\n")

    d, params, gbuf, faces = small_inputs()
    olv = orc.hashgrid_levels(16, SMALL["log2_hashmap_size"], 16, 1.5)
    oscene = orc.Scene.from_synth(olv, d, params)
    coef = orc.envmap_haar(faces)
    idx, st = orc.topk(coef, SMALL["k"])
    ind, wd = oscene.relight(idx, st, gbuf, threads=1)
    np.savez_compressed(
        OUT / "relight_small.npz",
        position=gbuf["position"], direct=gbuf["direct"], env_flag=gbuf["env_flag"],
        faces=faces, hash_table=params["hash_table"], wavelet_features=params["wavelet_features"],
        projection=params["projection"], layers=params["layers"], coef=coef, topk_idx=idx,
        topk_strength=st, indirect=ind, with_direct=wd)
    print("wrote", OUT / "hashgrid_levels.json", OUT / "relight_small.npz")


if __name__ == "__main__":
    main()
