"""Seeded synthetic workloads shared by the benchmark and the tests.

batch_jobs(): 32 relight jobs of uneven size, the batched-jobs workload --
job seed s in 0..31 draws its resolution from {256x256, 512x512, 1280x720,
1920x1080} and its K from {32, 64, 128, 300} with numpy's default_rng(s),
uses scene s % 2, its own 6 x 64 x 64 cubemap (synth_cubemap(s, 64)) and
its own G-buffer (synth_gbuffer(s, w, h, 0)); the two scenes are
synth_scene(config_scene_desc(64, FP16), BATCH_SCENE_SEEDS[i]).

The served stream: 4 resident scenes synth_scene(config_scene_desc(128,
FP16), SERVED_SCENE_SEEDS[i]), resident base cubemaps
synth_cubemap(SERVED_BASE_SEEDS[i], 128) that requests turn by a yaw, and
fresh cubemaps synth_cubemap(SERVED_FRESH_SEEDS[j], 128).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List

import numpy as np

BATCH_SCENE_SEEDS = (1000, 1001)
SERVED_SCENE_SEEDS = (2000, 2001, 2002, 2003)
SERVED_BASE_SEEDS = (3000, 3001, 3002, 3003)
SERVED_FRESH_SEEDS = tuple(range(4000, 4008))

RESOLUTIONS = [(256, 256), (512, 512), (1280, 720), (1920, 1080)]
KS = [32, 64, 128, 300]


@dataclass(frozen=True)
class JobSpec:
    seed: int
    width: int
    height: int
    k: int
    scene: int


def batch_jobs(n_jobs: int = 32, n_scenes: int = 2) -> List[JobSpec]:
    out = []
    for s in range(n_jobs):
        rng = np.random.default_rng(s)
        w, h = RESOLUTIONS[int(rng.integers(len(RESOLUTIONS)))]
        k = KS[int(rng.integers(len(KS)))]
        out.append(JobSpec(s, w, h, k, s % n_scenes))
    return out
