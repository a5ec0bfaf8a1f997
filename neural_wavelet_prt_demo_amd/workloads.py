"""Seeded synthetic workloads shared by the benchmark and the tests.

batch_jobs(): 32 relight jobs of uneven size, the batched-jobs workload --
job seed s in 0..31 draws its resolution from {256x256, 512x512, 1280x720,
1920x1080} and its K from {32, 64, 128, 300} with numpy's default_rng(s),
uses scene s % 2, its own 6 x 64 x 64 cubemap (synth_cubemap(s, 64)) and
its own G-buffer (synth_gbuffer(s, w, h, 0)).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List

import numpy as np

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
