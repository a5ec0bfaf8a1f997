#!/usr/bin/env python3
"""Relight benchmark: (pixel, wavelet) MLP evals/s at 1920x1080, K = 300.

A step is one whole relight request on resident inputs: Haar transform and
top-K of a 6 x 128 x 128 cubemap already in HBM, per-pixel hash-grid
encoding of the 1080p G-buffer, the fused per-pair MLP over all
1920 * 1080 * 300 pairs, the clamped compensated K-sum and the RGBA8 tone
map.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

FLOP_PER_PAIR = 2 * (32 * 64 + 64 * 64 + 64 * 64 + 64 * 3)  # 20,864
# Encode pass, per pixel: position float4 (16 B) + hit flag (1 B) read,
# 16 levels x 8 corners x 2 FP16 features gathered (512 B), 32 projected
# features written in the MLP element type (64 B for bf16 / fp16).
ENCODE_BYTES_PER_PIXEL = 16 + 1 + 16 * 8 * 2 * 2 + 32 * 2
# Denoiser, per pixel and a-trous pass: colour, normal and albedo read once
# (16 B each; the 25 taps are cache reuse) and the pass result written
# (16 B); plus one colour read (16 B) for the HDR normalisation.
DENOISE_BYTES_PER_PIXEL_PASS = 4 * 16
# G-buffer producer: bytes written per pixel (position, normal, reflect,
# kd, ks float4; roughness float; env flag byte).
GBUFFER_WRITE_BYTES_PER_PIXEL = 5 * 16 + 4 + 1
HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"bf16": 2500.0, "fp16": 2500.0, "fp8": 5000.0, "fp32": 157.3}
METRIC = "(pixel,wavelet) MLP evals/s at 1080p K=300; p50/p99 relight latency vs plain"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", default="bf16", choices=["bf16", "fp16", "fp8"])
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--face-size", type=int, default=128)
    ap.add_argument("--k", type=int, default=300)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="target duration of the CPU-baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the oracle check of the GPU results (part of the CPU-baseline "
                         "leg; --no-cpu-baseline skips it too)")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched uneven-jobs leg")
    ap.add_argument("--serving-requests", type=int, default=600,
                    help="requests per Poisson rate in the served-stream leg")
    ap.add_argument("--no-fp8", action="store_true",
                    help="skip the FP8 legs (steps, batch, served stream) of a BF16 run")
    ap.add_argument("--no-serving", action="store_true",
                    help="skip the served-stream leg")
    ap.add_argument("--no-pipeline", action="store_true",
                    help="skip the two-slot frame pipeline leg")
    ap.add_argument("--no-denoise", action="store_true",
                    help="skip the denoiser leg")
    ap.add_argument("--no-gbuffer", action="store_true",
                    help="skip the G-buffer producer and SH-PRT legs")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="after the run, one device-to-device copy of this many MiB: the "
                         "known-size kernel that scales the FETCH_SIZE / WRITE_SIZE counters")
    return ap.parse_args()


def measured_traffic(workload_key):
    """HBM bytes per launch of the fused kernel from the committed PMC
    summary of this workload (FETCH_SIZE + WRITE_SIZE, each scaled by the
    same counter's reading on a copy of known size), or None."""
    path = ROOT / "profiles" / f"relight_mfma_kernel_{workload_key}_pmc.txt"
    if not path.exists():
        return {}, None
    kv = {}
    for line in path.read_text().splitlines():
        if ":" in line and not line.startswith("#"):
            k, v = line.split(":", 1)
            kv[k.strip()] = v.strip()
    out = {}
    for key in ("traffic_bytes_per_launch", "fused_traffic_bytes_per_launch"):
        try:
            out[key] = float(kv[key])
        except (KeyError, ValueError):
            out[key] = None
    return out, str(path.relative_to(ROOT))


def sustained_mfma_ceiling():
    """Sustained v_mfma_f32_16x16x32_bf16 rate with random operands, from the
    committed summary of `neural_wavelet_prt_demo_amd/mfma_ceiling`, or None."""
    path = ROOT / "profiles" / "mfma_bf16_sustained_ceiling.txt"
    if not path.exists():
        return None, None
    for line in path.read_text().splitlines():
        if line.startswith("random_operands_tflops:"):
            try:
                return float(line.split(":", 1)[1]), str(path.relative_to(ROOT))
            except ValueError:
                return None, None
    return None, None


def percentile(xs, q):
    return float(np.percentile(np.asarray(xs, np.float64), q))


def bench_serving(nw, ctx, prec, n_requests=600, rates=(4, 8, 20, 30, 60, 120),
                  mix_4k=True):
    """Served stream: Poisson arrivals, 4 resident scenes (face 128, FP16 hash
    grid 16 x 2^19 x 2), K = 300; 70 % of requests turn a resident 6x128x128
    cubemap by a random yaw, 30 % bring a fresh cubemap from host memory.
    Latency is measured from each request's scheduled arrival to the poll
    that sees it complete; throughput over the span first arrival -> last
    completion.  Memory budget 200 GB of request state.  Each rate runs
    n_requests requests, or about 40 s of arrivals at the low rates (at
    least 300 requests, so p99 rests on at least three samples above it);
    max_rate_p99_le_16ms is the highest 1080p rate whose p99 is at most
    16 ms."""
    from neural_wavelet_prt_demo_amd import DeviceBuffer, FrameOut
    from neural_wavelet_prt_demo_amd import workloads
    desc = nw.config_scene_desc(128, nw.DTYPE_F16)
    scenes = [ctx.scene_create(desc, nw.synth_scene(desc, s)) for s in workloads.SERVED_SCENE_SEEDS]
    bases = [DeviceBuffer.from_array(nw.synth_cubemap(s, 128)) for s in workloads.SERVED_BASE_SEEDS]
    fresh = [nw.synth_cubemap(s, 128) for s in workloads.SERVED_FRESH_SEEDS]
    views = {"1080p": (1920, 1080), "4k": (3840, 2160)}
    gbufs = {v: nw.DeviceFrame(w, h, nw.synth_gbuffer(5000 + i, w, h, 0.0))
             for i, (v, (w, h)) in enumerate(views.items())}
    ring = 64
    outs = {v: [(DeviceBuffer(w * h * 16), DeviceBuffer(w * h * 16), DeviceBuffer(w * h * 4))
                for _ in range(ring)] for v, (w, h) in views.items()}
    ctx.sync()

    def stream(rate, p4k, seed):
        n = min(n_requests, max(300, int(rate * 40)))
        rng = np.random.default_rng(seed)
        arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
        srv = ctx.server_create(budget_bytes=200 << 30, max_active=ring, precision=prec)
        plan = []
        for i in range(n):
            v = "4k" if rng.random() < p4k else "1080p"
            rotated = rng.random() < 0.7
            plan.append((v, int(rng.integers(4)), rotated, float(rng.uniform(0, 2 * np.pi)),
                         int(rng.integers(len(fresh)))))
        lat, arr, pairs = {}, {}, 0
        t0 = time.perf_counter()
        i = 0
        while len(lat) < n:
            now = time.perf_counter() - t0
            while i < n and arrivals[i] <= now:
                v, s, rotated, yaw, fj = plan[i]
                o = outs[v][i % ring]
                fo = FrameOut(o[0].ptr, o[1].ptr, o[2].ptr)
                if rotated:
                    rid = ctx.server_submit(srv, scenes[s], gbufs[v], faces=bases[s], yaw=yaw,
                                            k=300, face_size=128, out=fo)
                else:
                    rid = ctx.server_submit(srv, scenes[s], gbufs[v], faces=fresh[fj], k=300,
                                            face_size=128, out=fo)
                arr[rid] = t0 + arrivals[i]
                pairs += views[v][0] * views[v][1] * 300
                i += 1
            for rid, _ in ctx.server_poll(srv, wait=0):
                lat[rid] = (time.perf_counter() - arr[rid]) * 1e3
        span = time.perf_counter() - (t0 + arrivals[0])
        st = ctx.server_stats(srv)
        ctx.server_destroy(srv)
        ms = np.array(list(lat.values()))
        return {"rate_req_s": rate, "views": "1080p" if p4k == 0 else f"1080p + {int(p4k * 100)}% 4K",
                "requests": n, "throughput_req_s": n / span,
                "evals_per_s": pairs / span, "p50_ms": percentile(ms, 50),
                "p99_ms": percentile(ms, 99), "max_ms": float(ms.max()),
                "peak_active": st["peak_active"], "peak_state_gb": st["peak_bytes"] / 1e9,
                "ticks": st["ticks"]}

    def closed_loop(n, seed):
        """One request in flight at a time (the next is submitted when the
        previous completes): the server's own request latency, no queueing."""
        rng = np.random.default_rng(seed)
        srv = ctx.server_create(budget_bytes=200 << 30, max_active=ring, precision=prec)
        lat = []
        for i in range(n + 2):
            s = int(rng.integers(4))
            o = outs["1080p"][i % ring]
            fo = FrameOut(o[0].ptr, o[1].ptr, o[2].ptr)
            t = time.perf_counter()
            if rng.random() < 0.7:
                ctx.server_submit(srv, scenes[s], gbufs["1080p"], faces=bases[s],
                                  yaw=float(rng.uniform(0, 2 * np.pi)), k=300, face_size=128,
                                  out=fo)
            else:
                ctx.server_submit(srv, scenes[s], gbufs["1080p"],
                                  faces=fresh[int(rng.integers(len(fresh)))], k=300,
                                  face_size=128, out=fo)
            got = ctx.server_poll(srv, wait=2)
            if i >= 2:  # the first two warm the pool
                lat.append((time.perf_counter() - t) * 1e3)
            assert len(got) == 1
        ctx.server_destroy(srv)
        ms = np.array(lat)
        return {"mode": "closed loop, one request in flight", "views": "1080p", "requests": n,
                "p50_ms": percentile(ms, 50), "p99_ms": percentile(ms, 99),
                "max_ms": float(ms.max())}

    single = closed_loop(40, 11)
    runs = []
    for r in rates:
        runs.append(stream(r, 0.0, r))
        print(f"[bench] served {prec} {r} req/s: p99 {runs[-1]['p99_ms']:.1f} ms",
              file=sys.stderr, flush=True)
    if mix_4k:
        runs.append(stream(60, 0.25, 7))
    ok = [r["rate_req_s"] for r in runs if r["views"] == "1080p" and r["p99_ms"] <= 16.0]
    return {"workload": "Poisson arrivals, 4 resident scenes, K=300, 70% yaw-rotated resident "
                        f"6x128x128 cubemap / 30% fresh host cubemap, up to {n_requests} "
                        "requests per rate (40 s of arrivals, at least 300, at the low rates)",
            "memory_budget_gb": 200, "closed_loop": single, "runs": runs,
            "max_rate_p99_le_16ms": max(ok) if ok else None}


def gbuffer_bench_mesh():
    """Synthetic scene for the G-buffer producer leg: a 768 x 672 heightfield
    (1,032,192 triangles, per-vertex normals) under a 512 x 256 UV sphere
    (262,144 triangles), three cookTorrance materials; ~1.29 M triangles in
    all."""
    gx, gz = 768, 672
    xs = np.linspace(-8.0, 8.0, gx + 1)
    zs = np.linspace(-9.0, 5.0, gz + 1)
    X, Z = np.meshgrid(xs, zs)
    Y = -1.5 + 0.35 * np.sin(0.9 * X) * np.cos(0.7 * Z) + 0.1 * np.sin(3.1 * X + 2.3 * Z)
    tv = np.stack([X, Y, Z], -1).reshape(-1, 3)
    dYdx = 0.35 * 0.9 * np.cos(0.9 * X) * np.cos(0.7 * Z) + 0.31 * np.cos(3.1 * X + 2.3 * Z)
    dYdz = -0.35 * 0.7 * np.sin(0.9 * X) * np.sin(0.7 * Z) + 0.23 * np.cos(3.1 * X + 2.3 * Z)
    tn = np.stack([-dYdx, np.ones_like(X), -dYdz], -1).reshape(-1, 3)
    tn /= np.linalg.norm(tn, axis=1, keepdims=True)
    i0 = (np.arange(gz)[:, None] * (gx + 1) + np.arange(gx)[None, :]).reshape(-1)
    ti = np.concatenate([np.stack([i0, i0 + gx + 1, i0 + 1], 1),
                         np.stack([i0 + 1, i0 + gx + 1, i0 + gx + 2], 1)])
    nu, nv = 512, 256
    th = np.linspace(0, np.pi, nv + 1)
    ph = np.linspace(0, 2 * np.pi, nu, endpoint=False)
    T, PH = np.meshgrid(th, ph, indexing="ij")
    sn = np.stack([np.sin(T) * np.cos(PH), np.cos(T), np.sin(T) * np.sin(PH)], -1).reshape(-1, 3)
    sv = np.array([0.5, 0.2, -2.0]) + 1.6 * sn
    a = (np.arange(nv)[:, None] * nu + np.arange(nu)[None, :]).reshape(-1)
    b = (np.arange(nv)[:, None] * nu + (np.arange(nu)[None, :] + 1) % nu).reshape(-1)
    si = np.concatenate([np.stack([a, a + nu, b], 1), np.stack([b, a + nu, b + nu], 1)]) + len(tv)
    verts = np.concatenate([tv, sv]).astype(np.float32)
    norms = np.concatenate([tn, sn]).astype(np.float32)
    idx = np.concatenate([ti, si]).astype(np.int32)
    mids = np.concatenate([(np.arange(len(ti)) // 4096) % 2,
                           np.full(len(si), 2)]).astype(np.int32)
    mats = [[0.45, 0.4, 0.3, 0.04, 0.04, 0.04, 0.6, 2], [0.3, 0.45, 0.25, 0.04, 0.04, 0.04, 0.8, 2],
            [0.8, 0.2, 0.2, 0.5, 0.5, 0.5, 0.15, 2]]
    return verts, norms, idx, mids, mats


def bench_gbuffer(nw, ctx, W, H, stream, dfaces, n_face):
    """G-buffer producer leg: one primary ray per pixel at W x H over the
    ~1.29 M-triangle synthetic scene (BVH built on the host once)."""
    from neural_wavelet_prt_demo_amd import Event
    verts, norms, idx, mids, mats = gbuffer_bench_mesh()
    t0 = time.perf_counter()
    mesh = ctx.mesh_create(verts, idx, mats, normals=norms, material_ids=mids)
    build_s = time.perf_counter() - t0
    info = ctx.mesh_info(mesh)
    gframe = nw.DeviceFrame(W, H)
    cam = [[0.0, 2.5, 7.5], [0.0, -0.8, -2.0], [0.0, 1.0, 0.0]]
    cast = lambda: ctx.gbuffer_cast(mesh, cam, gframe, env_faces=dfaces, env_n=n_face,
                                    stream=stream)
    for _ in range(3):
        cast()
    ctx.sync()
    ev = [Event() for _ in range(11)]
    ev[0].record(stream)
    for i in range(10):
        cast()
        ev[i + 1].record(stream)
    ctx.sync()
    ms = [ev[i].elapsed_ms(ev[i + 1]) for i in range(10)]
    cast_ms = float(np.mean(ms))
    flags = gframe.bufs["env_flag"].download(np.uint8, (W * H,))
    wb = W * H * GBUFFER_WRITE_BYTES_PER_PIXEL
    out = {"workload": f"primary rays {W}x{H} over {info['triangles']} triangles "
                       f"(heightfield + sphere), BVH {info['bvh_nodes']} nodes, depth "
                       f"{info['bvh_depth']}",
           "bvh_build_s": build_s, "ms": cast_ms, "p99_ms": percentile(ms, 99),
           "rays_per_s": W * H / (cast_ms / 1e3), "hit_fraction": float(flags.mean()),
           "write_bytes_per_pixel": GBUFFER_WRITE_BYTES_PER_PIXEL,
           "write_gbs": wb / (cast_ms / 1e3) / 1e9}
    # The same cast with the direct-light term: per pixel one environment
    # and one BSDF sample, each with a shadow ray, lit by the bench cubemap.
    for spp in (1, 4):
        castd = lambda: ctx.gbuffer_cast(mesh, cam, gframe, env_faces=dfaces, env_n=n_face,
                                         light_faces=dfaces, light_n=n_face, direct_samples=spp,
                                         seed=7, stream=stream)
        castd()
        ctx.sync()
        ev[0].record(stream)
        for i in range(5):
            castd()
            ev[i + 1].record(stream)
        ctx.sync()
        dms = float(np.mean([ev[i].elapsed_ms(ev[i + 1]) for i in range(5)]))
        out[f"with_direct_{spp}spp_ms"] = dms
    ctx.mesh_destroy(mesh)
    del gframe
    return out


def bench_sh(nw, ctx, W, H, stream, dfaces, n_face, samples=64, tm_samples=128, basis_n=32):
    """SH-PRT leg (the reference's shComputeTMatrix -> shRender chain):
    degree 8 (81 x 81 transfer matrices, the reference's tested degree) over
    a 256 x 128 UV sphere on a ground quad (65,538 triangles, 33,028
    vertices: 2.6 GB of transfer matrices).  The matrices are computed here
    from the 81 real-SH basis maps (basis_n^2 texels per face) with
    `tm_samples` paths per vertex (the reference uses 4096; the figure for
    4096 is the measured time scaled by 4096 / tm_samples, as the work is
    linear in the samples), then rendered at 1080p with `samples` BSDF
    samples per distinct (vertex, material) pair of the triangle corners."""
    from neural_wavelet_prt_demo_amd import DeviceBuffer, Event
    nu, nv = 256, 128
    th = np.linspace(0, np.pi, nv + 1)
    ph = np.linspace(0, 2 * np.pi, nu, endpoint=False)
    T_, P_ = np.meshgrid(th, ph, indexing="ij")
    sn = np.stack([np.sin(T_) * np.cos(P_), np.cos(T_), np.sin(T_) * np.sin(P_)], -1).reshape(-1, 3)
    sv = np.array([0.0, 0.2, -2.0]) + 1.6 * sn
    a = (np.arange(nv)[:, None] * nu + np.arange(nu)[None, :]).reshape(-1)
    b = (np.arange(nv)[:, None] * nu + (np.arange(nu)[None, :] + 1) % nu).reshape(-1)
    si = np.concatenate([np.stack([a, a + nu, b], 1), np.stack([b, a + nu, b + nu], 1)])
    g = np.array([[-8, -1.4, -9], [8, -1.4, -9], [8, -1.4, 5], [-8, -1.4, 5]])
    verts = np.concatenate([sv, g]).astype(np.float32)
    norms = np.concatenate([sn, np.tile([0.0, 1.0, 0.0], (4, 1))]).astype(np.float32)
    base = len(sv)
    idx = np.concatenate([si, np.array([[0, 2, 1], [0, 3, 2]]) + base]).astype(np.int32)
    mids = np.concatenate([np.zeros(len(si)), np.ones(2)]).astype(np.int32)
    mats = [[0.8, 0.3, 0.2, 0.04, 0.04, 0.04, 0.3, 2], [0.5, 0.5, 0.5, 0.04, 0.04, 0.04, 0.8, 0]]
    mesh = ctx.mesh_create(verts, idx, mats, normals=norms, material_ids=mids)
    nc = 81
    dT = DeviceBuffer(len(verts) * nc * nc * 12)
    dbasis = DeviceBuffer.from_array(nw.sh_basis_faces(8, basis_n))
    tm = lambda: ctx.sh_compute_transfer(mesh, dbasis, basis_n, 8, dT, samples=tm_samples,
                                         seed=1, stream=stream)
    tm()  # warm-up (work buffers, code objects)
    ctx.sync()
    tev = [Event() for _ in range(2)]
    tev[0].record(stream)
    tm()
    tev[1].record(stream)
    ctx.sync()
    tm_ms = tev[0].elapsed_ms(tev[1])
    triples = len(verts) * tm_samples * nc
    rng = np.random.default_rng(5)
    L = rng.normal(0, 0.2, (nc, 3)).astype(np.float32)
    L[0] = [2.0, 1.8, 1.6]
    dL = DeviceBuffer.from_array(L)
    out = DeviceBuffer(W * H * 16)
    cam = [[0.0, 2.5, 7.5], [0.0, -0.8, -2.0], [0.0, 1.0, 0.0]]
    run = lambda: ctx.sh_render(mesh, cam, W, H, dT, dL, 8, out, brdf_samples=samples, seed=1,
                                env_faces=dfaces, env_n=n_face, stream=stream)
    run()
    ctx.sync()
    ev = [Event() for _ in range(6)]
    ev[0].record(stream)
    for i in range(5):
        run()
        ev[i + 1].record(stream)
    ctx.sync()
    ms = float(np.mean([ev[i].elapsed_ms(ev[i + 1]) for i in range(5)]))
    res = {"workload": f"SH-PRT degree 8, {W}x{H}, {len(idx)} triangles, {len(verts)} vertices, "
                       f"{samples} BSDF samples per vertex; transfer matrices computed from 81 "
                       f"basis maps ({basis_n}^2 texels per face), {tm_samples} paths per vertex",
           "ms": ms, "transfer_bytes": len(verts) * nc * nc * 12,
           "transfer_read_gbs": len(verts) * nc * nc * 12 / (ms / 1e3) / 1e9,
           "transfer_compute_ms": tm_ms, "transfer_samples": tm_samples,
           "transfer_paths_x_maps_per_s": triples / (tm_ms / 1e3),
           "transfer_compute_ms_at_4096_samples": tm_ms * 4096 / tm_samples}
    ctx.mesh_destroy(mesh)
    del dT
    return res


def bench_pipelined(nw, ctx, scene, dfaces, frame, prec, k, n, steps, warmup):
    """Frames through the two-slot pipeline of the reference's interactive
    loop (Renderer/OutputVisualizer.cpp:71-123): per frame an envmap update,
    p1 and p2 -- the work of one sequential step -- with the update and p1
    of frame f + 1 on a second stream while p2 of frame f runs.  Timed from
    before the first frame's update to the end of the last p2."""
    from neural_wavelet_prt_demo_amd import Event
    from neural_wavelet_prt_demo_amd.pipeline import FramePipeline
    envs = [ctx.envmap_from_faces(dfaces, k, face_size=n) for _ in range(2)]
    pipe = FramePipeline(ctx, scene, envs, prec)
    for _ in range(max(2, warmup)):
        pipe.submit(frame, dfaces)
    pipe.drain()
    start, end = Event(), Event()
    start.record(pipe.side.handle)
    for _ in range(steps):
        pipe.submit(frame, dfaces)
    end.record(ctx.stream)
    pipe.drain()
    ms = start.elapsed_ms(end)
    for e in envs:
        ctx.envmap_destroy(e)
    pairs = frame.width * frame.height * k
    return {"frames": steps, "ms_per_frame": ms / steps,
            "evals_per_s": pairs * steps / (ms / 1e3),
            "streams": "map update + p1 on a side stream, p2 on the context stream"}


def bench_batch(nw, ctx, prec, reps=3):
    """32 uneven relight jobs (workloads.batch_jobs: 256^2..1080p, K 32..300,
    own 6x64x64 cubemaps, 2 scenes) as one captured graph with the persistent
    kernel, against the same jobs relit one call at a time."""
    from neural_wavelet_prt_demo_amd import workloads
    from neural_wavelet_prt_demo_amd import Event
    specs = workloads.batch_jobs()
    desc = nw.config_scene_desc(64, nw.DTYPE_F16)
    scenes = [ctx.scene_create(desc, nw.synth_scene(desc, s)) for s in workloads.BATCH_SCENE_SEEDS]
    jobs, keep = [], []
    for sp in specs:
        env = ctx.envmap_from_faces(nw.synth_cubemap(sp.seed, 64), sp.k)
        frame = nw.DeviceFrame(sp.width, sp.height, nw.synth_gbuffer(sp.seed, sp.width, sp.height, 0.0))
        jobs.append((scenes[sp.scene], env, frame))
        keep.append(frame)
    ctx.sync()
    bid = ctx.batch_create(jobs, prec)
    best, _ = ctx.batch_autotune(bid)
    n_items, n_pairs, variant = ctx.batch_info(bid)
    stream = ctx.stream

    def timed(fn, n):
        fn()
        ctx.sync()
        ev = [Event() for _ in range(n + 1)]
        ev[0].record(stream)
        for i in range(n):
            fn()
            ev[i + 1].record(stream)
        ctx.sync()
        return [ev[i].elapsed_ms(ev[i + 1]) for i in range(n)]

    batch_ms = timed(lambda: ctx.batch_launch(bid), reps)

    def sequential():
        for scene, env, frame in jobs:
            ctx.relight(scene, env, frame, prec, nw.TONEMAP_FULL, stream)

    seq_ms = timed(sequential, max(1, reps - 1))
    ctx.batch_destroy(bid)
    b = float(np.mean(batch_ms))
    s = float(np.mean(seq_ms))
    return {"workload": "32 jobs: 256^2..1920x1080, K in {32,64,128,300}, 6x64x64 cubemap each, "
                        "2 scenes", "jobs": len(specs),
            "pixels": int(sum(sp.width * sp.height for sp in specs)), "pairs": int(n_pairs),
            "work_items": int(n_items),
            "tile_variant": ("fp8 persistent" if prec == nw.PREC_FP8
                             else nw.Context.relight_variants()[variant]),
            "graph_ms": b, "evals_per_s": n_pairs / (b / 1e3),
            "one_call_per_job_ms": s, "one_call_per_job_evals_per_s": n_pairs / (s / 1e3),
            "speedup_vs_one_call_per_job": s / b}


def bench_haar(nw, ctx, stream, n=128, batch_maps=64, reps=20):
    """The Haar transform alone (nwprt_envmap_haar): one 6 x n x n map, as a
    served request runs it, and batch_maps maps in one call.  Algorithmic
    bytes: the FP32 texels read once and the coefficients written once,
    2 x 6 n^2 x 3 x 4 B per map."""
    from neural_wavelet_prt_demo_amd import DeviceBuffer, Event
    rng = np.random.default_rng(9)
    faces = rng.lognormal(0, 1.5, (batch_maps, 6, n, n, 3)).astype(np.float32)
    dfaces, dcoef = DeviceBuffer.from_array(faces), DeviceBuffer(faces.nbytes)
    per_map = 2 * 6 * n * n * 3 * 4

    def timed(n_maps):
        ctx.envmap_haar(dfaces, n, n_maps, dcoef, stream)
        ctx.sync()
        ev = [Event() for _ in range(reps + 1)]
        ev[0].record(stream)
        for i in range(reps):
            ctx.envmap_haar(dfaces, n, n_maps, dcoef, stream)
            ev[i + 1].record(stream)
        ctx.sync()
        return float(np.mean([ev[i].elapsed_ms(ev[i + 1]) for i in range(reps)]))

    one, many = timed(1), timed(batch_maps)
    del dfaces, dcoef
    return {"bound": "hbm", "face_size": n, "bytes_per_map": per_map,
            "one_map_us": one * 1e3, "one_map_gbs": per_map / (one / 1e3) / 1e9,
            "one_map_frac": per_map / (one / 1e3) / 1e9 / HBM_PEAK_GBS,
            "maps_per_call": batch_maps, "batch_ms": many,
            "batch_gbs": batch_maps * per_map / (many / 1e3) / 1e9,
            "batch_frac": batch_maps * per_map / (many / 1e3) / 1e9 / HBM_PEAK_GBS}


def cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; the relight path shards by "
                                   "pixel rows across GPUs only by design (no multi-GPU run)",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import neural_wavelet_prt_demo_amd as nw
    from neural_wavelet_prt_demo_amd import DeviceBuffer, Event

    W, H, N, K = args.width, args.height, args.face_size, args.k
    P = W * H
    pairs = P * K
    prec = nw.PRECISIONS[args.precision]

    desc = nw.config_scene_desc(N, nw.DTYPE_F16)
    params = nw.synth_scene(desc, 0)
    gbuf = nw.synth_gbuffer(0, W, H, 0.0)
    faces = nw.synth_cubemap(0, N)

    ctx = nw.Context(0)
    stream = ctx.stream
    scene = ctx.scene_create(desc, params)
    dfaces = DeviceBuffer.from_array(faces)
    env = ctx.envmap_from_faces(dfaces, K, face_size=N)
    frame = nw.DeviceFrame(W, H, gbuf)
    ctx.sync()

    def step(precision=prec):
        ctx.envmap_update_faces(env, dfaces, stream)
        ctx.relight(scene, env, frame, precision, nw.TONEMAP_FULL, stream)

    # Tile autotuning for this shape (records the fastest fused-kernel
    # variant for the (precision, K, pixels) bucket; steps then use it).
    variant, _ = ctx.autotune_relight(scene, env, frame, prec)
    names = (nw.Context.relight_fp8_variants() if prec == nw.PREC_FP8
             else nw.Context.relight_variants())
    variant_name = names[variant]

    # --- fused MFMA path: throughput and per-step latency ------------------
    def timed_steps(precision, steps, warmup):
        for _ in range(warmup):
            step(precision)
        ctx.sync()
        evs = [Event() for _ in range(steps + 1)]
        mids = [Event() for _ in range(steps)]  # map updated, relight starts
        t0 = time.perf_counter()
        evs[0].record(stream)
        for i in range(steps):
            ctx.envmap_update_faces(env, dfaces, stream)
            mids[i].record(stream)
            ctx.relight(scene, env, frame, precision, nw.TONEMAP_FULL, stream)
            evs[i + 1].record(stream)
        ctx.sync()
        wall = time.perf_counter() - t0
        step_ms = [evs[i].elapsed_ms(evs[i + 1]) for i in range(steps)]
        total_ms = evs[0].elapsed_ms(evs[-1])
        # The timed steps' dominant kernel: the fused-encode relight (plus
        # its selection prep, and for FP8 the calibration), timed live
        # between the events of every step.
        live = float(np.mean([mids[i].elapsed_ms(evs[i + 1]) for i in range(steps)]))
        return step_ms, total_ms, wall, live

    step_ms, total_ms, wall, relight_live_ms = timed_steps(prec, args.steps, args.warmup)
    value = pairs * args.steps / (total_ms / 1e3)

    # --- the FP8 form's steps (same work, FP8-E4M3 hidden layers) ----------
    fp8_leg = None
    if prec != nw.PREC_FP8 and not args.no_fp8:
        v8, _ = ctx.autotune_relight(scene, env, frame, nw.PREC_FP8)
        s8, t8, _, live8 = timed_steps(nw.PREC_FP8, args.steps, args.warmup)
        fp8_leg = {"ms_per_step": t8 / args.steps, "evals_per_s": pairs * args.steps / (t8 / 1e3),
                   "latency_ms": {"p50": percentile(s8, 50), "p99": percentile(s8, 99)},
                   "relight_live_ms": live8,
                   "relight_tflops": pairs * FLOP_PER_PAIR / (live8 / 1e3) / 1e12,
                   "relight_frac_of_fp8_peak": pairs * FLOP_PER_PAIR / (live8 / 1e3) / 1e12 /
                   PEAK_TFLOPS["fp8"],
                   "note": "relight interval = selection tiles + calibration (16K pixels x K) + "
                           "relight_fp8_kernel with the fused encode",
                   "tile_variant": nw.Context.relight_fp8_variants()[v8],
                   "speedup_vs_" + args.precision: (total_ms / args.steps) / (t8 / args.steps)}

    # --- two-slot frame pipeline: the same per-frame work, with the map
    # update and p1 of frame f + 1 overlapping p2 of frame f ---------------
    pipelined = None
    if not args.no_pipeline:
        pipelined = bench_pipelined(nw, ctx, scene, dfaces, frame, prec, K, N, args.steps,
                                    args.warmup)
        pipelined["speedup_vs_sequential"] = (total_ms / args.steps) / pipelined["ms_per_frame"]

    # --- hash-grid encode pass (p1) alone: random gathers, HBM-side -------
    pev = [Event() for _ in range(6)]
    p1ms = []
    for i in range(5):
        pev[i].record(stream)
        ctx.TCNNPass_p1(scene, 0, frame, prec, stream)
        pev[i + 1].record(stream)
        ctx.sync()
        p1ms.append(pev[i].elapsed_ms(pev[i + 1]))
    encode_ms = float(np.mean(p1ms))
    encode_gbs = P * ENCODE_BYTES_PER_PIXEL / (encode_ms / 1e3) / 1e9

    # --- denoiser on the relit frame (with-direct output, normal + kd guides)
    denoise_leg = None
    if not args.no_denoise:
        dn_out = DeviceBuffer(P * 16)
        dev = [Event() for _ in range(6)]
        dms = []
        for i in range(5):
            dev[i].record(stream)
            ctx.denoise(frame.out["indirect_with_direct"], frame.bufs["normal"],
                        frame.bufs["kd"], dn_out, W, H, 0.0, None, stream)
            dev[i + 1].record(stream)
            ctx.sync()
            dms.append(dev[i].elapsed_ms(dev[i + 1]))
        dn_ms = float(np.mean(dms[1:]))
        it = nw.Context.denoise_default_params().iterations
        dn_bytes = P * (DENOISE_BYTES_PER_PIXEL_PASS * it + 16)
        denoise_leg = {"bound": "hbm", "passes": it, "ms": dn_ms,
                       "bytes_per_pixel": DENOISE_BYTES_PER_PIXEL_PASS * it + 16,
                       "achieved_gbs": dn_bytes / (dn_ms / 1e3) / 1e9,
                       "frac": dn_bytes / (dn_ms / 1e3) / 1e9 / HBM_PEAK_GBS}
        del dn_out

    # --- G-buffer producer (the step before the path) ----------------------
    gbuffer_leg = sh_leg = None
    if not args.no_gbuffer:
        gbuffer_leg = bench_gbuffer(nw, ctx, W, H, stream, dfaces, N)
        sh_leg = bench_sh(nw, ctx, W, H, stream, dfaces, N)

    # --- dominant kernel alone: the per-pair pass (p2) with HIP events -----
    kev = [Event() for _ in range(6)]
    kms = []
    for i in range(5):
        kev[i].record(stream)
        ctx.TCNNPass_p2(scene, env, 0, frame, prec, nw.TONEMAP_FULL, stream)
        kev[i + 1].record(stream)
        ctx.sync()
        kms.append(kev[i].elapsed_ms(kev[i + 1]))
    kernel_ms = float(np.mean(kms))
    achieved = pairs * FLOP_PER_PAIR / (kernel_ms / 1e3) / 1e12

    # --- plain FP32 form --------------------------------------------------
    plain_ms = []
    step(nw.PREC_FP32)
    ctx.sync()
    for _ in range(max(1, args.plain_steps)):
        a, b = Event(), Event()
        a.record(stream)
        step(nw.PREC_FP32)
        b.record(stream)
        ctx.sync()
        plain_ms.append(a.elapsed_ms(b))
    plain_value = pairs / (np.mean(plain_ms) / 1e3)

    if args.copy_calibration_mib > 0:
        nb = args.copy_calibration_mib << 20
        src, dst = DeviceBuffer(nb), DeviceBuffer(nb)
        dst.copy_from(src, nb, stream)
        ctx.sync()
        del src, dst

    workload_key = f"{W}x{H}_k{K}_n{N}_{args.precision}"
    traffic, traffic_src = measured_traffic(workload_key)
    ceiling, ceiling_src = sustained_mfma_ceiling()
    peak = PEAK_TFLOPS[args.precision]
    mlp_only = {"kernel": "relight_mfma_kernel, features from pass p1 (pass p2 alone)",
                "kernel_ms": kernel_ms, "achieved": achieved, "frac": achieved / peak,
                "traffic": traffic.get("traffic_bytes_per_launch")}
    if prec != nw.PREC_FP8:
        # BF16 / FP16 steps run the fused-encode relight: that is the kernel
        # timed live in the steps and the one the roofline describes.
        live = pairs * FLOP_PER_PAIR / (relight_live_ms / 1e3) / 1e12
        roofline = {"bound": "mfma",
                    "kernel": "relight_mfma_kernel with the fused encode (timed steps)",
                    "achieved": live, "peak": peak, "unit": "TFLOP/s", "frac": live / peak,
                    "traffic": traffic.get("fused_traffic_bytes_per_launch"),
                    "traffic_source": traffic_src, "kernel_ms": relight_live_ms,
                    "flop_per_pair": FLOP_PER_PAIR, "mlp_only": mlp_only}
    else:
        roofline = {"bound": "mfma", "kernel": "relight_fp8_kernel (pass p2)",
                    "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                    "frac": achieved / peak, "traffic": traffic.get("traffic_bytes_per_launch"),
                    "traffic_source": traffic_src, "kernel_ms": kernel_ms,
                    "flop_per_pair": FLOP_PER_PAIR}
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic (seeded G-buffer, log-normal cubemap with sun, random-init network)",
        "config": {"workload": f"relight {W}x{H}, cubemap 6x{N}x{N}, K={K}, hash grid 16x2 "
                               f"2^19 FP16, MLP 32-64-64-64-3 {args.precision}",
                   "pixels": P, "k": K, "pairs_per_step": pairs,
                   "tile_variant": variant_name},
        "latency_ms": {"p50": percentile(step_ms, 50), "p99": percentile(step_ms, 99),
                       "max": float(max(step_ms))},
        "wall_s": wall,
        "plain_fp32": {"evals_per_s": plain_value, "ms_per_step": float(np.mean(plain_ms)),
                       "p50_ms": percentile(plain_ms, 50), "p99_ms": percentile(plain_ms, 99)},
        "speedup_vs_plain_fp32": value / plain_value,
        "roofline": roofline,
        "encode_pass": {"bound": "hbm", "ms": encode_ms, "bytes_per_pixel": ENCODE_BYTES_PER_PIXEL,
                        "achieved_gbs": encode_gbs, "frac": encode_gbs / HBM_PEAK_GBS},
    }
    if pipelined is not None:
        result["pipelined_two_slot"] = pipelined
    if denoise_leg is not None:
        result["denoise_pass"] = denoise_leg
    if gbuffer_leg is not None:
        result["gbuffer_cast"] = gbuffer_leg
    if sh_leg is not None:
        result["sh_prt_render"] = sh_leg
    if ceiling is not None and args.precision in ("bf16", "fp16"):
        # The published peak assumes the nominal clock; with real (toggling)
        # operands a pure MFMA loop on this GPU runs power-limited below it.
        result["roofline"]["sustained_mfma_ceiling"] = {
            "tflops": ceiling, "frac": result["roofline"]["achieved"] / ceiling,
            "source": ceiling_src}

    if fp8_leg is not None:
        result["fp8_form"] = fp8_leg
    result["haar"] = bench_haar(nw, ctx, stream, N)

    # --- batched uneven jobs: one graph, one persistent kernel ------------
    with_fp8 = prec != nw.PREC_FP8 and not args.no_fp8
    if not args.no_batch:
        result["batched"] = bench_batch(nw, ctx, prec)
        if with_fp8:
            result["batched_fp8"] = bench_batch(nw, ctx, nw.PREC_FP8)
    if not args.no_serving:
        result["served"] = bench_serving(nw, ctx, prec, args.serving_requests)
        if with_fp8:
            result["served_fp8"] = bench_serving(nw, ctx, nw.PREC_FP8, args.serving_requests,
                                                 rates=(8, 20, 60), mix_4k=False)

    # --- CPU-baseline leg, part 1: the FP64 oracle checks the GPU results on
    # a pixel sample (outside every timed region) ---------------------------
    if not args.no_cpu_baseline and not args.no_accuracy:
        from oracle import binding as orc
        idx, st = ctx.envmap_selection(env)
        ctx.relight(scene, env, frame, nw.PREC_FP32, nw.TONEMAP_FULL, stream)
        ctx.sync()
        res = frame.results()
        ctx.relight(scene, env, frame, prec, nw.TONEMAP_FULL, stream)
        ctx.sync()
        res_fast = frame.results()
        forms = [("fp32", res), (args.precision, res_fast)]
        if with_fp8:
            ctx.relight(scene, env, frame, nw.PREC_FP8, nw.TONEMAP_FULL, stream)
            ctx.sync()
            forms.append(("fp8", frame.results()))
        rng = np.random.default_rng(0)
        pix = rng.choice(P, 1000, replace=False)
        lv = orc.hashgrid_levels()
        oscene = orc.Scene.from_synth(lv, desc, params)
        ref, _ = oscene.relight(idx, st, gbuf, pixels=pix, threads=min(16, os.cpu_count() or 1))
        acc = {}
        for name, r in forms:
            got = r["indirect"][pix, :3].astype(np.float64)
            mse = float(np.mean((got - ref) ** 2))
            den = np.maximum(np.abs(ref), 1e-3 * np.abs(ref).max())
            acc[name] = {"max_rel_err": float(np.max(np.abs(got - ref) / den)),
                         "psnr_db": float(10 * np.log10(ref.max() ** 2 / mse)) if mse > 0 else None}
        result["accuracy_vs_fp64_oracle"] = {"sample_pixels": len(pix), **acc}

    # --- CPU-baseline leg, part 2: the oracle's FP32 port timed on the host
    # cores ------------------------------------------------------------------
    if not args.no_cpu_baseline:
        from oracle import binding as orc
        idx, st = ctx.envmap_selection(env)
        lv = orc.hashgrid_levels()
        oscene = orc.Scene.from_synth(lv, desc, params)
        threads = max(1, min(16, len(os.sched_getaffinity(0))))
        _, secs = oscene.cpu_relight_fp32(idx, st, gbuf, 0, W, threads)  # one row, calibration
        rows = int(max(1, min(H, args.cpu_seconds / max(secs, 1e-6))))
        _, secs = oscene.cpu_relight_fp32(idx, st, gbuf, 0, rows * W, threads)
        result["cpu_baseline"] = {"value": rows * W * K / secs, "unit": "evals/s",
                                  "cores": threads, "kind": "port", "cpu_model": cpu_model(),
                                  "sample": f"first {rows} rows x {W} px of the same frame, "
                                            f"K={K}, FP32 fused per pixel, {secs:.1f} s"}
    ctx.close()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
