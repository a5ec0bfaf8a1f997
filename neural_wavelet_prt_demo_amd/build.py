"""Build recipe for libnwprt.so (gfx950).

Every source under csrc/ is compiled with hipcc for gfx950 (host-only
sources too, so that one toolchain and one set of flags covers the library)
and linked into ``libnwprt.so`` next to this file.  Run as
``python -m neural_wavelet_prt_demo_amd.build`` or through ``build()``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libnwprt.so"
OFFLINE = PKG / "nwprt_offline"
CEILING = PKG / "mfma_ceiling"
VALU_RATE = PKG / "valu_rate"

SOURCES = [
    "host_util.cpp",
    "formats.cpp",
    "image.cpp",
    "synth.cpp",
    "model_folder.cpp",
    "host_abi.cpp",
    "engine.cpp",
    "server.cpp",
    "mesh.cpp",
    "envmap_kernels.hip",
    "encode_kernels.hip",
    "relight_kernels.hip",
    "relight_fp8_kernels.hip",
    "denoise_kernels.hip",
    "gbuffer_kernels.hip",
    "noalloc.hip",
]

# Packed-FP32 VALU forms are disabled for the device code (the host pass
# warns that it does not know the feature; that is expected).
HIPCC_FLAGS = [
    "-x", "hip",
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the HIP toolchain is required to build libnwprt")


def _compile(src: str) -> Path:
    out = OBJ / (src + ".o")
    cmd = [hipcc(), *HIPCC_FLAGS, "-c", str(CSRC / src), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in r.stderr.splitlines() if noise not in l)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{err}")
    if err.strip():
        print(f"[build] {src}:\n{err}", file=sys.stderr)
    return out


def build(verbose: bool = True) -> Path:
    OBJ.mkdir(exist_ok=True)
    jobs = min(len(SOURCES), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objs),
           "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stderr}")
    os.replace(tmp, LIB)
    # The offline trajectory driver, linked against the library next to it.
    exe_tmp = OFFLINE.with_suffix(".tmp")
    cmd = [hipcc(), "--offload-arch=gfx950", "-std=c++17", "-O2", "-Wall", "-x", "hip",
           str(CSRC / "offline_main.cpp"), "-x", "none", str(LIB), "-Wl,-rpath,$ORIGIN",
           "-o", str(exe_tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"offline driver build failed\n{' '.join(cmd)}\n{r.stderr}")
    os.replace(exe_tmp, OFFLINE)
    # The sustained-MFMA-rate probe (standalone; no library dependency).
    exe_tmp = CEILING.with_suffix(".tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, str(CSRC / "mfma_ceiling.hip"), "-o", str(exe_tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"mfma_ceiling build failed\n{' '.join(cmd)}\n{r.stderr}")
    os.replace(exe_tmp, CEILING)
    # The vector-ALU issue-cost probe (standalone; no library dependency).
    exe_tmp = VALU_RATE.with_suffix(".tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, str(CSRC / "valu_rate.hip"), "-o", str(exe_tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"valu_rate build failed\n{' '.join(cmd)}\n{r.stderr}")
    os.replace(exe_tmp, VALU_RATE)
    if verbose:
        print(f"[build] {LIB}\n[build] {OFFLINE}\n[build] {CEILING}\n[build] {VALU_RATE}")
    return LIB


if __name__ == "__main__":
    build()
