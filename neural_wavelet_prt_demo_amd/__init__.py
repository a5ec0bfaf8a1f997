"""MI355X (gfx950) neural wavelet PRT relighting path.

Python binding of ``libnwprt.so`` (C ABI declared in ``include/nwprt.h``)
over ctypes.  The compute path is the HIP library; this module only moves
numpy arrays in and out.  If the library is missing, every entry point
raises -- there is no Python or CPU fallback.

Names mirror the reference renderer where one exists: ``initTCNN`` loads a
weight folder (Renderer/tinycudann/tcnn.cu:49), ``TCNNPass_p1`` /
``TCNNPass_p2`` are the per-pixel and per-pair passes (tcnn.cu:83, :111),
``computeFinalPixelColors`` is the tone map (Renderer/toneMap.cu:73).
"""
from __future__ import annotations

import ctypes
import os
import re
from ctypes import (POINTER, Structure, c_char_p, c_double, c_float, c_int, c_int32, c_int64,
                    c_size_t, c_uint8, c_uint32, c_uint64, c_void_p)
from pathlib import Path
from typing import Dict, Optional

import numpy as np

from . import _hip
from ._hip import DeviceBuffer, Event, Stream

LIB_PATH = Path(__file__).resolve().with_name("libnwprt.so")

MAX_K = 1024
MAX_FACE_SIZE = 128
MAX_LEVELS = 32

# nwprt_status
OK, ERR_INVALID_ARGUMENT, ERR_IO, ERR_FORMAT, ERR_DEVICE, ERR_OUT_OF_MEMORY, \
    ERR_UNSUPPORTED, ERR_STATE, ERR_NOT_FOUND = range(9)
# nwprt_precision
PREC_FP32, PREC_FP16, PREC_BF16, PREC_FP8 = range(4)
PRECISIONS = {"fp32": PREC_FP32, "fp16": PREC_FP16, "bf16": PREC_BF16, "fp8": PREC_FP8}
# nwprt_dtype
DTYPE_F32, DTYPE_F16 = 0, 1
# nwprt_tonemap_source
TONEMAP_FULL, TONEMAP_INDIRECT, TONEMAP_NONE = 0, 1, 2
HOST, DEVICE = 0, 1


class NwprtError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"nwprt status {status}: {message}")
        self.status = status


class HashgridDesc(Structure):
    _fields_ = [("n_levels", c_int32), ("n_features_per_level", c_int32),
                ("log2_hashmap_size", c_int32), ("base_resolution", c_int32),
                ("per_level_scale", c_float)]


class HashgridLevels(Structure):
    _fields_ = [("n_levels", c_int32), ("scale", c_float * MAX_LEVELS),
                ("resolution", c_uint32 * MAX_LEVELS), ("dense", c_uint32 * MAX_LEVELS),
                ("size", c_uint32 * MAX_LEVELS), ("offset", c_uint32 * (MAX_LEVELS + 1))]


class SceneDesc(Structure):
    _fields_ = [("grid", HashgridDesc), ("hash_dtype", c_int32), ("face_size", c_int32),
                ("feature_dim", c_int32), ("hidden_width", c_int32),
                ("n_hidden_layers", c_int32), ("has_projection", c_int32),
                ("output_softplus", c_int32), ("aux_dims", c_int32),
                ("aux_pad_value", c_float)]


class SceneParams(Structure):
    _fields_ = [("hash_table", c_void_p), ("wavelet_features", c_void_p),
                ("projection", c_void_p), ("layers", c_void_p)]


class GBuffer(Structure):
    _fields_ = [("position", c_void_p), ("normal", c_void_p), ("reflect", c_void_p),
                ("roughness", c_void_p), ("kd", c_void_p), ("ks", c_void_p),
                ("direct", c_void_p), ("env_flag", c_void_p)]


class FrameOut(Structure):
    _fields_ = [("indirect", c_void_p), ("indirect_with_direct", c_void_p),
                ("rgba8", c_void_p)]


class BatchJob(Structure):
    _fields_ = [("scene_id", c_int32), ("env_id", c_int32), ("gbuffer", GBuffer),
                ("width", c_int32), ("height", c_int32), ("out", FrameOut)]


ENV_RESIDENT, ENV_ROTATED, ENV_FRESH = range(3)


class ServerConfig(Structure):
    _fields_ = [("memory_budget_bytes", c_uint64), ("max_active", c_int32),
                ("tick_pairs", c_int64), ("precision", c_int), ("tonemap", c_int),
                ("variant", c_int32)]


class Request(Structure):
    _fields_ = [("scene_id", c_int32), ("gbuffer", GBuffer), ("width", c_int32),
                ("height", c_int32), ("out", FrameOut), ("env_source", c_int),
                ("env_id", c_int32), ("faces", c_void_p), ("faces_location", c_int),
                ("face_size", c_int32), ("k", c_int32), ("yaw", c_double)]


class ServerStats(Structure):
    _fields_ = [("submitted", c_int64), ("completed", c_int64), ("queued", c_int64),
                ("active", c_int64), ("ticks", c_int64), ("items", c_int64),
                ("bytes_in_use", c_uint64), ("peak_bytes", c_uint64), ("peak_active", c_int32),
                ("failed", c_int64)]


class DenoiseParams(Structure):
    _fields_ = [("iterations", c_int32), ("sigma_color", c_float), ("sigma_normal", c_float),
                ("sigma_albedo", c_float)]


class Camera(Structure):
    _fields_ = [("from_", c_float * 3), ("at", c_float * 3), ("up", c_float * 3)]


class Material(Structure):
    _fields_ = [("kd", c_float * 3), ("ks", c_float * 3), ("roughness", c_float),
                ("type", c_int32), ("specular_transmission", c_float), ("metallic", c_float),
                ("subsurface", c_float), ("specular", c_float), ("specular_tint", c_float),
                ("anisotropic", c_float), ("sheen", c_float), ("sheen_tint", c_float),
                ("clearcoat", c_float), ("clearcoat_gloss", c_float), ("eta", c_float)]


# Material rows as arrays: kd[3], ks[3], roughness, type, then the
# DisneyBSDF parameters and eta in the struct's order.  Rows of 8 take
# Material::init's values for the rest (Renderer/Material.h:79-97).
MATERIAL_ROW = 19
MATERIAL_TAIL_DEFAULTS = (0.0, 0.0, 0.0, 0.5, 0.0, 0.0, 0.0, 0.5, 0.0, 1.0,
                          float(np.float32(1.49) / np.float32(1.000277)))


def material_row(m) -> np.ndarray:
    """A material row of 8 or 19 values as the full 19 (float64)."""
    r = np.asarray(m, np.float64).reshape(-1)
    if r.size == 8:
        r = np.concatenate([r, MATERIAL_TAIL_DEFAULTS])
    if r.size != MATERIAL_ROW:
        raise ValueError(f"a material row has 8 or {MATERIAL_ROW} values, not {r.size}")
    return r


class MeshInfo(Structure):
    _fields_ = [("lower", c_float * 3), ("upper", c_float * 3), ("triangles", c_int64),
                ("bvh_nodes", c_int64), ("bvh_depth", c_int32), ("vertices", c_int64)]


class CastParams(Structure):
    _fields_ = [("max_bounce", c_int32), ("bounds", c_void_p), ("view_env_faces", c_void_p),
                ("view_env_n", c_int32), ("light_env_faces", c_void_p), ("light_env_n", c_int32),
                ("direct_samples", c_int32), ("seed", c_uint32)]


MATERIAL_MIRROR = 3

GBUFFER_FIELDS = {
    "position": (np.float32, 4), "normal": (np.float32, 4), "reflect": (np.float32, 4),
    "roughness": (np.float32, 1), "kd": (np.float32, 4), "ks": (np.float32, 4),
    "direct": (np.float32, 4), "env_flag": (np.uint8, 1),
}

# Every symbol include/nwprt.h declares, with (restype, argtypes).
_SIGNATURES = {
    "nwprt_abi_version": (c_int32, []),
    "nwprt_last_error": (c_char_p, []),
    "nwprt_status_string": (c_char_p, [c_int]),
    "nwprt_create": (c_int, [c_int, POINTER(c_void_p)]),
    "nwprt_destroy": (c_int, [c_void_p]),
    "nwprt_sync": (c_int, [c_void_p]),
    "nwprt_get_stream": (c_int, [c_void_p, POINTER(c_void_p)]),
    "nwprt_hashgrid_level_table": (c_int, [POINTER(HashgridDesc), POINTER(HashgridLevels)]),
    "nwprt_scene_param_counts": (c_int, [POINTER(SceneDesc), POINTER(c_size_t), POINTER(c_size_t),
                                         POINTER(c_size_t), POINTER(c_size_t)]),
    "nwprt_synth_scene": (c_int, [POINTER(SceneDesc), c_uint64, c_float, c_float, c_void_p,
                                  c_void_p, c_void_p, c_void_p]),
    "nwprt_synth_gbuffer": (c_int, [c_uint64, c_int32, c_int32, c_float, POINTER(GBuffer)]),
    "nwprt_synth_cubemap": (c_int, [c_uint64, c_int32, c_void_p]),
    "nwprt_free_host": (None, [c_void_p]),
    "nwprt_npy_read_f32": (c_int, [c_char_p, POINTER(c_void_p), POINTER(c_int64),
                                   POINTER(c_int32), POINTER(c_int32)]),
    "nwprt_npy_read_i32": (c_int, [c_char_p, POINTER(c_void_p), POINTER(c_int64),
                                   POINTER(c_int32), POINTER(c_int32)]),
    "nwprt_npy_write_f32": (c_int, [c_char_p, c_void_p, POINTER(c_int64), c_int32]),
    "nwprt_npy_write_i32": (c_int, [c_char_p, c_void_p, POINTER(c_int64), c_int32]),
    "nwprt_hdr_read": (c_int, [c_char_p, POINTER(c_void_p), POINTER(c_int32), POINTER(c_int32)]),
    "nwprt_hdr_write": (c_int, [c_char_p, c_void_p, c_int32, c_int32, c_int32]),
    "nwprt_cubemap_cross_to_faces": (c_int, [c_void_p, c_int32, c_int32, c_void_p]),
    "nwprt_scene_create": (c_int, [c_void_p, POINTER(SceneDesc), POINTER(SceneParams),
                                   POINTER(c_int32)]),
    "nwprt_scene_load_folder": (c_int, [c_void_p, c_char_p, POINTER(c_int32)]),
    "nwprt_scene_destroy": (c_int, [c_void_p, c_int32]),
    "nwprt_scene_get_desc": (c_int, [c_void_p, c_int32, POINTER(SceneDesc)]),
    "nwprt_envmap_from_faces": (c_int, [c_void_p, c_void_p, c_int, c_int32, c_int32,
                                        POINTER(c_int32)]),
    "nwprt_envmap_update_faces": (c_int, [c_void_p, c_int32, c_void_p, c_int, c_void_p]),
    "nwprt_envmap_from_selection": (c_int, [c_void_p, c_void_p, c_void_p, c_int32, c_int32,
                                            POINTER(c_int32)]),
    "nwprt_envmap_load_folder": (c_int, [c_void_p, c_char_p, c_char_p, c_int32,
                                         POINTER(c_int32)]),
    "nwprt_envmap_haar": (c_int, [c_void_p, c_void_p, c_int32, c_int32, c_void_p, c_void_p]),
    "nwprt_envmap_destroy": (c_int, [c_void_p, c_int32]),
    "nwprt_envmap_info": (c_int, [c_void_p, c_int32, POINTER(c_int32), POINTER(c_int32)]),
    "nwprt_envmap_get_selection": (c_int, [c_void_p, c_int32, c_void_p, c_void_p]),
    "nwprt_envmap_get_coefficients": (c_int, [c_void_p, c_int32, c_void_p]),
    "nwprt_relight": (c_int, [c_void_p, c_int32, c_int32, POINTER(GBuffer), c_int32, c_int32,
                              POINTER(FrameOut), c_int, c_int, c_void_p]),
    "nwprt_relight_noalloc": (c_int, [c_void_p, c_int32, c_int32, POINTER(GBuffer), c_int32,
                                      c_int32, POINTER(FrameOut), c_int, c_int, c_void_p]),
    "nwprt_pass_p1": (c_int, [c_void_p, c_int32, c_int32, POINTER(GBuffer), c_int32, c_int32,
                              c_int, c_void_p]),
    "nwprt_pass_p2": (c_int, [c_void_p, c_int32, c_int32, c_int32, POINTER(GBuffer), c_int32,
                              c_int32, POINTER(FrameOut), c_int, c_int, c_void_p]),
    "nwprt_relight_variant_count": (c_int32, []),
    "nwprt_relight_variant_name": (c_char_p, [c_int32]),
    "nwprt_relight_fp8_variant_count": (c_int32, []),
    "nwprt_relight_fp8_variant_name": (c_char_p, [c_int32]),
    "nwprt_set_relight_variant": (c_int, [c_void_p, c_int32]),
    "nwprt_autotune_relight": (c_int, [c_void_p, c_int32, c_int32, POINTER(GBuffer), c_int32,
                                       c_int32, c_int, POINTER(c_int32), POINTER(c_float)]),
    "nwprt_selected_relight_variant": (c_int, [c_void_p, c_int, c_int32, c_int32,
                                               POINTER(c_int32)]),
    "nwprt_tonemap": (c_int, [c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_void_p]),
    "nwprt_denoise_default_params": (None, [POINTER(DenoiseParams)]),
    "nwprt_denoise": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_int32,
                              c_float, POINTER(DenoiseParams), c_void_p]),
    "nwprt_mesh_create": (c_int, [c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_int64,
                                  c_void_p, POINTER(Material), c_int32, POINTER(c_int32)]),
    "nwprt_mesh_get_info": (c_int, [c_void_p, c_int32, POINTER(MeshInfo)]),
    "nwprt_obj_read": (c_int, [c_char_p, POINTER(c_void_p), POINTER(c_int64), POINTER(c_void_p),
                               POINTER(c_void_p), POINTER(c_int64), POINTER(c_void_p),
                               POINTER(c_void_p), POINTER(c_int32)]),
    "nwprt_obj_read_textured": (c_int, [c_char_p, POINTER(c_void_p), POINTER(c_int64),
                                        POINTER(c_void_p), POINTER(c_void_p), POINTER(c_void_p),
                                        POINTER(c_int64), POINTER(c_void_p), POINTER(c_void_p),
                                        POINTER(c_int32), POINTER(c_void_p), POINTER(c_int64)]),
    "nwprt_ply_read": (c_int, [c_char_p, POINTER(c_void_p), POINTER(c_int64), POINTER(c_void_p),
                               POINTER(c_void_p), POINTER(c_int64), POINTER(c_void_p),
                               POINTER(c_void_p), POINTER(c_int32)]),
    "nwprt_mesh_destroy": (c_int, [c_void_p, c_int32]),
    "nwprt_texture_create": (c_int, [c_void_p, c_void_p, c_int32, c_int32, POINTER(c_int32)]),
    "nwprt_texture_destroy": (c_int, [c_void_p, c_int32]),
    "nwprt_mesh_set_textures": (c_int, [c_void_p, c_int32, c_void_p, c_void_p]),
    "nwprt_gbuffer_cast": (c_int, [c_void_p, c_int32, POINTER(Camera), c_int32, c_int32,
                                   POINTER(CastParams), POINTER(GBuffer), c_void_p]),
    "nwprt_sh_render": (c_int, [c_void_p, c_int32, POINTER(Camera), c_int32, c_int32, c_void_p,
                                c_void_p, c_int32, c_int32, c_uint32, c_void_p, c_int32, c_void_p,
                                c_void_p]),
    "nwprt_sh_compute_transfer": (c_int, [c_void_p, c_int32, c_void_p, c_int32, c_int32, c_int32,
                                          c_int32, c_uint32, ctypes.c_int64, ctypes.c_int64,
                                          c_void_p, c_void_p]),
    "nwprt_material_defaults": (c_int, [c_int32, c_void_p]),
    "nwprt_image_read_rgba8": (c_int, [c_char_p, POINTER(POINTER(ctypes.c_uint8)),
                                       POINTER(c_int32), POINTER(c_int32)]),
    "nwprt_debug_hash_encode": (c_int, [c_void_p, c_int32, c_void_p, c_int32, c_void_p, c_void_p,
                                        c_void_p]),
    "nwprt_debug_transport": (c_int, [c_void_p, c_int32, c_int32, POINTER(GBuffer), c_int32,
                                      c_int32, c_int32, c_int32, c_void_p, c_int, c_void_p]),
    "nwprt_envmap_update_rotated": (c_int, [c_void_p, c_int32, c_void_p, c_double, c_void_p]),
    "nwprt_batch_create": (c_int, [c_void_p, POINTER(BatchJob), c_int32, c_int, c_int,
                                   POINTER(c_int32)]),
    "nwprt_batch_launch": (c_int, [c_void_p, c_int32, c_void_p]),
    "nwprt_batch_autotune": (c_int, [c_void_p, c_int32, POINTER(c_int32), POINTER(c_float)]),
    "nwprt_batch_info": (c_int, [c_void_p, c_int32, POINTER(c_int64), POINTER(c_int64),
                                 POINTER(c_int32)]),
    "nwprt_batch_destroy": (c_int, [c_void_p, c_int32]),
    "nwprt_server_create": (c_int, [c_void_p, POINTER(ServerConfig), POINTER(c_int32)]),
    "nwprt_server_submit": (c_int, [c_void_p, c_int32, POINTER(Request), POINTER(c_int64)]),
    "nwprt_server_poll": (c_int, [c_void_p, c_int32, c_int32, c_void_p, c_void_p, c_int32,
                                  POINTER(c_int32)]),
    "nwprt_server_get_stats": (c_int, [c_void_p, c_int32, POINTER(ServerStats)]),
    "nwprt_server_destroy": (c_int, [c_void_p, c_int32]),
}

_lib = None


def lib() -> ctypes.CDLL:
    """Load libnwprt.so (raises if it was not built: no fallback exists)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(
                f"{LIB_PATH} is missing: build it with `python -m neural_wavelet_prt_demo_amd.build`")
        l = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def _check(status: int) -> None:
    if status != OK:
        msg = lib().nwprt_last_error()
        raise NwprtError(status, msg.decode() if msg else "")


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(c_void_p)


# ------------------------------------------------------------- host side --

def config_scene_desc(face_size: int, hash_dtype: int = DTYPE_F16, *, levels: int = 16,
                      features: int = 2, log2_hashmap_size: int = 19, base_resolution: int = 16,
                      per_level_scale: float = 1.5, projection: bool = True,
                      softplus: bool = True, aux: bool = False) -> SceneDesc:
    """The network of the relighting workloads: 16 x 2 hash grid (2^19
    entries per level, N_min 16, growth 1.5), 32 x 32 projection, MLP
    32 -> 64 -> 64 -> 64 -> 3 with ReLU and a softplus output.  With `aux`,
    the 13 G-buffer encodings (+3 pads) also enter the second layer, as in
    the reference graph's cmp_net input."""
    d = SceneDesc()
    d.grid = HashgridDesc(levels, features, log2_hashmap_size, base_resolution, per_level_scale)
    d.hash_dtype = hash_dtype
    d.face_size = face_size
    d.feature_dim = levels * features
    d.hidden_width = 64
    d.n_hidden_layers = 3
    d.has_projection = int(projection)
    d.output_softplus = int(softplus)
    d.aux_dims = 13 if aux else 0
    d.aux_pad_value = 1.0
    return d


def hashgrid_levels(grid: HashgridDesc) -> Dict[str, np.ndarray]:
    lv = HashgridLevels()
    _check(lib().nwprt_hashgrid_level_table(ctypes.byref(grid), ctypes.byref(lv)))
    n = lv.n_levels
    return {"scale": np.array(lv.scale[:n], np.float32),
            "resolution": np.array(lv.resolution[:n], np.uint32),
            "dense": np.array(lv.dense[:n], np.uint32),
            "size": np.array(lv.size[:n], np.uint32),
            "offset": np.array(lv.offset[:n + 1], np.uint32)}


def scene_param_counts(desc: SceneDesc):
    h, w, p, l = c_size_t(), c_size_t(), c_size_t(), c_size_t()
    _check(lib().nwprt_scene_param_counts(ctypes.byref(desc), ctypes.byref(h), ctypes.byref(w),
                                          ctypes.byref(p), ctypes.byref(l)))
    return h.value, w.value, p.value, l.value


def synth_scene(desc: SceneDesc, seed: int, hash_std: float = 0.1,
                feature_std: float = 0.1) -> Dict[str, np.ndarray]:
    """Seeded scene parameters (host arrays)."""
    entries, nwf, nproj, nlay = scene_param_counts(desc)
    F = desc.grid.n_features_per_level
    hdt = np.uint16 if desc.hash_dtype == DTYPE_F16 else np.float32
    out = {"hash_table": np.empty((entries, F), hdt),
           "wavelet_features": np.empty((desc.feature_dim, nwf // desc.feature_dim), np.float32),
           "projection": np.empty((desc.feature_dim, desc.feature_dim), np.float32)
           if nproj else None,
           "layers": np.empty(nlay, np.float32)}
    _check(lib().nwprt_synth_scene(ctypes.byref(desc), seed, hash_std, feature_std,
                                   _ptr(out["hash_table"]), _ptr(out["wavelet_features"]),
                                   _ptr(out["projection"]), _ptr(out["layers"])))
    return out


def synth_gbuffer(seed: int, width: int, height: int,
                  background_fraction: float = 0.0) -> Dict[str, np.ndarray]:
    """Seeded G-buffer (host arrays in the reference layout)."""
    P = width * height
    g = {k: np.empty((P, c) if c > 1 else (P,), dt) for k, (dt, c) in GBUFFER_FIELDS.items()}
    gb = GBuffer(**{k: _ptr(v) for k, v in g.items()})
    _check(lib().nwprt_synth_gbuffer(seed, width, height, background_fraction, ctypes.byref(gb)))
    return g


def obj_read(path: str) -> Dict[str, np.ndarray]:
    """Wavefront OBJ (+ MTL) as mesh arrays: vertices [n][3], normals [n][3]
    or None, indices [t][3], material_ids [t], materials [m][19] rows
    (kd[3], ks[3], roughness, type, the DisneyBSDF parameters, eta)."""
    return _mesh_read(lib().nwprt_obj_read, path)


def obj_read_textured(path: str) -> Dict[str, np.ndarray]:
    """obj_read plus texcoords [n][2] (or None) and maps: per material the
    kd / ks / roughness map files (map_Kd, map_Ks, map_Ns; "" for none)."""
    v, nr, tc, ix, mid, mats, mp = (c_void_p() for _ in range(7))
    nv, nt, nm, nb = c_int64(), c_int64(), c_int32(), c_int64()
    _check(lib().nwprt_obj_read_textured(str(path).encode(), ctypes.byref(v), ctypes.byref(nv),
                                         ctypes.byref(nr), ctypes.byref(tc), ctypes.byref(ix),
                                         ctypes.byref(nt), ctypes.byref(mid), ctypes.byref(mats),
                                         ctypes.byref(nm), ctypes.byref(mp), ctypes.byref(nb)))
    out = {"vertices": _take_host(v, nv.value * 3, np.float32).reshape(-1, 3),
           "normals": (_take_host(nr, nv.value * 3, np.float32).reshape(-1, 3)
                       if nr.value else None),
           "texcoords": (_take_host(tc, nv.value * 2, np.float32).reshape(-1, 2)
                         if tc.value else None),
           "indices": _take_host(ix, nt.value * 3, np.int32).reshape(-1, 3),
           "material_ids": _take_host(mid, nt.value, np.int32)}
    raw = _take_host(mats, nm.value * MATERIAL_ROW, np.float32).reshape(-1, MATERIAL_ROW)
    out["materials"] = np.concatenate(
        [raw[:, :7], raw[:, 7:8].view(np.int32).astype(np.float32), raw[:, 8:]], axis=1)
    names = _take_host(mp, nb.value, np.uint8).tobytes().split(b"\0")[:-1]
    out["maps"] = [[n.decode() for n in names[3 * i:3 * i + 3]] for i in range(nm.value)]
    return out


def ply_read(path: str) -> Dict[str, np.ndarray]:
    """Stanford PLY as the same mesh arrays as obj_read (one default
    cookTorrance material)."""
    return _mesh_read(lib().nwprt_ply_read, path)


def _mesh_read(fn, path) -> Dict[str, np.ndarray]:
    v, nr, ix, mid, mats = (c_void_p() for _ in range(5))
    nv, nt, nm = c_int64(), c_int64(), c_int32()
    _check(fn(str(path).encode(), ctypes.byref(v), ctypes.byref(nv),
                                ctypes.byref(nr), ctypes.byref(ix), ctypes.byref(nt),
                                ctypes.byref(mid), ctypes.byref(mats), ctypes.byref(nm)))
    out = {"vertices": _take_host(v, nv.value * 3, np.float32).reshape(-1, 3),
           "normals": (_take_host(nr, nv.value * 3, np.float32).reshape(-1, 3)
                       if nr.value else None),
           "indices": _take_host(ix, nt.value * 3, np.int32).reshape(-1, 3),
           "material_ids": _take_host(mid, nt.value, np.int32)}
    raw = _take_host(mats, nm.value * MATERIAL_ROW, np.float32).reshape(-1, MATERIAL_ROW)
    # The struct's eighth field is an int32 type: reinterpret that column.
    out["materials"] = np.concatenate(
        [raw[:, :7], raw[:, 7:8].view(np.int32).astype(np.float32), raw[:, 8:]], axis=1)
    return out


def synth_cubemap(seed: int, face_size: int) -> np.ndarray:
    faces = np.empty((6, face_size, face_size, 3), np.float32)
    _check(lib().nwprt_synth_cubemap(seed, face_size, _ptr(faces)))
    return faces


def _take_host(ptr: c_void_p, count: int, dtype) -> np.ndarray:
    try:
        buf = (ctypes.c_char * (count * np.dtype(dtype).itemsize)).from_address(ptr.value)
        return np.frombuffer(bytes(buf), dtype=dtype).copy()
    finally:
        lib().nwprt_free_host(ptr)


def npy_read(path: str, dtype=np.float32):
    data = c_void_p()
    shape = (c_int64 * 8)()
    nd, fo = c_int32(), c_int32()
    fn = lib().nwprt_npy_read_f32 if dtype == np.float32 else lib().nwprt_npy_read_i32
    _check(fn(str(path).encode(), ctypes.byref(data), shape, ctypes.byref(nd), ctypes.byref(fo)))
    shp = tuple(shape[i] for i in range(nd.value))
    arr = _take_host(data, int(np.prod(shp)) if shp else 1, dtype)
    return arr, shp, bool(fo.value)


def npy_write(path: str, a: np.ndarray) -> None:
    a = np.ascontiguousarray(a)
    shape = (c_int64 * max(1, a.ndim))(*a.shape)
    if a.dtype == np.float32:
        _check(lib().nwprt_npy_write_f32(str(path).encode(), _ptr(a), shape, a.ndim))
    elif a.dtype == np.int32:
        _check(lib().nwprt_npy_write_i32(str(path).encode(), _ptr(a), shape, a.ndim))
    else:
        raise TypeError("npy_write supports float32 and int32")


def hdr_read(path: str) -> np.ndarray:
    data = c_void_p()
    w, h = c_int32(), c_int32()
    _check(lib().nwprt_hdr_read(str(path).encode(), ctypes.byref(data), ctypes.byref(w),
                                ctypes.byref(h)))
    return _take_host(data, w.value * h.value * 3, np.float32).reshape(h.value, w.value, 3)


def hdr_write(path: str, img: np.ndarray) -> None:
    img = np.ascontiguousarray(img, np.float32)
    _check(lib().nwprt_hdr_write(str(path).encode(), _ptr(img), img.shape[1], img.shape[0],
                                 img.shape[2]))


def material_defaults(mat_type: int) -> np.ndarray:
    """The reference's default material of a type (nwprt_material_defaults)
    as a 19-value row."""
    m = Material()
    _check(lib().nwprt_material_defaults(mat_type, ctypes.byref(m)))
    return np.array(list(m.kd) + list(m.ks) + [m.roughness, m.type, m.specular_transmission,
                    m.metallic, m.subsurface, m.specular, m.specular_tint, m.anisotropic,
                    m.sheen, m.sheen_tint, m.clearcoat, m.clearcoat_gloss, m.eta], np.float64)


# Material texture fields (NWPRT_TEX_*).
TEX_FIELD_NAMES = ("kd", "ks", "roughness", "specular_transmission", "metallic", "subsurface",
                   "specular", "specular_tint", "anisotropic", "sheen", "sheen_tint", "clearcoat",
                   "clearcoat_gloss", "eta")
TEX_FIELDS = len(TEX_FIELD_NAMES)


def image_read_rgba8(path: str) -> np.ndarray:
    """An 8-bit texture image (nwprt_image_read_rgba8) as uint8
    [height][width][4], row 0 at the bottom of the picture."""
    p = POINTER(ctypes.c_uint8)()
    w, h = c_int32(), c_int32()
    _check(lib().nwprt_image_read_rgba8(str(path).encode(), ctypes.byref(p), ctypes.byref(w),
                                        ctypes.byref(h)))
    try:
        return np.ctypeslib.as_array(p, shape=(h.value, w.value, 4)).copy()
    finally:
        lib().nwprt_free_host(p)


def cross_to_faces(cross: np.ndarray) -> np.ndarray:
    cross = np.ascontiguousarray(cross, np.float32)
    h, w = cross.shape[:2]
    faces = np.empty((6, h // 3, h // 3, 3), np.float32)
    _check(lib().nwprt_cubemap_cross_to_faces(_ptr(cross), w, h, _ptr(faces)))
    return faces


def _natural_key(name: str):
    """compareNat (Renderer/files.h:27-58): digit runs compare as numbers
    and sort before letters; letters compare case-insensitively."""
    key = []
    for part in re.split(r"(\d+)", name):
        if not part:
            continue
        if part.isdigit():
            key.append((0, int(part), ""))
        else:
            key.extend((1, 0, ch.upper()) for ch in part)
    return key


def sh_eval_real(dirs: np.ndarray, degree: int = 8) -> np.ndarray:
    """pbrt's real SH basis (Renderer/device_functions/shaderUtils.h:230-337)
    at unit directions dirs [..., 3]: [..., (degree + 1)^2] float64, index
    l (l + 1) + m; associated Legendre functions with the Condon-Shortley
    phase, sqrt(2) K P_l^|m| sin(|m| phi) for m < 0, K P_l^0 for m = 0,
    sqrt(2) K P_l^m cos(m phi) for m > 0."""
    d = np.asarray(dirs, np.float64)
    x, y, z = d[..., 0], d[..., 1], d[..., 2]
    L = degree
    s = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    P = {}
    P[(0, 0)] = np.ones_like(z)
    for l in range(1, L + 1):
        P[(l, l)] = -(2 * l - 1) * s * P[(l - 1, l - 1)]
    for l in range(1, L + 1):
        P[(l, l - 1)] = (2 * l - 1) * z * P[(l - 1, l - 1)]
    for l in range(2, L + 1):
        for m in range(0, l - 1):
            P[(l, m)] = ((2 * l - 1) * z * P[(l - 1, m)] - (l - 1 + m) * P[(l - 2, m)]) / (l - m)
    phi = np.arctan2(y, x)
    out = np.empty(d.shape[:-1] + ((L + 1) ** 2,), np.float64)
    for l in range(L + 1):
        for m in range(-l, l + 1):
            am = abs(m)
            k = np.sqrt((2 * l + 1) / (4 * np.pi) * np.prod(
                np.arange(l - am + 1, l + am + 1, dtype=np.float64)) ** -1.0)
            if m < 0:
                v = np.sqrt(2.0) * k * P[(l, am)] * np.sin(am * phi)
            elif m == 0:
                v = k * P[(l, 0)]
            else:
                v = np.sqrt(2.0) * k * P[(l, m)] * np.cos(m * phi)
            out[..., l * l + l + m] = v
    return out


def cube_texel_dirs(n: int) -> np.ndarray:
    """Unit directions of the texel centres of [6][n][n] cube faces (face
    order +X, -X, +Y, -Y, -Z, +Z; EnvUVToDirec, Renderer/device_functions/
    cubeMap.cu:94-110, at ((x + 0.5) / n, (y + 0.5) / n)): [6][n][n][3]."""
    g = (np.arange(n) + 0.5) / n * 2.0 - 1.0
    vc, uc = np.meshgrid(g, g, indexing="ij")
    one = np.ones_like(uc)
    faces = [(one, vc, uc), (-one, vc, -uc), (uc, one, vc), (uc, -one, -vc), (uc, vc, -one),
             (-uc, vc, one)]
    d = np.stack([np.stack(f, -1) for f in faces])
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


def sh_basis_faces(degree: int, n: int) -> np.ndarray:
    """Basis maps for the transfer computation: map k holds the real SH
    basis function Y_k at every texel centre, in all three channels:
    float32 [(degree + 1)^2][6][n][n][3]."""
    y = sh_eval_real(cube_texel_dirs(n), degree)          # [6][n][n][nc]
    return np.ascontiguousarray(np.repeat(np.moveaxis(y, -1, 0)[..., None], 3, -1),
                                np.float32)


def load_sh_basis_folder(folder: str, degree: int) -> np.ndarray:
    """The basis maps of SHRenderer (Renderer/SHRenderer.cpp:12-22): the
    .npy cubemap crosses of `folder` in natural order, each read as
    Environment_Map reads one (EnvironmentMap.cpp:90-103, rows top first);
    returns the first (degree + 1)^2 as float32 [nc][6][n][n][3]."""
    nc = (degree + 1) ** 2
    names = sorted((f for f in os.listdir(folder) if f.endswith(".npy")),
                   key=lambda f: _natural_key(os.path.join(folder, f)))
    if len(names) < nc:
        raise NwprtError(ERR_INVALID_ARGUMENT,
                         f"{folder}: {len(names)} basis maps, degree {degree} needs {nc}")
    maps = []
    for f in names[:nc]:
        data, shape, _ = npy_read(os.path.join(folder, f))
        maps.append(cross_to_faces(data.reshape(shape)))
    return np.stack(maps)


# ------------------------------------------------------------ device side --

class DeviceFrame:
    """G-buffer and outputs of one frame in device memory (reference layout,
    Renderer/frames.h:143-183)."""

    def __init__(self, width: int, height: int, host: Optional[Dict[str, np.ndarray]] = None):
        self.width, self.height = width, height
        P = width * height
        self.bufs = {k: DeviceBuffer(P * np.dtype(dt).itemsize * c)
                     for k, (dt, c) in GBUFFER_FIELDS.items()}
        self.out = {"indirect": DeviceBuffer(P * 16), "indirect_with_direct": DeviceBuffer(P * 16),
                    "rgba8": DeviceBuffer(P * 4)}
        if host is not None:
            self.upload(host)

    def upload(self, host: Dict[str, np.ndarray]) -> None:
        for k, v in host.items():
            self.bufs[k].upload(v)

    def gbuffer(self) -> GBuffer:
        return GBuffer(**{k: b.ptr for k, b in self.bufs.items()})

    def frame_out(self, indirect=True, with_direct=True, rgba8=True) -> FrameOut:
        return FrameOut(self.out["indirect"].ptr if indirect else None,
                        self.out["indirect_with_direct"].ptr if with_direct else None,
                        self.out["rgba8"].ptr if rgba8 else None)

    def results(self) -> Dict[str, np.ndarray]:
        P = self.width * self.height
        return {"indirect": self.out["indirect"].download(np.float32, (P, 4)),
                "indirect_with_direct": self.out["indirect_with_direct"].download(np.float32, (P, 4)),
                "rgba8": self.out["rgba8"].download(np.uint32, (P,))}


class Context:
    """One nwprt context (device, stream, resident scenes and env maps)."""

    def __init__(self, device: int = 0):
        self.h = c_void_p()
        _check(lib().nwprt_create(device, ctypes.byref(self.h)))

    def close(self) -> None:
        if self.h:
            lib().nwprt_destroy(self.h)
            self.h = c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self) -> int:
        s = c_void_p()
        _check(lib().nwprt_get_stream(self.h, ctypes.byref(s)))
        return s.value or 0

    def sync(self) -> None:
        _check(lib().nwprt_sync(self.h))

    # scenes -------------------------------------------------------------
    def scene_create(self, desc: SceneDesc, params: Dict[str, np.ndarray]) -> int:
        keep = {k: (None if v is None else np.ascontiguousarray(v)) for k, v in params.items()}
        p = SceneParams(_ptr(keep["hash_table"]), _ptr(keep["wavelet_features"]),
                        _ptr(keep.get("projection")), _ptr(keep["layers"]))
        sid = c_int32()
        _check(lib().nwprt_scene_create(self.h, ctypes.byref(desc), ctypes.byref(p),
                                        ctypes.byref(sid)))
        return sid.value

    def initTCNN(self, model_dir: str) -> int:
        """RenderEngine::initTCNN (Renderer/tinycudann/tcnn.cu:49-81)."""
        sid = c_int32()
        _check(lib().nwprt_scene_load_folder(self.h, str(model_dir).encode(), ctypes.byref(sid)))
        return sid.value

    def scene_desc(self, scene: int) -> SceneDesc:
        d = SceneDesc()
        _check(lib().nwprt_scene_get_desc(self.h, scene, ctypes.byref(d)))
        return d

    def scene_destroy(self, scene: int) -> None:
        _check(lib().nwprt_scene_destroy(self.h, scene))

    # environment maps ----------------------------------------------------
    def envmap_from_faces(self, faces, k: int, face_size: Optional[int] = None) -> int:
        eid = c_int32()
        if isinstance(faces, DeviceBuffer):
            _check(lib().nwprt_envmap_from_faces(self.h, faces.ptr, DEVICE, face_size, k,
                                                 ctypes.byref(eid)))
        else:
            faces = np.ascontiguousarray(faces, np.float32)
            _check(lib().nwprt_envmap_from_faces(self.h, _ptr(faces), HOST, faces.shape[1], k,
                                                 ctypes.byref(eid)))
        return eid.value

    def envmap_update_faces(self, env: int, faces, stream: Optional[int] = None) -> None:
        """Recompute a faces-built map from new faces in place (no allocation)."""
        s = c_void_p(stream) if stream else None
        if isinstance(faces, DeviceBuffer):
            _check(lib().nwprt_envmap_update_faces(self.h, env, faces.ptr, DEVICE, s))
        else:
            faces = np.ascontiguousarray(faces, np.float32)
            _check(lib().nwprt_envmap_update_faces(self.h, env, _ptr(faces), HOST, s))

    def envmap_from_selection(self, indices: np.ndarray, strengths: np.ndarray,
                              face_size: int) -> int:
        idx = np.ascontiguousarray(indices, np.int32)
        st = np.ascontiguousarray(strengths, np.float32)
        eid = c_int32()
        _check(lib().nwprt_envmap_from_selection(self.h, _ptr(idx), _ptr(st), len(idx), face_size,
                                                 ctypes.byref(eid)))
        return eid.value

    def envmap_load_folder(self, folder: str, name: str, k: int) -> int:
        """Environment_Map(folder, name) (Renderer/EnvironmentMap.cpp:20-68)."""
        eid = c_int32()
        _check(lib().nwprt_envmap_load_folder(self.h, str(folder).encode(), str(name).encode(), k,
                                              ctypes.byref(eid)))
        return eid.value

    def envmap_info(self, env: int):
        n, k = c_int32(), c_int32()
        _check(lib().nwprt_envmap_info(self.h, env, ctypes.byref(n), ctypes.byref(k)))
        return n.value, k.value

    def envmap_selection(self, env: int):
        n, k = self.envmap_info(env)
        idx = np.empty(k, np.int32)
        st = np.empty((k, 3), np.float32)
        _check(lib().nwprt_envmap_get_selection(self.h, env, _ptr(idx), _ptr(st)))
        return idx, st

    def envmap_coefficients(self, env: int) -> np.ndarray:
        n, _ = self.envmap_info(env)
        c = np.empty((6, n, n, 3), np.float32)
        _check(lib().nwprt_envmap_get_coefficients(self.h, env, _ptr(c)))
        return c

    def envmap_update_rotated(self, env: int, faces: DeviceBuffer, yaw: float,
                              stream: Optional[int] = None) -> None:
        """Recompute a faces-built map from `faces` (device) turned by `yaw`
        radians about +y."""
        _check(lib().nwprt_envmap_update_rotated(self.h, env, faces.ptr, float(yaw),
                                                 c_void_p(stream) if stream else None))

    def envmap_haar(self, faces: "DeviceBuffer", face_size: int, n_maps: int,
                    coef: "DeviceBuffer", stream: Optional[int] = None) -> None:
        """Haar transform of n_maps device cubemaps into `coef` (same layout)."""
        _check(lib().nwprt_envmap_haar(self.h, faces.ptr, face_size, n_maps, coef.ptr,
                                       c_void_p(stream) if stream else None))

    def envmap_destroy(self, env: int) -> None:
        _check(lib().nwprt_envmap_destroy(self.h, env))

    # relight -------------------------------------------------------------
    def relight(self, scene: int, env: int, frame: DeviceFrame, precision: int = PREC_BF16,
                tonemap: int = TONEMAP_FULL, stream: Optional[int] = None,
                out: Optional[FrameOut] = None) -> None:
        gb = frame.gbuffer()
        fo = out if out is not None else frame.frame_out()
        _check(lib().nwprt_relight(self.h, scene, env, ctypes.byref(gb), frame.width,
                                   frame.height, ctypes.byref(fo), precision, tonemap,
                                   c_void_p(stream) if stream else None))

    def relight_noalloc(self, scene: int, env: int, frame: DeviceFrame,
                        precision: int = PREC_BF16, tonemap: int = TONEMAP_FULL,
                        stream: Optional[int] = None, out: Optional[FrameOut] = None) -> None:
        """nwprt_relight_noalloc: the fused-encode relight with the library's
        per-device scratch array as its only work memory."""
        gb = frame.gbuffer()
        fo = out if out is not None else frame.frame_out()
        _check(lib().nwprt_relight_noalloc(self.h, scene, env, ctypes.byref(gb), frame.width,
                                           frame.height, ctypes.byref(fo), precision, tonemap,
                                           c_void_p(stream) if stream else None))

    def TCNNPass_p1(self, scene: int, slot: int, frame: DeviceFrame,
                    precision: int = PREC_BF16, stream: Optional[int] = None) -> None:
        gb = frame.gbuffer()
        _check(lib().nwprt_pass_p1(self.h, scene, slot, ctypes.byref(gb), frame.width,
                                   frame.height, precision, c_void_p(stream) if stream else None))

    def TCNNPass_p2(self, scene: int, env: int, slot: int, frame: DeviceFrame,
                    precision: int = PREC_BF16, tonemap: int = TONEMAP_FULL,
                    stream: Optional[int] = None) -> None:
        gb = frame.gbuffer()
        fo = frame.frame_out()
        _check(lib().nwprt_pass_p2(self.h, scene, env, slot, ctypes.byref(gb), frame.width,
                                   frame.height, ctypes.byref(fo), precision, tonemap,
                                   c_void_p(stream) if stream else None))

    # tile tuning ---------------------------------------------------------
    @staticmethod
    def relight_variants():
        n = lib().nwprt_relight_variant_count()
        return [lib().nwprt_relight_variant_name(i).decode() for i in range(n)]

    @staticmethod
    def relight_fp8_variants():
        n = lib().nwprt_relight_fp8_variant_count()
        return [lib().nwprt_relight_fp8_variant_name(i).decode() for i in range(n)]

    def set_relight_variant(self, variant: int) -> None:
        _check(lib().nwprt_set_relight_variant(self.h, variant))

    def autotune_relight(self, scene: int, env: int, frame: DeviceFrame,
                         precision: int = PREC_BF16):
        gb = frame.gbuffer()
        best, ms = c_int32(), c_float()
        _check(lib().nwprt_autotune_relight(self.h, scene, env, ctypes.byref(gb), frame.width,
                                            frame.height, precision, ctypes.byref(best),
                                            ctypes.byref(ms)))
        return best.value, ms.value

    def selected_relight_variant(self, precision: int, k: int, n_pixels: int) -> int:
        v = c_int32()
        _check(lib().nwprt_selected_relight_variant(self.h, precision, k, n_pixels,
                                                    ctypes.byref(v)))
        return v.value

    def computeFinalPixelColors(self, src: DeviceBuffer, dst: DeviceBuffer, width: int,
                                height: int, stream: Optional[int] = None) -> None:
        _check(lib().nwprt_tonemap(self.h, src.ptr, dst.ptr, width, height,
                                   c_void_p(stream) if stream else None))

    # denoiser ------------------------------------------------------------
    @staticmethod
    def denoise_default_params() -> DenoiseParams:
        p = DenoiseParams()
        lib().nwprt_denoise_default_params(ctypes.byref(p))
        return p

    def denoise(self, color, normal, albedo, out, width: int, height: int,
                blend: float = 0.0, params: Optional[DenoiseParams] = None,
                stream: Optional[int] = None) -> None:
        """OptixBackend::denoise (Renderer/OptixBackend.cpp:841-906) on device
        float4 images (DeviceBuffer or raw device address)."""
        ptr = lambda b: b.ptr if isinstance(b, DeviceBuffer) else b
        _check(lib().nwprt_denoise(self.h, ptr(color), ptr(normal), ptr(albedo), ptr(out),
                                   width, height, blend,
                                   ctypes.byref(params) if params is not None else None,
                                   c_void_p(stream) if stream else None))

    # G-buffer producer ---------------------------------------------------
    def mesh_create(self, vertices: np.ndarray, indices: np.ndarray, materials,
                    normals: Optional[np.ndarray] = None,
                    material_ids: Optional[np.ndarray] = None) -> int:
        """Uploads a world-space triangle mesh and builds its BVH.  materials:
        sequence of Material or rows (kd[3], ks[3], roughness, type[, the
        DisneyBSDF parameters, eta]) of 8 or 19 values."""
        v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
        ix = np.ascontiguousarray(indices, np.int32).reshape(-1, 3)
        nr = None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
        mid = None if material_ids is None else np.ascontiguousarray(material_ids, np.int32)
        mats = (Material * len(materials))()
        for i, m in enumerate(materials):
            if isinstance(m, Material):
                mats[i] = m
            else:
                m = material_row(m)
                mats[i] = Material((c_float * 3)(*m[0:3]), (c_float * 3)(*m[3:6]), float(m[6]),
                                   int(m[7]), *[float(x) for x in m[8:]])
        out = c_int32()
        _check(lib().nwprt_mesh_create(self.h, _ptr(v), len(v), _ptr(nr), _ptr(ix), len(ix),
                                       _ptr(mid), mats, len(materials), ctypes.byref(out)))
        return out.value

    def mesh_info(self, mesh: int) -> Dict[str, object]:
        info = MeshInfo()
        _check(lib().nwprt_mesh_get_info(self.h, mesh, ctypes.byref(info)))
        return {"lower": list(info.lower), "upper": list(info.upper),
                "triangles": info.triangles, "bvh_nodes": info.bvh_nodes,
                "bvh_depth": info.bvh_depth, "vertices": info.vertices}

    def mesh_destroy(self, mesh: int) -> None:
        _check(lib().nwprt_mesh_destroy(self.h, mesh))

    def texture_create(self, rgba: np.ndarray) -> int:
        """A material texture (nwprt_texture_create) from uint8 [height]
        [width][4] texels, row 0 at v = 0 (as image_read_rgba8 returns
        them)."""
        t = np.ascontiguousarray(rgba, np.uint8)
        assert t.ndim == 3 and t.shape[2] == 4, t.shape
        tid = c_int32()
        _check(lib().nwprt_texture_create(self.h, t.ctypes.data, t.shape[1], t.shape[0],
                                          ctypes.byref(tid)))
        return tid.value

    def texture_destroy(self, texture: int) -> None:
        _check(lib().nwprt_texture_destroy(self.h, texture))

    def mesh_set_textures(self, mesh: int, texcoords: Optional[np.ndarray],
                          material_textures: Optional[np.ndarray]) -> None:
        """texcoords: [n_vertices][2]; material_textures: int [n_materials]
        [TEX_FIELDS] texture ids or -1 (None for either: untextured)."""
        tc = None if texcoords is None else np.ascontiguousarray(texcoords, np.float32)
        mt = None if material_textures is None else np.ascontiguousarray(material_textures,
                                                                         np.int32)
        _check(lib().nwprt_mesh_set_textures(self.h, mesh, None if tc is None else tc.ctypes.data,
                                             None if mt is None else mt.ctypes.data))

    def gbuffer_cast(self, mesh: int, camera, frame, *, max_bounce: int = 1, bounds=None,
                     env_faces=None, env_n: int = 0, light_faces=None, light_n: int = 0,
                     direct_samples: int = 1, seed: int = 0,
                     stream: Optional[int] = None) -> None:
        """Primary-ray G-buffer of `frame` (DeviceFrame) from `mesh` seen by
        camera = (from, at, up).  env_faces: view environment, device
        [6][n][n][3] FP32 faces (DeviceBuffer or address).  light_faces: the
        lighting environment of the direct term (written to the frame's
        `direct` buffer), or None to leave that buffer as is."""
        f, a, u = (np.asarray(x, np.float32).reshape(3) for x in camera)
        cam = Camera((c_float * 3)(*f), (c_float * 3)(*a), (c_float * 3)(*u))
        b = None if bounds is None else np.ascontiguousarray(bounds, np.float32).reshape(6)
        env = env_faces.ptr if isinstance(env_faces, DeviceBuffer) else env_faces
        light = light_faces.ptr if isinstance(light_faces, DeviceBuffer) else light_faces
        params = CastParams(max_bounce, _ptr(b), env, env_n, light, light_n, direct_samples,
                            seed & 0xFFFFFFFF)
        gb = frame.gbuffer()
        _check(lib().nwprt_gbuffer_cast(self.h, mesh, ctypes.byref(cam), frame.width,
                                        frame.height, ctypes.byref(params), ctypes.byref(gb),
                                        c_void_p(stream) if stream else None))

    def sh_render(self, mesh: int, camera, width: int, height: int, transfer, light_sh,
                  degree: int, out, *, brdf_samples: int = 64, seed: int = 0, env_faces=None,
                  env_n: int = 0, stream: Optional[int] = None) -> None:
        """SH-PRT render (RenderEngine::shRender) into `out` (device float4
        [height][width]); transfer / light_sh / env_faces are device buffers
        (DeviceBuffer or address)."""
        ptr = lambda b: b.ptr if isinstance(b, DeviceBuffer) else b
        f, a, u = (np.asarray(x, np.float32).reshape(3) for x in camera)
        cam = Camera((c_float * 3)(*f), (c_float * 3)(*a), (c_float * 3)(*u))
        _check(lib().nwprt_sh_render(self.h, mesh, ctypes.byref(cam), width, height,
                                     ptr(transfer), ptr(light_sh), degree, brdf_samples,
                                     seed & 0xFFFFFFFF, ptr(env_faces), env_n, ptr(out),
                                     c_void_p(stream) if stream else None))

    def sh_compute_transfer(self, mesh: int, basis, face_n: int, degree: int, out=None, *,
                            samples: int = 4096, max_bounce: int = 12, seed: int = 0,
                            vertex_begin: int = 0, vertex_count: int = -1,
                            stream: Optional[int] = None) -> "DeviceBuffer":
        """Transfer matrices (RenderEngine::shComputeTMatrix) of the mesh's
        vertices under the (degree + 1)^2 basis maps `basis` (device buffer
        [nc][6][n][n][3] float32, or a host array that is uploaded) into
        `out` (device [n_vertices][nc * nc][3] float32; allocated when
        None), which is returned."""
        ptr = lambda b: b.ptr if isinstance(b, DeviceBuffer) else b
        nc = (degree + 1) ** 2
        if isinstance(basis, np.ndarray):
            basis = DeviceBuffer.from_array(np.ascontiguousarray(basis, np.float32))
        if out is None:
            out = DeviceBuffer(self.mesh_info(mesh)["vertices"] * nc * nc * 3 * 4)
        _check(lib().nwprt_sh_compute_transfer(self.h, mesh, ptr(basis), face_n, degree, samples,
                                               max_bounce, seed & 0xFFFFFFFF, vertex_begin,
                                               vertex_count, ptr(out),
                                               c_void_p(stream) if stream else None))
        return out

    # debug taps ----------------------------------------------------------
    def debug_hash_encode(self, scene: int, positions: np.ndarray):
        """positions: [n, 4] float32 (host).  Returns (indices [n, L, 8],
        features [n, 32]) as host arrays."""
        pos = np.ascontiguousarray(positions, np.float32)
        n = pos.shape[0]
        L = self.scene_desc(scene).grid.n_levels
        dpos = DeviceBuffer.from_array(pos)
        didx = DeviceBuffer(n * L * 8 * 4)
        dfeat = DeviceBuffer(n * 32 * 4)
        _check(lib().nwprt_debug_hash_encode(self.h, scene, dpos.ptr, n, didx.ptr, dfeat.ptr, None))
        self.sync()
        return didx.download(np.uint32, (n, L, 8)), dfeat.download(np.float32, (n, 32))

    def debug_transport(self, scene: int, env: int, frame: DeviceFrame, begin: int, count: int,
                        precision: int = PREC_BF16) -> np.ndarray:
        _, k = self.envmap_info(env)
        dt = DeviceBuffer(count * k * 3 * 4)
        dt.zero()
        gb = frame.gbuffer()
        _check(lib().nwprt_debug_transport(self.h, scene, env, ctypes.byref(gb), frame.width,
                                           frame.height, begin, count, dt.ptr, precision, None))
        self.sync()
        return dt.download(np.float32, (count, k, 3))

    # ---- batches: many relight jobs as one captured graph --------------
    def batch_create(self, jobs, precision: int = PREC_BF16,
                     tonemap: int = TONEMAP_FULL) -> int:
        """jobs: sequence of (scene_id, env_id, DeviceFrame); each frame's
        own output buffers receive the job's results."""
        arr = (BatchJob * len(jobs))()
        for i, (scene, env, frame) in enumerate(jobs):
            arr[i] = BatchJob(scene, env, frame.gbuffer(), frame.width, frame.height,
                              frame.frame_out())
        bid = c_int32()
        _check(lib().nwprt_batch_create(self.h, arr, len(jobs), precision, tonemap,
                                        ctypes.byref(bid)))
        return bid.value

    def batch_launch(self, batch: int, stream: Optional[int] = None) -> None:
        _check(lib().nwprt_batch_launch(self.h, batch, c_void_p(stream) if stream else None))

    def batch_autotune(self, batch: int):
        best, ms = c_int32(), c_float()
        _check(lib().nwprt_batch_autotune(self.h, batch, ctypes.byref(best), ctypes.byref(ms)))
        return best.value, ms.value

    def batch_info(self, batch: int):
        items, pairs, variant = c_int64(), c_int64(), c_int32()
        _check(lib().nwprt_batch_info(self.h, batch, ctypes.byref(items), ctypes.byref(pairs),
                                      ctypes.byref(variant)))
        return items.value, pairs.value, variant.value

    def batch_destroy(self, batch: int) -> None:
        _check(lib().nwprt_batch_destroy(self.h, batch))

    # ---- request server: continuous batching -----------------------------
    def server_create(self, budget_bytes: int = 0, max_active: int = 64, tick_pairs: int = 0,
                      precision: int = PREC_BF16, tonemap: int = TONEMAP_FULL,
                      variant: int = -1) -> int:
        cfg = ServerConfig(budget_bytes, max_active, tick_pairs, precision, tonemap, variant)
        sid = c_int32()
        _check(lib().nwprt_server_create(self.h, ctypes.byref(cfg), ctypes.byref(sid)))
        if not hasattr(self, "_server_keep"):
            self._server_keep = {}
        return sid.value

    def server_submit(self, server: int, scene: int, frame: DeviceFrame, *, env: int = None,
                      faces=None, yaw: Optional[float] = None, k: int = 0,
                      face_size: int = 0, out: Optional[FrameOut] = None) -> int:
        """One request: env=<map id> (resident), or faces=<DeviceBuffer> with
        yaw (rotated), or faces=<DeviceBuffer or host array> (fresh)."""
        req = Request()
        req.scene_id = scene
        req.gbuffer = frame.gbuffer()
        req.width, req.height = frame.width, frame.height
        req.out = out if out is not None else frame.frame_out()
        keep = None
        if env is not None:
            req.env_source, req.env_id = ENV_RESIDENT, env
        else:
            if isinstance(faces, DeviceBuffer):
                req.faces, req.faces_location = faces.ptr, DEVICE
            else:
                keep = np.ascontiguousarray(faces, np.float32)  # read at admission
                req.faces, req.faces_location = keep.ctypes.data, HOST
            req.env_source = ENV_ROTATED if yaw is not None else ENV_FRESH
            req.yaw = float(yaw or 0.0)
            req.k, req.face_size = k, face_size
        rid = c_int64()
        _check(lib().nwprt_server_submit(self.h, server, ctypes.byref(req), ctypes.byref(rid)))
        if keep is not None:
            self._server_keep[(server, rid.value)] = keep
        return rid.value

    def server_poll(self, server: int, wait: int = 0, max_done: int = 4096):
        ids = np.zeros(max_done, np.int64)
        lat = np.zeros(max_done, np.float64)
        n = c_int32()
        _check(lib().nwprt_server_poll(self.h, server, wait, ids.ctypes.data, lat.ctypes.data,
                                       max_done, ctypes.byref(n)))
        done = [(int(ids[i]), float(lat[i])) for i in range(n.value)]
        for rid, _ in done:
            self._server_keep.pop((server, rid), None)
        return done

    def server_stats(self, server: int) -> Dict[str, int]:
        st = ServerStats()
        _check(lib().nwprt_server_get_stats(self.h, server, ctypes.byref(st)))
        return {name: getattr(st, name) for name, _ in ServerStats._fields_}

    def server_destroy(self, server: int) -> None:
        _check(lib().nwprt_server_destroy(self.h, server))
        for key in [k for k in getattr(self, "_server_keep", {}) if k[0] == server]:
            self._server_keep.pop(key)


__all__ = [n for n in dir() if not n.startswith("_")]
