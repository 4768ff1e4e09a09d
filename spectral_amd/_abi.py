"""ctypes mirror of include/spectral_hip.h and the loader of libspectral_hip.so.

The library is built in-tree (``spectral_amd/libspectral_hip.so``).  There is
no fallback: if the shared library is missing, importing the bindings raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libspectral_hip.so"

SPX_OK = 0
SPX_EINVAL = -1
SPX_ENOMEM = -2
SPX_EDEVICE = -3
SPX_ECAP = -4
SPX_EBUSY = -5

PRIM_SPHERE, PRIM_SOLID = 0, 1
MAT_LAMBERTIAN, MAT_DIELECTRIC = 0, 1
GLASS_BK7, GLASS_SF11, GLASS_WATER = 0, 1, 2
ILLUM_D65, ILLUM_A, ILLUM_E = 0, 1, 2
EST_POINT, EST_BEAM, EST_LEGACY = 0, 1, 2
ENGINE_PLAIN, ENGINE_WAVEFRONT = 0, 1
JOB_CAMERA, JOB_VOLUME, JOB_LIGHTTRACE = 0, 1, 2
KT_CAMERA, KT_PHOTON, KT_BINS = 0, 1, 2

F3 = C.c_float * 3


class Prim(C.Structure):
    _fields_ = [("center", F3), ("radius", C.c_float), ("plane_start", C.c_uint32),
                ("plane_count", C.c_uint32), ("kind", C.c_uint32), ("material", C.c_uint32)]


class Plane(C.Structure):
    _fields_ = [("normal", F3), ("d", C.c_float)]


class Material(C.Structure):
    _fields_ = [("reflectance", C.c_float), ("glass", C.c_uint32), ("kind", C.c_uint32),
                ("pad", C.c_uint32)]


class Env(C.Structure):
    _fields_ = [("background", C.c_float), ("horizon", C.c_float), ("has_horizon", C.c_uint32),
                ("sun_radiance", C.c_float), ("sun_dir", F3), ("sun_cos", C.c_float)]


class CameraBasis(C.Structure):
    _fields_ = [("origin", F3), ("lower_left", F3), ("horizontal", F3), ("vertical", F3)]


class BeamDesc(C.Structure):
    _fields_ = [("corner", F3), ("u", F3), ("v", F3), ("dir", F3)]


class ScreenDesc(C.Structure):
    _fields_ = [("corner", F3), ("u", F3), ("v", F3), ("width", C.c_uint32),
                ("height", C.c_uint32)]


class Opts(C.Structure):
    _fields_ = [("illuminant", C.c_uint32), ("dispersion", C.c_uint32),
                ("max_depth", C.c_uint32), ("rr_depth", C.c_uint32), ("hero_lanes", C.c_uint32)]


class VolParams(C.Structure):
    _fields_ = [("sigma_s", C.c_float), ("sigma_t", C.c_float), ("g", C.c_float),
                ("max_dist", C.c_float)]


class PhotonRecord(C.Structure):
    _fields_ = [("num_states", C.c_uint32), ("lambda_", C.c_float), ("tir_count", C.c_uint32),
                ("pad", C.c_uint32), ("states", ((C.c_float * 4) * 2) * 9)]


class Job(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32),
                ("spp", C.c_uint32), ("sample_base", C.c_uint32), ("n_photons", C.c_uint32),
                ("photon_base", C.c_uint32), ("seed", C.c_uint32), ("estimator", C.c_uint32),
                ("pad", C.c_uint32), ("opts", Opts), ("camera", CameraBasis), ("beam", BeamDesc),
                ("screen", ScreenDesc), ("vol", VolParams), ("scene", C.c_void_p),
                ("film", C.c_void_p), ("zbuf", C.c_void_p)]


class Tuning(C.Structure):
    _fields_ = [("max_replicas", C.c_uint32), ("replica_mib", C.c_uint32),
                ("splat_mode", C.c_uint32), ("chunk_log2", C.c_uint32), ("bin_log2", C.c_uint32),
                ("tile_wlog2", C.c_uint32), ("batch_blocks", C.c_uint32), ("chunk_paths", C.c_uint32),
                ("chunk_photons", C.c_uint32), ("overlap_bins", C.c_uint32),
                ("camera_blocks", C.c_uint32), ("photon_threads", C.c_uint32)]

    def as_dict(self) -> dict:
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class BatchCfg(C.Structure):
    """spx_batch_cfg: per-call overrides of the context's tuning (0 = tuned)."""
    _fields_ = [("splat_mode", C.c_uint32), ("blocks", C.c_uint32), ("chunk_paths", C.c_uint32),
                ("chunk_photons", C.c_uint32)]


class Request(C.Structure):
    _fields_ = [("job", Job), ("frames", C.c_uint32), ("orbit_deg", C.c_float),
                ("orbit_center", F3)]


class RequestInfo(C.Structure):
    _fields_ = [("state", C.c_uint32), ("frames_done", C.c_uint32), ("footprint", C.c_uint64),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double)]


REQ_PENDING, REQ_ACTIVE, REQ_DONE = 0, 1, 2

STRUCT_SIZES = {"prim": 32, "plane": 16, "material": 16, "env": 32, "camera": 48, "beam": 48,
                "screen": 44, "opts": 20, "vol_params": 16, "photon_record": 304}

_vp, _u32, _f32, _i32, _sz = C.c_void_p, C.c_uint32, C.c_float, C.c_int, C.c_size_t
_P = C.POINTER

# name -> (restype, argtypes).  Every function declared in include/spectral_hip.h.
SIGNATURES = {
    "spx_ctx_create": (_i32, [_i32, _P(_vp)]),
    "spx_ctx_destroy": (None, [_vp]),
    "spx_last_error": (C.c_char_p, [_vp]),
    "spx_ctx_num_cus": (_i32, [_vp]),
    "spx_ctx_set_tuning": (_i32, [_vp, _P(Tuning)]),
    "spx_ctx_get_tuning": (_i32, [_vp, _P(Tuning)]),
    "spx_ctx_kernel_timing": (_i32, [_vp, _i32]),
    "spx_ctx_kernel_time": (_i32, [_vp, _P(C.c_double), _P(_u32)]),
    "spx_ctx_kernel_times": (_i32, [_vp, _P(C.c_double), _P(_u32)]),
    "spx_atomic_probe": (_i32, [_vp, _vp, _sz, _u32, _u32, _vp]),
    "spx_camera_look_at": (_i32, [_P(_f32), _P(_f32), _P(_f32), _f32, _f32, _P(CameraBasis)]),
    "spx_scene_create": (_i32, [_vp, _P(Env), _P(Prim), _u32, _P(Plane), _u32, _P(Material), _u32,
                                _P(_vp)]),
    "spx_scene_destroy": (None, [_vp]),
    "spx_camera_accumulate": (_i32, [_vp, _vp, _P(CameraBasis), _u32, _u32, _u32, _u32, _u32,
                                     _P(Opts), _u32, _vp, _vp]),
    "spx_lighttrace": (_i32, [_vp, _vp, _P(ScreenDesc), _P(BeamDesc), _u32, _u32, _u32, _P(Opts),
                              _u32, _vp, _vp]),
    "spx_volume": (_i32, [_vp, _vp, _P(CameraBasis), _P(BeamDesc), _u32, _u32, _u32, _u32,
                          _P(VolParams), _vp, _u32, _u32, _P(Opts), _u32, _vp, _vp]),
    "spx_photon_records": (_i32, [_vp, _vp, _P(BeamDesc), _P(ScreenDesc), _u32, _u32, _u32, _f32,
                                  _u32, _P(Opts), _P(PhotonRecord)]),
    "spx_surface_rim": (_i32, [_vp, _vp, _P(CameraBasis), _u32, _u32, _vp, _vp, _vp]),
    "spx_resolve_srgb8": (_i32, [_vp, _vp, _u32, _u32, _f32, _f32, _vp, _vp]),
    "spx_box_blur": (_i32, [_vp, _vp, _vp, _vp, _u32, _u32, _u32, _vp]),
    "spx_floor_surface": (_i32, [_vp, _vp, _P(CameraBasis), _P(ScreenDesc), _vp, _f32, _P(_f32), _P(_f32), _P(_f32),
                                 _u32, _u32, _vp, _vp, _vp]),
    "spx_composite_srgb8": (_i32, [_vp, _vp, _vp, _u32, _u32, _f32, _vp, _vp]),
    "spx_display_resolve": (_i32, [_vp, _vp, _vp, _u32, _u32, _u32, _f32, _f32, _f32, _vp, _vp]),
    "spx_display_quantile": (_i32, [_vp, _vp, _vp, _u32, _u32, _u32, _f32, _f32, _P(_f32), _vp]),
    "spx_display_quantile_noalloc": (_i32, [_vp, _vp, _vp, _u32, _u32, _u32, _f32, _f32, _P(_f32), _vp]),
    "spx_rng_draws": (_i32, [_vp, _u32, _u32, _u32, _P(_u32)]),
    "spx_glass_n": (_i32, [_vp, _u32, _P(_f32), _u32, _P(_f32)]),
    "spx_batch_run": (_i32, [_vp, _P(Job), _u32, _P(BatchCfg), _vp]),
    "spx_batch_plan_create": (_i32, [_vp, _P(Job), _u32, _P(BatchCfg), _vp, _P(_vp)]),
    "spx_batch_plan_update": (_i32, [_vp, _P(Job), _u32, _vp]),
    "spx_batch_plan_launch": (_i32, [_vp, _vp]),
    "spx_batch_plan_info": (_i32, [_vp, _P(_u32), _P(_u32)]),
    "spx_batch_plan_destroy": (None, [_vp]),
    "spx_autotune": (_i32, [_vp, _P(Job), _u32, _u32, _vp, _P(Tuning), _P(C.c_double)]),
    "spx_server_create": (_i32, [_vp, C.c_uint64, _P(_vp)]),
    "spx_server_destroy": (None, [_vp]),
    "spx_server_submit": (_i32, [_vp, _P(Request), _P(C.c_uint64)]),
    "spx_server_step": (_i32, [_vp, _vp, _P(_u32)]),
    "spx_server_info": (_i32, [_vp, C.c_uint64, _P(RequestInfo)]),
    "spx_server_frame_camera": (_i32, [_vp, C.c_uint64, _u32, _P(CameraBasis)]),
    "spx_server_film": (_i32, [_vp, C.c_uint64, _P(_f32), _sz]),
    "spx_server_release": (_i32, [_vp, C.c_uint64]),
    "spx_server_bytes_in_use": (C.c_uint64, [_vp]),
    "spx_server_clock": (C.c_double, [_vp]),
    "spx_server_set_graphs": (_i32, [_vp, _i32]),
    "spx_server_stats": (_i32, [_vp, _P(C.c_uint64), _P(C.c_uint64)]),
    "spx_malloc": (_i32, [_vp, _sz, _P(_vp)]),
    "spx_free": (_i32, [_vp, _vp]),
    "spx_memcpy_h2d": (_i32, [_vp, _vp, _vp, _sz, _vp]),
    "spx_memcpy_d2h": (_i32, [_vp, _vp, _vp, _sz, _vp]),
    "spx_fill_f32": (_i32, [_vp, _vp, _f32, _sz, _vp]),
    "spx_copy_f32": (_i32, [_vp, _vp, _vp, _sz, _vp]),
    "spx_stream_create": (_i32, [_vp, _P(_vp)]),
    "spx_stream_destroy": (_i32, [_vp, _vp]),
    "spx_stream_sync": (_i32, [_vp, _vp]),
    "spx_event_create": (_i32, [_vp, _P(_vp)]),
    "spx_event_destroy": (_i32, [_vp, _vp]),
    "spx_event_record": (_i32, [_vp, _vp, _vp]),
    "spx_event_elapsed_ms": (_i32, [_vp, _vp, _vp, _P(_f32)]),
}

_lib = None


def lib() -> C.CDLL:
    """Load libspectral_hip.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(make -C spectral_amd/csrc); there is no CPU fallback")
        h = C.CDLL(os.fspath(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(h, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


class SpxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"spectral_hip error {code}: {msg}")
        self.code = code


def f3(v) -> F3:
    return F3(*(float(x) for x in v))
