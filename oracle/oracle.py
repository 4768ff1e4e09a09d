"""ctypes binding of the CPU oracle (oracle/build/libspectral_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  Struct arguments are passed as
pointers to the byte-identical ctypes structures of spectral_amd._abi.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "build" / "libspectral_oracle.so"
DATA_DIR = HERE.parent / "spectral_amd" / "data"

_vp, _u32, _f32, _i32 = C.c_void_p, C.c_uint32, C.c_float, C.c_int
_P = C.POINTER


class OrScene(C.Structure):
    """or_scene: pointers to the prim / plane / material arrays plus the env."""
    _fields_ = [("prims", _vp), ("n_prims", _u32), ("planes", _vp), ("n_planes", _u32),
                ("mats", _vp), ("n_mats", _u32), ("env", C.c_uint8 * 32)]


class Stats(C.Structure):
    _fields_ = [("segments", C.c_uint64), ("splats", C.c_uint64), ("paths", C.c_uint64)]


_SIG = {
    "or_pcg32": (_u32, [_u32]),
    "or_mix2": (_u32, [_u32, _u32]),
    "or_rng_new": (_u32, [_u32, _u32]),
    "or_rng_draws_u32": (None, [_u32, _u32, _u32, _P(_u32)]),
    "or_rng_draws_f32": (None, [_u32, _u32, _u32, _P(_f32)]),
    "or_hero_lambdas": (None, [_f32, _P(_f32)]),
    "or_glass_n": (_f32, [_u32, _f32]),
    "or_load_tables": (_i32, [C.c_char_p]),
    "or_cmf": (None, [_f32, _P(_f32)]),
    "or_illuminant": (_f32, [_u32, _f32]),
    "or_integrate_illuminant": (None, [_u32, _P(_f32)]),
    "or_illuminant_scale": (_f32, [_u32]),
    "or_xyz_to_linear_srgb": (None, [_P(_f32), _P(_f32)]),
    "or_tonemap_to_u8": (None, [_P(_f32), _f32, _P(C.c_uint8)]),
    "or_sphere_intersect": (_i32, [_P(_f32), _f32, _P(_f32), _P(_f32), _f32, _f32, _P(_f32)]),
    "or_solid_intersect": (_i32, [_vp, _u32, _P(_f32), _P(_f32), _f32, _f32, _P(_f32)]),
    "or_scene_intersect": (_i32, [_P(OrScene), _P(_f32), _P(_f32), _P(_f32), _P(_u32)]),
    "or_reflect": (None, [_P(_f32), _P(_f32), _P(_f32)]),
    "or_refract": (_i32, [_P(_f32), _P(_f32), _f32, _f32, _P(_f32)]),
    "or_fresnel_reflectance": (_f32, [_f32, _f32, _f32]),
    "or_material_scatter": (_i32, [_vp, _P(_f32), _P(_f32), _f32, _P(_u32), _P(_f32), _P(_f32)]),
    "or_camera_look_at": (None, [_P(_f32), _P(_f32), _P(_f32), _f32, _f32, _vp]),
    "or_camera_primary_ray": (None, [_vp, _f32, _f32, _P(_f32)]),
    "or_camera_project": (_i32, [_vp, _P(_f32), _P(_f32)]),
    "or_background_radiance": (_f32, [_vp, _P(_f32)]),
    "or_camera_accumulate": (None, [_P(OrScene), _vp, _u32, _u32, _u32, _u32, _u32, _vp, _vp,
                                    _u32, _P(Stats)]),
    "or_camera_sample": (None, [_P(OrScene), _vp, _u32, _u32, _u32, _u32, _u32, _u32, _vp,
                                _P(_f32)]),
    "or_lighttrace": (None, [_P(OrScene), _vp, _vp, _u32, _u32, _u32, _vp, _vp, _u32, _P(Stats)]),
    "or_lighttrace_photon": (_i32, [_P(OrScene), _vp, _P(_f32), _P(_f32), _f32, _P(_u32), _vp,
                                    _P(_f32)]),
    "or_volume": (None, [_P(OrScene), _vp, _vp, _u32, _u32, _u32, _u32, _vp, _vp, _u32, _u32, _vp,
                         _vp, _u32, _P(Stats)]),
    "or_phase_hg": (_f32, [_f32, _f32]),
    "or_volume_beam_max_box": (C.c_uint64, [_P(OrScene), _vp, _vp, _u32, _u32, _u32, _u32, _vp, _vp, _u32, _vp,
                                            _u32, _vp]),
    "or_photon_record_states": (None, [_P(OrScene), _vp, _vp, _u32, _u32, _f32, _u32, _vp, _vp]),
    "or_surface_rim": (None, [_P(OrScene), _vp, _u32, _u32, _vp, _vp]),
    "or_display_resolve": (None, [_vp, _vp, _u32, _u32, _u32, _f32, _f32, _f32, _vp]),
    "or_display_quantile": (_f32, [_vp, _vp, _u32, _u32, _u32, _f32, _f32]),
    "or_box_blur": (None, [_vp, _vp, _u32, _u32, C.c_int]),
    "or_floor_surface": (None, [_P(OrScene), _vp, _vp, _vp, _f32, _P(_f32), _P(_f32), _P(_f32), _u32, _u32, _vp,
                                _vp]),
    "or_composite_srgb8": (None, [_vp, _vp, _u32, _u32, _f32, _vp]),
    "or_struct_sizes": (_u32, [_u32]),
}

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} missing: run make -C oracle")
        h = C.CDLL(os.fspath(LIB_PATH))
        for k, (r, a) in _SIG.items():
            f = getattr(h, k)
            f.restype, f.argtypes = r, a
        rc = h.or_load_tables(os.fspath(DATA_DIR).encode())
        if rc != 0:
            raise RuntimeError(f"or_load_tables failed ({rc}) for {DATA_DIR}")
        _lib = h
    return _lib


def f3(v):
    return (C.c_float * 3)(*(float(x) for x in v))


def fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def scene_view(scene):
    """An or_scene over a spectral_amd.core.Scene's ctypes arrays (kept alive on the view)."""
    P, L, M = scene.arrays()
    env = scene.env()
    s = OrScene(C.cast(P, _vp), len(scene.prims), C.cast(L, _vp), len(scene.planes),
                C.cast(M, _vp), len(scene.mats))
    C.memmove(C.addressof(s) + OrScene.env.offset, C.addressof(env), 32)
    s._keep = (P, L, M, env)
    return s


def camera_accumulate(scene, cam, w, h, spp, sample_base, seed, opts, threads=1, sum_=None):
    sv = scene_view(scene)
    out = sum_ if sum_ is not None else np.zeros((h, w, 3), np.float32)
    st = Stats()
    lib().or_camera_accumulate(C.byref(sv), C.byref(cam.basis), w, h, spp, sample_base, seed,
                               C.byref(opts), out.ctypes.data, threads, C.byref(st))
    return out, st


def lighttrace(scene, scr, bm, n, photon_base, seed, opts, threads=1, sum_=None):
    sv = scene_view(scene)
    out = sum_ if sum_ is not None else np.zeros((scr.height, scr.width, 3), np.float32)
    st = Stats()
    lib().or_lighttrace(C.byref(sv), C.byref(scr), C.byref(bm), n, photon_base, seed,
                        C.byref(opts), out.ctypes.data, threads, C.byref(st))
    return out, st


def volume(scene, cam, bm, w, h, n, photon_base, vp, zbuf, seed, estimator, opts, threads=1,
           img=None):
    sv = scene_view(scene)
    out = img if img is not None else np.zeros((h, w, 3), np.float32)
    z = None if zbuf is None else np.ascontiguousarray(zbuf, np.float32)
    st = Stats()
    lib().or_volume(C.byref(sv), C.byref(cam.basis), C.byref(bm), w, h, n, photon_base,
                    C.byref(vp), None if z is None else z.ctypes.data, seed, estimator,
                    C.byref(opts), out.ctypes.data, threads, C.byref(st))
    return out, st


def photon_record(scene, bm, idx, seed, max_dist, mode, opts, scr=None):
    from spectral_amd import _abi as A
    sv = scene_view(scene)
    rec = A.PhotonRecord()
    lib().or_photon_record_states(C.byref(sv), C.byref(bm),
                                  None if scr is None else C.byref(scr), idx, seed, max_dist, mode,
                                  C.byref(opts), C.byref(rec))
    return rec


def display_resolve(vol, surf, total_photons, target_n, exposure, kradius):
    """The viewer's display pass on (h, w, 3) float32 films -> (h, w, 4) uint8."""
    vol = np.ascontiguousarray(vol, np.float32)
    h, w = vol.shape[:2]
    sv = None if surf is None else np.ascontiguousarray(surf, np.float32)
    out = np.zeros((h, w, 4), np.uint8)
    lib().or_display_resolve(vol.ctypes.data, None if sv is None else sv.ctypes.data, w, h, total_photons,
                             target_n, exposure, kradius, out.ctypes.data)
    return out


def display_quantile(vol, surf, total_photons, target_n, quantile):
    vol = np.ascontiguousarray(vol, np.float32)
    h, w = vol.shape[:2]
    sv = None if surf is None else np.ascontiguousarray(surf, np.float32)
    return lib().or_display_quantile(vol.ctypes.data, None if sv is None else sv.ctypes.data, w, h,
                                     total_photons, target_n, quantile)


def box_blur(img, radius):
    img = np.ascontiguousarray(img, np.float32)
    out = np.empty_like(img)
    lib().or_box_blur(img.ctypes.data, out.ctypes.data, img.shape[1], img.shape[0], radius)
    return out


def floor_surface(scene, cam, scr, floor_sum, gain, ambient, background, tint, w, h):
    sv = scene_view(scene)
    fs = np.ascontiguousarray(floor_sum, np.float32)
    surface = np.zeros((h, w, 3), np.float32)
    depth = np.zeros((h, w), np.float32)
    f3 = lambda v: (C.c_float * 3)(*v)
    lib().or_floor_surface(C.byref(sv), C.byref(cam.basis), C.byref(scr), fs.ctypes.data, gain, f3(ambient),
                           f3(background), f3(tint), w, h, surface.ctypes.data, depth.ctypes.data)
    return surface, depth


def composite_srgb8(surface, vol, exposure):
    surface = np.ascontiguousarray(surface, np.float32)
    vol = np.ascontiguousarray(vol, np.float32)
    out = np.zeros(vol.shape, np.uint8)
    lib().or_composite_srgb8(surface.ctypes.data, vol.ctypes.data, vol.shape[1], vol.shape[0], exposure,
                             out.ctypes.data)
    return out


def surface_rim(scene, cam, w, h):
    sv = scene_view(scene)
    rim = np.zeros((h, w, 3), np.float32)
    depth = np.zeros((h, w), np.float32)
    lib().or_surface_rim(C.byref(sv), C.byref(cam.basis), w, h, rim.ctypes.data, depth.ctypes.data)
    return rim, depth


def volume_beam_max_box(scene, cam, bm, w, h, n, photon_base, vp, zbuf, seed, opts, group=64):
    """Largest bounding box (pixels) of `group` consecutive beam-estimator
    steps of one segment, and the film the same photons give."""
    sv = scene_view(scene)
    img = np.zeros((h, w, 3), np.float32)
    z = None if zbuf is None else np.ascontiguousarray(zbuf, np.float32)
    a = lib().or_volume_beam_max_box(C.byref(sv), C.byref(cam.basis), C.byref(bm), w, h, n, photon_base,
                                     C.byref(vp), None if z is None else z.ctypes.data, seed, C.byref(opts), group,
                                     img.ctypes.data)
    return int(a), img
