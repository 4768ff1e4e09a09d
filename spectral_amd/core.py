"""Host-side mirror of spectral-core's public types over the C ABI.

Names and argument meanings follow the reference crate:
``Scene`` (scene.rs:19-79), ``Material`` (material.rs:11-16), ``Camera``
(camera.rs:6-45), ``Beam`` / ``Screen`` (lighttrace.rs:15-121),
``HipTracer`` (the ``Renderer`` implementation, tracer.rs:12-150 / lib.rs:30-34),
``lighttrace_trace`` (lighttrace.rs:157-166) and the three volume estimators
(volume.rs:22-329).  Device films stay resident on the GPU; ``download()``
copies them to host numpy arrays in the reference's row-0-top layout.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _abi as A


def _check(ctx: Optional["Context"], rc: int) -> None:
    if rc != A.SPX_OK:
        msg = A.lib().spx_last_error(ctx.handle if ctx else None)
        raise A.SpxError(rc, msg.decode() if msg else "")


class Context:
    """GpuContext::new (spectral-gpu/src/lib.rs:21-62) over one HIP device."""

    def __init__(self, device: int = 0):
        self._lib = A.lib()
        h = C.c_void_p()
        rc = self._lib.spx_ctx_create(device, C.byref(h))
        if rc != A.SPX_OK:
            raise A.SpxError(rc, f"spx_ctx_create(device={device}) failed (is a GPU visible?)")
        self.handle = h
        self.device = device

    @property
    def num_cus(self) -> int:
        return self._lib.spx_ctx_num_cus(self.handle)

    def tuning(self) -> A.Tuning:
        t = A.Tuning()
        _check(self, self._lib.spx_ctx_get_tuning(self.handle, C.byref(t)))
        return t

    def set_tuning(self, **kw) -> None:
        t = self.tuning()
        for k, v in kw.items():
            setattr(t, k, int(v))
        _check(self, self._lib.spx_ctx_set_tuning(self.handle, C.byref(t)))

    def kernel_timing(self, enable: bool) -> None:
        """Bracket each batched trace-kernel launch with HIP events."""
        _check(self, self._lib.spx_ctx_kernel_timing(self.handle, int(bool(enable))))

    def kernel_time(self):
        """(summed ms, launches) of the trace-kernel launches since the last read."""
        ms, n = C.c_double(0.0), C.c_uint32(0)
        _check(self, self._lib.spx_ctx_kernel_time(self.handle, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def kernel_times(self) -> dict:
        """{group: (summed ms, launches)} for the camera trace kernel, the photon
        trace kernel and the bin passes since the last read."""
        ms, n = (C.c_double * 3)(), (C.c_uint32 * 3)()
        _check(self, self._lib.spx_ctx_kernel_times(self.handle, ms, n))
        return {name: (ms[k], n[k]) for k, name in ((A.KT_CAMERA, "camera"), (A.KT_PHOTON, "photon"),
                                                    (A.KT_BINS, "bins"))}

    def close(self) -> None:
        if self.handle:
            self._lib.spx_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass

    # device memory -----------------------------------------------------
    def malloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        _check(self, self._lib.spx_malloc(self.handle, nbytes, C.byref(p)))
        return p.value

    def free(self, ptr: int) -> None:
        _check(self, self._lib.spx_free(self.handle, ptr))

    def upload(self, dst: int, arr: np.ndarray, stream=None) -> None:
        arr = np.ascontiguousarray(arr)
        _check(self, self._lib.spx_memcpy_h2d(self.handle, dst, arr.ctypes.data, arr.nbytes, stream))

    def download(self, src: int, arr: np.ndarray, stream=None) -> np.ndarray:
        _check(self, self._lib.spx_memcpy_d2h(self.handle, arr.ctypes.data, src, arr.nbytes, stream))
        return arr

    def fill(self, dst: int, value: float, n: int, stream=None) -> None:
        _check(self, self._lib.spx_fill_f32(self.handle, dst, value, n, stream))

    def sync(self, stream=None) -> None:
        _check(self, self._lib.spx_stream_sync(self.handle, stream))

    def stream(self) -> int:
        s = C.c_void_p()
        _check(self, self._lib.spx_stream_create(self.handle, C.byref(s)))
        return s.value

    def event(self) -> int:
        e = C.c_void_p()
        _check(self, self._lib.spx_event_create(self.handle, C.byref(e)))
        return e.value

    def record(self, ev: int, stream=None) -> None:
        _check(self, self._lib.spx_event_record(self.handle, ev, stream))

    def elapsed_ms(self, a: int, b: int) -> float:
        ms = C.c_float()
        _check(self, self._lib.spx_event_elapsed_ms(self.handle, a, b, C.byref(ms)))
        return ms.value


class DeviceBuffer:
    """A float32 device array (films, z-buffers).  The initial fill runs on
    `stream`; with the null stream it is waited for, so that work queued later
    on a non-blocking stream (which does not wait for the null stream) sees
    the filled buffer."""

    def __init__(self, ctx: Context, n: int, fill: Optional[float] = 0.0, stream=None):
        self.ctx, self.n = ctx, int(n)
        self.ptr = ctx.malloc(max(4, 4 * self.n))
        if fill is not None:
            ctx.fill(self.ptr, fill, self.n, stream)
            if stream is None:
                ctx.sync(None)

    def clear(self, value: float = 0.0, stream=None) -> None:
        self.ctx.fill(self.ptr, value, self.n, stream)

    def download(self, stream=None) -> np.ndarray:
        """Copies the buffer to the host after the work queued on `stream`."""
        out = np.empty(self.n, np.float32)
        self.ctx.download(self.ptr, out, stream)
        if stream is not None:
            self.ctx.sync(stream)
        return out

    def upload(self, arr: np.ndarray) -> None:
        a = np.ascontiguousarray(arr, np.float32).ravel()
        assert a.size == self.n
        self.ctx.upload(self.ptr, a)

    def free(self) -> None:
        if self.ptr:
            self.ctx.free(self.ptr)
            self.ptr = 0


class Film(DeviceBuffer):
    """3*w*h interleaved XYZ sums, row 0 = top (AccumBuffer::sum, lib.rs:7-27)."""

    def __init__(self, ctx: Context, width: int, height: int, stream=None):
        super().__init__(ctx, 3 * width * height, stream=stream)
        self.width, self.height = width, height

    def image(self, stream=None) -> np.ndarray:
        return self.download(stream).reshape(self.height, self.width, 3)


# ---------------------------------------------------------------------------
# Scene description (scene.rs, geom.rs, material.rs)
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class Lambertian:
    reflectance: float


@dataclass(frozen=True)
class Dielectric:
    glass: int  # GLASS_BK7 / GLASS_SF11 / GLASS_WATER


def _mat(m) -> A.Material:
    if isinstance(m, Lambertian):
        return A.Material(float(m.reflectance), 0, A.MAT_LAMBERTIAN, 0)
    return A.Material(0.0, int(m.glass), A.MAT_DIELECTRIC, 0)


@dataclass
class Scene:
    """Scene::new (scene.rs:35-37): background 1.0, no horizon."""

    background: float = 1.0
    horizon: Optional[float] = None
    sun_radiance: float = 0.0
    sun_dir: Sequence[float] = (0.0, 1.0, 0.0)
    sun_cos: float = 1.0
    prims: list = field(default_factory=list)
    planes: list = field(default_factory=list)
    mats: list = field(default_factory=list)

    def add_sphere(self, center, radius: float, material) -> None:
        self.mats.append(_mat(material))
        self.prims.append(A.Prim(A.f3(center), float(radius), 0, 0, A.PRIM_SPHERE, len(self.mats) - 1))

    def add_solid(self, planes: Sequence[tuple], material) -> None:
        """planes: ((nx, ny, nz), d) half-spaces n.x + d <= 0 (geom.rs:60-70)."""
        start = len(self.planes)
        for n, d in planes:
            self.planes.append(A.Plane(A.f3(n), float(d)))
        self.mats.append(_mat(material))
        self.prims.append(A.Prim(A.f3((0, 0, 0)), 0.0, start, len(planes), A.PRIM_SOLID,
                                 len(self.mats) - 1))

    def env(self) -> A.Env:
        return A.Env(float(self.background), float(self.horizon or 0.0),
                     1 if self.horizon is not None else 0, float(self.sun_radiance),
                     A.f3(self.sun_dir), float(self.sun_cos))

    def arrays(self):
        P = (A.Prim * max(1, len(self.prims)))(*self.prims)
        L = (A.Plane * max(1, len(self.planes)))(*self.planes)
        M = (A.Material * max(1, len(self.mats)))(*self.mats)
        return P, L, M

    def upload(self, ctx: Context) -> "DeviceScene":
        return DeviceScene(ctx, self)


class DeviceScene:
    def __init__(self, ctx: Context, scene: Scene):
        self.ctx, self.scene = ctx, scene
        P, L, M = scene.arrays()
        env = scene.env()
        h = C.c_void_p()
        _check(ctx, A.lib().spx_scene_create(ctx.handle, C.byref(env), P, len(scene.prims), L,
                                             len(scene.planes), M, len(scene.mats), C.byref(h)))
        self.handle = h

    def close(self) -> None:
        if self.handle:
            A.lib().spx_scene_destroy(self.handle)
            self.handle = None


def axis_box(mn, mx) -> list:
    """ConvexSolid::axis_box, geom.rs:73-84."""
    return [((1, 0, 0), -mx[0]), ((-1, 0, 0), mn[0]), ((0, 1, 0), -mx[1]), ((0, -1, 0), mn[1]),
            ((0, 0, 1), -mx[2]), ((0, 0, -1), mn[2])]


def triangular_prism(size: float, depth: float) -> list:
    """ConvexSolid::triangular_prism, geom.rs:89-103: three side planes with
    outward normals at pi/2 + k*2pi/3 (angle, cos and sin in f32 as the
    reference), d = -size (size is the inradius), and caps at z = +-depth/2."""
    f = np.float32
    sides = []
    for k in range(3):
        ang = f(np.pi / 2) + f(k) * f(2.0) * f(np.pi) / f(3.0)
        sides.append(((float(np.cos(ang)), float(np.sin(ang)), 0.0), -float(size)))
    half = float(f(depth) / f(2.0))
    return sides + [((0, 0, 1), -half), ((0, 0, -1), -half)]


def wedge(apex_deg: float, size: float, depth: float) -> list:
    """ConvexSolid::wedge, geom.rs:110-121 (cos/sin in f32 as the reference)."""
    a = np.float32(apex_deg) * np.float32(np.pi / 180.0)
    return [((-1.0, 0.0, 0.0), -size), ((float(np.cos(a)), float(np.sin(a)), 0.0), -size),
            ((0.0, -1.0, 0.0), -size), ((0, 0, 1), -depth / 2.0), ((0, 0, -1), -depth / 2.0)]


class Camera:
    """Camera::look_at / view_basis (camera.rs:15-38), computed by the library's host code."""

    def __init__(self, basis: A.CameraBasis):
        self.basis = basis

    @classmethod
    def look_at(cls, origin, target, up, vfov_deg: float, aspect: float) -> "Camera":
        b = A.CameraBasis()
        fo, ft, fu = A.f3(origin), A.f3(target), A.f3(up)
        rc = A.lib().spx_camera_look_at(fo, ft, fu, float(vfov_deg), float(aspect), C.byref(b))
        if rc != A.SPX_OK:
            raise A.SpxError(rc, "spx_camera_look_at: non-finite camera")
        return cls(b)

    def origin(self) -> np.ndarray:
        return np.array(self.basis.origin[:], np.float32)


def beam(corner, u, v, d) -> A.BeamDesc:
    return A.BeamDesc(A.f3(corner), A.f3(u), A.f3(v), A.f3(d))


def screen(corner, u, v, width: int, height: int) -> A.ScreenDesc:
    return A.ScreenDesc(A.f3(corner), A.f3(u), A.f3(v), int(width), int(height))


def opts(illuminant: int = A.ILLUM_D65, dispersion: bool = True, max_depth: int = 8,
         rr_depth: int = 0, hero_lanes: int = 0) -> A.Opts:
    """spx_opts; hero_lanes = 4 gives photons the stratified 4-wavelength comb."""
    return A.Opts(int(illuminant), 1 if dispersion else 0, int(max_depth), int(rr_depth), int(hero_lanes))


# ---------------------------------------------------------------------------
# Camera render: the Renderer trait (lib.rs:30-34) over the device.
# ---------------------------------------------------------------------------
class HipTracer:
    """CpuTracer's interface (tracer.rs:12-150) with a device-resident AccumBuffer."""

    def __init__(self, ctx: Context, scene: Scene, camera: Camera, width: int, height: int,
                 illuminant: int = A.ILLUM_D65, seed: int = 0, engine: int = A.ENGINE_PLAIN,
                 dispersion: bool = True, max_depth: int = 8, rr_depth: int = 0):
        self.ctx = ctx
        self.dscene = scene.upload(ctx) if isinstance(scene, Scene) else scene
        self.camera = camera
        self.width, self.height = width, height
        self.seed = seed
        self.engine = engine
        self.opts = opts(illuminant, dispersion, max_depth, rr_depth)
        self.film = Film(ctx, width, height)
        self.samples = 0

    def accumulate(self, samples_per_pixel: int, stream=None) -> None:
        _check(self.ctx, A.lib().spx_camera_accumulate(
            self.ctx.handle, self.dscene.handle, C.byref(self.camera.basis), self.width,
            self.height, samples_per_pixel, self.samples, self.seed, C.byref(self.opts),
            self.engine, self.film.ptr, stream))
        self.samples += samples_per_pixel

    def sum(self) -> np.ndarray:
        return self.film.image()

    def buffer(self) -> np.ndarray:
        """AccumBuffer::mean (lib.rs:23-26)."""
        return self.sum() / np.float32(max(self.samples, 1))

    def reset(self) -> None:
        self.film.clear()
        self.samples = 0

    def set_spectral(self, on: bool) -> None:
        self.opts.dispersion = 1 if on else 0


def lighttrace_trace(ctx: Context, scene, scr: A.ScreenDesc, bm: A.BeamDesc, n: int, seed: int,
                     film: Optional[Film] = None, photon_base: int = 0, illuminant: int = A.ILLUM_E,
                     dispersion: bool = True, engine: int = A.ENGINE_PLAIN, stream=None,
                     hero_lanes: int = 0) -> Film:
    """lighttrace::trace (lighttrace.rs:157-166): adds n photons into the screen film."""
    ds = scene.upload(ctx) if isinstance(scene, Scene) else scene
    film = film or Film(ctx, scr.width, scr.height, stream)
    o = opts(illuminant, dispersion, hero_lanes=hero_lanes)
    _check(ctx, A.lib().spx_lighttrace(ctx.handle, ds.handle, C.byref(scr), C.byref(bm), n,
                                       photon_base, seed, C.byref(o), engine, film.ptr, stream))
    return film


def render_volume(ctx: Context, scene, cam: Camera, bm: A.BeamDesc, width: int, height: int,
                  n_photons: int, sigma_s: float, sigma_t: float, g: float, max_dist: float,
                  zbuffer: Optional[DeviceBuffer], seed: int, estimator: int = A.EST_POINT,
                  film: Optional[Film] = None, photon_base: int = 0, illuminant: int = A.ILLUM_E,
                  dispersion: bool = True, engine: int = A.ENGINE_PLAIN, stream=None,
                  hero_lanes: int = 0) -> Film:
    ds = scene.upload(ctx) if isinstance(scene, Scene) else scene
    film = film or Film(ctx, width, height, stream)
    vp = A.VolParams(sigma_s, sigma_t, g, max_dist)
    o = opts(illuminant, dispersion, hero_lanes=hero_lanes)
    _check(ctx, A.lib().spx_volume(ctx.handle, ds.handle, C.byref(cam.basis), C.byref(bm), width,
                                   height, n_photons, photon_base, C.byref(vp),
                                   zbuffer.ptr if zbuffer else None, seed, estimator, C.byref(o),
                                   engine, film.ptr, stream))
    return film


def render_volumetric_scene_point(ctx, scene, cam, bm, width, height, n_photons, sigma_s, sigma_t,
                                  g, max_dist, zbuffer, seed, **kw) -> Film:
    """volume::render_volumetric_scene_point (volume.rs:99-168)."""
    return render_volume(ctx, scene, cam, bm, width, height, n_photons, sigma_s, sigma_t, g,
                         max_dist, zbuffer, seed, A.EST_POINT, **kw)


def render_volumetric_scene(ctx, scene, cam, bm, width, height, n_photons, sigma_s, sigma_t, g,
                            max_dist, zbuffer, seed, **kw) -> Film:
    """volume::render_volumetric_scene, the beam estimator (volume.rs:233-329)."""
    return render_volume(ctx, scene, cam, bm, width, height, n_photons, sigma_s, sigma_t, g,
                         max_dist, zbuffer, seed, A.EST_BEAM, **kw)


def render_volumetric(ctx, scene, cam, bm, width, height, n_photons, sigma_s, max_dist, seed,
                      **kw) -> Film:
    """volume::render_volumetric, the legacy isotropic estimator (volume.rs:22-85)."""
    return render_volume(ctx, scene, cam, bm, width, height, n_photons, sigma_s, 0.0, 0.0,
                         max_dist, None, seed, A.EST_LEGACY, **kw)


def photon_records(ctx: Context, scene, bm: A.BeamDesc, first: int, count: int, seed: int,
                   max_dist: float, mode: int = 0, scr: Optional[A.ScreenDesc] = None,
                   illuminant: int = A.ILLUM_E, dispersion: bool = True):
    """GpuPhotonFilm::trace_debug (vol_photons.rs:727-779): per-photon ray states."""
    ds = scene.upload(ctx) if isinstance(scene, Scene) else scene
    recs = (A.PhotonRecord * count)()
    o = opts(illuminant, dispersion)
    _check(ctx, A.lib().spx_photon_records(ctx.handle, ds.handle, C.byref(bm),
                                           C.byref(scr) if scr is not None else None, first, count,
                                           seed, max_dist, mode, C.byref(o), recs))
    return recs


def surface_rim(ctx: Context, scene, cam: Camera, width: int, height: int, stream=None):
    """GpuPhotonFilm::recompute_surface (vol_photons.rs:685-714): rim XYZ + depth."""
    ds = scene.upload(ctx) if isinstance(scene, Scene) else scene
    rim = Film(ctx, width, height, stream)
    depth = DeviceBuffer(ctx, width * height, stream=stream)
    _check(ctx, A.lib().spx_surface_rim(ctx.handle, ds.handle, C.byref(cam.basis), width, height,
                                        rim.ptr, depth.ptr, stream))
    return rim, depth


def resolve_srgb8(ctx: Context, film: "Film", scale: float, exposure: float, stream=None) -> np.ndarray:
    """xyz_to_linear_srgb + tonemap_to_u8 (cie.rs:121-144) of film * scale:
    (h, w, 3) uint8."""
    out = DeviceBuffer(ctx, (3 * film.width * film.height + 3) // 4, fill=None)
    try:
        _check(ctx, A.lib().spx_resolve_srgb8(ctx.handle, film.ptr, film.width, film.height, scale, exposure,
                                              out.ptr, stream))
        raw = out.download(stream)
    finally:
        out.free()
    return raw.view(np.uint8)[:3 * film.width * film.height].reshape(film.height, film.width, 3).copy()


def box_blur(ctx: Context, film: "Film", radius: int, out: Optional["Film"] = None, stream=None) -> "Film":
    """box_blur (prism_on_plane.rs:40-74) of a device film into `out` (new if None)."""
    out = out or Film(ctx, film.width, film.height, stream)
    tmp = Film(ctx, film.width, film.height, stream)
    try:
        _check(ctx, A.lib().spx_box_blur(ctx.handle, film.ptr, tmp.ptr, out.ptr, film.width, film.height, radius,
                                         stream))
        ctx.sync(stream)
    finally:
        tmp.free()
    return out


def floor_surface(ctx: Context, scene, cam: Camera, floor: A.ScreenDesc, floor_sum: "Film", gain: float,
                  ambient, background, tint, width: int, height: int, stream=None):
    """prism_on_plane's surface pass + z-buffer (prism_on_plane.rs:16-38, 125-154)."""
    ds = scene.upload(ctx) if isinstance(scene, Scene) else scene
    surface = Film(ctx, width, height, stream)
    depth = DeviceBuffer(ctx, width * height, stream=stream)
    f3 = lambda v: (C.c_float * 3)(*v)
    _check(ctx, A.lib().spx_floor_surface(ctx.handle, ds.handle, C.byref(cam.basis), C.byref(floor), floor_sum.ptr,
                                          gain, f3(ambient), f3(background), f3(tint), width, height, surface.ptr,
                                          depth.ptr, stream))
    return surface, depth


def composite_srgb8(ctx: Context, surface: "Film", vol: "Film", exposure: float, stream=None) -> np.ndarray:
    """tonemap_to_u8(xyz_to_linear_srgb(surface + vol), exposure): (h, w, 3) uint8."""
    out = DeviceBuffer(ctx, (3 * vol.width * vol.height + 3) // 4, fill=None)
    try:
        _check(ctx, A.lib().spx_composite_srgb8(ctx.handle, surface.ptr, vol.ptr, vol.width, vol.height, exposure,
                                                out.ptr, stream))
        raw = out.download(stream)
    finally:
        out.free()
    return raw.view(np.uint8)[:3 * vol.width * vol.height].reshape(vol.height, vol.width, 3).copy()


def display_resolve(ctx: Context, vol: "Film", surface: Optional["Film"], total_photons: int, target_n: float,
                    exposure: float, kradius: float = 2.0, stream=None) -> np.ndarray:
    """The viewer's display pass (spectral-viewer main.rs:131-174) of a device
    film: (h, w, 4) uint8 RGBA."""
    out = DeviceBuffer(ctx, vol.width * vol.height, fill=None)
    try:
        _check(ctx, A.lib().spx_display_resolve(ctx.handle, vol.ptr, surface.ptr if surface is not None else None,
                                                vol.width, vol.height, total_photons, target_n, exposure,
                                                kradius, out.ptr, stream))
        raw = out.download(stream)
    finally:
        out.free()
    return raw.view(np.uint8).reshape(vol.height, vol.width, 4).copy()


def display_quantile(ctx: Context, vol: "Film", surface: Optional["Film"], total_photons: int, target_n: float,
                     quantile: float = 0.997, stream=None) -> float:
    """The exposure rules' luminance quantile (main.rs:614-644): exposure = c / this.
    Runs after the work queued on `stream`."""
    q = C.c_float(0.0)
    _check(ctx, A.lib().spx_display_quantile(ctx.handle, vol.ptr, surface.ptr if surface is not None else None,
                                             vol.width, vol.height, total_photons, target_n, quantile,
                                             C.byref(q), stream))
    return q.value


def rng_draws(ctx: Context, key: int, seed: int, n: int) -> np.ndarray:
    out = (C.c_uint32 * n)()
    _check(ctx, A.lib().spx_rng_draws(ctx.handle, key, seed, n, out))
    return np.frombuffer(out, np.uint32).copy()


def glass_n(ctx: Context, glass: int, lambdas: np.ndarray) -> np.ndarray:
    lam = np.ascontiguousarray(lambdas, np.float32)
    out = np.empty_like(lam)
    _check(ctx, A.lib().spx_glass_n(ctx.handle, glass,
                                    lam.ctypes.data_as(C.POINTER(C.c_float)), lam.size,
                                    out.ctypes.data_as(C.POINTER(C.c_float))))
    return out
