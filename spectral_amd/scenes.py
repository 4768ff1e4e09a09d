"""Scene, beam and camera builders: the reference's fixtures and the benchmark
configurations, as plain PODs that the device path and the oracle both consume.

Reference fixtures:
  dsotm_scene      spectral-gpu/src/vol_photons.rs:188-222 (prism_dsotm.rs:22-52)
  tir_beam         spectral-gpu/src/vol_photons.rs:1259-1274
  diff_gate_scene  spectral-gpu/tests/diff_gate.rs:11-17
  lambertian_scene spectral-gpu/tests/lambertian_gate.rs:20-26
  wedge_screen / wedge_beam  spectral-core/src/lighttrace.rs:188-207
  scene_a_scene    spectral-core/tests/scene_a_prism.rs:66-77 (horizon sky)

Benchmark workloads (sizes in each docstring):
  prism_scene / prism_beam / prism_camera  -- the DSOTM prism workload: equilateral
      SF11 prism of side 1.0 centred on the origin, a 0.02-wide collimated
      beam at SF11 minimum deviation aimed at the left face's midpoint,
      320x180 pinhole camera at 40 deg vertical FOV.
  rainbow_scene / rainbow_camera -- unit water sphere, D65 sun disk 42 deg
      behind the camera plus a dim constant sky.
  floor variants -- the same prism plus a Lambertian floor (albedo 0.8).
"""
from __future__ import annotations

import math

import numpy as np

from . import _abi as A
from .core import Camera, Dielectric, Lambertian, Scene, axis_box, beam, screen, triangular_prism, wedge

F = np.float32


def _norm(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


# ---------------------------------------------------------------------------
# Reference fixtures
# ---------------------------------------------------------------------------
def dsotm_scene(aspect: float = 1.0):
    """The exact DSOTM scene, beam and camera (vol_photons.rs:188-222)."""
    k = 0.9
    planes = [((-0.8660254, 0.5, 0.0), -k), ((0.8660254, 0.5, 0.0), -k), ((0.0, -1.0, 0.0), -1.0),
              ((0.0, 0.0, 1.0), -1.0), ((0.0, 0.0, -1.0), -1.0)]
    sc = Scene(background=0.0)
    sc.add_solid(planes, Dielectric(A.GLASS_SF11))
    bdir = _unit32((0.84, 0.54, 0.0))
    bperp = _scale32(_unit32((-0.54, 0.84, 0.0)), 0.05)
    bm = beam((-2.6, -0.95, -0.6), (0.0, 0.0, 1.2), bperp, bdir)
    cam = Camera.look_at((0.5, -0.5, 12.0), (0.5, -0.5, 0.0), (0, 1, 0), 42.0, aspect)
    return sc, bm, cam


def tir_beam():
    """A steep beam whose violet photons TIR at the exit face (vol_photons.rs:1259-1274)."""
    bdir = _unit32((0.934, 0.358, 0.0))
    bperp = _scale32(_unit32((-0.358, 0.934, 0.0)), 0.6)
    return beam((-3.0, -1.1, -0.6), (0.0, 0.0, 1.2), bperp, bdir)


def diff_gate_scene() -> Scene:
    """diff_gate.rs:11-17."""
    s = Scene(background=1.0)
    s.add_sphere((0.0, 0.0, -3.0), 1.0, Dielectric(A.GLASS_SF11))
    s.add_sphere((0.0, 0.0, -6.0), 2.0, Lambertian(0.7))
    return s


def lambertian_scene() -> Scene:
    """lambertian_gate.rs:20-26."""
    s = Scene(background=1.0)
    s.add_sphere((0.0, 0.0, -3.0), 1.0, Lambertian(0.5))
    s.add_sphere((0.0, 0.0, -6.0), 2.0, Lambertian(0.7))
    return s


def gate_camera():
    """The diff-gate camera (diff_gate.rs:102)."""
    return Camera.look_at((0, 0, 0), (0, 0, -1), (0, 1, 0), 50.0, 1.0)


def scene_a_scene(dispersive: bool = True) -> Scene:
    """scene_a_prism.rs:66-77: a triangular prism (inradius 1, depth 2), SF11
    glass or a 0.8 Lambertian control, under a dim background (0.3) and a
    bright sky above the horizon (4.0)."""
    s = Scene(background=0.3, horizon=4.0)
    s.add_solid(triangular_prism(1.0, 2.0), Dielectric(A.GLASS_SF11) if dispersive else Lambertian(0.8))
    return s


def scene_a_camera():
    """scene_a_prism.rs:71-77."""
    return Camera.look_at((0.0, 0.0, 4.0), (0.0, 0.0, 0.0), (0, 1, 0), 50.0, 1.0)


def wedge_scene() -> Scene:
    s = Scene()
    s.add_solid(wedge(30.0, 1.0, 2.0), Dielectric(A.GLASS_SF11))
    return s


def wedge_screen(w: int = 64, h: int = 64):
    """lighttrace.rs:188-198."""
    return screen((4.0, -3.0, -2.0), (0.0, 0.0, 4.0), (0.0, 8.0, 0.0), w, h)


def wedge_beam():
    """lighttrace.rs:199-207."""
    return beam((-3.0, -0.05, -0.05), (0.0, 0.0, 0.1), (0.0, 0.1, 0.0), (1.0, 0.0, 0.0))


def slab_scene() -> Scene:
    """tracer.rs:221-226."""
    s = Scene(background=1.0)
    s.add_solid(axis_box((-5.0, -5.0, -0.1), (5.0, 5.0, 0.1)), Dielectric(A.GLASS_BK7))
    return s


# ---------------------------------------------------------------------------
# Benchmark workloads
# ---------------------------------------------------------------------------
PRISM_R = float(F(1.0 / (2.0 * math.sqrt(3.0))))  # inradius of the side-1.0 triangle
PRISM_NORMALS = ((-0.8660254, 0.5, 0.0), (0.8660254, 0.5, 0.0), (0.0, -1.0, 0.0))
FLOOR_Y = -1.6  # below the beam (which starts near y = -1.49)


def prism_planes(rotation_deg: float = 0.0):
    """Side-1.0 equilateral prism centred on the origin, caps at z = +-1, rotated about z."""
    c, s = math.cos(math.radians(rotation_deg)), math.sin(math.radians(rotation_deg))
    out = []
    for nx, ny, nz in PRISM_NORMALS:
        out.append(((float(F(c * nx - s * ny)), float(F(s * nx + c * ny)), nz), -PRISM_R))
    out += [((0.0, 0.0, 1.0), -1.0), ((0.0, 0.0, -1.0), -1.0)]
    return out


def prism_scene(rotation_deg: float = 0.0, floor: bool = False) -> Scene:
    sc = Scene(background=0.0)
    sc.add_solid(prism_planes(rotation_deg), Dielectric(A.GLASS_SF11))
    if floor:
        # A thick bounded slab, so that no ray ever starts inside an unbounded
        # half-space (geom.rs:157-171 would then report a hit at t = +inf).
        sc.add_solid(axis_box((-40.0, FLOOR_Y - 1.0, -40.0), (40.0, FLOOR_Y, 40.0)), Lambertian(0.8))
    return sc


def prism_beam(width: float = 0.02, offset: float = 0.0):
    """White collimated beam at SF11 minimum deviation (direction (0.84, 0.54, 0),
    prism_dsotm.rs:36), centred on the left face's midpoint, a 1.2-deep sheet in z,
    `width` wide across the beam, shifted by `offset` along its perpendicular."""
    d = _unit32((0.84, 0.54, 0.0))
    p = _unit32((-0.54, 0.84, 0.0))
    mid = np.array([-math.sqrt(3.0) * PRISM_R / 2.0, PRISM_R / 2.0, 0.0])
    start = mid - 3.0 * np.asarray(d, np.float64) + (offset - width / 2.0) * np.asarray(p, np.float64)
    corner = (start[0], start[1], -0.6)
    return beam(corner, (0.0, 0.0, 1.2), _scale32(p, width), d)


def prism_camera(width: int = 320, height: int = 180, orbit_deg: float = 0.0):
    """Face-on pinhole camera at 40 deg vertical FOV; orbit_deg turns it about +y."""
    tgt = np.array([0.4, -0.2, 0.0])
    r = 4.2
    a = math.radians(orbit_deg)
    org = tgt + np.array([r * math.sin(a), 0.0, r * math.cos(a)])
    return Camera.look_at(org, tgt, (0, 1, 0), 40.0, width / height)


def floor_camera(width: int = 1920, height: int = 1080, orbit_deg: float = 0.0):
    tgt = np.array([0.8, -0.8, 0.0])
    a = math.radians(orbit_deg)
    org = tgt + np.array([6.0 * math.sin(a), 2.3, 6.0 * math.cos(a)])
    return Camera.look_at(org, tgt, (0, 1, 0), 40.0, width / height)


def floor_screen(w: int = 1024, h: int = 1024):
    """The floor as a lighttrace Screen (prism_on_plane.rs:81-87 pattern): normal +y."""
    return screen((-2.0, FLOOR_Y, -3.0), (0.0, 0.0, 6.0), (8.0, 0.0, 0.0), w, h)


HAZE = dict(sigma_s=0.05, sigma_t=0.05, g=0.0, max_dist=14.0)


# prism_on_plane.rs: the floor-caustic capstone (the wedge_scene() prism on
# a floor screen, a beam high up its face, an oblique camera).
def plane_floor_screen(w: int = 512, h: int = 512):
    """prism_on_plane.rs:81-87: the floor, 10 x 14 at y = -1, normal +Y."""
    return screen((-5.0, -1.0, -5.0), (0.0, 0.0, 10.0), (14.0, 0.0, 0.0), w, h)


def plane_beam():
    """prism_on_plane.rs:98-103: a 1.5 x 0.12 shaft along +X entering at y ~ 0.6."""
    return beam((-3.0, 0.55, -0.75), (0.0, 0.0, 1.5), (0.0, 0.12, 0.0), (1.0, 0.0, 0.0))


def plane_camera(width: int = 800, height: int = 600):
    """prism_on_plane.rs:113-119."""
    return Camera.look_at((0.0, 4.5, 9.0), (4.0, -1.2, 0.0), (0.0, 1.0, 0.0), 50.0, width / height)


# prism_on_plane.rs:121-123 and 157-164: shading constants and the haze.
PLANE_SHADE = dict(background=(0.005, 0.005, 0.008), ambient=(0.03, 0.03, 0.035), tint=(0.05, 0.07, 0.09))
PLANE_HAZE = dict(sigma_s=0.4, sigma_t=0.12, g=0.6, max_dist=22.0)


def rainbow_scene() -> Scene:
    """Unit water sphere; D65 sun (radius 3 deg, radiance 20) 42 deg behind the
    camera at +z, elevated 42 deg; constant sky 0.05."""
    sc = Scene(background=0.05)
    a = math.radians(42.0)
    sc.sun_dir = tuple(float(x) for x in _unit32((0.0, math.sin(a), math.cos(a))))
    sc.sun_cos = float(F(math.cos(math.radians(3.0))))
    sc.sun_radiance = 20.0
    sc.add_sphere((0.0, 0.0, 0.0), 1.0, Dielectric(A.GLASS_WATER))
    return sc


def rainbow_camera(width: int = 512, height: int = 512, orbit_deg: float = 0.0):
    a = math.radians(orbit_deg)
    org = (4.0 * math.sin(a), 0.0, 4.0 * math.cos(a))
    return Camera.look_at(org, (0, 0, 0), (0, 1, 0), 40.0, width / height)


def _unit32(v):
    """glam Vec3::normalize in f32: v * (1 / sqrt(dot))."""
    x, y, z = (F(c) for c in v)
    ln = F(np.sqrt(F(F(x * x) + F(y * y)) + F(z * z)))
    r = F(F(1.0) / ln)
    return (float(F(x * r)), float(F(y * r)), float(F(z * r)))


def _scale32(v, s):
    s = F(s)
    return tuple(float(F(F(c) * s)) for c in v)
