"""The floor-caustic product image of spectral-core's capstone example
(prism_on_plane.rs:76-185), every pass on the device: the caustic traced
onto the floor screen, its reconstruction blur, the surface pass with its
z-buffer, the z-buffered beam-estimator volume, the 99.5th-percentile
exposure and the sRGB8 tonemap of surface + volume.  Sizes default to the
example's (800x600, 4e7 caustic and 3.5e7 volume photons)."""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _abi as A
from . import core as S
from . import scenes


@dataclass
class PlaneImage:
    rgb: np.ndarray        # (h, w, 3) uint8
    exposure: float
    floor_gain: float
    floor: np.ndarray      # blurred floor screen sums, (512, 512, 3)
    surface: np.ndarray    # (h, w, 3)
    depth: np.ndarray      # (h, w)
    volume: np.ndarray     # (h, w, 3)


def prism_on_plane(ctx: S.Context, width: int = 800, height: int = 600, n_caustic: int = 40_000_000,
                   n_volume: int = 35_000_000, floor_side: int = 512, engine: int = A.ENGINE_WAVEFRONT,
                   volume_engine: int = A.ENGINE_WAVEFRONT, stream=None) -> PlaneImage:
    sc = scenes.wedge_scene().upload(ctx)       # prism_on_plane.rs:91-93
    scr = scenes.plane_floor_screen(floor_side, floor_side)
    bm = scenes.plane_beam()
    films = []
    try:
        # PASS 1, prism_on_plane.rs:107-110: caustic, blur, gain from the brightest Y
        floor = S.lighttrace_trace(ctx, sc, scr, bm, n_caustic, 0x0F1, engine=engine, stream=stream)
        films.append(floor)
        S.box_blur(ctx, floor, 2, out=floor, stream=stream)
        floor_max = S.display_quantile(ctx, floor, None, 1, 1.0, 1.0, stream)
        floor_gain = float(np.float32(6.0) / np.float32(floor_max))
        # surface pass + z-buffer, prism_on_plane.rs:113-154
        cam = scenes.plane_camera(width, height)
        sh = scenes.PLANE_SHADE
        surface, depth = S.floor_surface(ctx, sc, cam, scr, floor, floor_gain, sh["ambient"], sh["background"],
                                         sh["tint"], width, height, stream)
        films += [surface, depth]
        # volumetric pass, prism_on_plane.rs:157-164 (the beam estimator, z-buffered)
        hz = scenes.PLANE_HAZE
        vol = S.render_volumetric_scene(ctx, sc, cam, bm, width, height, n_volume, hz["sigma_s"], hz["sigma_t"],
                                        hz["g"], hz["max_dist"], depth, 0xBEE5, engine=volume_engine,
                                        stream=stream)
        films.append(vol)
        # exposure and tonemap, prism_on_plane.rs:166-182
        p = S.display_quantile(ctx, vol, surface, 1, 1.0, 0.995, stream)
        exposure = float(np.float32(1.1) / np.float32(p))
        rgb = S.composite_srgb8(ctx, surface, vol, exposure, stream)
        return PlaneImage(rgb, exposure, floor_gain, floor.image(stream), surface.image(stream),
                          depth.download(stream).reshape(height, width), vol.image(stream))
    finally:
        for f in films:
            f.free()
        sc.close()


def main(argv=None) -> int:
    """python -m spectral_amd.product [--out image.ppm] [--small]: renders the
    floor-caustic product image on device 0 and reports the wall time."""
    import argparse
    import time
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None, help="write the image as a binary PPM")
    ap.add_argument("--small", action="store_true", help="2^22 + 2^22 photons instead of 4e7 + 3.5e7")
    a = ap.parse_args(argv)
    ctx = S.Context(0)
    n_c, n_v = (1 << 22, 1 << 22) if a.small else (40_000_000, 35_000_000)
    t0 = time.perf_counter()
    img = prism_on_plane(ctx, 800, 600, n_c, n_v)
    dt = time.perf_counter() - t0
    print(f"prism_on_plane 800x600: {n_c} caustic + {n_v} volume photons in {dt * 1e3:.1f} ms, "
          f"exposure {img.exposure:.4f}", flush=True)
    if a.out:
        h, w = img.rgb.shape[:2]
        with open(a.out, "wb") as f:
            f.write(f"P6 {w} {h} 255\n".encode())
            f.write(img.rgb.tobytes())
    ctx.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
