"""Batched jobs: the mixed job stream and the host side of spx_batch_run.

``mixed_jobs`` draws the uneven job mix (generator seed 0, job seed = job id):
half camera renders of the rainbow scene (square film, side log-uniform in
[128, 1024], spp uniform in 1..64, dispersion on/off with p = 0.5, Russian
roulette after depth 3), half volume renders of the prism-on-floor scene
(prism rotation uniform in [0, 60) deg, beam offset uniform in +-0.1, photon
count log-uniform in [2^14, 2^24], dispersion on/off, D65 or A beam,
sigma_s = 0.05 isotropic haze).  ``JobSpec`` keeps the plain-data
description; ``BatchRunner`` uploads scenes and films once and runs the table
in one persistent launch, or job by job ("plain", one launch per job).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import _abi as A
from . import scenes
from .core import Context, DeviceScene, Film, _check, opts


# spx_autotune's pick on the 512-job mixed stream (bench.py --autotune, one
# coordinate sweep): deposit mode 3, three record buffers cycling over
# overlapped rounds, a 128-block concurrent camera round, four persistent
# blocks per compute unit with 32768-path and 8192-photon items (217.2 ms
# per step and a 60.3 ms camera half, against 219.6 and 67.4 ms with the
# previous pick of three blocks, 16384 and 4096).
TUNED_MIXED512 = {"max_replicas": 16, "replica_mib": 64, "splat_mode": 3, "chunk_log2": 22, "bin_log2": 30,
                  "tile_wlog2": 5, "batch_blocks": 1024, "chunk_paths": 32768, "chunk_photons": 8192,
                  "overlap_bins": 2, "camera_blocks": 128, "photon_threads": 256}


@dataclass
class JobSpec:
    kind: int
    size: int            # square film side
    spp: int = 0
    n_photons: int = 0
    dispersion: bool = True
    illuminant: int = A.ILLUM_D65
    rotation: float = 0.0
    offset: float = 0.0
    seed: int = 0
    orbit: float = 0.0   # camera orbit angle (deg), used by progressive frames

    @property
    def units(self) -> int:
        """Paths (camera) or photons (volume) this job traces."""
        return self.size * self.size * self.spp if self.kind == A.JOB_CAMERA else self.n_photons

    def film_bytes(self) -> int:
        return 12 * self.size * self.size


def mixed_jobs(n: int = 512, seed: int = 0, max_side: int = 1024, max_photons_log2: int = 24,
               max_spp: int = 64) -> List[JobSpec]:
    rng = np.random.default_rng(seed)
    kinds = np.array([A.JOB_CAMERA] * (n // 2) + [A.JOB_VOLUME] * (n - n // 2))
    rng.shuffle(kinds)
    out = []
    for j, k in enumerate(kinds):
        side = int(round(128 * (max_side / 128) ** rng.uniform()))
        disp = bool(rng.uniform() < 0.5)
        if k == A.JOB_CAMERA:
            out.append(JobSpec(A.JOB_CAMERA, side, spp=int(rng.integers(1, max_spp + 1)),
                               dispersion=disp, seed=j))
        else:
            nph = int(2 ** rng.uniform(14, max_photons_log2))
            out.append(JobSpec(A.JOB_VOLUME, side, n_photons=nph, dispersion=disp,
                               illuminant=int(rng.integers(0, 2)),
                               rotation=float(rng.uniform(0, 60)), offset=float(rng.uniform(-0.1, 0.1)),
                               seed=j))
    return out


def job_geometry(spec: JobSpec):
    """Scene, camera, beam for a spec (host PODs, shared by device and oracle)."""
    if spec.kind == A.JOB_CAMERA:
        return scenes.rainbow_scene(), scenes.rainbow_camera(spec.size, spec.size, spec.orbit), None
    sc = scenes.prism_scene(spec.rotation, floor=True)
    cam = scenes.floor_camera(spec.size, spec.size, spec.orbit)
    return sc, cam, scenes.prism_beam(offset=spec.offset)


def job_opts(spec: JobSpec) -> A.Opts:
    if spec.kind == A.JOB_CAMERA:
        return opts(A.ILLUM_D65, spec.dispersion, 8, 3)
    return opts(spec.illuminant, spec.dispersion, 8, 0)


class BatchRunner:
    """Device-resident scenes and films for a list of job specs."""

    def __init__(self, ctx: Context, specs: List[JobSpec]):
        self.ctx, self.specs = ctx, specs
        self.scenes: List[DeviceScene] = []
        self.films: List[Film] = []
        self.table = (A.Job * max(1, len(specs)))()
        rainbow = None
        for j, sp in enumerate(specs):
            sc, cam, bm = job_geometry(sp)
            if sp.kind == A.JOB_CAMERA:
                rainbow = rainbow or sc.upload(ctx)
                ds = rainbow
            else:
                ds = sc.upload(ctx)
                self.scenes.append(ds)
            film = Film(ctx, sp.size, sp.size)
            self.films.append(film)
            jb = self.table[j]
            jb.kind, jb.width, jb.height = sp.kind, sp.size, sp.size
            jb.spp, jb.n_photons, jb.seed = sp.spp, sp.n_photons, sp.seed
            jb.estimator = A.EST_POINT
            jb.opts = job_opts(sp)
            jb.camera = cam.basis
            if bm is not None:
                jb.beam = bm
            jb.vol = A.VolParams(**scenes.HAZE)
            jb.scene = ds.handle.value
            jb.film = film.ptr
            jb.zbuf = None
        if rainbow is not None:
            self.scenes.append(rainbow)

    def set_bases(self, sample_bases=None, photon_bases=None) -> None:
        for j in range(len(self.specs)):
            if sample_bases is not None:
                self.table[j].sample_base = int(sample_bases[j])
            if photon_bases is not None:
                self.table[j].photon_base = int(photon_bases[j])

    def clear(self, stream=None) -> None:
        for f in self.films:
            f.clear(0.0, stream)

    def run_batched(self, stream=None, cfg: Optional[A.BatchCfg] = None, subset=None) -> None:
        tbl, n = self._table(subset)
        _check(self.ctx, A.lib().spx_batch_run(self.ctx.handle, tbl, n,
                                               C.byref(cfg) if cfg is not None else None, stream))

    def _table(self, subset=None):
        if subset is None:
            return self.table, len(self.specs)
        return (A.Job * max(1, len(subset)))(*[self.table[j] for j in subset]), len(subset)

    def plan(self, stream=None, cfg: Optional[A.BatchCfg] = None, subset=None) -> "BatchPlan":
        """The table's schedule captured as a hipGraph (spx_batch_plan_create)."""
        tbl, n = self._table(subset)
        return BatchPlan(self.ctx, tbl, n, cfg, stream)

    def autotune(self, stream=None, sweeps: int = 1, subset=None):
        """spx_autotune over the table (or a subset): returns (Tuning, best ms)."""
        tbl, n = self._table(subset)
        t, ms = A.Tuning(), C.c_double(0.0)
        _check(self.ctx, A.lib().spx_autotune(self.ctx.handle, tbl, n, sweeps, stream, C.byref(t),
                                              C.byref(ms)))
        return t, ms.value

    def run_plain(self, stream=None, engine: int = A.ENGINE_PLAIN) -> None:
        """One launch per job through the single-job entry points."""
        L = A.lib()
        for j, sp in enumerate(self.specs):
            jb = self.table[j]
            if sp.kind == A.JOB_CAMERA:
                rc = L.spx_camera_accumulate(self.ctx.handle, C.c_void_p(jb.scene), C.byref(jb.camera),
                                             jb.width, jb.height, jb.spp, jb.sample_base, jb.seed,
                                             C.byref(jb.opts), engine, jb.film, stream)
            else:
                rc = L.spx_volume(self.ctx.handle, C.c_void_p(jb.scene), C.byref(jb.camera),
                                  C.byref(jb.beam), jb.width, jb.height, jb.n_photons, jb.photon_base,
                                  C.byref(jb.vol), None, jb.seed, jb.estimator, C.byref(jb.opts),
                                  engine, jb.film, stream)
            _check(self.ctx, rc)

    def close(self) -> None:
        for f in self.films:
            f.free()
        for s in self.scenes:
            s.close()


class BatchPlan:
    """A job table's launch schedule captured once as a hipGraph; ``launch``
    replays it, ``update`` rewrites per-job values (RNG bases, cameras, films)
    that keep the same schedule."""

    def __init__(self, ctx: Context, table, n: int, cfg: Optional[A.BatchCfg] = None, stream=None):
        self.ctx = ctx
        h = C.c_void_p()
        _check(ctx, A.lib().spx_batch_plan_create(ctx.handle, table, n,
                                                  C.byref(cfg) if cfg is not None else None, stream,
                                                  C.byref(h)))
        self.handle = h

    def update(self, table, n: int, stream=None) -> None:
        _check(self.ctx, A.lib().spx_batch_plan_update(self.handle, table, n, stream))

    def launch(self, stream=None) -> None:
        _check(self.ctx, A.lib().spx_batch_plan_launch(self.handle, stream))

    def info(self):
        r, nodes = C.c_uint32(0), C.c_uint32(0)
        _check(self.ctx, A.lib().spx_batch_plan_info(self.handle, C.byref(r), C.byref(nodes)))
        return {"rounds": r.value, "graph_nodes": nodes.value}

    def close(self) -> None:
        if self.handle:
            A.lib().spx_batch_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass
