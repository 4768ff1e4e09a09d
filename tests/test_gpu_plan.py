"""Batch plans (a job table's schedule captured once as a hipGraph) and the
autotuner: a replayed graph gives the films of spx_batch_run on every replay,
an updated plan gives the films of spx_batch_run at the new RNG bases, and an
autotuned context still gives every job its alone film."""
import numpy as np
import pytest

from spectral_amd import _abi as A
from spectral_amd import jobs

from _metrics import rel_l1_per_channel

pytestmark = pytest.mark.gpu

ALONE_TOL = 1e-4  # float summation order only


def _small_mix(n=16, seed=3):
    return jobs.mixed_jobs(n=n, seed=seed, max_side=160, max_photons_log2=16, max_spp=12)


def _images(br):
    return [f.image().copy() for f in br.films]


@pytest.mark.parametrize("splat_mode,overlap,threads", [(1, 0, 256), (2, 0, 256), (3, 0, 256), (2, 1, 256),
                                                         (3, 1, 256), (3, 1, 1024), (2, 2, 256), (3, 2, 256)])
def test_plan_replays_equal_batch_run(ctx, splat_mode, overlap, threads):
    # overlap = 1 / 2: the graph forks the bin passes onto a second branch
    # (two / three record buffers; bin_log2 = 21 makes several rounds without
    # overflowing the record pools, whose fallback atomics would reorder sums);
    # threads = 1024: the photon kernel in compute-unit-wide blocks
    specs = _small_mix()
    br = jobs.BatchRunner(ctx, specs)
    cfg = A.BatchCfg(splat_mode, 0, 4096, 2048)
    plan = None
    old = ctx.tuning()
    ctx.set_tuning(photon_threads=threads)
    if overlap:  # and the camera round on a third branch
        ctx.set_tuning(overlap_bins=overlap, bin_log2=21, camera_blocks=256)
    try:
        br.clear()
        br.run_batched(cfg=cfg)
        eager = _images(br)
        plan = br.plan(cfg=cfg)
        info = plan.info()
        print("plan", info)
        assert info["rounds"] >= (4 if overlap else 2) and info["graph_nodes"] >= 3
        # several replays, with eager work and synchronisation between them
        for rep in range(3):
            br.clear()
            ctx.sync()
            plan.launch()
            ctx.sync()
            got = _images(br)
            for j, (a, b) in enumerate(zip(got, eager)):
                assert a.sum() > 0, f"replay {rep} job {j} empty"
                if splat_mode == 2 and specs[j].kind == A.JOB_VOLUME:
                    # binned deposits: bit-reproducible volume films (camera
                    # films sum their samples with LDS float atomics)
                    assert np.array_equal(a, b), f"replay {rep} job {j} differs"
                else:
                    err = rel_l1_per_channel(a, b)
                    assert (err < ALONE_TOL).all(), f"replay {rep} job {j}: {err}"
            br.clear()
            br.run_batched(cfg=cfg)  # eager run between replays
            ctx.sync()
    finally:
        if plan is not None:
            plan.close()
        ctx.set_tuning(**old.as_dict())
        br.close()


def test_plan_update_new_bases_and_rejects_new_shape(ctx):
    specs = _small_mix(12, seed=5)
    br = jobs.BatchRunner(ctx, specs)
    plan = None
    try:
        plan = br.plan()
        sb = [sp.spp for sp in specs]
        pb = [sp.n_photons for sp in specs]
        br.set_bases(sb, pb)  # the second frame of every job
        plan.update(br.table, len(specs))
        br.clear()
        plan.launch()
        ctx.sync()
        got = _images(br)
        br.clear()
        br.run_batched()
        want = _images(br)
        for j, (a, b) in enumerate(zip(got, want)):
            assert a.sum() > 0
            err = rel_l1_per_channel(a, b)
            assert (err < ALONE_TOL).all(), f"job {j}: {err}"
        # a frame-0 film differs from the frame-1 film: the bases reached the device
        br.set_bases([0] * len(specs), [0] * len(specs))
        br.clear()
        br.run_batched()
        first = _images(br)
        assert any(not np.array_equal(a, b) for a, b in zip(first, want))
        # a different photon count needs a different schedule
        k = next(j for j, sp in enumerate(specs) if sp.kind == A.JOB_VOLUME)
        br.table[k].n_photons = br.table[k].n_photons + 100000
        with pytest.raises(A.SpxError):
            plan.update(br.table, len(specs))
        br.table[k].n_photons = specs[k].n_photons
    finally:
        if plan is not None:
            plan.close()
        br.close()


def test_autotune_installs_a_tuning_that_keeps_results(ctx):
    specs = _small_mix(12, seed=7)
    br = jobs.BatchRunner(ctx, specs)
    old = ctx.tuning()
    try:
        br.clear()
        br.run_plain()
        alone = _images(br)
        t, ms = br.autotune(sweeps=1)
        print("autotuned", t.as_dict(), f"{ms:.3f} ms")
        assert ms > 0
        cur = ctx.tuning()
        assert cur.as_dict() == t.as_dict()
        assert t.splat_mode in (1, 2, 3) and 2 <= t.tile_wlog2 <= 8 and t.batch_blocks > 0
        # the films the autotuner ran into were its own: ours are unchanged
        for j, (a, f) in enumerate(zip(alone, br.films)):
            assert np.array_equal(a, f.image()), f"job {j} film touched"
        br.clear()
        br.run_batched()
        for j, (a, b) in enumerate(zip(_images(br), alone)):
            err = rel_l1_per_channel(a, b)
            assert (err < ALONE_TOL).all(), f"job {j}: {err}"
    finally:
        ctx.set_tuning(**old.as_dict())
        br.close()


def test_tile_shapes_give_equal_films(ctx):
    specs = [sp for sp in _small_mix(16, seed=9) if sp.kind == A.JOB_VOLUME][:4]
    br = jobs.BatchRunner(ctx, specs)
    old = ctx.tuning()
    try:
        ref = None
        for tw in (2, 5, 8):
            ctx.set_tuning(tile_wlog2=tw)
            br.clear()
            br.run_batched(cfg=A.BatchCfg(2, 0, 0, 0))
            got = _images(br)
            if ref is None:
                ref = got
                continue
            for j, (a, b) in enumerate(zip(got, ref)):
                err = rel_l1_per_channel(a, b)
                assert (err < ALONE_TOL).all(), f"tile width 2^{tw} job {j}: {err}"
    finally:
        ctx.set_tuning(**old.as_dict())
        br.close()
