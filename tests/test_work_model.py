"""The work model behind the roofline: segments per camera path, segments
and film deposits per photon of the mixed stream's two job classes, counted
by the oracle on the same samples bench.py's CPU leg runs (smaller here),
agree with the constants bench.py falls back to when that leg is skipped."""
import importlib.util

from spectral_amd import _abi as A
from spectral_amd import jobs, scenes


def _bench():
    spec = importlib.util.spec_from_file_location("bench", jobs.__file__.rsplit("/", 2)[0] + "/bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_pinned_work_model_matches_oracle_counts(oracle):
    B = _bench()
    cam_spec = jobs.JobSpec(A.JOB_CAMERA, 128, spp=4, seed=1)
    sc, cam, _ = jobs.job_geometry(cam_spec)
    _, st = oracle.camera_accumulate(sc, cam, 128, 128, 4, 2, 1, jobs.job_opts(cam_spec), threads=8)
    seg_per_path = st.segments / st.paths
    segs = splats = n = 0
    vp = A.VolParams(**scenes.HAZE)
    for rot, off in B.VOLUME_SAMPLES:
        sp = jobs.JobSpec(A.JOB_VOLUME, 512, n_photons=1 << 13, rotation=rot, offset=off, seed=2)
        sc, cam, bm = jobs.job_geometry(sp)
        _, s2 = oracle.volume(sc, cam, bm, 512, 512, sp.n_photons, 0, vp, None, 2, A.EST_POINT,
                              jobs.job_opts(sp), threads=8)
        segs, splats, n = segs + s2.segments, splats + s2.splats, n + s2.paths
    got = {"seg_per_path": seg_per_path, "seg_per_photon": segs / n, "splat_per_photon": splats / n}
    for k, v in B.WORK_MODEL_PINNED.items():
        assert abs(got[k] - v) / v < 0.03, (k, got[k], v)
