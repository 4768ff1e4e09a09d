"""The C-ABI library loads on the CPU and exports every function that
include/spectral_hip.h declares; struct layouts agree with the header, the
oracle and the reference's POD sizes (upload.rs:34-43, vol_photons.rs:868-874).
No compute call is made here."""
import ctypes as C
import re
from pathlib import Path

import pytest

from spectral_amd import _abi as A

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "spectral_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(spx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = A.lib()
    names = declared_functions()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # every declared function has a ctypes signature in the mirror
    assert not [n for n in names if n not in A.SIGNATURES], [n for n in names if n not in A.SIGNATURES]


def test_struct_layouts(oracle):
    sizes = {"prim": A.Prim, "plane": A.Plane, "material": A.Material, "env": A.Env,
             "camera": A.CameraBasis, "beam": A.BeamDesc, "screen": A.ScreenDesc, "opts": A.Opts,
             "vol_params": A.VolParams, "photon_record": A.PhotonRecord}
    for i, (k, cls) in enumerate(sizes.items()):
        assert C.sizeof(cls) == A.STRUCT_SIZES[k], k
        assert oracle.lib().or_struct_sizes(i) == C.sizeof(cls), k
    # reference PODs: GpuPrimitive/GpuMaterial multiples of 16, GpuPlane 16, DebugPhotonGpu 304
    assert C.sizeof(A.Prim) % 16 == 0 and C.sizeof(A.Material) % 16 == 0
    assert C.sizeof(A.Plane) == 16 and C.sizeof(A.PhotonRecord) == 304


def test_context_without_device_fails_cleanly():
    lib = A.lib()
    h = C.c_void_p()
    rc = lib.spx_ctx_create(0, C.byref(h))
    if rc == A.SPX_OK:  # a GPU is visible: nothing to check here
        lib.spx_ctx_destroy(h)
        pytest.skip("a GPU is visible")
    assert rc in (A.SPX_EDEVICE, A.SPX_EINVAL) and not h.value
    assert lib.spx_ctx_create(-1, C.byref(h)) != A.SPX_OK


def test_null_arguments_rejected_without_device():
    lib = A.lib()
    assert lib.spx_ctx_create(0, None) == A.SPX_EINVAL
    assert lib.spx_camera_look_at(None, None, None, 40.0, 1.0, None) == A.SPX_EINVAL
    assert lib.spx_camera_accumulate(None, None, None, 1, 1, 1, 0, 0, None, 0, None, None) == A.SPX_EINVAL
    assert lib.spx_batch_run(None, None, 0, None, None) == A.SPX_EINVAL
    assert lib.spx_server_step(None, None, None) == A.SPX_EINVAL
    assert lib.spx_batch_plan_create(None, None, 0, None, None, None) == A.SPX_EINVAL
    assert lib.spx_batch_plan_update(None, None, 0, None) == A.SPX_EINVAL
    assert lib.spx_batch_plan_launch(None, None) == A.SPX_EINVAL
    assert lib.spx_batch_plan_info(None, None, None) == A.SPX_EINVAL
    lib.spx_batch_plan_destroy(None)
    assert lib.spx_autotune(None, None, 0, 1, None, None, None) == A.SPX_EINVAL
    assert lib.spx_ctx_kernel_times(None, None, None) == A.SPX_EINVAL


def test_tuning_layout_matches_header():
    text = (ROOT / "include" / "spectral_hip.h").read_text()
    body = text[text.index("typedef struct {\n    uint32_t max_replicas;"):]
    body = body[:body.index("} spx_tuning;")]
    fields = re.findall(r"uint32_t (\w+);", body)
    assert fields == [n for n, _ in A.Tuning._fields_]


def _header_fields(struct_name):
    """(type, name) of each field of a typedef struct in the public header."""
    text = (ROOT / "include" / "spectral_hip.h").read_text()
    end = text.index("} " + struct_name + ";")
    body = text[text.rindex("typedef struct {", 0, end) + len("typedef struct {"):end]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    return re.findall(r"([\w ]+?\**)\s*\b(\w+)(?:\[\d+\])?;", body)


def test_batch_structs_match_header():
    # the ctypes mirrors of spx_batch_cfg and spx_job: same fields, in order
    for name, cls in (("spx_batch_cfg", A.BatchCfg), ("spx_job", A.Job)):
        names = [n for _, n in _header_fields(name)]
        assert names == [n for n, _ in cls._fields_], name
    # spx_job: 10 u32 + opts 20 + camera 48 + beam 48 + screen 44 + vol 16 + 3 pointers
    assert C.sizeof(A.Job) == 40 + 20 + 48 + 48 + 44 + 16 + 3 * 8
    assert C.sizeof(A.BatchCfg) == 16
