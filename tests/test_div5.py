"""The device computes the CIE table position (nm - 380) / 5 (cie.rs:47-89)
with a multiply and two FMAs instead of an IEEE division; this checks, for
every float in [0, 350] (every argument a wavelength in [380, 730] gives),
that the two are bit-identical.  Compiled and run on the CPU."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_div5_equals_ieee_division_on_every_float(tmp_path):
    exe = tmp_path / "div5_check"
    subprocess.run(["gcc", "-O2", "-fopenmp", "-ffp-contract=off", str(HERE / "div5_check.c"), "-o", str(exe),
                    "-lm"], check=True)
    env = dict(os.environ, OMP_NUM_THREADS=str(min(8, os.cpu_count() or 1)))
    out = subprocess.run([str(exe), "0", "350"], check=True, capture_output=True, text=True, env=env).stdout
    n, bad = (int(v) for v in out.split())
    assert n == 1135542273  # every float in [0, 350]
    assert bad == 0
