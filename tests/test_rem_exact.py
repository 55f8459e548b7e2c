"""The minimum-image remainder of the device code (rem_exact / f_modulo in
tgx_device.h) is bit-identical to the C library's fmod and to the reference's
MODULO formulation: a host build of the same header function checked on 2
million operands (random, exact and near multiples of the period, tiny, far
away, both signs)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_rem_exact_matches_fmod(tmp_path):
    exe = tmp_path / "rem_exact_check"
    src = os.path.join(ROOT, "tests", "native", "rem_exact_check.cpp")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "--offload-arch=gfx950",
                    "-I" + os.path.join(ROOT, "turbogap_2.0_amd", "csrc"),
                    "-I" + os.path.join(ROOT, "include"), src, "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    bad, n = map(int, out.stdout.split())
    assert n == 4000000 and bad == 0
