"""The C ABI without Python or torch: pencil-rs_amd/examples/r2c_round_trip
(a C program linked against libpencilfft_hip.so, built with the library)
runs a 64 x 48 x 40 f32 R2C forward + normalized C2R inverse on device
buffers from the HIP runtime and checks the round trip (<= 2e-6) and the DC
bin (= the sum of the field)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "pencil-rs_amd", "examples", "r2c_round_trip")


def test_c_program_round_trip(gpu):
    assert os.path.exists(EXE), f"{EXE} is missing: build the library first (make -C pencil-rs_amd)"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("r2c_round_trip rel_l2=")
