"""Planner checks without a GPU: the library's host code
(pencil-rs_amd/csrc/pfh_api.cpp) is compiled with g++ against a host
emulation of the HIP runtime and of the kernel launchers (tests/emu/hip_emu.cpp).
Every launch of C2C / R2C / R2R routes (3D and 4D, extras, odd and prime
lengths, slab-pipelined stage pairs) and of the batched path is checked for
in-bounds addresses, single writes per element, launch shapes that the
kernels cover, and the route's layout semantics end to end (stage transforms
replaced by the identity on their points). The transforms' arithmetic is
checked by the GPU tests."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def emu(tmp_path_factory):
    if shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"):
        pytest.skip("needs g++ and the ROCm headers")
    out = tmp_path_factory.mktemp("emu") / "emu"
    cmd = ["g++", "-std=c++17", "-O2", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           os.path.join(ROOT, "pencil-rs_amd", "csrc", "pfh_api.cpp"),
           os.path.join(ROOT, "pencil-rs_amd", "csrc", "pfh_tune_cache.cpp"),
           os.path.join(ROOT, "tests", "emu", "hip_emu.cpp"), "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True)
    return str(out)


def test_routes_and_batch_in_bounds_with_layout_semantics(emu):
    r = subprocess.run([emu], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    assert "failures 0" in r.stdout
    # every scenario reported ok
    lines = [l for l in r.stdout.splitlines() if l.endswith(": ok") or l.endswith(": FAILED")]
    assert lines and all(l.endswith(": ok") for l in lines)


def test_boundary_contract_refuses_before_device_work(emu):
    """WorkspaceMismatch for a workspace bound to another plan
    (distributed.rs:2188-2191, :4015-4017), overlapping src / dst / workspace
    refused, and the in-place Poisoned state after a failure that follows
    the first write (run_in_place_transaction, distributed.rs:3776-3794;
    many.rs:234), synchronous and deferred: every refusal launches nothing
    and writes nothing."""
    r = subprocess.run([emu, "contract"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    assert "contract: ok" in r.stdout


def test_batch_tuning_cache(emu, tmp_path):
    """The autotuner's choice is kept per (descriptors, precision, device):
    reused in the process, read back from the cache file by a later
    creation, malformed records skipped, the last good record of a key
    winning, another batch measured, and no file written when the cache is
    switched off."""
    path = tmp_path / "tune" / "batch_tuning.txt"
    r = subprocess.run([emu, "tunecache", str(path)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    assert "tune cache: ok" in r.stdout


def test_long_lines_against_naive_sums(emu):
    """Lines split into pieces, with the arithmetic emulated in long double
    (small thresholds: lines above 32 points split, prime factors above 7
    through Bluestein): local C2C in every direction, R2C / C2R, the eight
    R2R kinds and the DHT, 3D C2C routes with the long axis in each position,
    against naive sums; then real-size lines (2^15, 2^16, 2^20, 1009, 10007)
    for launch shapes, thread mappings and every address."""
    r = subprocess.run([emu, "long"], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    assert "long lines: ok" in r.stdout
