"""The C-ABI library loads on a machine without a GPU and exports every
function that include/pencil_fft_hip.h declares (no compute calls here)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pencil_fft_hip.h")
LIB = os.path.join(ROOT, "pencil-rs_amd", "libpencilfft_hip.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pfh_[a-z_0-9]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        subprocess.run(["make", "-C", os.path.join(ROOT, "pencil-rs_amd"), "-j8"], check=True)
    return ctypes.CDLL(LIB)


def test_header_declares_the_plan_api():
    names = declared_functions()
    for required in ("pfh_plan_create", "pfh_execute", "pfh_local_execute", "pfh_execute_in_place",
                     "pfh_plan_workspace_bytes", "pfh_plan_layout", "pfh_batch_execute"):
        assert required in names


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_mirror_binds_every_exported_symbol():
    import pencil_fft_hip as pf

    assert set(pf.EXPORTED_SYMBOLS) == set(declared_functions())


def test_version_and_status_strings_need_no_device(lib):
    lib.pfh_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.pfh_version()
    lib.pfh_status_string.restype = ctypes.c_char_p
    lib.pfh_status_string.argtypes = [ctypes.c_int]
    assert lib.pfh_status_string(17) == b"inverse spectrum has an invalid constrained boundary plane"
