import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "pencil-rs_amd"), os.path.join(ROOT, "oracle"), os.path.dirname(__file__)):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "slow: full-size case (seconds to a minute)")


@pytest.fixture(scope="session")
def gpu():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)
    import pencil_fft_hip

    pencil_fft_hip.load_library()
    return torch.device("cuda:0")


@pytest.fixture(scope="session", autouse=True)
def _batch_tuning_cache(tmp_path_factory):
    """Autotuned batches keep their tunings in a file of this session, not in
    the user's cache directory, unless PFH_TUNE_CACHE is set already."""
    if "PFH_TUNE_CACHE" not in os.environ:
        os.environ["PFH_TUNE_CACHE"] = str(tmp_path_factory.mktemp("tune") / "batch_tuning.txt")
    yield
