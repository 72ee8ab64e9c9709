import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the HIP library")
    config.addinivalue_line("markers", "slow: large inputs (minutes on the CPU)")


def _have_gpu() -> bool:
    try:
        import graphreduce_amd as gr

        return gr.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Skips nothing: a GPU test on a machine without a device must fail loudly."""
    import graphreduce_amd as gr

    n = gr.device_count()
    assert n > 0, "no HIP device visible for a test marked gpu"
    return gr
