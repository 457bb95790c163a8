"""Test configuration: the `gpu` marker and shared fixtures.

`-m "not gpu"` tests run on any host (oracle, host logic, library exports);
`-m gpu` tests run the HIP engine on cuda:0 and compare it with the oracle.
"""
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine on cuda:0)")


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    # The engine must load: no CPU fallback exists, so a missing library fails here.
    import magpackage_amd._lib as L  # noqa: F401
    torch.cuda.set_device(0)
    return torch.device("cuda", 0)
