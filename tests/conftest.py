import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (runs the HIP library)')


@pytest.fixture(scope='session')
def gpu():
    """The GPU device; a gpu-marked test fails (never silently passes) when
    no GPU or no built library is present."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail('gpu test selected but torch.cuda.is_available() is False')
    from afsi_gpu_amd.mac import cuda_ib
    cuda_ib.build()
    return torch.device('cuda:0')
