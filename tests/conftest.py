import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs a ROCm GPU (MI355X); parity of the HIP kernels against the oracle')
    config.addinivalue_line('markers', 'slow: long-running CPU check')


@pytest.fixture(scope='session')
def oracle_lib():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope='session')
def cuda():
    import torch
    assert torch.cuda.is_available(), 'GPU tests need a ROCm device'
    import diffrp_amd._lib as L
    L.lib()  # fails loudly when the HIP library is missing
    return torch.device('cuda:0')
