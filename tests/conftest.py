import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")


@pytest.fixture(scope="session")
def golden():
    import json

    import numpy as np
    here = os.path.join(ROOT, "tests", "golden")
    arrays = np.load(os.path.join(here, "golden.npz"), allow_pickle=False)
    with open(os.path.join(here, "plan_tables.json")) as fh:
        meta = json.load(fh)
    return arrays, meta


@pytest.fixture(scope="session")
def device():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import msosm_gpu_amd
    msosm_gpu_amd.lib()  # fail loudly if the HIP library is missing
    return torch.device("cuda:0")
