import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (HERE, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU machine)")


@pytest.fixture(scope="session")
def gpu_aligner():
    """One device context for the GPU tests; fails (does not skip) without a GPU."""
    import mhm2_staging_amd as m
    n = m.get_num_node_gpus()
    assert n > 0, "no GPU visible: GPU tests must run on the GPU machine"
    al = m.Aligner(device=0)
    yield al
    al.close()
