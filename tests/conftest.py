import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs the kernels through the C ABI")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.build()
    return O


@pytest.fixture(scope="session")
def star_spec():
    from dwost_amd import scenes

    return scenes.star_scene_2d()


@pytest.fixture(scope="session")
def holes_spec():
    from dwost_amd import scenes

    return scenes.holes_scene_2d(n_query=2048)


@pytest.fixture(scope="session")
def ico_spec():
    from dwost_amd import scenes

    return scenes.icosphere_scene(level=4, n_query=2048)


def has_gpu():
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False
