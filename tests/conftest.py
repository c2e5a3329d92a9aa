import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


@pytest.fixture(scope="session")
def golden():
    import numpy as np

    def load(name):
        return np.load(os.path.join(GOLDEN, name), allow_pickle=False)

    return load


@pytest.fixture(scope="session")
def problems():
    """Host problems of the two film workloads (inverse-CDF reservoir sampling)."""
    from phonomc_amd import runtime as rt

    cache = {}

    def get(key, **overrides):
        name = {"c1": "c1_cross_plane.toml", "c2": "c2_inplane_film.toml"}[key]
        k = (key, tuple(sorted(overrides.items())))
        if k not in cache:
            cache[k] = rt.Problem(rt.config_text(name, **overrides), mode_sampling=1)
        return cache[k]

    return get
