import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


# Geometries used by the tests.
#  tiny   : 16 rings, 96 bins x 96 views, span 3 (two segment groups), no rebinning
#  tinyrb : 16 rings, 128 bins x 128 views read at -X 64 (2x2 rebinning), span 3
#  c1     : HRRT span 9, -X 128 (the CPU reference configuration)
#  c2     : HRRT span 3, 256^2 (the GPU headline configuration)
GEOMS = {
    "tiny": dict(rings=16, def_elements=96, def_angles=96, x_pixels=96, span=3, maxdel=7, z_min=3, z_max=28),
    "tinyrb": dict(rings=16, def_elements=128, def_angles=128, x_pixels=64, span=3, maxdel=7, z_min=3, z_max=28,
                   rel_fov=0.9),
    "tiny9": dict(rings=16, def_elements=96, def_angles=96, x_pixels=96, span=9, maxdel=13, z_min=3, z_max=28),
    "c1": dict(x_pixels=128, span=9, maxdel=67),
    "c2": dict(x_pixels=256, span=3, maxdel=67),
}


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle
    oracle.build()
    return oracle
