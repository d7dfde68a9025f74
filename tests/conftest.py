import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle_api():
    import oracle
    if not os.path.exists(oracle.LIB):
        oracle.build(ref=False)
    return oracle.Oracle()


def golden_names():
    return sorted(f[len("golden_"):-4] for f in os.listdir(GOLDEN)
                  if f.startswith("golden_") and f.endswith(".npz")
                  and f not in ("golden_aca.npz", "golden_mbls.npz", "golden_transposed.npz", "golden_hlu.npz",
                                "golden_acar.npz", "golden_heh.npz"))


def load_golden(name):
    return dict(np.load(os.path.join(GOLDEN, f"golden_{name}.npz"), allow_pickle=False))


def golden_points(g):
    from ahmed_amd import geometry
    geo, N, seed = str(g["geometry"]), int(g["N"]), int(g["seed"])
    if geo == "sphere":
        return geometry.sphere(N, seed)
    if geo == "cube":
        return geometry.cube_surface(N, seed)
    return geometry.ellipsoid(N, (1.0, 2.5, 3.7), seed)
