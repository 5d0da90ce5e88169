import importlib
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def tdp():
    """The product package (native libraries must be built)."""
    return importlib.import_module("3dpipe_amd")


@pytest.fixture(scope="session")
def small_tiles(tdp, tmp_path_factory):
    """Small tiles: 10 vessels x 20 nuclei in a 2000^3 cube, seed 0,
    written as RAW .dt files so that the oracle reads the same bytes."""
    d = tmp_path_factory.mktemp("tiles")
    n, v = tdp.generate(seed=0, n_vessels=10, nuclei_per_vessel=20, cube=2000.0)
    pn, pv = str(d / "nuclei.dt"), str(d / "vessels.dt")
    n.save(pn)
    v.save(pv)
    return pn, pv


@pytest.fixture(scope="session")
def tiny_tiles(tdp, tmp_path_factory):
    """Small meshes (subdiv-2 nuclei, 8-sided vessels) for brute-force checks."""
    d = tmp_path_factory.mktemp("tiny")
    n, v = tdp.generate(seed=3, n_vessels=4, nuclei_per_vessel=5, cube=600.0, subdiv=2,
                        vessel_sides=8, rings_min=2, rings_max=2, prototypes=8, voxel_tris=20,
                        place_min=20.0, place_max=120.0)
    pn, pv = str(d / "nuclei.dt"), str(d / "vessels.dt")
    n.save(pn)
    v.save(pv)
    return pn, pv


def random_triangles(rng, n, spread=10.0, offset=20.0):
    """Random triangle pairs plus degenerate families (shared vertex, shared
    edge, coplanar, parallel offset, collinear, coincident)."""
    S = rng.uniform(-spread, spread, (n, 9)).astype(np.float32)
    T = (rng.uniform(-spread, spread, (n, 9)) + rng.uniform(-offset, offset, (n, 1))).astype(np.float32)
    k = n // 16
    S[0:k, 3:6] = S[0:k, 0:3]                        # degenerate edge in S
    T[k:2 * k, 0:3] = S[k:2 * k, 0:3]                # shared vertex
    T[2 * k:3 * k, 0:6] = S[2 * k:3 * k, 3:9]        # shared edge
    T[3 * k:4 * k] = S[3 * k:4 * k] + np.float32(1.5)  # translated copy
    S[4 * k:5 * k, 2::3] = 0
    T[4 * k:5 * k, 2::3] = 0                         # coplanar
    S[5 * k:6 * k, 2::3] = 0
    T[5 * k:6 * k, 2::3] = 3                         # parallel planes
    S[6 * k:7 * k, 6:9] = S[6 * k:7 * k, 0:3] * 0.5 + S[6 * k:7 * k, 3:6] * 0.5  # collinear
    T[7 * k:8 * k] = S[7 * k:8 * k]                  # identical
    return S, T
