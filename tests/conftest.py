import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")
    config.addinivalue_line("markers", "slow: full-size workloads")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle

    oracle.load()
    return oracle


@pytest.fixture(scope="session")
def gs():
    import gamesplat_amd

    gamesplat_amd.load()
    return gamesplat_amd


@pytest.fixture(scope="session")
def ctx(gs):
    c = gs.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def c1_scene(gs):
    from gamesplat_amd import scenes

    g, cfg = scenes.scene_config("c1")
    cam = scenes.random_view_camera(cfg["seed"], cfg["width"], cfg["height"])
    return g, cfg, cam
