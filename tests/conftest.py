import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")
    config.addinivalue_line("markers", "slow: longer CPU checks")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def cfg():
    from minotert_amd import configs
    return configs


@pytest.fixture(scope="session")
def ctx():
    import minotert_amd as mrt
    c = mrt.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def scene0(cfg, orc):
    """Config-0 scene data: library soup, oracle BVH."""
    import minotert_amd as mrt
    verts, mat, alb = mrt.generate_scene(cfg.scene_desc(0))
    return {"verts": verts, "mat": mat, "alb": alb, "bvh": orc.BVH(verts)}


@pytest.fixture(scope="session")
def scene1(cfg, orc):
    import minotert_amd as mrt
    verts, mat, alb = mrt.generate_scene(cfg.scene_desc(1))
    return {"verts": verts, "mat": mat, "alb": alb, "bvh": orc.BVH(verts)}
