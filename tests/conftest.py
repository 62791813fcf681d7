import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def scene_dir(tmp_path_factory):
    return tmp_path_factory.mktemp("scenes")


@pytest.fixture(scope="session")
def menger8(scene_dir):
    """Level-4 Menger sponge at depth 8 in one treelet, written to disk."""
    from libsvo_amd import workloads
    sc = workloads.menger_scene()
    base = str(scene_dir / "menger8")
    sc.write(base)
    return sc, base


@pytest.fixture(scope="session")
def terrain8(scene_dir):
    """Depth-8 terrain in 512-node treelets (many treelets), dense packing."""
    from libsvo_amd import workloads
    sc = workloads.terrain_scene(depth=8, treelet_bytes=4096)
    base = str(scene_dir / "terrain8")
    sc.write(base)
    return sc, base


@pytest.fixture(scope="session")
def terrain8_bfs(scene_dir):
    """Depth-8 terrain in 512-node treelets, reference BFS packing."""
    from libsvo_amd import workloads
    sc = workloads.terrain_scene(depth=8, treelet_bytes=4096, packing="bfs")
    base = str(scene_dir / "terrain8_bfs")
    sc.write(base)
    return sc, base
