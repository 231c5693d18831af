import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
REFERENCE_SCENES = "/root/reference/scenes"  # present only in the build container


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


aqua = _load("aqua_amd", os.path.join(ROOT, "aqua-engine_amd", "aqua.py"))
oracle = _load("aqua_oracle", os.path.join(ROOT, "oracle", "oracle.py"))


@pytest.fixture(scope="session")
def aq():
    return aqua


@pytest.fixture(scope="session")
def orc():
    return oracle


@pytest.fixture(scope="session")
def box_room_dir(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("box_room"))
    path = aqua.procgen(aqua.PROCGEN_BOX_ROOM, d)
    return path


@pytest.fixture(scope="session")
def box_room(box_room_dir):
    return aqua.Scene.load(box_room_dir)


def oracle_scene_of(scene):
    """The oracle's copy of a product scene (same flattened arrays, texcoords
    and decoded images)."""
    v, m = scene.triangles()
    info = scene.info()
    uvs = scene.texcoords() if info.has_texcoords else None
    images = [scene.image(k) for k in range(info.n_images)]
    return oracle.Scene(v, m, scene.materials(), scene.point_lights(), uvs=uvs, images=images)


@pytest.fixture(scope="session")
def box_room_oracle(box_room):
    return oracle_scene_of(box_room)


def has_gpu():
    try:
        return aqua.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not has_gpu():
        pytest.fail("a GPU test ran without a visible HIP device")
    return 0


def small_camera(cam, w, h):
    """Same pose and fov, other resolution."""
    c = aqua.Camera()
    c.cam_to_world[:] = cam.cam_to_world[:]
    c.tan_half_fov = cam.tan_half_fov
    c.width = w
    c.height = h
    c.aspect = np.float32(np.float32(w) / np.float32(h))
    return c
