"""The C ABI library: it loads, exports every function include/aqua/aqua.h
declares, rejects a CPU device, and fails cleanly without a GPU."""
import ctypes as C
import os
import re

import pytest

from conftest import ROOT, aqua, has_gpu

HEADER = os.path.join(ROOT, "include", "aqua", "aqua.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(aqua_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 25
    L = C.CDLL(aqua.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_struct_sizes_match_header():
    text = open(HEADER).read()
    sizes = dict((m[1], int(m[0])) for m in re.findall(r"/\* (\d+) bytes.*?\*/\s*typedef struct (\w+)", text, re.S))
    py = {"aqua_load_opts": aqua.LoadOpts, "aqua_material": aqua.Material, "aqua_camera": aqua.Camera,
          "aqua_scene_info": aqua.SceneInfo, "aqua_procgen_params": aqua.ProcgenParams,
          "aqua_integrator_cfg": aqua.IntegratorCfg, "aqua_frame_stats": aqua.FrameStats,
          "aqua_renderer_opts": aqua.RendererOpts, "aqua_point_light": aqua.PointLight,
          "aqua_nrc_cfg": aqua.NrcCfg, "aqua_nrc_stats": aqua.NrcStats}
    assert set(sizes) <= set(py)
    for k, s in sizes.items():
        assert C.sizeof(py[k]) == s, k


def test_abi_version():
    assert aqua.lib().aqua_abi_version() == 1


def test_negative_device_is_rejected(box_room):
    with pytest.raises(aqua.AquaError) as e:
        aqua.Renderer(box_room, aqua.integrator_default(), device=-1)
    assert e.value.code == aqua.ERR_INVALID_ARG


def test_no_device_error_without_gpu(box_room):
    if has_gpu():
        pytest.skip("a GPU is visible")
    with pytest.raises(aqua.AquaError) as e:
        aqua.Renderer(box_room, aqua.integrator_default(), device=0)
    assert e.value.code == aqua.ERR_NO_DEVICE


def test_scene_from_arrays_validation():
    cam = aqua.Camera()
    cam.width = cam.height = 8
    m = aqua.Material()
    import numpy as np
    v = np.zeros((1, 9), np.float32)
    with pytest.raises(aqua.AquaError) as e:
        aqua.Scene.from_arrays(v, np.uint32([3]), [m], cam)
    assert e.value.code == aqua.ERR_INVALID_ARG
    v[0, 0] = np.nan
    with pytest.raises(aqua.AquaError) as e:
        aqua.Scene.from_arrays(v, np.uint32([0]), [m], cam)
    assert e.value.code == aqua.ERR_NONFINITE
