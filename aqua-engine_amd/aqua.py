"""Python host binding of the C ABI in include/aqua/aqua.h (ctypes).

The reference crate (`arukas`, Cargo.toml:2-5) has no code and therefore no
Python API; its boundary is its data: scene JSON (scenes/cbox.json:2-627),
integrator JSON (scenes/integrator.json:2-8) and BSON meshes
(scenes/CornellBox-Original_floor_0.mesh). This module loads exactly those
through libaqua.so and drives the gfx950 renderer. It has no compute of its
own and no CPU fallback: without the built library or without a GPU, the
calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libaqua.so")

OK = 0
ERR_NAMES = {
    1: "IO", 2: "PARSE", 3: "MISSING_MESH", 4: "NO_DEVICE", 5: "OOM", 6: "HIP",
    7: "INVALID_ARG", 8: "NONFINITE", 9: "UNSUPPORTED",
}
ERR_IO, ERR_PARSE, ERR_MISSING_MESH, ERR_NO_DEVICE = 1, 2, 3, 4
ERR_OOM, ERR_HIP, ERR_INVALID_ARG, ERR_NONFINITE, ERR_UNSUPPORTED = 5, 6, 7, 8, 9

INTEGRATOR_PT, INTEGRATOR_NRC = 0, 1
SAMPLER_INDEPENDENT, SAMPLER_STRATIFIED = 0, 1
PROCGEN_BOX_ROOM, PROCGEN_INTERIOR = 1, 2
MISS = 0xFFFFFFFF


class AquaError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"aqua error {ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


class LoadOpts(C.Structure):
    _fields_ = [("skip_missing", C.c_uint32), ("image_as_constant", C.c_uint32),
                ("emission_from_mtl", C.c_uint32), ("reserved", C.c_uint32 * 5)]


class Material(C.Structure):
    _fields_ = [("base_color", C.c_float * 3), ("metallic", C.c_float),
                ("emission", C.c_float * 3), ("roughness", C.c_float),
                ("specular", C.c_float), ("base_color_tex", C.c_uint32), ("pad", C.c_float * 2)]


class Camera(C.Structure):
    _fields_ = [("cam_to_world", C.c_float * 12), ("tan_half_fov", C.c_float),
                ("aspect", C.c_float), ("width", C.c_uint32), ("height", C.c_uint32)]


class PointLight(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("intensity", C.c_float * 3)]


class SceneInfo(C.Structure):
    _fields_ = [("n_triangles", C.c_uint32), ("n_materials", C.c_uint32),
                ("n_shapes", C.c_uint32), ("n_skipped_shapes", C.c_uint32),
                ("n_emissive_triangles", C.c_uint32), ("n_point_lights", C.c_uint32),
                ("n_image_textures", C.c_uint32), ("n_vertices", C.c_uint32),
                ("bounds_lo", C.c_float * 3), ("bounds_hi", C.c_float * 3),
                ("n_images", C.c_uint32), ("has_texcoords", C.c_uint32)]


class ProcgenParams(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("seed", C.c_uint32), ("n_meshes", C.c_uint32),
                ("target_tris", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32),
                ("reserved", C.c_uint32 * 2)]


class IntegratorCfg(C.Structure):
    _fields_ = [("type", C.c_uint32), ("spp", C.c_uint32), ("max_depth", C.c_uint32),
                ("batch_size", C.c_uint32), ("training_iters", C.c_uint32),
                ("learning_rate", C.c_float), ("visualize_cache", C.c_uint32),
                ("seed", C.c_uint32), ("rr_depth", C.c_uint32), ("nrc_precision", C.c_uint32),
                ("nrc_train_paths", C.c_uint32), ("nrc_query_vertex", C.c_uint32),
                ("sampler", C.c_uint32), ("reserved", C.c_uint32 * 3)]


class FrameStats(C.Structure):
    _fields_ = [("camera_rays", C.c_uint64), ("extension_rays", C.c_uint64),
                ("shadow_rays", C.c_uint64), ("nonfinite_samples", C.c_uint64),
                ("ms_frame", C.c_double), ("ms_intersect", C.c_double),
                ("ms_shade", C.c_double), ("ms_shadow", C.c_double),
                ("n_intersect_launches", C.c_uint32), ("n_passes", C.c_uint32),
                ("nodes_visited", C.c_uint64), ("tris_tested", C.c_uint64),
                ("shadow_nodes_visited", C.c_uint64), ("shadow_tris_tested", C.c_uint64),
                ("nrc_queries", C.c_uint32), ("nrc_train_records", C.c_uint32),
                ("nrc_loss", C.c_float), ("ms_nrc", C.c_float), ("spill_node_visits", C.c_uint64)]

    @property
    def rays(self) -> int:
        return self.camera_rays + self.extension_rays + self.shadow_rays


class RendererOpts(C.Structure):
    _fields_ = [("samples_per_pass", C.c_uint32), ("count_traversal", C.c_uint32),
                ("timing", C.c_uint32), ("sort_materials", C.c_uint32), ("train_ahead", C.c_uint32),
                ("reserved", C.c_uint32 * 3)]


NRC_F16, NRC_BF16, NRC_FP8 = 0, 1, 2


class NrcCfg(C.Structure):
    _fields_ = [("n_levels", C.c_uint32), ("log2_table", C.c_uint32), ("base_res", C.c_uint32),
                ("level_scale", C.c_float), ("learning_rate", C.c_float), ("beta1", C.c_float),
                ("beta2", C.c_float), ("eps", C.c_float), ("train_precision", C.c_uint32),
                ("seed", C.c_uint32), ("max_batch", C.c_uint32), ("reserved", C.c_uint32 * 5)]


class NrcStats(C.Structure):
    _fields_ = [("ms_total", C.c_double), ("ms_fwd_bwd", C.c_double), ("ms_wgrad", C.c_double),
                ("ms_adam", C.c_double), ("loss", C.c_float), ("steps", C.c_uint32),
                ("rows", C.c_uint32), ("ms_scatter", C.c_float), ("reserved", C.c_uint32 * 4)]


assert C.sizeof(NrcCfg) == 64 and C.sizeof(NrcStats) == 64
assert C.sizeof(LoadOpts) == 32 and C.sizeof(Material) == 48 and C.sizeof(Camera) == 64
assert C.sizeof(SceneInfo) == 64 and C.sizeof(ProcgenParams) == 32
assert C.sizeof(IntegratorCfg) == 64 and C.sizeof(FrameStats) == 128 and C.sizeof(RendererOpts) == 32

# name -> (restype, argtypes)
_P = C.c_void_p
_SIGS = {
    "aqua_last_error": (C.c_char_p, []),
    "aqua_abi_version": (C.c_int, []),
    "aqua_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "aqua_scene_load": (C.c_int, [C.c_char_p, C.POINTER(LoadOpts), C.POINTER(_P)]),
    "aqua_scene_from_arrays": (C.c_int, [C.c_uint32, _P, _P, C.c_uint32, _P, C.POINTER(Camera),
                                         C.c_uint32, _P, C.POINTER(_P)]),
    "aqua_scene_get_info": (C.c_int, [_P, C.POINTER(SceneInfo)]),
    "aqua_scene_get_triangles": (C.c_int, [_P, _P, _P]),
    "aqua_scene_get_materials": (C.c_int, [_P, _P]),
    "aqua_scene_get_camera": (C.c_int, [_P, C.POINTER(Camera)]),
    "aqua_scene_get_point_lights": (C.c_int, [_P, _P]),
    "aqua_scene_get_texcoords": (C.c_int, [_P, _P]),
    "aqua_scene_get_image": (C.c_int, [_P, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), _P]),
    "aqua_scene_set_texcoords": (C.c_int, [_P, _P]),
    "aqua_scene_add_image": (C.c_int, [_P, C.c_uint32, C.c_uint32, _P, C.POINTER(C.c_uint32)]),
    "aqua_image_load": (C.c_int, [C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), _P]),
    "aqua_image_write_png": (C.c_int, [C.c_char_p, C.c_uint32, C.c_uint32, C.c_uint32, _P]),
    "aqua_image_write_pfm": (C.c_int, [C.c_char_p, C.c_uint32, C.c_uint32, _P]),
    "aqua_film_tonemap": (C.c_int, [_P, C.c_float, _P]),
    "aqua_scene_destroy": (None, [_P]),
    "aqua_scene_bvh": (C.c_int, [_P, C.POINTER(C.c_uint32), _P, _P, C.POINTER(C.c_uint32)]),
    "aqua_scene_bvh_device": (C.c_int, [_P, C.POINTER(C.c_uint32), _P]),
    "aqua_procgen_write": (C.c_int, [C.POINTER(ProcgenParams), C.c_char_p, C.c_char_p, C.c_size_t]),
    "aqua_integrator_default": (None, [C.POINTER(IntegratorCfg)]),
    "aqua_integrator_load": (C.c_int, [C.c_char_p, C.POINTER(IntegratorCfg)]),
    "aqua_renderer_create": (C.c_int, [_P, C.POINTER(IntegratorCfg), C.c_int, C.POINTER(_P)]),
    "aqua_renderer_set_opts": (C.c_int, [_P, C.POINTER(RendererOpts)]),
    "aqua_render_frame": (C.c_int, [_P, C.POINTER(Camera), C.c_uint32, C.c_uint32,
                                    C.POINTER(FrameStats)]),
    "aqua_film_read": (C.c_int, [_P, _P, C.POINTER(C.c_uint32)]),
    "aqua_render_oracle": (C.c_int, [_P, C.POINTER(Camera), C.c_uint32, C.c_uint32, C.c_uint32]),
    "aqua_oracle_read": (C.c_int, [_P, _P, C.POINTER(C.c_uint32)]),
    "aqua_film_error": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "aqua_primary_hits": (C.c_int, [_P, C.POINTER(Camera), C.c_uint32, C.c_uint32, _P, _P]),
    "aqua_trace_closest": (C.c_int, [_P, C.c_uint32, _P, _P, _P]),
    "aqua_trace_closest_noalloc": (C.c_int, [_P, C.c_uint32, _P, _P, _P]),
    "aqua_trace_occluded": (C.c_int, [_P, C.c_uint32, _P, _P]),
    "aqua_renderer_destroy": (None, [_P]),
    "aqua_fp_probe": (C.c_int, [C.c_uint32, _P, _P, _P, _P, _P]),
    "aqua_nrc_default_cfg": (None, [C.POINTER(NrcCfg)]),
    "aqua_nrc_create": (C.c_int, [C.c_int, C.POINTER(NrcCfg), C.POINTER(_P)]),
    "aqua_nrc_destroy": (None, [_P]),
    "aqua_nrc_param_counts": (C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]),
    "aqua_nrc_get_params": (C.c_int, [_P, _P, _P]),
    "aqua_nrc_set_params": (C.c_int, [_P, _P, _P]),
    "aqua_nrc_fp8_scales": (C.c_int, [_P, _P, _P]),
    "aqua_nrc_query_order": (C.c_int, [_P, C.c_uint32, _P, _P]),
    "aqua_nrc_infer": (C.c_int, [_P, C.c_uint32, _P, C.c_uint32, _P]),
    "aqua_nrc_train": (C.c_int, [_P, C.c_uint32, _P, _P, C.c_uint32, C.POINTER(NrcStats)]),
    "aqua_nrc_train_device": (C.c_int, [_P, C.c_uint32, _P, _P, C.c_uint32, C.POINTER(NrcStats)]),
    "aqua_nrc_gradients": (C.c_int, [_P, C.c_uint32, _P, _P, _P, _P, C.POINTER(C.c_float)]),
    "aqua_renderer_nrc": (_P, [_P]),
    "aqua_renderer_set_nrc_enabled": (C.c_int, [_P, C.c_int]),
    "aqua_renderer_set_nrc_precision": (C.c_int, [_P, C.c_uint32]),
    "aqua_renderer_nrc_checkpoint": (C.c_int, [_P, C.c_int]),
    "aqua_nrc_checkpoint": (C.c_int, [_P, C.c_int]),
    "aqua_renderer_nrc_batch": (C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(_P), C.POINTER(_P)]),
    "aqua_camera_flythrough": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_uint32,
                                         C.c_uint32, C.POINTER(Camera)]),
}

_lib: Optional[C.CDLL] = None


def lib() -> C.CDLL:
    """Loads libaqua.so (built in-tree by `make -C aqua-engine_amd/csrc`)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `make -C aqua-engine_amd/csrc`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(code: int) -> None:
    if code != OK:
        raise AquaError(code, lib().aqua_last_error().decode(errors="replace"))


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def device_count() -> int:
    n = C.c_int(0)
    _check(lib().aqua_device_count(C.byref(n)))
    return n.value


def camera_to_array(cam: Camera) -> np.ndarray:
    """16 float32: matrix, tan_half, aspect, width bits, height bits."""
    return np.frombuffer(bytes(cam), dtype=np.float32).copy()


class Scene:
    """A loaded scene (aqua_scene): the flattened triangles, materials, camera
    and point lights of a reference-schema scene JSON."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle

    @classmethod
    def load(cls, path: str, skip_missing: bool = False, image_as_constant: bool = False,
             emission_from_mtl: bool = False) -> "Scene":
        o = LoadOpts(int(skip_missing), int(image_as_constant), int(emission_from_mtl))
        h = C.c_void_p()
        _check(lib().aqua_scene_load(path.encode(), C.byref(o), C.byref(h)))
        return cls(h)

    @classmethod
    def from_arrays(cls, verts: np.ndarray, mat_ids: np.ndarray, materials: list, camera: Camera,
                    point_lights: Optional[list] = None) -> "Scene":
        v = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 9)
        m = np.ascontiguousarray(mat_ids, dtype=np.uint32)
        mats = (Material * max(1, len(materials)))(*materials)
        pls = point_lights or []
        pla = (PointLight * max(1, len(pls)))(*pls)
        h = C.c_void_p()
        _check(lib().aqua_scene_from_arrays(len(m), _ptr(v), _ptr(m), len(materials),
                                            C.cast(mats, C.c_void_p), C.byref(camera), len(pls),
                                            C.cast(pla, C.c_void_p), C.byref(h)))
        return cls(h)

    def close(self) -> None:
        if self._h:
            lib().aqua_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def info(self) -> SceneInfo:
        i = SceneInfo()
        _check(lib().aqua_scene_get_info(self._h, C.byref(i)))
        return i

    def triangles(self):
        n = self.info().n_triangles
        v = np.zeros((n, 9), np.float32)
        m = np.zeros(n, np.uint32)
        _check(lib().aqua_scene_get_triangles(self._h, _ptr(v), _ptr(m)))
        return v, m

    def materials(self) -> np.ndarray:
        n = self.info().n_materials
        a = np.zeros((n, 12), np.float32)
        _check(lib().aqua_scene_get_materials(self._h, _ptr(a)))
        return a

    def camera(self) -> Camera:
        c = Camera()
        _check(lib().aqua_scene_get_camera(self._h, C.byref(c)))
        return c

    def point_lights(self) -> np.ndarray:
        n = self.info().n_point_lights
        a = np.zeros((n, 6), np.float32)
        _check(lib().aqua_scene_get_point_lights(self._h, _ptr(a) if n else None))
        return a

    def texcoords(self) -> np.ndarray:
        """float32[n_triangles, 6]: (u, v) of each corner, load order."""
        uv = np.zeros((self.info().n_triangles, 6), np.float32)
        _check(lib().aqua_scene_get_texcoords(self._h, _ptr(uv)))
        return uv

    def image(self, k: int) -> np.ndarray:
        """uint8[h, w, 4]: image k's sRGB RGBA texels, row 0 = top."""
        w, h = C.c_uint32(), C.c_uint32()
        _check(lib().aqua_scene_get_image(self._h, k, C.byref(w), C.byref(h), None))
        px = np.zeros((h.value, w.value, 4), np.uint8)
        _check(lib().aqua_scene_get_image(self._h, k, C.byref(w), C.byref(h), _ptr(px)))
        return px

    def set_texcoords(self, uv: np.ndarray) -> None:
        a = np.ascontiguousarray(uv, np.float32).reshape(-1, 6)
        if len(a) != self.info().n_triangles:
            raise ValueError("texcoords: one row of 6 per triangle")
        _check(lib().aqua_scene_set_texcoords(self._h, _ptr(a)))

    def add_image(self, rgba: np.ndarray) -> int:
        """Adds an 8-bit sRGB RGBA image (h, w, 4); materials name it as base_color_tex = index + 1."""
        a = np.ascontiguousarray(rgba, np.uint8)
        if a.ndim != 3 or a.shape[2] != 4:
            raise ValueError("image must be (h, w, 4) uint8")
        idx = C.c_uint32()
        _check(lib().aqua_scene_add_image(self._h, a.shape[1], a.shape[0], _ptr(a), C.byref(idx)))
        return idx.value

    def bvh_device(self) -> np.ndarray:
        """The same tree as stored on the device: uint32[n, 16] quantised 64-byte nodes."""
        n = C.c_uint32(0)
        _check(lib().aqua_scene_bvh_device(self._h, C.byref(n), None))
        words = np.zeros((n.value, 16), np.uint32)
        _check(lib().aqua_scene_bvh_device(self._h, C.byref(n), _ptr(words)))
        return words

    def bvh(self):
        """(nodes float32[n,32]: the exact 4-wide tree, order uint32[ntri], max_depth)."""
        n = C.c_uint32(0)
        d = C.c_uint32(0)
        _check(lib().aqua_scene_bvh(self._h, C.byref(n), None, None, C.byref(d)))
        nodes = np.zeros((n.value, 32), np.float32)
        order = np.zeros(self.info().n_triangles, np.uint32)
        _check(lib().aqua_scene_bvh(self._h, C.byref(n), _ptr(nodes), _ptr(order), C.byref(d)))
        return nodes, order, d.value


def image_load(path: str) -> np.ndarray:
    """uint8[h, w, 4]: a JPEG file decoded to RGBA, row 0 = top."""
    w, h = C.c_uint32(), C.c_uint32()
    _check(lib().aqua_image_load(path.encode(), C.byref(w), C.byref(h), None))
    px = np.zeros((h.value, w.value, 4), np.uint8)
    _check(lib().aqua_image_load(path.encode(), C.byref(w), C.byref(h), _ptr(px)))
    return px


def image_write_png(path: str, pixels: np.ndarray) -> None:
    """Writes uint8 (h, w, 3) or (h, w, 4) pixels as PNG."""
    a = np.ascontiguousarray(pixels, np.uint8)
    _check(lib().aqua_image_write_png(path.encode(), a.shape[1], a.shape[0], a.shape[2], _ptr(a)))


def image_write_pfm(path: str, rgb: np.ndarray) -> None:
    """Writes float (h, w, 3) radiance as PFM."""
    a = np.ascontiguousarray(rgb, np.float32)
    _check(lib().aqua_image_write_pfm(path.encode(), a.shape[1], a.shape[0], _ptr(a)))


def procgen(kind: int, out_dir: str, seed: int = 0, n_meshes: int = 0, target_tris: int = 0,
            width: int = 0, height: int = 0) -> str:
    """Writes a procedural scene (reference schema: scene.json + .mesh files)."""
    p = ProcgenParams(kind, seed, n_meshes, target_tris, width, height)
    buf = C.create_string_buffer(4096)
    _check(lib().aqua_procgen_write(C.byref(p), out_dir.encode(), buf, 4096))
    return buf.value.decode()


def camera_flythrough(scene: "Scene", frame: int, n_frames: int, seed: int = 0, fov: float = 60.0,
                      width: int = 1920, height: int = 1080) -> Camera:
    c = Camera()
    _check(lib().aqua_camera_flythrough(scene.handle, seed, frame, n_frames, fov, width, height, C.byref(c)))
    return c


def integrator_default() -> IntegratorCfg:
    c = IntegratorCfg()
    lib().aqua_integrator_default(C.byref(c))
    return c


def integrator_load(path: str) -> IntegratorCfg:
    c = IntegratorCfg()
    _check(lib().aqua_integrator_load(path.encode(), C.byref(c)))
    return c


class Renderer:
    """aqua_renderer: the scene resident in HBM plus path queues and films."""

    def __init__(self, scene: Scene, cfg: IntegratorCfg, device: int = 0):
        self._h = C.c_void_p()
        _check(lib().aqua_renderer_create(scene.handle, C.byref(cfg), device, C.byref(self._h)))
        self.scene_camera = scene.camera()
        self.cfg = cfg

    def close(self) -> None:
        if self._h:
            lib().aqua_renderer_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_opts(self, samples_per_pass: int = 0, count_traversal: bool = False, timing: bool = True,
                 sort_materials: bool = False, train_ahead: bool = False):
        o = RendererOpts(samples_per_pass, int(count_traversal), int(timing), int(sort_materials),
                         int(train_ahead))
        _check(lib().aqua_renderer_set_opts(self._h, C.byref(o)))

    def render(self, spp: int, frame: int = 0, camera: Optional[Camera] = None) -> FrameStats:
        st = FrameStats()
        _check(lib().aqua_render_frame(self._h, C.byref(camera) if camera else None, frame, spp,
                                       C.byref(st)))
        return st

    def film(self, camera: Optional[Camera] = None) -> np.ndarray:
        cam = camera or self.scene_camera
        out = np.zeros((cam.height, cam.width, 3), np.float32)
        spp = C.c_uint32(0)
        _check(lib().aqua_film_read(self._h, _ptr(out), C.byref(spp)))
        return out

    def tonemap(self, exposure: float = 0.0, camera: Optional[Camera] = None) -> np.ndarray:
        """uint8[h, w, 4]: the last frame x 2^exposure, clamped, sRGB-encoded (GPU)."""
        cam = camera or self.scene_camera
        out = np.zeros((cam.height, cam.width, 4), np.uint8)
        _check(lib().aqua_film_tonemap(self._h, exposure, _ptr(out)))
        return out

    def render_oracle(self, spp_begin: int, spp_count: int, frame: int = 0,
                      camera: Optional[Camera] = None) -> None:
        _check(lib().aqua_render_oracle(self._h, C.byref(camera) if camera else None, frame,
                                        spp_begin, spp_count))

    def oracle_film(self, camera: Optional[Camera] = None) -> np.ndarray:
        cam = camera or self.scene_camera
        out = np.zeros((cam.height, cam.width, 3), np.float64)
        spp = C.c_uint32(0)
        _check(lib().aqua_oracle_read(self._h, _ptr(out), C.byref(spp)))
        return out

    def film_error(self):
        a, b = C.c_double(), C.c_double()
        _check(lib().aqua_film_error(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def primary_hits(self, sample: int = 0, frame: int = 0, camera: Optional[Camera] = None):
        cam = camera or self.scene_camera
        n = cam.width * cam.height
        ids = np.zeros(n, np.uint32)
        t = np.zeros(n, np.float32)
        _check(lib().aqua_primary_hits(self._h, C.byref(camera) if camera else None, frame, sample,
                                       _ptr(ids), _ptr(t)))
        return ids.reshape(cam.height, cam.width), t.reshape(cam.height, cam.width)

    def trace_closest(self, rays: np.ndarray):
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        ids = np.zeros(len(r), np.uint32)
        t = np.zeros(len(r), np.float32)
        _check(lib().aqua_trace_closest(self._h, len(r), _ptr(r), _ptr(ids), _ptr(t)))
        return ids, t

    def trace_closest_noalloc(self, n: int, d_rays: int, d_prim_ids: int, d_t_hit: int) -> None:
        """aqua_trace_closest_noalloc on caller-owned device arrays (addresses)."""
        _check(lib().aqua_trace_closest_noalloc(self._h, n, d_rays, d_prim_ids, d_t_hit))

    def set_nrc_enabled(self, enabled: bool) -> None:
        _check(lib().aqua_renderer_set_nrc_enabled(self._h, int(enabled)))

    def set_nrc_precision(self, precision: int) -> None:
        """Inference precision of the cache from the next frame on (same weights)."""
        _check(lib().aqua_renderer_set_nrc_precision(self._h, precision))

    def nrc_checkpoint(self, restore: bool = False) -> None:
        """Saves (restore False) or puts back (True) the cache's training state."""
        _check(lib().aqua_renderer_nrc_checkpoint(self._h, int(restore)))

    def nrc_handle(self):
        return lib().aqua_renderer_nrc(self._h)

    def nrc_batch(self):
        """(n, device pointer of records, device pointer of targets) of the last frame."""
        n = C.c_uint32()
        a, b = C.c_void_p(), C.c_void_p()
        _check(lib().aqua_renderer_nrc_batch(self._h, C.byref(n), C.byref(a), C.byref(b)))
        return n.value, a.value, b.value

    def nrc_train_batch(self, steps: int = 1) -> "NrcStats":
        """Adam steps of the cache on the last frame's training batch (device-resident)."""
        n, a, b = self.nrc_batch()
        st = NrcStats()
        _check(lib().aqua_nrc_train_device(self.nrc_handle(), n, a, b, steps, C.byref(st)))
        return st

    def trace_occluded(self, rays: np.ndarray) -> np.ndarray:
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        occ = np.zeros(len(r), np.uint8)
        _check(lib().aqua_trace_occluded(self._h, len(r), _ptr(r), _ptr(occ)))
        return occ


def nrc_default_cfg() -> NrcCfg:
    c = NrcCfg()
    lib().aqua_nrc_default_cfg(C.byref(c))
    return c


class Nrc:
    """aqua_nrc: the neural radiance cache on one device."""

    def __init__(self, cfg: Optional[NrcCfg] = None, device: int = 0):
        self.cfg = cfg or nrc_default_cfg()
        self._h = C.c_void_p()
        _check(lib().aqua_nrc_create(device, C.byref(self.cfg), C.byref(self._h)))

    def close(self) -> None:
        if self._h:
            lib().aqua_nrc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def counts(self):
        t, m = C.c_uint64(), C.c_uint32()
        _check(lib().aqua_nrc_param_counts(self._h, C.byref(t), C.byref(m)))
        return t.value, m.value

    def params(self):
        nt, nm = self.counts()
        table = np.zeros(nt, np.float32)
        mlp = np.zeros(nm, np.float32)
        _check(lib().aqua_nrc_get_params(self._h, _ptr(table), _ptr(mlp)))
        return table, mlp

    def set_params(self, table, mlp):
        table = np.ascontiguousarray(table, np.float32)
        mlp = np.ascontiguousarray(mlp, np.float32)
        _check(lib().aqua_nrc_set_params(self._h, _ptr(table), _ptr(mlp)))

    def fp8_scales(self):
        """(wscale[5], ascale[5]) of the FP8 query path."""
        w = np.zeros(5, np.float32)
        a = np.zeros(5, np.float32)
        _check(lib().aqua_nrc_fp8_scales(self._h, _ptr(w), _ptr(a)))
        return w, a

    def query_order(self, records) -> np.ndarray:
        """The order the cache evaluates the records in (stable Morton sort)."""
        rec = np.ascontiguousarray(records, np.float32).reshape(-1, 16)
        out = np.zeros(len(rec), np.uint32)
        _check(lib().aqua_nrc_query_order(self._h, len(rec), _ptr(rec), _ptr(out)))
        return out

    def infer(self, records, precision=NRC_F16) -> np.ndarray:
        rec = np.ascontiguousarray(records, np.float32).reshape(-1, 16)
        out = np.zeros((len(rec), 3), np.float32)
        _check(lib().aqua_nrc_infer(self._h, len(rec), _ptr(rec), precision, _ptr(out)))
        return out

    def gradients(self, records, targets):
        rec = np.ascontiguousarray(records, np.float32).reshape(-1, 16)
        tgt = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
        nt, nm = self.counts()
        gt = np.zeros(nt, np.float32)
        gm = np.zeros(nm, np.float32)
        loss = C.c_float()
        _check(lib().aqua_nrc_gradients(self._h, len(rec), _ptr(rec), _ptr(tgt), _ptr(gt), _ptr(gm),
                                        C.byref(loss)))
        return loss.value, gt, gm

    def train(self, records, targets, steps=1) -> NrcStats:
        rec = np.ascontiguousarray(records, np.float32).reshape(-1, 16)
        tgt = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
        st = NrcStats()
        _check(lib().aqua_nrc_train(self._h, len(rec), _ptr(rec), _ptr(tgt), steps, C.byref(st)))
        return st


def fp_probe(a, b, c, u) -> np.ndarray:
    a, b, c, u = (np.ascontiguousarray(x, dtype=np.float32) for x in (a, b, c, u))
    out = np.zeros((len(a), 5), np.float32)
    _check(lib().aqua_fp_probe(len(a), _ptr(a), _ptr(b), _ptr(c), _ptr(u), _ptr(out)))
    return out
