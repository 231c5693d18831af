"""ctypes binding of the CPU restatement in oracle.c.

TEST INFRASTRUCTURE: only tests/, the smoke check and the CPU-baseline leg of
bench.py import this module, and only as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


class Cfg(C.Structure):
    _fields_ = [("max_depth", C.c_uint32), ("rr_depth", C.c_uint32), ("seed", C.c_uint32),
                ("threads", C.c_uint32), ("sampler", C.c_uint32)]


class Stats(C.Structure):
    _fields_ = [("camera_rays", C.c_uint64), ("extension_rays", C.c_uint64),
                ("shadow_rays", C.c_uint64), ("nonfinite", C.c_uint64),
                ("nodes", C.c_uint64), ("tris", C.c_uint64)]

    @property
    def rays(self) -> int:
        return self.camera_rays + self.extension_rays + self.shadow_rays


_P = C.c_void_p
_SIGS = {
    "orc_scene_create": (C.c_int, [C.c_uint32, _P, _P, C.c_uint32, _P, C.c_uint32, _P, C.POINTER(_P)]),
    "orc_scene_destroy": (None, [_P]),
    "orc_scene_set_textures": (C.c_int, [_P, _P, C.c_uint32, _P, _P]),
    "orc_scene_bvh_info": (None, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "orc_render": (C.c_int, [_P, _P, C.POINTER(Cfg), C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                             _P, C.POINTER(Stats)]),
    "orc_render64": (C.c_int, [_P, _P, C.POINTER(Cfg), C.c_uint32, C.c_uint32, _P]),
    "orc_primary_hits": (C.c_int, [_P, _P, C.POINTER(Cfg), C.c_uint32, C.c_uint32, _P, _P]),
    "orc_primary_hits_rows": (C.c_int, [_P, _P, C.POINTER(Cfg), C.c_uint32, C.c_uint32, C.c_uint32,
                                        C.c_uint32, _P, _P]),
    "orc_trace_closest": (C.c_int, [_P, C.c_uint32, _P, C.c_int, _P, _P]),
    "orc_trace_occluded": (C.c_int, [_P, C.c_uint32, _P, C.c_int, _P]),
    "orc_pcg32_seq": (None, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32, _P]),
    "orc_strat2_seq": (None, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _P]),
    "orc_sincos_2pi": (None, [C.c_uint32, _P, _P, _P]),
    "orc_bsdf_eval": (None, [_P, _P, _P, _P, _P]),
    "orc_bsdf_sample": (None, [_P, _P, _P, C.c_float, C.c_float, C.c_float, _P]),
    "orc_light_table": (C.c_uint32, [_P, _P, _P, _P, _P]),
    "orc_ggx": (None, [C.c_float, C.c_float, C.c_float, _P]),
    "orc_camera_ray": (None, [_P, C.c_uint32, C.c_uint32, C.c_float, C.c_float, _P]),
}

_lib: Optional[C.CDLL] = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Scene:
    def __init__(self, verts, mat_ids, materials, point_lights=None, uvs=None, images=()):
        """uvs: float32 (n, 6); images: uint8 (h, w, 4) sRGB RGBA arrays that
        materials name as base_color_tex = index + 1 (slot 9 as uint32 bits)."""
        self.verts = f32(verts).reshape(-1, 9)
        self.mat_ids = np.ascontiguousarray(mat_ids, dtype=np.uint32)
        self.materials = f32(materials).reshape(-1, 12)
        pl = np.zeros((0, 6), np.float32) if point_lights is None else f32(point_lights).reshape(-1, 6)
        self.points = pl
        self._h = C.c_void_p()
        rc = lib().orc_scene_create(len(self.mat_ids), _p(self.verts), _p(self.mat_ids),
                                    len(self.materials), _p(self.materials), len(pl),
                                    _p(pl) if len(pl) else None, C.byref(self._h))
        if rc != 0:
            raise MemoryError("orc_scene_create failed")
        if uvs is not None:
            self.uvs = f32(uvs).reshape(-1, 6)
            dims = np.array([[im.shape[1], im.shape[0]] for im in images], np.uint32).reshape(-1)
            tex = np.concatenate([np.ascontiguousarray(im, np.uint8).reshape(-1) for im in images]) \
                if len(images) else np.zeros(4, np.uint8)
            if lib().orc_scene_set_textures(self._h, _p(self.uvs), len(images), _p(dims), _p(tex)) != 0:
                raise MemoryError("orc_scene_set_textures failed")

    def __del__(self):
        try:
            if self._h:
                lib().orc_scene_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def bvh_info(self):
        n, d = C.c_uint32(), C.c_uint32()
        lib().orc_scene_bvh_info(self._h, C.byref(n), C.byref(d))
        return n.value, d.value

    def render(self, camera16, spp, max_depth, rr_depth=3, seed=0, frame=0, y0=0, y1=None, threads=0, sampler=0):
        cam = f32(camera16)
        w = int(cam[14:15].view(np.uint32)[0])
        h = int(cam[15:16].view(np.uint32)[0])
        y1 = h if y1 is None else y1
        out = np.zeros((y1 - y0, w, 3), np.float32)
        st = Stats()
        cfg = Cfg(max_depth, rr_depth, seed, threads, sampler)
        rc = lib().orc_render(self._h, _p(cam), C.byref(cfg), frame, spp, y0, y1, _p(out), C.byref(st))
        if rc != 0:
            raise ValueError("orc_render: bad arguments")
        return out, st

    def render64(self, camera16, spp, max_depth, rr_depth=3, seed=0, frame=0, threads=0, sampler=0):
        cam = f32(camera16)
        w = int(cam[14:15].view(np.uint32)[0])
        h = int(cam[15:16].view(np.uint32)[0])
        out = np.zeros((h, w, 3), np.float64)
        cfg = Cfg(max_depth, rr_depth, seed, threads, sampler)
        if lib().orc_render64(self._h, _p(cam), C.byref(cfg), frame, spp, _p(out)) != 0:
            raise ValueError("orc_render64: bad arguments")
        return out

    def primary_hits(self, camera16, sample=0, frame=0, max_depth=5, seed=0, y0=0, y1=None, threads=0,
                     sampler=0):
        """First-hit ids and distances of rows [y0, y1) (all rows by default)."""
        cam = f32(camera16)
        w = int(cam[14:15].view(np.uint32)[0])
        h = int(cam[15:16].view(np.uint32)[0])
        y1 = h if y1 is None else y1
        ids = np.zeros((y1 - y0) * w, np.uint32)
        t = np.zeros((y1 - y0) * w, np.float32)
        cfg = Cfg(max_depth, 3, seed, threads, sampler)
        if lib().orc_primary_hits_rows(self._h, _p(cam), C.byref(cfg), frame, sample, y0, y1, _p(ids), _p(t)):
            raise ValueError("orc_primary_hits_rows: bad arguments")
        return ids.reshape(y1 - y0, w), t.reshape(y1 - y0, w)

    def trace_closest(self, rays, brute=False):
        r = f32(rays).reshape(-1, 8)
        ids = np.zeros(len(r), np.uint32)
        t = np.zeros(len(r), np.float32)
        lib().orc_trace_closest(self._h, len(r), _p(r), int(brute), _p(ids), _p(t))
        return ids, t

    def trace_occluded(self, rays, brute=False):
        r = f32(rays).reshape(-1, 8)
        occ = np.zeros(len(r), np.uint8)
        lib().orc_trace_occluded(self._h, len(r), _p(r), int(brute), _p(occ))
        return occ

    def light_table(self):
        n_all = len(self.points) + len(self.mat_ids)
        pmf = np.zeros(n_all + 1, np.float32)
        cdf = np.zeros(n_all + 1, np.float32)
        area = np.zeros(n_all + 1, np.float32)
        mp = np.zeros(len(self.materials), np.float32)
        n = lib().orc_light_table(self._h, _p(pmf), _p(cdf), _p(area), _p(mp))
        return n, pmf, cdf, area, mp


def strat2(seed, frame, pixel, pair, n, first=0):
    """Points first .. first + n - 1 of the stratified sampler's 2-D pattern
    `pair` of one pixel and frame, as an (n, 2) float32 array."""
    out = np.zeros((n, 2), np.float32)
    lib().orc_strat2_seq(seed, frame, pixel, pair, first, n, _p(out))
    return out


def pcg32(initstate, initseq, n, advance=0):
    out = np.zeros(n, np.uint32)
    lib().orc_pcg32_seq(initstate, initseq, advance, n, _p(out))
    return out


def sincos_2pi(u):
    u = f32(u)
    s = np.zeros_like(u)
    c = np.zeros_like(u)
    lib().orc_sincos_2pi(len(u), _p(u), _p(s), _p(c))
    return s, c


def bsdf_eval(mat12, n, wo, wi):
    out = np.zeros(5, np.float32)
    lib().orc_bsdf_eval(_p(f32(mat12)), _p(f32(n)), _p(f32(wo)), _p(f32(wi)), _p(out))
    return out


def ggx(roughness, cos_h, cos_v):
    out = np.zeros(2, np.float32)
    lib().orc_ggx(roughness, cos_h, cos_v, _p(out))
    return float(out[0]), float(out[1])


def camera_ray(camera16, px, py, u0, u1):
    out = np.zeros(6, np.float32)
    lib().orc_camera_ray(_p(f32(camera16)), px, py, u0, u1, _p(out))
    return out


def bsdf_sample(mat12, n, wo, ul, u1, u2):
    out = np.zeros(10, np.float32)
    lib().orc_bsdf_sample(_p(f32(mat12)), _p(f32(n)), _p(f32(wo)), ul, u1, u2, _p(out))
    return out
