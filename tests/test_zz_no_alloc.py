"""aqua_trace_closest_noalloc (the allocation-free closest-hit entry point)
against aqua_trace_closest on the same rays: bit-exact ids and distances, on
two calls in a row (the second smaller, so stale scratch from the first must
not leak into it), with caller-owned device arrays allocated through the
same HIP runtime the library uses."""
import ctypes as C

import numpy as np
import pytest

from conftest import aqua, oracle_scene_of

pytestmark = pytest.mark.gpu


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime libaqua.so links."""

    H2D, D2H = 1, 2

    def __init__(self):
        aqua.lib()
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipSetDevice.argtypes = [C.c_int]

    def check(self, rc, what):
        assert rc == 0, f"{what} failed: {rc}"

    def alloc(self, nbytes):
        p = C.c_void_p()
        self.check(self.L.hipMalloc(C.byref(p), nbytes), "hipMalloc")
        return p.value

    def upload(self, ptr, a):
        a = np.ascontiguousarray(a)
        self.check(self.L.hipMemcpy(ptr, a.ctypes.data, a.nbytes, self.H2D), "hipMemcpy H2D")

    def download(self, ptr, a):
        self.check(self.L.hipMemcpy(a.ctypes.data, ptr, a.nbytes, self.D2H), "hipMemcpy D2H")
        return a

    def free(self, ptr):
        self.check(self.L.hipFree(ptr), "hipFree")


def _rays(scene, n, seed):
    info = scene.info()
    lo, hi = np.array(info.bounds_lo), np.array(info.bounds_hi)
    rng = np.random.default_rng(seed)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = rng.uniform(lo, hi, (n, 3))
    r[:, 3] = np.where(rng.random(n) < 0.2, rng.uniform(0.01, 1.0, n), 3.0e38)
    d = rng.standard_normal((n, 3))
    d[: n // 50, 1:] = 0.0
    r[:, 4:7] = d
    return r


def test_trace_closest_noalloc_matches_trace_closest(gpu, tmp_path):
    path = aqua.procgen(aqua.PROCGEN_INTERIOR, str(tmp_path), seed=2, n_meshes=40, target_tris=200000,
                        width=64, height=64)
    sc = aqua.Scene.load(path)
    c = aqua.integrator_default()
    r = aqua.Renderer(sc, c)
    hip = Hip()
    hip.check(hip.L.hipSetDevice(0), "hipSetDevice")
    n1, n2 = 300000, 1234
    d_rays = hip.alloc(n1 * 32)
    d_ids = hip.alloc(n1 * 4)
    d_t = hip.alloc(n1 * 4)
    try:
        for n, seed in ((n1, 1), (n2, 2)):
            rays = _rays(sc, n, seed)
            ids_ref, t_ref = r.trace_closest(rays)
            hip.upload(d_rays, rays)
            # sentinels: the call must write exactly its n answers
            hip.upload(d_ids, np.full(n1, 0xA5A5A5A5, np.uint32))
            hip.upload(d_t, np.full(n1, -7.0, np.float32))
            r.trace_closest_noalloc(n, d_rays, d_ids, d_t)
            ids = hip.download(d_ids, np.zeros(n1, np.uint32))
            t = hip.download(d_t, np.zeros(n1, np.float32))
            np.testing.assert_array_equal(ids[:n], ids_ref)
            np.testing.assert_array_equal(t[:n], t_ref)
            assert np.all(ids[n:] == 0xA5A5A5A5) and np.all(t[n:] == -7.0)
            assert (ids_ref != aqua.MISS).mean() > 0.5
        # the same answers as the CPU restatement
        rays = _rays(sc, n2, 2)
        ids_c, t_c = oracle_scene_of(sc).trace_closest(rays)
        np.testing.assert_array_equal(ids[:n2], ids_c)
        np.testing.assert_array_equal(t[:n2], t_c)
        # n = 0 is a no-op; too many rays or a misaligned array is refused
        r.trace_closest_noalloc(0, d_rays, d_ids, d_t)
        with pytest.raises(aqua.AquaError) as e:
            r.trace_closest_noalloc((1 << 30) + 1, d_rays, d_ids, d_t)
        assert e.value.code == aqua.ERR_INVALID_ARG
        with pytest.raises(aqua.AquaError) as e:
            r.trace_closest_noalloc(4, d_rays + 8, d_ids, d_t)
        assert e.value.code == aqua.ERR_INVALID_ARG
        # a second renderer on the same device shares the scratch: calls in
        # turn stay correct
        r2 = aqua.Renderer(sc, c)
        rays = _rays(sc, 5000, 3)
        hip.upload(d_rays, rays)
        r2.trace_closest_noalloc(5000, d_rays, d_ids, d_t)
        ids_ref, t_ref = r.trace_closest(rays)
        np.testing.assert_array_equal(hip.download(d_ids, np.zeros(5000, np.uint32)), ids_ref)
        np.testing.assert_array_equal(hip.download(d_t, np.zeros(5000, np.float32)), t_ref)
    finally:
        for p in (d_rays, d_ids, d_t):
            hip.free(p)
