"""GPU kernels vs the CPU restatement (oracle/) on identical inputs and seeds.

Bar: primary-hit triangle ids bit-exact; FP32 film bit-exact (relRMSE
tolerance 1e-4 written below is the contract, the kernels meet it with 0).
"""
import os
from fractions import Fraction

import numpy as np
import pytest

from conftest import GOLDEN, aqua, oracle, oracle_scene_of, small_camera

pytestmark = pytest.mark.gpu

RELRMSE_TOL = 1e-4


def relrmse(x, y):
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    den = np.sum(y * y)
    return float(np.sqrt(np.sum((x - y) ** 2) / den)) if den > 0 else float(np.sqrt(np.sum((x - y) ** 2)))


def round_f32(r: Fraction) -> np.float32:
    """Nearest-even FP32 value of an exact rational."""
    f = np.float32(float(r))
    best = f
    for cand in (np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))):
        if not np.isfinite(cand):
            continue
        d, bd = abs(Fraction(float(cand)) - r), abs(Fraction(float(best)) - r)
        if d < bd or (d == bd and (int(cand.view(np.uint32)) & 1) == 0):
            best = cand
    return best


def cfg(spp=4, max_depth=5, seed=0, sampler=0):
    c = aqua.integrator_default()
    c.spp = spp
    c.max_depth = max_depth
    c.seed = seed
    c.sampler = sampler
    return c


def test_device_fp_arithmetic_is_ieee(gpu):
    rng = np.random.default_rng(1)
    n = 4096
    a = (rng.random(n) * 10 ** rng.uniform(-6, 6, n)).astype(np.float32)
    b = (rng.random(n) * 10 ** rng.uniform(-6, 6, n) + 1e-30).astype(np.float32)
    c = (rng.standard_normal(n) * 10 ** rng.uniform(-6, 6, n)).astype(np.float32)
    u = rng.random(n).astype(np.float32)
    u[:4] = [0.0, 0.25, 0.5, np.float32(0.99999994)]
    out = aqua.fp_probe(a, b, c, u)
    np.testing.assert_array_equal(out[:, 0], a / b)          # correctly rounded divide
    np.testing.assert_array_equal(out[:, 1], np.sqrt(a))     # correctly rounded sqrt
    for i in range(0, n, 16):                                # single-rounding fma
        exact = Fraction(float(a[i])) * Fraction(float(b[i])) + Fraction(float(c[i]))
        assert out[i, 2] == round_f32(exact), i
    s, co = oracle.sincos_2pi(u)
    np.testing.assert_array_equal(out[:, 3], s)
    np.testing.assert_array_equal(out[:, 4], co)


def _hits_match(scene, orc_scene, cam=None, samples=(0, 1)):
    r = aqua.Renderer(scene, cfg())
    c = cam or scene.camera()
    cam16 = aqua.camera_to_array(c)
    for s in samples:
        ids_g, t_g = r.primary_hits(sample=s, camera=cam)
        ids_c, t_c = orc_scene.primary_hits(cam16, sample=s)
        frac = float(np.mean(ids_g == ids_c))
        assert frac == 1.0, f"sample {s}: {frac:.6f} of primary ids match"
        hit = ids_c != aqua.MISS
        np.testing.assert_array_equal(t_g[hit], t_c[hit])


def test_primary_hits_box_room(gpu, box_room, box_room_oracle):
    _hits_match(box_room, box_room_oracle)


def test_primary_hits_shipped_cbox(gpu):
    sc = aqua.Scene.load(os.path.join(GOLDEN, "cbox", "cbox.json"))
    _hits_match(sc, oracle_scene_of(sc))


def test_trace_closest_random_rays_vs_brute_force(gpu, box_room, box_room_oracle):
    rng = np.random.default_rng(7)
    n = 20000
    o = rng.uniform([-0.99, 0.01, -0.99], [0.99, 1.99, 0.99], (n, 3))
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 3] = 3.0e38
    rays[:, 4:7] = d
    r = aqua.Renderer(box_room, cfg())
    ids_g, t_g = r.trace_closest(rays)
    ids_b, t_b = box_room_oracle.trace_closest(rays, brute=True)
    ids_c, t_c = box_room_oracle.trace_closest(rays)
    np.testing.assert_array_equal(ids_c, ids_b)
    np.testing.assert_array_equal(ids_g, ids_b)
    np.testing.assert_array_equal(t_g, t_b)
    # any-hit with tmax just short of the closest hit: never occluded
    short = rays.copy()
    short[:, 3] = np.where(ids_b != aqua.MISS, t_b * np.float32(0.999), 3.0e38)
    occ_g = r.trace_occluded(short)
    occ_c = box_room_oracle.trace_occluded(short, brute=True)
    np.testing.assert_array_equal(occ_g, occ_c)


def test_trace_occluded_sizes_and_mixed_segments(gpu, box_room, box_room_oracle):
    # the persistent any-hit kernel on ray counts around its 64-ray chunks,
    # half of the segments ending just past their closest hit (occluded), half
    # just short of it (clear)
    rng = np.random.default_rng(8)
    r = aqua.Renderer(box_room, cfg())
    for n in (1, 63, 65, 4097):
        o = rng.uniform([-0.99, 0.01, -0.99], [0.99, 1.99, 0.99], (n, 3))
        d = rng.standard_normal((n, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        rays = np.zeros((n, 8), np.float32)
        rays[:, 0:3] = o
        rays[:, 3] = 3.0e38
        rays[:, 4:7] = d
        ids_b, t_b = box_room_oracle.trace_closest(rays, brute=True)
        past = rng.random(n) < 0.5
        seg = rays.copy()
        seg[:, 3] = np.where(ids_b != aqua.MISS, t_b * np.where(past, np.float32(1.001), np.float32(0.999)),
                             np.float32(3.0e38))
        occ_g = r.trace_occluded(seg)
        np.testing.assert_array_equal(occ_g, box_room_oracle.trace_occluded(seg, brute=True))
        np.testing.assert_array_equal(occ_g, ((ids_b != aqua.MISS) & past).astype(np.uint8))


def _render_match(scene, orc_scene, cam, spp, max_depth, seed=0, frame=0, sampler=0):
    r = aqua.Renderer(scene, cfg(spp, max_depth, seed, sampler))
    st = r.render(spp, frame=frame, camera=cam)
    img_g = r.film(cam)
    img_c, ost = orc_scene.render(aqua.camera_to_array(cam), spp, max_depth, seed=seed, frame=frame,
                                  sampler=sampler)
    err = relrmse(img_g, img_c)
    assert err <= RELRMSE_TOL, f"relRMSE {err:.3e}"
    assert st.camera_rays == ost.camera_rays
    assert st.extension_rays == ost.extension_rays
    assert st.shadow_rays == ost.shadow_rays
    return img_g, img_c, st


def test_render_box_room_parity(gpu, box_room, box_room_oracle):
    cam = small_camera(box_room.camera(), 96, 96)
    g, c, st = _render_match(box_room, box_room_oracle, cam, spp=16, max_depth=5)
    np.testing.assert_array_equal(g, c)
    assert g.mean() > 0.05


def test_stratified_sampler_film_and_hits_match_oracle(gpu, box_room, box_room_oracle):
    # stratified camera / first-hit dimensions: the same integer sampler on
    # both sides, so films and primary hits stay bit-exact; samples 8.. of
    # frame 3 (a later pass) use the same patterns
    cam = small_camera(box_room.camera(), 80, 64)
    g, c, _ = _render_match(box_room, box_room_oracle, cam, spp=16, max_depth=5, seed=7, frame=3, sampler=1)
    np.testing.assert_array_equal(g, c)
    g0, _, _ = _render_match(box_room, box_room_oracle, cam, spp=16, max_depth=5, seed=7, frame=3)
    assert not np.array_equal(g, g0)
    r = aqua.Renderer(box_room, cfg(sampler=1, seed=7))
    r.set_opts(samples_per_pass=4)
    r.render(16, frame=3, camera=cam)
    np.testing.assert_array_equal(r.film(cam), c)
    cam16 = aqua.camera_to_array(cam)
    for s in (0, 5, 1023):
        ids_g, t_g = r.primary_hits(sample=s, frame=3, camera=cam)
        ids_c, t_c = box_room_oracle.primary_hits(cam16, sample=s, frame=3, seed=7, sampler=1)
        np.testing.assert_array_equal(ids_g, ids_c)
        np.testing.assert_array_equal(t_g, t_c)


def test_render_shipped_cbox_point_light_parity(gpu):
    sc = aqua.Scene.load(os.path.join(GOLDEN, "cbox", "cbox.json"))
    cam = small_camera(sc.camera(), 64, 64)
    g, c, _ = _render_match(sc, oracle_scene_of(sc), cam, spp=8, max_depth=5)
    np.testing.assert_array_equal(g, c)


def test_render_interior_ggx_parity(gpu, tmp_path):
    path = aqua.procgen(aqua.PROCGEN_INTERIOR, str(tmp_path), seed=3, n_meshes=24, target_tris=60000,
                        width=80, height=48)
    sc = aqua.Scene.load(path)
    g, c, _ = _render_match(sc, oracle_scene_of(sc), sc.camera(), spp=4, max_depth=8)
    np.testing.assert_array_equal(g, c)


def test_overlapped_kernels_and_passes_do_not_change_the_film(gpu, box_room, box_room_oracle):
    # without per-kernel timing the shadow rays of bounce b run on a second
    # stream beside the closest-hit rays of bounce b + 1, and consecutive
    # passes run concurrently on two sets of queues; the film must be the one
    # of the serialized order, bit for bit (and the oracle's), for the FP32
    # and the FP64 film, and the cache (single lane) must still work
    cam = small_camera(box_room.camera(), 61, 35)
    r = aqua.Renderer(box_room, cfg(spp=6))
    r.set_opts(timing=True)
    r.render(6, frame=3, camera=cam)
    serial = r.film(cam)
    r.set_opts(timing=False)
    for _ in range(3):
        r.render(6, frame=3, camera=cam)
        np.testing.assert_array_equal(r.film(cam), serial)
    # several passes: consecutive passes run concurrently on two sets of
    # queues and streams, their films added in pass order
    for spp_pass in (1, 2, 4):
        r.set_opts(samples_per_pass=spp_pass, timing=False)
        st = r.render(6, frame=3, camera=cam)
        assert st.n_passes == (6 + spp_pass - 1) // spp_pass
        np.testing.assert_array_equal(r.film(cam), serial)
    img_c, _ = box_room_oracle.render(aqua.camera_to_array(cam), 6, 5, frame=3)
    np.testing.assert_array_equal(serial, img_c)
    # large enough that both traversal grids are full and really run together
    big = small_camera(box_room.camera(), 512, 384)
    r2 = aqua.Renderer(box_room, cfg(spp=4, max_depth=8))
    r2.render(4, frame=7, camera=big)
    serial = r2.film(big)
    st_serial = r2.render(4, frame=7, camera=big)
    r2.set_opts(samples_per_pass=1, timing=False)
    st = r2.render(4, frame=7, camera=big)
    np.testing.assert_array_equal(r2.film(big), serial)
    assert (st.camera_rays, st.extension_rays, st.shadow_rays) == \
        (st_serial.camera_rays, st_serial.extension_rays, st_serial.shadow_rays)
    # the default pass size puts a 64-spp frame of this size in one pass of
    # 12.6 M paths; 16-spp passes (3.1 M paths) alternate lanes: same film
    r2.set_opts(samples_per_pass=0, timing=False)
    st = r2.render(64, frame=9, camera=big)
    assert st.n_passes == 1
    one_pass = r2.film(big)
    r2.set_opts(samples_per_pass=16, timing=False)
    st = r2.render(64, frame=9, camera=big)
    assert st.n_passes == 4
    np.testing.assert_array_equal(r2.film(big), one_pass)
    # FP64 film, resumed across calls, 2-sample passes
    r3 = aqua.Renderer(box_room, cfg(spp=4))
    r3.render_oracle(0, 5, frame=2, camera=cam)
    r3.render_oracle(5, 4, frame=2, camera=cam)
    serial64 = r3.oracle_film(cam)
    r3.set_opts(samples_per_pass=2, timing=False)
    r3.render_oracle(0, 5, frame=2, camera=cam)
    r3.render_oracle(5, 4, frame=2, camera=cam)
    np.testing.assert_array_equal(r3.oracle_film(cam), serial64)
    # the cache: without timing, the queries' sort, inference and apply run on
    # the second stream beside the training paths' last bounces, and passes
    # alternate lanes. A fresh renderer's first frame depends only on the
    # initial weights (training follows the film), so it must match the
    # serialized frame bit for bit, in one pass and in several.
    c = cfg(spp=4, max_depth=8)
    c.type = aqua.INTEGRATOR_NRC
    c.nrc_train_paths = 256
    for view in (cam, big):
        for spp_pass in (0, 1):
            ref = aqua.Renderer(box_room, c)
            ref.set_opts(samples_per_pass=spp_pass, timing=True)
            st_ref = ref.render(4, frame=1, camera=view)
            n = aqua.Renderer(box_room, c)
            n.set_opts(samples_per_pass=spp_pass, timing=False)
            st = n.render(4, frame=1, camera=view)
            assert st.nrc_queries > 0 and st.nrc_queries == st_ref.nrc_queries
            img = n.film(view)
            assert np.all(np.isfinite(img))
            np.testing.assert_array_equal(img, ref.film(view))


def test_pass_split_does_not_change_the_film(gpu, box_room, box_room_oracle):
    # sample s of pixel p draws from a fixed position of stream p, so the film
    # is the same however the samples are split into passes (here 3 + 3 + 1
    # against 7 in one pass), at a resolution that is not a multiple of 64
    cam = small_camera(box_room.camera(), 37, 23)
    r = aqua.Renderer(box_room, cfg(spp=7))
    r.set_opts(samples_per_pass=3)
    st = r.render(7, frame=5, camera=cam)
    assert st.n_passes == 3
    split = r.film(cam)
    r.set_opts(samples_per_pass=7)
    st = r.render(7, frame=5, camera=cam)
    assert st.n_passes == 1
    np.testing.assert_array_equal(r.film(cam), split)
    img_c, _ = box_room_oracle.render(aqua.camera_to_array(cam), 7, 5, frame=5)
    np.testing.assert_array_equal(split, img_c)
    # same frame again: identical; another frame: a different noise pattern
    r.render(7, frame=5, camera=cam)
    np.testing.assert_array_equal(r.film(cam), split)
    r.render(7, frame=6, camera=cam)
    assert not np.array_equal(r.film(cam), split)


def test_single_triangle_and_empty_scenes(gpu):
    m = aqua.Material()
    m.base_color[:] = [0.8, 0.8, 0.8]
    m.roughness = 0.5
    cam = aqua.Camera()
    cam.cam_to_world[:] = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 3]
    cam.tan_half_fov = np.float32(0.5)
    cam.width, cam.height, cam.aspect = 33, 17, np.float32(33 / 17)
    pl = aqua.PointLight()
    pl.pos[:] = [0.0, 0.0, 2.0]
    pl.intensity[:] = [4.0, 4.0, 4.0]
    tri = np.array([[-1, -1, 0, 1, -1, 0, 0, 1, 0]], np.float32)
    sc = aqua.Scene.from_arrays(tri, np.zeros(1, np.uint32), [m], cam, [pl])
    orc = oracle_scene_of(sc)
    _hits_match(sc, orc)
    g, c, _ = _render_match(sc, orc, cam, spp=4, max_depth=3)
    np.testing.assert_array_equal(g, c)
    assert g.max() > 0.0
    # no triangles at all: every primary ray misses and the film is black
    empty = aqua.Scene.from_arrays(np.zeros((0, 9), np.float32), np.zeros(0, np.uint32), [m], cam, [pl])
    r = aqua.Renderer(empty, cfg())
    ids, _ = r.primary_hits()
    assert np.all(ids == aqua.MISS)
    st = r.render(2)
    assert st.extension_rays == 0 and st.shadow_rays == 0
    assert not np.any(r.film())


def test_trace_degenerate_and_edge_rays(gpu, box_room, box_room_oracle):
    v, _ = box_room.triangles()
    v = v.reshape(-1, 3, 3)
    rng = np.random.default_rng(11)
    rays = []
    inside = np.array([0.3, 1.5, 0.6], np.float32)  # above the short box, clear of the tall one
    for t in v[::3]:
        for p in (t[0], t[1], t[2], (t[0] + t[1]) / 2, (t[1] + t[2]) / 2, t.mean(0)):
            d = p - inside
            rays.append([*inside, 3.0e38, *(d / np.linalg.norm(d)), 0.0])   # at vertices / edges
    for k in range(3):                                                  # axis-aligned both ways
        for s in (-1.0, 1.0):
            d = np.zeros(3, np.float32)
            d[k] = s
            rays.append([*inside, 3.0e38, *d, 0.0])
    o = rng.uniform(-0.9, 0.9, (64, 3)).astype(np.float32)
    o[:, 1] += 1.0
    d = rng.standard_normal((64, 3)).astype(np.float32)
    for i in range(64):
        rays.append([*o[i], 0.0, *d[i], 0.0])                           # tmax 0: always a miss
        rays.append([*o[i], 1e-6, *(d[i] * 1e-3), 0.0])                 # short, unnormalised
    rays = np.array(rays, np.float32)
    r = aqua.Renderer(box_room, cfg())
    ids_g, t_g = r.trace_closest(rays)
    ids_b, t_b = box_room_oracle.trace_closest(rays, brute=True)
    np.testing.assert_array_equal(ids_g, ids_b)
    np.testing.assert_array_equal(t_g, t_b)
    assert np.all(ids_g[len(rays) - 128::2] == aqua.MISS)
    np.testing.assert_array_equal(r.trace_occluded(rays), box_room_oracle.trace_occluded(rays, brute=True))


def test_closed_cube_is_watertight(gpu):
    # Rays from inside a closed cube aimed exactly at its vertices and edge
    # midpoints (where two or three triangles meet) must all hit: a crack in
    # the ray/triangle test would let some through.
    c = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], np.float32)
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    tris = []
    for a, b, cc, d in quads:
        tris += [c[[a, b, cc]].reshape(-1), c[[a, cc, d]].reshape(-1)]
    m = aqua.Material()
    m.base_color[:] = [0.5, 0.5, 0.5]
    cam = aqua.Camera()
    cam.cam_to_world[:] = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]
    cam.tan_half_fov = np.float32(0.5)
    cam.width, cam.height, cam.aspect = 8, 8, np.float32(1)
    sc = aqua.Scene.from_arrays(np.array(tris), np.zeros(len(tris), np.uint32), [m], cam)
    targets = list(c)
    for i in range(8):
        for j in range(i + 1, 8):
            if np.sum(c[i] != c[j]) == 1:
                targets.append((c[i] + c[j]) / 2)
    rng = np.random.default_rng(5)
    rays = []
    for o in rng.uniform(-0.9, 0.9, (50, 3)).astype(np.float32):
        for p in targets:
            d = np.asarray(p, np.float32) - o
            rays.append([*o, 3.0e38, *d, 0.0])
    rays = np.array(rays, np.float32)
    r = aqua.Renderer(sc, cfg())
    ids, t = r.trace_closest(rays)
    assert np.all(ids != aqua.MISS), int(np.sum(ids == aqua.MISS))
    orc = oracle_scene_of(sc)
    ids_b, t_b = orc.trace_closest(rays, brute=True)
    np.testing.assert_array_equal(ids, ids_b)
    np.testing.assert_array_equal(t, t_b)
