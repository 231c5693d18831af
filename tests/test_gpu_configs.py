"""GPU parity at the benchmark configurations themselves, not scaled-down
stand-ins:

* the 150-mesh / 1.09 M-triangle interior at 1920x1080 (deep BVH: spilled
  traversal stacks, 29-bit child references, 8-bit node planes over a
  1 M-triangle extent): every primary hit of the full frame, a band of rows of
  a 4-spp depth-8 film, and 10^5 random plus vertex / edge rays, bit-exact
  against the CPU restatement;
* the 400-mesh / 5.4 M-triangle interior at 3840x2160: primary hits on bands
  of rows bit-exact, and the FP8 and FP16 online caches finite, switched on,
  and better than the plain tracer at equal samples;
* the cache's training step at its real table size (2^19 entries per level)
  on box-room surface rows, gradients against the float64 restatement;
* online rendering at 1080p with 2^16 training paths per frame on the
  box-room fly-through;
* the FP64 film (the oracle accumulation on the GPU) and the FP64 error
  reductions against the CPU and numpy.

Bars: ids / distances / FP32 films bit-exact; FP64 films bit-exact when the
sample count is a power of two (both sides then scale a sum of the same
doubles exactly), else within one rounding; cache gradients and images as
stated in each test.
"""
import importlib.util
import math
import os

import numpy as np
import pytest

from conftest import ROOT, aqua, oracle, oracle_scene_of, small_camera

pytestmark = pytest.mark.gpu

spec = importlib.util.spec_from_file_location("nrc_ref", os.path.join(ROOT, "oracle", "nrc_ref.py"))
nref = importlib.util.module_from_spec(spec)
spec.loader.exec_module(nref)


def cfg(spp=4, max_depth=8, seed=0):
    c = aqua.integrator_default()
    c.spp, c.max_depth, c.seed = spp, max_depth, seed
    return c


def relmse(x, y):
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    return float(np.mean((x - y) ** 2 / (y * y + 1e-2)))


# ------------------------------------------- 1.09 M triangles, 1080p -------

@pytest.fixture(scope="module")
def interior1m(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("interior1m"))
    path = aqua.procgen(aqua.PROCGEN_INTERIOR, d, seed=0, n_meshes=150, target_tris=1000000, width=1920,
                        height=1080)
    sc = aqua.Scene.load(path)
    assert sc.info().n_triangles > 1000000
    print(f"interior1m: {sc.info().n_triangles} triangles", flush=True)
    return sc, oracle_scene_of(sc)


def test_interior1m_primary_hits_full_frame(gpu, interior1m):
    sc, orc = interior1m
    r = aqua.Renderer(sc, cfg())
    cam16 = aqua.camera_to_array(sc.camera())
    for s in (0, 5):
        ids_g, t_g = r.primary_hits(sample=s)
        ids_c, t_c = orc.primary_hits(cam16, sample=s, max_depth=8)
        assert ids_g.shape == (1080, 1920)
        np.testing.assert_array_equal(ids_g, ids_c)
        hit = ids_c != aqua.MISS
        np.testing.assert_array_equal(t_g[hit], t_c[hit])


def test_interior1m_film_band_and_spilled_stacks(gpu, interior1m):
    # The full 1080p frame at 4 spp, depth 8, on the GPU; the CPU restatement
    # renders two bands of rows of the same frame: the films agree bit for
    # bit there. Rendered again in 1-sample passes (two lanes, overlapped
    # streams): the same film.
    sc, orc = interior1m
    r = aqua.Renderer(sc, cfg(spp=4, max_depth=8))
    r.set_opts(timing=False)
    st = r.render(4, frame=2)
    img = r.film()
    cam16 = aqua.camera_to_array(sc.camera())
    for y0, y1 in ((532, 548), (0, 8)):
        band, _ = orc.render(cam16, 4, 8, frame=2, y0=y0, y1=y1)
        np.testing.assert_array_equal(img[y0:y1], band)
    assert img.mean() > 0.0 and st.nonfinite_samples == 0
    r.set_opts(samples_per_pass=1, timing=False)
    r.render(4, frame=2)
    np.testing.assert_array_equal(r.film(), img)
    # the tree is deep enough that traversal stacks overflow their 16 LDS
    # entries into the global spill area, and those rays still agree above
    nodes, _, depth = sc.bvh()
    assert 3 * (depth + 1) > 16, depth
    r.set_opts(count_traversal=True, timing=True)
    st = r.render(1, frame=2)
    assert st.spill_node_visits > 0, "no traversal stack spilled"
    print(f"BVH4 depth {depth}, {len(nodes)} nodes; {st.spill_node_visits} of "
          f"{st.nodes_visited + st.shadow_nodes_visited} node visits with a spilled stack")


def _edge_rays(verts, rng, n, origin_lo, origin_hi):
    tri = verts[rng.integers(0, len(verts), n)].reshape(n, 3, 3).astype(np.float64)
    k = rng.integers(0, 6, n)
    tgt = np.where((k < 3)[:, None], tri[np.arange(n), np.minimum(k, 2)],
                   0.5 * (tri[np.arange(n), k % 3] + tri[np.arange(n), (k + 1) % 3]))
    o = rng.uniform(origin_lo, origin_hi, (n, 3))
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 3] = 3.0e38
    rays[:, 4:7] = tgt - o
    return rays


def test_interior1m_random_and_edge_rays(gpu, interior1m):
    sc, orc = interior1m
    info = sc.info()
    lo, hi = np.array(info.bounds_lo), np.array(info.bounds_hi)
    rng = np.random.default_rng(21)
    n = 100000
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = rng.uniform(lo, hi, (n, 3))
    rays[:, 3] = 3.0e38
    d = rng.standard_normal((n, 3))
    d[:5000, 1:] = 0.0     # axis-aligned
    rays[:, 4:7] = d
    verts, _ = sc.triangles()
    rays = np.concatenate([rays, _edge_rays(verts, rng, 20000, lo, hi)])
    r = aqua.Renderer(sc, cfg())
    ids_g, t_g = r.trace_closest(rays)
    ids_c, t_c = orc.trace_closest(rays)
    np.testing.assert_array_equal(ids_g, ids_c)
    np.testing.assert_array_equal(t_g, t_c)
    assert (ids_g != aqua.MISS).mean() > 0.9
    short = rays.copy()
    short[:, 3] = np.where(ids_c != aqua.MISS, t_c * np.float32(0.999), 3.0e38)
    np.testing.assert_array_equal(r.trace_occluded(short), orc.trace_occluded(short))
    np.testing.assert_array_equal(r.trace_occluded(rays), (ids_c != aqua.MISS).astype(np.uint8))


# ------------------------------------------- 5.4 M triangles, 2160p --------

@pytest.fixture(scope="module")
def interior5m(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("interior5m"))
    path = aqua.procgen(aqua.PROCGEN_INTERIOR, d, seed=0, n_meshes=400, target_tris=5000000, width=3840,
                        height=2160)
    sc = aqua.Scene.load(path)
    assert sc.info().n_triangles > 5000000
    print(f"interior5m: {sc.info().n_triangles} triangles", flush=True)
    return sc


def test_interior5m_primary_hits_bands(gpu, interior5m):
    sc = interior5m
    r = aqua.Renderer(sc, cfg())
    ids_g, t_g = r.primary_hits(sample=0)
    del r
    orc = oracle_scene_of(sc)
    cam16 = aqua.camera_to_array(sc.camera())
    for y0, y1 in ((1064, 1096), (0, 16), (2144, 2160)):
        ids_c, t_c = orc.primary_hits(cam16, sample=0, max_depth=8, y0=y0, y1=y1)
        np.testing.assert_array_equal(ids_g[y0:y1], ids_c)
        hit = ids_c != aqua.MISS
        np.testing.assert_array_equal(t_g[y0:y1][hit], t_c[hit])


def test_interior5m_fp8_and_fp16_caches(gpu, interior5m):
    # 1 spp per frame with an online cache along the fly-through, FP8 (e4m3)
    # queries with BF16 training and FP16 throughout; then the last frame of
    # each against a 16-spp FP64 render on the GPU, beside the plain tracer's
    # 1-spp frame on the same samples.
    sc = interior5m
    frames = 24
    imgs = {}
    for name, prec in (("fp8", aqua.NRC_FP8), ("f16", aqua.NRC_F16)):
        c = cfg(spp=1, max_depth=8)
        c.type = aqua.INTEGRATOR_NRC
        c.learning_rate = 1e-2
        c.nrc_precision = prec
        c.nrc_train_paths = 1 << 16
        r = aqua.Renderer(sc, c)
        r.set_opts(timing=False)
        for f in range(frames):
            cam = aqua.camera_flythrough(sc, f, frames, width=3840, height=2160)
            st = r.render(1, frame=f, camera=cam)
            assert st.nrc_queries > 0 and math.isfinite(st.nrc_loss), (name, f)
        imgs[name] = r.film(cam)
        assert np.all(np.isfinite(imgs[name])), name
        del r
    pt = aqua.Renderer(sc, cfg(spp=1, max_depth=8))
    pt.render(1, frame=frames - 1, camera=cam)
    pt1 = pt.film(cam)
    pt.render_oracle(0, 16, frame=frames - 1 + 50000, camera=cam)
    ref = pt.oracle_film(cam)
    e = {k: relmse(v, ref) for k, v in imgs.items()}
    e_pt = relmse(pt1, ref)
    gap = relmse(imgs["fp8"], imgs["f16"])
    print(f"4K 5.4M tris, frame {frames - 1}: relMSE vs 16-spp FP64: PT 1 spp {e_pt:.4f}, "
          f"cache FP8 {e['fp8']:.4f}, FP16 {e['f16']:.4f}; FP8 vs FP16 {gap:.4f}")
    assert e["fp8"] < e_pt and e["f16"] < e_pt
    assert gap < 0.25 * e_pt
    assert abs(e["fp8"] - e["f16"]) < 0.5 * e["f16"]


# ------------------------------------- cache training at 2^19 entries ------

def _surface_rows(n, seed):
    """Cache records at box-room surface points seen from its camera (first
    hits of jittered camera rays), the directions towards the camera, the
    hit triangles' normals and albedos; positions normalised to the scene
    bounds as the renderer does."""
    import tempfile
    rng = np.random.default_rng(seed)
    with tempfile.TemporaryDirectory() as d:
        sc = aqua.Scene.load(aqua.procgen(aqua.PROCGEN_BOX_ROOM, d, width=256, height=256))
    orc = oracle_scene_of(sc)
    cam = sc.camera()
    cam16 = aqua.camera_to_array(cam)
    px = rng.integers(0, 256, (n, 2))
    u = rng.random((n, 2)).astype(np.float32)
    o = np.zeros((n, 3), np.float32)
    dirs = np.zeros((n, 3), np.float32)
    for i in range(n):
        ray = oracle.camera_ray(cam16, int(px[i, 0]), int(px[i, 1]), u[i, 0], u[i, 1])
        o[i], dirs[i] = ray[:3], ray[3:]
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3], rays[:, 3], rays[:, 4:7] = o, 3.0e38, dirs
    ids, t = orc.trace_closest(rays)
    keep = ids != aqua.MISS
    verts, mids = sc.triangles()
    mats = sc.materials()
    info = sc.info()
    lo, hi = np.array(info.bounds_lo), np.array(info.bounds_hi)
    p = o[keep] + t[keep, None] * dirs[keep]
    tri = verts[ids[keep]].reshape(-1, 3, 3)
    nrm = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    nrm = np.where((np.einsum("ij,ij->i", nrm, dirs[keep]) > 0)[:, None], -nrm, nrm)
    rec = np.zeros((keep.sum(), 16), np.float32)
    rec[:, 0:3] = np.clip((p - lo) / (hi - lo), 0, 1)
    rec[:, 3] = mats[mids[ids[keep]], 7]
    rec[:, 4:7] = -dirs[keep]
    rec[:, 8:11] = nrm
    rec[:, 12:15] = mats[mids[ids[keep]], 0:3]
    return rec


@pytest.mark.parametrize("prec", [aqua.NRC_F16, aqua.NRC_BF16])
def test_training_gradients_at_full_table_size(gpu, prec):
    # 2^19 entries per level (16.8 M parameters), 2^14 box-room surface rows.
    # Gradient tolerances against the restatement that rounds to 16 bits
    # where the kernels do (table FP16, layer inputs and backward values in
    # the training precision): levels 0-3 accumulate FP32, levels 4-15 BF16
    # pairs (8-bit mantissa), whose sums drift with the number of rows an
    # entry collects -- bounded here per level.
    rec = _surface_rows(1 << 14, seed=3)
    n = len(rec)
    rng = np.random.default_rng(4)
    tgt = rng.uniform(0.0, 2.0, (n, 3)).astype(np.float32)
    c = aqua.nrc_default_cfg()
    c.log2_table = 19
    c.train_precision = prec
    c.max_batch = 1 << 15
    nrc = aqua.Nrc(c)
    nt, nm = nrc.counts()
    table = rng.uniform(-1e-1, 1e-1, nt).astype(np.float32)
    mlp = rng.uniform(-0.2, 0.2, nm).astype(np.float32)
    nrc.set_params(table, mlp)
    loss_g, gt, gm = nrc.gradients(rec, tgt)
    q = "f16" if prec == aqua.NRC_F16 else "bf16"
    loss_r, gt_r, gm_r = nref.loss_and_grads(rec, tgt, table.reshape(16, 1 << 19, 2), mlp, quant=q)
    assert abs(loss_g - loss_r) <= 1e-3 * abs(loss_r)
    gt = gt.reshape(16, -1)
    gt_r = gt_r.reshape(16, -1)
    errs = []
    for lv in range(16):
        nz = np.abs(gt_r[lv]) > 0
        e = np.linalg.norm(gt[lv] - gt_r[lv]) / max(np.linalg.norm(gt_r[lv]), 1e-30)
        errs.append(e)
        assert np.all(gt[lv][~nz] == 0.0), lv      # untouched entries stay zero
    print("table-gradient rel. L2 per level:", " ".join(f"{e:.1e}" for e in errs))
    tol_mlp = 3e-3 if prec == aqua.NRC_F16 else 1e-2
    e_mlp = np.linalg.norm(gm - gm_r) / np.linalg.norm(gm_r)
    assert e_mlp < tol_mlp, e_mlp
    tight = 3e-3 if prec == aqua.NRC_F16 else 1e-2
    assert max(errs[:8]) < tight, errs[:8]         # FP32 accumulation
    assert max(errs[8:]) < 2e-2, errs[8:]          # BF16-pair accumulation


def test_bf16_pair_gradients_with_many_rows_per_entry(gpu):
    # 2^18 rows packed into a cube of side 0.03: on the coarse levels every
    # entry collects thousands of rows. Levels 0-7 accumulate FP32 pairs and
    # must stay tight; the BF16 pairs of levels 8-15 (an 8-bit mantissa) drift
    # with the rows per entry -- about a hundred at level 8 here, against the
    # few a batch spread over a scene gives them. The bound is on the relative
    # L2 of each level's gradient against the float64 restatement.
    rng = np.random.default_rng(6)
    n = 1 << 18
    rec = np.zeros((n, 16), np.float32)
    rec[:, 0:3] = 0.4 + 0.03 * rng.random((n, 3))
    rec[:, 3] = rng.random(n)
    d = rng.standard_normal((n, 3))
    rec[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rec[:, 8:11] = [0.0, 1.0, 0.0]
    rec[:, 12:15] = rng.random((n, 3))
    tgt = rng.uniform(0.0, 2.0, (n, 3)).astype(np.float32)
    c = aqua.nrc_default_cfg()
    c.log2_table = 19
    c.train_precision = aqua.NRC_BF16
    c.max_batch = n
    nrc = aqua.Nrc(c)
    nt, nm = nrc.counts()
    table = rng.uniform(-1e-1, 1e-1, nt).astype(np.float32)
    mlp = rng.uniform(-0.2, 0.2, nm).astype(np.float32)
    nrc.set_params(table, mlp)
    _, gt, _ = nrc.gradients(rec, tgt)
    _, gt_r, _ = nref.loss_and_grads(rec, tgt, table.reshape(16, 1 << 19, 2), mlp, quant="bf16")
    gt = gt.reshape(16, -1)
    gt_r = gt_r.reshape(16, -1)
    errs = [np.linalg.norm(gt[lv] - gt_r[lv]) / max(np.linalg.norm(gt_r[lv]), 1e-30) for lv in range(16)]
    print("clustered rows, table-gradient rel. L2 per level:", " ".join(f"{e:.1e}" for e in errs))
    assert max(errs[:8]) < 1e-2, errs[:8]
    assert max(errs[8:]) < 8e-2, errs[8:]


# ------------------------------------- online cache at 1080p, 2^16 paths ---

def test_online_cache_1080p_flythrough(gpu, box_room):
    frames, n_frames = 12, 240
    c = cfg(spp=1, max_depth=8)
    c.type = aqua.INTEGRATOR_NRC
    c.learning_rate = 1e-2
    c.nrc_train_paths = 1 << 16
    r = aqua.Renderer(box_room, c)
    r.set_opts(timing=False)
    for f in range(frames):
        cam = aqua.camera_flythrough(box_room, f, n_frames)
        st = r.render(1, frame=f, camera=cam)
        assert st.nrc_queries > 0 and st.nrc_train_records > 0 and math.isfinite(st.nrc_loss), f
    assert cam.width == 1920 and cam.height == 1080
    cam = aqua.camera_flythrough(box_room, frames, n_frames)
    pt = aqua.Renderer(box_room, cfg(spp=4, max_depth=8))
    pt.render_oracle(0, 256, frame=frames, camera=cam)
    ref = pt.oracle_film(cam)
    st = r.render(4, frame=frames + 200000, camera=cam)
    img = r.film(cam)
    assert np.all(np.isfinite(img)) and st.nrc_queries > 0
    pt.render(4, frame=frames + 100000, camera=cam)
    e_c, e_p = relmse(img, ref), relmse(pt.film(cam), ref)
    print(f"1080p box room, frame {frames}, 4 spp: relMSE cache {e_c:.4f} vs plain {e_p:.4f}")
    assert e_c < e_p


def test_cache_first_frame_overlap_at_1080p(gpu, box_room):
    # Inference long enough to overlap the training step's forward/backward:
    # 1080p, 1-sample passes, 2^16 training paths. The first frame depends
    # only on the initial weights, so the overlapped run must equal the
    # serialized one bit for bit; a second frame stays finite.
    c = cfg(spp=4, max_depth=8)
    c.type = aqua.INTEGRATOR_NRC
    c.nrc_train_paths = 1 << 16
    cam = aqua.camera_flythrough(box_room, 30, 240)
    ref = aqua.Renderer(box_room, c)
    ref.set_opts(samples_per_pass=1, timing=True)
    ref.render(4, frame=1, camera=cam)
    want = ref.film(cam)
    del ref
    r = aqua.Renderer(box_room, c)
    r.set_opts(samples_per_pass=1, timing=False)
    st = r.render(4, frame=1, camera=cam)
    assert st.n_passes == 4 and st.nrc_queries > 1000000
    np.testing.assert_array_equal(r.film(cam), want)
    r.render(4, frame=2, camera=cam)
    assert np.all(np.isfinite(r.film(cam)))


# ------------------------------------------ FP64 film, error reductions ----

@pytest.mark.parametrize("scene_kind", ["box_room", "interior"])
def test_fp64_film_matches_cpu_render64(gpu, box_room, tmp_path, scene_kind):
    if scene_kind == "box_room":
        sc = box_room
        cam = small_camera(sc.camera(), 64, 48)
        depth = 5
    else:
        path = aqua.procgen(aqua.PROCGEN_INTERIOR, str(tmp_path), seed=5, n_meshes=16, target_tris=20000,
                            width=48, height=32)
        sc = aqua.Scene.load(path)
        cam = sc.camera()
        depth = 8
    orc = oracle_scene_of(sc)
    r = aqua.Renderer(sc, cfg(spp=4, max_depth=depth))
    cam16 = aqua.camera_to_array(cam)
    # 16 samples (a power of two): both sides scale the same double sums
    # exactly, resumed or not
    r.render_oracle(0, 16, frame=3, camera=cam)
    want = orc.render64(cam16, 16, depth, frame=3)
    np.testing.assert_array_equal(r.oracle_film(cam), want)
    r.render_oracle(0, 7, frame=3, camera=cam)
    r.render_oracle(7, 9, frame=3, camera=cam)
    np.testing.assert_array_equal(r.oracle_film(cam), want)
    # 12 samples: x * (1/12) against x / 12, within one rounding
    r.render_oracle(0, 12, frame=3, camera=cam)
    np.testing.assert_allclose(r.oracle_film(cam), orc.render64(cam16, 12, depth, frame=3), rtol=2.3e-16,
                               atol=0)


def test_film_error_reductions_match_numpy(gpu, box_room):
    cam = small_camera(box_room.camera(), 97, 61)
    r = aqua.Renderer(box_room, cfg(spp=4, max_depth=5))
    r.render_oracle(0, 64, frame=1, camera=cam)
    for spp in (1, 3, 8):
        r.render(spp, frame=1000 + spp, camera=cam)
        x = r.film(cam).astype(np.float64)
        y = r.oracle_film(cam)
        want_mse = np.mean((x - y) ** 2 / (y * y + 1e-2))
        want_rmse = math.sqrt(np.sum((x - y) ** 2) / np.sum(y * y))
        got_mse, got_rmse = r.film_error()
        assert abs(got_mse - want_mse) <= 1e-12 * want_mse, (spp, got_mse, want_mse)
        assert abs(got_rmse - want_rmse) <= 1e-12 * want_rmse, (spp, got_rmse, want_rmse)
    # identical films: zero error
    r.render(64, frame=1, camera=cam)
    got_mse, got_rmse = r.film_error()
    assert got_mse < 1e-12 and got_rmse < 1e-6


# ------------------------------------------------------ point light --------

@pytest.mark.parametrize("look_at,light", [((0.0, 0.0), (0.0, 2.0, 0.0)), ((0.0, 0.0), (1.5, 2.0, 0.5))])
def test_point_light_closed_form_on_gpu(gpu, look_at, light):
    # rho / pi * I * cos / r^2 over a Lambertian floor, as in the oracle pin
    rho, I = 0.6, 5.0
    m = aqua.Material()
    m.base_color[:] = [rho] * 3
    m.roughness = 0.5
    floor = np.float32([[-50, 0, -50, -50, 0, 50, 50, 0, 50], [-50, 0, -50, 50, 0, 50, 50, 0, -50]])
    cam = aqua.Camera()
    x, z = look_at
    cam.cam_to_world[:] = [1, 0, 0, x, 0, 0, 1, 3, 0, -1, 0, z]
    cam.tan_half_fov = float(np.float32(1e-5))
    cam.aspect, cam.width, cam.height = 1.0, 2, 2
    pl = aqua.PointLight()
    pl.pos[:] = light
    pl.intensity[:] = [I, I, I]
    sc = aqua.Scene.from_arrays(floor, np.zeros(2, np.uint32), [m], cam, [pl])
    r = aqua.Renderer(sc, cfg(spp=4, max_depth=1))
    r.render(4)
    dv = np.array(light) - np.array([x, 0.0, z])
    r2 = float(dv @ dv)
    expect = rho / math.pi * I * (dv[1] / math.sqrt(r2)) / r2
    np.testing.assert_allclose(r.film().astype(np.float64), expect, rtol=5e-5)
