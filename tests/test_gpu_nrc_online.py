"""Online neural radiance cache inside the wavefront renderer."""
import math

import numpy as np
import pytest

from conftest import aqua, small_camera

pytestmark = pytest.mark.gpu


def nrc_cfg(max_depth=5, train_paths=1024, precision=aqua.NRC_F16, query_vertex=2, lr=1e-2):
    c = aqua.integrator_default()
    c.type = aqua.INTEGRATOR_NRC
    c.max_depth = max_depth
    c.nrc_train_paths = train_paths
    c.nrc_precision = precision
    c.nrc_query_vertex = query_vertex
    c.learning_rate = lr
    return c


def furnace_scene(rho, w=48, h=48):
    """Closed cube [-1,1]^3, every wall a Lambertian emitter (Le = 1, albedo
    rho) facing inward; the camera sits inside."""
    tris = []
    for axis in range(3):
        for s in (-1.0, 1.0):
            def P(a, b):
                p = [0.0, 0.0, 0.0]
                p[axis] = s
                p[(axis + 1) % 3] = a
                p[(axis + 2) % 3] = b
                return p
            q = [P(-1, -1), P(1, -1), P(1, 1), P(-1, 1)]
            for t in ((q[0], q[1], q[2]), (q[0], q[2], q[3])):
                v = np.array(t, np.float32)
                nrm = np.cross(v[1] - v[0], v[2] - v[0])
                if nrm @ (-v.mean(0)) < 0:
                    v = v[[0, 2, 1]]
                tris.append(v.reshape(-1))
    m = aqua.Material()
    m.base_color[:] = [rho] * 3
    m.emission[:] = [1.0] * 3
    m.roughness = 1.0
    cam = aqua.Camera()
    cam.cam_to_world[:] = [1, 0, 0, 0.1, 0, 1, 0, -0.2, 0, 0, 1, 0.3]
    cam.tan_half_fov = float(np.float32(math.tan(math.radians(40))))
    cam.width, cam.height, cam.aspect = w, h, 1.0
    return aqua.Scene.from_arrays(np.array(tris), np.zeros(len(tris), np.uint32), [m], cam)


def test_cache_disabled_is_the_plain_tracer(gpu, box_room):
    cam = small_camera(box_room.camera(), 64, 64)
    pt = aqua.integrator_default()
    pt.max_depth = 5
    a = aqua.Renderer(box_room, pt)
    a.render(4, camera=cam)
    r = aqua.Renderer(box_room, nrc_cfg())
    r.set_nrc_enabled(False)
    st = r.render(4, camera=cam)
    assert st.nrc_queries == 0
    np.testing.assert_array_equal(r.film(cam), a.film(cam))


def test_furnace_through_the_cache(gpu):
    # Query paths: emission at vertex 1, NEE / BSDF to vertex 2, emission
    # there, then the cache for the rest; the cache must learn the reflected
    # radiance sum_{k=1}^{D-1} rho^k, so the image converges to sum_{k<=D} rho^k.
    rho, D = 0.5, 5
    sc = furnace_scene(rho)
    r = aqua.Renderer(sc, nrc_cfg(max_depth=D, train_paths=512, lr=1e-2))
    for f in range(60):
        st = r.render(1, frame=f)
    assert st.nrc_queries > 0 and st.nrc_train_records > 0
    imgs = []
    for f in range(60, 70):
        r.render(1, frame=f)
        imgs.append(r.film())
    mean = float(np.mean(imgs))
    expect = sum(rho ** k for k in range(D + 1))
    assert abs(mean - expect) < 0.04 * expect, (mean, expect)


def test_training_targets_in_the_furnace(gpu):
    # query vertex 1, depth 1: every training sample sits at a first hit and
    # its target is the one-bounce reflected radiance, rho in expectation.
    rho = 0.6
    sc = furnace_scene(rho)
    r = aqua.Renderer(sc, nrc_cfg(max_depth=1, train_paths=2048, query_vertex=1))
    r.render(1, frame=0)
    n, _, _ = r.nrc_batch()
    assert n > 1000
    st = r.nrc_train_batch(steps=0)
    assert st.rows == 0  # zero steps: nothing trained
    # the cache then learns targets whose mean is rho
    for f in range(1, 40):
        r.render(1, frame=f)
    nrc = aqua.Nrc.__new__(aqua.Nrc)
    nrc._h = aqua.C.c_void_p(r.nrc_handle())
    # states the cache was trained on: points on the walls (the renderer maps
    # the bounds [-1, 1]^3, padded by 1e-3 of the extent, to [0, 1]^3), the
    # inward normal, an outgoing direction in the normal's hemisphere
    rec = np.zeros((256, 16), np.float32)
    rng = np.random.default_rng(0)
    rec[:, 0:3] = rng.uniform(0.05, 0.95, (256, 3))
    axis = rng.integers(0, 3, 256)
    side = rng.integers(0, 2, 256)
    pad = 1e-3 / (1.0 + 2e-3)
    rows = np.arange(256)
    rec[rows, axis] = np.where(side == 0, pad, 1.0 - pad)
    nrm = np.zeros((256, 3))
    nrm[rows, axis] = np.where(side == 0, 1.0, -1.0)
    d = rng.standard_normal((256, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d *= np.sign(np.sum(d * nrm, axis=1, keepdims=True))
    rec[:, 3] = 1.0
    rec[:, 4:7] = d
    rec[:, 8:11] = nrm
    rec[:, 12:15] = rho
    y = nrc.infer(rec)
    nrc._h = None  # owned by the renderer
    assert abs(float(np.median(y)) - rho) < 0.15, float(np.median(y))


@pytest.mark.parametrize("prec", [aqua.NRC_F16, aqua.NRC_BF16, aqua.NRC_FP8])
def test_box_room_cache_reduces_error(gpu, box_room, prec):
    cam = small_camera(box_room.camera(), 96, 96)
    pt = aqua.integrator_default()
    pt.max_depth = 5
    ref = aqua.Renderer(box_room, pt)
    ref.render_oracle(0, 512, camera=cam)
    ref.render(1, camera=cam, frame=1000)
    pt_err, _ = ref.film_error()
    r = aqua.Renderer(box_room, nrc_cfg(max_depth=5, train_paths=2048, precision=prec))
    for f in range(40):
        st = r.render(1, camera=cam, frame=f)
        assert np.isfinite(st.nrc_loss)
    img = r.film(cam)
    assert np.all(np.isfinite(img))
    oracle = ref.oracle_film(cam)
    rel = float(np.mean((img - oracle) ** 2 / (oracle ** 2 + 1e-2)))
    assert rel < pt_err, (rel, pt_err)


def test_diverged_cache_falls_back_to_plain_tracing(gpu, box_room):
    # A learning rate of 1e30 overflows the FP16 weights on the first step, so
    # the next step's loss is not finite: the renderer switches its cache off
    # and later frames are the plain tracer's, bit for bit.
    cam = small_camera(box_room.camera(), 64, 64)
    r = aqua.Renderer(box_room, nrc_cfg(max_depth=5, train_paths=1024, lr=1e30))
    losses = []
    for f in range(4):
        st = r.render(1, camera=cam, frame=f)
        losses.append(st.nrc_loss)
        if st.nrc_queries == 0:
            break
    assert st.nrc_queries == 0, losses
    assert not all(np.isfinite(losses)), losses
    pt = aqua.integrator_default()
    pt.max_depth = 5
    a = aqua.Renderer(box_room, pt)
    a.render(4, camera=cam, frame=9)
    st = r.render(4, camera=cam, frame=9)
    assert st.nrc_queries == 0 and st.nrc_train_records == 0
    np.testing.assert_array_equal(r.film(cam), a.film(cam))


def test_cache_checkpoint_restores_the_training_state(gpu, box_room):
    # A frame depends only on the cache state before it (its own training
    # step follows the film), so after a checkpoint, more training and a
    # restore, the same frame renders bit for bit as it did at the
    # checkpoint; without the restore it does not.
    cam = small_camera(box_room.camera(), 96, 64)
    c = aqua.integrator_default()
    c.type = aqua.INTEGRATOR_NRC
    c.spp, c.max_depth, c.learning_rate, c.nrc_train_paths = 2, 8, 1e-2, 2048
    r = aqua.Renderer(box_room, c)
    for f in range(6):
        r.render(2, frame=f, camera=cam)
    r.nrc_checkpoint()
    r.render(2, frame=100, camera=cam)
    at_ckpt = r.film(cam)
    for f in range(6, 12):
        r.render(2, frame=f, camera=cam)
    r.render(2, frame=100, camera=cam)
    assert not np.array_equal(r.film(cam), at_ckpt)
    r.nrc_checkpoint(restore=True)
    r.render(2, frame=100, camera=cam)
    np.testing.assert_array_equal(r.film(cam), at_ckpt)
    # restoring twice from the same checkpoint gives the same frame again
    r.render(2, frame=50, camera=cam)
    r.nrc_checkpoint(restore=True)
    r.render(2, frame=100, camera=cam)
    np.testing.assert_array_equal(r.film(cam), at_ckpt)


def test_cache_query_precision_switches_on_one_state(gpu, box_room):
    # The query precision is a property of inference only: the same saved
    # state renders a frame in FP8 and in BF16 from one set of weights, the
    # two frames differ by FP8's rounding alone, and switching back and
    # restoring gives the FP8 frame again bit for bit.
    cam = small_camera(box_room.camera(), 96, 64)
    c = nrc_cfg(max_depth=8, train_paths=2048, precision=aqua.NRC_FP8)
    r = aqua.Renderer(box_room, c)
    for f in range(8):
        r.render(2, frame=f, camera=cam)
    r.nrc_checkpoint()
    r.render(2, frame=100, camera=cam)
    fp8 = r.film(cam).astype(np.float64)
    r.nrc_checkpoint(restore=True)
    r.set_nrc_precision(aqua.NRC_BF16)
    st = r.render(2, frame=100, camera=cam)
    assert st.nrc_queries > 0
    bf16 = r.film(cam).astype(np.float64)
    assert not np.array_equal(fp8, bf16)
    rel = np.mean((fp8 - bf16) ** 2 / (bf16 ** 2 + 1e-2))
    assert rel < 0.05, rel
    r.nrc_checkpoint(restore=True)
    r.set_nrc_precision(aqua.NRC_FP8)
    r.render(2, frame=100, camera=cam)
    np.testing.assert_array_equal(r.film(cam), fp8.astype(np.float32))
    with pytest.raises(aqua.AquaError):
        r.set_nrc_precision(7)
    pt = aqua.Renderer(box_room, aqua.integrator_default())
    with pytest.raises(aqua.AquaError):
        pt.set_nrc_precision(aqua.NRC_F16)


def test_visualize_cache_queries_at_the_first_hit(gpu, box_room):
    # visualize_cache renders the cache itself: the first frame (initial
    # weights, training after the film) is bit for bit the frame of a cache
    # queried at vertex 1, and not the default vertex-2 frame
    cam = small_camera(box_room.camera(), 64, 48)
    films = []
    for vis, qv in ((1, 2), (0, 1), (0, 2)):
        c = nrc_cfg(max_depth=5, query_vertex=qv)
        c.visualize_cache = vis
        r = aqua.Renderer(box_room, c)
        st = r.render(2, frame=0, camera=cam)
        assert st.nrc_queries > 0
        films.append(r.film(cam))
    np.testing.assert_array_equal(films[0], films[1])
    assert not np.array_equal(films[0], films[2])
    assert np.all(np.isfinite(films[0]))


def test_training_stops_after_training_iters(gpu, box_room):
    # training_iters caps the cache's Adam steps: the warm-up takes four per
    # frame, so a cap of 6 ends training within the second frame (two steps
    # there); later frames trace no training paths and leave the cache as it
    # was
    cam = small_camera(box_room.camera(), 48, 32)
    c = nrc_cfg(max_depth=5)
    c.training_iters = 6
    r = aqua.Renderer(box_room, c)
    got = [r.render(1, frame=f, camera=cam).nrc_train_records for f in range(4)]
    assert got[0] > 0 and got[1] > 0 and got[2] == 0 and got[3] == 0, got
    r.render(1, frame=50, camera=cam)
    before = r.film(cam)
    r.render(1, frame=7, camera=cam)
    r.render(1, frame=50, camera=cam)
    np.testing.assert_array_equal(r.film(cam), before)


def test_queued_training_sort_matches_the_in_step_sort(gpu, box_room):
    # Once the cache takes one Adam step per frame (after 128), a frame queues
    # its training rows' sort in the pass, ahead of the host's count read;
    # with per-kernel timing on it sorts them inside the step instead. From
    # one state both must train on the same rows: equal loss up to the float
    # atomics' summation order, and next frames that agree just as closely.
    cam = small_camera(box_room.camera(), 64, 48)
    c = aqua.integrator_default()
    c.type = aqua.INTEGRATOR_NRC
    c.spp, c.max_depth, c.learning_rate, c.nrc_train_paths = 1, 6, 1e-2, 2048
    r = aqua.Renderer(box_room, c)
    r.set_opts(timing=False)
    for f in range(34):  # 4 steps a frame: past the first 128
        r.render(1, frame=f, camera=cam)
    r.nrc_checkpoint()
    out = {}
    for timing in (False, True, False):
        r.nrc_checkpoint(restore=True)
        r.set_opts(timing=timing)
        st = r.render(2, frame=40, camera=cam)
        assert st.nrc_train_records > 0
        r.set_opts(timing=False)
        r.render(2, frame=41, camera=cam)
        out.setdefault(timing, []).append((st.nrc_loss, r.film(cam)))
    (l_q, f_q), (l_q2, f_q2) = out[False]
    (l_s, f_s), = out[True]
    assert np.isfinite(l_q) and l_q > 0
    for l, f in ((l_s, f_s), (l_q2, f_q2)):
        assert abs(l - l_q) <= 1e-4 * abs(l_q), (l, l_q)
        assert np.sqrt(np.sum((f - f_q) ** 2) / np.sum(f_q ** 2)) < 1e-3


def _rel(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


def test_train_ahead_queries_see_the_same_weights(gpu, box_room):
    # train_ahead trains a frame's batch at the start of the next frame (its
    # queries wait for the update), so every frame queries the weights it
    # would have queried with the step at the end of the previous frame.
    # Two runs of the same renderer already differ a little (training rows
    # are appended in no fixed order, and the young cache's four steps take
    # a quarter of them each; float atomics): train_ahead must stay within
    # that spread. Frame 0 (no training before it) is the same bit for bit.
    cam = small_camera(box_room.camera(), 64, 48)
    c = nrc_cfg(max_depth=5, train_paths=2048)
    a, a2, b = (aqua.Renderer(box_room, c) for _ in range(3))
    for r in (a, a2):
        r.set_opts(timing=False)
    b.set_opts(timing=False, train_ahead=True)
    spread, gap = [], []
    for f in list(range(8)) + [20]:
        spp = 2 if f == 20 else 1  # frame 20: the last batch is trained first
        sa, _, sb = (r.render(spp, frame=f, camera=cam) for r in (a, a2, b))
        fa, fa2, fb = a.film(cam), a2.film(cam), b.film(cam)
        if f == 0:
            np.testing.assert_array_equal(fa, fb)
        assert sb.nrc_queries == sa.nrc_queries and sb.nrc_train_records > 0
        spread.append(_rel(fa2, fa))
        gap.append(_rel(fb, fa))
    assert max(gap) < max(4 * max(spread), 1e-4), (gap, spread)


def test_train_ahead_checkpoint_keeps_the_waiting_batch(gpu, box_room):
    # the batch that waits for the next frame is part of the checkpoint: after
    # a restore the next frame trains it again before its queries (float
    # atomics: equal up to the spread of two such restores)
    cam = small_camera(box_room.camera(), 64, 48)
    r = aqua.Renderer(box_room, nrc_cfg(max_depth=5, train_paths=2048))
    r.set_opts(timing=False, train_ahead=True)
    for f in range(5):
        r.render(1, frame=f, camera=cam)
    r.nrc_checkpoint()
    r.render(2, frame=100, camera=cam)
    at_ckpt = r.film(cam)
    for f in range(5, 9):
        r.render(1, frame=f, camera=cam)
    r.render(2, frame=100, camera=cam)
    moved = r.film(cam)
    backs = []
    for _ in range(2):
        r.nrc_checkpoint(restore=True)
        r.render(2, frame=100, camera=cam)
        backs.append(r.film(cam))
    spread = _rel(backs[1], backs[0])
    assert _rel(backs[0], at_ckpt) < max(5 * spread, 1e-4), (_rel(backs[0], at_ckpt), spread)
    assert _rel(moved, at_ckpt) > 5 * max(_rel(backs[0], at_ckpt), 1e-4)


def test_train_ahead_diverged_cache_falls_back(gpu, box_room):
    cam = small_camera(box_room.camera(), 64, 64)
    r = aqua.Renderer(box_room, nrc_cfg(max_depth=5, train_paths=1024, lr=1e30))
    r.set_opts(timing=False, train_ahead=True)
    for f in range(6):
        st = r.render(1, camera=cam, frame=f)
        if st.nrc_queries == 0:
            break
    assert st.nrc_queries == 0
    pt = aqua.integrator_default()
    pt.max_depth = 5
    a = aqua.Renderer(box_room, pt)
    a.render(4, camera=cam, frame=9)
    st = r.render(4, camera=cam, frame=9)
    assert st.nrc_queries == 0 and st.nrc_train_records == 0
    np.testing.assert_array_equal(r.film(cam), a.film(cam))
