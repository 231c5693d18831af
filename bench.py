"""Benchmark: plain GPU path tracer on the 1M-triangle interior at 1080p, and
the neural-radiance-cache legs the headline metric is quoted for.

Workload (one step = one frame): procedural interior of 150 seeded meshes
(~1.09 M FP32 triangles, GGX metallic-roughness materials, 3 quad area
lights), 1920x1080, 64 spp, max depth 8, binned-SAH BVH (32 bins). The scene
is written in the reference's schema (scene JSON + BSON .mesh), loaded, and
uploaded once; every timed frame traces all 64 samples of every pixel with
the scene already resident in HBM.

Prints one JSON line: Mrays/s (camera + extension + shadow rays per second of
wall time), ms/frame, the frame's relMSE against a 256-spp FP64 render of the
same view (GPU, FP64 accumulation), the roofline of the closest-hit traversal
kernel, and the CPU restatement's rate on a bounded sample of the same frame
on all host cores. Unless --no-nrc, further legs follow:
  box_room       the 36-triangle box room, 512x512, 16 spp, depth 5, with its
                 relMSE against a 4096-spp FP64 render;
  material_sort  shading in material order vs path order (textured room
                 subset, interior), 1080p, 16 spp;
  nrc_infer      4-spp cache frames at 1080p: frame time and the cache's own work;
  nrc_train      one cache training step on 2^18 tracer samples;
  nrc_eval       ms/frame to relMSE <= 1e-2 against a 4096-spp FP64 render,
                 plain tracer vs cache (FP16 and BF16, query at vertex 2; FP16
                 at vertex 1 beside it) on 24 frames of the box-room
                 fly-through at 1080p, with independent sample draws and
                 again with stratified camera / first-hit dimensions;
  large          the 400-mesh / 5.4 M-triangle interior at 3840x2160, 1 spp
                 with an FP8 (e4m3) and an FP16 cache over a 60-frame
                 fly-through, error against a 64-spp FP64 render.
"""
import argparse
import gc
import hashlib
import importlib.util
import json
import os
import platform
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))

METRIC = "ms/frame to relMSE 1e-2 vs FP64 oracle (NRC vs plain PT); Mrays/s"
HBM_PEAK_GBS = 8000.0
# VALU issue peak of MI355X in wave64 instructions: the published FP32 vector
# rate, 157.3 TFLOP/s, is 128 FMA lanes per CU per clock (256 CUs, 2.4 GHz),
# i.e. one wave64 instruction per SIMD every 2 clocks. (One every 4 clocks
# cannot be the limit: the traversal kernel's counters show 0.35 instructions
# per SIMD-clock.)
VALU_PEAK_GINSTR = round(157.3e3 / (2 * 64), 1)
# Algorithmic bytes of one closest-hit ray: the ray (origin+tmax, direction:
# 32 B) read once, the hit record (slot, t: 8 B) written once, one 64-byte
# quantised 4-wide node per node visit and one 48-byte triangle per triangle
# test.
RAY_BYTES, HIT_BYTES, NODE_BYTES, TRI_BYTES = 32, 8, 64, 48
# The sources the traversal kernels' instruction streams come from (the
# kernels, their work queue, the node / triangle tests, the vector math and
# the queue and BVH structures): a committed counter profile counts only
# while their digest matches.
TRAVERSAL_SOURCES = ("aqua-engine_amd/csrc/device/traverse.h", "aqua-engine_amd/csrc/device/trace.h",
                     "aqua-engine_amd/csrc/device/common.h", "aqua-engine_amd/csrc/device/kernels.h")
VALU_PROFILE = "profiles/k_intersect_interior1m_1080p_64spp_valu.json"
TRAFFIC_PROFILE = "profiles/k_intersect_interior1m_1080p_64spp_traffic.json"
SHADOW_VALU_PROFILE = "profiles/k_shadow_interior1m_1080p_64spp_valu.json"
# the hash-grid gradient scatter's atomic requests per training step (PMC),
# valid while the cache sources are the ones it was counted on
CACHE_SOURCES = ("aqua-engine_amd/csrc/nrc.hip",)
SCATTER_PROFILE = "profiles/k_nrc_scatter_boxroom_1080p_2e18rows_atomics.json"


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def _digest(sources):
    h = hashlib.sha256()
    for rel in sources:
        with open(os.path.join(ROOT, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def traversal_digest():
    return _digest(TRAVERSAL_SOURCES)


def cache_digest():
    return _digest(CACHE_SOURCES)


def _profile(rel, digest=traversal_digest):
    """A committed counter summary, or None when absent or measured on other
    sources than these (the traversal sources unless `digest` says otherwise)."""
    try:
        with open(os.path.join(ROOT, rel)) as f:
            p = json.load(f)
    except (OSError, ValueError):
        return None
    return p if p.get("source_digest") == digest() else None


def atomic_ceiling_here(n=1 << 24, table=1 << 23, reps=10):
    """G float atomics per second with every lane on its own random word of a
    32 MiB table (torch index_add_: one global_atomic_add_f32 per element, so
    one 64-byte segment per lane): the rate the hash-grid scatter is held to.
    Its streamed reads (index 8 B + value 4 B per atomic) are ~1/20 of the
    time at HBM speed. Needs a process in which torch is the first to use the
    GPU (torch carries its own HIP runtime, which finds no device once the
    library's runtime holds it): see atomic_ceiling."""
    import torch
    dev = torch.device("cuda", 0)
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    t = torch.zeros(table, dtype=torch.float32, device=dev)
    idx = torch.randint(0, table, (n,), device=dev, generator=g)
    src = torch.ones(n, dtype=torch.float32, device=dev)
    for _ in range(2):
        t.index_add_(0, idx, src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        t.index_add_(0, idx, src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    return n / (ms * 1e-3) / 1e9


def atomic_ceiling():
    """atomic_ceiling_here in a child process (bench.py --atomic-ceiling), or
    None when that process fails (no GPU for torch, say)."""
    import subprocess
    try:
        out = subprocess.run([sys.executable, os.path.abspath(__file__), "--atomic-ceiling"], capture_output=True,
                             text=True, timeout=300, check=True)
        return float(out.stdout.strip().splitlines()[-1])
    except (subprocess.SubprocessError, ValueError, IndexError) as e:
        print(f"bench: atomic rate not measured ({type(e).__name__})", file=sys.stderr, flush=True)
        return None


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(scene, aqua, max_depth, target_s=12.0):
    """CPU restatement (oracle/) on a band of rows of the same frame, on every
    host core the process may use."""
    orc = _load("aqua_oracle", os.path.join(ROOT, "oracle", "oracle.py"))
    v, m = scene.triangles()
    t0 = time.perf_counter()
    o = orc.Scene(v, m, scene.materials(), scene.point_lights())
    build_s = time.perf_counter() - t0
    cam = scene.camera()
    cam16 = aqua.camera_to_array(cam)
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    mid = cam.height // 2
    rows, spp = 8, 1
    while True:
        y0 = max(0, mid - rows // 2)
        y1 = min(cam.height, y0 + rows)
        t0 = time.perf_counter()
        _, st = o.render(cam16, spp, max_depth, frame=0, y0=y0, y1=y1, threads=threads)
        dt = time.perf_counter() - t0
        if dt >= 0.7 * target_s:
            break
        if y1 - y0 < cam.height:
            rows = min(cam.height, max(rows * 2, int(rows * target_s / max(dt, 1e-3) * 1.1)))
        else:
            spp = max(spp + 1, int(spp * target_s / max(dt, 1e-3)))
    return {
        "value": round(st.rays / dt / 1e6, 3),
        "unit": "Mrays/s",
        "cores": threads,
        "nproc": os.cpu_count(),
        "cpu": _cpu_model(),
        "kind": "port",
        "sample": f"rows {y0}-{y1} of the {cam.width}x{cam.height} frame at {spp} spp, max depth {max_depth} "
                  f"({st.rays} rays in {dt:.2f} s on {threads} threads; its own BVH build, {build_s:.2f} s, "
                  f"not timed)",
    }


def _relmse(img, ref):
    import numpy as np
    return float(np.mean((img.astype(np.float64) - ref) ** 2 / (ref ** 2 + 1e-2)))


def box_room_leg(aqua, oracle_spp=4096, steps=10):
    """The 36-triangle box room at 512x512, 16 spp, depth 5 (PCG seed 0)."""
    with tempfile.TemporaryDirectory() as d:
        scene = aqua.Scene.load(aqua.procgen(aqua.PROCGEN_BOX_ROOM, d))
    cfg = aqua.integrator_default()
    cfg.spp, cfg.max_depth = 16, 5
    r = aqua.Renderer(scene, cfg)
    r.set_opts(timing=False)
    for w in range(2):
        r.render(16, frame=w)
    rays, t0 = 0, time.perf_counter()
    for k in range(steps):
        st = r.render(16, frame=0)
        rays += st.rays
    wall = time.perf_counter() - t0
    r.render_oracle(0, oracle_spp, frame=777)
    e, _ = r.film_error()
    # the same frame with the stratified sampler (camera and first-hit
    # dimensions), against the same FP64 render
    import numpy as np
    ref = r.oracle_film()
    cfg.sampler = aqua.SAMPLER_STRATIFIED
    rs = aqua.Renderer(scene, cfg)
    rs.set_opts(timing=False)
    rs.render(16, frame=0)
    x = rs.film().astype(np.float64)
    e_s = float(np.mean((x - ref) ** 2 / (ref * ref + 1e-2)))
    return {"scene": "box room, 36 triangles, 512x512, 16 spp, max depth 5",
            "ms_per_frame": round(wall * 1e3 / steps, 3), "mrays_per_s": round(rays / wall / 1e6, 1),
            "relmse_vs_fp64": round(e, 6), "relmse_vs_fp64_stratified": round(e_s, 6), "oracle_spp": oracle_spp}


def material_sort_leg(aqua, spp=16, steps=5, width=1920, height=1080):
    """Shading in material order (aqua_renderer_opts.sort_materials) against
    path order, plain tracer, depth 8: the textured living-room subset
    (16 materials, 8 images; the fixture under tests/golden) at its own
    camera, and the 150-mesh interior. ms/frame (kernels overlapped) and the
    shading time per frame (per-kernel events; with the sort on it includes
    the sort's kernels), both ways."""
    import numpy as np
    out = {}
    with tempfile.TemporaryDirectory() as d:
        scenes = {
            "room_subset": aqua.Scene.load(os.path.join(ROOT, "tests", "golden", "room_subset", "scene.json"),
                                           emission_from_mtl=True),
            "interior": aqua.Scene.load(aqua.procgen(aqua.PROCGEN_INTERIOR, d, seed=0, n_meshes=150,
                                                     target_tris=1000000, width=width, height=height)),
        }
    for name, scene in scenes.items():
        base = scene.camera()
        cam = aqua.Camera()
        cam.cam_to_world[:] = base.cam_to_world[:]
        cam.tan_half_fov = base.tan_half_fov
        cam.width, cam.height = width, height
        cam.aspect = np.float32(np.float32(width) / np.float32(height))
        cfg = aqua.integrator_default()
        cfg.spp, cfg.max_depth = spp, 8
        r = aqua.Renderer(scene, cfg)
        res = {"materials": scene.info().n_materials}
        films = {}
        for sort in (False, True):
            r.set_opts(timing=False, sort_materials=sort)
            r.render(spp, frame=0, camera=cam)
            t0 = time.perf_counter()
            for k in range(steps):
                r.render(spp, frame=k, camera=cam)
            ms = (time.perf_counter() - t0) * 1e3 / steps
            films[sort] = r.film(cam)
            r.set_opts(timing=True, sort_materials=sort)
            st = r.render(spp, frame=steps - 1, camera=cam)
            res["sorted" if sort else "path_order"] = {"ms_per_frame": round(ms, 3),
                                                       "ms_shade": round(st.ms_shade, 3)}
        res["same_film"] = bool(np.array_equal(films[False], films[True]))
        out[name] = res
    out["config"] = f"{width}x{height}, {spp} spp, max depth 8"
    return out


def nrc_train_leg(aqua, rows=1 << 18, steps=20, precision=0, ceiling=None):
    """Cache training step alone on samples the path tracer produced: box room,
    1920x1080, training suffix paths only, then `steps` Adam steps on 2^18 of
    those samples resident in HBM."""
    with tempfile.TemporaryDirectory() as d:
        scene = aqua.Scene.load(aqua.procgen(aqua.PROCGEN_BOX_ROOM, d, width=1920, height=1080))
    cfg = aqua.integrator_default()
    cfg.type = aqua.INTEGRATOR_NRC
    cfg.max_depth = 8
    cfg.learning_rate = 1e-2
    cfg.nrc_precision = precision
    cfg.nrc_train_paths = 1 << 18
    r = aqua.Renderer(scene, cfg)
    n = 0
    for f in range(4):
        r.render(1, frame=f)
        n, _, _ = r.nrc_batch()
        if n >= rows:
            break
    n_use = min(n, rows)
    r.nrc_train_batch(steps=2)  # warm-up
    n0, a, b = r.nrc_batch()
    st = aqua.NrcStats()
    aqua._check(aqua.lib().aqua_nrc_train_device(r.nrc_handle(), n_use, a, b, steps, aqua.C.byref(st)))
    ms = st.ms_total / steps
    table_params = 16 * (1 << 19) * 2
    mlp_params = 4 * 64 * 64 + 64 * 16
    # bytes per sample: record 64 B + target 16 B in; 16 levels x 8 corners x
    # 4 B gathered; 5 layer inputs and 5 pre-activation gradients (128 B each)
    # written, then read by the weight gradients; 16 feature-pair gradients
    # (128 B) written, then read with the position (16 B) by the scatter, which
    # adds 8 levels x 8 corners x 8 B (FP32 pairs) and 8 x 8 x 4 B (BF16
    # pairs). Adam reads param/grad/m/v (16 B) and writes param/grad/m/v
    # (16 B) + the FP16 copy (2 B) per table parameter.
    per_sample = 80 + 16 * 8 * 4 + 2 * 10 * 128 + 128 + 144 + 8 * 8 * 8 + 8 * 8 * 4
    bytes_step = n_use * per_sample + table_params * 34 + mlp_params * 34
    flops_step = 3 * 2 * n_use * (4 * 64 * 64 + 64 * 16)
    return {
        "rows": n_use,
        "precision": ["f16", "bf16"][min(precision, 1)],
        "ms_per_step": round(ms, 4),
        "samples_per_s": round(n_use / (ms * 1e-3), 1),
        "ms_fwd_bwd": round(st.ms_fwd_bwd / steps, 4),
        "ms_wgrad": round(st.ms_wgrad / steps, 4),
        "ms_adam": round(st.ms_adam / steps, 4),
        "ms_scatter": round(st.ms_scatter / steps, 4),
        "loss": round(float(st.loss), 5),
        "roofline_hbm": {"achieved": round(bytes_step / (ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "bytes_per_step": bytes_step,
                         "note": "not the bound: the hash-grid scatter's atomics are (one 64-byte segment "
                                 "per instruction and segment touched)"},
        "mfma_tflops": round(flops_step / (ms * 1e-3) / 1e12, 3),
        "roofline_atomics": scatter_roofline(st.ms_scatter / steps, n_use, ceiling),
    }


def scatter_roofline(ms_scatter, rows, ceiling):
    """The scatter against the atomic rate: its 64-byte atomic segments per
    step (L1-to-L2 atomic requests, counted once with rocprofv3 on this batch)
    over its HIP-event time, against the rate measured live for one lane per
    random word (atomic_ceiling)."""
    if ceiling is None:
        return {"note": "atomic rate not measured in this run"}
    out = {"unit": "G 64-byte atomic segments/s", "peak": round(ceiling, 2),
           "peak_source": "measured here: torch index_add_ of 2^24 ones into a 32 MiB FP32 table at random "
                          "indices (one atomic per lane and segment)"}
    p = _profile(SCATTER_PROFILE, cache_digest)
    if p is None or p.get("rows") != rows:
        out["note"] = "no counted segments for these cache sources or this batch"
        return out
    seg = p["atomic_requests_per_step"]
    out.update({"segments_per_step": seg, "segments_source": SCATTER_PROFILE,
                "achieved": round(seg / (ms_scatter * 1e-3) / 1e9, 2)})
    out["frac"] = round(out["achieved"] / ceiling, 3)
    return out


def _ladder(render, oracle, max_spp, before=None):
    """Smallest spp on the ladder 1, 2, 4, ... whose frame reaches relMSE <= 1e-2.
    `before` runs ahead of every rung (the cache puts its state back there)."""
    res = None
    spp = 1
    while spp <= max_spp:
        if before:
            before()
        st, img = render(spp)
        e = _relmse(img, oracle)
        res = {"spp": spp, "relmse": round(e, 5), "ms_per_frame": round(st.ms_frame, 3)}
        if st.nrc_queries:
            res["queries"] = int(st.nrc_queries)
        if e <= 1e-2:
            break
        spp *= 2
    return res


def nrc_infer_leg(aqua, warm_frames=40, frames=5, spp=4, width=1920, height=1080):
    """Cache frames alone: the box room fly-through at 1080p with an FP16 cache
    (query at vertex 2, 2^16 training paths), trained online over
    `warm_frames` 1-spp frames, then `frames` frames of `spp` samples. ms per
    frame with kernels overlapped, and with per-kernel events the cache's own
    work in a frame (query sort, inference, apply, the frame's training step)."""
    import numpy as np
    with tempfile.TemporaryDirectory() as d:
        scene = aqua.Scene.load(aqua.procgen(aqua.PROCGEN_BOX_ROOM, d, width=width, height=height))
    c = aqua.integrator_default()
    c.type = aqua.INTEGRATOR_NRC
    c.max_depth = 8
    c.learning_rate = 1e-2
    c.nrc_precision = aqua.NRC_F16
    c.nrc_train_paths = 1 << 16
    r = aqua.Renderer(scene, c)
    r.set_opts(timing=False)
    for f in range(warm_frames):
        r.render(1, frame=f, camera=aqua.camera_flythrough(scene, f, 240, width=width, height=height))
    ms, nq = [], 0
    for f in range(warm_frames, warm_frames + frames):
        st = r.render(spp, frame=f, camera=aqua.camera_flythrough(scene, f, 240, width=width, height=height))
        ms.append(st.ms_frame)
        nq = int(st.nrc_queries)
    r.set_opts(timing=True)
    mc, mf = [], []
    for f in range(warm_frames + frames, warm_frames + 2 * frames):
        st = r.render(spp, frame=f, camera=aqua.camera_flythrough(scene, f, 240, width=width, height=height))
        mc.append(st.ms_nrc)
        mf.append(st.ms_frame)
    return {"scene": f"box room, {width}x{height}, {spp} spp, max depth 8, FP16 cache after {warm_frames} frames",
            "queries_per_frame": nq, "ms_per_frame": round(float(np.mean(ms)), 3),
            "ms_per_frame_kernels_serial": round(float(np.mean(mf)), 3),
            "ms_cache_work": round(float(np.mean(mc)), 3)}


def nrc_eval_leg(aqua, frames=tuple(range(5, 240, 10)), n_frames=240, oracle_spp=4096, width=1920,
                 height=1080, max_spp=256, sampler=0, oracles=None, train_ahead=False):
    """ms/frame to relMSE <= 1e-2 against an FP64 render of `oracle_spp` samples,
    plain path tracer vs cache, on the box room fly-through: max depth 8, 2^16
    training suffix paths per frame, online training at 1 spp through every
    frame before an evaluated one. Every tenth frame of the 240 is evaluated
    (24 frames; the FP64 render takes ~3.7 s a frame, so all 240 would take a
    quarter of an hour); the figures are means over them.

    Each evaluated frame uses three disjoint sample sets: the oracle renders
    frame index f, the plain tracer f + 100000, the cache f + 200000. The
    cache's state is checkpointed after the frames before f and restored
    before every rung of the spp ladder, so each rung sees exactly the cache
    those earlier frames trained -- and pays for its own training step, as
    in online use. The caches query at path vertex 2 in FP16 and in BF16 (the
    configuration the ratio is quoted for); an FP16 cache querying at vertex
    1 is reported beside them.

    `sampler` is the tracer's and the caches' sample generator (independent
    PCG32 draws, or stratified camera / first-hit dimensions). The FP64
    renders are filled into `oracles` (frame -> film) on the first call and
    reused by a later one. With `train_ahead` the caches train each frame's
    batch at the start of the next frame, beside its first bounces (the
    checkpoint keeps the batch that waits, so every rung trains it)."""
    with tempfile.TemporaryDirectory() as d:
        scene = aqua.Scene.load(aqua.procgen(aqua.PROCGEN_BOX_ROOM, d, width=width, height=height))
    pt = aqua.integrator_default()
    pt.max_depth = 8
    pt.sampler = sampler
    ref = aqua.Renderer(scene, pt)
    ref.set_opts(timing=False)  # ms/frame from the frame's own events; kernels overlapped
    caches = {}
    for key, qv, prec in (("nrc_v2_f16", 2, aqua.NRC_F16), ("nrc_v2_bf16", 2, aqua.NRC_BF16),
                          ("nrc_v1_f16", 1, aqua.NRC_F16)):
        c = aqua.integrator_default()
        c.type = aqua.INTEGRATOR_NRC
        c.max_depth = 8
        c.learning_rate = 1e-2
        c.nrc_precision = prec
        c.nrc_train_paths = 1 << 16
        c.nrc_query_vertex = qv
        c.sampler = sampler
        caches[key] = aqua.Renderer(scene, c)
        caches[key].set_opts(timing=False, train_ahead=train_ahead)
    out = []
    next_frame = 0
    t_oracle = 0.0
    oracles = {} if oracles is None else oracles
    for fe in frames:
        cam = aqua.camera_flythrough(scene, fe, n_frames, width=width, height=height)
        for f in range(next_frame, fe):
            cf = aqua.camera_flythrough(scene, f, n_frames, width=width, height=height)
            for r in caches.values():
                r.render(1, frame=f, camera=cf)
        next_frame = fe
        print(f"bench: nrc_eval frame {fe}", file=sys.stderr, flush=True)
        if fe not in oracles:
            t0 = time.perf_counter()
            ref.render_oracle(0, oracle_spp, frame=fe, camera=cam)
            oracles[fe] = ref.oracle_film(cam)
            t_oracle += time.perf_counter() - t0
        oracle = oracles[fe]
        row = {"frame": fe}

        def pt_render(spp):
            st = ref.render(spp, frame=fe + 100000, camera=cam)
            return st, ref.film(cam)
        row["pt"] = _ladder(pt_render, oracle, max_spp)
        for key, r in caches.items():
            r.nrc_checkpoint()

            def nrc_render(spp, r=r):
                st = r.render(spp, frame=fe + 200000, camera=cam)
                return st, r.film(cam)
            row[key] = _ladder(nrc_render, oracle, max_spp, before=lambda r=r: r.nrc_checkpoint(restore=True))
            r.nrc_checkpoint(restore=True)   # the online run goes on from the frames before fe
        out.append(row)
    pt_ms = sum(o["pt"]["ms_per_frame"] for o in out) / len(out)
    res = {
        "scene": "box room (36 tris), fly-through of %d frames, %dx%d, max depth 8" % (n_frames, width, height),
        "frames_evaluated": list(frames),
        "sampler": ("stratified camera / first-hit dimensions (shuffled Owen-scrambled Sobol)" if sampler
                    else "independent PCG32 draws"),
        "oracle_spp": oracle_spp,
        "oracle_s_per_frame": round(t_oracle / len(frames), 2) if t_oracle > 0 else "reused",
        "sample_sets": "oracle frame f, plain tracer f + 100000, cache f + 200000; cache state restored per rung",
        "frames": out,
        "pt_ms_to_1e-2": round(pt_ms, 3),
        "pt_reached": all(o["pt"]["relmse"] <= 1e-2 for o in out),
    }
    for key in caches:
        ms = sum(o[key]["ms_per_frame"] for o in out) / len(out)
        res[key + "_ms_to_1e-2"] = round(ms, 3)
        res[key + "_over_pt"] = round(ms / pt_ms, 4) if pt_ms > 0 else None
        res[key + "_reached"] = all(o[key]["relmse"] <= 1e-2 for o in out)
    res["nrc_over_pt"] = res["nrc_v2_f16_over_pt"]  # the configured default: FP16, vertex 2
    return res


def large_leg(aqua, meshes=400, tris=5000000, width=3840, height=2160, frames=60, oracle_spp=64,
              eval_frames=(50, 55, 59)):
    """Large interior (about 5.4 M triangles, 3840x2160): 1 spp per frame with
    an online cache over a `frames`-frame fly-through, once with FP8 (e4m3)
    queries + BF16 training and once with FP16 throughout. At the evaluated
    frames each cache's frame is compared with a 64-spp FP64 render and with
    the other cache's; the plain tracer's 1-spp frame on the same samples
    is compared too. The caches take the same number of Adam steps (the
    frames' schedule); relMSE figures are means over the evaluated frames.
    The two caches are different models (BF16 against FP16 training, and
    float atomics in the gradient), so their images drift apart; the error
    of FP8 queries alone is measured on the FP8 cache itself: at each
    evaluated frame its state is saved, the frame rendered with BF16 queries
    (the training precision), the state put back and the frame rendered
    with FP8 queries, from the same weights and samples."""
    import numpy as np
    t0 = time.perf_counter()
    with tempfile.TemporaryDirectory() as d:
        path = aqua.procgen(aqua.PROCGEN_INTERIOR, d, seed=0, n_meshes=meshes, target_tris=tris,
                            width=width, height=height)
        scene = aqua.Scene.load(path)
    load_s = time.perf_counter() - t0
    info = scene.info()
    pt = aqua.integrator_default()
    pt.max_depth = 8
    t0 = time.perf_counter()
    ref = aqua.Renderer(scene, pt)
    build_s = time.perf_counter() - t0
    ref.set_opts(timing=False)
    res = {"scene": "procedural interior, %d meshes, %d triangles, %dx%d, max depth 8" % (
               meshes, info.n_triangles, width, height),
           "load_s": round(load_s, 2), "bvh_build_s": round(build_s, 2), "frames": frames,
           "eval_frames": list(eval_frames), "oracle_spp": oracle_spp}
    cams = [aqua.camera_flythrough(scene, f, frames, width=width, height=height) for f in range(frames)]
    imgs = {}
    for name, prec in (("fp8", aqua.NRC_FP8), ("f16", aqua.NRC_F16)):
        print(f"bench: large, {name} cache", file=sys.stderr, flush=True)
        c = aqua.integrator_default()
        c.type = aqua.INTEGRATOR_NRC
        c.max_depth = 8
        c.learning_rate = 1e-2
        c.nrc_precision = prec
        c.nrc_train_paths = 1 << 16
        r = aqua.Renderer(scene, c)
        r.set_opts(timing=False)
        ms, ms_cache, rays, wall = [], [], 0, 0.0
        imgs[name] = {}
        for f in range(frames):
            if name == "fp8" and f in eval_frames:
                r.nrc_checkpoint()
                r.set_nrc_precision(aqua.NRC_BF16)
                r.render(1, frame=f, camera=cams[f])
                imgs.setdefault("fp8_as_bf16", {})[f] = r.film(cams[f]).astype(np.float64)
                r.nrc_checkpoint(restore=True)
                r.set_nrc_precision(aqua.NRC_FP8)
            t1 = time.perf_counter()
            st = r.render(1, frame=f, camera=cams[f])
            wall += time.perf_counter() - t1
            if f >= 2:
                ms.append(st.ms_frame)
            rays += st.rays
            if f in eval_frames:
                imgs[name][f] = r.film(cams[f]).astype(np.float64)
        res[name] = {"ms_per_frame": round(float(np.mean(ms)), 3),
                     "mrays_per_s": round(rays / wall / 1e6, 1),
                     "queries_last_frame": int(st.nrc_queries),
                     "loss_last_frame": round(float(st.nrc_loss), 5)}
        del r
    e = {"pt_1spp": [], "nrc_fp8": [], "nrc_f16": [], "fp8_vs_f16": [], "fp8_vs_bf16_same_weights": []}
    pt_ms = []
    for f in eval_frames:
        st = ref.render(1, frame=f, camera=cams[f])
        pt_ms.append(st.ms_frame)
        pt1 = ref.film(cams[f]).astype(np.float64)
        ref.render_oracle(0, oracle_spp, frame=f + 300000, camera=cams[f])
        oracle = ref.oracle_film(cams[f])
        e["pt_1spp"].append(_relmse(pt1, oracle))
        e["nrc_fp8"].append(_relmse(imgs["fp8"][f], oracle))
        e["nrc_f16"].append(_relmse(imgs["f16"][f], oracle))
        e["fp8_vs_f16"].append(_relmse(imgs["fp8"][f], imgs["f16"][f]))
        e["fp8_vs_bf16_same_weights"].append(_relmse(imgs["fp8"][f], imgs["fp8_as_bf16"][f]))
    res["pt_1spp_ms"] = round(float(np.mean(pt_ms)), 3)
    res["relmse_vs_fp64"] = {k: round(float(np.mean(v)), 5) for k, v in e.items() if k.startswith(("pt", "nrc"))}
    res["relmse_fp8_vs_f16"] = round(float(np.mean(e["fp8_vs_f16"])), 6)
    res["relmse_fp8_vs_bf16_same_weights"] = round(float(np.mean(e["fp8_vs_bf16_same_weights"])), 6)
    res["relmse_per_frame"] = {k: [round(x, 5) for x in v] for k, v in e.items()}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--meshes", type=int, default=150)
    ap.add_argument("--tris", type=int, default=1000000)
    ap.add_argument("--max-depth", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-nrc", action="store_true",
                    help="skip the extra legs (box room, cache training step, cache evaluation, large scene)")
    ap.add_argument("--only-nrc", action="store_true", help="skip the interior workload, run the extra legs")
    ap.add_argument("--no-large", action="store_true", help="skip the 5.4 M-triangle 3840x2160 leg")
    ap.add_argument("--legs", default="box,sort,infer,train,eval,large",
                    help="the extra legs to run, a comma list of box, sort, infer, train, eval, large")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run a plain 256 MiB device copy (the reference kernel for HBM counters)")
    ap.add_argument("--atomic-ceiling", action="store_true",
                    help="print the atomic rate of one lane per random word (G/s) and exit")
    ap.add_argument("--no-atomic-ceiling", action="store_true",
                    help="do not measure that rate for the training leg (runs under a profiler, whose "
                         "preloaded library has the GPU before the child process could start)")
    ap.add_argument("--profile-frame", action="store_true",
                    help="render frame 0 of the interior workload once (kernels overlapped, as in the timed "
                         "frames), print its ray counts as JSON and exit: the run counter profiles use")
    args = ap.parse_args()
    if args.atomic_ceiling:
        print(atomic_ceiling_here(), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the path tracer would shard by interleaved pixel tiles, one process per GPU)", file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy:
        # A copy of known size: 256 MiB read + 256 MiB written. Profiled with
        # the workload, its FETCH_SIZE / WRITE_SIZE turn the counters of the
        # traversal kernel into bytes.
        import torch
        x = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        del x, y

    legs = set(args.legs.split(","))
    if args.no_large:
        legs.discard("large")
    # the atomic rate for the training leg's scatter, measured in a child
    # process before this one touches the GPU
    ceiling = None
    if "train" in legs and not args.no_nrc and not args.profile_frame and not args.no_atomic_ceiling:
        ceiling = atomic_ceiling()

    aqua = _load("aqua_amd", os.path.join(ROOT, "aqua-engine_amd", "aqua.py"))
    if aqua.device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(1)

    def extra_legs(out):
        if "box" in legs:
            print("bench: box room leg", file=sys.stderr, flush=True)
            out["box_room"] = box_room_leg(aqua)
        if "sort" in legs:
            print("bench: material sort leg", file=sys.stderr, flush=True)
            out["material_sort"] = material_sort_leg(aqua)
        if "infer" in legs:
            print("bench: cache inference leg", file=sys.stderr, flush=True)
            out["nrc_infer"] = nrc_infer_leg(aqua)
        if "train" in legs:
            print("bench: cache training leg", file=sys.stderr, flush=True)
            out["nrc_train"] = nrc_train_leg(aqua, ceiling=ceiling)
        if "eval" in legs:
            print("bench: cache evaluation leg", file=sys.stderr, flush=True)
            oracles = {}
            out["nrc_eval"] = nrc_eval_leg(aqua, oracles=oracles)
            gc.collect()  # the first set of renderers gives its GPU queues back
            print("bench: cache evaluation leg, stratified sampler", file=sys.stderr, flush=True)
            out["nrc_eval_stratified"] = nrc_eval_leg(aqua, sampler=aqua.SAMPLER_STRATIFIED, oracles=oracles)
            del oracles
            ev, es = out["nrc_eval"], out["nrc_eval_stratified"]
            out["cache_over_tracer"] = {
                "metric": "cache ms/frame to relMSE 1e-2 over the plain tracer's (query at vertex 2, 24 frames)",
                "target": 0.25,
                "stratified": {"f16": es["nrc_v2_f16_over_pt"], "bf16": es["nrc_v2_bf16_over_pt"]},
                "independent": {"f16": ev["nrc_v2_f16_over_pt"], "bf16": ev["nrc_v2_bf16_over_pt"]},
            }
        if "large" in legs:
            print("bench: large-scene leg", file=sys.stderr, flush=True)
            out["large"] = large_leg(aqua)

    if args.only_nrc:
        out = {}
        extra_legs(out)
        print(json.dumps(out))
        return

    with tempfile.TemporaryDirectory() as d:
        path = aqua.procgen(aqua.PROCGEN_INTERIOR, d, seed=0, n_meshes=args.meshes,
                            target_tris=args.tris, width=args.width, height=args.height)
        scene = aqua.Scene.load(path)
    info = scene.info()
    cfg = aqua.integrator_default()
    cfg.spp, cfg.max_depth, cfg.seed = args.spp, args.max_depth, 0
    t0 = time.perf_counter()
    r = aqua.Renderer(scene, cfg, device=0)
    setup_s = time.perf_counter() - t0

    if args.profile_frame:
        r.set_opts(count_traversal=False, timing=False)
        st = r.render(args.spp, frame=0)
        print(json.dumps({"frame": 0, "closest_rays": st.camera_rays + st.extension_rays,
                          "shadow_rays": st.shadow_rays, "launches": st.n_intersect_launches,
                          "ms_frame": round(st.ms_frame, 3), "source_digest": traversal_digest()}))
        return

    # Instrumented frame (untimed): nodes / triangles per closest-hit ray.
    r.set_opts(count_traversal=True, timing=False)
    cst = r.render(args.spp, frame=0)
    closest_rays = cst.camera_rays + cst.extension_rays
    nodes_per_ray = cst.nodes_visited / closest_rays
    tris_per_ray = cst.tris_tested / closest_rays
    bytes_per_ray = RAY_BYTES + HIT_BYTES + NODE_BYTES * nodes_per_ray + TRI_BYTES * tris_per_ray

    # Timed frames: no per-kernel events, so the shadow rays of each bounce
    # overlap the next bounce's closest-hit rays (wall clock only).
    print("bench: interior, timed frames", file=sys.stderr, flush=True)
    r.set_opts(count_traversal=False, timing=False)
    for w in range(args.warmup):
        r.render(args.spp, frame=w)
    rays = 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        st = r.render(args.spp, frame=k)
        rays += st.rays
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    mrays = rays / wall / 1e6
    # One more frame with HIP events around every launch (kernels serialized)
    # gives the closest-hit kernel's own launch time for the roofline.
    r.set_opts(count_traversal=False, timing=True)
    st = r.render(args.spp, frame=0)
    ms_isect = st.ms_intersect
    launches = st.n_intersect_launches
    isect_rays = st.camera_rays + st.extension_rays
    ms_shadow, shadow_rays = st.ms_shadow, st.shadow_rays
    # The frame's error against a 256-spp FP64 render of the same view
    # (its own sample set), reduced on the GPU in FP64.
    r.render_oracle(0, 256, frame=777)
    r.set_opts(count_traversal=False, timing=False)
    r.render(args.spp, frame=0)
    relmse, relrmse = r.film_error()

    avg_launch_ms = ms_isect / max(1, launches)
    rays_per_launch = isect_rays / max(1, launches)
    bytes_per_launch = bytes_per_ray * rays_per_launch
    alg_gbs = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9 if avg_launch_ms > 0 else 0.0
    node_rate = nodes_per_ray * rays_per_launch / (avg_launch_ms * 1e-3) if avg_launch_ms > 0 else 0.0

    valu = _profile(VALU_PROFILE)
    traffic = _profile(TRAFFIC_PROFILE)
    roof = {"kernel": "k_intersect (closest hit)", "bound": "valu", "unit": "G VALU instr/s",
            "peak": VALU_PEAK_GINSTR, "achieved": None, "frac": None,
            "peak_derivation": "157.3 TFLOP/s FP32 vector / (2 FLOP x 64 lanes) = 256 CUs x 4 SIMDs x 2.4 GHz / "
                               "2 clocks per wave64 VALU instruction",
            "source_digest": traversal_digest()}
    if valu:
        # wave64 VALU instructions per closest-hit ray (PMC SQ_INSTS_VALU of
        # the same workload) x rays per launch / the launch time measured here
        ginstr = valu["valu_insts_per_ray"] * rays_per_launch / (avg_launch_ms * 1e-3) / 1e9
        roof.update(achieved=round(ginstr, 1), frac=round(ginstr / VALU_PEAK_GINSTR, 4),
                    valu_insts_per_ray=valu["valu_insts_per_ray"],
                    lane_utilisation=valu["lane_utilisation"], valu_source=VALU_PROFILE,
                    # clock-independent cross-check: instructions per SIMD per
                    # GPU-active clock of the profiled launches, over 0.5
                    frac_per_clock=round(valu["valu_issue_per_simd_cycle"] / 0.5, 4))
    roof["traffic"] = traffic["hbm_bytes_per_launch"] if traffic else None
    roof.update({
        "traffic_source": TRAFFIC_PROFILE if traffic else None,
        "hbm_achieved": round(roof["traffic"] / (avg_launch_ms * 1e-3) / 1e9, 1) if traffic else None,
        "hbm_peak": HBM_PEAK_GBS,
        "algorithmic_bytes_per_launch": round(bytes_per_launch),
        "algorithmic_gbs": round(alg_gbs, 1),
        "bytes_per_ray": round(bytes_per_ray, 1),
        "nodes_per_ray": round(nodes_per_ray, 2),
        "tris_per_ray": round(tris_per_ray, 2),
        "avg_launch_ms": round(avg_launch_ms, 4),
        "launches": launches,
        "rays_per_launch": round(rays_per_launch),
        "node_fetch_rate": {"nodes_per_s": round(node_rate),
                            "hbm_fed_nodes_per_s": round(HBM_PEAK_GBS * 1e9 / NODE_BYTES)},
        "note": "traversal is bound by VALU issue (divergent node / leaf steps); most node and triangle "
                "bytes come from LDS and L2, so the algorithmic byte rate can pass HBM's peak while the "
                "measured DRAM traffic (traffic / launch time) stays well below it",
    })
    if traffic:
        roof["hbm_frac"] = round(roof["hbm_achieved"] / HBM_PEAK_GBS, 4)
    shadow = _profile(SHADOW_VALU_PROFILE)
    if shadow and ms_shadow > 0:
        # the any-hit kernel, same bound: its own counters and launch time
        ginstr = shadow["valu_insts_per_ray"] * shadow_rays / (ms_shadow * 1e-3) / 1e9
        roof["k_shadow"] = {"achieved": round(ginstr, 1), "frac": round(ginstr / VALU_PEAK_GINSTR, 4),
                            "frac_per_clock": round(shadow["valu_issue_per_simd_cycle"] / 0.5, 4),
                            "valu_insts_per_ray": shadow["valu_insts_per_ray"],
                            "lane_utilisation": shadow["lane_utilisation"], "ms_per_frame": round(ms_shadow, 3),
                            "rays_per_frame": shadow_rays, "valu_source": SHADOW_VALU_PROFILE}
    out = {
        "metric": METRIC,
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "plain path tracer, procedural interior",
            "meshes": args.meshes,
            "triangles": info.n_triangles,
            "width": args.width,
            "height": args.height,
            "spp": args.spp,
            "max_depth": args.max_depth,
            "bvh": "binned SAH (32 bins) collapsed to 4-wide nodes",
            "rays_per_frame": rays // args.steps,
            "ms_per_frame": round(ms_per_step, 3),
            "setup_s": round(setup_s, 3),
            "relmse_vs_fp64_256spp": round(relmse, 6),
            "relrmse_vs_fp64_256spp": round(relrmse, 6),
        },
        "roofline": roof,
    }
    if not args.no_cpu_baseline:
        print("bench: CPU baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(scene, aqua, args.max_depth)
    del r
    if not args.no_nrc:
        extra_legs(out)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
