"""Scene / integrator / .mesh loading: the reference's data formats.

KATs use the shipped Cornell box (tests/golden/cbox: scenes/cbox.json, the
eight CornellBox-Original_*.mesh, CornellBox-Original.obj/.mtl and
scenes/integrator.json, copied as fixtures).
"""
import hashlib
import json
import math
import os
import shutil
import struct

import numpy as np
import pytest

from conftest import GOLDEN, REFERENCE_SCENES, aqua

CBOX = os.path.join(GOLDEN, "cbox")


# ---- an independent minimal BSON reader (bsonspec.org 1.1) for the KATs ----

def _bson_doc(b, o):
    n = struct.unpack_from("<i", b, o)[0]
    end, o, out = o + n, o + 4, {}
    while o < end - 1:
        t = b[o]
        k_end = b.index(0, o + 1)
        k = b[o + 1:k_end].decode()
        o = k_end + 1
        if t == 0x01:
            v = struct.unpack_from("<d", b, o)[0]
            o += 8
        elif t == 0x12:
            v = struct.unpack_from("<q", b, o)[0]
            o += 8
        elif t == 0x10:
            v = struct.unpack_from("<i", b, o)[0]
            o += 4
        elif t == 0x02:
            sl = struct.unpack_from("<i", b, o)[0]
            v = b[o + 4:o + 3 + sl].decode()
            o += 4 + sl
        elif t in (0x03, 0x04):
            v, o2 = _bson_doc(b, o)
            if t == 0x04:
                v = [v[str(i)] for i in range(len(v))]
            o = o2
        else:
            raise ValueError(hex(t))
        out[k] = v
    return out, end


def read_mesh(path):
    b = open(path, "rb").read()
    d, end = _bson_doc(b, 0)
    assert end == len(b)
    return d


def _obj_vertices(path):
    vs = []
    for line in open(path):
        p = line.split()
        if p and p[0] == "v":
            vs.append([float(x) for x in p[1:4]])
    return np.array(vs)


# --------------------------------------------------------------- KATs ------

def test_srgb_decode_matches_mtl_kd():
    # cbox.json stores Srgb colours; IEC 61966-2-1 decoding gives the MTL Kd.
    sc = aqua.Scene.load(os.path.join(CBOX, "cbox.json"))
    names = list(json.load(open(os.path.join(CBOX, "cbox.json")))["named_bsdfs"])
    mats = sc.materials()
    want = {"shortBox": (0.725, 0.71, 0.68), "leftWall": (0.63, 0.065, 0.05),
            "rightWall": (0.14, 0.45, 0.091), "light": (0.78, 0.78, 0.78)}
    for name, kd in want.items():
        np.testing.assert_allclose(mats[names.index(name), :3], kd, atol=2e-7)


def test_cbox_meshes_match_bson_and_obj():
    sc = aqua.Scene.load(os.path.join(CBOX, "cbox.json"))
    info = sc.info()
    assert (info.n_triangles, info.n_shapes, info.n_point_lights) == (36, 8, 1)
    verts, mat_ids = sc.triangles()
    doc = json.load(open(os.path.join(CBOX, "cbox.json")))
    names = list(doc["named_bsdfs"])
    counts = []
    row = 0
    for shape in doc["shapes"]:
        fname, named = shape["Mesh"]
        m = read_mesh(os.path.join(CBOX, fname))
        V = np.array(m["vertices"], np.float64)
        I = np.array(m["indices"], np.int64)
        assert np.all(V.astype(np.float32).astype(np.float64) == V)   # FP32-exact widening
        assert len(m["normals"]) == len(m["vertices"])
        got = verts[row:row + len(I)].reshape(-1, 3, 3)
        np.testing.assert_array_equal(got, V[I].astype(np.float32))
        assert np.all(mat_ids[row:row + len(I)] == names.index(named["Named"]))
        counts.append(len(I))
        row += len(I)
    # floor, ceiling, backWall, rightWall, leftWall, shortBox, tallBox, light
    assert counts == [2, 2, 2, 2, 2, 12, 12, 2]
    # every mesh vertex is an OBJ `v` line (scenes/CornellBox-Original.obj:15-168)
    obj = _obj_vertices(os.path.join(CBOX, "CornellBox-Original.obj"))
    for p in verts.reshape(-1, 3):
        assert np.min(np.abs(obj - p).max(axis=1)) < 1e-6


def test_cbox_camera_and_point_light():
    sc = aqua.Scene.load(os.path.join(CBOX, "cbox.json"))
    c = sc.camera()
    assert (c.width, c.height) == (512, 512)
    np.testing.assert_allclose(list(c.cam_to_world), [1, 0, 0, 0, 0, 1, 0, 1, 0, 0, 1, 9], atol=0)
    assert c.tan_half_fov == np.float32(math.tan(math.radians(7.5)))
    pl = sc.point_lights()
    np.testing.assert_allclose(pl, [[0.0, 1.7, 0.1, 1.0, 1.0, 1.0]], atol=1e-7)


def test_camera_rotation_convention(tmp_path):
    doc = json.load(open(os.path.join(CBOX, "cbox.json")))
    doc["camera"]["Perspective"]["transform"]["rotate"] = [0.0, 90.0, 0.0]
    d = tmp_path / "s"
    shutil.copytree(CBOX, d)
    json.dump(doc, open(d / "cbox.json", "w"))
    c = aqua.Scene.load(str(d / "cbox.json")).camera()
    m = np.array(c.cam_to_world).reshape(3, 4)
    # forward (-Z column) turns to -X under +90 degrees about y
    np.testing.assert_allclose(-m[:, 2], [-1, 0, 0], atol=1e-7)


def test_integrator_json():
    c = aqua.integrator_load(os.path.join(CBOX, "integrator.json"))
    assert (c.type, c.spp, c.batch_size, c.max_depth, c.training_iters, c.visualize_cache) == \
        (aqua.INTEGRATOR_NRC, 4, 512, 5, 2048, 0)
    assert c.learning_rate == np.float32(0.001)
    assert (c.seed, c.rr_depth) == (0, 3)


def test_integrator_pt_and_unknown_keys(tmp_path):
    p = tmp_path / "i.json"
    p.write_text('{"type": "pt", "spp": 16, "max_depth": 8, "future_key": [1, 2], "seed": 7}')
    c = aqua.integrator_load(str(p))
    assert (c.type, c.spp, c.max_depth, c.seed) == (aqua.INTEGRATOR_PT, 16, 8, 7)
    p.write_text('{"type": "bdpt"}')
    with pytest.raises(aqua.AquaError) as e:
        aqua.integrator_load(str(p))
    assert e.value.code == aqua.ERR_PARSE


# ------------------------------------------------------------ failures -----

def _copy_cbox(tmp_path):
    d = tmp_path / "cb"
    shutil.copytree(CBOX, d)
    return d


def _expect(code, path, **kw):
    with pytest.raises(aqua.AquaError) as e:
        aqua.Scene.load(str(path), **kw)
    assert e.value.code == code, str(e.value)


def test_missing_mesh_and_skip(tmp_path):
    d = _copy_cbox(tmp_path)
    os.remove(d / "CornellBox-Original_tallBox_6.mesh")
    _expect(aqua.ERR_MISSING_MESH, d / "cbox.json")
    sc = aqua.Scene.load(str(d / "cbox.json"), skip_missing=True)
    assert (sc.info().n_triangles, sc.info().n_skipped_shapes) == (24, 1)


def test_truncated_and_corrupt_bson(tmp_path):
    d = _copy_cbox(tmp_path)
    f = d / "CornellBox-Original_floor_0.mesh"
    b = f.read_bytes()
    f.write_bytes(b[:200])
    _expect(aqua.ERR_PARSE, d / "cbox.json")
    # index out of range: last int64 of the file is an index
    bad = bytearray(b)
    struct.pack_into("<q", bad, len(bad) - 11, 99)
    f.write_bytes(bytes(bad))
    _expect(aqua.ERR_PARSE, d / "cbox.json")


@pytest.mark.parametrize("mutate,code", [
    (lambda j: j["named_bsdfs"]["floor"]["Principled"].update(color={"Gradient": 1}), aqua.ERR_PARSE),
    (lambda j: j["lights"].append({"Spot": {}}), aqua.ERR_UNSUPPORTED),
    (lambda j: j["named_bsdfs"]["floor"]["Principled"].update(transmission={"Float": 1.0}),
     aqua.ERR_UNSUPPORTED),
    (lambda j: j["shapes"][0]["Mesh"][1].update(Named="nope"), aqua.ERR_PARSE),
    (lambda j: j["camera"]["Perspective"].update(lens_radius=0.1), aqua.ERR_UNSUPPORTED),
])
def test_schema_errors(tmp_path, mutate, code):
    d = _copy_cbox(tmp_path)
    j = json.load(open(d / "cbox.json"))
    mutate(j)
    json.dump(j, open(d / "cbox.json", "w"))
    _expect(code, d / "cbox.json")


def test_malformed_json_and_missing_file(tmp_path):
    p = tmp_path / "x.json"
    p.write_text('{"named_bsdfs": {,}')
    _expect(aqua.ERR_PARSE, p)
    _expect(aqua.ERR_IO, tmp_path / "absent.json")


@pytest.mark.skipif(not os.path.isdir(REFERENCE_SCENES), reason="reference scenes not mounted")
def test_room_json_loads_with_skip_missing():
    # room.json:3021 names the mesh the snapshot withholds (.MISSING_LARGE_BLOBS:1)
    path = os.path.join(REFERENCE_SCENES, "room.json")
    _expect(aqua.ERR_MISSING_MESH, path)
    _expect(aqua.ERR_MISSING_MESH, path, image_as_constant=True)
    sc = aqua.Scene.load(path, skip_missing=True, image_as_constant=True)
    i = sc.info()
    assert (i.n_shapes, i.n_skipped_shapes, i.n_image_textures, i.n_images) == (152, 1, 16, 0)
    assert i.n_triangles == 394269


@pytest.mark.skipif(not os.path.isdir(REFERENCE_SCENES), reason="reference scenes not mounted")
def test_room_json_as_shipped_textures_and_lamps():
    # The 16 Image colours (room.json:6, 194, 382, ...) name 16 distinct JPEGs,
    # decoded with the texcoords of every mesh; with emission_from_mtl the
    # lamp quads (black, emission 0 in the JSON: room.json:251-258) take Ke
    # from living_room.mtl (BackLight 2 2 2 at :11-13; Left / Middle / Right
    # Light 5 5 5 at :167-168, :193-194, :287-288); the point light stays
    # (room.json:2656-2672).
    path = os.path.join(REFERENCE_SCENES, "room.json")
    sc = aqua.Scene.load(path, skip_missing=True, emission_from_mtl=True)
    i = sc.info()
    assert (i.n_image_textures, i.n_images, i.has_texcoords, i.n_point_lights) == (16, 16, 1, 1)
    mats = sc.materials()
    names = json.load(open(path))["named_bsdfs"]
    by = dict(zip(names.keys(), mats))
    for lamp, ke in (("BackLight", 2.0), ("LeftLight", 5.0), ("MiddleLight", 5.0), ("RightLight", 5.0)):
        np.testing.assert_array_equal(by[lamp][4:7], [ke, ke, ke])
    assert i.n_emissive_triangles > 0
    tex = mats[:, 9].view(np.uint32)
    assert (tex > 0).sum() == 16 and sorted(tex[tex > 0]) == list(range(1, 17))
    assert np.all(mats[tex == 0][:, 4:7].max(1) >= 0)
    # without the option the lamps stay dark, as written
    plain = aqua.Scene.load(path, skip_missing=True).materials()
    assert plain[:, 4:7].max() == 0.0
    uv = sc.texcoords()
    assert uv.shape == (i.n_triangles, 6) and np.all(np.isfinite(uv))
    # a decoded texture is the JPEG itself
    wood = [k for k, v in names.items() if v["Principled"]["color"].get("Image", "").endswith("wood4.jpg")][0]
    img = sc.image(int(by[wood][9:10].view(np.uint32)[0]) - 1)
    np.testing.assert_array_equal(img, aqua.image_load(os.path.join(REFERENCE_SCENES, "textures", "wood4.jpg")))


# ------------------------------------------------------------- procgen -----

def _digest(d):
    h = hashlib.sha256()
    for f in sorted(os.listdir(d)):
        h.update(f.encode())
        h.update(open(os.path.join(d, f), "rb").read())
    return h.hexdigest()


def test_box_room_is_deterministic_and_shaped(tmp_path):
    a, b = tmp_path / "a", tmp_path / "b"
    aqua.procgen(aqua.PROCGEN_BOX_ROOM, str(a))
    aqua.procgen(aqua.PROCGEN_BOX_ROOM, str(b))
    assert _digest(a) == _digest(b)
    # the committed digest of the emitted files (names + bytes, sorted by
    # name): the same scene on every machine and every build
    assert _digest(a) == "72cf38d4d185f4ea1a677e1fa2fc15d7b067c1dd2a252ef05401a0e7cf6c74e7"
    sc = aqua.Scene.load(str(a / "scene.json"))
    i = sc.info()
    assert (i.n_triangles, i.n_emissive_triangles, i.n_shapes) == (36, 2, 8)
    np.testing.assert_allclose(list(i.bounds_lo), [-1, 0, -1], atol=1e-7)
    np.testing.assert_allclose(list(i.bounds_hi), [1, 2, 1], atol=1e-7)
    c = sc.camera()
    assert (c.width, c.height) == (512, 512)
    assert c.tan_half_fov == np.float32(math.tan(math.radians(20)))
    mats = sc.materials()
    assert np.any(np.all(mats[:, 4:7] == np.float32([17, 12, 4]), axis=1))
    # the light faces down
    v, m = sc.triangles()
    em = mats[m, 4] > 0
    n = np.cross(v[em, 3:6] - v[em, 0:3], v[em, 6:9] - v[em, 0:3])
    assert np.all(n[:, 1] < 0)
    # walls (10 tris), short box (12), tall box (12), light (2); boxes 0.6 and
    # 1.2 tall, their side faces turned 17 degrees about y
    short, tall = v[10:22].reshape(-1, 3), v[22:34].reshape(-1, 3)
    assert np.isclose(short[:, 1].max(), 0.6) and np.isclose(tall[:, 1].max(), 1.2)
    for pts in (short, tall):
        base = pts[pts[:, 1] == 0.0]
        e = base[None, :, [0, 2]] - base[:, None, [0, 2]]
        ang = np.degrees(np.arctan2(e[..., 1], e[..., 0]))[np.hypot(e[..., 0], e[..., 1]) > 0.5]
        assert np.any(np.isclose(np.abs(ang) % 90, 17.0, atol=1e-3) | np.isclose(np.abs(ang) % 90, 73.0, atol=1e-3))


def test_interior_generator(tmp_path):
    p = aqua.procgen(aqua.PROCGEN_INTERIOR, str(tmp_path), seed=1, n_meshes=20, target_tris=50000,
                     width=64, height=32)
    sc = aqua.Scene.load(p)
    i = sc.info()
    assert i.n_shapes == 1 + 3 + 20 and i.n_emissive_triangles == 6
    assert 25000 < i.n_triangles < 100000
    mats = sc.materials()[4:]
    assert np.all((mats[:, 7] >= 0.05) & (mats[:, 7] <= 1.0))
    assert set(np.unique(mats[:, 3])) <= {0.0, 1.0}
    assert np.all((mats[:, :3] >= 0.1) & (mats[:, :3] <= 0.9))


def _check_bvh4(nodes, order, verts):
    """Every triangle in exactly one leaf; every leaf triangle and every
    child node inside its parent's box; nodes in preorder; the leaf children
    of a node take one run of slots, in child order."""
    ntri = len(verts)
    assert sorted(order.tolist()) == list(range(ntri))
    idx = nodes[:, 24:28].view(np.int32)
    cnt = nodes[:, 28:32].view(np.int32)
    covered = np.zeros(ntri, int)
    for n in range(len(nodes)):
        leaves = [(idx[n, k], cnt[n, k]) for k in range(4) if idx[n, k] >= 0 and cnt[n, k] > 0]
        for (s0, c0), (s1, _) in zip(leaves, leaves[1:]):
            assert s1 == s0 + c0
        for k in range(4):
            if idx[n, k] < 0:
                continue
            lo = nodes[n, [0 + k, 8 + k, 16 + k]]
            hi = nodes[n, [4 + k, 12 + k, 20 + k]]
            if cnt[n, k] > 0:
                for s in range(idx[n, k], idx[n, k] + cnt[n, k]):
                    covered[s] += 1
                    p = verts[order[s]].reshape(3, 3)
                    assert np.all(p >= lo) and np.all(p <= hi)
            else:
                child = idx[n, k]
                assert child > n   # preorder
                used = idx[child] >= 0
                clo = np.stack([nodes[child, 0:4], nodes[child, 8:12], nodes[child, 16:20]], 1)[used]
                chi = np.stack([nodes[child, 4:8], nodes[child, 12:16], nodes[child, 20:24]], 1)[used]
                assert np.all(clo >= lo) and np.all(chi <= hi)
    assert np.all(covered == 1)
    # the top three levels come first, breadth-first (the kernels stage a
    # prefix of the node array in LDS)
    depth = {0: 0}
    order = [0]
    for n in order:
        for k in range(4):
            if idx[n, k] >= 0 and cnt[n, k] == 0:
                depth[int(idx[n, k])] = depth[n] + 1
                order.append(int(idx[n, k]))
    top = sorted(v for v, d in depth.items() if d <= 3)
    assert top == list(range(len(top)))
    by_bfs = [v for v in order if depth[v] <= 3]
    assert by_bfs == sorted(by_bfs)


def test_scene_bvh_invariants(box_room, tmp_path):
    verts, _ = box_room.triangles()
    nodes, order, depth = box_room.bvh()
    _check_bvh4(nodes, order, verts)
    p = aqua.procgen(aqua.PROCGEN_INTERIOR, str(tmp_path), seed=2, n_meshes=6, target_tris=4000,
                     width=8, height=8)
    sc = aqua.Scene.load(p)
    verts, _ = sc.triangles()
    nodes, order, depth = sc.bvh()
    _check_bvh4(nodes, order, verts)
    # 4-wide: far fewer nodes than triangles, and the tree stays shallow
    assert len(nodes) < len(verts) / 2 and depth < 20


def _decode_device_nodes(words):
    """Float64 decode of the 64-byte device nodes: (lo[n,4,3], hi[n,4,3], ref[n,4])."""
    origin = words[:, 0:3].view(np.float32).astype(np.float64)
    e = _device_exponents(words)
    scale = np.ldexp(1.0, e)                                                     # [n, 3]
    q = np.stack([(words[:, 4:10] >> (8 * k)) & 0xFF for k in range(4)], 1)     # [n, 4, 6]
    lo = origin[:, None, :] + q[:, :, 0::2] * scale[:, None, :]
    hi = origin[:, None, :] + q[:, :, 1::2] * scale[:, None, :]
    return lo, hi, words[:, 10:14]


def _device_exponents(words):
    """Word 3: signed exponent bytes (two's complement), x in bits 0-7."""
    return np.stack([((words[:, 3] >> (8 * a)) & 0xFF).astype(np.uint8).view(np.int8) for a in range(3)],
                    1).astype(np.int64)


def test_device_nodes_enclose_the_exact_boxes(box_room, tmp_path):
    # the quantised device nodes: same children and references, and every
    # decoded child box encloses the exact one (planes rounded outward), so
    # culling against them stays conservative
    p = aqua.procgen(aqua.PROCGEN_INTERIOR, str(tmp_path), seed=4, n_meshes=12, target_tris=20000,
                     width=8, height=8)
    for sc in (box_room, aqua.Scene.load(p)):
        nodes, _, _ = sc.bvh()
        words = sc.bvh_device()
        assert words.shape == (len(nodes), 16)
        lo, hi, ref = _decode_device_nodes(words)
        idx = nodes[:, 24:28].view(np.int32)
        cnt = nodes[:, 28:32].view(np.int32)
        used = idx >= 0
        flag = np.uint32(0x80000000)
        # an inner child is its node index; an empty child is a leaf of no
        # triangles (the leaf flag alone)
        inner = used & (cnt == 0)
        np.testing.assert_array_equal(ref[inner], idx[inner].astype(np.uint32))
        np.testing.assert_array_equal(ref[~used], flag)
        assert np.all(ref[inner] < flag)
        # a leaf child is the flag plus its slots as a mask over the node's
        # leaf run, which starts at word 14 (word 15 is 0)
        leaf = used & (cnt > 0)
        run = np.where(leaf, idx, np.iinfo(np.int32).max).min(1)
        has = leaf.any(1)
        np.testing.assert_array_equal(words[has, 14], run[has].astype(np.uint32))
        np.testing.assert_array_equal(words[~has, 14], 0)
        np.testing.assert_array_equal(words[:, 15], 0)
        off = np.where(leaf, idx - run[:, None], 0).astype(np.int64)
        want = flag | (((1 << cnt.astype(np.int64)) - 1) << off).astype(np.uint32)
        np.testing.assert_array_equal(ref[leaf], want[leaf])
        # the masks of one node's leaves are disjoint and cover its run
        masks = np.where(leaf, ref & ~flag, 0)
        np.testing.assert_array_equal(np.bitwise_or.reduce(masks, 1),
                                      ((1 << np.where(leaf, cnt, 0).sum(1).astype(np.int64)) - 1).astype(np.uint32))
        assert np.all(np.sum([bin(int(v)).count("1") for v in masks.ravel()]) == np.where(leaf, cnt, 0).sum())
        exact_lo = np.stack([nodes[:, 0:4], nodes[:, 8:12], nodes[:, 16:20]], 2).astype(np.float64)
        exact_hi = np.stack([nodes[:, 4:8], nodes[:, 12:16], nodes[:, 20:24]], 2).astype(np.float64)
        assert np.all(lo[used] <= exact_lo[used]) and np.all(hi[used] >= exact_hi[used])
        # and tight: at most one quantisation step (extent / 255) of slack
        e = _device_exponents(words)
        step = np.broadcast_to(np.ldexp(1.0, e)[:, None, :], lo.shape)
        assert np.all((exact_lo - lo)[used] <= step[used]) and np.all((hi - exact_hi)[used] <= step[used])
