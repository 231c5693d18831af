"""Image textures and display output on the GPU against the CPU restatement:

* a box room whose walls and floor carry a random non-power-of-two texture
  with texcoords far outside [0, 1] (repeat), rendered bit-exact against the
  oracle, the cache's training records included (finite films);
* the shipped part of the reference's living room (tests/golden/room_subset:
  floor, walls, lamps with emission from living_room.mtl, pictures, carpet,
  books, TV -- 8 JPEG textures), bit-exact films and primary hits;
* the GPU tonemap against a float64 reference of the sRGB encoding, and its
  PNG read back by PIL.
"""
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN, aqua, oracle_scene_of, small_camera

pytestmark = pytest.mark.gpu


def cfg(spp=4, max_depth=5):
    c = aqua.integrator_default()
    c.spp, c.max_depth = spp, max_depth
    return c


def _textured_box_room(box_room, seed=1):
    verts, mids = box_room.triangles()
    mats = box_room.materials()
    rng = np.random.default_rng(seed)
    img = rng.integers(0, 256, (23, 37, 4), dtype=np.uint8)
    img[..., 3] = 255
    ms = []
    for k, row in enumerate(mats):
        m = aqua.Material.from_buffer_copy(row.tobytes())
        if row[4:7].max() == 0 and k % 2 == 0:   # some non-emissive materials get the image
            m.base_color_tex = 1
        ms.append(m)
    sc = aqua.Scene.from_arrays(verts, mids, ms, box_room.camera(), [])
    assert sc.add_image(img) == 0
    v = verts.reshape(-1, 3, 3).astype(np.float64)
    uv = np.stack([1.7 * v[..., 0] + 0.9 * v[..., 2] - 3.2, 2.3 * v[..., 1] - 1.1 * v[..., 2] + 5.6], -1)
    sc.set_texcoords(uv.reshape(-1, 6).astype(np.float32))
    assert any(m.base_color_tex for m in ms)
    return sc


def test_textured_box_room_parity(gpu, box_room):
    sc = _textured_box_room(box_room)
    orc = oracle_scene_of(sc)
    cam = small_camera(sc.camera(), 80, 56)
    r = aqua.Renderer(sc, cfg(spp=8, max_depth=5))
    r.render(8, frame=4, camera=cam)
    img_c, _ = orc.render(aqua.camera_to_array(cam), 8, 5, frame=4)
    np.testing.assert_array_equal(r.film(cam), img_c)
    # the texture changes the image (against the untextured room)
    plain = aqua.Renderer(box_room, cfg(spp=8, max_depth=5))
    plain.render(8, frame=4, camera=cam)
    assert np.abs(plain.film(cam) - img_c).max() > 0.05
    # the cache sees textured albedos in its records and stays finite
    c = cfg(spp=2, max_depth=8)
    c.type = aqua.INTEGRATOR_NRC
    c.nrc_train_paths = 4096
    n = aqua.Renderer(sc, c)
    for f in range(3):
        st = n.render(2, frame=f, camera=cam)
        assert st.nrc_queries > 0 and math.isfinite(st.nrc_loss)
    assert np.all(np.isfinite(n.film(cam)))


def test_textured_scene_without_texcoords_is_refused(gpu, box_room):
    verts, mids = box_room.triangles()
    m = aqua.Material()
    m.base_color[:] = [0.5, 0.5, 0.5]
    m.base_color_tex = 1
    sc = aqua.Scene.from_arrays(verts, mids * 0, [m], box_room.camera(), [])
    with pytest.raises(aqua.AquaError):
        aqua.Renderer(sc, cfg())                       # names an image the scene lacks
    sc.add_image(np.full((2, 2, 4), 200, np.uint8))
    with pytest.raises(aqua.AquaError) as e:
        aqua.Renderer(sc, cfg())                       # and has no texcoords
    assert e.value.code == aqua.ERR_INVALID_ARG


@pytest.fixture(scope="module")
def room_subset():
    sc = aqua.Scene.load(os.path.join(GOLDEN, "room_subset", "scene.json"), emission_from_mtl=True)
    i = sc.info()
    assert (i.n_images, i.has_texcoords, i.n_point_lights) == (8, 1, 1) and i.n_emissive_triangles > 0
    return sc, oracle_scene_of(sc)


def test_room_subset_primary_hits(gpu, room_subset):
    sc, orc = room_subset
    cam = small_camera(sc.camera(), 320, 180)
    r = aqua.Renderer(sc, cfg())
    ids_g, t_g = r.primary_hits(camera=cam)
    ids_c, t_c = orc.primary_hits(aqua.camera_to_array(cam))
    np.testing.assert_array_equal(ids_g, ids_c)
    np.testing.assert_array_equal(t_g[ids_c != aqua.MISS], t_c[ids_c != aqua.MISS])


@pytest.mark.parametrize("fov_deg", [None, 44.0])
def test_room_subset_film_parity(gpu, room_subset, fov_deg):
    # the room's own 120-degree camera, and a narrower one on the textured
    # back wall, pictures and floor; depth 5, lamps + point light
    sc, orc = room_subset
    cam = small_camera(sc.camera(), 128, 72)
    if fov_deg:
        cam.tan_half_fov = float(np.float32(math.tan(math.radians(fov_deg / 2))))
    r = aqua.Renderer(sc, cfg(spp=8, max_depth=5))
    st = r.render(8, frame=1, camera=cam)
    img_c, ost = orc.render(aqua.camera_to_array(cam), 8, 5, frame=1)
    np.testing.assert_array_equal(r.film(cam), img_c)
    assert (st.camera_rays, st.extension_rays, st.shadow_rays) == \
        (ost.camera_rays, ost.extension_rays, ost.shadow_rays)
    assert img_c.mean() > 0.01


def test_material_sort_keeps_the_film(gpu, room_subset, box_room):
    # ordering each bounce's hits by material before shading changes which
    # wave shades a path, never what it computes: the film of the textured
    # room (16 materials, 8 images) matches the oracle with the sort on, in
    # one pass and in several overlapped ones, and the cache's first frame
    # (queries, training records) is unchanged too
    sc, orc = room_subset
    cam = small_camera(sc.camera(), 160, 90)
    r = aqua.Renderer(sc, cfg(spp=6, max_depth=6))
    r.set_opts(timing=False, sort_materials=True)
    st = r.render(6, frame=2, camera=cam)
    img_c, ost = orc.render(aqua.camera_to_array(cam), 6, 6, frame=2)
    np.testing.assert_array_equal(r.film(cam), img_c)
    assert (st.camera_rays, st.extension_rays, st.shadow_rays) == \
        (ost.camera_rays, ost.extension_rays, ost.shadow_rays)
    r.set_opts(samples_per_pass=2, timing=False, sort_materials=True)
    r.render(6, frame=2, camera=cam)
    np.testing.assert_array_equal(r.film(cam), img_c)
    c = cfg(spp=2, max_depth=6)
    c.type = aqua.INTEGRATOR_NRC
    c.nrc_train_paths = 1024
    films = []
    for sort in (False, True):
        rn = aqua.Renderer(box_room, c)
        rn.set_opts(timing=False, sort_materials=sort)
        big = small_camera(box_room.camera(), 96, 64)
        st = rn.render(2, frame=0, camera=big)
        assert st.nrc_queries > 0
        films.append(rn.film(big))
    np.testing.assert_array_equal(films[0], films[1])


def _srgb_code(x):
    x = np.clip(np.asarray(x, np.float64), 0.0, 1.0)
    e = np.where(x <= 0.0031308, 12.92 * x, 1.055 * x ** (1.0 / 2.4) - 0.055)
    return np.floor(e * 255.0 + 0.5).astype(np.int64)


@pytest.mark.parametrize("exposure", [0.0, 1.0, -2.5])
def test_tonemap_and_png(gpu, box_room, tmp_path, exposure):
    PIL = pytest.importorskip("PIL.Image")
    cam = small_camera(box_room.camera(), 97, 61)
    r = aqua.Renderer(box_room, cfg(spp=3, max_depth=5))
    r.render(3, frame=2, camera=cam)
    film = r.film(cam)
    out = r.tonemap(exposure, camera=cam)
    assert out.shape == (61, 97, 4) and np.all(out[..., 3] == 255)
    scale = np.float32(2.0 ** exposure)
    want = _srgb_code(film * scale)        # film = sum * (1 / spp), as the kernel forms it
    got = out[..., :3].astype(np.int64)
    diff = np.abs(got - want)
    # a pixel may sit within double rounding of a code boundary: one level
    assert diff.max() <= 1 and (diff > 0).mean() < 1e-4
    p = str(tmp_path / "frame.png")
    aqua.image_write_png(p, out)
    np.testing.assert_array_equal(np.asarray(PIL.open(p)), out)
