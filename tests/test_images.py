"""Image I/O without a GPU: the JPEG decoder against PIL (an independent
libjpeg-based decoder) on the reference's textures and on every variant the
format allows here (baseline / progressive, 4:4:4 / 4:2:2 / 4:2:0, grey,
restart markers, sizes off the block grid); PNG and PFM writers read back by
independent readers; and the CPU restatement's texture lookup pinned by a
closed form (a textured Lambertian floor under a point light, texel values
interpolated here in float64).

Decoder bar: within 3 levels of PIL at every texel and 0.3 levels on
average (the inverse DCTs round differently: libjpeg's is an integer one).
"""
import glob
import io
import math
import os

import numpy as np
import pytest

from conftest import REFERENCE_SCENES, aqua, oracle

PIL = pytest.importorskip("PIL.Image")


def _pil_rgba(data):
    return np.asarray(PIL.open(io.BytesIO(data)).convert("RGBA"))


def _compare(ours, theirs):
    d = np.abs(ours.astype(int) - theirs.astype(int))
    return int(d.max()), float(d.mean())


@pytest.mark.skipif(not os.path.isdir(REFERENCE_SCENES), reason="reference scenes not mounted")
def test_jpeg_decoder_matches_pil_on_the_room_textures():
    files = sorted(glob.glob(os.path.join(REFERENCE_SCENES, "textures", "*.jpg")))
    assert len(files) == 17
    for f in files:
        ours = aqua.image_load(f)
        theirs = np.asarray(PIL.open(f).convert("RGBA"))
        assert ours.shape == theirs.shape, f
        mx, mean = _compare(ours, theirs)
        assert mx <= 3 and mean < 0.3, (os.path.basename(f), mx, mean)


def _test_image(w, h, seed):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    img = np.stack([128 + 100 * np.sin(x / 7.0 + seed), 128 + 90 * np.cos(y / 5.0),
                    (x * 3 + y * 2) % 256], -1)
    img = img + rng.normal(0, 12, img.shape)
    return np.clip(img, 0, 255).astype(np.uint8)


@pytest.mark.parametrize("progressive", [False, True])
@pytest.mark.parametrize("subsampling", [0, 1, 2])          # 4:4:4, 4:2:2, 4:2:0
@pytest.mark.parametrize("size", [(67, 45), (16, 16), (1, 1), (257, 9)])
def test_jpeg_variants_match_pil(tmp_path, progressive, subsampling, size):
    img = _test_image(*size, seed=sum(size))
    b = io.BytesIO()
    PIL.fromarray(img).save(b, "JPEG", quality=90, subsampling=subsampling, progressive=progressive)
    p = tmp_path / "t.jpg"
    p.write_bytes(b.getvalue())
    ours = aqua.image_load(str(p))
    mx, mean = _compare(ours, _pil_rgba(b.getvalue()))
    assert ours.shape == (size[1], size[0], 4)
    assert mx <= 3 and mean < 0.3, (mx, mean)
    assert np.all(ours[..., 3] == 255)


@pytest.mark.parametrize("progressive", [False, True])
def test_jpeg_grey_and_restart_markers(tmp_path, progressive):
    img = _test_image(75, 41, seed=3)
    grey = PIL.fromarray(img).convert("L")
    for kw in ({}, {"restart_marker_blocks": 3}, {"restart_marker_rows": 1}):
        b = io.BytesIO()
        src = PIL.fromarray(img) if kw else grey
        src.save(b, "JPEG", quality=85, progressive=progressive, **kw)
        p = tmp_path / "t.jpg"
        p.write_bytes(b.getvalue())
        if kw:
            assert b"\xff\xdd" in b.getvalue()     # a DRI segment is present
        mx, mean = _compare(aqua.image_load(str(p)), _pil_rgba(b.getvalue()))
        assert mx <= 3 and mean < 0.3, (kw, mx, mean)


def test_jpeg_errors_are_reported(tmp_path):
    p = tmp_path / "bad.jpg"
    p.write_bytes(b"not an image at all")
    with pytest.raises(aqua.AquaError) as e:
        aqua.image_load(str(p))
    assert e.value.code == aqua.ERR_UNSUPPORTED
    b = io.BytesIO()
    PIL.fromarray(_test_image(40, 30, 1)).save(b, "JPEG")
    data = b.getvalue()
    p.write_bytes(data[:200])                      # headers only, no scan data
    with pytest.raises(aqua.AquaError) as e:
        aqua.image_load(str(p))
    assert e.value.code == aqua.ERR_PARSE
    sos = data.index(b"\xff\xda")
    p.write_bytes(data[: (sos + len(data)) // 2])  # scan cut short: decodes, garbage tail
    assert aqua.image_load(str(p)).shape == (30, 40, 4)
    with pytest.raises(aqua.AquaError) as e:
        aqua.image_load(str(tmp_path / "absent.jpg"))
    assert e.value.code == aqua.ERR_IO


@pytest.mark.parametrize("channels,size", [(3, (13, 7)), (4, (1, 1)), (3, (640, 3))])
def test_png_writer_round_trip(tmp_path, channels, size):
    rng = np.random.default_rng(channels)
    px = rng.integers(0, 256, (size[1], size[0], channels), dtype=np.uint8)
    p = str(tmp_path / "o.png")
    aqua.image_write_png(p, px)
    back = np.asarray(PIL.open(p))
    np.testing.assert_array_equal(back, px)


def test_pfm_writer_round_trip(tmp_path):
    rng = np.random.default_rng(2)
    rgb = (rng.random((5, 9, 3)) * 100).astype(np.float32)
    rgb[0, 0] = [0.0, 1e-30, 3.0e38]
    p = tmp_path / "o.pfm"
    aqua.image_write_pfm(str(p), rgb)
    data = p.read_bytes()
    head = data.split(b"\n", 3)
    assert head[0] == b"PF" and head[1] == b"9 5" and float(head[2]) < 0   # little endian
    px = np.frombuffer(head[3], "<f4").reshape(5, 9, 3)[::-1]           # stored bottom row first
    np.testing.assert_array_equal(px, rgb)


# ---------------------------------------------------- oracle texture pin ---

def _srgb_to_linear(c):
    c = np.asarray(c, np.float64) / 255.0
    return np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)


def _bilinear(img, u, v):
    """float64 reference lookup: repeat, texel centres at half-integers,
    v = 0 at the bottom row."""
    h, w = img.shape[:2]
    u, v = u - math.floor(u), v - math.floor(v)
    x, y = u * w - 0.5, (1.0 - v) * h - 0.5
    x0, y0 = math.floor(x), math.floor(y)
    fx, fy = x - x0, y - y0
    lin = _srgb_to_linear(img[..., :3])
    c = lambda xx, yy: lin[yy % h, xx % w]
    return (c(x0, y0) * (1 - fx) + c(x0 + 1, y0) * fx) * (1 - fy) + \
        (c(x0, y0 + 1) * (1 - fx) + c(x0 + 1, y0 + 1) * fx) * fy


@pytest.mark.parametrize("look_at", [(0.3, -0.2), (-1.7, 2.45), (0.0, 0.0), (3.1, 0.6)])
def test_oracle_texture_lookup_closed_form(look_at):
    # A floor y = 0 whose texcoords are (u, v) = (x / 2, -z / 3) + offsets
    # (so u and v leave [0, 1] and the texture repeats), a random 7 x 5
    # texture, a point light above: radiance = albedo(u, v) / pi * I cos / r^2
    # with albedo the float64 bilinear lookup of the decoded texels.
    rng = np.random.default_rng(9)
    img = rng.integers(0, 256, (5, 7, 4), dtype=np.uint8)
    img[..., 3] = 255
    X = 8.0
    corners = np.array([[-X, 0, -X], [-X, 0, X], [X, 0, X], [X, 0, -X]], np.float64)
    uv_of = lambda p: (p[0] / 2.0 + 0.25, -p[2] / 3.0 + 0.1)
    tris = np.float32([[*corners[0], *corners[1], *corners[2]], [*corners[0], *corners[2], *corners[3]]])
    uvs = np.float32([[*uv_of(corners[0]), *uv_of(corners[1]), *uv_of(corners[2])],
                      [*uv_of(corners[0]), *uv_of(corners[2]), *uv_of(corners[3])]])
    mat = np.array([1, 1, 1, 0, 0, 0, 0, 0.5, 0, 0, 0, 0], np.float32)
    mat[9:10] = np.uint32([1]).view(np.float32)            # image 0
    light, I = np.array([0.5, 2.0, -0.3]), 3.0
    sc = oracle.Scene(tris, np.zeros(2, np.uint32), mat[None], np.float32([[*light, I, I, I]]), uvs=uvs,
                      images=[img])
    c = aqua.Camera()
    x, z = look_at
    c.cam_to_world[:] = [1, 0, 0, x, 0, 0, 1, 3, 0, -1, 0, z]
    c.tan_half_fov = float(np.float32(1e-6))
    c.aspect, c.width, c.height = 1.0, 2, 2
    out, _ = sc.render(aqua.camera_to_array(c), 4, 1)
    p = np.array([x, 0.0, z])
    dv = light - p
    r2 = float(dv @ dv)
    expect = _bilinear(img, *uv_of(p)) / math.pi * I * (dv[1] / math.sqrt(r2)) / r2
    np.testing.assert_allclose(out.reshape(-1, 3).astype(np.float64), np.tile(expect, (4, 1)), rtol=2e-4)
