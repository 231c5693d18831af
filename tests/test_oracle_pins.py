"""Pins of the CPU restatement (oracle/) against published vectors, closed
forms, invariants and brute force -- checks that fail on a dropped term, a
wrong sign or index, or a transposed operand in the oracle itself.
"""
import math

import numpy as np
import pytest

from conftest import aqua, oracle

ULP32 = np.float32(2.0 ** -23)


# ---------------------------------------------------------------- PCG32 ----

def test_pcg32_published_vector():
    # pcg32-demo (pcg-c-basic) with initstate 42, initseq 54.
    got = oracle.pcg32(42, 54, 6)
    want = [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]
    assert [int(x) for x in got] == want


@pytest.mark.parametrize("n", [0, 1, 2, 3, 17, 1000, 65537, 99991])
def test_pcg32_advance_equals_repeated_next(n):
    seq = oracle.pcg32(7, 1234567, n + 4)
    jumped = oracle.pcg32(7, 1234567, 4, advance=n)
    np.testing.assert_array_equal(jumped, seq[n:])


# -------------------------------------------------------------- sincos ----

def test_sincos_2pi_against_float64():
    u = np.linspace(0.0, 1.0, 1 << 20, endpoint=False).astype(np.float32)
    u = np.concatenate([u, np.float32([0.125, 0.25, 0.375, 0.5, 0.75, np.nextafter(np.float32(1), 0)])])
    s, c = oracle.sincos_2pi(u)
    ang = 2.0 * np.pi * u.astype(np.float64)
    for got, ref in ((s, np.sin(ang)), (c, np.cos(ang))):
        tol = 2.0 * np.maximum(np.abs(ref), 2.0 ** -24) * 2.0 ** -23
        err = np.abs(got.astype(np.float64) - ref)
        assert np.all(err <= tol), float((err / tol).max())


# -------------------------------------------------------------- camera ----

def _cam16(origin, tan_half, w, h):
    c = aqua.Camera()
    m = [1, 0, 0, origin[0], 0, 1, 0, origin[1], 0, 0, 1, origin[2]]
    c.cam_to_world[:] = [float(x) for x in m]
    c.tan_half_fov = tan_half
    c.aspect = np.float32(w) / np.float32(h)
    c.width, c.height = w, h
    return aqua.camera_to_array(c)


def test_camera_centre_and_corner_rays():
    t = float(np.float32(math.tan(math.radians(20.0))))
    cam = _cam16((0.0, 1.0, 3.8), t, 512, 256)
    r = oracle.camera_ray(cam, 256, 128, 0.0, 0.0)
    np.testing.assert_allclose(r[:3], [0.0, 1.0, 3.8], rtol=0, atol=1e-7)
    np.testing.assert_allclose(r[3:], [0.0, 0.0, -1.0], atol=1e-7)
    r = oracle.camera_ray(cam, 0, 0, 0.0, 0.0)   # top-left corner of the image
    d = np.array([-t * 2.0, t, -1.0])
    d /= np.linalg.norm(d)
    np.testing.assert_allclose(r[3:], d, atol=3e-7)
    r = oracle.camera_ray(cam, 511, 255, 0.999999, 0.999999)  # bottom-right
    d = np.array([t * 2.0, -t, -1.0])
    d /= np.linalg.norm(d)
    np.testing.assert_allclose(r[3:], d, atol=3e-6)


# ------------------------------------------------------------ traversal ----

def _soup(rng, n):
    c = rng.uniform(-5, 5, (n, 1, 3))
    return (c + rng.normal(0, 0.6, (n, 3, 3))).astype(np.float32).reshape(n, 9)


def _mat12(base=(0.5, 0.5, 0.5), metallic=0.0, roughness=0.5, specular=0.0, emission=(0, 0, 0)):
    return np.array([*base, metallic, *emission, roughness, specular, 0, 0, 0], np.float32)


def _rays(o, d, tmax=3.0e38):
    r = np.zeros((len(o), 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = tmax
    r[:, 4:7] = d
    return r


def test_bvh_closest_and_any_hit_equal_brute_force():
    rng = np.random.default_rng(11)
    tris = _soup(rng, 3000)
    sc = oracle.Scene(tris, np.zeros(len(tris), np.uint32), _mat12()[None])
    o = rng.uniform(-7, 7, (20000, 3))
    d = rng.standard_normal((20000, 3))
    d[:2000, 1:] = 0.0      # axis-aligned directions
    d[2000:3000, 0] = 0.0
    rays = _rays(o, d)
    ids_b, t_b = sc.trace_closest(rays, brute=True)
    ids, t = sc.trace_closest(rays)
    np.testing.assert_array_equal(ids, ids_b)
    np.testing.assert_array_equal(t, t_b)
    assert (ids != aqua.MISS).mean() > 0.2
    short = rays.copy()
    short[:, 3] = np.where(ids_b != aqua.MISS, t_b, 3.0e38)
    np.testing.assert_array_equal(sc.trace_occluded(short), sc.trace_occluded(short, brute=True))
    # tmax == the closest t: the hit itself is excluded (t < tmax), so any
    # other blocker must lie strictly in front -- there is none.
    assert not sc.trace_occluded(short[ids_b != aqua.MISS], brute=True).any()


def _moller_trumbore(o, d, tris):
    """float64 closest hit over all triangles (Moller & Trumbore 1997)."""
    v0, v1, v2 = tris[:, 0:3], tris[:, 3:6], tris[:, 6:9]
    e1, e2 = v1 - v0, v2 - v0
    p = np.cross(d[:, None, :], e2[None])
    det = np.einsum("nj,rnj->rn", e1, p)
    inv = 1.0 / np.where(np.abs(det) < 1e-300, np.nan, det)
    s = o[:, None, :] - v0[None]
    u = np.einsum("rnj,rnj->rn", s, p) * inv
    q = np.cross(s, e1[None])
    vv = np.einsum("rj,rnj->rn", d, q) * inv
    t = np.einsum("nj,rnj->rn", e2, q) * inv
    ok = (u >= 0) & (vv >= 0) & (u + vv <= 1) & (t > 0)
    margin = np.minimum(np.minimum(u, vv), 1 - u - vv)
    t = np.where(ok, t, np.inf)
    best = np.argmin(t, axis=1)
    tb = t[np.arange(len(o)), best]
    mb = margin[np.arange(len(o)), best]
    t_sorted = np.sort(t, axis=1)
    gap = t_sorted[:, 1] - t_sorted[:, 0] if t.shape[1] > 1 else np.full(len(o), np.inf)
    return np.where(np.isfinite(tb), best, -1), tb, mb, gap


def test_watertight_test_agrees_with_float64_moller_trumbore():
    rng = np.random.default_rng(5)
    tris = _soup(rng, 300).astype(np.float64)
    o = rng.uniform(-7, 7, (4000, 3))
    d = rng.standard_normal((4000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    sc = oracle.Scene(tris.astype(np.float32), np.zeros(300, np.uint32), _mat12()[None])
    ids, t = sc.trace_closest(_rays(o, d), brute=True)
    mt_id, mt_t, margin, gap = _moller_trumbore(o.astype(np.float32).astype(np.float64),
                                                d.astype(np.float32).astype(np.float64),
                                                tris.astype(np.float32).astype(np.float64))
    clear = (mt_id < 0) | ((margin > 1e-4) & (gap > 1e-4 * np.abs(mt_t)))
    hit = clear & (mt_id >= 0)
    assert clear.mean() > 0.95
    np.testing.assert_array_equal(ids[clear & (mt_id < 0)], aqua.MISS)
    np.testing.assert_array_equal(ids[hit].astype(np.int64), mt_id[hit])
    np.testing.assert_allclose(t[hit], mt_t[hit], rtol=2e-5)


def _closed_cube(n):
    """Closed cube [-1,1]^3 of 12 n^2 triangles whose shared edges use
    identical vertex values."""
    g = np.linspace(-1.0, 1.0, n + 1).astype(np.float32)
    tris = []
    for axis in range(3):
        for s in (-1.0, 1.0):
            for j in range(n):
                for i in range(n):
                    def P(a, b):
                        p = [0.0, 0.0, 0.0]
                        p[axis] = s
                        p[(axis + 1) % 3] = g[a]
                        p[(axis + 2) % 3] = g[b]
                        return p
                    a, b, c, d = P(i, j), P(i + 1, j), P(i + 1, j + 1), P(i, j + 1)
                    tris.append(a + b + c)
                    tris.append(a + c + d)
    return np.array(tris, np.float32)


def test_watertight_no_ray_escapes_through_shared_edges_or_vertices():
    tris = _closed_cube(6)
    sc = oracle.Scene(tris, np.zeros(len(tris), np.uint32), _mat12()[None])
    verts = tris.reshape(-1, 3)
    mids = 0.5 * (tris.reshape(-1, 3, 3) + np.roll(tris.reshape(-1, 3, 3), 1, axis=1)).reshape(-1, 3)
    targets = np.unique(np.concatenate([verts, mids]), axis=0)
    rng = np.random.default_rng(3)
    origins = rng.uniform(-0.9, 0.9, (8, 3)).astype(np.float32)
    o = np.repeat(origins, len(targets), axis=0)
    d = np.tile(targets, (len(origins), 1)) - o
    rays = _rays(o, d)
    for brute in (True, False):
        ids, _ = sc.trace_closest(rays, brute=brute)
        assert int((ids == aqua.MISS).sum()) == 0


# ---------------------------------------------------------------- BSDF -----

def _dirs(rng, n):
    v = rng.standard_normal((n, 3))
    v[:, 2] = np.abs(v[:, 2]) + 1e-3
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)


def test_lambert_is_rho_over_pi():
    rng = np.random.default_rng(2)
    m = _mat12(base=(0.8, 0.4, 0.1), metallic=0.0, specular=0.0)
    n = np.float32([0, 0, 1])
    for wo, wi in zip(_dirs(rng, 200), _dirs(rng, 200)):
        f = oracle.bsdf_eval(m, n, wo, wi)
        np.testing.assert_allclose(f[:3], np.array([0.8, 0.4, 0.1]) / np.pi, rtol=3e-7)
        np.testing.assert_allclose(f[3], wi[2] / np.pi, rtol=3e-7)


@pytest.mark.parametrize("rough", [0.1, 0.3, 0.6, 1.0])
def test_ggx_d_normalisation(rough):
    # 2 pi int_0^1 D(mu) mu dmu = 1 (projected area of the microsurface).
    a = max(rough * rough, 1e-3)
    x = np.concatenate([[0.0], np.logspace(-14, 0, 6000)])   # x = 1 - mu
    mu = 1.0 - x
    D = np.array([oracle.ggx(rough, float(m), 0.5)[0] for m in mu.astype(np.float32)], np.float64)
    val = 2 * np.pi * np.trapezoid(D * mu, -mu)
    assert abs(val - 1.0) < 2e-3, (rough, a, val)


@pytest.mark.parametrize("rough,theta_o", [(0.3, 0.3), (0.6, 1.0), (1.0, 1.3), (0.45, 0.8)])
def test_visible_normals_normalisation(rough, theta_o):
    # int G1(wo) max(0, wo.h) D(h) / cos_o dh = 1 (Heitz 2014, eq. 18): ties
    # Lambda to D.
    th = np.linspace(0, np.pi / 2, 800)
    ph = np.linspace(0, 2 * np.pi, 800, endpoint=False)
    T, Pp = np.meshgrid(th, ph, indexing="ij")
    h = np.stack([np.sin(T) * np.cos(Pp), np.sin(T) * np.sin(Pp), np.cos(T)], -1)
    wo = np.array([np.sin(theta_o), 0.0, np.cos(theta_o)])
    D = np.array([oracle.ggx(rough, float(c), 0.5)[0] for c in np.cos(th).astype(np.float32)])[:, None]
    lam = oracle.ggx(rough, 1.0, float(np.float32(np.cos(theta_o))))[1]
    g1 = 1.0 / (1.0 + lam)
    integrand = np.maximum(0.0, h @ wo) * D * np.sin(T)
    val = g1 / np.cos(theta_o) * np.trapezoid(np.trapezoid(integrand, ph, axis=1), th)
    assert abs(val - 1.0) < 5e-3, val


MATS = [
    _mat12(base=(0.9, 0.6, 0.3), metallic=0.0, roughness=0.5, specular=0.5),
    _mat12(base=(0.9, 0.6, 0.3), metallic=1.0, roughness=0.3),
    _mat12(base=(0.2, 0.7, 0.5), metallic=0.0, roughness=0.9, specular=0.5),
]


@pytest.mark.parametrize("mi", range(len(MATS)))
def test_bsdf_reciprocity(mi):
    rng = np.random.default_rng(mi)
    n = np.float32([0, 0, 1])
    for wo, wi in zip(_dirs(rng, 300), _dirs(rng, 300)):
        a = oracle.bsdf_eval(MATS[mi], n, wo, wi)[:3]
        b = oracle.bsdf_eval(MATS[mi], n, wi, wo)[:3]
        np.testing.assert_allclose(a, b, rtol=2e-5, atol=1e-9)


def _hemisphere_grid(nt=300, npp=300):
    th = (np.arange(nt) + 0.5) / nt * (np.pi / 2)
    ph = (np.arange(npp) + 0.5) / npp * 2 * np.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    w = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    dw = (np.pi / 2 / nt) * (2 * np.pi / npp) * np.sin(T).reshape(-1)
    return w.astype(np.float32), dw


@pytest.mark.parametrize("mi", range(len(MATS)))
def test_bsdf_energy_and_pdf_mass(mi):
    w, dw = _hemisphere_grid(160, 160)
    n = np.float32([0, 0, 1])
    wo = np.float32([np.sin(0.6), 0.0, np.cos(0.6)])
    out = np.array([oracle.bsdf_eval(MATS[mi], n, wo, wi) for wi in w])
    albedo = (out[:, :3] * out[:, 4:5] * dw[:, None]).sum(0)
    mass = (out[:, 3] * dw).sum()
    assert np.all(albedo <= 1.0 + 1e-2), albedo
    assert 0.9 < mass <= 1.0 + 1e-2, mass


@pytest.mark.parametrize("mi", range(len(MATS)))
def test_bsdf_sampling_matches_pdf_chi2(mi):
    rng = np.random.default_rng(100 + mi)
    n = np.float32([0, 0, 1])
    wo = np.float32([np.sin(0.5), 0.0, np.cos(0.5)])
    N = 60000
    u = rng.random((N, 3)).astype(np.float32)
    mu_bins, ph_bins = 10, 12
    counts = np.zeros((mu_bins, ph_bins))
    for k in range(N):
        s = oracle.bsdf_sample(MATS[mi], n, wo, u[k, 0], u[k, 1], u[k, 2])
        if s[8] == 0:
            continue
        wi = s[:3]
        ev = oracle.bsdf_eval(MATS[mi], n, wo, wi)
        assert abs(ev[3] - s[6]) <= 1e-5 * max(1.0, ev[3])
        mu = min(max(wi[2], 0.0), 0.999999)
        phi = math.atan2(wi[1], wi[0]) % (2 * math.pi)
        counts[int(mu * mu_bins), int(phi / (2 * math.pi) * ph_bins)] += 1
    # expected counts: pdf integrated over each (mu, phi) bin (dw = dmu dphi)
    sub = 6
    exp = np.zeros_like(counts)
    for i in range(mu_bins):
        for j in range(ph_bins):
            mus = (i + (np.arange(sub) + 0.5) / sub) / mu_bins
            phs = (j + (np.arange(sub) + 0.5) / sub) / ph_bins * 2 * np.pi
            M, PH = np.meshgrid(mus, phs, indexing="ij")
            st = np.sqrt(1 - M * M)
            W = np.stack([st * np.cos(PH), st * np.sin(PH), M], -1).reshape(-1, 3).astype(np.float32)
            pdf = np.array([oracle.bsdf_eval(MATS[mi], n, wo, wi)[3] for wi in W])
            exp[i, j] = pdf.mean() * (1.0 / mu_bins) * (2 * np.pi / ph_bins) * N
    keep = exp > 5
    chi2 = float((((counts - exp) ** 2) / exp)[keep].sum())
    dof = int(keep.sum()) - 1
    # 1e-4 upper tail of chi^2 with dof degrees of freedom (Wilson-Hilferty)
    z = 3.719
    crit = dof * (1 - 2 / (9 * dof) + z * math.sqrt(2 / (9 * dof))) ** 3
    assert chi2 < crit, (chi2, crit, dof)


# ---------------------------------------------------------- integrator -----

def _box_inward(emission, albedo):
    """Closed cube [-1,1]^3, windings (hence emission) facing inward."""
    tris = _closed_cube(1)
    c = tris.reshape(-1, 3, 3).mean(1)
    nrm = np.cross(tris[:, 3:6] - tris[:, 0:3], tris[:, 6:9] - tris[:, 0:3])
    outward = np.einsum("ij,ij->i", nrm, c) > 0
    tris[outward] = tris[outward].reshape(-1, 3, 3)[:, [0, 2, 1], :].reshape(-1, 9)
    nrm = np.cross(tris[:, 3:6] - tris[:, 0:3], tris[:, 6:9] - tris[:, 0:3])
    assert np.all(np.einsum("ij,ij->i", nrm, -c) > 0)
    mats = _mat12(base=(albedo,) * 3, metallic=0.0, specular=0.0, emission=(emission,) * 3)[None]
    return oracle.Scene(tris, np.zeros(len(tris), np.uint32), mats)


@pytest.mark.parametrize("rho,depth", [(0.5, 5), (0.8, 8), (0.0, 3)])
def test_white_furnace_closed_form(rho, depth):
    sc = _box_inward(1.0, rho)
    t = float(np.float32(math.tan(math.radians(40))))
    cam = _cam16((0.1, -0.2, 0.3), t, 24, 24)
    img, st = sc.render(cam, 128, depth)
    expect = sum(rho ** k for k in range(depth + 1))
    px = img.reshape(-1, 3)[:, 0].astype(np.float64)
    se = px.std() / math.sqrt(len(px))
    assert abs(px.mean() - expect) < 6 * se + 1e-6, (px.mean(), expect, se)
    assert st.nonfinite == 0


def _lambert_polygon_irradiance(x, n, poly, L):
    """E(x) = L/2 sum_i Theta_i (n . Gamma_i) for a uniform Lambertian polygon
    (Lambert 1760; e.g. Arvo 1995, eq. 3)."""
    E = 0.0
    k = len(poly)
    for i in range(k):
        a = poly[i] - x
        b = poly[(i + 1) % k] - x
        a /= np.linalg.norm(a)
        b /= np.linalg.norm(b)
        theta = math.acos(max(-1.0, min(1.0, float(a @ b))))
        g = np.cross(a, b)
        g /= np.linalg.norm(g)
        E += theta * float(g @ n)
    return abs(E) * L / 2.0


def test_direct_light_matches_lambert_polygon_irradiance():
    # Diffuse floor (rho 0.5) at y = 0; a downward quad light (L = 4) at
    # y = 1 offset in x; a narrow camera looking straight down at x = 0.
    rho, L = 0.5, 4.0
    floor = np.float32([[-50, 0, -50, -50, 0, 50, 50, 0, 50], [-50, 0, -50, 50, 0, 50, 50, 0, -50]])
    quad = np.float64([[0.5, 1, -0.5], [1.5, 1, -0.5], [1.5, 1, 0.5], [0.5, 1, 0.5]])
    light = np.float32([[*quad[0], *quad[1], *quad[2]], [*quad[0], *quad[2], *quad[3]]])
    ln = np.cross(light[:, 3:6] - light[:, 0:3], light[:, 6:9] - light[:, 0:3])
    assert np.all(ln[:, 1] < 0)   # one-sided emitter faces the floor
    fn = np.cross(floor[:, 3:6] - floor[:, 0:3], floor[:, 6:9] - floor[:, 0:3])
    assert np.all(fn[:, 1] > 0)
    tris = np.concatenate([floor, light])
    mats = np.stack([_mat12(base=(rho,) * 3, specular=0.0), _mat12(base=(0, 0, 0), emission=(L,) * 3)])
    sc = oracle.Scene(tris, np.uint32([0, 0, 1, 1]), mats)
    c = aqua.Camera()
    c.cam_to_world[:] = [1, 0, 0, 0, 0, 0, 1, 3, 0, -1, 0, 0]   # looking down -y, up = -z
    c.tan_half_fov = float(np.float32(math.tan(math.radians(0.25))))
    c.aspect, c.width, c.height = 1.0, 4, 4
    img, _ = sc.render(aqua.camera_to_array(c), 4096, 1)
    E = _lambert_polygon_irradiance(np.zeros(3), np.array([0.0, 1.0, 0.0]), quad, L)
    expect = rho / math.pi * E
    got = img[..., 0].astype(np.float64)
    assert abs(got.mean() - expect) < 0.01 * expect, (got.mean(), expect)


def test_fp32_film_agrees_with_fp64_film(box_room_oracle, box_room):
    cam = aqua.camera_to_array(box_room.camera())
    c = cam.copy()
    c[14:16] = np.uint32([64, 64]).view(np.float32)
    a, _ = box_room_oracle.render(c, 16, 5)
    b = box_room_oracle.render64(c, 16, 5)
    np.testing.assert_allclose(a, b, rtol=1e-5, atol=1e-6)


def test_light_table_by_hand(box_room, box_room_oracle):
    n, pmf, cdf, area, mat_pdf = box_room_oracle.light_table()
    assert n == 2
    # the light quad obj:152-156: 0.47 x 0.38, split in two equal triangles
    np.testing.assert_allclose(area[:2], 0.47 * 0.38 / 2, rtol=1e-5)
    np.testing.assert_allclose(pmf[:2], 0.5, rtol=1e-6)
    assert cdf[1] == 1.0
    lum = 0.2126 * 17 + 0.7152 * 12 + 0.0722 * 4
    total = 2 * math.pi * (0.47 * 0.38 / 2) * lum
    mats = box_room.materials()
    li = int(np.argmax(mats[:, 4]))
    np.testing.assert_allclose(mat_pdf[li], math.pi * lum / total, rtol=1e-4)


# ------------------------------------------------------ point lights ------

def _floor_and_camera(rho, look_at, tan_half=1e-5):
    """A large diffuse floor at y = 0 facing +y and a 2x2 camera 3 units
    above `look_at` looking straight down, so narrow that every pixel sees
    the same floor point to ~1e-5."""
    floor = np.float32([[-50, 0, -50, -50, 0, 50, 50, 0, 50], [-50, 0, -50, 50, 0, 50, 50, 0, -50]])
    c = aqua.Camera()
    x, z = look_at
    c.cam_to_world[:] = [1, 0, 0, x, 0, 0, 1, 3, 0, -1, 0, z]
    c.tan_half_fov = float(np.float32(tan_half))
    c.aspect, c.width, c.height = 1.0, 2, 2
    return floor, _mat12(base=(rho,) * 3, specular=0.0)[None], aqua.camera_to_array(c)


@pytest.mark.parametrize("look_at,light", [((0.0, 0.0), (0.0, 2.0, 0.0)),      # on axis
                                           ((0.0, 0.0), (1.5, 2.0, 0.5)),      # oblique
                                           ((0.7, -0.4), (-1.0, 0.8, 1.2))])   # grazing-ish
def test_point_light_radiance_closed_form(look_at, light):
    # A delta light of intensity I (W/sr) over a Lambertian floor: the
    # reflected radiance is rho / pi * I * cos(theta) / r^2 (inverse square
    # law). Depth 1 = direct light only; the floor is the only geometry, so
    # the continuation ray escapes.
    rho, I = 0.6, 5.0
    floor, mats, cam = _floor_and_camera(rho, look_at)
    pl = np.float32([[*light, I, I, I]])
    sc = oracle.Scene(floor, np.zeros(2, np.uint32), mats, pl)
    img, st = sc.render(cam, 4, 1)
    p = np.array([look_at[0], 0.0, look_at[1]])
    dv = np.array(light) - p
    r2 = float(dv @ dv)
    cos = dv[1] / math.sqrt(r2)
    expect = rho / math.pi * I * cos / r2
    np.testing.assert_allclose(img.astype(np.float64), expect, rtol=5e-5)
    assert st.shadow_rays == 4 * 4


def test_two_point_lights_selected_by_power():
    # Two grey point lights at mirror positions about the shaded point: each
    # sample picks one light with probability proportional to its power and
    # divides by that probability. With intensities 2 and 6 the two
    # estimates I_i cos / (r^2 p_i) coincide (p_i = I_i / 8), so every sample
    # equals the sum of both lights' radiance exactly. A missing 1 / p_sel
    # (or a uniform pick) gives a different value on every sample.
    rho = 0.5
    floor, mats, cam = _floor_and_camera(rho, (0.0, 0.0))
    pl = np.float32([[1.0, 2.0, 0.0, 2.0, 2.0, 2.0], [-1.0, 2.0, 0.0, 6.0, 6.0, 6.0]])
    sc = oracle.Scene(floor, np.zeros(2, np.uint32), mats, pl)
    n, pmf, *_ = sc.light_table()
    assert n == 2
    np.testing.assert_allclose(pmf[:2], [0.25, 0.75], rtol=1e-6)
    img, _ = sc.render(cam, 64, 1)
    r2, cos = 5.0, 2.0 / math.sqrt(5.0)
    expect = rho / math.pi * (2.0 + 6.0) * cos / r2
    np.testing.assert_allclose(img.astype(np.float64), expect, rtol=5e-5)


# ----------------------------------------------------------- Fresnel ------

def _fresnel_dielectric(cos_i, eta):
    """Unpolarised Fresnel reflectance of a smooth dielectric (Fresnel's
    equations; e.g. Born & Wolf, sec. 1.5.2)."""
    sin_t2 = (1.0 - cos_i * cos_i) / (eta * eta)
    cos_t = math.sqrt(max(0.0, 1.0 - sin_t2))
    rs = (cos_i - eta * cos_t) / (cos_i + eta * cos_t)
    rp = (eta * cos_i - cos_t) / (eta * cos_i + cos_t)
    return 0.5 * (rs * rs + rp * rp)


def _mirror_fresnel(mat, rough, cos_o):
    """F(wo.h) the BSDF applies, read back from f in the mirror configuration
    (wi = reflect(wo), so h = n and wo.h = cos_o): f = F D(1) G2 / (4 cos^2)
    with D and Lambda the GGX terms pinned by the normalisation tests."""
    s = math.sqrt(1.0 - cos_o * cos_o)
    wo = np.float32([s, 0.0, cos_o])
    wi = np.float32([-s, 0.0, cos_o])
    n = np.float32([0, 0, 1])
    f = oracle.bsdf_eval(mat, n, wo, wi)[:3].astype(np.float64)
    D1, _ = oracle.ggx(rough, 1.0, 0.5)
    _, lam = oracle.ggx(rough, 1.0, float(np.float32(cos_o)))
    G2 = 1.0 / (1.0 + 2.0 * lam)
    c = float(np.float32(cos_o))
    return f * (4.0 * c * c) / (D1 * G2)


def test_dielectric_fresnel_matches_ior_1_5():
    # specular 0.5 is the usual dielectric default: F0 = 0.08 * 0.5 = 0.04,
    # the normal-incidence reflectance of glass-like IOR 1.5,
    # ((1.5 - 1) / (1.5 + 1))^2. With a black base only the specular lobe is
    # left. Schlick's curve stays within 0.04 of Fresnel's equations at every
    # angle (its largest gap, 0.036, is near 85 degrees) and reaches 1 at
    # grazing incidence.
    rough = 0.3
    mat = _mat12(base=(0.0, 0.0, 0.0), metallic=0.0, roughness=rough, specular=0.5)
    F_n = _mirror_fresnel(mat, rough, 1.0)
    np.testing.assert_allclose(F_n, 0.04, rtol=2e-5)
    assert abs(_fresnel_dielectric(1.0, 1.5) - 0.04) < 1e-15
    for deg in (10, 30, 45, 60, 70, 80, 85, 88):
        c = math.cos(math.radians(deg))
        F = _mirror_fresnel(mat, rough, c)
        assert np.all(np.abs(F - _fresnel_dielectric(c, 1.5)) < 0.04), (deg, F, _fresnel_dielectric(c, 1.5))
    assert np.all(_mirror_fresnel(mat, rough, 0.002) > 0.98)


@pytest.mark.parametrize("rough", [0.2, 0.5, 0.9])
def test_metal_reflects_its_base_colour_at_normal_incidence(rough):
    # Metallic = 1: the base colour IS the normal-incidence reflectance (F0),
    # with no diffuse lobe, so f(n, n) = base * D(1) / 4 = base / (4 pi a^2)
    # (G2 = 1 at normal incidence). Towards grazing, F rises to 1.
    base = (0.9, 0.6, 0.3)
    mat = _mat12(base=base, metallic=1.0, roughness=rough, specular=0.5)
    n = np.float32([0, 0, 1])
    f = oracle.bsdf_eval(mat, n, n, n)[:3].astype(np.float64)
    a = max(float(np.float32(rough) * np.float32(rough)), 1e-3)
    # D's FP32 form ((a^2 - 1) + 1 at n.h = 1) cancels: ~2^-23 / a^2 relative
    tol = 1e-5 + 2.0 ** -22 / (a * a)
    np.testing.assert_allclose(f, np.array(base) / (4 * math.pi * a * a), rtol=tol)
    np.testing.assert_allclose(_mirror_fresnel(mat, rough, 1.0), base, rtol=2e-5)
    Fg = _mirror_fresnel(mat, rough, 0.01)
    assert np.all(Fg > 0.95) and np.all(Fg <= 1.0 + 1e-5), Fg
    # half metallic: F0 is the average of the dielectric 0.04 and the base
    half = _mat12(base=base, metallic=0.5, roughness=rough, specular=0.5)
    spec_only = _mirror_fresnel(half, rough, 1.0)
    diffuse = 0.5 * np.array(base) / math.pi * (1.0 - 0.04)   # (1 - metallic)(1 - Fd) base / pi
    D1, _ = oracle.ggx(rough, 1.0, 0.5)
    spec_only = spec_only - diffuse * 4.0 / D1
    np.testing.assert_allclose(spec_only, 0.5 * 0.04 + 0.5 * np.array(base), rtol=2e-4)
