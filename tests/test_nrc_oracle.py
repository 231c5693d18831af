"""Pins of the NRC restatement (oracle/nrc_ref.py): SH orthonormality and the
addition theorem, hash-grid interpolation reproducing linear fields, dense
indexing bijectivity, finite-difference gradients, Adam vs torch.optim.Adam."""
import importlib.util
import os

import numpy as np
import pytest
import torch

from conftest import ROOT

spec = importlib.util.spec_from_file_location("nrc_ref", os.path.join(ROOT, "oracle", "nrc_ref.py"))
ref = importlib.util.module_from_spec(spec)
spec.loader.exec_module(ref)


def _sphere_grid(n_t=200, n_p=400):
    # Gauss-Legendre in cos(theta) x uniform phi: exact for these polynomials
    mu, wmu = np.polynomial.legendre.leggauss(n_t)
    ph = (np.arange(n_p) + 0.5) / n_p * 2 * np.pi
    M, P = np.meshgrid(mu, ph, indexing="ij")
    st = np.sqrt(1 - M * M)
    d = np.stack([st * np.cos(P), st * np.sin(P), M], -1).reshape(-1, 3)
    w = (wmu[:, None] * np.full(n_p, 2 * np.pi / n_p)[None, :]).reshape(-1)
    return d, w


def test_sh_orthonormal():
    d, w = _sphere_grid()
    Y = ref.sh4(d)
    G = (Y * w[:, None]).T @ Y
    np.testing.assert_allclose(G, np.eye(16), atol=1e-10)


def test_sh_addition_theorem():
    # sum_m Y_lm(d)^2 = (2l + 1) / (4 pi) for every direction and band
    rng = np.random.default_rng(0)
    d = rng.standard_normal((1000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    Y = ref.sh4(d)
    for l, (a, b) in enumerate([(0, 1), (1, 4), (4, 9), (9, 16)]):
        np.testing.assert_allclose((Y[:, a:b] ** 2).sum(1), (2 * l + 1) / (4 * np.pi), rtol=1e-12)


def test_dense_level_indexing_is_bijective():
    for l in range(ref.N_LEVELS):
        if not ref.level_is_dense(l):
            continue
        n = ref.level_res(l) + 1
        c = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3)
        idx = ref.corner_index(l, c)
        assert sorted(idx.tolist()) == list(range(n ** 3))


def test_hash_level_indices_in_range():
    rng = np.random.default_rng(1)
    for l in range(ref.N_LEVELS):
        c = rng.integers(0, ref.level_res(l) + 1, (5000, 3))
        idx = ref.corner_index(l, c, 12)
        assert idx.min() >= 0 and idx.max() < (1 << 12)


def test_hash_levels_keep_vertex_blocks_together():
    # hashed levels: the 32 vertices of a 4 x 4 x 2 block take one aligned
    # 32-entry run, each half (16 entries, 64 bytes) a 4 x 2 x 2 brick: x in
    # bits 0-1, y's low bit in bit 2, z in bit 3, y's high bit in bit 4;
    # different blocks hash apart
    l = next(l for l in range(ref.N_LEVELS) if not ref.level_is_dense(l))
    rng = np.random.default_rng(3)
    base = rng.integers(0, ref.level_res(l) // 4, (200, 3)) * np.array([4, 4, 2])
    off = np.stack(np.meshgrid(np.arange(4), np.arange(4), np.arange(2), indexing="ij"), -1).reshape(-1, 3)
    idx = ref.corner_index(l, base[:, None, :] + off[None, :, :])
    assert np.all(idx // 32 == idx[:, :1] // 32)
    np.testing.assert_array_equal(idx % 32, np.broadcast_to(off[:, 0] + 4 * (off[:, 1] & 1) + 8 * off[:, 2] +
                                                            16 * (off[:, 1] >> 1), idx.shape))
    # the block part is the spatial hash of the block coordinates
    bc = (base // np.array([4, 4, 2])).astype(np.uint64)
    h = (bc[:, 0] * np.uint64(1)) ^ (bc[:, 1] * np.uint64(2654435761)) ^ (bc[:, 2] * np.uint64(805459861))
    np.testing.assert_array_equal(idx[:, 0] // 32, ((h & np.uint64(0xFFFFFFFF)) % np.uint64(1 << 14)).astype(np.int64))


def test_trilinear_reproduces_linear_field():
    # On a dense level, table values that are linear in the vertex position
    # interpolate to the same linear function of x.
    l = 1
    n = ref.level_res(l) + 1
    N = ref.level_res(l)
    table = np.zeros((ref.N_LEVELS, 1 << 19, 2))
    c = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3)
    idx = ref.corner_index(l, c)
    table[l, idx, 0] = c[:, 0] / N + 2 * c[:, 1] / N
    table[l, idx, 1] = c[:, 2] / N
    rng = np.random.default_rng(2)
    x = rng.random((500, 3)).astype(np.float32)
    x[:5] = [[0, 0, 0], [1, 1, 1], [0.5, 0.25, 1.0], [1, 0, 0.999], [0.0, 1.0, 0.5]]
    rec = np.zeros((500, 16), np.float32)
    rec[:, 0:3] = x
    rec[:, 4:7] = [0, 0, 1]
    f = ref.features_torch(rec, torch.from_numpy(table)).numpy()
    np.testing.assert_allclose(f[:, 2 * l], x[:, 0] + 2 * x[:, 1], atol=1e-6)
    np.testing.assert_allclose(f[:, 2 * l + 1], x[:, 2], atol=1e-6)
    np.testing.assert_allclose(f[:, 55], 1.0)


def _problem(n=40, log2=12, seed=0):
    rng = np.random.default_rng(seed)
    rec = np.zeros((n, 16), np.float32)
    rec[:, 0:3] = rng.random((n, 3))
    rec[:, 3] = rng.random(n)
    d = rng.standard_normal((n, 3))
    rec[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    nn = rng.standard_normal((n, 3))
    rec[:, 8:11] = nn / np.linalg.norm(nn, axis=1, keepdims=True)
    rec[:, 12:15] = rng.random((n, 3))
    table = rng.uniform(-0.5, 0.5, (ref.N_LEVELS, 1 << log2, 2))
    mlp = rng.uniform(-0.3, 0.3, 4 * 4096 + 1024)
    mlp[4 * 4096:].reshape(64, 16)[:, 3:] = 0.0
    tgt = rng.random((n, 3))
    return rec, table, mlp, tgt


def test_gradients_match_finite_differences():
    rec, table, mlp, tgt = _problem()
    loss, gt, gm = ref.loss_and_grads(rec, tgt, table, mlp)

    def L(tb, mp):
        # the stop-gradient denominator is held at the unperturbed outputs
        return ref.loss_and_grads(rec, tgt, tb, mp)[0]

    y0 = ref.infer(rec, table, mlp)

    def L_fixed(tb, mp):
        y = ref.infer(rec, tb, mp)
        return float((((y - tgt) ** 2) / (y0 ** 2 + 0.01)).mean())

    assert abs(L_fixed(table, mlp) - loss) < 1e-12
    rng = np.random.default_rng(4)
    h = 1e-6
    for k in rng.choice(np.flatnonzero(np.abs(gm) > 1e-6), 12, replace=False):
        mp, mm = mlp.copy(), mlp.copy()
        mp[k] += h
        mm[k] -= h
        fd = (L_fixed(table, mp) - L_fixed(table, mm)) / (2 * h)
        assert abs(fd - gm[k]) <= 1e-5 * max(1.0, abs(gm[k])), (k, fd, gm[k])
    flat = np.flatnonzero(np.abs(gt.reshape(-1)) > 1e-6)
    for k in rng.choice(flat, 12, replace=False):
        tp, tm = table.copy().reshape(-1), table.copy().reshape(-1)
        tp[k] += h
        tm[k] -= h
        fd = (L_fixed(tp.reshape(table.shape), mlp) - L_fixed(tm.reshape(table.shape), mlp)) / (2 * h)
        assert abs(fd - gt.reshape(-1)[k]) <= 1e-5 * max(1.0, abs(fd)), (k, fd, gt.reshape(-1)[k])


def test_adam_matches_torch():
    rng = np.random.default_rng(5)
    p = rng.standard_normal(1000)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    tp = torch.tensor(p.copy(), requires_grad=True)
    opt = torch.optim.Adam([tp], lr=1e-2, betas=(0.9, 0.99), eps=1e-15)
    for step in range(1, 4):
        g = rng.standard_normal(1000)
        p, m, v = ref.adam_step(p, g, m, v, step, 1e-2)
        opt.zero_grad()
        tp.grad = torch.tensor(g)
        opt.step()
        np.testing.assert_allclose(p, tp.detach().numpy(), rtol=1e-12, atol=1e-14)


def _e4m3_values():
    """Every finite OCP e4m3 value from its bit fields: sign, 4-bit exponent
    (bias 7), 3-bit mantissa; exponent 0 is subnormal (m / 8 * 2^-6); the
    code S.1111.111 is NaN, so the largest finite value is 1.75 * 2^8 = 448."""
    vals = []
    for code in range(256):
        s, e, m = code >> 7, (code >> 3) & 15, code & 7
        if e == 15 and m == 7:
            continue
        mag = (m / 8.0) * 2.0 ** -6 if e == 0 else (1.0 + m / 8.0) * 2.0 ** (e - 7)
        vals.append(-mag if s else mag)
    return np.array(sorted(set(vals)))


def test_e4m3_rounding_is_nearest_even_over_every_code():
    grid = _e4m3_values()
    assert len(grid) == 253 and grid.max() == 448.0 and grid[grid > 0].min() == 2.0 ** -9
    # representable values round to themselves
    np.testing.assert_array_equal(ref.e4m3(torch.tensor(grid)).numpy(), grid)
    # random values: the nearest grid value, ties to the even mantissa
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(-448, 448, 20000), rng.uniform(-1, 1, 20000) * 2.0 ** -5,
                        (grid[:-1] + grid[1:]) / 2])  # exact midpoints
    x = x.astype(np.float32).astype(np.float64)
    got = ref.e4m3(torch.tensor(x)).numpy()
    d = np.abs(x[:, None] - grid[None, :])
    best = d.min(1)
    assert np.all(np.abs(got - x) == best)
    mid = np.isclose(d, best[:, None], rtol=0, atol=0).sum(1) == 2
    assert mid.sum() >= 250
    code_of = {}
    for code in range(256):
        e, m = (code >> 3) & 15, code & 7
        if not (e == 15 and m == 7) and code < 128:
            code_of[(m / 8.0) * 2.0 ** -6 if e == 0 else (1.0 + m / 8.0) * 2.0 ** (e - 7)] = code
    for xv, g in zip(x[mid], got[mid]):
        assert code_of[abs(float(g))] & 1 == 0, (xv, g)   # the even neighbour
    # saturation: beyond the largest finite value
    np.testing.assert_array_equal(ref.e4m3(torch.tensor([500.0, -1e6])).numpy(), [448.0, -448.0])


def test_fp8_forward_is_exact_when_quantisation_is_exact():
    # Weights and inputs that are e4m3 values after scaling lose nothing to
    # the quantisation, so the FP8 path must equal the exact forward pass.
    rng = np.random.default_rng(4)
    grid = _e4m3_values()
    small = grid[(np.abs(grid) >= 2.0 ** -3) & (np.abs(grid) <= 4.0)]
    Ws = [torch.tensor(rng.choice(small, (64, 64)) / 8.0) for _ in range(4)]
    Ws.append(torch.tensor(rng.choice(small, (64, 16)) / 8.0))
    X = torch.tensor(rng.choice(small[small > 0], (32, 64)) / 4.0)
    wscale = [8.0] * 5
    ascale = [4.0, 1.0, 1.0, 1.0, 1.0]
    got = ref.forward_fp8(X, Ws, wscale, ascale).numpy()
    # layers 1..4 see the previous layer's output: keep them exact by
    # quantising that output the same way in the exact reference
    h = X
    for L, W in enumerate(Ws):
        hq = ref.e4m3(h.to(torch.float32) * ascale[L]) / ascale[L]
        out = hq @ W
        h = torch.relu(out) if L < 4 else out
    np.testing.assert_allclose(got, h[:, :3].numpy(), rtol=1e-6, atol=1e-9)
    # a power-of-two change of a scale keeps the quantisation lossless, and
    # the dequantisation 1 / (ascale wscale) undoes it
    np.testing.assert_allclose(ref.forward_fp8(X, Ws, [2.0] * 5, ascale).numpy(), got, rtol=1e-6, atol=1e-9)
    # a scale that pushes the weights past 448 saturates them: results change
    sat = ref.forward_fp8(X, Ws, [4096.0] * 5, ascale).numpy()
    assert not np.allclose(sat, got, rtol=1e-2)


def test_morton_codes_match_a_bit_by_bit_interleave():
    # the magic-number spread against the definition: bit b of the x / y / z
    # cell goes to bit 3b / 3b + 1 / 3b + 2
    rng = np.random.default_rng(5)
    pos = rng.uniform(-0.2, 1.2, (4000, 3)).astype(np.float32)
    pos[:6] = [[0, 0, 0], [1, 1, 1], [1 / 1024, 0, 0], [0, 1 / 1024, 0], [0, 0, 1 / 1024], [0.5, 0.25, 0.125]]
    got = ref.morton30(pos)
    cells = [[min(int(np.float32(min(max(c, 0.0), 1.0)) * np.float32(1024.0)), 1023) for c in p] for p in pos]
    want = [sum(((c[a] >> b) & 1) << (3 * b + a) for b in range(10) for a in range(3)) for c in cells]
    np.testing.assert_array_equal(got, np.array(want, np.uint32))
    assert list(got[:6]) == [0, (1 << 30) - 1, 1, 2, 4, (1 << 27) | (1 << 25) | (1 << 23)]


def test_query_order_is_a_stable_sort_of_cells():
    rng = np.random.default_rng(6)
    rec = np.zeros((3000, 16), np.float32)
    rec[:, 0:3] = rng.integers(0, 8, (3000, 3)) / 8.0 + 1e-4  # many equal cells
    order = ref.query_order(rec)
    assert sorted(order.tolist()) == list(range(3000))
    key = ref.morton30(rec[order, 0:3]) >> 6
    assert np.all(np.diff(key.astype(np.int64)) >= 0)
    same = np.diff(key.astype(np.int64)) == 0
    assert np.all(np.diff(order.astype(np.int64))[same] > 0)  # ties keep record order
