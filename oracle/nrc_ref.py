"""CPU restatement of the neural radiance cache (test infrastructure only:
loaded by tests/ as the checker, never by the product).

The reference names the integrator (`"type": "nrc"`, scenes/integrator.json:2,
with batch_size / training_iters / learning_rate at :4,6,7) but ships no code,
so this restates the published methods the GPU kernels implement:

* multiresolution hash encoding -- Mueller, Evans, Schied, Keller 2022,
  "Instant Neural Graphics Primitives": L levels of F features, resolution
  N_l = floor(N_min * b^l), trilinear interpolation of the 8 cell corners,
  dense indexing when (N_l + 1)^3 <= T, otherwise the spatial hash with
  its primes applied to 4 x 4 x 2 blocks of vertices, the vertex's place in
  its block in the low 5 bits (x, y's low bit, z, y's high bit: each 64-byte
  half of a block a 4 x 2 x 2 brick):
  ((((x >> 2) * 1) ^ ((y >> 2) * 2654435761) ^ ((z >> 1) * 805459861)) << 5
   | (x & 3) | (y & 1) << 2 | (z & 1) << 3 | (y & 2) << 3) mod T;
* real spherical harmonics up to degree 4 (16 coefficients, l = 0..3) of the
  direction (the standard graphics table, e.g. Sloan 2008);
* an MLP of 4 hidden ReLU layers of width 64, no biases (a constant-1 input
  feature plays that role), 3 linear outputs;
* the relative L2 loss of Mueller, Rousselle, Novak, Keller 2021 ("Real-time
  Neural Radiance Caching for Path Tracing"):
  mean over samples and channels of (y - t)^2 / (sg(y)^2 + 0.01);
* Adam (Kingma & Ba 2015) with bias correction;
* the order the cache evaluates records in: a stable sort by the Morton code
  (Z-order, x lowest) of the position's 10-bit cells, its low 6 bits
  dropped. This order has no counterpart in the reference; it is defined here
  and pinned by a bit-by-bit interleave in the tests.

Feature order of the 64-wide input (the GPU kernels use the same order):
  [0, 32)  hash features, level-major (level l -> 2l, 2l + 1)
  [32, 48) SH coefficients 0..15
  [48, 64) normal xyz, roughness, albedo rgb, 1.0, then 8 zeros
"""
from __future__ import annotations

import numpy as np

try:
    import torch
except Exception:  # pragma: no cover - torch is present in this image
    torch = None

N_LEVELS = 16
N_FEATURES = 2
LOG2_TABLE = 19
BASE_RES = 16
LEVEL_SCALE = 1.5
WIDTH = 64
N_HIDDEN = 4
N_OUT_PAD = 16
PRIMES = (1, 2654435761, 805459861)


def level_res(l: int, base=BASE_RES, scale=LEVEL_SCALE) -> int:
    # exact in double for these arguments (1.5^l = 3^l / 2^l)
    return int(np.floor(float(base) * float(scale) ** l))


def level_is_dense(l: int, log2_table: int = LOG2_TABLE) -> bool:
    n = level_res(l) + 1
    return n * n * n <= (1 << log2_table)


def corner_index(l: int, c: np.ndarray, log2_table=LOG2_TABLE) -> np.ndarray:
    """c: int64 (..., 3) vertex coordinates of level l -> table index."""
    T = 1 << log2_table
    if level_is_dense(l, log2_table):
        n = level_res(l) + 1
        return (c[..., 0] + n * (c[..., 1] + n * c[..., 2])).astype(np.int64)
    c = c.astype(np.uint64)
    x, y, z = c[..., 0], c[..., 1], c[..., 2]
    b = ((x >> np.uint64(2)) * np.uint64(PRIMES[0])) ^ ((y >> np.uint64(2)) * np.uint64(PRIMES[1])) ^ \
        ((z >> np.uint64(1)) * np.uint64(PRIMES[2]))
    h = ((b & np.uint64(0xFFFFFFFF)) << np.uint64(5)) | (x & np.uint64(3)) | ((y & np.uint64(1)) << np.uint64(2)) | \
        ((z & np.uint64(1)) << np.uint64(3)) | ((y & np.uint64(2)) << np.uint64(3))
    return (h & np.uint64(0xFFFFFFFF) & np.uint64(T - 1)).astype(np.int64)


def corners(x: np.ndarray, l: int, log2_table=LOG2_TABLE):
    """Indices (n, 8) and trilinear weights (n, 8) of level l for positions
    x (n, 3) in [0, 1]^3, computed in FP32 as on the GPU (the cell is clamped
    so that x = 1 stays in the last cell)."""
    x = np.clip(x.astype(np.float32), 0.0, 1.0)
    N = np.float32(level_res(l))
    s = x * N
    c0 = np.minimum(np.floor(s), N - np.float32(1))
    f = (s - c0).astype(np.float32)
    c0 = c0.astype(np.int64)
    idx = np.zeros((len(x), 8), np.int64)
    w = np.zeros((len(x), 8), np.float32)
    for k in range(8):
        bit = np.array([(k >> 0) & 1, (k >> 1) & 1, (k >> 2) & 1])
        c = c0 + bit
        idx[:, k] = corner_index(l, c, log2_table)
        wk = np.ones(len(x), np.float32)
        for d in range(3):
            wk = wk * (f[:, d] if bit[d] else (np.float32(1) - f[:, d]))
        w[:, k] = wk
    return idx, w


SH_C = [0.28209479177387814, 0.48860251190291987, 1.0925484305920792, 0.31539156525251999,
        0.54627421529603959, 0.59004358992664352, 2.8906114426405538, 0.45704579946446572,
        0.3731763325901154, 1.4453057213202769]


def sh4(d):
    """Real SH basis l = 0..3 of unit directions d (n, 3) -> (n, 16)."""
    x, y, z = d[:, 0], d[:, 1], d[:, 2]
    xx, yy, zz = x * x, y * y, z * z
    one = np.ones_like(x) if isinstance(d, np.ndarray) else torch.ones_like(x)
    out = [
        SH_C[0] * one,
        -SH_C[1] * y,
        SH_C[1] * z,
        -SH_C[1] * x,
        SH_C[2] * x * y,
        -SH_C[2] * y * z,
        0.94617469575755997 * zz - SH_C[3],
        -SH_C[2] * x * z,
        SH_C[4] * (xx - yy),
        SH_C[5] * y * (-3.0 * xx + yy),
        SH_C[6] * x * y * z,
        SH_C[7] * y * (1.0 - 5.0 * zz),
        SH_C[8] * z * (5.0 * zz - 3.0),
        SH_C[7] * x * (1.0 - 5.0 * zz),
        SH_C[9] * z * (xx - yy),
        SH_C[5] * x * (-xx + 3.0 * yy),
    ]
    if isinstance(d, np.ndarray):
        return np.stack(out, 1)
    return torch.stack(out, 1)


def split_inputs(rec: np.ndarray):
    """Records (n, 16): pos xyz, rough, dir xyz, -, normal xyz, -, albedo rgb, -."""
    rec = np.asarray(rec, np.float32).reshape(-1, 16)
    return rec[:, 0:3], rec[:, 3], rec[:, 4:7], rec[:, 8:11], rec[:, 12:15]


def features_torch(rec: np.ndarray, table, dtype=None):
    """64-wide input features; `table` is a torch tensor (L, T, F) so autograd
    reaches the table entries."""
    pos, rough, dirs, nrm, alb = split_inputs(rec)
    n = len(pos)
    dtype = dtype or table.dtype
    log2_table = int(round(np.log2(table.shape[1])))
    cols = []
    for l in range(N_LEVELS):
        idx, w = corners(pos, l, log2_table)
        g = table[l][torch.from_numpy(idx)]            # (n, 8, F)
        wt = torch.from_numpy(w.astype(np.float64)).to(dtype)[:, :, None]
        cols.append((g * wt).sum(1))
    hashf = torch.cat(cols, 1)                          # level-major (n, 32)
    shf = sh4(torch.from_numpy(dirs.astype(np.float64)).to(dtype))
    extra = torch.zeros((n, 16), dtype=dtype)
    extra[:, 0:3] = torch.from_numpy(nrm.astype(np.float64)).to(dtype)
    extra[:, 3] = torch.from_numpy(rough.astype(np.float64)).to(dtype)
    extra[:, 4:7] = torch.from_numpy(alb.astype(np.float64)).to(dtype)
    extra[:, 7] = 1.0
    return torch.cat([hashf, shf, extra], 1)


def unpack_mlp(flat: np.ndarray):
    """Flat MLP parameters -> [W0..W3 (64x64), W4 (64x16)], each W[in][out]."""
    flat = np.asarray(flat, np.float64)
    Ws = [flat[i * 4096:(i + 1) * 4096].reshape(64, 64) for i in range(4)]
    Ws.append(flat[4 * 4096:4 * 4096 + 1024].reshape(64, 16))
    return Ws


def _rounder(quant):
    if quant is None:
        return lambda x: x
    dt = {"f16": torch.float16, "bf16": torch.bfloat16}[quant]
    return lambda x: x.to(dt).to(x.dtype)


def _round_ops(quant):
    """(round forward value, round backward gradient) autograd ops that model
    where the MFMA kernels convert to 16 bits: layer inputs and weights on the
    way forward, pre-activation gradients on the way back."""
    rnd = _rounder(quant)

    class FwdRound(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x):
            return rnd(x)

        @staticmethod
        def backward(ctx, g):
            return g

    class BwdRound(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x):
            return x

        @staticmethod
        def backward(ctx, g):
            return rnd(g)

    return FwdRound.apply, BwdRound.apply


def forward_torch(X, Ws, quant=None):
    """quant None: exact; "f16" / "bf16": inputs of every layer and the
    weights rounded forward, pre-activation gradients rounded backward."""
    fr, br = _round_ops(quant)
    if quant is not None:
        Ws = [fr(W) for W in Ws]
    h = X
    for W in Ws[:N_HIDDEN]:
        h = torch.relu(br(fr(h) @ W))
    return br(fr(h) @ Ws[N_HIDDEN])[:, :3]


E4M3_MAX = 448.0


def e4m3(x):
    """Round float32 values to OCP FP8 e4m3 (round to nearest even, subnormals
    kept; inputs are pre-saturated to +-448, where e4m3 has no infinity) and
    back to float64."""
    x32 = torch.as_tensor(x).to(torch.float32)
    x32 = torch.clamp(x32, -E4M3_MAX, E4M3_MAX)
    return x32.to(torch.float8_e4m3fn).to(torch.float64)


def forward_fp8(X, Ws, wscale, ascale):
    """The FP8 query path: per layer L, inputs e4m3(sat(x * ascale[L])) and
    weights e4m3(sat(w * wscale[L])) multiplied exactly and accumulated (here
    in float64; FP32 on the GPU), the sum scaled by float32 1 / (ascale[L] *
    wscale[L]); ReLU between the hidden layers. Every product and scale is
    formed in float32 as the kernels form it."""
    h = X
    for L, W in enumerate(Ws):
        sa = np.float32(ascale[L])
        sw = np.float32(wscale[L])
        deq = float(np.float32(1.0) / np.float32(sa * sw))
        hq = e4m3(h.to(torch.float32) * torch.tensor(sa, dtype=torch.float32))
        Wq = e4m3(W.to(torch.float32) * torch.tensor(sw, dtype=torch.float32))
        out = (hq @ Wq) * deq
        h = torch.relu(out).to(torch.float32).to(torch.float64) if L < N_HIDDEN else out
    return h[:, :3]


def relative_l2(y, t):
    return (((y - t) ** 2) / (y.detach() ** 2 + 0.01)).mean()


def _table_values(table_np, quant):
    t = np.asarray(table_np, np.float64)
    # the kernels gather from an FP16 copy of the table in every mode
    # (the FP8 path included)
    return t.astype(np.float16).astype(np.float64) if quant is not None else t


def infer(rec, table_np, mlp_np, quant=None, fp8_scales=None):
    """float64 forward pass -> (n, 3). quant "fp8" needs fp8_scales =
    (wscale[5], ascale[5]), the scales the query kernel was given."""
    table = torch.from_numpy(_table_values(table_np, quant))
    Ws = [torch.from_numpy(W.astype(np.float64)) for W in unpack_mlp(mlp_np)]
    with torch.no_grad():
        X = features_torch(rec, table)
        if quant == "fp8":
            wscale, ascale = fp8_scales
            return forward_fp8(X, Ws, wscale, ascale).numpy()
        return forward_torch(X, Ws, quant).numpy()


def loss_and_grads(rec, targets, table_np, mlp_np, dtype=torch.float64 if torch else None, quant=None):
    table = torch.tensor(_table_values(table_np, quant), dtype=dtype, requires_grad=True)
    Ws = [torch.tensor(W, dtype=dtype, requires_grad=True) for W in unpack_mlp(mlp_np)]
    y = forward_torch(features_torch(rec, table, dtype), Ws, quant)
    t = torch.from_numpy(np.asarray(targets, np.float64).reshape(-1, 3)).to(dtype)
    loss = relative_l2(y, t)
    loss.backward()
    g_mlp = np.concatenate([W.grad.numpy().reshape(-1) for W in Ws])
    return float(loss.detach()), table.grad.numpy(), g_mlp


def morton30(pos):
    """Morton codes of positions [n, 3]: 10-bit cells per axis
    (min(floor(clamp(p, 0, 1) * 1024), 1023) in float32), interleaved with x in
    bit 0, y in bit 1, z in bit 2 of each group of three."""
    p = np.asarray(pos, np.float32)
    cell = np.minimum((np.clip(p, 0.0, 1.0).astype(np.float32) * np.float32(1024.0)).astype(np.uint32), 1023)
    key = np.zeros(len(p), np.uint64)
    for axis in range(3):
        v = cell[:, axis].astype(np.uint64)
        v = (v | (v << 16)) & 0x030000FF
        v = (v | (v << 8)) & 0x0300F00F
        v = (v | (v << 4)) & 0x030C30C3
        v = (v | (v << 2)) & 0x09249249
        key |= v << axis
    return key.astype(np.uint32)


def query_order(rec):
    """Record order of a cache pass: stable argsort of morton30 >> 6."""
    rec = np.asarray(rec, np.float32).reshape(-1, 16)
    return np.argsort(morton30(rec[:, 0:3]) >> 6, kind="stable").astype(np.uint32)


def adam_step(params, grads, m, v, step, lr, beta1=0.9, beta2=0.99, eps=1e-15):
    """One Adam update (Kingma & Ba 2015, Algorithm 1) on flat arrays; returns
    new (params, m, v). `step` is 1 for the first update."""
    m = beta1 * m + (1.0 - beta1) * grads
    v = beta2 * v + (1.0 - beta2) * grads * grads
    mh = m / (1.0 - beta1 ** step)
    vh = v / (1.0 - beta2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v
