"""CPU oracle of the neural radiance cache (TEST INFRASTRUCTURE ONLY: tests/
import it as the checker; the product never does).

The cache is an addition to the reference (which has no learned part),
so the oracle restates the method as include/ornament_hip.h defines it, with
the same rounding points as the gfx950 kernels:

  encoding  Instant-NGP style multiresolution hash grid (Mueller et al. 2022):
            16 levels; position p in the scene box -> u = clamp((p - lo) *
            inv_extent, 0, 1); x = u * res; c = min(floor(x), res - 1);
            corner index dense (cx + (res+1) (cy + (res+1) cz)) when
            (res+1)^3 fits the table, else (cx ^ cy*2654435761 ^
            cz*805459861) mod 2^log2_table; trilinear weights (wx wy) wz;
            features summed corner by corner in f32 from the fp16 table.
  inputs    32 grid features, dir*0.5+0.5, normal*0.5+0.5, albedo,
            roughness, one-hot type (4), 1, zeros (64 in all)
  network   x0 = bf16(inputs); x(l+1) = bf16(relu(x(l) W(l))), l = 0..4;
            y = x5 W5 (products of bf16 values accumulated in f64 here, in f32
            by the MFMA units)
  loss      mean over samples and RGB of (y - t)^2 / (y^2 + 0.01), denominator
            held constant
  backward  dY5 = bf16(dL/dy); dW(l) = x(l)^T dY(l); dY(l-1) = bf16((dY(l)
            W(l)^T) * [x(l) > 0]); grid gradient from dx0 = dY0 W0^T (f32)
            scattered to the 8 corners with the trilinear weights
  Adam      m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2; p -= lr (m / (1-b1^t)) /
            (sqrt(v / (1-b2^t)) + eps); hash-grid parameters whose gradient
            is exactly 0 skip the step (value and moments kept, t still
            counts every step: Mueller et al. 2022, sec. 4)

Independent checks in tests/test_nrc_oracle.py pin this restatement: the
hash and dense indices against closed forms, the encoding's linearity and
partition of unity, the analytic backward against torch float64 autograd
(straight-through rounding) and finite differences, Adam against
torch.optim.Adam and the skipping form against torch.optim.SparseAdam.
"""
import numpy as np

LEVELS = 16
IN = 64
W = 64
OUT = 16
HIDDEN = 5
MLP_PARAMS = HIDDEN * IN * W + W * OUT
P1, P2 = np.uint32(2654435761), np.uint32(805459861)


def bf16(x):
    """float32 -> bf16 (round to nearest even) -> float32."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    return np.where(np.isnan(x), np.float32(np.nan), out).astype(np.float32)


def unpack_mlp(p):
    """Parameter vector -> [W0..W4 (64x64), W5 (64x16)] in [in][out] layout."""
    p = np.asarray(p, np.float32)
    ws = [p[l * IN * W:(l + 1) * IN * W].reshape(IN, W) for l in range(HIDDEN)]
    ws.append(p[HIDDEN * IN * W:].reshape(W, OUT))
    return ws


def unit_position(info, pos):
    lo = np.asarray(info["bounds_lo"], np.float32)
    ie = np.asarray(info["inv_extent"], np.float32)
    u = (np.asarray(pos, np.float32) - lo) * ie
    return np.minimum(np.maximum(u, np.float32(0.0)), np.float32(1.0)).astype(np.float32)


def level_corners(info, level, u):
    """(indices (N, 8) uint32, weights (N, 8) float32) of one level."""
    res = np.uint32(info["resolution"][level])
    x = u * np.float32(res)
    c0 = np.minimum(np.floor(x).astype(np.uint32), res - np.uint32(1))
    fr = (x - c0.astype(np.float32)).astype(np.float32)
    mask = np.uint32((1 << info["log2_table"]) - 1)
    idx = np.zeros((len(u), 8), np.uint32)
    w = np.zeros((len(u), 8), np.float32)
    for c in range(8):
        b = np.array([c & 1, (c >> 1) & 1, (c >> 2) & 1], np.uint32)
        cc = c0 + b
        if info["dense"][level]:
            r1 = res + np.uint32(1)
            idx[:, c] = cc[:, 0] + r1 * (cc[:, 1] + r1 * cc[:, 2])
        else:
            idx[:, c] = (cc[:, 0] ^ (cc[:, 1] * P1) ^ (cc[:, 2] * P2)) & mask
        wx = np.where(b[0], fr[:, 0], np.float32(1.0) - fr[:, 0]).astype(np.float32)
        wy = np.where(b[1], fr[:, 1], np.float32(1.0) - fr[:, 1]).astype(np.float32)
        wz = np.where(b[2], fr[:, 2], np.float32(1.0) - fr[:, 2]).astype(np.float32)
        w[:, c] = (wx * wy) * wz
    return idx, w


def grid_table(info, grid_master):
    """The fp16 working table (as float32), [level][entry][feature]."""
    t = 1 << info["log2_table"]
    return np.asarray(grid_master, np.float32).astype(np.float16).astype(np.float32).reshape(LEVELS, t, 2)


def encode(info, table, pos):
    """Grid features (N, 32) float32."""
    u = unit_position(info, pos)
    out = np.zeros((len(u), 2 * LEVELS), np.float32)
    for l in range(LEVELS):
        idx, w = level_corners(info, l, u)
        f0 = np.zeros(len(u), np.float32)
        f1 = np.zeros(len(u), np.float32)
        for c in range(8):
            v = table[l, idx[:, c]]
            f0 = (f0 + w[:, c] * v[:, 0]).astype(np.float32)
            f1 = (f1 + w[:, c] * v[:, 1]).astype(np.float32)
        out[:, 2 * l] = f0
        out[:, 2 * l + 1] = f1
    return out


def inputs(info, table, rec):
    """Network inputs (N, 64) float32 before bf16 rounding."""
    x = np.zeros((len(rec), IN), np.float32)
    x[:, :32] = encode(info, table, rec["pos"])
    h = np.float32(0.5)
    x[:, 32:35] = rec["dir"] * h + h
    x[:, 35:38] = rec["normal"] * h + h
    x[:, 38:41] = rec["albedo"]
    x[:, 41] = rec["roughness"]
    for t in range(4):
        x[:, 42 + t] = (rec["type"] == t).astype(np.float32)
    x[:, 46] = 1.0
    return x


def forward(mlp, x_in):
    """(activations [x0..x5] as float64 bf16 values, y (N, 16) float64)."""
    ws = [bf16(w).astype(np.float64) for w in unpack_mlp(mlp)]
    xs = [bf16(x_in).astype(np.float64)]
    for l in range(HIDDEN):
        z = xs[-1] @ ws[l]
        xs.append(bf16(np.maximum(z, 0.0).astype(np.float32)).astype(np.float64))
    return xs, xs[-1] @ ws[HIDDEN]


def loss_and_grad(y, target):
    """Relative L2 loss and dL/dy (N, 16) float64 (zeros past RGB)."""
    n = len(y)
    d = y[:, :3] - target
    den = y[:, :3] ** 2 + 0.01
    loss = float(np.sum(d * d / den) / (3 * n))
    g = np.zeros_like(y)
    g[:, :3] = 2 * d / den / (3 * n)
    return loss, g


def backward(mlp, xs, gy):
    """(weight gradient vector (21504,), dx0 (N, 64)) with bf16 rounding of
    each layer's output gradient, as the kernels do."""
    ws = [bf16(w).astype(np.float64) for w in unpack_mlp(mlp)]
    grads = [None] * (HIDDEN + 1)
    dy = bf16(gy.astype(np.float32)).astype(np.float64)
    for l in range(HIDDEN, -1, -1):
        grads[l] = xs[l].T @ dy
        dx = dy @ ws[l].T
        if l > 0:
            dy = bf16((dx * (xs[l] > 0)).astype(np.float32)).astype(np.float64)
    return np.concatenate([g.reshape(-1) for g in grads]), dx


def grid_gradient(info, pos, dx0):
    """Dense grid gradient [level][entry][feature] from d(inputs 0..31)."""
    t = 1 << info["log2_table"]
    g = np.zeros((LEVELS, t, 2), np.float64)
    u = unit_position(info, pos)
    for l in range(LEVELS):
        idx, w = level_corners(info, l, u)
        for c in range(8):
            np.add.at(g[l, :, 0], idx[:, c], w[:, c].astype(np.float64) * dx0[:, 2 * l])
            np.add.at(g[l, :, 1], idx[:, c], w[:, c].astype(np.float64) * dx0[:, 2 * l + 1])
    return g.reshape(-1)


def adam(p, g, m, v, t, lr, b1, b2, eps, skip_zero=False):
    """One Adam step in float64; returns (p, m, v). With skip_zero, entries
    whose gradient is exactly 0 keep p, m and v (the hash grid's rule)."""
    m1 = b1 * m + (1 - b1) * g
    v1 = b2 * v + (1 - b2) * g * g
    p1 = p - lr * (m1 / (1 - b1 ** t)) / (np.sqrt(v1 / (1 - b2 ** t)) + eps)
    if not skip_zero:
        return p1, m1, v1
    keep = np.asarray(g) == 0
    return (np.where(keep, p, p1), np.where(keep, m, m1), np.where(keep, v, v1))


def torch_reference(mlp, x_in, target, straight_through=True):
    """torch float64 autograd of the same network and loss: returns (loss,
    weight gradient vector, d inputs). With straight_through the forward
    rounds like the kernels and the backward passes gradients through the
    rounding unchanged (no rounding of gradients)."""
    import torch

    def rnd(t):
        if not straight_through:
            return t
        r = torch.from_numpy(bf16(t.detach().numpy().astype(np.float32)).astype(np.float64))
        return t + (r - t).detach()

    ws = [torch.tensor(bf16(w).astype(np.float64), requires_grad=True) for w in unpack_mlp(mlp)]
    x = torch.tensor(np.asarray(x_in, np.float64), requires_grad=True)
    h = rnd(x)
    for l in range(HIDDEN):
        h = rnd(torch.relu(h @ ws[l]))
    y = h @ ws[HIDDEN]
    tt = torch.tensor(np.asarray(target, np.float64))
    d = y[:, :3] - tt
    den = (y[:, :3] ** 2 + 0.01).detach()
    loss = (d * d / den).sum() / (3 * len(tt))
    loss.backward()
    g = np.concatenate([w.grad.numpy().reshape(-1) for w in ws])
    return loss.item(), g, x.grad.numpy()
