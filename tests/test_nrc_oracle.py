"""The neural radiance cache's CPU oracle, pinned without a GPU: the grid
encoding by closed forms (dense levels reproduce linear functions exactly,
weights are a partition of unity, dense indices are a bijection), the
analytic backward against torch float64 autograd and finite differences,
and the Adam restatement against torch.optim.Adam (and its zero-skipping
form, used for the hash grid, against torch.optim.SparseAdam)."""
import os
import sys

import numpy as np
import pytest
import torch

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import nrc_oracle as R  # noqa: E402
from zig_ornament_amd.nrc import RECORD_DTYPE  # noqa: E402


def make_info(log2_table=12, base=4, top=64):
    growth = np.exp((np.log(top) - np.log(base)) / 15)
    res = [max(1, int(np.floor(base * growth ** l))) for l in range(16)]
    return dict(levels=16, log2_table=log2_table, resolution=res,
                dense=[1 if (r + 1) ** 3 <= (1 << log2_table) else 0 for r in res],
                bounds_lo=[-1.0, -2.0, 0.5], inv_extent=[0.5, 0.25, 1.0])


def test_dense_level_indices_are_a_bijection():
    info = make_info()
    l = 0
    r = info["resolution"][l]
    assert info["dense"][l]
    # the lower corner of every cell plus the far faces cover the lattice once
    g = np.stack(np.meshgrid(*[np.arange(r) + 0.5] * 3, indexing="ij"), -1).reshape(-1, 3) / r
    idx, _ = R.level_corners(info, l, g.astype(np.float32))
    assert len(np.unique(idx)) == (r + 1) ** 3
    assert idx.max() == (r + 1) ** 3 - 1


def test_weights_are_a_partition_of_unity():
    info = make_info()
    u = np.random.default_rng(0).random((500, 3)).astype(np.float32)
    u[:5] = [[0, 0, 0], [1, 1, 1], [1, 0, 0.5], [0.999999, 0.5, 0], [0.25, 0.5, 0.75]]
    for l in (0, 7, 15):
        _, w = R.level_corners(info, l, u)
        assert np.all(w >= 0)
        assert np.allclose(w.sum(1), 1.0, atol=2e-6)


def test_dense_levels_reproduce_linear_functions():
    """A table holding a*x + b*y + c*z + d at every lattice point makes the
    trilinear feature equal that function at any point of the level."""
    info = make_info()
    t = 1 << info["log2_table"]
    table = np.zeros((16, t, 2), np.float32)
    a = np.array([0.25, -0.5, 0.125], np.float32)
    for l in range(16):
        if not info["dense"][l]:
            continue
        r1 = info["resolution"][l] + 1
        cz, cy, cx = np.meshgrid(np.arange(r1), np.arange(r1), np.arange(r1), indexing="ij")
        idx = (cx + r1 * (cy + r1 * cz)).reshape(-1)
        table[l, idx, 0] = (a[0] * cx + a[1] * cy + a[2] * cz).reshape(-1) + 1.0
        table[l, idx, 1] = cx.reshape(-1).astype(np.float32)
    lo = np.asarray(info["bounds_lo"], np.float32)
    ext = 1 / np.asarray(info["inv_extent"], np.float32)
    u = np.random.default_rng(1).random((300, 3)).astype(np.float32)
    pos = lo + u * ext
    f = R.encode(info, table, pos)
    uu = R.unit_position(info, pos)
    for l in range(16):
        if not info["dense"][l]:
            continue
        x = uu * info["resolution"][l]
        assert np.allclose(f[:, 2 * l], x @ a + 1.0, atol=1e-4), l
        assert np.allclose(f[:, 2 * l + 1], x[:, 0], atol=1e-4), l


def test_hashed_levels_use_the_published_primes():
    """Instant-NGP's spatial hash (pi = 1, 2654435761, 805459861): a table that
    is 1 at one entry only lights exactly the cells whose corners hash to it
    (found by brute force over the level's lattice)."""
    info = make_info()
    l = max(i for i in range(16) if not info["dense"][i])
    r = info["resolution"][l]
    t = 1 << info["log2_table"]
    target = 1234
    table = np.zeros((16, t, 2), np.float32)
    table[l, target, 0] = 1.0
    hit = set()
    for cx in range(r + 1):
        for cy in range(r + 1):
            for cz in range(r + 1):
                h = (cx ^ ((cy * 2654435761) & 0xFFFFFFFF) ^ ((cz * 805459861) & 0xFFFFFFFF)) % t
                if h == target:
                    hit.add((cx, cy, cz))
    assert hit
    cx, cy, cz = sorted(hit)[0]
    # a point just inside the cell whose upper corner is (cx, cy, cz)
    c = np.array([max(cx - 1, 0), max(cy - 1, 0), max(cz - 1, 0)], np.float32)
    u = ((c + 0.5) / r).astype(np.float32)[None]
    lo = np.asarray(info["bounds_lo"], np.float32)
    ext = 1 / np.asarray(info["inv_extent"], np.float32)
    f = R.encode(info, table, lo + u * ext)
    assert f[0, 2 * l] > 0


def records(n, seed):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, RECORD_DTYPE)
    r["pos"] = rng.uniform(-1, 2, (n, 3))
    d = rng.normal(size=(n, 3))
    r["dir"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    nn = rng.normal(size=(n, 3))
    r["normal"] = nn / np.linalg.norm(nn, axis=1, keepdims=True)
    r["albedo"] = rng.random((n, 3))
    r["roughness"] = rng.random(n) * 0.5
    r["type"] = rng.integers(0, 3, n)
    r["weight"] = rng.random((n, 3)) * 2
    return r


def test_analytic_backward_matches_autograd_and_finite_differences():
    info = make_info()
    rng = np.random.default_rng(2)
    mlp = (rng.normal(size=R.MLP_PARAMS) * 0.2).astype(np.float32)
    table = (rng.normal(size=(16, 1 << info["log2_table"], 2)) * 0.5).astype(np.float32)
    rec = records(96, 3)
    x = R.inputs(info, table, rec)
    xs, y = R.forward(mlp, x)
    loss, gy = R.loss_and_grad(y, rec["weight"].astype(np.float64))
    g, dx0 = R.backward(mlp, xs, gy)
    # torch autograd with the same forward rounding, gradients unrounded
    tl, tg, tdx = R.torch_reference(mlp, x, rec["weight"])
    assert abs(tl - loss) < 1e-12 * max(1.0, abs(loss))
    assert np.linalg.norm(g - tg) < 2e-2 * np.linalg.norm(tg)
    assert np.linalg.norm(dx0 - tdx) < 2e-2 * np.linalg.norm(tdx)
    for l in range(6):  # every layer's block agrees in direction
        a = g[l * 4096:(l + 1) * 4096] if l < 5 else g[5 * 4096:]
        b = tg[l * 4096:(l + 1) * 4096] if l < 5 else tg[5 * 4096:]
        assert a @ b / (np.linalg.norm(a) * np.linalg.norm(b)) > 0.999
    # the float64 network without rounding: autograd equals finite differences
    _, g64, _ = R.torch_reference(mlp, x, rec["weight"], straight_through=False)
    base = R.bf16(mlp).astype(np.float64)
    y0 = unrounded_output(x, base)
    for j in rng.choice(R.MLP_PARAMS, 12, replace=False):
        e = 1e-6
        p1, p0 = base.copy(), base.copy()
        p1[j] += e
        p0[j] -= e
        fd = (f_exact(x, rec, p1, y0) - f_exact(x, rec, p0, y0)) / (2 * e)
        assert abs(fd - g64[j]) <= 1e-7 + 1e-5 * abs(g64[j]), j


def unrounded_output(x, p):
    ws = [p[l * 4096:(l + 1) * 4096].reshape(64, 64) for l in range(5)]
    ws.append(p[5 * 4096:].reshape(64, 16))
    h = x.astype(np.float64)
    for l in range(5):
        h = np.maximum(h @ ws[l], 0)
    return h @ ws[5]


def f_exact(x, rec, p, y_ref):
    """Loss of the unrounded float64 network with weights p (denominator from
    y_ref, held constant)."""
    yy = unrounded_output(x, p)
    d = yy[:, :3] - rec["weight"]
    den = y_ref[:, :3] ** 2 + 0.01
    return np.sum(d * d / den) / (3 * len(d))


def test_grid_gradient_matches_autograd_of_the_encoding():
    info = make_info()
    rng = np.random.default_rng(4)
    t = 1 << info["log2_table"]
    table = rng.normal(size=(16, t, 2)).astype(np.float32)
    rec = records(64, 5)
    coef = rng.normal(size=(64, 32))
    g = R.grid_gradient(info, rec["pos"], coef)
    # torch: features as gathers of the table, loss = sum(coef * features)
    tt = torch.tensor(table.astype(np.float64), requires_grad=True)
    u = R.unit_position(info, rec["pos"])
    feats = []
    for l in range(16):
        idx, w = R.level_corners(info, l, u)
        v = tt[l][torch.from_numpy(idx.astype(np.int64))]  # (N, 8, 2)
        feats.append((torch.from_numpy(w.astype(np.float64))[:, :, None] * v).sum(1))
    f = torch.cat(feats, 1)
    (f * torch.from_numpy(coef)).sum().backward()
    assert np.allclose(g, tt.grad.numpy().reshape(-1), atol=1e-12)


def test_adam_matches_torch():
    rng = np.random.default_rng(6)
    p0 = rng.normal(size=50)
    pt = torch.tensor(p0.copy(), requires_grad=True)
    opt = torch.optim.Adam([pt], lr=1e-2, betas=(0.9, 0.99), eps=1e-15)
    p, m, v = p0.copy(), np.zeros(50), np.zeros(50)
    for t in range(1, 6):
        g = rng.normal(size=50)
        pt.grad = torch.tensor(g)
        opt.step()
        p, m, v = R.adam(p, g, m, v, t, 1e-2, 0.9, 0.99, 1e-15)
    assert np.allclose(p, pt.detach().numpy(), rtol=0, atol=1e-12)


def test_skipping_adam_matches_sparse_adam():
    """Entries with an exactly-zero gradient keep value and moments while the
    step count advances: torch.optim.SparseAdam's update of the entries a
    sparse gradient holds. (SparseAdam adds eps before the bias correction
    of the denominator; at eps = 1e-15 against moments ~1e-2 that is far
    below the tolerance.)"""
    rng = np.random.default_rng(7)
    n = 64
    p0 = rng.normal(size=n)
    pt = torch.tensor(p0.copy(), requires_grad=True)
    opt = torch.optim.SparseAdam([pt], lr=1e-2, betas=(0.9, 0.99), eps=1e-15)
    p, m, v = p0.copy(), np.zeros(n), np.zeros(n)
    skipped_with_moments = 0
    for t in range(1, 8):
        g = rng.normal(size=n)
        g[rng.random(n) < 0.5] = 0.0
        nz = np.nonzero(g)[0]
        skipped_with_moments += int(np.sum((g == 0) & (m != 0)))
        pt.grad = torch.sparse_coo_tensor(torch.from_numpy(nz[None]), torch.from_numpy(g[nz]), (n,))
        opt.step()
        p, m, v = R.adam(p, g, m, v, t, 1e-2, 0.9, 0.99, 1e-15, skip_zero=True)
        assert np.allclose(p, pt.detach().numpy(), rtol=1e-12, atol=1e-12), t
    assert skipped_with_moments > 0  # the skip was exercised where dense Adam would move
    # and the dense form differs there
    q, _, _ = R.adam(p0, np.where(np.arange(n) % 2, 1.0, 0.0), np.ones(n), np.ones(n), 2, 1e-2, 0.9, 0.99, 1e-15)
    assert not np.allclose(q[::2], p0[::2])
