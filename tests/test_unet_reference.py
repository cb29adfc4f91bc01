"""Pins of the PyTorch reference of the learned preconditioner's U-net (oracle/unet_ref.py):
each layer type against a direct loop over its definition, the full network's autograd
gradient against central differences, and the positive homogeneity N(c x) = c N(x) of a
bias-free ReLU network.  CPU only."""
import numpy as np
import pytest
import torch

import oracle.unet_ref as U

rng = np.random.default_rng(7)


def _rand(*s):
    return rng.standard_normal(s)


def _to_t(x):  # (nz, ny, nx, C) -> (1, C, nz, ny, nx)
    return torch.tensor(x).permute(3, 0, 1, 2)[None]


def _from_t(y):
    return y[0].permute(1, 2, 3, 0).numpy()


def test_c3_matches_direct_loop():
    nz, ny, nx, ci, co = 4, 5, 3, 2, 3
    x = _rand(nz, ny, nx, ci)
    W = _rand(27, ci, co)
    y = _from_t(U.conv_c3(_to_t(x), torch.tensor(W)))
    ref = np.zeros((nz, ny, nx, co))
    for z in range(nz):
        for yy in range(ny):
            for xx in range(nx):
                for t in range(27):
                    dx, dy, dz = t % 3 - 1, t // 3 % 3 - 1, t // 9 - 1
                    a, b, c = xx + dx, yy + dy, z + dz
                    if 0 <= a < nx and 0 <= b < ny and 0 <= c < nz:
                        ref[z, yy, xx] += x[c, b, a] @ W[t]
    np.testing.assert_allclose(y, ref, rtol=1e-12, atol=1e-12)


def test_c2s2_matches_direct_loop_on_odd_grid():
    nz, ny, nx, ci, co = 5, 4, 3, 2, 3
    x = _rand(nz, ny, nx, ci)
    W = _rand(8, ci, co)
    y = _from_t(U.conv_c2s2(_to_t(x), torch.tensor(W)))
    cz, cy, cx = (nz + 1) // 2, (ny + 1) // 2, (nx + 1) // 2
    assert y.shape == (cz, cy, cx, co)
    ref = np.zeros((cz, cy, cx, co))
    for z in range(cz):
        for yy in range(cy):
            for xx in range(cx):
                for t in range(8):
                    a, b, c = 2 * xx + (t & 1), 2 * yy + (t >> 1 & 1), 2 * z + (t >> 2)
                    if a < nx and b < ny and c < nz:
                        ref[z, yy, xx] += x[c, b, a] @ W[t]
    np.testing.assert_allclose(y, ref, rtol=1e-12, atol=1e-12)


def test_upc3_and_c1_match_direct_loops():
    fz, fy, fx, ci, co = 5, 3, 4, 2, 2
    cz, cy, cx = 3, 2, 2
    xc = _rand(cz, cy, cx, ci)
    W = _rand(27, ci, co)
    up = U.upsample(_to_t(xc), (fz, fy, fx))
    assert tuple(up.shape[2:]) == (fz, fy, fx)
    upn = _from_t(up)
    for z in range(fz):
        for yy in range(fy):
            for xx in range(fx):
                np.testing.assert_array_equal(upn[z, yy, xx], xc[z // 2, yy // 2, xx // 2])
    W1 = _rand(1, ci, co)
    y1 = _from_t(U.conv_c1(_to_t(upn), torch.tensor(W1)))
    np.testing.assert_allclose(y1, upn @ W1[0], rtol=1e-12)


def _small_params(channels, seed=3):
    r = np.random.default_rng(seed)
    ps = []
    for mode, cin, cout, *_ in U.layer_spec(channels):
        n = U.ntaps(mode)
        w = r.standard_normal((n, cin, cout)) * np.sqrt(2.0 / (n * cin))
        if mode == U.C3 and cin == 16 and len(ps) == 0:
            w[:, 3:, :] = 0.0
        ps.append(w)
    ps[-1][:, :, 3:] = 0.0
    return np.concatenate([p.reshape(-1) for p in ps])


def test_network_gradient_matches_central_differences():
    grid = (5, 4, 6)  # (nz, ny, nx)
    ch = (16, 16, 16, 16)
    w = _small_params(ch)
    n = int(np.prod(grid))
    x = rng.standard_normal((n, 3))
    tgt = rng.standard_normal((n, 3))
    loss, g = U.loss_and_grad(w, x, tgt, grid, ch, bf16=False, dtype=torch.float64)

    def f(wv):
        ps = [torch.tensor(p) for p in U.split_params(wv, ch)]
        return float(U.cos_loss(U.forward(ps, x, grid, ch, bf16=False, dtype=torch.float64), tgt))

    assert abs(f(w) - loss) < 1e-14
    idx = rng.choice(np.nonzero(np.abs(g) > 1e-8)[0], size=25, replace=False)
    for k in idx:
        h = 1e-6 * max(1.0, abs(w[k]))
        wp, wm = w.copy(), w.copy()
        wp[k] += h
        wm[k] -= h
        fd = (f(wp) - f(wm)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])) + 1e-8, (k, fd, g[k])


def test_network_is_positively_homogeneous_and_uses_three_channels():
    grid = (4, 5, 3)
    ch = (16, 32, 16, 16)
    w = _small_params(ch, seed=5)
    ps = [torch.tensor(p) for p in U.split_params(w, ch)]
    x = rng.standard_normal((int(np.prod(grid)), 3))
    y1 = U.forward(ps, x, grid, ch, bf16=False, dtype=torch.float64)
    y2 = U.forward(ps, 3.5 * x, grid, ch, bf16=False, dtype=torch.float64)
    np.testing.assert_allclose(y2.numpy(), 3.5 * y1.numpy(), rtol=1e-12, atol=1e-14)
    assert y1.shape == (int(np.prod(grid)), 3)


@pytest.mark.parametrize("ch", [(16, 32, 64, 64)])
def test_parameter_count(ch):
    n = sum(U.ntaps(m) * ci * co for m, ci, co, *_ in U.layer_spec(ch))
    assert n == 27 * 16 * 16 * 2 + 8 * 16 * 32 + 27 * 32 * 32 + 8 * 32 * 64 + 27 * 64 * 64 + 8 * 64 * 64 + \
        27 * 64 * 64 * 2 + 27 * 64 * 32 + 27 * 32 * 16 + 16 * 16
