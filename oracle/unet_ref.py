"""PyTorch (CPU, FP32/FP64) reference of the learned preconditioner's U-net.

TEST INFRASTRUCTURE ONLY: the checker for the HIP network kernels
(plasticity_amd/csrc/unet.hip).  Only tests/ may use it.

The network has no counterpart in the reference solver; this module defines it
independently of the HIP code with torch.nn.functional convolutions, from the layer
definitions below (checked against direct loops in tests/test_unet_reference.py):

  C3    y[v][co] = sum_t sum_ci W[t][ci][co] x[v + d(t)][ci], d(t) = (t%3-1, t//3%3-1, t//9-1)
        (x, y, z offsets), zero outside the grid
  C2S2  y[c][co] = sum_t sum_ci W[t][ci][co] x[2c + o(t)][ci], o(t) = (t&1, t>>1&1, t>>2)
  UPC3  C3 applied to the nearest-neighbour upsample up(x)[w] = x[w // 2] on the finer grid
  C1    y[v][co] = sum_ci W[0][ci][co] x[v][ci]

Grids are (nz, ny, nx) node arrays, x fastest; level l+1 has ceil(n/2) nodes per axis.
Weights are the flat FP32 vector of pfem_net_get_params: per layer [tap][c_in][c_out].
With `bf16=True` the weights, the input and every layer output are rounded to BF16 at
the places where the HIP kernels store BF16, so forward results agree to FP32
accumulation-order differences.
"""
from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F

C3, C1, C2S2, UPC3 = 0, 1, 2, 3


def layer_spec(channels=(16, 32, 64, 64)):
    """(mode, cin, cout, level_in, level_out, relu, skip) per layer, as in unet.cpp."""
    c = channels
    return [
        (C3, 16, c[0], 0, 0, 1, -1), (C3, c[0], c[0], 0, 0, 1, -1),
        (C2S2, c[0], c[1], 0, 1, 1, -1), (C3, c[1], c[1], 1, 1, 1, -1),
        (C2S2, c[1], c[2], 1, 2, 1, -1), (C3, c[2], c[2], 2, 2, 1, -1),
        (C2S2, c[2], c[3], 2, 3, 1, -1), (C3, c[3], c[3], 3, 3, 1, -1),
        (UPC3, c[3], c[2], 3, 2, 1, 5), (UPC3, c[2], c[1], 2, 1, 1, 3),
        (UPC3, c[1], c[0], 1, 0, 1, 1), (C1, c[0], 16, 0, 0, 0, -1),
    ]


def ntaps(mode):
    return 27 if mode in (C3, UPC3) else (8 if mode == C2S2 else 1)


def split_params(w, channels=(16, 32, 64, 64)):
    """Flat vector -> list of arrays [taps, cin, cout]."""
    out, off = [], 0
    for mode, cin, cout, *_ in layer_spec(channels):
        n = ntaps(mode) * cin * cout
        out.append(np.asarray(w[off:off + n]).reshape(ntaps(mode), cin, cout))
        off += n
    assert off == len(w)
    return out


def _q(t, bf16):
    return t.to(torch.bfloat16).to(t.dtype) if bf16 else t


def _w3(W):
    """[27][ci][co] -> torch (co, ci, 3, 3, 3) with kernel index (dz+1, dy+1, dx+1)."""
    co, ci = W.shape[2], W.shape[1]
    w = W.reshape(3, 3, 3, ci, co)  # t = (dz+1)*9 + (dy+1)*3 + (dx+1)
    return w.permute(4, 3, 0, 1, 2).contiguous()


def _w2(W):
    """[8][ci][co] -> (co, ci, 2, 2, 2) with kernel index (oz, oy, ox), t = ox + 2 oy + 4 oz."""
    co, ci = W.shape[2], W.shape[1]
    w = W.reshape(2, 2, 2, ci, co)
    return w.permute(4, 3, 0, 1, 2).contiguous()


def conv_c3(h, W):
    return F.conv3d(h, _w3(W), padding=1)


def conv_c2s2(h, W):
    D, H, Wd = h.shape[2:]
    h = F.pad(h, (0, Wd % 2, 0, H % 2, 0, D % 2))  # zero node beyond an odd edge
    return F.conv3d(h, _w2(W), stride=2)


def upsample(h, shape):
    u = h.repeat_interleave(2, dim=2).repeat_interleave(2, dim=3).repeat_interleave(2, dim=4)
    return u[:, :, :shape[0], :shape[1], :shape[2]]


def conv_c1(h, W):
    return F.conv3d(h, W[0].t().reshape(W.shape[2], W.shape[1], 1, 1, 1))


def level_shapes(grid):
    nz, ny, nx = grid
    s = [(nz, ny, nx)]
    for _ in range(3):
        a = s[-1]
        s.append(tuple((n + 1) // 2 for n in a))
    return s


def forward(params, x3, grid, channels=(16, 32, 64, 64), bf16=True, dtype=torch.float32):
    """params: list of [taps, cin, cout] tensors (requires_grad allowed); x3: (nz*ny*nx, 3)
    input, node-major; returns (nvox, 3)."""
    shapes = level_shapes(grid)
    nz, ny, nx = grid
    x = torch.as_tensor(x3, dtype=dtype).reshape(nz, ny, nx, 3).permute(3, 0, 1, 2)[None]
    x = F.pad(x, (0, 0, 0, 0, 0, 0, 0, 13))  # 3 -> 16 channels
    x = _q(x, bf16)
    acts = []
    h = x
    for k, (mode, cin, cout, lin, lout, relu, skip) in enumerate(layer_spec(channels)):
        W = _q(params[k].to(dtype), bf16)
        if mode == C3:
            y = conv_c3(h, W)
        elif mode == C2S2:
            y = conv_c2s2(h, W)
        elif mode == UPC3:
            y = conv_c3(upsample(h, shapes[lout]), W)
        else:
            y = conv_c1(h, W)
        if skip >= 0:
            y = y + acts[skip]
        if relu:
            y = torch.relu(y)
        if k < 11:
            y = _q(y, bf16)
        acts.append(y)
        h = y
    return h[0, :3].permute(1, 2, 3, 0).reshape(-1, 3)


def cos_loss(out, target):
    o = out.reshape(-1)
    z = torch.as_tensor(target, dtype=out.dtype).reshape(-1)
    return 1.0 - (o @ z) / (torch.linalg.norm(o) * torch.linalg.norm(z))


def loss_and_grad(w_flat, x3, target3, grid, channels=(16, 32, 64, 64), bf16=True, dtype=torch.float32):
    """Loss 1 - cos(N(x), target) and its gradient w.r.t. the flat weight vector (with
    the weights and forward activations rounded as on the device when bf16)."""
    ps = [torch.tensor(p, dtype=dtype, requires_grad=True) for p in split_params(w_flat, channels)]
    tgt = torch.as_tensor(target3, dtype=dtype)
    if bf16:
        tgt = _q(tgt, True)
    loss = cos_loss(forward(ps, x3, grid, channels, bf16, dtype), tgt)
    loss.backward()
    g = np.concatenate([p.grad.detach().numpy().reshape(-1) for p in ps])
    return float(loss.detach()), g
