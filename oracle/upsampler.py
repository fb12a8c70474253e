"""ORACLE — test infrastructure only: float64 CPU restatement (torch autograd as
the differentiator) of the learned flow upsampler, used to check the device
forward pass and the device gradients.  Only tests/ may import it.

Input: the reference's kernel_bilinear mapping (flowconsistency.cu:58-74) of
the low-resolution flow, times the scale, plus the frame's luma; three 3x3
convs (3->16 ReLU, 16->16 ReLU, 16->2) added to the bilinear flow.  Loss: mean
Charbonnier end-point error sqrt(|out - gt|^2 + eps^2), eps = 1e-3.
"""
from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F

SHAPES = [(16, 3, 3, 3), (16,), (16, 16, 3, 3), (16,), (2, 16, 3, 3), (2,)]


def unpack(flat):
    out, o = [], 0
    for s in SHAPES:
        n = int(np.prod(s))
        out.append(torch.from_numpy(flat[o:o + n].astype(np.float64).reshape(s)))
        o += n
    assert o == flat.size
    return out


def bilinear_taps(out_n, in_n):
    """kernel_bilinear source index / fraction along one axis (float32 as the device)."""
    o = np.arange(out_n, dtype=np.float32)
    xx = (o * np.float32(in_n)) / np.float32(out_n)
    i0 = np.floor(xx).astype(np.int64)
    fr = (xx - i0.astype(np.float32)).astype(np.float64)
    i1 = np.minimum(i0 + 1, in_n - 1)
    return i0, i1, fr


def prep(lo, rgba, scale):
    """lo [B, lh, lw, 2], rgba [B, H, W, 4] -> x [B, 3, H, W], base [B, 2, H, W] (float64)."""
    B, lh, lw, _ = lo.shape
    H, W = rgba.shape[1:3]
    x0, x1, fx = bilinear_taps(W, lw)
    y0, y1, fy = bilinear_taps(H, lh)
    L = lo.astype(np.float64)
    fx = fx[None, None, :, None]
    fy = fy[None, :, None, None]
    v00 = L[:, y0][:, :, x0]
    v10 = L[:, y0][:, :, x1]
    v01 = L[:, y1][:, :, x0]
    v11 = L[:, y1][:, :, x1]
    base = (v00 * (1 - fx) * (1 - fy) + v10 * fx * (1 - fy) + v01 * (1 - fx) * fy + v11 * fx * fy) * scale
    luma = (0.299 * rgba[..., 0] + 0.587 * rgba[..., 1] + 0.114 * rgba[..., 2]) / 255.0
    x = np.concatenate([base, luma[..., None]], -1)
    return (torch.from_numpy(np.ascontiguousarray(x.transpose(0, 3, 1, 2))),
            torch.from_numpy(np.ascontiguousarray(base.transpose(0, 3, 1, 2))))


def forward(P, x, base):
    w1, b1, w2, b2, w3, b3 = P
    a1 = torch.relu(F.conv2d(x, w1, b1, padding=1))
    a2 = torch.relu(F.conv2d(a1, w2, b2, padding=1))
    return base + F.conv2d(a2, w3, b3, padding=1)


def loss_and_grad(flat, lo, rgba, gt, scale=2.0, eps=1e-3):
    P = [p.clone().requires_grad_(True) for p in unpack(flat)]
    x, base = prep(lo, rgba, scale)
    out = forward(P, x, base)
    g = torch.from_numpy(gt.astype(np.float64)).permute(0, 3, 1, 2)
    rho = torch.sqrt(((out - g) ** 2).sum(1) + eps * eps)
    loss = rho.mean()
    loss.backward()
    return float(loss.detach()), np.concatenate([p.grad.numpy().ravel() for p in P])
