"""Pins the oracle's ORT Warp (warp.cc:71-134) by closed forms and by an
independent formulation: torch.nn.functional.grid_sample (align_corners=True,
zero padding) times the 0.999 in-bounds mask, the definition the reference's
own test uses (model-conversion/test.py:99-136)."""
import numpy as np
import pytest


def test_zero_flow_identity(orc, rng):
    x = rng.standard_normal((2, 3, 7, 9)).astype(np.float32)
    out = orc.warp(x, np.zeros((2, 2, 7, 9), np.float32))
    assert np.array_equal(out, x)


def test_integer_flow_shift_with_zero_padding(orc, rng):
    x = rng.standard_normal((1, 2, 8, 10)).astype(np.float32)
    dx, dy = 3, -2
    fl = np.zeros((1, 2, 8, 10), np.float32)
    fl[:, 0] = dx
    fl[:, 1] = dy
    out = orc.warp(x, fl)
    ref = np.zeros_like(x)
    for yy in range(8):
        for xx in range(10):
            sy, sx = yy + dy, xx + dx
            if 0 <= sy < 8 and 0 <= sx < 10:
                ref[:, :, yy, xx] = x[:, :, sy, sx]
    assert np.array_equal(out, ref)


def test_half_pixel_average_and_edge_mask(orc, rng):
    x = rng.standard_normal((1, 1, 4, 6)).astype(np.float32)
    fl = np.zeros((1, 2, 4, 6), np.float32)
    fl[:, 0] = 0.5
    out = orc.warp(x, fl)
    np.testing.assert_allclose(out[0, 0, :, :5], 0.5 * (x[0, 0, :, :5] + x[0, 0, :, 1:]), rtol=1e-6)
    assert np.all(out[0, 0, :, 5] == 0.0)   # half the bilinear mass is outside -> zeroed


def test_matches_grid_sample(orc, rng):
    torch = pytest.importorskip("torch")
    N, C, H, W = 2, 4, 17, 23
    x = rng.standard_normal((N, C, H, W)).astype(np.float32)
    fl = rng.uniform(-4, 4, (N, 2, H, W)).astype(np.float32)
    out = orc.warp(x, fl)
    xs = torch.from_numpy(x).double()
    gx = (torch.arange(W).double()[None, None, :] + torch.from_numpy(fl[:, 0]).double()) * 2 / (W - 1) - 1
    gy = (torch.arange(H).double()[None, :, None] + torch.from_numpy(fl[:, 1]).double()) * 2 / (H - 1) - 1
    grid = torch.stack([gx, gy], dim=-1)
    F = torch.nn.functional
    val = F.grid_sample(xs, grid, mode="bilinear", padding_mode="zeros", align_corners=True)
    mask = F.grid_sample(torch.ones_like(xs), grid, mode="bilinear", padding_mode="zeros", align_corners=True)
    keep = mask > 0.999
    ref = (val * keep).numpy()
    near = (torch.abs(mask - 0.999) < 1e-5).numpy()
    diff = np.abs(out - ref)
    assert np.all(diff[~near] <= 1e-5 * (1 + np.abs(ref[~near])))
    assert near.mean() < 0.01


def test_f32_close_to_f64(orc, rng):
    x = rng.standard_normal((1, 8, 20, 30))
    fl = rng.uniform(-4, 4, (1, 2, 20, 30))
    a = orc.warp(x.astype(np.float32), fl.astype(np.float32))
    b = orc.warp(x, fl.astype(np.float32).astype(np.float64), dtype=np.float64)
    assert np.max(np.abs(a - b)) < 1e-5
