"""GPU parity of the ORT custom ops (Correlation, Warp) through the C ABI
against the CPU oracle (correlation.cc / warp.cc restatements)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# pyramid levels of a 1920x1088 frame (PWCNet-light feature extractor)
LEVELS = [(32, 272, 480), (64, 136, 240), (96, 68, 120), (128, 34, 60), (196, 17, 30)]


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _smooth_flow(rng, H, W):
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    f = rng.uniform(-4, 4, (2, H, W))
    for k in range(3):
        ax, ay, ph = rng.uniform(0.5, 3, 3)
        f[0] += 2 / 3 * np.sin(2 * np.pi * (ax * xs / W + ay * ys / H) + ph)
        f[1] += 2 / 3 * np.cos(2 * np.pi * (ay * xs / W + ax * ys / H) + ph)
    return f.astype(np.float32)


@pytest.mark.parametrize("C,H,W", LEVELS)
def test_correlation_f32_levels_bit_exact(vsc, orc, cuda, C, H, W):
    print(f"corr f32 C={C} H={H} W={W}", flush=True)
    rng = np.random.default_rng(C)
    f1 = rng.standard_normal((1, C, H, W)).astype(np.float32)
    f2 = rng.standard_normal((1, C, H, W)).astype(np.float32)
    got = vsc.correlation(_t(f1, cuda), _t(f2, cuda), 4, 0).cpu().numpy()
    ref = orc.correlation(f1, f2, 4)
    assert got.shape == (1, 9, 9, H, W)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("md", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("shape", [(2, 3, 5, 7), (1, 1, 1, 1), (1, 9, 70, 3), (1, 17, 4, 131)])
def test_correlation_f32_shapes(vsc, orc, cuda, md, shape):
    rng = np.random.default_rng(md * 100 + sum(shape))
    f1 = rng.standard_normal(shape).astype(np.float32)
    f2 = rng.standard_normal(shape).astype(np.float32)
    got = vsc.correlation(_t(f1, cuda), _t(f2, cuda), md, 0).cpu().numpy()
    assert np.array_equal(got, orc.correlation(f1, f2, md))


@pytest.mark.parametrize("md", [1, 4])
def test_correlation_legacy(vsc, orc, cuda, md):
    rng = np.random.default_rng(7)
    f1 = rng.standard_normal((2, 70, 19, 33)).astype(np.float32)
    f2 = rng.standard_normal((2, 70, 19, 33)).astype(np.float32)
    got = vsc.correlation(_t(f1, cuda), _t(f2, cuda), md, 1).cpu().numpy()
    ref = orc.correlation(f1, f2, md, legacy=True)
    assert got.shape == ref.shape == (2, (2 * md + 1) ** 2, 19, 33)
    bound = orc.correlation(np.abs(f1), np.abs(f2), md, legacy=True)
    assert np.all(np.abs(got - ref) <= 2 * 70 * 2 ** -24 * bound + 1e-30)


def test_correlation_f64(vsc, orc, cuda):
    rng = np.random.default_rng(8)
    f1 = rng.standard_normal((1, 12, 9, 14))
    f2 = rng.standard_normal((1, 12, 9, 14))
    got = vsc.correlation(_t(f1, cuda), _t(f2, cuda), 4, 0).cpu().numpy()
    np.testing.assert_allclose(got, orc.correlation(f1, f2, 4, dtype=np.float64), rtol=1e-13, atol=1e-13)


def test_correlation_rejects_bad_attributes(vsc, cuda):
    import torch
    x = torch.zeros((1, 2, 3, 3), device=cuda)
    with pytest.raises(vsc.VscError):
        vsc.correlation(x, x, 5, 1)         # legacy pads by 4
    with pytest.raises(vsc.VscError):
        vsc.correlation(x.int(), x.int(), 4, 0)


def _warp_check(got, ref, mask_near):
    diff = np.abs(got - ref)
    tol = 1e-5 * np.maximum(1.0, np.abs(ref))
    bad = (diff > tol) & ~mask_near
    assert not bad.any(), (diff[bad].max(), bad.sum())


def _mask_near(flow, H, W):
    """Pixels whose in-bounds bilinear mass is within 1e-6 of 0.999 (the
    FP32 weights may decide them differently from the double weights)."""
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    xf = xs + flow[:, 0].astype(np.float64)
    yf = ys + flow[:, 1].astype(np.float64)
    x0, y0 = np.floor(xf), np.floor(yf)
    a, b = xf - x0, yf - y0
    m = np.zeros_like(xf)
    for dx, dy, w in ((0, 0, (1 - a) * (1 - b)), (1, 0, a * (1 - b)), (0, 1, (1 - a) * b), (1, 1, a * b)):
        ok = (x0 + dx >= 0) & (x0 + dx <= W - 1) & (y0 + dy >= 0) & (y0 + dy <= H - 1)
        m += w * ok
    return np.abs(m - 0.999) < 1e-6


@pytest.mark.parametrize("C,H,W", LEVELS)
def test_warp_f32_levels(vsc, orc, cuda, C, H, W):
    rng = np.random.default_rng(C + 1)
    x = rng.standard_normal((1, C, H, W)).astype(np.float32)
    fl = _smooth_flow(rng, H, W)[None]
    got = vsc.warp(_t(x, cuda), _t(fl, cuda)).cpu().numpy()
    ref = orc.warp(x, fl)
    _warp_check(got, ref, _mask_near(fl, H, W)[:, None])


def test_warp_f32_batch_and_edges(vsc, orc, cuda):
    rng = np.random.default_rng(3)
    for shape in [(3, 5, 7, 9), (1, 1, 1, 1), (2, 2, 1, 17), (1, 3, 23, 1)]:
        N, C, H, W = shape
        x = rng.standard_normal(shape).astype(np.float32)
        fl = rng.uniform(-2, 2, (N, 2, H, W)).astype(np.float32)
        fl[:, :, :1, :1] = np.round(fl[:, :, :1, :1])   # integer displacements too
        got = vsc.warp(_t(x, cuda), _t(fl, cuda)).cpu().numpy()
        _warp_check(got, orc.warp(x, fl), _mask_near(fl, H, W)[:, None])
    # zero flow is the identity, bit for bit
    x = rng.standard_normal((1, 4, 10, 12)).astype(np.float32)
    z = np.zeros((1, 2, 10, 12), np.float32)
    assert np.array_equal(vsc.warp(_t(x, cuda), _t(z, cuda)).cpu().numpy(), x)


def test_warp_f64(vsc, orc, cuda):
    rng = np.random.default_rng(4)
    x = rng.standard_normal((1, 6, 15, 21))
    fl = rng.uniform(-3, 3, (1, 2, 15, 21))
    got = vsc.warp(_t(x, cuda), _t(fl, cuda)).cpu().numpy()
    np.testing.assert_allclose(got, orc.warp(x, fl, dtype=np.float64), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("shape,legacy", [((1, 8, 136, 240), 0), ((1, 8, 136, 240), 1), ((1, 5, 17, 30), 0)])
def test_correlation_f32_unaligned_output(vsc, cuda, shape, legacy):
    """The FP32 kernel writes 4 outputs of a plane row with one 16-byte store
    when the output is 16-byte aligned and W a multiple of 4; an output
    4 bytes off that alignment takes the per-pixel stores, with the same bits
    (the 240-pixel shape uses the 4-pixel tile, the 30-pixel one the
    1-pixel tile)."""
    import torch
    rng = np.random.default_rng(sum(shape) + legacy)
    a = _t(rng.standard_normal(shape).astype(np.float32), cuda)
    b = _t(rng.standard_normal(shape).astype(np.float32), cuda)
    ref = vsc.correlation(a, b, 4, legacy)
    buf = torch.empty(ref.numel() + 1, dtype=torch.float32, device=cuda)
    out = buf[1:].view(ref.shape)
    assert out.data_ptr() % 16 == 4
    vsc.correlation(a, b, 4, legacy, out=out)
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
