"""3x3 NHWC convolution on MFMA (the flow network's Conv2d layers) against a
float64 torch reference on the CPU of the same (dtype-rounded) operands.

Bounds: FP32 accumulation of exact products in some order:
|err| <= 2 * K * 2^-24 * sum |w * x| (K = 9 * Cin terms); a 16-bit output adds
one rounding of the result, 2^-8 (BF16) / 2^-11 (FP16) relative.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DT = {"f32": "float32", "bf16": "bfloat16", "f16": "float16"}
OUT_U = {"f32": 0.0, "bf16": 2.0 ** -8, "f16": 2.0 ** -11}


def _ref(x, w, b, stride, dil, pad, act, add=None):
    import torch
    xt = x.permute(0, 3, 1, 2).double()
    wt = w.permute(0, 3, 1, 2).double()
    y = torch.nn.functional.conv2d(xt, wt, b.double() if b is not None else None, stride=stride, padding=pad,
                                   dilation=dil)
    if act == 1:
        y = torch.nn.functional.leaky_relu(y, 0.1)
    elif act == 2:
        y = torch.relu(y)
    y = y.permute(0, 2, 3, 1)
    if add is not None:
        y = y + add.double()
    yabs = torch.nn.functional.conv2d(xt.abs(), wt.abs(), b.double().abs() if b is not None else None,
                                      stride=stride, padding=pad, dilation=dil).permute(0, 2, 3, 1)
    return y, yabs


CASES = [
    # N, H, W, Cin, Cout, stride, dil, act
    (1, 20, 33, 8, 16, 1, 1, 1),
    (2, 17, 30, 16, 32, 1, 1, 0),
    (1, 34, 60, 96, 64, 1, 1, 1),
    (1, 37, 41, 40, 96, 2, 1, 1),
    (1, 64, 48, 32, 128, 2, 1, 2),
    (1, 40, 72, 40, 128, 1, 2, 1),
    (1, 40, 72, 128, 96, 1, 4, 1),
    (1, 45, 70, 64, 32, 1, 8, 1),
    (1, 50, 66, 32, 16, 1, 16, 1),
    (1, 30, 50, 160, 2, 1, 1, 0),
    (1, 136, 240, 256, 96, 1, 1, 1),
    # 16-bit layers with 8 / 16 input channels stage narrow chunks and pack
    # 4 / 2 taps into one MFMA K-step (stride 2 and dilation included)
    (1, 37, 41, 8, 32, 2, 1, 1),
    (1, 30, 50, 16, 16, 1, 2, 1),
    (2, 19, 45, 8, 16, 1, 1, 2),
]


@pytest.mark.parametrize("kind", ["f32", "bf16", "f16"])
@pytest.mark.parametrize("case", CASES)
def test_conv3x3(vsc, cuda, kind, case):
    import torch
    N, H, W, cin, cout, stride, dil, act = case
    print(f"conv {kind} {case}", flush=True)
    g = torch.Generator().manual_seed(hash(case) % 1000 + len(kind))
    dt = getattr(torch, DT[kind])
    x = torch.randn((N, H, W, cin), generator=g).to(dt)
    w = (torch.randn((cout, 3, 3, cin), generator=g) / (3 * cin ** 0.5)).to(dt)
    b = torch.randn(cout, generator=g).float()
    pad = dil
    out = vsc.conv3x3_nhwc(x.to(cuda), w.to(cuda), b.to(cuda), stride=stride, dilation=dil, pad=pad, act=act,
                           out_dtype=torch.float32).cpu().double()
    ref, rabs = _ref(x, w, b, stride, dil, pad, act)
    assert out.shape == ref.shape
    K = 9 * cin + 1
    err = (out - ref).abs()
    bound = 2 * K * 2.0 ** -24 * rabs + 1e-30
    print(f"  max err/sum|wx| = {float((err / (rabs + 1e-30)).max()):.3e}")
    assert bool((err <= bound).all())
    # 16-bit output: one more rounding
    if kind != "f32":
        out16 = vsc.conv3x3_nhwc(x.to(cuda), w.to(cuda), b.to(cuda), stride=stride, dilation=dil, pad=pad,
                                 act=act).cpu().double()
        assert bool(((out16 - ref).abs() <= OUT_U[kind] * ref.abs() + 1.01 * bound + 2.0 ** -24).all())


@pytest.mark.parametrize("kind", ["f32", "bf16"])
def test_conv3x3_slices_and_residual(vsc, cuda, kind):
    """Input and output are channel slices of wider pixels (the decoders'
    DenseNet-style concatenation buffers); the residual is added after the
    activation; channels outside the output slice are untouched."""
    import torch
    g = torch.Generator().manual_seed(5)
    dt = getattr(torch, DT[kind])
    N, H, W = 1, 23, 37
    buf = torch.randn((N, H, W, 64), generator=g).to(dt)
    w = (torch.randn((20, 3, 3, 24), generator=g) / 10).to(dt)
    b = torch.randn(20, generator=g).float()
    out = torch.full((N, H, W, 48), 7.0, dtype=torch.float32)
    add = torch.randn((N, H, W, 28), generator=g).float()
    got = vsc.conv3x3_nhwc(buf.to(cuda), w.to(cuda), b.to(cuda), act=1, out=out.to(cuda), out_dtype=torch.float32,
                           add=add.to(cuda), in_slice=(16, 24), out_slice=(12, 20), add_slice=(8, 20)).cpu()
    ref, rabs = _ref(buf[..., 16:40], w, b, 1, 1, 1, 1, add=add[..., 8:28])
    err = (got[..., 12:32].double() - ref).abs()
    assert bool((err <= 2 * 217 * 2.0 ** -24 * (rabs + add[..., 8:28].double().abs()) + 1e-30).all())
    assert bool((got[..., :12] == 7.0).all()) and bool((got[..., 32:] == 7.0).all())


def test_conv3x3_rejects_bad_slices(vsc, cuda):
    import torch
    x = torch.zeros((1, 8, 8, 12), device=cuda)
    w = torch.zeros((4, 3, 3, 6), device=cuda)
    with pytest.raises(vsc.VscError):
        vsc.conv3x3_nhwc(x, w, in_slice=(0, 6))          # not whole 16-byte vectors
    with pytest.raises(vsc.VscError):
        vsc.conv3x3_nhwc(x, torch.zeros((4, 3, 3, 12), device=cuda), stride=2, dilation=2)
