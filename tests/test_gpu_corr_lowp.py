"""Reduced-precision Correlation on MFMA (BF16 / FP16 / FP8-E4M3, FP32
accumulation) against the FP64 oracle.

Tolerances (stated per element, normalised by B = sum_c |f1_c * f2_c|):
* against the FP64 correlation of the *quantised* operands, only the FP32
  accumulation differs: |err| <= 2 * C * 2^-24 * B (products of BF16/FP16/FP8
  values are exact in FP32);
* against the FP64 correlation of the original FP32 features the input
  rounding dominates: |err| <= 2u(1 + u) * B with u = 2^-9 (BF16), 2^-12
  (FP16), 2^-4 (E4M3; plus the subnormal floor of the scaled operands).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LEVELS = [(32, 272, 480), (64, 136, 240), (96, 68, 120), (128, 34, 60), (196, 17, 30)]
U = {"bf16": 2.0 ** -9, "f16": 2.0 ** -12}


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _bound(orc, f1, f2, md, legacy=False):
    return orc.correlation(np.abs(f1).astype(np.float64), np.abs(f2).astype(np.float64), md,
                           legacy=legacy, dtype=np.float64)


def test_conversions_match_torch(vsc, cuda):
    import torch
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(20000) * 10, rng.uniform(-500, 500, 5000),
                        [0.0, -0.0, 1e-8, 2 ** -10, 2 ** -9, 3 * 2 ** -11, 448.0, 464.0, 500.0, -1e30]]).astype(np.float32)
    xt = _t(x, cuda)
    assert torch.equal(vsc.convert(xt, "bf16").view(torch.int16), xt.to(torch.bfloat16).view(torch.int16))
    assert torch.equal(vsc.convert(xt, "f16").view(torch.int16), xt.to(torch.float16).view(torch.int16))
    q = vsc.convert(xt, "e4m3")
    ref = torch.clamp(xt, -448.0, 448.0).to(torch.float8_e4m3fn).view(torch.uint8)
    assert torch.equal(q, ref)


@pytest.mark.parametrize("kind", ["bf16", "f16"])
@pytest.mark.parametrize("C,H,W", LEVELS)
def test_corr_16bit_levels(vsc, orc, cuda, kind, C, H, W):
    import torch
    print(f"corr {kind} C={C} H={H} W={W}", flush=True)
    rng = np.random.default_rng(C + (0 if kind == "bf16" else 1))
    f1 = rng.standard_normal((1, C, H, W)).astype(np.float32)
    f2 = rng.standard_normal((1, C, H, W)).astype(np.float32)
    a = vsc.convert(_t(f1, cuda), kind)
    b = vsc.convert(_t(f2, cuda), kind)
    got = vsc.correlation(a, b, 4, 0).cpu().numpy().astype(np.float64)
    qa = a.float().cpu().numpy()
    qb = b.float().cpu().numpy()
    ref_q = orc.correlation(qa.astype(np.float64), qb.astype(np.float64), 4, dtype=np.float64)
    bq = _bound(orc, qa, qb, 4)
    assert np.all(np.abs(got - ref_q) <= 2 * C * 2.0 ** -24 * bq + 1e-30)
    ref = orc.correlation(f1.astype(np.float64), f2.astype(np.float64), 4, dtype=np.float64)
    b0 = _bound(orc, f1, f2, 4)
    u = U[kind]
    assert np.all(np.abs(got - ref) <= (2 * u * (1 + u) + 2 * C * 2.0 ** -24) * b0 + 1e-30)


@pytest.mark.parametrize("md", [0, 1, 3, 4])
@pytest.mark.parametrize("shape", [(2, 5, 7, 9), (1, 40, 1, 33), (1, 33, 21, 1), (1, 70, 13, 50),
                                   (2, 5, 7, 12), (1, 70, 13, 52), (1, 33, 3, 4)])
def test_corr_bf16_shapes_and_legacy(vsc, orc, cuda, md, shape):
    rng = np.random.default_rng(md + sum(shape))
    f1 = rng.standard_normal(shape).astype(np.float32)
    f2 = rng.standard_normal(shape).astype(np.float32)
    a = vsc.convert(_t(f1, cuda), "bf16")
    b = vsc.convert(_t(f2, cuda), "bf16")
    qa = a.float().cpu().numpy().astype(np.float64)
    qb = b.float().cpu().numpy().astype(np.float64)
    C = shape[1]
    for legacy in (0, 1):
        if legacy and md > 4:
            continue
        got = vsc.correlation(a, b, md, legacy).cpu().numpy().astype(np.float64)
        ref = orc.correlation(qa, qb, md, legacy=bool(legacy), dtype=np.float64)
        bq = _bound(orc, qa, qb, md, legacy=bool(legacy))
        assert got.shape == ref.shape
        assert np.all(np.abs(got - ref) <= 4 * C * 2.0 ** -24 * bq + 1e-30)


def test_corr_lowp_rejects_large_displacement(vsc, cuda):
    import torch
    x = torch.zeros((1, 4, 5, 5), dtype=torch.bfloat16, device=cuda)
    with pytest.raises(vsc.VscError):
        vsc.correlation(x, x, 5, 0)


@pytest.mark.parametrize("C,H,W", LEVELS)
def test_corr_fp8_equalised(vsc, orc, cuda, C, H, W):
    import torch
    print(f"corr fp8 C={C} H={H} W={W}", flush=True)
    rng = np.random.default_rng(C + 7)
    # channels with different magnitudes: what the equalisation is for
    g1 = np.exp(rng.uniform(-2, 2, (1, C, 1, 1))).astype(np.float32)
    g2 = np.exp(rng.uniform(-2, 2, (1, C, 1, 1))).astype(np.float32)
    f1 = (rng.standard_normal((1, C, H, W)) * g1).astype(np.float32)
    f2 = (rng.standard_normal((1, C, H, W)) * g2).astype(np.float32)
    q1, q2, sc = vsc.quantize_fp8_pair(_t(f1, cuda), _t(f2, cuda))
    # quantiser: equalisation factors, per-tensor scales, RNE E4M3 (torch decodes)
    m1 = np.abs(f1).max(axis=(0, 2, 3))
    m2 = np.abs(f2).max(axis=(0, 2, 3))
    r = np.sqrt(m2 / m1).astype(np.float32)
    s1 = np.float32((m1 * r).max() / np.float32(448))
    s2 = np.float32((m2 / r).max() / np.float32(448))
    e1 = ((f1 * r[None, :, None, None]) / s1).astype(np.float32)
    e2 = ((f2 / r[None, :, None, None]) / s2).astype(np.float32)
    ref_q1 = torch.from_numpy(e1).to(torch.float8_e4m3fn).view(torch.uint8).numpy()
    ref_q2 = torch.from_numpy(e2).to(torch.float8_e4m3fn).view(torch.uint8).numpy()
    assert np.array_equal(q1.cpu().numpy(), ref_q1)
    assert np.array_equal(q2.cpu().numpy(), ref_q2)
    assert abs(float(sc.cpu()[0]) - float(s1 * s2)) <= 1e-6 * float(s1 * s2)
    got = vsc.correlation_fp8(q1, q2, sc, 4, 0).cpu().numpy().astype(np.float64)
    d1 = q1.view(torch.float8_e4m3fn).float().cpu().numpy().astype(np.float64)
    d2 = q2.view(torch.float8_e4m3fn).float().cpu().numpy().astype(np.float64)
    s = float(sc.cpu()[0])
    ref_q = s * orc.correlation(d1, d2, 4, dtype=np.float64)
    bq = s * _bound(orc, d1, d2, 4)
    acc_err = float(np.max(np.abs(got - ref_q) / (bq + 1e-30)))
    print(f"fp8 accumulation: max |err|/sum|ab| vs FP64 of the quantised operands = {acc_err:.3e}")
    # the FP8 MFMA's internal sum of 32 exact products is not an FP32 rounding
    # sequence; stay within 2^-10 of the absolute-product sum
    assert acc_err <= 2.0 ** -10
    ref = orc.correlation(f1.astype(np.float64), f2.astype(np.float64), 4, dtype=np.float64)
    b0 = _bound(orc, f1, f2, 4)
    rel = np.max(np.abs(got - ref) / (b0 + 1e-30))
    print(f"fp8 max |err|/sum|ab| = {rel:.4f}")
    assert rel <= 2 * 2.0 ** -4 * (1 + 2.0 ** -4) + 0.02


@pytest.mark.parametrize("kind", ["bf16", "f16", "fp8"])
def test_corr_workspace_form_equals_allocating_form(vsc, cuda, kind):
    """vsc_correlation_forward_ws (caller's workspace, no allocation) gives the
    allocating entry point's bits at every pyramid level, with one workspace of
    the largest level's size reused across calls; likewise the FP8 quantiser."""
    import torch
    rng = np.random.default_rng(9)
    code = {"bf16": vsc.VSC_BF16, "f16": vsc.VSC_F16, "fp8": vsc.VSC_FP8E4M3}[kind]
    need = max(vsc.correlation_workspace_bytes((1, C, H, W), code) for C, H, W in LEVELS)
    ws = torch.empty(need, dtype=torch.uint8, device=cuda)
    qws = torch.empty(vsc.lib().vsc_quantize_fp8_pair_workspace_size(196), dtype=torch.uint8, device=cuda)
    for C, H, W in LEVELS:
        a = _t(rng.standard_normal((1, C, H, W)).astype(np.float32), cuda)
        b = _t(rng.standard_normal((1, C, H, W)).astype(np.float32), cuda)
        if kind == "fp8":
            q1, q2, sc = vsc.quantize_fp8_pair(a, b)
            w1, w2, wsc = vsc.quantize_fp8_pair(a, b, workspace=qws)
            assert torch.equal(q1, w1) and torch.equal(q2, w2) and torch.equal(sc, wsc)
            ref = vsc.correlation_fp8(q1, q2, sc, 4, 0)
            got = vsc.correlation_fp8(q1, q2, sc, 4, 0, workspace=ws)
        else:
            ca, cb = vsc.convert(a, kind), vsc.convert(b, kind)
            ref = vsc.correlation(ca, cb, 4, 0)
            got = vsc.correlation(ca, cb, 4, 0, workspace=ws)
        torch.cuda.synchronize()
        assert torch.equal(ref, got), (kind, C, H, W)


@pytest.mark.parametrize("kind", ["bf16", "f16", "fp8"])
@pytest.mark.parametrize("shape", [(1, 32, 272, 480), (2, 70, 13, 52), (1, 196, 17, 28)])
def test_corr_nchw_staging_equals_channels_last_copy(vsc, cuda, kind, shape):
    """Widths that are multiples of 4 with aligned operands are read in their
    NCHW layout by the MFMA kernel; misaligned operands take the channels-last
    copy.  Both stage the same LDS tiles, so the outputs are the same bits
    (normal and legacy)."""
    import torch
    rng = np.random.default_rng(sum(shape))
    f1 = _t(rng.standard_normal(shape).astype(np.float32), cuda)
    f2 = _t(rng.standard_normal(shape).astype(np.float32), cuda)
    sc = None
    if kind == "fp8":
        a, b, sc = vsc.quantize_fp8_pair(f1, f2)
    else:
        a, b = vsc.convert(f1, kind), vsc.convert(f2, kind)

    def misaligned(x):
        # the same values one element past an aligned boundary (pointer not 4/8-byte aligned)
        buf = torch.empty(x.numel() + 1, dtype=x.dtype, device=cuda)
        v = buf[1:].view(x.shape)
        v.copy_(x)
        return v

    ma, mb = misaligned(a), misaligned(b)
    assert ma.data_ptr() % 4 != 0
    for legacy in (0, 1):
        if kind == "fp8":
            direct = vsc.correlation_fp8(a, b, sc, 4, legacy)
            copy = vsc.correlation_fp8(ma, mb, sc, 4, legacy)
        else:
            direct = vsc.correlation(a, b, 4, legacy)
            copy = vsc.correlation(ma, mb, 4, legacy)
        torch.cuda.synchronize()
        assert torch.equal(direct, copy), (kind, shape, legacy)


@pytest.mark.parametrize("kind", ["bf16", "fp8"])
@pytest.mark.parametrize("shape", [(1, 32, 136, 256), (1, 40, 17, 28)])
def test_corr_lowp_unaligned_output(vsc, cuda, kind, shape):
    """The MFMA kernel writes its staged planes with 16-byte stores when the
    output is 16-byte aligned and W a multiple of 4; an output 4 bytes off
    that alignment takes per-pixel stores with the same bits.  The 136 x 256
    shape (272 blocks) runs the two-blocks-per-CU form, the 17 x 28 one the
    other."""
    import torch
    rng = np.random.default_rng(sum(shape))
    a = _t(rng.standard_normal(shape).astype(np.float32), cuda)
    b = _t(rng.standard_normal(shape).astype(np.float32), cuda)
    if kind == "fp8":
        q1, q2, sc = vsc.quantize_fp8_pair(a, b)
        run = lambda out=None: vsc.correlation_fp8(q1, q2, sc, 4, 0, out=out)
    else:
        ca, cb = vsc.convert(a, kind), vsc.convert(b, kind)
        run = lambda out=None: vsc.correlation(ca, cb, 4, 0, out=out)
    ref = run()
    buf = torch.empty(ref.numel() + 1, dtype=torch.float32, device=cuda)
    out = buf[1:].view(ref.shape)
    assert out.data_ptr() % 16 == 4
    run(out)
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
