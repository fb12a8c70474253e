"""vsc_correlation_forward_noalloc: the Correlation op whose reduced-precision
operand copies live in a static device scratch instead of an allocation.  Its
outputs equal vsc_correlation_forward's bit for bit on two calls in a row (the
second with a smaller input, so the scratch holds the first call's larger
contents), for BF16, FP16 and FP8 (scaled) operands and for FP32 (no scratch).
The widths are not multiples of 4, so the reduced-precision operands take the
channels-last copies that live in the scratch; inputs needing more scratch
than the static size are refused before any GPU work, and inputs the kernel
reads in their NCHW layout need none, whatever their size."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(vsc, fn, a, b, md, scale=None):
    import torch
    N, C, H, W = a.shape
    P = 2 * md + 1
    out = torch.full((N, P, P, H, W), float("nan"), dtype=torch.float32, device=a.device)
    code = {torch.bfloat16: vsc.VSC_BF16, torch.float16: vsc.VSC_F16, torch.uint8: vsc.VSC_FP8E4M3,
            torch.float32: vsc.VSC_F32}[a.dtype]
    sp = None if scale is None else scale.data_ptr()
    rc = getattr(vsc.lib(), fn)(a.data_ptr(), b.data_ptr(), out.data_ptr(), N, C, H, W, md, 0, code, sp,
                                torch.cuda.current_stream().cuda_stream)
    vsc.check(rc)
    return out


@pytest.mark.parametrize("kind", ["bf16", "f16", "fp8", "f32"])
def test_noalloc_equals_allocating_entry_point(vsc, cuda, kind):
    import torch
    rng = np.random.default_rng(17)
    # first call: about the largest level of the 1920x1088 feature pyramid;
    # second: a smaller one (widths 2 mod 4: channels-last copies in the scratch)
    assert vsc.correlation_workspace_bytes((1, 32, 272, 478), vsc.VSC_BF16) > 0
    for C, H, W in ((32, 272, 478), (96, 68, 118)):
        a = torch.from_numpy(rng.standard_normal((1, C, H, W)).astype(np.float32)).to(cuda)
        b = torch.from_numpy(rng.standard_normal((1, C, H, W)).astype(np.float32)).to(cuda)
        sc = None
        if kind == "fp8":
            a, b, sc = vsc.quantize_fp8_pair(a, b)
        elif kind != "f32":
            a, b = vsc.convert(a, kind), vsc.convert(b, kind)
        ref = _call(vsc, "vsc_correlation_forward", a, b, 4, sc)
        got = _call(vsc, "vsc_correlation_forward_noalloc", a, b, 4, sc)
        torch.cuda.synchronize()
        assert torch.equal(ref, got), (kind, C, H, W)


def test_noalloc_refuses_oversized_operands(vsc):
    L = vsc.lib()
    dummy = ctypes.c_void_p(256)
    # 1 x 256 x 256 x 258 BF16 needs 2 * 32.25 MiB of channels-last copies
    # (width 2 mod 4: not read in place)
    assert L.vsc_correlation_workspace_size(1, 256, 256, 258, vsc.VSC_BF16) > (32 << 20)
    assert L.vsc_correlation_forward_noalloc(dummy, dummy, dummy, 1, 256, 256, 258, 4, 0, vsc.VSC_BF16, None,
                                             None) == vsc.VSC_E_INVALID
    assert b"scratch" in L.vsc_last_error()


def test_noalloc_reads_aligned_operands_in_place(vsc, cuda):
    """Operands whose channels-last copies would exceed the static scratch are
    accepted when their width is a multiple of 4 (read in their NCHW layout),
    with the allocating entry point's bits."""
    import torch
    rng = np.random.default_rng(5)
    C, H, W = 160, 256, 256
    assert vsc.correlation_workspace_bytes((1, C, H, W), vsc.VSC_BF16) > (32 << 20)
    a = vsc.convert(torch.from_numpy(rng.standard_normal((1, C, H, W)).astype(np.float32)).to(cuda), "bf16")
    b = vsc.convert(torch.from_numpy(rng.standard_normal((1, C, H, W)).astype(np.float32)).to(cuda), "bf16")
    ref = _call(vsc, "vsc_correlation_forward", a, b, 4)
    got = _call(vsc, "vsc_correlation_forward_noalloc", a, b, 4)
    torch.cuda.synchronize()
    assert torch.equal(ref, got)
