"""Pins the oracle's Correlation (correlation.cc:147-275, correlation_cuda.cu:182-331)
with an independent FP64 formulation (shifted slices), closed forms and the
float summation order."""
import numpy as np
import pytest


def corr_fp64_slices(f1, f2, md):
    """Independent FP64 restatement: for each displacement, multiply f1 by a
    zero-padded shifted view of f2 and sum over channels."""
    N, C, H, W = f1.shape
    P = 2 * md + 1
    f1 = f1.astype(np.float64)
    pad = np.zeros((N, C, H + 2 * md, W + 2 * md))
    pad[:, :, md:md + H, md:md + W] = f2
    out = np.empty((N, P, P, H, W))
    for ph in range(P):
        for pw in range(P):
            out[:, ph, pw] = (f1 * pad[:, :, ph:ph + H, pw:pw + W]).sum(axis=1)
    return out


def test_fp64_matches_slices(orc, rng):
    f1 = rng.standard_normal((2, 5, 9, 11))
    f2 = rng.standard_normal((2, 5, 9, 11))
    for md in (0, 1, 4):
        got = orc.correlation(f1, f2, md, dtype=np.float64)
        np.testing.assert_allclose(got, corr_fp64_slices(f1, f2, md), rtol=1e-12, atol=1e-12)


def test_f32_sequential_order_is_bit_exact(orc, rng):
    """The reference accumulates sequentially in c in float: replay that order
    element-wise with numpy float32 ops (no fused multiply-add)."""
    f1 = rng.standard_normal((1, 7, 6, 10)).astype(np.float32)
    f2 = rng.standard_normal((1, 7, 6, 10)).astype(np.float32)
    md = 2
    got = orc.correlation(f1, f2, md)
    P = 2 * md + 1
    pad = np.zeros((1, 7, 6 + 2 * md, 10 + 2 * md), np.float32)
    pad[:, :, md:-md, md:-md] = f2
    for ph in range(P):
        for pw in range(P):
            acc = np.zeros((6, 10), np.float32)
            sh = pad[0, :, ph:ph + 6, pw:pw + 10]
            inside = np.zeros((6 + 2 * md, 10 + 2 * md), bool)
            inside[md:-md, md:-md] = True
            ok = inside[ph:ph + 6, pw:pw + 10]
            for c in range(7):
                acc = np.where(ok, np.float32(acc + np.float32(f1[0, c] * sh[c])), acc)
            assert np.array_equal(got[0, ph, pw], acc)


def test_f32_within_forward_error_bound(orc, rng):
    f1 = rng.standard_normal((1, 96, 12, 13)).astype(np.float32)
    f2 = rng.standard_normal((1, 96, 12, 13)).astype(np.float32)
    got = orc.correlation(f1, f2, 4).astype(np.float64)
    ref = corr_fp64_slices(f1, f2, 4)
    bound = corr_fp64_slices(np.abs(f1), np.abs(f2), 4)
    assert np.all(np.abs(got - ref) <= 96 * 2 ** -24 * bound + 1e-30)


def test_special_cases(orc, rng):
    f1 = rng.standard_normal((1, 4, 8, 9)).astype(np.float32)
    # md = 0: channel dot product per pixel
    np.testing.assert_allclose(orc.correlation(f1, f1, 0)[0, 0, 0], (f1[0] ** 2).sum(0), rtol=1e-5)
    # C = 1: product of shifted planes, exact
    a = rng.standard_normal((1, 1, 5, 6)).astype(np.float32)
    b = rng.standard_normal((1, 1, 5, 6)).astype(np.float32)
    out = orc.correlation(a, b, 1)
    assert out[0, 1, 2, 2, 3] == np.float32(a[0, 0, 2, 3] * b[0, 0, 2, 4])
    assert out[0, 0, 0, 0, 0] == 0.0          # shifted above/left of the image
    assert out[0, 2, 2, 4, 5] == 0.0          # shifted below/right
    # f2 = f1: the centre displacement is the squared norm, >= every other
    c = orc.correlation(f1, f1, 2)[0]
    np.testing.assert_allclose(c[2, 2], (f1[0].astype(np.float64) ** 2).sum(0), rtol=1e-5)


def test_translation_argmax(orc, rng):
    base = rng.standard_normal((1, 64, 20, 24)).astype(np.float32)
    dy, dx = 2, -3
    f2 = np.roll(base, (dy, dx), axis=(2, 3))
    c = orc.correlation(base, f2, 4)[0].reshape(81, 20, 24)
    inner = c[:, 6:14, 6:18].reshape(81, -1)
    idx = inner.argmax(0)
    assert np.all(idx == (dy + 4) * 9 + (dx + 4))


def test_zero_border_exact(orc, rng):
    f1 = rng.standard_normal((1, 3, 5, 5)).astype(np.float32)
    f2 = rng.standard_normal((1, 3, 5, 5)).astype(np.float32)
    c = orc.correlation(f1, f2, 4)[0]
    # displacement (+4, +4) only reaches the pixel (0, 0)
    m = np.zeros((5, 5), bool)
    m[0, 0] = True
    assert np.all(c[8, 8][~m] == 0.0)


@pytest.mark.parametrize("md", [1, 4])
def test_legacy_is_nonlegacy_over_c(orc, rng, md):
    f1 = rng.standard_normal((2, 40, 7, 9)).astype(np.float32)
    f2 = rng.standard_normal((2, 40, 7, 9)).astype(np.float32)
    leg = orc.correlation(f1, f2, md, legacy=True)
    P = 2 * md + 1
    assert leg.shape == (2, P * P, 7, 9)
    ref = corr_fp64_slices(f1, f2, md).reshape(2, P * P, 7, 9) / 40
    np.testing.assert_allclose(leg, ref, rtol=2e-5, atol=2e-6)
