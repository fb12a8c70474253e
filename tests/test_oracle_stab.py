"""Pins the oracle's stabilization stages (flowconsistency.cu:50-258,
gpuimage.cu:39-67, flowIO.cpp:31-78) by closed forms, invariants and an
independent sparse direct solve of the solver's fixed point."""
import os
import struct

import numpy as np
import pytest

DEF = dict(alpha=6800.0, beta=6800.0, gamma=2.0)


# ---- warp (kernel_warp) --------------------------------------------------------

def test_stab_warp_integer_shift_with_clamp(orc, rng):
    H, W = 9, 12
    img = rng.random((H, W, 3)).astype(np.float32)
    fl = np.zeros((H, W, 2), np.float32)
    fl[..., 0] = 2
    fl[..., 1] = -1
    out = orc.stab_warp(img, fl)
    ys, xs = np.mgrid[0:H, 0:W]
    sx = np.clip(xs + 2, 0, W - 3)
    sy = np.clip(ys - 1, 0, H - 3)
    assert np.array_equal(out, img[sy, xs * 0 + sx])


def test_stab_warp_reproduces_linear_ramp(orc, rng):
    H, W = 16, 20
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    img = np.stack([0.01 * xs + 0.02 * ys, 0.03 * xs, 0.5 - 0.01 * ys], -1).astype(np.float32)
    fl = rng.uniform(-3, 3, (H, W, 3)).astype(np.float32)   # 3-channel flow (network output)
    out = orc.stab_warp(img, fl)
    mx = np.clip(xs + fl[..., 0], 0, W - 3)
    my = np.clip(ys + fl[..., 1], 0, H - 3)
    ref = np.stack([0.01 * mx + 0.02 * my, 0.03 * mx, 0.5 - 0.01 * my], -1)
    np.testing.assert_allclose(out, ref, atol=2e-6)


# ---- adaptive combination / consistency weights -------------------------------

def _adap_ref64(ci, cp, pi, pp, ni, np_, ls, alpha):
    wp = np.exp(-alpha * (ci - pi) ** 2)
    wn = np.exp(-alpha * (ci - ni) ** 2)
    wp = np.where(wp > 0.45, 0.45, wp)
    wn = np.where(wn > 0.3, 0.3, wn)
    wp = np.where(wp < 0.001, 0, wp)
    wn = np.where(wn < 0.001, 0, wn)
    adIn = wp * pi + wn * ni + (1 - (wp + wn)) * ci
    ap = wp * pp + wn * np_ + (1 - (wp + wn)) * cp
    return adIn, wp * ls + (1 - wp) * ap


def test_adap_comb_limits_exact(orc, rng):
    H, W = 4, 5
    ci, cp, pp, np_, ls = (rng.random((H, W, 3)).astype(np.float32) for _ in range(5))
    far = (ci + 0.5).astype(np.float32)            # weight exp(-6800*0.25) -> 0
    adIn, adPr = orc.stab_adap_comb(ci, cp, far, pp, far, np_, ls, 6800.0)
    assert np.array_equal(adIn, ci) and np.array_equal(adPr, cp)
    adIn, adPr = orc.stab_adap_comb(ci, cp, ci, pp, ci, np_, ls, 6800.0)   # weights capped 0.45 / 0.3
    w = np.float32
    ap = (w(0.45) * pp + w(0.3) * np_) + (w(1) - (w(0.45) + w(0.3))) * cp
    np.testing.assert_allclose(adPr, w(0.45) * ls + (w(1) - w(0.45)) * ap, rtol=1e-6)
    np.testing.assert_allclose(adIn, ci, rtol=1e-6)


def test_adap_comb_and_weights_vs_fp64(orc, rng):
    H, W = 20, 30
    ci = rng.random((H, W, 3)).astype(np.float32)
    d = rng.normal(0, 0.02, (6, H, W, 3)).astype(np.float32)
    pi, ni, cp, pp, np_, ls = (np.clip(ci + d[k], 0, 1).astype(np.float32) for k in range(6))
    adIn, adPr = orc.stab_adap_comb(ci, cp, pi, pp, ni, np_, ls, 6800.0)
    rIn, rPr = _adap_ref64(*(a.astype(np.float64) for a in (ci, cp, pi, pp, ni, np_, ls)), 6800.0)
    # away from the clamp thresholds the float result is within a few ulp
    wp = np.exp(-6800.0 * (ci.astype(np.float64) - pi) ** 2)
    wn = np.exp(-6800.0 * (ci.astype(np.float64) - ni) ** 2)
    ok = (np.abs(wp - 0.001) > 1e-5) & (np.abs(wn - 0.001) > 1e-5)
    assert ok.mean() > 0.95
    np.testing.assert_allclose(adIn[ok], rIn[ok], atol=2e-6)
    np.testing.assert_allclose(adPr[ok], rPr[ok], atol=2e-6)
    wt = orc.stab_consist_wt(adIn, ci, 6800.0, 2.0)
    rw = 2.0 * np.exp(-6800.0 * (ci.astype(np.float64) - adIn) ** 2)
    rw = np.where(rw < 0.001, 0, rw)
    ok = np.abs(rw - 0.001) > 1e-6
    np.testing.assert_allclose(wt[ok], rw[ok], rtol=1e-5, atol=1e-7)
    assert np.all((wt == 0) | (wt >= 0.001))
    assert np.all(wt <= 2.0)


# ---- consistency solver -----------------------------------------------------------

def test_solver_1x1_heavy_ball_recurrence(orc):
    """No neighbours: O' = O - s*w*(O-A) + m*U, U' = -s*w*(O-A), in float."""
    P = np.array([[[0.3, 0.6, 0.9]]], np.float32)
    A = np.array([[[0.7, 0.1, 0.5]]], np.float32)
    Wt = np.array([[[1.5, 0.2, 2.0]]], np.float32)
    O = orc.stab_consist_out(P, A, Wt, P, 37)
    f = np.float32
    o = P.copy()
    u = np.zeros_like(o)
    for k in range(37):
        g = (f(0) - f(0)) - Wt * (o - A)   # lapO - lapP = 0 - 0
        sg = f(0.15) * g
        o = (o + sg) + f(0.15) * u if k > 0 else o + sg
        u = sg
    assert np.array_equal(O, o)
    O = orc.stab_consist_out(P, A, Wt, P, 1500)
    np.testing.assert_allclose(O, A, atol=1e-5)  # converges to the anchor


def test_solver_gamma0_identity_bit_exact(orc, rng):
    P = rng.random((13, 17, 3)).astype(np.float32)
    A = rng.random((13, 17, 3)).astype(np.float32)
    O = orc.stab_consist_out(P, A, np.zeros_like(P), P, 150)
    assert np.array_equal(O, P)


def _sparse_fixed_point(P, A, Wt):
    sp = pytest.importorskip("scipy.sparse")
    spl = pytest.importorskip("scipy.sparse.linalg")
    H, W, _ = P.shape
    n = H * W
    rows, cols, vals = [], [], []
    for y in range(H):
        for x in range(W):
            i = y * W + x
            nb = []
            if x + 1 < W - 1:
                nb.append(i + 1)
            if x - 1 >= 0:
                nb.append(i - 1)
            if y + 1 < H - 1:
                nb.append(i + W)
            if y - 1 >= 0:
                nb.append(i - W)
            for j in nb:
                rows.append(i); cols.append(j); vals.append(1.0)
            rows.append(i); cols.append(i); vals.append(-float(len(nb)))
    L = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    out = np.empty((H, W, 3))
    for c in range(3):
        p = P[..., c].astype(np.float64).ravel()
        a = A[..., c].astype(np.float64).ravel()
        w = Wt[..., c].astype(np.float64).ravel()
        M = L - sp.diags(w)
        out[..., c] = spl.spsolve(M.tocsc(), L @ p - w * a).reshape(H, W)
    return out


def test_solver_converges_to_sparse_direct_solution(orc, rng):
    """The sweep is gradient descent with momentum on
    (L O - L P) - w*(O - A) = 0, L = the reference's asymmetric neighbour
    Laplacian (right only if x+1 < W-1, down only if y+1 < H-1)."""
    H, W = 9, 12
    P = rng.random((H, W, 3)).astype(np.float32)
    A = rng.random((H, W, 3)).astype(np.float32)
    Wt = rng.uniform(0.5, 2.0, (H, W, 3)).astype(np.float32)
    ref = _sparse_fixed_point(P, A, Wt)
    O64 = orc.stab_consist_out(P, A, Wt, P, 3000, dtype=np.float64)
    np.testing.assert_allclose(O64, ref, atol=1e-9)
    O32 = orc.stab_consist_out(P, A, Wt, P, 3000)
    np.testing.assert_allclose(O32, ref, atol=2e-5)


def test_solver_one_sweep_order(orc, rng):
    """One Jacobi sweep replayed with numpy float32 element-wise ops in the
    reference's order (right, left, down, up; then - cnt*center)."""
    H, W = 6, 7
    P, A, O = (rng.random((H, W, 3)).astype(np.float32) for _ in range(3))
    Wt = rng.uniform(0, 2, (H, W, 3)).astype(np.float32)
    got = orc.stab_consist_out(P, A, Wt, O, 1)
    f = np.float32
    ref = np.empty_like(O)
    for y in range(H):
        for x in range(W):
            lp = np.zeros(3, f); lo = np.zeros(3, f); cnt = 0
            for cond, (yy, xx) in ((x + 1 < W - 1, (y, x + 1)), (x - 1 >= 0, (y, x - 1)),
                                   (y + 1 < H - 1, (y + 1, x)), (y - 1 >= 0, (y - 1, x))):
                if cond:
                    lp = lp + P[yy, xx]; lo = lo + O[yy, xx]; cnt += 1
            lp = lp - f(cnt) * P[y, x]
            lo = lo - f(cnt) * O[y, x]
            g = (lo - lp) - Wt[y, x] * (O[y, x] - A[y, x])
            ref[y, x] = O[y, x] + f(0.15) * g
    assert np.array_equal(got, ref)


# ---- the FMA form of the solver ---------------------------------------------------------

def test_fma_solver_1x1_against_fp64_recurrence(orc):
    """No neighbours: the FMA form follows the heavy-ball recurrence
    O' = O - s*w*(O-A) + m*U evaluated in float64 to float rounding, and
    converges to the anchor."""
    P = np.array([[[0.3, 0.6, 0.9]]], np.float32)
    A = np.array([[[0.7, 0.1, 0.5]]], np.float32)
    Wt = np.array([[[1.5, 0.2, 2.0]]], np.float32)
    O = orc.stab_consist_out(P, A, Wt, P, 37, solver="fma")
    o = P.astype(np.float64)
    u = np.zeros_like(o)
    for k in range(37):
        sg = -0.15 * Wt * (o - A)
        o = o + sg + (0.15 * u if k > 0 else 0.0)
        u = sg
    np.testing.assert_allclose(O, o, rtol=0, atol=2e-7)
    O = orc.stab_consist_out(P, A, Wt, P, 1500, solver="fma")
    np.testing.assert_allclose(O, A, atol=1e-5)


def test_fma_solver_converges_to_sparse_direct_solution(orc, rng):
    """Same fixed point as the reference order: neighbour sets, cnt + w and
    lapP - w*A enter with the right signs and positions."""
    H, W = 9, 12
    P = rng.random((H, W, 3)).astype(np.float32)
    A = rng.random((H, W, 3)).astype(np.float32)
    Wt = rng.uniform(0.5, 2.0, (H, W, 3)).astype(np.float32)
    ref = _sparse_fixed_point(P, A, Wt)
    O32 = orc.stab_consist_out(P, A, Wt, P, 3000, solver="fma")
    np.testing.assert_allclose(O32, ref, atol=2e-5)


def test_fma_solver_sweeps_track_reference_order(orc, rng):
    """Sweep by sweep the two forms differ by float rounding only (both within
    a few ulp of the FP64 sweeps), including momentum, edges and w = 0."""
    H, W = 11, 14
    P = rng.random((H, W, 3)).astype(np.float32)
    A = rng.random((H, W, 3)).astype(np.float32)
    Wt = rng.uniform(0, 2, (H, W, 3)).astype(np.float32)
    Wt[2:5, 3:9] = 0
    for iters in (1, 2, 7, 40):
        o64 = orc.stab_consist_out(P, A, Wt, P, iters, dtype=np.float64)
        ex = orc.stab_consist_out(P, A, Wt, P, iters)
        fm = orc.stab_consist_out(P, A, Wt, P, iters, solver="fma")
        assert np.abs(fm - o64).max() <= 2e-6, iters
        assert np.abs(ex - o64).max() <= 2e-6, iters
    O = orc.stab_consist_out(P, A, np.zeros_like(P), P, 150, solver="fma")
    np.testing.assert_allclose(O, P, rtol=0, atol=2e-6)   # gamma = 0: the processed frame stays


def test_fma_frames_within_one_lsb_of_reference_order(orc, vsc):
    """Whole frames (2-level pyramid, 150 sweeps) of a synthetic sequence: the
    FMA form is within 1 LSB of the reference order everywhere and as close to
    the FP64 twin."""
    orig, proc, fwd, bwd = vsc.synth_video(width=96, height=64, nframes=5, nblobs=40, motion=3.0, npatches=1,
                                           patch_size=16, seed=4)
    ex, exf = orc.stabilize_sequence(orig, proc, fwd, bwd)
    fm, fmf = orc.stabilize_sequence(orig, proc, fwd, bwd, solver="fma")
    _, f64 = orc.stabilize_sequence(orig, proc, fwd, bwd, dtype=np.float64)
    for i in range(1, 4):
        d = (fm[i][..., :3].astype(int) - ex[i][..., :3].astype(int)) % 256
        assert np.minimum(d, 256 - d).max() <= 1
        e_fma = np.abs(fmf[i] - f64[i]).max()
        e_ex = np.abs(exf[i] - f64[i]).max()
        assert e_fma <= max(2 * e_ex, 1e-6), (i, e_fma, e_ex)


# ---- bilinear resample --------------------------------------------------------------

def test_bilinear_decimation_and_midpoints(orc, rng):
    img = rng.random((10, 14, 3)).astype(np.float32)
    down = orc.stab_bilinear(img, 7, 5)
    assert np.array_equal(down, img[::2, ::2])
    up = orc.stab_bilinear(down, 14, 10)
    assert np.array_equal(up[::2, ::2], down)
    np.testing.assert_allclose(up[::2, 1:-1:2], 0.5 * (down[:, :-1] + down[:, 1:]), rtol=1e-6)
    assert np.array_equal(up[::2, -1], down[:, -1])   # clamp to the last column


def test_bilinear_linear_ramp_and_channels(orc):
    H, W = 11, 13
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    img = np.stack([xs, ys, xs + ys], -1).astype(np.float32)
    out = orc.stab_bilinear(img, 29, 23, out_c=2)      # arbitrary factor, 2 of 3 channels
    oy, ox = np.mgrid[0:23, 0:29].astype(np.float64)
    xx = np.minimum(ox * W / 29, W - 1)
    yy = np.minimum(oy * H / 23, H - 1)
    inner = (ox * W / 29 <= W - 1) & (oy * H / 23 <= H - 1)
    np.testing.assert_allclose(out[..., 0][inner], xx[inner], atol=1e-4)
    np.testing.assert_allclose(out[..., 1][inner], yy[inner], atol=1e-4)
    assert out.shape == (23, 29, 2)


# ---- u8 conversions ----------------------------------------------------------------

def test_u8_to_float_exhaustive(orc):
    v = np.arange(256, dtype=np.uint8)
    rgba = np.stack([v, v[::-1], v, np.full(256, 7, np.uint8)], -1).reshape(16, 16, 4)
    out = orc.rgba_to_float(rgba)
    exp = (np.arange(256, dtype=np.float64) / 255.0).astype(np.float32).reshape(16, 16)
    assert np.array_equal(out[..., 0], exp)
    assert np.array_equal(out[..., 1], exp.ravel()[::-1].reshape(16, 16))


def test_float_to_u8_rounding_and_wrap(orc):
    k = np.arange(0, 300, dtype=np.float64)
    vals = []
    for kk in k:
        b = kk / 255.0
        vals += [np.nextafter(np.float32(b), np.float32(-1)), np.float32(b), np.nextafter(np.float32(b), np.float32(9))]
    vals += [-0.5, -1.0, 2.0, 1e12, np.nan]
    img = np.array(vals, np.float32)
    pad = (-len(img)) % 3
    img = np.concatenate([img, np.zeros(pad, np.float32)]).reshape(1, -1, 3)
    out = orc.float_to_rgba(img)
    f = (np.abs(img.astype(np.float32)) * np.float32(255)).ravel()
    exp = []
    for v in f:
        if np.isnan(v):
            exp.append(0)
        elif v >= 2.0 ** 32:
            exp.append((2 ** 32 - 1) % 256)
        else:
            exp.append(int(np.floor(np.float64(v))) % 256)
    assert np.array_equal(out[..., :3].ravel(), np.array(exp, np.uint8))
    assert np.all(out[..., 3] == 1)
    # the wrap: 1.0 -> 255, 256/255 -> 0
    probe = orc.float_to_rgba(np.array([[[1.0, 256 / 255 + 1e-6, 2.0]]], np.float32))
    assert list(probe[0, 0, :3]) == [255, 0, 254]


# ---- .flo files -----------------------------------------------------------------------

def _write_flo(path, w, h, data, tag=202021.25, extra=b""):
    with open(path, "wb") as fh:
        fh.write(struct.pack("<f", tag) + struct.pack("<ii", w, h) + data.astype("<f4").tobytes() + extra)


def test_read_flo_roundtrip_and_rejections(orc, rng, tmp_path):
    fl = rng.standard_normal((5, 7, 2)).astype(np.float32)
    p = str(tmp_path / "a.flo")
    _write_flo(p, 7, 5, fl)
    got, rc = orc.read_flo(p)
    assert rc == 0 and np.array_equal(got, fl)
    _write_flo(p, 7, 5, fl, tag=1.0)
    assert orc.read_flo(p)[1] == -3
    _write_flo(p, 0, 5, fl)
    assert orc.read_flo(p)[1] == -4
    _write_flo(p, 7, 100000, fl)
    assert orc.read_flo(p)[1] == -5
    _write_flo(p, 7, 5, fl[:4])
    assert orc.read_flo(p)[1] == -6
    _write_flo(p, 7, 5, fl, extra=b"\0")
    assert orc.read_flo(p)[1] == -7
    assert orc.read_flo(str(tmp_path / "missing.flo"))[1] == -1
