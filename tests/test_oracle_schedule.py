"""Pins the oracle's frame schedule, the stabilization body of
VideoStabilizer::doOneStep (videostabilizer.cpp:182-228), without reusing
the oracle's own stage functions:

  (a) num_iter = 0: the output is the pyramid's up/down resampling of the
      processed frame alone, checked against the midpoint closed form of an
      exact 2x resample (the finer level is overwritten by the upsampled
      coarse solution, videostabilizer.cpp:219-224);
  (b) an integer-translation window with exact flows: the five warps are
      array slices, the weights and the adapted anchor follow the
      reference's formulas in float64, and the level-0 solution after many
      sweeps must equal the sparse direct solve of the solver's fixed point
      (catches which flow warps which frame, and what `last` is warped with);
  (c) uniform frames, zero flow, processed = original: output = input within
      1 LSB;
  (d) a 2x2 frame whose pixel (0, 0) has no neighbours at either level: its
      value after the two levels is the heavy-ball recurrence's closed form
      (characteristic roots) over num_iter/2 and then num_iter sweeps.

Each pin is also run against planted schedule mistakes (oracle
`stab_frame_mutant`): every mistake must fail at least one pin."""
import numpy as np
import pytest

from test_oracle_stab import _sparse_fixed_point

ALPHA = BETA = 6800.0
GAMMA = 2.0
STEP = MOM = float(np.float32(0.15))      # the oracle takes step and momentum as float


# ---- independent helpers ------------------------------------------------------------

def _up2_midpoints(d, out_w, out_h):
    """Exact 2x upsample characterised by its samples: even output positions
    copy the input, odd ones are the mean of the two (or four) neighbours,
    the neighbour past the last row/column being the last one."""
    h, w, _ = d.shape
    assert (out_w, out_h) == (2 * w, 2 * h)
    d = d.astype(np.float64)
    nx = np.minimum(np.arange(w) + 1, w - 1)
    ny = np.minimum(np.arange(h) + 1, h - 1)
    out = np.empty((out_h, out_w, d.shape[2]))
    out[0::2, 0::2] = d
    out[0::2, 1::2] = 0.5 * (d + d[:, nx])
    out[1::2, 0::2] = 0.5 * (d + d[ny, :])
    out[1::2, 1::2] = 0.25 * (d + d[:, nx] + d[ny, :] + d[ny][:, nx])
    return out


def _slice_warp(img, fx, fy):
    """kernel_warp for a constant integer flow: the sample position is
    clamped to [0, W-3] x [0, H-3] (flowconsistency.cu:93-96) and needs no
    interpolation."""
    H, W, _ = img.shape
    ys = np.clip(np.arange(H) + fy, 0, H - 3)
    xs = np.clip(np.arange(W) + fx, 0, W - 3)
    return img[ys][:, xs]


# the kernels' literals are floats (0.45f, 0.3f, 0.001f, flowconsistency.cu:146-150,186)
CAP_P, CAP_N, TINY = (float(np.float32(v)) for v in (0.45, 0.3, 0.001))


def _weights64(ci, pi, ni):
    """kernel_adap_comb's weights (flowconsistency.cu:142-150) in float64."""
    wp = np.exp(-ALPHA * (ci - pi) ** 2)
    wn = np.exp(-ALPHA * (ci - ni) ** 2)
    wp = np.where(wp > CAP_P, CAP_P, wp)
    wn = np.where(wn > CAP_N, CAP_N, wn)
    wp = np.where(wp < TINY, 0.0, wp)
    wn = np.where(wn < TINY, 0.0, wn)
    return wp, wn


def _anchor64(ci, cp, pi, pp, ni, np_, last_w):
    """Adapted processed anchor and consistency weight (flowconsistency.cu:
    152-157, 181-187) in float64."""
    wp, wn = _weights64(ci, pi, ni)
    ad_in = wp * pi + wn * ni + (1 - (wp + wn)) * ci
    ad_pr = wp * pp + wn * np_ + (1 - (wp + wn)) * cp
    ad_pr = wp * last_w + (1 - wp) * ad_pr
    w = GAMMA * np.exp(-BETA * (ci - ad_in) ** 2)
    w = np.where(w < TINY, 0.0, w)
    return ad_pr, w


def _heavy_ball_closed_form(e0, a, m, n):
    """e_k of e_{k+1} = (1-a) e_k - m a e_{k-1} (k >= 1), e_1 = (1-a) e_0:
    the error O - A of an isolated pixel under n sweeps O' = O + s g + m U,
    U' = s g, g = -w (O - A), a = s w, with no momentum on the first sweep.
    Solved through the roots of r^2 - (1-a) r + m a = 0."""
    if n == 0:
        return e0
    disc = complex((1 - a) ** 2 - 4 * m * a)
    r1 = ((1 - a) + np.sqrt(disc)) / 2
    r2 = ((1 - a) - np.sqrt(disc)) / 2
    # e_0 = c1 + c2, e_1 = c1 r1 + c2 r2
    e1 = (1 - a) * e0
    c1 = (e1 - r2 * e0) / (r1 - r2)
    c2 = e0 - c1
    return float((c1 * r1 ** n + c2 * r2 ** n).real)


def _frame(orc, ims, fwd, bwd, last, params, dtype=np.float64, solver="exact", mutation=None):
    if mutation is not None:
        return orc.stab_frame_mutant(*ims, fwd, bwd, last, params=params, mutation=mutation)
    return orc.stab_frame(*ims, fwd, bwd, last, params=params, dtype=dtype, solver=solver)


# ---- (a) num_iter = 0: resampling only ---------------------------------------------

def _case_a(rng):
    H, W = 12, 16
    ims = [rng.random((H, W, 3)) for _ in range(6)]
    fwd = rng.uniform(-2, 2, (H, W, 2))
    bwd = rng.uniform(-2, 2, (H, W, 2))
    last = rng.random((H, W, 3))
    P = ims[4]
    dec = P[::2, ::2]                                     # exact 2x down = decimation
    expect = {2: _up2_midpoints(dec, W, H),
              3: _up2_midpoints(_up2_midpoints(dec[::2, ::2], W // 2, H // 2), W, H)}
    return ims, fwd, bwd, last, expect


@pytest.mark.parametrize("levels", [2, 3])
def test_zero_iterations_is_pyramid_resampling(orc, rng, levels):
    ims, fwd, bwd, last, expect = _case_a(rng)
    prm = dict(pyramid_levels=levels, num_iter=0)
    out64 = _frame(orc, ims, fwd, bwd, last, prm)
    np.testing.assert_allclose(out64, expect[levels], rtol=0, atol=1e-15)
    out32 = _frame(orc, ims, fwd, bwd, last, prm, dtype=np.float32)
    np.testing.assert_allclose(out32, expect[levels], rtol=0, atol=4e-7)
    one = _frame(orc, ims, fwd, bwd, last, dict(pyramid_levels=1, num_iter=0))
    assert np.array_equal(one, ims[4])                    # one level: the processed frame


@pytest.mark.parametrize("mutation", [5, 6])
def test_zero_iterations_pin_detects_mistakes(orc, rng, mutation):
    ims, fwd, bwd, last, expect = _case_a(rng)
    out = _frame(orc, ims, fwd, bwd, last, dict(pyramid_levels=2, num_iter=0), mutation=mutation)
    assert np.abs(out - expect[2]).max() > 1e-3


# ---- (b) integer translation with exact flows ------------------------------------------

def _case_b(rng):
    """Originals are windows of one texture whose values are multiples of 1/8,
    so a warped pixel either equals the current one (weight at its cap) or
    differs by >= 1/8 (weight exp(-106) -> 0): no weight lies near a clamp
    threshold, and float32 and float64 agree on every weight."""
    H, W = 24, 32
    tex = np.floor(rng.random((H + 30, W + 30, 3)) * 8) / 8
    off = {"prev": (12, 10), "cur": (10, 13), "next": (13, 15)}        # (oy, ox) of each window
    win = {k: tex[oy:oy + H, ox:ox + W] for k, (oy, ox) in off.items()}
    # cur(x) = next(x + fwd) = prev(x + bwd)
    fwd = (off["cur"][1] - off["next"][1], off["cur"][0] - off["next"][0])   # (fx, fy) = (-2, -3)
    bwd = (off["cur"][1] - off["prev"][1], off["cur"][0] - off["prev"][0])   # (3, -2)
    # float32-representable, so the float32 and float64 runs see the same inputs
    proc = {k: rng.random((H, W, 3)).astype(np.float32).astype(np.float64) for k in off}
    last = rng.random((H, W, 3)).astype(np.float32).astype(np.float64)
    flow = lambda f: np.broadcast_to(np.array(f, np.float64), (H, W, 2)).copy()
    ims = [win["prev"], win["cur"], win["next"], proc["prev"], proc["cur"], proc["next"]]
    ci, cp = win["cur"], proc["cur"]
    pi, pp = _slice_warp(win["prev"], *bwd), _slice_warp(proc["prev"], *bwd)
    ni, np_ = _slice_warp(win["next"], *fwd), _slice_warp(proc["next"], *fwd)
    lw = _slice_warp(last, *bwd)
    A, w = _anchor64(ci, cp, pi, pp, ni, np_, lw)
    wp, wn = _weights64(ci, pi, ni)
    # the window is a real translation: most pixels find their match
    assert (wp == CAP_P).mean() > 0.7 and (wn == CAP_N).mean() > 0.7
    expect = _sparse_fixed_point(cp.astype(np.float32), A, w)
    return ims, flow(fwd), flow(bwd), last, expect


def test_translation_window_reaches_the_fixed_point(orc, rng):
    ims, fwd, bwd, last, expect = _case_b(rng)
    prm = dict(pyramid_levels=1, num_iter=3000)
    np.testing.assert_allclose(_frame(orc, ims, fwd, bwd, last, prm), expect, rtol=0, atol=1e-9)
    np.testing.assert_allclose(_frame(orc, ims, fwd, bwd, last, prm, dtype=np.float32), expect, rtol=0, atol=3e-5)
    fma = _frame(orc, ims, fwd, bwd, last, prm, dtype=np.float32, solver="fma")
    np.testing.assert_allclose(fma, expect, rtol=0, atol=3e-5)


@pytest.mark.parametrize("mutation", [1, 2, 3])
def test_translation_pin_detects_mistakes(orc, rng, mutation):
    ims, fwd, bwd, last, expect = _case_b(rng)
    out = _frame(orc, ims, fwd, bwd, last, dict(pyramid_levels=1, num_iter=3000), mutation=mutation)
    assert np.abs(out - expect).max() > 1e-2


# ---- (c) uniform frames ------------------------------------------------------------------

@pytest.mark.parametrize("solver", ["exact", "fma"])
def test_uniform_frames_are_kept(orc, solver):
    """Uniform frames, zero flow, processed = original: every warp returns the
    frame, the anchor equals it, and the solver has nothing to move: each
    output byte is within 1 LSB of the input (u8 -> float -> u8 can floor one
    step down), over a 5-frame sequence at 2 and 3 pyramid levels."""
    H, W = 40, 56
    colours = [(37, 200, 128), (0, 255, 1), (254, 3, 99), (128, 128, 128), (17, 17, 240)]
    rgba = np.stack([np.broadcast_to(np.array(c + (255,), np.uint8), (H, W, 4)) for c in colours])
    zero = np.zeros((5, H, W, 2), np.float32)
    for levels in (2, 3):
        outs, _ = orc.stabilize_sequence(rgba, rgba, zero, zero, params=dict(pyramid_levels=levels), solver=solver)
        for i in range(1, 4):
            d = np.abs(outs[i][..., :3].astype(int) - rgba[i][..., :3].astype(int))
            assert d.max() <= 1, (levels, i)


# ---- (d) level iteration counts at an isolated pixel ----------------------------------------

def _case_d(rng, num_iter):
    """2x2 frames: pixel (0, 0) has no neighbour at level 0 (right needs
    x+1 < W-1 = 1, down y+1 < H-1 = 1) nor at the 1x1 level 1; resampling
    between the levels moves pixel (0, 0) unchanged (sample position 0).  All
    warps of a 2x2 frame sample pixel (0, 0) (clamp to W-3 < 0), and the three
    originals are equal there, so the weights are at their caps."""
    o = rng.random((2, 2, 3))
    ims = [o, o, o] + [rng.random((2, 2, 3)) for _ in range(3)]
    fwd = np.zeros((2, 2, 2))
    bwd = np.zeros((2, 2, 2))
    last = rng.random((2, 2, 3))
    q = lambda a: a[0, 0]
    A, w = _anchor64(q(o), q(ims[4]), q(o), q(ims[3]), q(o), q(ims[5]), q(last))
    a = STEP * w
    e0 = q(ims[4]) - A

    def after(n1, n0, overwrite=True):
        e1 = np.array([_heavy_ball_closed_form(e, ac, MOM, n1) for e, ac in zip(e0, a)])
        start = e1 if overwrite else e0
        return A + np.array([_heavy_ball_closed_form(e, ac, MOM, n0) for e, ac in zip(start, a)])

    return ims, fwd, bwd, last, after


def test_level_iterations_at_an_isolated_pixel(orc, rng):
    n = 7                                             # odd: level 1 runs 7 // 2 = 3 sweeps
    ims, fwd, bwd, last, after = _case_d(rng, n)
    expect = after(n // 2, n)
    # the alternatives a schedule mistake would give are far apart
    for alt in (after(n, n), after(n // 2 + 1, n), after(n // 2, n, overwrite=False), after(0, n)):
        assert np.abs(alt - expect).max() > 1e-4
    prm = dict(pyramid_levels=2, num_iter=n)
    out = _frame(orc, ims, fwd, bwd, last, prm)
    np.testing.assert_allclose(out[0, 0], expect, rtol=0, atol=1e-13)
    out32 = _frame(orc, ims, fwd, bwd, last, prm, dtype=np.float32)
    np.testing.assert_allclose(out32[0, 0], expect, rtol=0, atol=2e-6)
    # one level runs num_iter sweeps and nothing else
    one = _frame(orc, ims, fwd, bwd, last, dict(pyramid_levels=1, num_iter=n))
    np.testing.assert_allclose(one[0, 0], after(0, n, overwrite=False), rtol=0, atol=1e-13)


@pytest.mark.parametrize("mutation", [4, 5, 6])
def test_isolated_pixel_pin_detects_mistakes(orc, rng, mutation):
    n = 7
    ims, fwd, bwd, last, after = _case_d(rng, n)
    out = _frame(orc, ims, fwd, bwd, last, dict(pyramid_levels=2, num_iter=n), mutation=mutation)
    assert np.abs(out[0, 0] - after(n // 2, n)).max() > 1e-4
