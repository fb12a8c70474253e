"""Host logic of the temporally blocked consistency solver (CPU only, no GPU
work): the launch plan chosen per pyramid level and the tile placement along
an axis, checked against the properties that make the blocked result equal to
one sweep per launch (flowconsistency.cu:193-258): every pixel is stored by
exactly one tile, and a stored pixel lies at least `halo` pixels (>= the
sweeps of a launch) from every tile edge that is not an image edge."""
import pytest


def test_plan_1080p(vsc):
    # FMA form: halo 19 at both levels of a 1920x1080 frame, full-height tiles
    assert vsc.blocked_plan(1920, 1080, 150, True) == (19, 128, 21, 12)
    assert vsc.blocked_plan(960, 540, 75, True) == (19, 128, 11, 6)
    halo, rows, _, _ = vsc.blocked_plan(1920, 1080, 150, False)
    assert rows in (112, 56) and 1 <= halo <= 30


@pytest.mark.parametrize("w,h,iters", [(1920, 1080, 150), (960, 540, 75), (1280, 720, 150), (640, 360, 75),
                                       (640, 360, 150), (3840, 2160, 150), (1920, 1080, 75), (255, 129, 21)])
@pytest.mark.parametrize("fma", [True, False])
def test_plan_properties(vsc, w, h, iters, fma):
    halo, rows, tx, ty = vsc.blocked_plan(w, h, iters, fma)
    assert rows in ((128, 64) if fma else (112, 56))
    assert rows - 2 * halo >= 8 and 128 - 2 * halo >= 8
    # the plan's tile grid is the placement's tile count along each axis
    assert vsc.tile_span(w, 128, halo, 0)[0] == tx
    assert vsc.tile_span(h, rows, halo, 0)[0] == ty
    # a launch of 209..256 tiles (one nearly full round) is avoided when
    # another halo fits, as it is at these sizes
    assert not (208 < 3 * tx * ty <= 256)


def _spans(vsc, L, T, h):
    n = vsc.tile_span(L, T, h, 0)[0]
    return [vsc.tile_span(L, T, h, i)[1:] for i in range(n)]


@pytest.mark.parametrize("T", [128, 64, 112, 56])
def test_tile_placement(vsc, T):
    for h in (4, 9, 17, 19, 25, 30):
        O = T - 2 * h
        if O < 8:
            continue
        lengths = sorted(set(list(range(1, 400, 13)) + [T - 1, T, T + 1, 2 * T - 2 * h, 2 * T - 2 * h + 1,
                                                        540, 960, 1080, 1920, 2160, 3840]))
        for L in lengths:
            spans = _spans(vsc, L, T, h)
            prev_hi = 0
            for i, (s, lo, hi) in enumerate(spans):
                assert lo == prev_hi and lo <= hi, (L, T, h, i)     # in order, no gap, no overlap
                prev_hi = hi
                if L >= T:
                    assert 0 <= s and s + T <= L, (L, T, h, i)   # flush: every tile pixel in the image
                else:
                    assert s == i * O - h, (L, T, h, i)          # padded placement
                # stored pixels are exact: h from every edge facing another tile
                assert lo - s >= h or (s == 0 and L >= T), (L, T, h, i)
                assert s + T - hi >= h or (s + T == L and L >= T), (L, T, h, i)
            assert spans[0][1] == 0 and prev_hi == L, (L, T, h)
            # flush placement never needs more tiles than the padded one
            assert len(spans) <= -(-L // O), (L, T, h)


def test_plan_and_span_reject_bad_requests(vsc):
    with pytest.raises(vsc.VscError):
        vsc.blocked_plan(1920, 1080, 0)
    with pytest.raises(vsc.VscError):
        vsc.blocked_plan(0, 1080, 150)
    with pytest.raises(vsc.VscError):
        vsc.tile_span(100, 128, 64, 0)       # 128 - 2*64 < 1: no output per tile
    n = vsc.tile_span(1000, 128, 19, 0)[0]
    with pytest.raises(vsc.VscError):
        vsc.tile_span(1000, 128, 19, n)      # index out of range
