"""Pins the oracle's exp (Cody-Waite reduction + degree-7 polynomial) against
libm double exp: it feeds every photometric weight and its clamps."""
import numpy as np


def _ulp_err(got, ref):
    got = np.float64(got)
    spacing = np.spacing(np.float32(ref)).astype(np.float64)
    return abs(got - ref) / spacing


def test_exp_special_values(orc):
    assert orc.expf(0.0) == 1.0
    assert orc.expf(-0.0) == 1.0
    assert orc.expf(89.0) == float("inf")
    assert orc.expf(-104.0) == 0.0
    assert np.isnan(orc.expf(float("nan")))
    assert orc.expf(1.0) == np.float32(np.e)


def test_exp_error_bound_normal_range(orc, rng):
    xs = np.concatenate([rng.uniform(-87.0, 88.0, 4000), rng.uniform(-1, 1, 2000),
                         -6800.0 * rng.uniform(0, 0.11, 2000) ** 2]).astype(np.float32)
    worst = 0.0
    for x in xs:
        ref = np.exp(np.float64(x))
        if ref < np.finfo(np.float32).tiny:
            continue
        worst = max(worst, _ulp_err(orc.expf(float(x)), ref))
    assert worst <= 2.0, worst


def test_exp_subnormal_and_threshold_range(orc):
    # results below 0.001 are clamped to 0 by the callers; still bounded error
    for x in np.linspace(-103.9, -87.4, 200, dtype=np.float32):
        ref = np.exp(np.float64(x))
        assert abs(orc.expf(float(x)) - ref) <= 2 * 2.0 ** -149 + 1e-6 * ref
    # monotone across the clamp thresholds used by the weights
    xs = np.float32(np.log(np.array([0.001, 0.3, 0.45]))) + np.float32(1e-7) * np.arange(-50, 51)[:, None]
    for col in xs.T:
        vals = [orc.expf(float(v)) for v in np.sort(col)]
        assert all(a <= b for a, b in zip(vals, vals[1:]))
