"""CPU checks of the flow-network oracle (oracle/pwcnet_light.py): its cost
volume and warp against the C restatement of the reference ops, its layer
table against the architecture, and closed forms of the final resize."""
import numpy as np
import pytest
import torch

from oracle import pwcnet_light as pw


def test_layer_table_matches_architecture():
    shapes = pw.layer_shapes()
    assert len(shapes) == 18 + 3 + 24 + 7
    # DecoderLight input channels: feature + 81 + 2 (pwcnet.py:466-469)
    assert [s[0][1] for s in shapes[21:45:6]] == [196 + 83, 128 + 83, 96 + 83, 64 + 83]
    # refiner on cat([x5, flow]) = 34 channels (pwcnet.py:880-882)
    assert shapes[45][0] == (128, 34, 3, 3)
    assert shapes[18][0] == (2, 2, 4, 4)


def test_correlation_matches_c_oracle(orc):
    rng = np.random.default_rng(0)
    f1 = rng.standard_normal((2, 7, 9, 13))
    f2 = rng.standard_normal((2, 7, 9, 13))
    got = pw.correlation(torch.from_numpy(f1), torch.from_numpy(f2)).numpy()
    ref = orc.correlation(f1, f2, 4, dtype=np.float64).reshape(2, 81, 9, 13)
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_warp_matches_c_oracle(orc):
    rng = np.random.default_rng(1)
    x = rng.standard_normal((1, 5, 11, 17))
    # flows on a 1/8 grid plus 1/64: taps never sit on the 0.999 mask boundary
    fl = (rng.integers(-24, 24, (1, 11, 17, 2)) / 8.0 + 1 / 64).astype(np.float32)
    got = pw.warp(torch.from_numpy(x), fl, 1.0).numpy()
    ref = orc.warp(x, np.ascontiguousarray(fl.transpose(0, 3, 1, 2)).astype(np.float64), dtype=np.float64)
    assert np.allclose(got, ref, rtol=1e-6, atol=1e-6)
    # scale multiplies the flow
    got2 = pw.warp(torch.from_numpy(x), fl * 2, 0.5).numpy()
    assert np.array_equal(got, got2)


def test_final_flow_constant_field():
    # a constant flow stays constant through both resizes, times 20 and the
    # wrapper's per-channel factors (H/Hp for channel 0, W/Wp for channel 1)
    fl = torch.ones((1, 2, 9, 13), dtype=torch.float64) * torch.tensor([0.5, -0.25], dtype=torch.float64)[None, :, None, None]
    out = pw.final_flow(fl, 72, 104, 70, 100)
    assert out.shape == (1, 70, 100, 2)
    assert np.allclose(out[..., 0].numpy(), 10.0 * 70 / 72)
    assert np.allclose(out[..., 1].numpy(), -5.0 * 100 / 104)


def test_preprocess_shapes_and_range():
    rgba = np.random.default_rng(2).integers(0, 256, (70, 100, 4), dtype=np.uint8)
    x, Hd, Wd = pw.preprocess(rgba, 1)
    assert (Hd, Wd) == (70, 100) and x.shape == (1, 3, 128, 128)
    assert float(x.min()) >= 0.0 and float(x.max()) <= 1.0
    x2, Hd2, Wd2 = pw.preprocess(rgba, 2)
    assert (Hd2, Wd2) == (35, 50) and x2.shape == (1, 3, 64, 64)
