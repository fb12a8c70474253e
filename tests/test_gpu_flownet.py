"""PWCNet-light flow network on the library's kernels against the float64
oracle network (oracle/pwcnet_light.py) with the same parameters.

Each decoder level is checked with teacher forcing at the warp: the oracle
warps with the device's upsampled flow, so both sides take the same bilinear
taps and the same 0.999 mask decisions (a discontinuous function of the flow)
and the comparison measures arithmetic error only.  Errors are normalised by
the largest magnitude of the reference tensor.

Tolerances: FP32 storage 2e-5 (float32 through ~30 conv layers; measured
errors ~3e-6, printed); BF16 storage 4e-2 (bfloat16 activations, 2^-9 relative
rounding per layer); FP32 storage with the cost volume's products on the BF16
MFMA 5e-3 (the features rounded to bfloat16 before the dot products of each
level, everything else FP32; measured errors up to 1.2e-3).
"""
import numpy as np
import pytest

from oracle import pwcnet_light as pw

pytestmark = pytest.mark.gpu

TOL = {"f32": 2e-5, "bf16": 4e-2, "f32+bf16corr": 5e-3}


def _frames(vsc, W, H, seed=3):
    orig, _, _, _ = vsc.synth_video(width=W, height=H, nframes=2, nblobs=80, motion=3.0, npatches=1,
                                    patch_size=max(8, W // 8), seed=seed)
    return orig


def _rel(dev, ref):
    ref = np.asarray(ref, dtype=np.float64)
    return float(np.max(np.abs(np.asarray(dev, np.float64) - ref)) / (np.max(np.abs(ref)) + 1e-3))


@pytest.mark.parametrize("kind", ["f32", "bf16", "f32+bf16corr"])
@pytest.mark.parametrize("W,H,down", [(192, 128, 1), (200, 136, 2), (320, 256, 1)])
def test_flownet_levels(vsc, cuda, kind, W, H, down):
    import torch
    print(f"flownet {kind} {W}x{H} down={down}", flush=True)
    orig = _frames(vsc, W, H)
    net = vsc.FlowNet(W, H, down, kind, seed=11, slots=3)
    fr = torch.from_numpy(orig).to(cuda)
    net.extract(fr[0], 0)
    net.extract(fr[1], 1)
    outs = net.flow([(0, 1), (1, 0)], out_channels=3)
    torch.cuda.synchronize()
    P = pw.unpack_params(net.params())
    x0, Hd, Wd = pw.preprocess(orig[0], down)
    x1, _, _ = pw.preprocess(orig[1], down)
    assert net.flow_size == (Wd, Hd)
    F0, F1 = pw.extractor(P, x0), pw.extractor(P, x1)
    Hp, Wp = x0.shape[2:]
    worst = 0.0
    for item, (fa, fb) in enumerate([(F0, F1), (F1, F0)]):
        ups = {lvl: net.intermediate(0, lvl, item).cpu().numpy()[None] for lvl in (5, 4, 3)}
        res = pw.flow_pair(P, fa, fb, ups)
        errs = {}
        for lvl in (6, 5, 4, 3):
            errs[f"flow{lvl}"] = _rel(net.intermediate(1, lvl, item).cpu().numpy(), res[f"flow{lvl}"][0])
        for lvl in (5, 4, 3):
            errs[f"up{lvl}"] = _rel(ups[lvl][0], res[f"up{lvl}"][0])
        errs["refined"] = _rel(net.intermediate(2, 3, item).cpu().numpy(), res["refined"][0])
        fin = pw.final_flow(res["refined_nchw"], Hp, Wp, Hd, Wd)[0].numpy()
        got = outs[item].cpu().numpy()
        errs["final"] = _rel(got[..., :2], fin)
        assert np.all(got[..., 2] == 0.0)
        print(f"  item {item}: " + " ".join(f"{k}={v:.2e}" for k, v in errs.items()) +
              f"  |flow|max={np.abs(fin).max():.2f}")
        worst = max(worst, max(errs.values()))
    assert worst <= TOL[kind]


def test_flownet_bf16_cost_volume_between_fp32_and_bf16(vsc, cuda):
    """The BF16 cost volume in an FP32 network changes the flows (its products
    are BF16), by less than the all-BF16 network does; a BF16 network refuses
    FP32 cost-volume products."""
    import torch
    W, H = 320, 256
    orig = _frames(vsc, W, H)
    fr = torch.from_numpy(orig).to(cuda)
    flows = {}
    for kind in ("f32", "f32+bf16corr", "bf16"):
        net = vsc.FlowNet(W, H, 1, kind, seed=11, slots=3)
        net.extract(fr[0], 0)
        net.extract(fr[1], 1)
        flows[kind] = net.flow([(0, 1)])[0].cpu().numpy().astype(np.float64)
        if kind == "bf16":
            assert vsc.lib().vsc_flownet_set_corr_dtype(net.h, vsc.VSC_F32) == vsc.VSC_E_DTYPE
        net.close()
    ref = flows["f32"]
    scale = np.abs(ref).max() + 1e-3
    e_mix = np.abs(flows["f32+bf16corr"] - ref).max() / scale
    e_bf = np.abs(flows["bf16"] - ref).max() / scale
    print(f"vs FP32 network: BF16 cost volume {e_mix:.2e}, BF16 network {e_bf:.2e}")
    assert 0.0 < e_mix < e_bf


@pytest.mark.parametrize("kind", ["f32", "bf16"])
def test_flownet_batch_matches_single(vsc, cuda, kind):
    """A pair decoded alone and batched with a second pair gives the same bits."""
    import torch
    orig = _frames(vsc, 256, 192, seed=5)
    net = vsc.FlowNet(256, 192, 1, kind, seed=2, slots=2)
    fr = torch.from_numpy(orig).to(cuda)
    net.extract(fr[0], 0)
    net.extract(fr[1], 1)
    both = [t.clone() for t in net.flow([(0, 1), (1, 0)])]
    alone = net.flow([(1, 0)])[0]
    torch.cuda.synchronize()
    assert torch.equal(both[1], alone)
    assert not torch.equal(both[0], both[1])


def test_flownet_params_roundtrip_and_errors(vsc, cuda):
    import torch
    net = vsc.FlowNet(128, 64, 1, "f32", seed=1, slots=2)
    p = net.params()
    assert p.size == sum(int(np.prod(s[0])) + s[1] for s in pw.layer_shapes())
    q = p * 0.5
    net.set_params(q)
    assert np.array_equal(net.params(), q)
    with pytest.raises(vsc.VscError):
        net.set_params(q[:-1])
    with pytest.raises(vsc.VscError):
        vsc.FlowNet(128, 64, 3, "f32")
    with pytest.raises(vsc.VscError):
        net.extract(torch.zeros((64, 128, 4), dtype=torch.uint8, device=cuda), 2)
    assert vsc.FlowNet.footprint(1920, 1080, 1, "bf16", 3) > vsc.FlowNet.footprint(1920, 1080, 2, "bf16", 3) > 0
