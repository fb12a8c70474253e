"""PWCNet-light flow network on the library's kernels against the float64
oracle network (oracle/pwcnet_light.py) with the same parameters.

Each decoder level is checked with teacher forcing at the warp: the oracle
warps with the device's upsampled flow, so both sides take the same bilinear
taps and the same 0.999 mask decisions (a discontinuous function of the flow)
and the comparison measures arithmetic error only.  Errors are normalised by
the largest magnitude of the reference tensor.

Tolerances: FP32 storage 2e-5 (float32 through ~30 conv layers; measured
errors ~3e-6, printed); BF16 storage 4e-2 (bfloat16 activations, 2^-9 relative
rounding per layer).
"""
import numpy as np
import pytest

from oracle import pwcnet_light as pw

pytestmark = pytest.mark.gpu

TOL = {"f32": 2e-5, "bf16": 4e-2}


def _frames(vsc, W, H, seed=3):
    orig, _, _, _ = vsc.synth_video(width=W, height=H, nframes=2, nblobs=80, motion=3.0, npatches=1,
                                    patch_size=max(8, W // 8), seed=seed)
    return orig


def _rel(dev, ref):
    ref = np.asarray(ref, dtype=np.float64)
    return float(np.max(np.abs(np.asarray(dev, np.float64) - ref)) / (np.max(np.abs(ref)) + 1e-3))


@pytest.mark.parametrize("kind", ["f32", "bf16"])
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
