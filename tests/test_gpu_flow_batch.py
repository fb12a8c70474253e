"""Flow batch size B (the reference's batchSize: FlowModel::run over B frame
pairs every B frames, videostabilizer.cpp:65-71,177,267-279): a pipeline
with flow_batch B stabilizes B consecutive windows per call, their forward
and backward flows decoded in one network batch of 2B pairs.  Each flow has
the same bits in any batch (the convolution split depends on one item's
geometry only), so every output frame, and the learned-upsampler gate, equals
the one-frame-per-call pipeline's, in both network precisions and with the FP32
network's cost volume on the BF16 MFMA, with frame pipelining, and with a
short last batch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _video(vsc, W, H, n, seed):
    orig, proc, _, _ = vsc.synth_video(width=W, height=H, nframes=n, nblobs=80, motion=3.0, npatches=1,
                                       patch_size=24, seed=seed)
    return orig, proc


def _window(o, p, t):
    return (o[t - 1], o[t], o[t + 1], p[t - 1], p[t], p[t + 1])


# the two named forms, and the FP32 network at half resolution with the cost
# volume's products in BF16 (learned upsampling)
FORM_KW = {
    "chosen": dict(form="chosen"),
    "plain": dict(form="plain"),
    "bf16corr": dict(flow_dtype="f32", downscale=2, upsample="learned", solver="fma", corr_dtype="bf16"),
}


def _single(vsc, cuda, o, p, W, H, form, frames, overlap=False):
    import torch
    pipe = vsc.Pipeline(W, H, seed=7, flow_head_scale=0.05, overlap=overlap, **FORM_KW[form])
    pipe.set_last_rgba(p[2])
    outs, gates = [], []
    for t in frames:
        of = torch.empty((H, W, 3), dtype=torch.float32, device=cuda)
        r = pipe.frame(_window(o, p, t), t, out_float=of)
        if overlap:
            pipe.sync()
        outs.append((r.clone(), of.clone()))
        gates.append(pipe.gate())
    pipe.close()
    return outs, gates


def _batched(vsc, cuda, o, p, W, H, form, frames, B, overlap=False):
    import torch
    pipe = vsc.Pipeline(W, H, seed=7, flow_head_scale=0.05, overlap=overlap, flow_batch=B, **FORM_KW[form])
    pipe.set_last_rgba(p[2])
    outs, gates = [], []
    for i in range(0, len(frames), B):
        ts = frames[i:i + B]
        of = [torch.empty((H, W, 3), dtype=torch.float32, device=cuda) for _ in ts]
        rs = pipe.frames([_window(o, p, t) for t in ts], ts[0], out_float=of)
        if overlap:
            pipe.sync()
        outs += [(r.clone(), f.clone()) for r, f in zip(rs, of)]
        gates.append(pipe.gate())   # the gate of the batch's last frame
    pipe.close()
    return outs, gates


@pytest.mark.parametrize("form,B,overlap", [("chosen", 4, False), ("chosen", 3, True), ("plain", 2, False),
                                             ("bf16corr", 2, True)])
def test_flow_batch_equals_one_frame_per_call(vsc, cuda, form, B, overlap):
    import torch
    W, H = (192, 128) if form == "plain" else (320, 192)
    n = 12
    print(f"flow batch {form} B={B} overlap={overlap} {W}x{H}", flush=True)
    orig, proc = _video(vsc, W, H, n, seed=31)
    o = torch.from_numpy(orig).to(cuda)
    p = torch.from_numpy(proc).to(cuda)
    frames = list(range(1, n - 1))          # 10 frames: the last batch is short unless B divides 10
    ref, ref_gates = _single(vsc, cuda, o, p, W, H, form, frames)
    got, got_gates = _batched(vsc, cuda, o, p, W, H, form, frames, B, overlap)
    torch.cuda.synchronize()
    assert len(got) == len(ref)
    for i, ((ra, fa), (rb, fb)) in enumerate(zip(ref, got)):
        assert torch.equal(ra, rb), (form, B, frames[i])
        assert torch.equal(fa, fb), (form, B, frames[i])
    last = [ref_gates[min(i + B, len(frames)) - 1] for i in range(0, len(frames), B)]
    assert got_gates == last


def test_flow_batch_refusals(vsc, cuda):
    import torch
    W, H = 192, 128
    orig, proc = _video(vsc, W, H, 6, seed=2)
    o = torch.from_numpy(orig).to(cuda)
    p = torch.from_numpy(proc).to(cuda)
    pipe = vsc.Pipeline(W, H, form="plain", seed=1, flow_batch=2)
    pipe.set_last_rgba(p[2])
    with pytest.raises(vsc.VscError) as e:
        pipe.frames([_window(o, p, t) for t in (1, 2, 3)], 1)
    assert e.value.code == vsc.VSC_E_INVALID
    # the refused call changed nothing: the next batch runs as a fresh one would
    a = pipe.frames([_window(o, p, t) for t in (1, 2)], 1)
    ref = vsc.Pipeline(W, H, form="plain", seed=1, flow_batch=2)
    ref.set_last_rgba(p[2])
    b = ref.frames([_window(o, p, t) for t in (1, 2)], 1)
    torch.cuda.synchronize()
    assert all(torch.equal(x, y) for x, y in zip(a, b))
    for bad in (0, 5):
        with pytest.raises(vsc.VscError):
            vsc.Pipeline(W, H, form="plain", flow_batch=bad)


def test_flow_batch_overlap_back_to_back_with_jump(vsc, cuda):
    """Frame-pipelined batches submitted back to back (no sync between calls,
    so a batch's feature extraction runs on its own stream beside the previous
    batch's decode), with short batches and a jump in the frame index (the
    jump re-extracts into slots the decode before it may still read), equal
    the one-frame-per-call pipeline frame for frame."""
    import torch
    W, H, n = 320, 192, 15
    orig, proc = _video(vsc, W, H, n, seed=37)
    o = torch.from_numpy(orig).to(cuda)
    p = torch.from_numpy(proc).to(cuda)
    groups = [[1, 2], [3, 4], [5], [8, 9], [10, 11], [12], [3, 4], [5, 6]]
    frames = [t for g in groups for t in g]
    ref, ref_gates = _single(vsc, cuda, o, p, W, H, "chosen", frames)
    pipe = vsc.Pipeline(W, H, form="chosen", seed=7, flow_head_scale=0.05, overlap=True, flow_batch=2)
    pipe.set_last_rgba(p[2])
    got = []
    for g in groups:
        of = [torch.empty((H, W, 3), dtype=torch.float32, device=cuda) for _ in g]
        rs = pipe.frames([_window(o, p, t) for t in g], g[0], out_float=of)
        got += list(zip(rs, of))
    pipe.sync()
    gate = pipe.gate()
    torch.cuda.synchronize()
    pipe.close()
    for i, ((ra, fa), (rb, fb)) in enumerate(zip(ref, got)):
        assert torch.equal(ra, rb), frames[i]
        assert torch.equal(fa, fb), frames[i]
    assert gate == ref_gates[-1]
