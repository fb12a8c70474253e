"""Per-stream frame pipeline (flow network -> FLOWDOWNSCALE -> upsampling ->
stabilization) against the same steps composed by hand through the public
entry points: the pipeline adds wiring only, so outputs must be bit-identical.
Also: feature caching across frames, the reference's backward-flow index
mode, and the consistency fallback (gate 0) reducing to scaled bilinear."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

W, H = 192, 128


def _video(vsc, n=6, seed=21):
    orig, proc, _, _ = vsc.synth_video(width=W, height=H, nframes=n, nblobs=60, motion=3.0, npatches=1,
                                       patch_size=24, seed=seed)
    return orig, proc


def _window(o, p, t):
    return (o[t - 1], o[t], o[t + 1], p[t - 1], p[t], p[t + 1])


@pytest.mark.parametrize("form", ["plain", "chosen"])
def test_pipeline_matches_manual_composition(vsc, cuda, form):
    import torch
    orig, proc = _video(vsc)
    o = torch.from_numpy(orig).to(cuda)
    p = torch.from_numpy(proc).to(cuda)
    kw = vsc.Pipeline.FORMS[form]
    pipe = vsc.Pipeline(W, H, form=form, seed=5)
    pipe.set_last_rgba(p[2])
    net = vsc.FlowNet(W, H, kw["downscale"], kw["flow_dtype"], seed=5, slots=3)
    stab = vsc.VideoStabilizer(W, H, mode=vsc.VSC_FRAME_FMA if kw["solver"] == "fma" else vsc.VSC_FRAME_FUSED)
    stab.set_last_rgba(p[2])
    ups = vsc.Upsampler(params=vsc.load_upsampler_weights()) if kw["upsample"] == "learned" else None
    prev_fwd = None
    for t in range(1, 5):
        got = pipe.frame(_window(o, p, t), t).clone()
        # by hand
        for i in (t - 1, t, t + 1):
            net.extract(o[i], i % 3)
        fwd, bwd = net.flow([(t % 3, (t + 1) % 3), (t % 3, (t - 1) % 3)])
        if ups is not None:
            if prev_fwd is None:
                gate = torch.ones(1, device=cuda)
            else:
                _, gate = vsc.flow_fb_gate(bwd, prev_fwd, 0.25)
            prev_fwd = fwd.clone()
            fwd_h = ups.run(fwd, o[t], scale=2.0, gate=gate)
            bwd_h = ups.run(bwd, o[t], scale=2.0, gate=gate)
        else:
            fwd_h, bwd_h = fwd, bwd
        ref = stab.do_one_step((o[t - 1], o[t], o[t + 1]), (p[t - 1], p[t], p[t + 1]), fwd_h, bwd_h)
        torch.cuda.synchronize()
        assert torch.equal(got, ref), f"frame {t}"


def test_pipeline_flows_and_reference_bwd_index(vsc, cuda):
    import torch
    orig, proc = _video(vsc, seed=3)
    o = torch.from_numpy(orig).to(cuda)
    p = torch.from_numpy(proc).to(cuda)
    net = vsc.FlowNet(W, H, 1, "f32", seed=9, slots=3)
    for i in range(3):
        net.extract(o[i], i)
    f_ref, b_ref = [x.clone() for x in net.flow([(1, 2), (1, 0)])]
    nb_ref = net.flow([(2, 1)])[0].clone()
    pipe = vsc.Pipeline(W, H, seed=9)
    f, b = pipe.flows(_window(o, p, 1), 1)
    assert torch.equal(f, f_ref) and torch.equal(b, b_ref)
    pr = vsc.Pipeline(W, H, seed=9, ref_bwd_index=True)
    f2, b2 = pr.flows(_window(o, p, 1), 1)
    assert torch.equal(f2, f_ref) and torch.equal(b2, nb_ref)


def test_pipeline_fallback_is_scaled_bilinear(vsc, cuda):
    """With the fallback forced (every frame inconsistent), the learned path
    produces exactly the bilinear x2 path's output."""
    import torch
    orig, proc = _video(vsc, seed=8)
    o = torch.from_numpy(orig).to(cuda)
    p = torch.from_numpy(proc).to(cuda)
    a = vsc.Pipeline(W, H, flow_dtype="bf16", downscale=2, upsample="learned", fb_max_inconsistent=0.0,
                     ref_bwd_index=True, seed=4)
    b = vsc.Pipeline(W, H, flow_dtype="bf16", downscale=2, upsample="bilinear", scale_vectors=True,
                     ref_bwd_index=True, seed=4)
    a.set_last_rgba(p[2])
    b.set_last_rgba(p[2])
    for t in range(1, 4):
        ra = a.frame(_window(o, p, t), t).clone()
        rb = b.frame(_window(o, p, t), t).clone()
        g, c = a.gate()
        torch.cuda.synchronize()
        if g == 0.0:
            assert torch.equal(ra, rb)
        assert c > 0 and g == 0.0   # random-weight flows are never perfectly consistent


def test_pipeline_errors_and_footprint(vsc, cuda):
    with pytest.raises(vsc.VscError):
        vsc.Pipeline(W, H, flow_dtype="f32", downscale=3)
    assert vsc.Pipeline.footprint(1920, 1080, "plain") > vsc.Pipeline.footprint(1920, 1080, "chosen") > 0


@pytest.mark.parametrize("form", ["plain", "chosen"])
def test_pipeline_overlap_matches_serial(vsc, cuda, form):
    """Frame-pipelined execution (flows of frame t+1 on one internal stream
    while frame t is stabilized on another) gives the serial pipeline's bits
    on every frame, float outputs and the fallback gate included."""
    import torch
    n = 9
    orig, proc = _video(vsc, n=n + 2, seed=13)
    o = torch.from_numpy(orig).to(cuda)
    p = torch.from_numpy(proc).to(cuda)
    res = {}
    for ov in (False, True):
        pipe = vsc.Pipeline(W, H, form=form, seed=4, overlap=ov)
        pipe.set_last_rgba(p[2])
        outs = [torch.empty((H, W, 4), dtype=torch.uint8, device=cuda) for _ in range(n)]
        flts = [torch.empty((H, W, 3), dtype=torch.float32, device=cuda) for _ in range(n)]
        for t in range(1, n):
            pipe.frame(_window(o, p, t), t, out_rgba=outs[t], out_float=flts[t])
        pipe.sync()
        gate = pipe.gate()
        torch.cuda.synchronize()
        res[ov] = ([x.cpu() for x in outs[1:]], [x.cpu() for x in flts[1:]], gate)
        pipe.close()
    for t in range(n - 1):
        assert torch.equal(res[False][0][t], res[True][0][t]), f"frame {t + 1}"
        assert torch.equal(res[False][1][t], res[True][1][t]), f"frame {t + 1} float"
    assert res[False][2] == res[True][2]
    if form == "plain":
        assert res[True][2] == (0.0, 0)     # no learned upsampler: bilinear, nothing counted


def test_pipeline_refused_params_leave_state_unchanged(vsc, cuda):
    """A frame with invalid parameters is refused before any state change
    (feature slots, previous flow, buffer parity): the next frame with valid
    parameters gives the same bits as a pipeline that never saw the refusal."""
    import torch
    orig, proc = _video(vsc, n=6, seed=23)
    o = torch.from_numpy(orig).to(cuda)
    p = torch.from_numpy(proc).to(cuda)
    ref_pipe = vsc.Pipeline(W, H, form="chosen", seed=2)
    ref_pipe.set_last_rgba(p[2])
    ref = {t: ref_pipe.frame(_window(o, p, t), t).clone() for t in range(1, 5)}
    pipe = vsc.Pipeline(W, H, form="chosen", seed=2)
    pipe.set_last_rgba(p[2])
    bad = vsc.HyperParams()
    bad.pyramidLevels = 9
    neg = vsc.HyperParams()
    neg.numIter = -1
    for t in range(1, 5):
        for prm in (bad, neg):
            with pytest.raises(vsc.VscError) as e:
                pipe.frame(_window(o, p, t), t, params=prm)
            assert e.value.code == vsc.VSC_E_INVALID
        got = pipe.frame(_window(o, p, t), t)
        torch.cuda.synchronize()
        assert torch.equal(got, ref[t]), f"frame {t}"


@pytest.mark.parametrize("size", [(200, 120, 100, 60), (37, 29, 37, 29), (999, 555, 500, 278)])
def test_stream_create_odd_sizes(vsc, cuda, size):
    """Streams whose level images are not multiples of 256 bytes: creation
    checks that its carve ends exactly at the footprint, and a frame at every
    pyramid depth runs (flows upsampled when they arrive downscaled)."""
    import torch
    Wf, Hf, fw, fh = size
    st = vsc.VideoStabilizer(Wf, Hf, fw, fh, mode=vsc.VSC_FRAME_FMA)
    orig, proc, fwd, bwd = vsc.synth_video(width=Wf, height=Hf, nframes=3, nblobs=20, seed=4)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
    st.set_last_rgba(T(proc[2]))
    iy, ix = (np.arange(fh) * Hf) // fh, (np.arange(fw) * Wf) // fw
    lo = lambda f: T(f[iy][:, ix])            # nearest resample to the flow size
    nlev = 0
    w, h = Wf, Hf
    while nlev < 4 and w >= 1 and h >= 1:
        nlev, w, h = nlev + 1, w // 2, h // 2
    for L in range(1, nlev + 1):
        hp = vsc.HyperParams()
        hp.pyramidLevels = L
        hp.numIter = 8
        out = st.do_one_step([T(x) for x in orig], [T(x) for x in proc], lo(fwd[1]), lo(bwd[1]), params=hp)
        torch.cuda.synchronize()
        assert out.shape == (Hf, Wf, 4)
    st.close()
