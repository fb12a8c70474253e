"""GPU parity of the stabilization path through the C ABI against the CPU
oracle: every per-stage entry point, the fused frame (front pass + temporally
blocked solver) and the staged frame, over single frames and sequences."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _img(vsc, a, dev):
    return vsc.GPUImage.from_tensor(_t(a.astype(np.float32), dev))


def test_rgba_float_conversions_exhaustive(vsc, orc, cuda):
    v = np.arange(256, dtype=np.uint8)
    rgba = np.stack([v, v[::-1], np.roll(v, 7), v], -1).reshape(16, 16, 4)
    im = vsc.GPUImage.from_rgba(_t(rgba, cuda))
    assert np.array_equal(im.data.cpu().numpy(), orc.rgba_to_float(rgba))
    vals = np.concatenate([np.linspace(-0.1, 1.2, 3000), [1e12, np.nan, 256 / 255, 1.0]]).astype(np.float32)
    vals = np.concatenate([vals, np.zeros((-len(vals)) % 3, np.float32)]).reshape(1, -1, 3)
    got = vsc.GPUImage.from_tensor(_t(vals, cuda)).to_rgba().cpu().numpy()
    assert np.array_equal(got, orc.float_to_rgba(vals))


@pytest.mark.parametrize("H,W,fc", [(36, 64, 2), (17, 29, 3), (360, 640, 2)])
def test_stab_warp(vsc, orc, cuda, H, W, fc):
    rng = np.random.default_rng(H)
    img = rng.random((H, W, 3)).astype(np.float32)
    fl = rng.uniform(-6, 6, (H, W, fc)).astype(np.float32)
    out = vsc.GPUImage(W, H, 3)
    vsc.get_warp_result(_img(vsc, img, cuda), _img(vsc, fl, cuda), out)
    assert np.array_equal(out.data.cpu().numpy(), orc.stab_warp(img, fl))


def test_adap_comb_and_consist_wt(vsc, orc, cuda):
    rng = np.random.default_rng(5)
    H, W = 45, 77
    ci = rng.random((H, W, 3)).astype(np.float32)
    others = [np.clip(ci + rng.normal(0, 0.02, ci.shape), 0, 1).astype(np.float32) for _ in range(6)]
    pi, ni, cp, pp, np_, ls = others
    adIn, adPr = vsc.GPUImage(W, H, 3), vsc.GPUImage(W, H, 3)
    g = lambda a: _img(vsc, a, cuda)
    for alpha in (6800.0, 0.0, 123.5):
        vsc.get_adap_comb(g(ci), g(cp), g(pi), g(pp), g(ni), g(np_), adIn, adPr, g(ls), alpha)
        rIn, rPr = orc.stab_adap_comb(ci, cp, pi, pp, ni, np_, ls, alpha)
        assert np.array_equal(adIn.data.cpu().numpy(), rIn)
        assert np.array_equal(adPr.data.cpu().numpy(), rPr)
    wt = vsc.GPUImage(W, H, 3)
    for beta, gamma in ((6800.0, 2.0), (10000.0, 10.0), (500.0, 0.0)):
        vsc.get_consist_wt(adIn, g(ci), wt, beta, gamma)
        assert np.array_equal(wt.data.cpu().numpy(), orc.stab_consist_wt(rIn, ci, beta, gamma))


@pytest.mark.parametrize("H,W,iters", [(1, 1, 9), (23, 31, 7), (90, 160, 40), (2, 5, 3)])
def test_consist_out_per_stage(vsc, orc, cuda, H, W, iters):
    rng = np.random.default_rng(W)
    P, A, O = (rng.random((H, W, 3)).astype(np.float32) for _ in range(3))
    Wt = rng.uniform(0, 2, (H, W, 3)).astype(np.float32)
    Wt[rng.random(Wt.shape) < 0.3] = 0
    out = _img(vsc, O, cuda)
    vsc.get_consist_out(_img(vsc, P, cuda), _img(vsc, A, cuda), _img(vsc, Wt, cuda), iters, 0.15, 0.15, out)
    assert np.array_equal(out.data.cpu().numpy(), orc.stab_consist_out(P, A, Wt, O, iters))


@pytest.mark.parametrize("src,dst", [((14, 10), (7, 5)), ((7, 5), (14, 10)), ((13, 11), (29, 23)),
                                     ((960, 540), (1920, 1080))])
def test_bilinear(vsc, orc, cuda, src, dst):
    rng = np.random.default_rng(src[0])
    img = rng.random((src[1], src[0], 3)).astype(np.float32)
    out = vsc.GPUImage(dst[0], dst[1], 3)
    vsc.get_bilinear(_img(vsc, img, cuda), out)
    assert np.array_equal(out.data.cpu().numpy(), orc.stab_bilinear(img, dst[0], dst[1]))


def test_average_frames(vsc, orc, cuda):
    rng = np.random.default_rng(9)
    a, b = (rng.random((12, 20, 3)).astype(np.float32) for _ in range(2))
    out = vsc.GPUImage(20, 12, 3)
    x = _img(vsc, a, cuda)
    # only processed_prev and processed_next are read (flowconsistency.cu:282)
    vsc.perform_consistency(x, _img(vsc, a, cuda), x, _img(vsc, b, cuda), x, x, x, out)
    assert np.array_equal(out.data.cpu().numpy(), orc.average_frames(a, b))


# ---------------------------------------------------------------------------
# whole frames
# ---------------------------------------------------------------------------

def _run_sequence(vsc, orig, proc, fwd, bwd, dev, params=None, mode=0, flow_size=None):
    """FileStabilizer frame loop (videostabilizer.cpp:136-164) on the GPU."""
    n, H, W, _ = orig.shape
    fw, fh = flow_size or (W, H)
    st = vsc.VideoStabilizer(W, H, fw, fh, mode=mode)
    if params:
        for k, v in params.items():
            setattr(st.controlParameters, k, v)
    O = [_t(o, dev) for o in orig]
    Pr = [_t(p, dev) for p in proc]
    F = [_t(f, dev) for f in fwd]
    B = [_t(b, dev) for b in bwd]
    st.set_last_rgba(Pr[2])
    outs, floats = [None] * n, [None] * n
    import torch
    for i in range(1, n - 1):
        of = torch.empty((H, W, 3), dtype=torch.float32, device=dev)
        outs[i] = st.do_one_step(O[i - 1:i + 2], Pr[i - 1:i + 2], F[i], B[i], out_float=of).cpu().numpy()
        floats[i] = of.cpu().numpy()
    st.close()
    return outs, floats


def _oracle_params(params):
    m = {"pyramidLevels": "pyramid_levels", "numIter": "num_iter", "stepSize": "step_size", "momFac": "mom_fac"}
    return {m.get(k, k): v for k, v in (params or {}).items()}


def _solver(mode):
    """The oracle solver a frame mode follows: FMA form for VSC_FRAME_FMA."""
    return "fma" if mode == 2 else "exact"


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_sequence_640x360_bit_exact(vsc, orc, cuda, mode):
    print(f"sequence 640x360 mode={mode}", flush=True)
    orig, proc, fwd, bwd = vsc.synth_video(nframes=6)
    outs, floats = _run_sequence(vsc, orig, proc, fwd, bwd, cuda, mode=mode)
    r_outs, r_floats = orc.stabilize_sequence(orig, proc, fwd, bwd, solver=_solver(mode))
    for i in range(1, 5):
        assert np.array_equal(floats[i], r_floats[i]), i
        assert np.array_equal(outs[i], r_outs[i]), i
    if mode == 2:
        # the FMA form stays within 1 LSB of the reference op order
        e_outs, _ = orc.stabilize_sequence(orig, proc, fwd, bwd)
        for i in range(1, 5):
            d = (outs[i][..., :3].astype(int) - e_outs[i][..., :3].astype(int)) % 256
            assert np.minimum(d, 256 - d).max() <= 1, i


@pytest.mark.parametrize("params", [
    {"pyramidLevels": 1, "numIter": 20},
    {"pyramidLevels": 3, "numIter": 33},
    {"gamma": 0.0, "pyramidLevels": 1, "numIter": 50},
    {"alpha": 100.0, "beta": 9000.0, "gamma": 7.5, "numIter": 17, "stepSize": 0.2, "momFac": 0.5},
    {"numIter": 0},
])
@pytest.mark.parametrize("mode", [0, 2])
def test_frame_params(vsc, orc, cuda, params, mode):
    orig, proc, fwd, bwd = vsc.synth_video(nframes=4, width=160, height=96, nblobs=40, seed=11)
    outs, floats = _run_sequence(vsc, orig, proc, fwd, bwd, cuda, params=params, mode=mode)
    r_outs, r_floats = orc.stabilize_sequence(orig, proc, fwd, bwd, params=_oracle_params(params),
                                              solver=_solver(mode))
    for i in (1, 2):
        assert np.array_equal(floats[i], r_floats[i])
        assert np.array_equal(outs[i], r_outs[i])
    if params.get("gamma") == 0.0 and mode == 0:
        # gamma = 0, one level: output = processed frame, exactly
        assert np.array_equal(outs[1][..., :3], proc[1][..., :3])


@pytest.mark.parametrize("W,H", [(37, 29), (130, 4), (4, 70), (255, 129), (128, 128), (300, 64)])
def test_frame_odd_sizes(vsc, orc, cuda, W, H):
    orig, proc, fwd, bwd = vsc.synth_video(nframes=4, width=W, height=H, nblobs=20, seed=W, patch_size=8)
    for mode in (0, 1, 2):
        outs, floats = _run_sequence(vsc, orig, proc, fwd, bwd, cuda, params={"numIter": 21}, mode=mode)
        r_outs, r_floats = orc.stabilize_sequence(orig, proc, fwd, bwd, params={"num_iter": 21},
                                                  solver=_solver(mode))
        for i in (1, 2):
            assert np.array_equal(floats[i], r_floats[i]), (mode, i)


def test_flowdownscale_upsampled_flows(vsc, orc, cuda):
    orig, proc, fwd, bwd = vsc.synth_video(nframes=4, width=200, height=120, nblobs=50, seed=5)
    fw, fh = 100, 60
    # low-resolution 3-channel flows (network output layout)
    lo_f = np.concatenate([fwd[:, ::2, ::2], np.zeros((4, fh, fw, 1), np.float32)], -1)
    lo_b = np.concatenate([bwd[:, ::2, ::2], np.zeros((4, fh, fw, 1), np.float32)], -1)
    outs, floats = _run_sequence(vsc, orig, proc, lo_f, lo_b, cuda, params={"numIter": 30}, flow_size=(fw, fh))
    up_f = [orc.stab_bilinear(f, 200, 120) for f in lo_f]
    up_b = [orc.stab_bilinear(b, 200, 120) for b in lo_b]
    r_outs, r_floats = orc.stabilize_sequence(orig, proc, up_f, up_b, params={"num_iter": 30})
    for i in (1, 2):
        assert np.array_equal(floats[i], r_floats[i])


@pytest.mark.parametrize("mode", [0, 2])
def test_frame_1080p_single(vsc, orc, cuda, mode):
    print(f"frame 1080p mode={mode}", flush=True)
    orig, proc, fwd, bwd = vsc.synth_video(**dict(vsc.SYNTH_1080P, nframes=3))
    outs, floats = _run_sequence(vsc, orig, proc, fwd, bwd, cuda, mode=mode)
    r_outs, r_floats = orc.stabilize_sequence(orig, proc, fwd, bwd, solver=_solver(mode))
    assert np.array_equal(floats[1], r_floats[1])
    assert np.array_equal(outs[1], r_outs[1])


def test_frame_4k_single_fma(vsc, orc, cuda):
    """A 3840x2160 frame with the FMA-form solver (its tile plan at this size
    differs from 1080p's) against the oracle."""
    print("frame 3840x2160 mode=2", flush=True)
    orig, proc, fwd, bwd = vsc.synth_video(**dict(vsc.SYNTH_1080P, width=3840, height=2160, nframes=3, nblobs=1600))
    outs, floats = _run_sequence(vsc, orig, proc, fwd, bwd, cuda, mode=2)
    r_outs, r_floats = orc.stabilize_sequence(orig, proc, fwd, bwd, solver="fma")
    assert np.array_equal(floats[1], r_floats[1])
    assert np.array_equal(outs[1], r_outs[1])


def test_frame_deterministic_and_stream_independent(vsc, cuda):
    orig, proc, fwd, bwd = vsc.synth_video(nframes=5, width=320, height=180, nblobs=60, seed=3)
    a, fa = _run_sequence(vsc, orig, proc, fwd, bwd, cuda)
    b, fb = _run_sequence(vsc, orig, proc, fwd, bwd, cuda)
    for i in range(1, 4):
        assert np.array_equal(fa[i], fb[i])


def test_frame_errors(vsc, cuda):
    import torch
    st = vsc.VideoStabilizer(64, 48)
    z = torch.zeros((48, 64, 4), dtype=torch.uint8, device=cuda)
    f = torch.zeros((48, 64, 2), dtype=torch.float32, device=cuda)
    with pytest.raises(vsc.VscError):        # no previous frame yet
        st.do_one_step([z, z, z], [z, z, z], f, f)
    st.set_last_rgba(z)
    st.controlParameters.pyramidLevels = 9
    with pytest.raises(vsc.VscError):
        st.do_one_step([z, z, z], [z, z, z], f, f)
    st.controlParameters.pyramidLevels = 2
    st.do_one_step([z, z, z], [z, z, z], f, f)
    st.close()


@pytest.mark.parametrize("H,W,iters", [(37, 53, 19), (300, 500, 8), (1080, 1920, 9), (3, 700, 5),
                                        (40, 70, 1), (40, 70, 2), (130, 260, 4), (128, 128, 30),
                                        (129, 257, 33), (200, 300, 45), (64, 700, 40), (540, 960, 75),
                                        (1080, 1920, 40), (2160, 3840, 25)])
def test_consist_out_planar_blocked(vsc, orc, cuda, H, W, iters):
    """The temporally blocked solver on planar images, checked against the
    oracle (HWC) over tile and launch boundaries.  Sweep counts 1, 2 and 4
    cover a launch with only the level's first sweep, the first sweep plus an
    odd remainder, and the first sweep plus ascending/descending pairs.  Sizes
    of one tile (128 x 128), one tile plus one pixel, two tiles and several,
    and a strip-high image (64 rows) cover the tile placements flush with the
    image edges and the padded placement; the larger sweep counts take
    several launches at the chosen halo (3840 x 2160: the largest serving
    size)."""
    import torch
    rng = np.random.default_rng(H + W)
    P, A, O = (rng.random((H, W, 3)).astype(np.float32) for _ in range(3))
    Wt = rng.uniform(0, 2, (H, W, 3)).astype(np.float32)
    Wt[rng.random(Wt.shape) < 0.2] = 0
    pl = lambda a: torch.from_numpy(np.ascontiguousarray(a.transpose(2, 0, 1))).to(cuda)
    out = pl(O)
    vsc.get_consist_out_planar(pl(P), pl(A), pl(Wt), iters, 0.15, 0.15, out)
    got = out.cpu().numpy().transpose(1, 2, 0)
    assert np.array_equal(got, orc.stab_consist_out(P, A, Wt, O, iters))


@pytest.mark.parametrize("mode_name", ["VSC_FRAME_FUSED", "VSC_FRAME_FMA"])
def test_frame_uniform_invariant_and_zero_iterations(vsc, cuda, mode_name):
    """Through the fused device frame at 640x360: uniform frames with zero flow
    and processed = original come out within 1 LSB of the input, at 2 and 3
    pyramid levels; with num_iter = 0 the output is the pyramid's 2x
    resampling of the processed frame (decimation, then midpoints)."""
    import torch
    W, H = 640, 360
    mode = getattr(vsc, mode_name)
    zero = torch.zeros((H, W, 2), dtype=torch.float32, device=cuda)
    for colour in ((37, 200, 128), (254, 3, 99)):
        rgba = torch.tensor(colour + (255,), dtype=torch.uint8, device=cuda).expand(H, W, 4).contiguous()
        for levels in (2, 3):
            st = vsc.VideoStabilizer(W, H, mode=mode)
            st.set_last_rgba(rgba)
            hp = vsc.HyperParams()
            hp.pyramidLevels = levels
            out = st.do_one_step([rgba] * 3, [rgba] * 3, zero, zero, params=hp).cpu().numpy().astype(int)
            st.close()
            assert np.abs(out[..., :3] - np.array(colour)).max() <= 1, (colour, levels)
    rng = np.random.default_rng(2)
    rgba = [_t(rng.integers(0, 256, (H, W, 4), dtype=np.uint8), cuda) for _ in range(6)]
    st = vsc.VideoStabilizer(W, H, mode=mode)
    st.set_last_rgba(rgba[5])
    hp = vsc.HyperParams()
    hp.numIter = 0
    of = torch.empty((H, W, 3), dtype=torch.float32, device=cuda)
    fl = _t(rng.uniform(-3, 3, (H, W, 2)).astype(np.float32), cuda)
    st.do_one_step(rgba[0:3], rgba[3:6], fl, fl, out_float=of, params=hp)
    got = of.cpu().numpy().astype(np.float64)
    st.close()
    d = (rgba[4].cpu().numpy()[..., :3].astype(np.float32) / np.float32(255.0)).astype(np.float64)[::2, ::2]
    nx = np.minimum(np.arange(d.shape[1]) + 1, d.shape[1] - 1)
    ny = np.minimum(np.arange(d.shape[0]) + 1, d.shape[0] - 1)
    ref = np.empty((H, W, 3))
    ref[0::2, 0::2] = d
    ref[0::2, 1::2] = 0.5 * (d + d[:, nx])
    ref[1::2, 0::2] = 0.5 * (d + d[ny, :])
    ref[1::2, 1::2] = 0.25 * (d + d[:, nx] + d[ny, :] + d[ny][:, nx])
    assert np.abs(got - ref).max() <= 3e-7


def test_flow_dir_frame_loop_matches_oracle(vsc, orc, cuda, tmp_path):
    """The -f frame loop (FileStabilizer with a flow directory): flows written
    as .flo files under the reference's names, frames from host memory; the
    outputs are the oracle's sequence bit for bit (frame 0 and the last frame
    processed, the others stabilized with `last` carried), and a flow of the
    wrong size is refused."""
    W, H, n = 160, 96, 6
    orig, proc, fwd, bwd = vsc.synth_video(width=W, height=H, nframes=n, nblobs=40, motion=3.0, npatches=1,
                                           patch_size=16, seed=8)
    for i in range(1, n - 1):
        vsc.write_flo(tmp_path / f"frame_{i + 1:06d}.flo", fwd[i])
        vsc.write_flo(tmp_path / f"frame_{i:06d}_bwd.flo", bwd[i])
    outs = vsc.stabilize_flow_dir(orig, proc, tmp_path, mode=vsc.VSC_FRAME_FUSED)
    ref, _ = orc.stabilize_sequence(orig, proc, fwd, bwd)
    assert len(outs) == n
    for i in range(n):
        assert np.array_equal(outs[i], ref[i]), f"frame {i}"
    fm = vsc.stabilize_flow_dir(orig, proc, tmp_path, mode=vsc.VSC_FRAME_FMA)
    reff, _ = orc.stabilize_sequence(orig, proc, fwd, bwd, solver="fma")
    for i in range(n):
        assert np.array_equal(fm[i], reff[i]), f"FMA frame {i}"
    vsc.write_flo(tmp_path / f"frame_{3:06d}.flo", np.zeros((H // 2, W // 2, 2), np.float32))
    vsc.write_flo(tmp_path / f"frame_{2:06d}_bwd.flo", np.zeros((H // 2, W // 2, 2), np.float32))
    with pytest.raises(vsc.VscError):
        vsc.stabilize_flow_dir(orig, proc, tmp_path)
