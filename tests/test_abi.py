"""CPU-side checks of the C ABI: the library loads, exports every function
include/vsc.h declares, and rejects invalid requests on the host with the
documented status codes (the reference throws for these, correlation.h:21-31,
correlation.cc:140-141) — no GPU work is issued."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "vsc.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(vsc_[a-z0-9_]+)\s*\(", src)))


def test_exports_every_declared_symbol(vsc):
    L = vsc.lib()
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(vsc.EXPORTED)


def test_version_and_defaults(vsc):
    assert "gfx950" in vsc.version()
    p = vsc.VscParams()
    assert vsc.lib().vsc_default_params(ctypes.byref(p)) == 0
    h = vsc.HyperParams()
    assert (p.alpha, p.beta, p.gamma, p.pyramid_levels, p.num_iter) == (h.alpha, h.beta, h.gamma,
                                                                       h.pyramidLevels, h.numIter)
    assert abs(p.step_size - 0.15) < 1e-7 and abs(p.mom_fac - 0.15) < 1e-7


def test_host_side_validation(vsc):
    L = vsc.lib()
    nul = None
    dummy = ctypes.c_void_p(16)
    # correlation: null pointers, empty tensors, legacy md > 4, unknown dtype
    assert L.vsc_correlation_forward(nul, dummy, dummy, 1, 1, 1, 1, 4, 0, 0, nul, nul) == -1
    assert L.vsc_correlation_forward(dummy, dummy, dummy, 0, 1, 1, 1, 4, 0, 0, nul, nul) == -1
    assert L.vsc_correlation_forward(dummy, dummy, dummy, 1, 1, 1, 1, 5, 1, 0, nul, nul) == -1
    assert L.vsc_correlation_forward(dummy, dummy, dummy, 1, 1, 1, 1, -1, 0, 0, nul, nul) == -1
    assert L.vsc_correlation_forward(dummy, dummy, dummy, 1, 1, 1, 1, 4, 0, 99, nul, nul) == -2
    assert b"Unsupported" in L.vsc_last_error() or b"unsupported" in L.vsc_last_error()
    assert L.vsc_warp_forward(dummy, dummy, dummy, 1, 1, 1, 1, 2, nul) == -2
    assert L.vsc_warp_forward(dummy, dummy, dummy, 1, 0, 1, 1, 0, nul) == -1
    img = vsc.VscImage(16, 8, 8, 3)
    bad = vsc.VscImage(16, 8, 7, 3)
    assert L.vsc_stab_warp(img, vsc.VscImage(16, 8, 8, 1), img, nul) == -1     # 1-channel flow
    assert L.vsc_stab_warp(img, img, bad, nul) == -1                             # size mismatch
    assert L.vsc_stab_consist_out(img, img, img, -1, 0.1, 0.1, img, nul, nul) == -1
    assert L.vsc_stab_consist_out(img, img, img, 0, 0.1, 0.1, img, nul, nul) == 0  # no sweeps: no work
    assert L.vsc_stab_bilinear(vsc.VscImage(16, 8, 8, 2), img, nul) == -1       # more out channels
    assert L.vsc_stab_consist_out_workspace_size(8, 8) == 2 * 8 * 8 * 3 * 4
    h = ctypes.c_void_p()
    assert L.vsc_stream_create(3, 3, 3, 3, ctypes.byref(h)) == -1
    assert L.vsc_stream_create(64, 64, 128, 64, ctypes.byref(h)) == -1
    assert L.vsc_stream_footprint(1920, 1080, 1920, 1080) > 7 * 3 * 1920 * 1080 * 4
    assert L.vsc_stream_footprint(1920, 1080, 960, 540) > L.vsc_stream_footprint(1920, 1080, 1920, 1080)
    assert L.vsc_stab_frame(None, None, None, None, None, 0, None) == -1
    assert L.vsc_stream_destroy(None) == 0


def test_python_errors_without_gpu(vsc):
    with pytest.raises(vsc.VscError) as e:
        vsc.check(-1)
    assert e.value.code == -1


def test_default_cfgs_fill_every_field(vsc):
    """vsc_pipeline_default_cfg / vsc_server_default_cfg set every field: a
    struct pre-filled with 0xFF bytes comes back equal to a zero-filled one."""
    L = vsc.lib()
    for T, fill in ((vsc.PipelineCfg, L.vsc_pipeline_default_cfg), (vsc.ServerCfg, L.vsc_server_default_cfg)):
        a, b = T(), T()
        ctypes.memset(ctypes.byref(a), 0xFF, ctypes.sizeof(T))
        ctypes.memset(ctypes.byref(b), 0x00, ctypes.sizeof(T))
        fill(ctypes.byref(a))
        fill(ctypes.byref(b))
        # compare field by field (padding bytes may differ)
        def fields(s):
            out = []
            for n, t in s._fields_:
                v = getattr(s, n)
                out.extend(fields(v) if isinstance(v, ctypes.Structure) else [(n, v)])
            return out
        assert fields(a) == fields(b), T.__name__
    c = vsc.PipelineCfg()
    ctypes.memset(ctypes.byref(c), 0xFF, ctypes.sizeof(c))
    L.vsc_pipeline_default_cfg(ctypes.byref(c))
    assert c.overlap == 0
    # the pipeline's footprint accepts the defaults (a field left unset would
    # make the configuration invalid and the footprint 0)
    assert L.vsc_pipeline_footprint(640, 360, ctypes.byref(c)) > 0


def test_pipeline_flow_batch_config(vsc):
    """flow_batch defaults to 1; 1..4 are accepted with a footprint that grows
    with the batch (more feature slots, flow buffer sets and decode room);
    other values make the configuration invalid; the server takes only 1."""
    L = vsc.lib()
    c = vsc.PipelineCfg()
    L.vsc_pipeline_default_cfg(ctypes.byref(c))
    assert c.flow_batch == 1
    c.flow_dtype, c.downscale, c.upsample = vsc.VSC_BF16, 2, 1
    sizes = []
    for b in (1, 2, 3, 4):
        c.flow_batch = b
        sizes.append(L.vsc_pipeline_footprint(1920, 1080, ctypes.byref(c)))
    assert all(x > 0 for x in sizes) and sizes == sorted(sizes) and len(set(sizes)) == 4
    for b in (0, 5):
        c.flow_batch = b
        assert L.vsc_pipeline_footprint(1920, 1080, ctypes.byref(c)) == 0
    s = vsc.ServerCfg()
    L.vsc_server_default_cfg(ctypes.byref(s))
    assert s.form.flow_batch == 1
    s.form.flow_batch = 2
    h = ctypes.c_void_p()
    ups = (ctypes.c_float * L.vsc_upsampler_param_count())()
    assert L.vsc_server_create(ctypes.byref(s), ups, ctypes.byref(h)) == -1
    assert b"flow_batch" in L.vsc_last_error()


def test_server_rejects_overlap_form(vsc):
    L = vsc.lib()
    c = vsc.ServerCfg()
    L.vsc_server_default_cfg(ctypes.byref(c))
    c.form.overlap = 1
    h = ctypes.c_void_p()
    ups = (ctypes.c_float * L.vsc_upsampler_param_count())()
    assert L.vsc_server_create(ctypes.byref(c), ups, ctypes.byref(h)) == -1
    assert b"overlap" in L.vsc_last_error()


@pytest.mark.parametrize("size", [(200, 120, 100, 60), (37, 29, 37, 29), (160, 96, 160, 96), (999, 555, 500, 278),
                                  (1279, 719, 640, 360), (1920, 1080, 960, 540)])
def test_stream_footprint_covers_the_carve(vsc, size):
    """The stream's block is carved into 8 planar images per pyramid level
    (plus two frame-size flows when the flows arrive downscaled), each starting
    on a 256-byte boundary.  The footprint must hold every piece with its
    padding: at least sum(ceil256(piece)) bytes, counted here from the level
    dimensions W>>j x H>>j down to 1 pixel."""
    L = vsc.lib()
    W, H, fw, fh = size
    need = 0
    w, h = W, H
    levels = 0
    while levels < 4 and w >= 1 and h >= 1:
        need += 8 * (-(-(3 * w * h * 4) // 256) * 256)
        w //= 2
        h //= 2
        levels += 1
    if (fw, fh) != (W, H):
        need += 2 * (-(-(3 * W * H * 4) // 256) * 256)
    got = L.vsc_stream_footprint(W, H, fw, fh)
    assert got >= need, (got, need)
    assert got % 256 == 0


def test_workspace_entry_points_check_on_the_host(vsc):
    """The workspace forms refuse a missing or short workspace before any GPU
    work; FP32 and FP64 correlation need none."""
    L = vsc.lib()
    dummy = ctypes.c_void_p(16)
    n = L.vsc_correlation_workspace_size(1, 33, 17, 30, vsc.VSC_BF16)
    assert n == 2 * (-(-(17 * 30 * 64 * 2) // 256) * 256)          # channels padded to 64, 2 tensors
    assert L.vsc_correlation_workspace_size(1, 32, 17, 30, vsc.VSC_FP8E4M3) == 2 * (-(-(17 * 30 * 32) // 256) * 256)
    assert L.vsc_correlation_workspace_size(1, 32, 17, 30, vsc.VSC_F32) == 0
    assert L.vsc_correlation_forward_ws(dummy, dummy, dummy, 1, 33, 17, 30, 4, 0, vsc.VSC_BF16, None, None, 0,
                                        None) == -1
    assert L.vsc_correlation_forward_ws(dummy, dummy, dummy, 1, 33, 17, 30, 4, 0, vsc.VSC_BF16, None, dummy, n - 1,
                                        None) == -1
    assert b"workspace" in L.vsc_last_error()
    assert L.vsc_quantize_fp8_pair_workspace_size(196) == (3 * 196 + 4) * 4
    assert L.vsc_quantize_fp8_pair_ws(dummy, dummy, dummy, dummy, dummy, 1, 196, 4, 4, dummy, 100, None) == -1
