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
