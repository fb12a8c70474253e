"""MI355X (gfx950) per-frame consistency hot path of video-stream-consistency.

Python mirror of the reference interfaces over the C ABI of ``libvsc.so``
(``include/vsc.h``):

* the ONNX Runtime custom ops ``Correlation`` and ``Warp``
  (src/ort_custom_ops/include/ort_custom_ops/opticalflow/correlation.h:10-84,
  warp.h:11-60) as :func:`correlation` and :func:`warp`;
* the stabilization stage functions of src/stabilization/flowconsistency.cuh:3-50
  (``get_warp_result``, ``get_adap_comb``, ``get_consist_wt``,
  ``get_consist_out``, ``get_bilinear``, ``perform_consistency``) on
  :class:`GPUImage` objects;
* :class:`VideoStabilizer` with ``hyperParams`` and ``doOneStep``
  (src/stabilization/videostabilizer.h:38-110).

torch supplies device memory and the current HIP stream; every computation
runs in the hand-written HIP kernels of ``libvsc.so``.  There is no fallback:
if the library is missing, loading fails with an error.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvsc.so")
SYNTH_PATH = os.path.join(_HERE, "libvsc_synth.so")

VSC_OK = 0
VSC_E_INVALID, VSC_E_DTYPE, VSC_E_HIP, VSC_E_NOMEM, VSC_E_UNSUPPORTED, VSC_E_AGAIN = -1, -2, -3, -4, -5, -6
VSC_F32, VSC_F64, VSC_BF16, VSC_F16, VSC_FP8E4M3 = 0, 1, 2, 3, 4
VSC_FRAME_FUSED, VSC_FRAME_STAGED, VSC_FRAME_FMA = 0, 1, 2
SOLVERS = {"exact": 0, "fma": 1}    # stabilization solver: reference op order | FMA form
VSC_MAX_LEVELS = 4

#: every function the C header declares (checked against the built library)
EXPORTED = (
    "vsc_last_error", "vsc_version", "vsc_default_params", "vsc_correlation_forward",
    "vsc_quantize_fp8_pair", "vsc_convert_f32", "vsc_warp_forward", "vsc_stab_warp",
    "vsc_stab_adap_comb", "vsc_stab_consist_wt", "vsc_stab_consist_out_workspace_size",
    "vsc_stab_consist_out", "vsc_stab_bilinear", "vsc_average_frames", "vsc_rgba_to_float",
    "vsc_float_to_rgba", "vsc_stream_create", "vsc_stream_destroy", "vsc_stream_footprint",
    "vsc_stream_set_last_rgba", "vsc_stream_set_last", "vsc_stream_copy_last", "vsc_stab_frame",
    "vsc_profile_begin", "vsc_profile_end", "vsc_stab_consist_out_planar_workspace_size",
    "vsc_stab_blocked_plan", "vsc_stab_tile_span",
    "vsc_stab_consist_out_planar", "vsc_conv3x3_nhwc", "vsc_flownet_create", "vsc_flownet_destroy",
    "vsc_flownet_footprint", "vsc_flownet_flow_size", "vsc_flownet_param_count", "vsc_flownet_get_params",
    "vsc_flownet_set_params", "vsc_flownet_extract", "vsc_flownet_flow", "vsc_flownet_intermediate",
    "vsc_flownet_scale_flow_heads", "vsc_flownet_set_corr_dtype",
    "vsc_upsampler_create", "vsc_upsampler_destroy", "vsc_upsampler_param_count", "vsc_upsampler_get_params",
    "vsc_upsampler_set_params", "vsc_upsampler_workspace_size", "vsc_upsampler_run", "vsc_upsampler_train_step",
    "vsc_upsampler_loss_grad", "vsc_flow_fb_gate", "vsc_pipeline_default_cfg", "vsc_pipeline_create",
    "vsc_pipeline_destroy", "vsc_pipeline_footprint", "vsc_pipeline_set_upsampler_params", "vsc_pipeline_flownet",
    "vsc_pipeline_stream", "vsc_pipeline_flows", "vsc_pipeline_frame", "vsc_pipeline_frames", "vsc_pipeline_gate",
    "vsc_pipeline_sync",
    "vsc_server_default_cfg", "vsc_server_create", "vsc_server_destroy", "vsc_server_stream_footprint",
    "vsc_server_used_bytes", "vsc_server_open", "vsc_server_close", "vsc_server_set_params", "vsc_server_submit",
    "vsc_server_tick", "vsc_server_poll", "vsc_server_drain", "vsc_server_read", "vsc_flo_read", "vsc_flo_write",
    "vsc_flow_dir_path", "vsc_flow_dir_read", "vsc_stabilize_flow_dir", "vsc_correlation_workspace_size",
    "vsc_correlation_forward_ws", "vsc_quantize_fp8_pair_workspace_size", "vsc_quantize_fp8_pair_ws",
    "vsc_correlation_forward_noalloc",
)


#: vsc_frame_sink: (user, index, rgba, width, height) -> int
FRAME_SINK = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_longlong, ctypes.c_void_p, ctypes.c_int,
                              ctypes.c_int)


class VscImage(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("width", ctypes.c_int), ("height", ctypes.c_int),
                ("channels", ctypes.c_int)]


class VscParams(ctypes.Structure):
    _fields_ = [("alpha", ctypes.c_float), ("beta", ctypes.c_float), ("gamma", ctypes.c_float),
                ("pyramid_levels", ctypes.c_int), ("num_iter", ctypes.c_int),
                ("step_size", ctypes.c_float), ("mom_fac", ctypes.c_float)]


class VscFrameIn(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("orig_prev", "orig_cur", "orig_next", "proc_prev",
                                                "proc_cur", "proc_next", "flow_fwd", "flow_bwd")] + \
               [("flow_channels", ctypes.c_int)]


class PipelineCfg(ctypes.Structure):
    _fields_ = [("flow_dtype", ctypes.c_int), ("downscale", ctypes.c_int), ("upsample", ctypes.c_int),
                ("scale_vectors", ctypes.c_int), ("fb_max_inconsistent", ctypes.c_float),
                ("ref_bwd_index", ctypes.c_int), ("seed", ctypes.c_ulonglong), ("solver", ctypes.c_int),
                ("overlap", ctypes.c_int), ("flow_batch", ctypes.c_int), ("corr_dtype", ctypes.c_int)]


class Window(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("orig_prev", "orig_cur", "orig_next", "proc_prev", "proc_cur",
                                                "proc_next")]


class ServerCfg(ctypes.Structure):
    _fields_ = [("max_streams", ctypes.c_int), ("hbm_budget", ctypes.c_size_t), ("hip_streams", ctypes.c_int),
                ("form", PipelineCfg), ("use_graphs", ctypes.c_int), ("batch", ctypes.c_int)]


class ServerDone(ctypes.Structure):
    _fields_ = [("stream", ctypes.c_int), ("tag", ctypes.c_longlong), ("frame", ctypes.c_longlong),
                ("out_rgba", ctypes.c_void_p)]


class SynthCfg(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int), ("height", ctypes.c_int), ("nframes", ctypes.c_int),
                ("nblobs", ctypes.c_int), ("sigma_min", ctypes.c_float), ("sigma_max", ctypes.c_float),
                ("motion", ctypes.c_float), ("npatches", ctypes.c_int), ("patch_size", ctypes.c_int),
                ("gain_min", ctypes.c_float), ("gain_max", ctypes.c_float), ("offset_max", ctypes.c_float),
                ("noise_sigma", ctypes.c_float), ("seed", ctypes.c_ulonglong)]


class VscError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"vsc error {code}: {msg}")
        self.code = code


_lib = None
_synth_lib = None


def lib():
    """Load libvsc.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(make -C video-stream-consistency_amd)")
        # torch ships its own HIP runtime under the same soname (libamdhip64.so.7);
        # loading torch first makes the library bind to that one, so torch's
        # streams and allocations and this library's launches share one runtime
        import torch  # noqa: F401
        L = ctypes.CDLL(LIB_PATH)
        vp, i, f = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
        img = VscImage
        L.vsc_last_error.restype = ctypes.c_char_p
        L.vsc_version.restype = ctypes.c_char_p
        L.vsc_default_params.argtypes = [ctypes.POINTER(VscParams)]
        L.vsc_correlation_forward.argtypes = [vp, vp, vp, i, i, i, i, i, i, i, vp, vp]
        L.vsc_quantize_fp8_pair.argtypes = [vp, vp, vp, vp, vp, i, i, i, i, vp]
        L.vsc_quantize_fp8_pair_workspace_size.argtypes = [i]
        L.vsc_quantize_fp8_pair_workspace_size.restype = ctypes.c_size_t
        L.vsc_quantize_fp8_pair_ws.argtypes = [vp, vp, vp, vp, vp, i, i, i, i, vp, ctypes.c_size_t, vp]
        L.vsc_correlation_workspace_size.argtypes = [i, i, i, i, i]
        L.vsc_correlation_workspace_size.restype = ctypes.c_size_t
        L.vsc_correlation_forward_ws.argtypes = [vp, vp, vp, i, i, i, i, i, i, i, vp, vp, ctypes.c_size_t, vp]
        L.vsc_correlation_forward_noalloc.argtypes = [vp, vp, vp, i, i, i, i, i, i, i, vp, vp]
        L.vsc_convert_f32.argtypes = [vp, vp, ctypes.c_size_t, i, vp]
        L.vsc_warp_forward.argtypes = [vp, vp, vp, i, i, i, i, i, vp]
        L.vsc_stab_warp.argtypes = [img, img, img, vp]
        L.vsc_stab_adap_comb.argtypes = [img] * 9 + [f, vp]
        L.vsc_stab_consist_wt.argtypes = [img, img, img, f, f, vp]
        L.vsc_stab_consist_out_workspace_size.argtypes = [i, i]
        L.vsc_stab_consist_out_workspace_size.restype = ctypes.c_size_t
        L.vsc_stab_consist_out.argtypes = [img, img, img, i, f, f, img, vp, vp]
        L.vsc_stab_consist_out_planar_workspace_size.argtypes = [i, i]
        L.vsc_stab_consist_out_planar_workspace_size.restype = ctypes.c_size_t
        L.vsc_stab_consist_out_planar.argtypes = [vp, vp, vp, i, i, i, f, f, vp, vp, vp]
        ip = ctypes.POINTER(ctypes.c_int)
        L.vsc_stab_blocked_plan.argtypes = [i, i, i, i, ip, ip, ip, ip]
        L.vsc_stab_tile_span.argtypes = [i, i, i, i, ip, ip, ip, ip]
        L.vsc_stab_bilinear.argtypes = [img, img, vp]
        L.vsc_average_frames.argtypes = [img, img, img, vp]
        L.vsc_rgba_to_float.argtypes = [vp, img, vp]
        L.vsc_float_to_rgba.argtypes = [img, vp, vp]
        L.vsc_stream_create.argtypes = [i, i, i, i, ctypes.POINTER(vp)]
        L.vsc_stream_destroy.argtypes = [vp]
        L.vsc_stream_footprint.argtypes = [i, i, i, i]
        L.vsc_stream_footprint.restype = ctypes.c_size_t
        L.vsc_stream_set_last_rgba.argtypes = [vp, vp, vp]
        L.vsc_stream_set_last.argtypes = [vp, img, vp]
        L.vsc_stream_copy_last.argtypes = [vp, img, vp]
        L.vsc_stab_frame.argtypes = [vp, ctypes.POINTER(VscFrameIn), ctypes.POINTER(VscParams), vp, vp, i, vp]
        L.vsc_conv3x3_nhwc.argtypes = [vp, i, i, i, i, i, i, vp, vp, vp, i, i, i, i, i, i, i, vp, i, i, i, i, vp]
        L.vsc_flownet_create.argtypes = [i, i, i, i, ctypes.c_ulonglong, i, ctypes.POINTER(vp)]
        L.vsc_flownet_destroy.argtypes = [vp]
        L.vsc_flownet_footprint.argtypes = [i, i, i, i, i]
        L.vsc_flownet_footprint.restype = ctypes.c_size_t
        L.vsc_flownet_flow_size.argtypes = [vp, ctypes.POINTER(i), ctypes.POINTER(i)]
        L.vsc_flownet_param_count.argtypes = [vp]
        L.vsc_flownet_param_count.restype = ctypes.c_size_t
        L.vsc_flownet_get_params.argtypes = [vp, vp, ctypes.c_size_t]
        L.vsc_flownet_set_params.argtypes = [vp, vp, ctypes.c_size_t]
        L.vsc_flownet_extract.argtypes = [vp, vp, i, vp]
        L.vsc_flownet_flow.argtypes = [vp, i, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(vp), i, vp]
        L.vsc_flownet_intermediate.argtypes = [vp, i, i, i, vp, vp]
        L.vsc_flownet_scale_flow_heads.argtypes = [vp, f]
        L.vsc_flownet_set_corr_dtype.argtypes = [vp, i]
        L.vsc_upsampler_create.argtypes = [ctypes.c_ulonglong, ctypes.POINTER(vp)]
        L.vsc_upsampler_destroy.argtypes = [vp]
        L.vsc_upsampler_param_count.argtypes = []
        L.vsc_upsampler_param_count.restype = ctypes.c_size_t
        L.vsc_upsampler_get_params.argtypes = [vp, vp]
        L.vsc_upsampler_set_params.argtypes = [vp, vp]
        L.vsc_upsampler_workspace_size.argtypes = [i, i]
        L.vsc_upsampler_workspace_size.restype = ctypes.c_size_t
        L.vsc_upsampler_run.argtypes = [vp, vp, i, i, i, vp, i, i, f, vp, i, vp, vp, vp]
        L.vsc_upsampler_train_step.argtypes = [vp, i, vp, i, i, vp, vp, i, i, f, f, vp, vp]
        L.vsc_upsampler_loss_grad.argtypes = [vp, i, vp, i, i, vp, vp, i, i, f, vp, vp, vp]
        L.vsc_flow_fb_gate.argtypes = [vp, vp, i, i, i, f, vp, vp, vp]
        L.vsc_pipeline_default_cfg.argtypes = [ctypes.POINTER(PipelineCfg)]
        L.vsc_pipeline_default_cfg.restype = None
        L.vsc_pipeline_create.argtypes = [i, i, ctypes.POINTER(PipelineCfg), ctypes.POINTER(vp)]
        L.vsc_pipeline_destroy.argtypes = [vp]
        L.vsc_pipeline_footprint.argtypes = [i, i, ctypes.POINTER(PipelineCfg)]
        L.vsc_pipeline_footprint.restype = ctypes.c_size_t
        L.vsc_pipeline_set_upsampler_params.argtypes = [vp, vp]
        L.vsc_pipeline_flownet.argtypes = [vp, ctypes.POINTER(vp)]
        L.vsc_pipeline_stream.argtypes = [vp, ctypes.POINTER(vp)]
        L.vsc_pipeline_flows.argtypes = [vp, ctypes.POINTER(Window), ctypes.c_longlong, vp, vp, vp]
        L.vsc_pipeline_frame.argtypes = [vp, ctypes.POINTER(Window), ctypes.c_longlong, ctypes.POINTER(VscParams),
                                         vp, vp, vp]
        L.vsc_pipeline_frames.argtypes = [vp, ctypes.POINTER(Window), i, ctypes.c_longlong, ctypes.POINTER(VscParams),
                                          ctypes.POINTER(vp), ctypes.POINTER(vp), vp]
        L.vsc_pipeline_gate.argtypes = [vp, vp, vp, vp]
        L.vsc_pipeline_sync.argtypes = [vp, vp]
        L.vsc_server_default_cfg.argtypes = [ctypes.POINTER(ServerCfg)]
        L.vsc_server_default_cfg.restype = None
        L.vsc_server_create.argtypes = [ctypes.POINTER(ServerCfg), vp, ctypes.POINTER(vp)]
        L.vsc_server_destroy.argtypes = [vp]
        L.vsc_server_stream_footprint.argtypes = [vp, i, i]
        L.vsc_server_stream_footprint.restype = ctypes.c_size_t
        L.vsc_server_used_bytes.argtypes = [vp]
        L.vsc_server_used_bytes.restype = ctypes.c_size_t
        L.vsc_server_open.argtypes = [vp, i, i, ctypes.POINTER(i)]
        L.vsc_server_close.argtypes = [vp, i]
        L.vsc_server_set_params.argtypes = [vp, i, ctypes.POINTER(VscParams)]
        L.vsc_server_submit.argtypes = [vp, i, vp, vp, ctypes.c_longlong]
        L.vsc_server_tick.argtypes = [vp, ctypes.POINTER(i)]
        L.vsc_server_poll.argtypes = [vp, ctypes.POINTER(ServerDone), i, ctypes.POINTER(i)]
        L.vsc_server_drain.argtypes = [vp]
        L.vsc_server_read.argtypes = [vp, i, ctypes.c_longlong, vp, vp]
        L.vsc_flo_read.argtypes = [ctypes.c_char_p, vp, ctypes.c_size_t, ctypes.POINTER(i), ctypes.POINTER(i)]
        L.vsc_flo_write.argtypes = [ctypes.c_char_p, vp, i, i]
        L.vsc_flow_dir_path.argtypes = [ctypes.c_char_p, ctypes.c_longlong, i, ctypes.c_char_p, ctypes.c_size_t]
        L.vsc_flow_dir_read.argtypes = [ctypes.c_char_p, ctypes.c_longlong, vp, vp, ctypes.c_size_t,
                                        ctypes.POINTER(i), ctypes.POINTER(i)]
        L.vsc_stabilize_flow_dir.argtypes = [vp, vp, i, i, i, ctypes.c_char_p, ctypes.POINTER(VscParams), i,
                                             FRAME_SINK, vp, vp]
        L.vsc_profile_begin.argtypes = [i]
        L.vsc_profile_end.argtypes = [ctypes.c_char_p, ctypes.POINTER(i), ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_double)]
        _lib = L
    return _lib


def check(rc):
    if rc != VSC_OK:
        raise VscError(rc, lib().vsc_last_error().decode())
    return rc


def version():
    return lib().vsc_version().decode()


def _torch():
    import torch
    return torch


def _stream():
    return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


# ---------------------------------------------------------------------------
# flow network building blocks (channels-last tensors)
# ---------------------------------------------------------------------------
def conv3x3_nhwc(x, wt, bias=None, stride=1, dilation=1, pad=None, act=0, out=None, out_dtype=None,
                 add=None, in_slice=None, out_slice=None, add_slice=None):
    """3x3 convolution of an NHWC device tensor on MFMA (vsc_conv3x3_nhwc).

    x: [N, H, W, Cs] float32 / bfloat16 / float16; in_slice = (c0, Cin) selects
    channels (default all).  wt: [Cout, 3, 3, Cin] of x's dtype; bias: float32
    [Cout].  act: 0 none, 1 LeakyReLU(0.1), 2 ReLU.  out: [N, Ho, Wo, Cs_out]
    (allocated when None) with out_slice = (c0, Cout); add: residual of the
    output's dtype, add_slice = (c0, Cout)."""
    torch = _torch()
    N, H, W, cs = x.shape
    c0, cin = in_slice if in_slice is not None else (0, cs)
    cout = wt.shape[0]
    if pad is None:
        pad = dilation
    Ho = (H + 2 * pad - 2 * dilation - 1) // stride + 1
    Wo = (W + 2 * pad - 2 * dilation - 1) // stride + 1
    odt = out_dtype if out_dtype is not None else x.dtype
    if out is None:
        out = torch.zeros((N, Ho, Wo, cout), dtype=odt, device=x.device)
    oc0, ocout = out_slice if out_slice is not None else (0, cout)
    if ocout != cout:
        raise VscError(-1, "out_slice width must equal Cout")
    ac0 = add_slice[0] if add_slice is not None else 0
    check(lib().vsc_conv3x3_nhwc(_ptr(x), cs, c0, N, H, W, cin, _ptr(wt), _ptr(bias) if bias is not None else None,
                                 _ptr(out), out.shape[3], oc0, cout, stride, dilation, pad, act,
                                 _ptr(add) if add is not None else None, add.shape[3] if add is not None else 0,
                                 ac0, _dtype_code(x), _dtype_code(out), _stream()))
    return out


class FlowNet:
    """PWCNet-light flow network on the library's kernels (vsc_flownet_*).

    Frames are device RGBA8 tensors [H, W, 4]; extract() caches a frame's
    features in a slot, flow() runs the decoders for one or two (first,
    second) slot pairs and returns HWC float flows at flow_size.  dtype
    "f32+bf16corr": an FP32 network whose cost volume multiplies on the BF16
    MFMA."""

    def __init__(self, width, height, downscale=1, dtype="f32", seed=0, slots=4):
        self.dtype = {"f32": VSC_F32, "bf16": VSC_BF16, "f32+bf16corr": VSC_F32}[dtype]
        h = ctypes.c_void_p()
        check(lib().vsc_flownet_create(width, height, downscale, self.dtype, seed, slots, ctypes.byref(h)))
        self.h = h
        if dtype == "f32+bf16corr":
            check(lib().vsc_flownet_set_corr_dtype(self.h, VSC_BF16))
        w_, h_ = ctypes.c_int(), ctypes.c_int()
        check(lib().vsc_flownet_flow_size(self.h, ctypes.byref(w_), ctypes.byref(h_)))
        self.flow_size = (w_.value, h_.value)
        self.width, self.height, self.downscale = width, height, downscale
        self.padded = ((self.flow_size[0] + 63) // 64 * 64, (self.flow_size[1] + 63) // 64 * 64)

    @staticmethod
    def footprint(width, height, downscale=1, dtype="f32", slots=4):
        return lib().vsc_flownet_footprint(width, height, downscale, {"f32": VSC_F32, "bf16": VSC_BF16}[dtype], slots)

    def params(self):
        import numpy as np
        n = lib().vsc_flownet_param_count(self.h)
        a = np.empty(n, dtype=np.float32)
        check(lib().vsc_flownet_get_params(self.h, a.ctypes.data, n))
        return a

    def set_params(self, a):
        import numpy as np
        a = np.ascontiguousarray(a, dtype=np.float32)
        check(lib().vsc_flownet_set_params(self.h, a.ctypes.data, a.size))

    def scale_flow_heads(self, factor):
        check(lib().vsc_flownet_scale_flow_heads(self.h, factor))

    def extract(self, rgba, slot):
        check(lib().vsc_flownet_extract(self.h, _ptr(rgba), slot, _stream()))

    def flow(self, pairs, out_channels=2, out=None):
        torch = _torch()
        n = len(pairs)
        if out is None:
            out = [torch.empty((self.flow_size[1], self.flow_size[0], out_channels), dtype=torch.float32,
                               device="cuda") for _ in range(n)]
        a = (ctypes.c_int * n)(*[p[0] for p in pairs])
        b = (ctypes.c_int * n)(*[p[1] for p in pairs])
        o = (ctypes.c_void_p * n)(*[t.data_ptr() for t in out])
        check(lib().vsc_flownet_flow(self.h, n, a, b, o, out_channels, _stream()))
        return out

    def intermediate(self, which, level, item=0):
        torch = _torch()
        h, w = self.padded[1] >> level, self.padded[0] >> level
        t = torch.empty((h, w, 2), dtype=torch.float32, device="cuda")
        check(lib().vsc_flownet_intermediate(self.h, which, level, item, _ptr(t), _stream()))
        return t

    def close(self):
        if self.h:
            lib().vsc_flownet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Upsampler:
    """Learned flow upsampler (vsc_upsampler_*): FP16 inference, FP32 training
    with the library's forward/backward/Adam kernels."""

    def __init__(self, seed=0, params=None):
        h = ctypes.c_void_p()
        check(lib().vsc_upsampler_create(seed, ctypes.byref(h)))
        self.h = h
        self._ws = None
        if params is not None:
            self.set_params(params)

    @staticmethod
    def param_count():
        return lib().vsc_upsampler_param_count()

    def params(self):
        import numpy as np
        a = np.empty(self.param_count(), dtype=np.float32)
        check(lib().vsc_upsampler_get_params(self.h, a.ctypes.data))
        return a

    def set_params(self, a):
        import numpy as np
        a = np.ascontiguousarray(a, dtype=np.float32)
        if a.size != self.param_count():
            raise VscError(-1, "wrong parameter count")
        check(lib().vsc_upsampler_set_params(self.h, a.ctypes.data))

    def workspace(self, width, height):
        torch = _torch()
        n = lib().vsc_upsampler_workspace_size(width, height)
        if self._ws is None or self._ws.numel() < n:
            self._ws = torch.empty(n, dtype=torch.uint8, device="cuda")
        return self._ws

    def run(self, flow_lo, rgba, scale=2.0, out_ch=2, gate=None, out=None):
        """flow_lo [lh, lw, 2|3] float32, rgba [H, W, 4] uint8 (device) -> flow [H, W, out_ch]."""
        torch = _torch()
        H, W = rgba.shape[:2]
        lh, lw, lch = flow_lo.shape
        if out is None:
            out = torch.empty((H, W, out_ch), dtype=torch.float32, device=rgba.device)
        ws = self.workspace(W, H)
        check(lib().vsc_upsampler_run(self.h, _ptr(flow_lo), lw, lh, lch, _ptr(rgba), W, H, scale, _ptr(out), out_ch,
                                      _ptr(gate) if gate is not None else None, _ptr(ws), _stream()))
        return out

    def train_step(self, flow_lo, rgba, flow_gt, lr=1e-3, scale=2.0, want_loss=False):
        """One Adam step on a batch: flow_lo [B, lh, lw, 2], rgba [B, H, W, 4], flow_gt [B, H, W, 2]."""
        B, H, W = rgba.shape[:3]
        loss = ctypes.c_float()
        check(lib().vsc_upsampler_train_step(self.h, B, _ptr(flow_lo), flow_lo.shape[2], flow_lo.shape[1], _ptr(rgba),
                                             _ptr(flow_gt), W, H, scale, lr,
                                             ctypes.byref(loss) if want_loss else None, _stream()))
        return loss.value if want_loss else None

    def loss_grad(self, flow_lo, rgba, flow_gt, scale=2.0):
        import numpy as np
        B, H, W = rgba.shape[:3]
        loss = ctypes.c_float()
        g = np.empty(self.param_count(), dtype=np.float32)
        check(lib().vsc_upsampler_loss_grad(self.h, B, _ptr(flow_lo), flow_lo.shape[2], flow_lo.shape[1], _ptr(rgba),
                                            _ptr(flow_gt), W, H, scale, ctypes.byref(loss), g.ctypes.data, _stream()))
        return loss.value, g

    def close(self):
        if self.h:
            lib().vsc_upsampler_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def flow_fb_gate(fwd, bwd, max_inconsistent=0.25):
    """Forward-backward consistency count and fallback gate (device scalars)."""
    torch = _torch()
    h, w, ch = fwd.shape
    count = torch.empty(1, dtype=torch.int32, device=fwd.device)
    gate = torch.empty(1, dtype=torch.float32, device=fwd.device)
    check(lib().vsc_flow_fb_gate(_ptr(fwd), _ptr(bwd), w, h, ch, max_inconsistent, _ptr(count), _ptr(gate),
                                 _stream()))
    return count, gate


# ---------------------------------------------------------------------------
# ORT custom ops
# ---------------------------------------------------------------------------
_DT = None


def _dtype_code(t):
    torch = _torch()
    global _DT
    if _DT is None:
        _DT = {torch.float32: VSC_F32, torch.float64: VSC_F64, torch.bfloat16: VSC_BF16,
               torch.float16: VSC_F16}
        if hasattr(torch, "float8_e4m3fn"):
            _DT[torch.float8_e4m3fn] = VSC_FP8E4M3
    if t.dtype not in _DT:
        raise VscError(-2, f"unsupported input type {t.dtype}")
    return _DT[t.dtype]


def correlation_workspace_bytes(shape, dtype_code):
    """Device bytes vsc_correlation_forward_ws needs for NCHW `shape` (an upper
    bound: operands read in their NCHW layout need none)."""
    N, C, H, W = shape
    return int(lib().vsc_correlation_workspace_size(N, C, H, W, int(dtype_code)))


def correlation(input1, input2, max_displacement: int = 4, legacy: int = 0, scale=None, out=None, workspace=None):
    """``custom::Correlation`` (correlation.h:60-84; correlation.cc:45-143).

    input1, input2: NCHW device tensors of float32 / float64 / bfloat16 /
    float16 / float8_e4m3fn.  Returns [N, 2md+1, 2md+1, H, W] (legacy=0) or
    [N, (2md+1)^2, H, W] (legacy=1), float32 (float64 for float64 inputs).
    workspace: a device uint8 tensor of correlation_workspace_bytes() bytes
    (vsc_correlation_forward_ws: no allocation in the call)."""
    torch = _torch()
    if input1.shape != input2.shape or input1.dim() != 4:
        raise VscError(-1, "Correlation expects two NCHW tensors of the same shape")
    a = input1.contiguous()
    b = input2.contiguous()
    N, C, H, W = a.shape
    P = 2 * int(max_displacement) + 1
    odt = torch.float64 if a.dtype == torch.float64 else torch.float32
    shape = (N, P * P, H, W) if legacy else (N, P, P, H, W)
    if out is None:
        out = torch.empty(shape, dtype=odt, device=a.device)
    cs = None if scale is None else _ptr(scale.contiguous())
    if workspace is not None:
        check(lib().vsc_correlation_forward_ws(_ptr(a), _ptr(b), _ptr(out), N, C, H, W, int(max_displacement),
                                               int(bool(legacy)), _dtype_code(a), cs, _ptr(workspace),
                                               workspace.numel() * workspace.element_size(), _stream()))
        return out
    check(lib().vsc_correlation_forward(_ptr(a), _ptr(b), _ptr(out), N, C, H, W, int(max_displacement),
                                        int(bool(legacy)), _dtype_code(a), cs, _stream()))
    return out


def warp(input, flow, out=None):
    """``custom::Warp`` (warp.h:38-60; warp.cc:25-134): NCHW input, N2HW flow."""
    torch = _torch()
    a = input.contiguous()
    fl = flow.contiguous()
    N, C, H, W = a.shape
    if tuple(fl.shape) != (N, 2, H, W):
        raise VscError(-1, "Warp expects flow of shape [N, 2, H, W]")
    if out is None:
        out = torch.empty_like(a)
    check(lib().vsc_warp_forward(_ptr(a), _ptr(fl), _ptr(out), N, C, H, W, _dtype_code(a), _stream()))
    return out


def quantize_fp8_pair(f1, f2, workspace=None):
    """Equalised E4M3 operands for the FP8 correlation: returns q1, q2 (uint8
    tensors of E4M3 bits) and scale (1 float): f1.f2 ~= scale * q1.q2.
    workspace: optional device uint8 tensor (vsc_quantize_fp8_pair_ws)."""
    torch = _torch()
    a = f1.contiguous().float()
    b = f2.contiguous().float()
    N, C, H, W = a.shape
    q1 = torch.empty(a.shape, dtype=torch.uint8, device=a.device)
    q2 = torch.empty(a.shape, dtype=torch.uint8, device=a.device)
    sc = torch.empty(1, dtype=torch.float32, device=a.device)
    if workspace is not None:
        check(lib().vsc_quantize_fp8_pair_ws(_ptr(a), _ptr(b), _ptr(q1), _ptr(q2), _ptr(sc), N, C, H, W,
                                             _ptr(workspace), workspace.numel() * workspace.element_size(),
                                             _stream()))
        return q1, q2, sc
    check(lib().vsc_quantize_fp8_pair(_ptr(a), _ptr(b), _ptr(q1), _ptr(q2), _ptr(sc), N, C, H, W, _stream()))
    return q1, q2, sc


def correlation_fp8(q1, q2, scale, max_displacement=4, legacy=0, workspace=None, out=None):
    """Correlation of e4m3 operands given as uint8 bit patterns (out: an
    optional float32 device tensor of the output shape)."""
    torch = _torch()
    N, C, H, W = q1.shape
    P = 2 * int(max_displacement) + 1
    shape = (N, P * P, H, W) if legacy else (N, P, P, H, W)
    if out is None:
        out = torch.empty(shape, dtype=torch.float32, device=q1.device)
    if workspace is not None:
        check(lib().vsc_correlation_forward_ws(_ptr(q1.contiguous()), _ptr(q2.contiguous()), _ptr(out), N, C, H, W,
                                               int(max_displacement), int(bool(legacy)), VSC_FP8E4M3,
                                               _ptr(scale.contiguous()), _ptr(workspace),
                                               workspace.numel() * workspace.element_size(), _stream()))
        return out
    check(lib().vsc_correlation_forward(_ptr(q1.contiguous()), _ptr(q2.contiguous()), _ptr(out), N, C, H, W,
                                        int(max_displacement), int(bool(legacy)), VSC_FP8E4M3,
                                        _ptr(scale.contiguous()), _stream()))
    return out


def convert(x, dtype):
    """FP32 device tensor -> "bf16" / "f16" (torch tensors of that type) or
    "e4m3" (uint8 bit patterns), round to nearest even."""
    torch = _torch()
    x = x.contiguous().float()
    code = {"bf16": VSC_BF16, "f16": VSC_F16, "e4m3": VSC_FP8E4M3}[dtype]
    tdt = {"bf16": torch.bfloat16, "f16": torch.float16, "e4m3": torch.uint8}[dtype]
    out = torch.empty(x.shape, dtype=tdt, device=x.device)
    check(lib().vsc_convert_f32(_ptr(x), _ptr(out), x.numel(), code, _stream()))
    return out


# ---------------------------------------------------------------------------
# GPUImage and the stabilization stages
# ---------------------------------------------------------------------------
class GPUImage:
    """Owning HWC float32 device image (gpuimage.h:8-31).  ``data`` is a torch
    tensor of shape (height, width, channels) on the current device."""

    def __init__(self, width: int, height: int, channels: int, data=None):
        torch = _torch()
        self.width, self.height, self.channels = int(width), int(height), int(channels)
        if data is None:
            data = torch.empty((self.height, self.width, self.channels), dtype=torch.float32, device="cuda")
        if tuple(data.shape) != (self.height, self.width, self.channels) or data.dtype != torch.float32:
            raise VscError(-1, "GPUImage data must be float32 (height, width, channels)")
        self.data = data.contiguous()

    @classmethod
    def from_tensor(cls, t):
        h, w, c = t.shape
        return cls(w, h, c, t)

    @classmethod
    def from_rgba(cls, rgba):
        """copyFromQImage (gpuimage.cpp:103-124): RGBA8 (H, W, 4) uint8 device tensor."""
        h, w, _ = rgba.shape
        im = cls(w, h, 3)
        check(lib().vsc_rgba_to_float(_ptr(rgba.contiguous()), im.c(), _stream()))
        return im

    def to_rgba(self):
        """copyToQImage (gpuimage.cpp:127-135): RGBA8 (H, W, 4) uint8 device tensor."""
        torch = _torch()
        out = torch.empty((self.height, self.width, 4), dtype=torch.uint8, device=self.data.device)
        check(lib().vsc_float_to_rgba(self.c(), _ptr(out), _stream()))
        return out

    def copy_from(self, other: "GPUImage"):
        if (other.width, other.height, other.channels) != (self.width, self.height, self.channels):
            raise VscError(-1, "Invalid dimensions")
        self.data.copy_(other.data)

    def c(self):
        return VscImage(self.data.data_ptr(), self.width, self.height, self.channels)


def get_warp_result(input: GPUImage, flow: GPUImage, input_warp: GPUImage):
    check(lib().vsc_stab_warp(input.c(), flow.c(), input_warp.c(), _stream()))


def get_adap_comb(crnt_in, crnt_pr, prev_warp_in, prev_warp_pr, next_warp_in, next_warp_pr, adap_cmb_in,
                  adap_cmb_pr, last_stab_warp, alpha):
    check(lib().vsc_stab_adap_comb(crnt_in.c(), crnt_pr.c(), prev_warp_in.c(), prev_warp_pr.c(),
                                   next_warp_in.c(), next_warp_pr.c(), adap_cmb_in.c(), adap_cmb_pr.c(),
                                   last_stab_warp.c(), float(alpha), _stream()))


def get_consist_wt(adap_cmb_in, crnt_in, consis_wt, beta, gamma):
    check(lib().vsc_stab_consist_wt(adap_cmb_in.c(), crnt_in.c(), consis_wt.c(), float(beta), float(gamma),
                                    _stream()))


def get_consist_out(crnt_pr, prev_stab_warp, cons_wt, num_iter, step_size, mom_fac, consis_out, workspace=None):
    ws = None if workspace is None else _ptr(workspace)
    check(lib().vsc_stab_consist_out(crnt_pr.c(), prev_stab_warp.c(), cons_wt.c(), int(num_iter),
                                     float(step_size), float(mom_fac), consis_out.c(), ws, _stream()))


def get_consist_out_planar(crnt_pr, prev_stab_warp, cons_wt, num_iter, step_size, mom_fac, consis_out,
                           workspace=None):
    """get_consist_out on planar (3, H, W) float32 device tensors (the stream
    driver's layout; temporally blocked solver)."""
    _, h, w = consis_out.shape
    ws = None if workspace is None else _ptr(workspace)
    check(lib().vsc_stab_consist_out_planar(_ptr(crnt_pr), _ptr(prev_stab_warp), _ptr(cons_wt), w, h,
                                            int(num_iter), float(step_size), float(mom_fac), _ptr(consis_out),
                                            ws, _stream()))


def blocked_plan(width, height, num_iter, fma=True):
    """Launch plan of the temporally blocked solver for one pyramid level:
    (sweeps per launch, tile rows, tiles across, tiles down)."""
    out = [ctypes.c_int() for _ in range(4)]
    check(lib().vsc_stab_blocked_plan(int(width), int(height), int(num_iter), int(bool(fma)),
                                      *[ctypes.byref(v) for v in out]))
    return tuple(v.value for v in out)


def tile_span(length, tile, halo, i):
    """(tile count, first computed pixel, stored range lo, hi) of tile i along
    an axis of the blocked solver."""
    out = [ctypes.c_int() for _ in range(4)]
    check(lib().vsc_stab_tile_span(int(length), int(tile), int(halo), int(i), *[ctypes.byref(v) for v in out]))
    return tuple(v.value for v in out)


def consist_out_workspace_bytes(width, height):
    return lib().vsc_stab_consist_out_workspace_size(int(width), int(height))


def get_bilinear(input: GPUImage, output: GPUImage):
    check(lib().vsc_stab_bilinear(input.c(), output.c(), _stream()))


def perform_consistency(original_curr, processed_prev, processed_curr, processed_next, flow_fwd, flow_bwd,
                        stabilized_prev, stabilized_out):
    """perform_consistency (flowconsistency.cu:269-286): only the average of the
    previous and next processed frames is computed, as in the reference."""
    check(lib().vsc_average_frames(processed_prev.c(), processed_next.c(), stabilized_out.c(), _stream()))


# ---------------------------------------------------------------------------
# VideoStabilizer
# ---------------------------------------------------------------------------
@dataclass
class HyperParams:
    """struct hyperParams (videostabilizer.h:38-46) with initHyperParams
    defaults (videostabilizer.cpp:104-113)."""
    alpha: float = 6800.0
    beta: float = 6800.0
    gamma: float = 2.0
    pyramidLevels: int = 2
    numIter: int = 150
    stepSize: float = 0.15
    momFac: float = 0.15

    def c(self):
        return VscParams(self.alpha, self.beta, self.gamma, self.pyramidLevels, self.numIter, self.stepSize,
                         self.momFac)


class VideoStabilizer:
    """Per-stream stabilizer (videostabilizer.h:48-110) over the device-resident
    state of ``vsc_stream``.  Frames are RGBA8 device tensors (H, W, 4); flows
    are HWC float32 device tensors (fh, fw, 2|3), fwd = cur->next, bwd = cur->prev.
    ``flow_width``/``flow_height`` below the frame size select the
    FLOWDOWNSCALE path (flows upsampled with get_bilinear)."""

    k = 1

    def __init__(self, width, height, flow_width=None, flow_height=None, mode=VSC_FRAME_FUSED):
        self.width, self.height = int(width), int(height)
        self.flow_width = int(flow_width or width)
        self.flow_height = int(flow_height or height)
        self.mode = mode
        self.controlParameters = HyperParams()
        h = ctypes.c_void_p()
        check(lib().vsc_stream_create(self.width, self.height, self.flow_width, self.flow_height, ctypes.byref(h)))
        self._h = h

    def getHyperParams(self) -> HyperParams:
        return self.controlParameters

    @staticmethod
    def footprint(width, height, flow_width=None, flow_height=None):
        return int(lib().vsc_stream_footprint(width, height, flow_width or width, flow_height or height))

    def set_last_rgba(self, processed_rgba):
        """preloadProcessedFrames: last <- processed frame 2k+B-1."""
        check(lib().vsc_stream_set_last_rgba(self._h, _ptr(processed_rgba), _stream()))

    def set_last(self, img: GPUImage):
        check(lib().vsc_stream_set_last(self._h, img.c(), _stream()))

    def last(self) -> GPUImage:
        im = GPUImage(self.width, self.height, 3)
        check(lib().vsc_stream_copy_last(self._h, im.c(), _stream()))
        return im

    def do_one_step(self, orig, proc, flow_fwd, flow_bwd, out_rgba=None, out_float=None, params=None):
        """The stabilization body of doOneStep (videostabilizer.cpp:177-238).
        orig, proc: sequences (prev, cur, next) of RGBA8 device tensors."""
        torch = _torch()
        fc = int(flow_fwd.shape[2])
        fin = VscFrameIn(*(t.data_ptr() for t in (orig[0], orig[1], orig[2], proc[0], proc[1], proc[2],
                                                     flow_fwd, flow_bwd)), fc)
        if out_rgba is None:
            out_rgba = torch.empty((self.height, self.width, 4), dtype=torch.uint8, device=orig[1].device)
        prm = (params or self.controlParameters).c()
        check(lib().vsc_stab_frame(self._h, ctypes.byref(fin), ctypes.byref(prm), _ptr(out_rgba),
                                   None if out_float is None else _ptr(out_float), int(self.mode), _stream()))
        return out_rgba

    def close(self):
        if getattr(self, "_h", None):
            lib().vsc_stream_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def load_upsampler_weights():
    """The upsampler weights trained by upsampler_train.py (float32, torch layout)."""
    path = os.path.join(_HERE, "upsampler_weights.npy")
    return np.load(path, allow_pickle=False)


class Pipeline:
    """Per-stream frame pipeline (vsc_pipeline_*): flow network, FLOWDOWNSCALE,
    bilinear or learned upsampling with the consistency fallback, stabilization.

    form "plain": FP32 network at full resolution; form "chosen": BF16 network
    at FLOWDOWNSCALE 2 with the learned upsampler.  Frames are RGBA8 device
    tensors; a window is (orig_prev, orig_cur, orig_next, proc_prev, proc_cur,
    proc_next) and `frame` is the index of orig_cur (features are cached by
    index).

    overlap=True pipelines frames on two internal HIP streams (the flows of a
    frame while the previous one is stabilized): frame t's outputs are then
    ordered on torch's current stream by the call for frame t+1, or by sync(),
    and the window tensors of a call must stay unchanged until then.

    flow_batch=B (1..4): frames() stabilizes up to B consecutive windows with
    their flows from one batched network decode (the reference's batchSize)."""

    # both forms stabilize with the FMA-form solver (FP32, <= 1 LSB from the
    # reference op order); solver="exact" selects the reference order
    FORMS = {
        "plain": dict(flow_dtype="f32", downscale=1, upsample="bilinear", solver="fma"),
        "chosen": dict(flow_dtype="bf16", downscale=2, upsample="learned", solver="fma"),
    }

    def __init__(self, width, height, form=None, flow_dtype="f32", downscale=1, upsample="bilinear",
                 scale_vectors=False, fb_max_inconsistent=0.25, ref_bwd_index=False, seed=0, upsampler_params=None,
                 flow_head_scale=None, solver=None, overlap=False, flow_batch=1, corr_dtype=None):
        if form is not None:
            kw = self.FORMS[form]
            flow_dtype, downscale, upsample = kw["flow_dtype"], kw["downscale"], kw["upsample"]
            solver = solver or kw["solver"]
        c = PipelineCfg()
        lib().vsc_pipeline_default_cfg(ctypes.byref(c))
        c.flow_dtype = {"f32": VSC_F32, "bf16": VSC_BF16}[flow_dtype]
        c.downscale = downscale
        c.upsample = {"bilinear": 0, "learned": 1}[upsample]
        c.scale_vectors = int(bool(scale_vectors))
        c.fb_max_inconsistent = fb_max_inconsistent
        c.ref_bwd_index = int(bool(ref_bwd_index))
        c.seed = seed
        c.solver = SOLVERS[solver or "exact"]
        c.overlap = int(bool(overlap))
        c.flow_batch = int(flow_batch)
        c.corr_dtype = {None: VSC_F32, "f32": VSC_F32, "bf16": VSC_BF16}[corr_dtype]
        self.cfg = c
        self.width, self.height = width, height
        h = ctypes.c_void_p()
        check(lib().vsc_pipeline_create(width, height, ctypes.byref(c), ctypes.byref(h)))
        self.h = h
        self.flow_size = (int(round(width / downscale)), int(round(height / downscale)))
        self.controlParameters = HyperParams()
        if c.upsample == 1:
            p = upsampler_params if upsampler_params is not None else load_upsampler_weights()
            p = np.ascontiguousarray(p, dtype=np.float32)
            check(lib().vsc_pipeline_set_upsampler_params(self.h, p.ctypes.data))
        if flow_head_scale is not None:
            net = ctypes.c_void_p()
            check(lib().vsc_pipeline_flownet(self.h, ctypes.byref(net)))
            check(lib().vsc_flownet_scale_flow_heads(net, flow_head_scale))

    @staticmethod
    def footprint(width, height, form="plain"):
        """Device bytes of a pipeline of this form (admission by HBM budget)."""
        fk = Pipeline.FORMS[form]
        c = PipelineCfg()
        lib().vsc_pipeline_default_cfg(ctypes.byref(c))
        c.flow_dtype = {"f32": VSC_F32, "bf16": VSC_BF16}[fk["flow_dtype"]]
        c.downscale = fk["downscale"]
        c.upsample = {"bilinear": 0, "learned": 1}[fk["upsample"]]
        c.solver = SOLVERS[fk["solver"]]
        return int(lib().vsc_pipeline_footprint(width, height, ctypes.byref(c)))

    def _stream_handle(self):
        s = ctypes.c_void_p()
        check(lib().vsc_pipeline_stream(self.h, ctypes.byref(s)))
        return s

    def set_last_rgba(self, processed_rgba):
        check(lib().vsc_stream_set_last_rgba(self._stream_handle(), _ptr(processed_rgba), _stream()))

    @staticmethod
    def _window(w):
        return Window(*(t.data_ptr() for t in w))

    def flows(self, window, frame):
        torch = _torch()
        fw, fh = self.flow_size
        f = torch.empty((fh, fw, 2), dtype=torch.float32, device="cuda")
        b = torch.empty((fh, fw, 2), dtype=torch.float32, device="cuda")
        check(lib().vsc_pipeline_flows(self.h, ctypes.byref(self._window(window)), frame, _ptr(f), _ptr(b), _stream()))
        return f, b

    def frame(self, window, frame, out_rgba=None, out_float=None, params=None):
        """Stabilize window `frame` into out_rgba (and out_float, HWC float);
        stream-ordered on torch's current stream (one frame later with overlap)."""
        torch = _torch()
        if out_rgba is None:
            out_rgba = torch.empty((self.height, self.width, 4), dtype=torch.uint8, device="cuda")
        prm = (params or self.controlParameters).c()
        check(lib().vsc_pipeline_frame(self.h, ctypes.byref(self._window(window)), frame, ctypes.byref(prm),
                                       _ptr(out_rgba), None if out_float is None else _ptr(out_float), _stream()))
        return out_rgba

    def frames(self, windows, first, out_rgba=None, out_float=None, params=None):
        """Stabilize len(windows) <= flow_batch consecutive windows (frames first,
        first+1, ...) with one batched flow decode; out_rgba / out_float: lists
        of tensors (or None).  Returns the RGBA8 outputs."""
        torch = _torch()
        n = len(windows)
        if out_rgba is None:
            out_rgba = [torch.empty((self.height, self.width, 4), dtype=torch.uint8, device="cuda") for _ in range(n)]
        ws = (Window * n)(*(self._window(w) for w in windows))
        ro = (ctypes.c_void_p * n)(*(None if t is None else t.data_ptr() for t in out_rgba))
        fo = None if out_float is None else (ctypes.c_void_p * n)(*(None if t is None else t.data_ptr()
                                                                    for t in out_float))
        prm = (params or self.controlParameters).c()
        check(lib().vsc_pipeline_frames(self.h, ws, n, first, ctypes.byref(prm), ro, fo, _stream()))
        return out_rgba

    def sync(self):
        """Order torch's current stream after every submitted frame (overlap mode:
        frame t's outputs are otherwise ordered only by the call for frame t+1)."""
        check(lib().vsc_pipeline_sync(self.h, _stream()))

    def gate(self):
        torch = _torch()
        g = torch.empty(1, dtype=torch.float32, device="cuda")
        c = torch.empty(1, dtype=torch.int32, device="cuda")
        check(lib().vsc_pipeline_gate(self.h, _ptr(g), _ptr(c), _stream()))
        return float(g.cpu()[0]), int(c.cpu()[0])

    def close(self):
        if getattr(self, "h", None):
            lib().vsc_pipeline_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Multi-stream frame server (vsc_server_*): admission by device-memory
    budget, continuous batching of ready windows: each stream's flow work on a
    pool of HIP streams, then (batch=True) the stabilization of every stream's
    window of a tick in one graph of persistent work-list launches."""

    def __init__(self, hbm_budget=64 << 30, max_streams=64, hip_streams=4, form="chosen", upsampler_params=None,
                 use_graphs=True, solver=None, batch=True):
        c = ServerCfg()
        lib().vsc_server_default_cfg(ctypes.byref(c))
        c.max_streams, c.hbm_budget, c.hip_streams = max_streams, hbm_budget, hip_streams
        c.use_graphs = int(bool(use_graphs))
        c.batch = int(bool(batch))
        fk = Pipeline.FORMS[form]
        c.form.flow_dtype = {"f32": VSC_F32, "bf16": VSC_BF16}[fk["flow_dtype"]]
        c.form.downscale = fk["downscale"]
        c.form.upsample = {"bilinear": 0, "learned": 1}[fk["upsample"]]
        c.form.solver = SOLVERS[solver or fk["solver"]]
        self._params = None
        if c.form.upsample == 1:
            self._params = np.ascontiguousarray(
                upsampler_params if upsampler_params is not None else load_upsampler_weights(), dtype=np.float32)
        h = ctypes.c_void_p()
        check(lib().vsc_server_create(ctypes.byref(c), None if self._params is None else self._params.ctypes.data,
                                      ctypes.byref(h)))
        self.h = h
        self.max_streams = max_streams
        self.sizes = {}

    def footprint(self, width, height):
        return int(lib().vsc_server_stream_footprint(self.h, width, height))

    def used_bytes(self):
        return int(lib().vsc_server_used_bytes(self.h))

    def open(self, width, height):
        i = ctypes.c_int()
        check(lib().vsc_server_open(self.h, width, height, ctypes.byref(i)))
        self.sizes[i.value] = (width, height)
        return i.value

    def close_stream(self, sid):
        check(lib().vsc_server_close(self.h, sid))
        self.sizes.pop(sid, None)

    def set_params(self, sid, params: "HyperParams"):
        prm = params.c()
        check(lib().vsc_server_set_params(self.h, sid, ctypes.byref(prm)))

    def submit(self, sid, orig, proc, tag=0):
        check(lib().vsc_server_submit(self.h, sid, _ptr(orig), _ptr(proc), int(tag)))

    def tick(self):
        n = ctypes.c_int()
        check(lib().vsc_server_tick(self.h, ctypes.byref(n)))
        return n.value

    def poll(self, max_items=1024):
        arr = (ServerDone * max_items)()
        n = ctypes.c_int()
        check(lib().vsc_server_poll(self.h, arr, max_items, ctypes.byref(n)))
        return [(arr[k].stream, arr[k].tag, arr[k].frame, arr[k].out_rgba) for k in range(n.value)]

    def read(self, sid, frame, out=None):
        """Stabilized frame `frame` of stream sid (stream-ordered on torch's current
        stream) into `out` (a contiguous uint8 [H, W, 4] device or pinned host
        tensor) or a new device tensor."""
        torch = _torch()
        w, h = self.sizes[sid]
        t = torch.empty((h, w, 4), dtype=torch.uint8, device="cuda") if out is None else out
        if t.shape != (h, w, 4) or t.dtype != torch.uint8 or not t.is_contiguous():
            raise VscError(-1, "out must be a contiguous uint8 [H, W, 4] tensor")
        if not t.is_cuda and not t.is_pinned():
            raise VscError(-1, "a host destination must be pinned memory")
        check(lib().vsc_server_read(self.h, sid, int(frame), _ptr(t), _stream()))
        return t

    def drain(self):
        check(lib().vsc_server_drain(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib().vsc_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_flo(path):
    """ReadFlowFile (flowIO.cpp:31-78): a .flo file as a float32 [H, W, 2] array."""
    w, h = ctypes.c_int(), ctypes.c_int()
    check(lib().vsc_flo_read(str(path).encode(), None, 0, ctypes.byref(w), ctypes.byref(h)))
    a = np.empty((h.value, w.value, 2), dtype=np.float32)
    check(lib().vsc_flo_read(str(path).encode(), a.ctypes.data, a.size, ctypes.byref(w), ctypes.byref(h)))
    return a


def write_flo(path, flow):
    a = np.ascontiguousarray(flow, dtype=np.float32)
    if a.ndim != 3 or a.shape[2] != 2:
        raise VscError(-1, "flow must be [H, W, 2]")
    check(lib().vsc_flo_write(str(path).encode(), a.ctypes.data, a.shape[1], a.shape[0]))


def flow_dir_path(flow_dir, frame, backward=False):
    """The .flo file the flow-file mode reads for window `frame`
    (stabilizefiles.cpp:134-149): frame_<frame+1>.flo (forward) or
    frame_<frame>_bwd.flo (backward)."""
    buf = ctypes.create_string_buffer(4096)
    check(lib().vsc_flow_dir_path(str(flow_dir).encode(), int(frame), int(bool(backward)), buf, 4096))
    return buf.value.decode()


def flow_dir_read(flow_dir, frame):
    """(forward, backward) flows of window `frame` from a flow directory."""
    w, h = ctypes.c_int(), ctypes.c_int()
    d = str(flow_dir).encode()
    check(lib().vsc_flow_dir_read(d, int(frame), None, None, 0, ctypes.byref(w), ctypes.byref(h)))
    f = np.empty((h.value, w.value, 2), np.float32)
    b = np.empty_like(f)
    check(lib().vsc_flow_dir_read(d, int(frame), f.ctypes.data, b.ctypes.data, f.size, ctypes.byref(w),
                                  ctypes.byref(h)))
    return f, b


def stabilize_flow_dir(orig, proc, flow_dir, params=None, mode=VSC_FRAME_FUSED):
    """The -f frame loop (FileStabilizer::stabilizeAll with flows from
    `flow_dir`): orig, proc uint8 [N, H, W, 4] host arrays.  Returns the N
    output RGBA8 frames (frame 0 and N-1 processed, 1..N-2 stabilized)."""
    orig = np.ascontiguousarray(orig, dtype=np.uint8)
    proc = np.ascontiguousarray(proc, dtype=np.uint8)
    n, H, W, _ = orig.shape
    po = (ctypes.c_void_p * n)(*[orig[i].ctypes.data for i in range(n)])
    pp = (ctypes.c_void_p * n)(*[proc[i].ctypes.data for i in range(n)])
    outs = [None] * n

    def sink(user, index, rgba, w, h):
        outs[index] = np.ctypeslib.as_array(ctypes.cast(rgba, ctypes.POINTER(ctypes.c_uint8)),
                                            shape=(h, w, 4)).copy()
        return 0

    cb = FRAME_SINK(sink)
    prm = (params or HyperParams()).c()
    check(lib().vsc_stabilize_flow_dir(po, pp, n, W, H, str(flow_dir).encode(), ctypes.byref(prm), int(mode), cb,
                                       None, _stream()))
    return outs


def profile_begin(capacity=4096):
    """Start timing the instrumented kernels with HIP events (see vsc.h)."""
    check(lib().vsc_profile_begin(int(capacity)))


def profile_end(kernel="consist_blocked"):
    """Return (launches, total_ms, algorithmic_bytes) for one kernel."""
    n = ctypes.c_int()
    ms = ctypes.c_double()
    by = ctypes.c_double()
    check(lib().vsc_profile_end(kernel.encode(), ctypes.byref(n), ctypes.byref(ms), ctypes.byref(by)))
    return n.value, ms.value, by.value


# ---------------------------------------------------------------------------
# synthetic inputs
# ---------------------------------------------------------------------------
def _synth():
    global _synth_lib
    if _synth_lib is None:
        if not os.path.exists(SYNTH_PATH):
            raise RuntimeError(f"{SYNTH_PATH} is missing: build with make -C video-stream-consistency_amd")
        L = ctypes.CDLL(SYNTH_PATH)
        L.vsc_synth_generate.argtypes = [ctypes.POINTER(SynthCfg), ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p]
        _synth_lib = L
    return _synth_lib


#: seed-generated workloads: small plain stabilization (640x360, 40 frames)
#: and the full 1080p form (1920x1080, 300 frames)
SYNTH_640x360 = dict(width=640, height=360, nframes=40, nblobs=200, sigma_min=5.0, sigma_max=40.0, motion=3.0,
                     npatches=1, patch_size=64, gain_min=0.8, gain_max=1.2, offset_max=20.0, noise_sigma=4.0,
                     seed=0)
SYNTH_1080P = dict(width=1920, height=1080, nframes=300, nblobs=800, sigma_min=5.0, sigma_max=40.0, motion=8.0,
                   npatches=4, patch_size=64, gain_min=0.8, gain_max=1.2, offset_max=20.0, noise_sigma=4.0,
                   seed=2)


#: the serving mix: 24 concurrent streams of four resolutions
SERVE_MIX = [(640, 360)] * 8 + [(1280, 720)] * 8 + [(1920, 1080)] * 6 + [(3840, 2160)] * 2


def serve_stream_video(i, nframes, size=None):
    """Frames of serving stream i (its own seed, the 1080p generator's content
    scaled to the stream's size): (orig, proc) uint8 [nframes, H, W, 4]."""
    w, h = size or SERVE_MIX[i]
    o, p, _, _ = synth_video(width=w, height=h, nframes=nframes, nblobs=max(100, int(800 * w * h / (1920 * 1080))),
                             motion=8.0, npatches=4, patch_size=max(16, w // 15), seed=1000 + i)
    return o, p


def synth_video(**kw):
    """Generate (orig, proc, fwd, bwd) numpy arrays: orig/proc uint8
    [N, H, W, 4] RGBA, fwd/bwd float32 [N, H, W, 2] (fwd[t] maps t -> t+1,
    bwd[t] maps t -> t-1; zero at the sequence ends)."""
    cfg = dict(SYNTH_640x360)
    cfg.update(kw)
    c = SynthCfg(**cfg)
    n, h, w = c.nframes, c.height, c.width
    orig = np.empty((n, h, w, 4), np.uint8)
    proc = np.empty((n, h, w, 4), np.uint8)
    fwd = np.empty((n, h, w, 2), np.float32)
    bwd = np.empty((n, h, w, 2), np.float32)
    rc = _synth().vsc_synth_generate(ctypes.byref(c), orig.ctypes.data, proc.ctypes.data, fwd.ctypes.data,
                                     bwd.ctypes.data)
    if rc != 0:
        raise VscError(rc, "synth_video: invalid configuration")
    return orig, proc, fwd, bwd
