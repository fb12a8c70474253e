"""ORACLE — test infrastructure only.

ctypes/numpy front end of ``libvsc_oracle.so`` (built from ``vsc_oracle.cpp`` by
the Makefile next to this file).  Only tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import this module; it is the
checker, never the thing measured or shipped.

Every wrapper takes and returns numpy arrays in the reference's own layouts:
NCHW for the ORT ops (``Correlation``, ``Warp``) and HWC interleaved float
images for the stabilization stages.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libvsc_oracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_i = ctypes.c_int
_f = ctypes.c_float
_d = ctypes.c_double


def build() -> str:
    """Compile the oracle library with its Makefile (g++, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.vsco_set_threads.restype = _i
        L.vsco_set_threads.argtypes = [_i]
        L.vsco_expf.restype = _f
        L.vsco_expf.argtypes = [_f]
        for name, p in (("f32", _f32p), ("f64", _f64p)):
            for fn in ("vsco_correlation_", "vsco_correlation_legacy_"):
                g = getattr(L, fn + name)
                g.restype = None
                g.argtypes = [p, p, p, _i, _i, _i, _i, _i]
            g = getattr(L, "vsco_warp_" + name)
            g.restype = None
            g.argtypes = [p, p, p, _i, _i, _i, _i]
            g = getattr(L, "vsco_stab_warp_" + name)
            g.restype = None
            g.argtypes = [p, p, _i, p, _i, _i]
            sc = _f if name == "f32" else _d
            g = getattr(L, "vsco_stab_adap_comb_" + name)
            g.restype = None
            g.argtypes = [p] * 9 + [_i, _i, sc]
            g = getattr(L, "vsco_stab_consist_wt_" + name)
            g.restype = None
            g.argtypes = [p, p, p, _i, _i, sc, sc]
            g = getattr(L, "vsco_stab_consist_out_" + name)
            g.restype = None
            g.argtypes = [p, p, p, _i, sc, sc, p, _i, _i]
            g = getattr(L, "vsco_stab_bilinear_" + name)
            g.restype = None
            g.argtypes = [p, _i, _i, _i, p, _i, _i, _i]
            g = getattr(L, "vsco_stab_frame_" + name)
            g.restype = None
            g.argtypes = [p] * 8 + [_i, p, _f, _f, _f, _i, _i, _f, _f, p, _i, _i]
        L.vsco_stab_frame_mutant_f64.restype = None
        L.vsco_stab_frame_mutant_f64.argtypes = [_f64p] * 8 + [_i, _f64p, _f, _f, _f, _i, _i, _f, _f, _f64p, _i,
                                                               _i, _i]
        L.vsco_stab_consist_out_fma_f32.restype = None
        L.vsco_stab_consist_out_fma_f32.argtypes = [_f32p, _f32p, _f32p, _i, _f, _f, _f32p, _i, _i]
        L.vsco_stab_frame_fma_f32.restype = None
        L.vsco_stab_frame_fma_f32.argtypes = [_f32p] * 8 + [_i, _f32p, _f, _f, _f, _i, _i, _f, _f, _f32p, _i, _i]
        L.vsco_average_frames_f32.restype = None
        L.vsco_average_frames_f32.argtypes = [_f32p, _f32p, _f32p, _i, _i, _i]
        L.vsco_rgba_to_float.restype = None
        L.vsco_rgba_to_float.argtypes = [_u8p, _f32p, _i, _i]
        L.vsco_float_to_rgba.restype = None
        L.vsco_float_to_rgba.argtypes = [_f32p, _u8p, _i, _i]
        L.vsco_double_to_rgba.restype = None
        L.vsco_double_to_rgba.argtypes = [_f64p, _u8p, _i, _i]
        L.vsco_read_flo.restype = _i
        L.vsco_read_flo.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_longlong,
                                    ctypes.POINTER(_i), ctypes.POINTER(_i)]
        _lib = L
    return _lib


def set_threads(n: int) -> int:
    """Set the oracle's OpenMP thread count; returns the count in effect."""
    return lib().vsco_set_threads(int(n))


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def _sfx(dt):
    return "f32" if np.dtype(dt) == np.float32 else "f64"


def expf(x: float) -> float:
    return lib().vsco_expf(float(x))


def expf_array(x: np.ndarray) -> np.ndarray:
    f = lib().vsco_expf
    flat = np.asarray(x, dtype=np.float32).ravel()
    return np.array([f(float(v)) for v in flat], dtype=np.float32).reshape(np.shape(x))


# ---- ORT ops (NCHW) -------------------------------------------------------

def correlation(f1, f2, max_displacement=4, legacy=False, dtype=np.float32):
    """Correlation op. Output [N, 2md+1, 2md+1, H, W] (non-legacy) or
    [N, (2md+1)^2, H, W] (legacy)."""
    f1 = _c(f1, dtype)
    f2 = _c(f2, dtype)
    N, C, H, W = f1.shape
    P = 2 * max_displacement + 1
    out = np.empty((N, P, P, H, W), dtype=dtype)
    fn = ("vsco_correlation_legacy_" if legacy else "vsco_correlation_") + _sfx(dtype)
    getattr(lib(), fn)(f1, f2, out, N, C, H, W, max_displacement)
    return out.reshape(N, P * P, H, W) if legacy else out


def warp(inp, flow, dtype=np.float32):
    """Warp op. inp [N,C,H,W], flow [N,2,H,W] -> [N,C,H,W]."""
    inp = _c(inp, dtype)
    flow = _c(flow, dtype)
    N, C, H, W = inp.shape
    out = np.empty_like(inp)
    getattr(lib(), "vsco_warp_" + _sfx(dtype))(inp, flow, out, N, C, H, W)
    return out


# ---- stabilization stages (HWC float images) --------------------------------

def stab_warp(img, flow, dtype=np.float32):
    img = _c(img, dtype)
    flow = _c(flow, dtype)
    H, W, _ = img.shape
    out = np.empty((H, W, 3), dtype=dtype)
    getattr(lib(), "vsco_stab_warp_" + _sfx(dtype))(img, flow, flow.shape[2], out, W, H)
    return out


def stab_adap_comb(cIn, cPr, pIn, pPr, nIn, nPr, last, alpha, dtype=np.float32):
    args = [_c(a, dtype) for a in (cIn, cPr, pIn, pPr, nIn, nPr)]
    last = _c(last, dtype)
    H, W, _ = args[0].shape
    adIn = np.empty((H, W, 3), dtype=dtype)
    adPr = np.empty((H, W, 3), dtype=dtype)
    getattr(lib(), "vsco_stab_adap_comb_" + _sfx(dtype))(*args, adIn, adPr, last, W, H, alpha)
    return adIn, adPr


def stab_consist_wt(adIn, cIn, beta, gamma, dtype=np.float32):
    adIn = _c(adIn, dtype)
    cIn = _c(cIn, dtype)
    H, W, _ = cIn.shape
    wt = np.empty((H, W, 3), dtype=dtype)
    getattr(lib(), "vsco_stab_consist_wt_" + _sfx(dtype))(adIn, cIn, wt, W, H, beta, gamma)
    return wt


def _solver_fn(base, solver, dtype):
    if solver == "fma":
        if dtype != np.float32:
            raise ValueError("the FMA-form solver is restated in float32 only")
        return getattr(lib(), base + "fma_f32")
    if solver != "exact":
        raise ValueError(f"unknown solver {solver!r}")
    return getattr(lib(), base + _sfx(dtype))


def stab_consist_out(P, A, Wt, O_init, iters, step=0.15, mom=0.15, dtype=np.float32, solver="exact"):
    """solver "exact": the reference op order; "fma": the FMA form."""
    P, A, Wt = (_c(a, dtype) for a in (P, A, Wt))
    O = _c(O_init, dtype).copy()
    H, W, _ = P.shape
    _solver_fn("vsco_stab_consist_out_", solver, dtype)(P, A, Wt, iters, step, mom, O, W, H)
    return O


def stab_bilinear(img, out_w, out_h, out_c=None, dtype=np.float32):
    img = _c(img, dtype)
    H, W, C = img.shape
    oc = C if out_c is None else out_c
    out = np.empty((out_h, out_w, oc), dtype=dtype)
    getattr(lib(), "vsco_stab_bilinear_" + _sfx(dtype))(img, W, H, C, out, out_w, out_h, oc)
    return out


def average_frames(a, b):
    a = _c(a, np.float32)
    b = _c(b, np.float32)
    H, W, C = a.shape
    out = np.empty_like(a)
    lib().vsco_average_frames_f32(a, b, out, W, H, C)
    return out


def rgba_to_float(rgba):
    rgba = _c(rgba, np.uint8)
    H, W, _ = rgba.shape
    out = np.empty((H, W, 3), dtype=np.float32)
    lib().vsco_rgba_to_float(rgba, out, W, H)
    return out


def float_to_rgba(img):
    img = np.ascontiguousarray(img)
    H, W, _ = img.shape
    out = np.empty((H, W, 4), dtype=np.uint8)
    if img.dtype == np.float64:
        lib().vsco_double_to_rgba(img, out, W, H)
    else:
        lib().vsco_float_to_rgba(_c(img, np.float32), out, W, H)
    return out


DEFAULT_PARAMS = dict(alpha=6800.0, beta=6800.0, gamma=2.0, pyramid_levels=2, num_iter=150,
                      step_size=0.15, mom_fac=0.15)


def stab_frame(o_prev, o_cur, o_next, p_prev, p_cur, p_next, fwd, bwd, last, params=None,
               dtype=np.float32, solver="exact"):
    """Stabilization body of one doOneStep call (float HWC frames); solver
    "exact" (reference op order) or "fma" (FMA form, float32)."""
    prm = dict(DEFAULT_PARAMS)
    if params:
        prm.update(params)
    ims = [_c(a, dtype) for a in (o_prev, o_cur, o_next, p_prev, p_cur, p_next)]
    fwd = _c(fwd, dtype)
    bwd = _c(bwd, dtype)
    last = _c(last, dtype)
    H, W, _ = ims[0].shape
    out = np.empty((H, W, 3), dtype=dtype)
    _solver_fn("vsco_stab_frame_", solver, dtype)(
        *ims, fwd, bwd, fwd.shape[2], last, prm["alpha"], prm["beta"], prm["gamma"],
        prm["pyramid_levels"], prm["num_iter"], prm["step_size"], prm["mom_fac"], out, W, H)
    return out


def stab_frame_mutant(o_prev, o_cur, o_next, p_prev, p_cur, p_next, fwd, bwd, last, params=None, mutation=0):
    """stab_frame in float64 with a planted schedule mistake (see
    FrameParams::mutation in vsc_oracle.cpp); for tests that show the
    schedule pins detect such mistakes."""
    prm = dict(DEFAULT_PARAMS)
    if params:
        prm.update(params)
    ims = [_c(a, np.float64) for a in (o_prev, o_cur, o_next, p_prev, p_cur, p_next)]
    fwd = _c(fwd, np.float64)
    bwd = _c(bwd, np.float64)
    last = _c(last, np.float64)
    H, W, _ = ims[0].shape
    out = np.empty((H, W, 3), dtype=np.float64)
    lib().vsco_stab_frame_mutant_f64(*ims, fwd, bwd, fwd.shape[2], last, prm["alpha"], prm["beta"], prm["gamma"],
                                     prm["pyramid_levels"], prm["num_iter"], prm["step_size"], prm["mom_fac"], out,
                                     W, H, int(mutation))
    return out


def stabilize_sequence(orig_rgba, proc_rgba, fwd_flows, bwd_flows, params=None, dtype=np.float32, solver="exact"):
    """The frame loop of FileStabilizer::stabilizeAll with k = 1, batch 1
    (videostabilizer.cpp:136-164,167-265): frames 0 and N-1 are emitted as the
    processed frames; frame i in 1..N-2 is stabilized with the window
    (i-1, i, i+1), fwd_flows[i] (i -> i+1) and bwd_flows[i] (i -> i-1); the
    initial `last` is processed frame 2 (= 2k+B-1).  Returns (rgba outputs,
    float stabilized frames)."""
    n = len(orig_rgba)
    to_f = (lambda a: rgba_to_float(a).astype(dtype))
    O = [to_f(a) for a in orig_rgba]
    Pr = [to_f(a) for a in proc_rgba]
    outs = [None] * n
    floats = [None] * n
    outs[0] = float_to_rgba(Pr[0])
    last = Pr[2].copy()
    for i in range(1, n - 1):
        st = stab_frame(O[i - 1], O[i], O[i + 1], Pr[i - 1], Pr[i], Pr[i + 1], fwd_flows[i], bwd_flows[i], last,
                        params, dtype, solver)
        outs[i] = float_to_rgba(st)
        floats[i] = st
        last = st
    outs[n - 1] = float_to_rgba(Pr[n - 1])
    return outs, floats


def read_flo(path):
    """ReadFlowFile (flowIO.cpp:31-78). Returns (flow HxWx2 float32, code)."""
    w = _i()
    h = _i()
    rc = lib().vsco_read_flo(path.encode(), None, 0, ctypes.byref(w), ctypes.byref(h))
    if rc != 0:
        return None, rc
    buf = np.empty((h.value, w.value, 2), dtype=np.float32)
    rc = lib().vsco_read_flo(path.encode(), buf.ctypes.data, buf.size, ctypes.byref(w), ctypes.byref(h))
    if rc != 0:
        return None, rc
    return buf, 0
