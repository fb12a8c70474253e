"""ORACLE — test infrastructure only: a float64 CPU restatement (torch as a
plain tensor library) of the flow network that the library runs, used to check
the HIP path.  Only tests/ may import it.

The network is "pwcnet-my-light4dec" (model_spec.py:32: decoder_light=True,
decoder_count=4, refinement on, legacy=False) wrapped by the ONNX export's
pre/post-processing (convert_onnx.py:16-35,63-88):

* preprocessing: RGBA8 -> RGB/255 -> bilinear resize (align_corners=False) to
  multiples of 64 (convert_onnx.py:16-20,66-71); with FLOWDOWNSCALE the frame is
  first resized by nearest sampling (flowmodel.cpp:126-133);
* Extractor (pwcnet.py:179-253): 6 modules of 3 Conv2d 3x3 (+LeakyReLU 0.1),
  the first of each with stride 2; channels 16, 32, 64, 96, 128, 196;
* DecoderLight at levels 6, 5, 4, 3 (pwcnet.py:458-611): upflow =
  ConvTranspose2d(2, 2, 4, 2, 1) of the previous flow, f2 warped by
  upflow * dblBackward (2.5 / 1.25 / 0.625 at levels 3 / 4 / 5), cost volume
  leaky(Correlation(f1, warped f2)), then the DenseNet-light conv stack;
* Refiner (pwcnet.py:625-659) on cat([x5, flow]) of level 3, added to the flow
  (pwcnet.py:981-986);
* final: bilinear x8 (align_corners=True) of flow * 20 (pwcnet.py:990-1003),
  then resize to the frame (align_corners=False) with channel 0 scaled by
  H/Hp and channel 1 by W/Wp, as the wrapper does (convert_onnx.py:27-34).

Parameters are supplied by the caller in the library's export order
(vsc_flownet_get_params), so this module never draws random numbers itself.

Pins: the building blocks are torch's conv2d / conv_transpose2d / interpolate
(independent implementations); the correlation and the warp are restated
here from the op definitions (correlation.cc:203-275; warp.cc:71-134) and are
checked against the C oracle of those ops in tests/test_oracle_flownet.py.
"""
from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F

FEAT_C = [3, 16, 32, 64, 96, 128, 196]
DBL = {3: 2.5, 4: 1.25, 5: 0.625}
DEC_OUT = [128, 128, 96, 64, 32, 2]
REF_OUT = [128, 128, 128, 96, 64, 32, 2]
REF_DIL = [1, 2, 4, 8, 16, 1, 1]


def layer_shapes():
    """(weight shape, bias length, kind) in export order."""
    out = []
    for k in range(1, 7):
        ci, co = FEAT_C[k - 1], FEAT_C[k]
        out += [((co, ci, 3, 3), co, "conv"), ((co, co, 3, 3), co, "conv"), ((co, co, 3, 3), co, "conv")]
    for _ in (5, 4, 3):
        out.append(((2, 2, 4, 4), 2, "convT"))
    for lvl in (6, 5, 4, 3):
        cin = [83 + FEAT_C[lvl], 128, 256, 224, 160, 96]
        for j in range(6):
            out.append(((DEC_OUT[j], cin[j], 3, 3), DEC_OUT[j], "conv"))
    ci = 34
    for j in range(7):
        out.append(((REF_OUT[j], ci, 3, 3), REF_OUT[j], "conv"))
        ci = REF_OUT[j]
    return out


def unpack_params(flat: np.ndarray):
    """Split the exported float32 vector into [(weight, bias)] float64 tensors."""
    res, o = [], 0
    for wshape, nb, _ in layer_shapes():
        nw = int(np.prod(wshape))
        w = torch.from_numpy(flat[o:o + nw].astype(np.float64).reshape(wshape))
        o += nw
        b = torch.from_numpy(flat[o:o + nb].astype(np.float64))
        o += nb
        res.append((w, b))
    if o != flat.size:
        raise ValueError(f"parameter count mismatch: {o} != {flat.size}")
    return res


def leaky(x):
    return F.leaky_relu(x, 0.1)


def preprocess(rgba: np.ndarray, down: int):
    """RGBA8 [H, W, 4] -> NCHW float64 [1, 3, Hp, Wp] (and Hd, Wd)."""
    H, W = rgba.shape[:2]
    Wd, Hd = int(round(W / down)), int(round(H / down))
    img = rgba[..., :3]
    if (Wd, Hd) != (W, H):
        # nearest by pixel centres, in float32 as the device computes it
        xs = np.minimum(W - 1, ((np.arange(Wd, dtype=np.float32) + np.float32(0.5)) *
                               np.float32(np.float32(W) / np.float32(Wd))).astype(np.int64))
        ys = np.minimum(H - 1, ((np.arange(Hd, dtype=np.float32) + np.float32(0.5)) *
                               np.float32(np.float32(H) / np.float32(Hd))).astype(np.int64))
        img = img[ys][:, xs]
    x = torch.from_numpy((img.astype(np.float32) / np.float32(255.0)).astype(np.float64))
    x = x.permute(2, 0, 1)[None]
    Hp, Wp = (Hd + 63) // 64 * 64, (Wd + 63) // 64 * 64
    return F.interpolate(x, size=(Hp, Wp), mode="bilinear", align_corners=False), Hd, Wd


def extractor(P, x):
    feats = []
    for k in range(6):
        for j in range(3):
            w, b = P[3 * k + j]
            x = leaky(F.conv2d(x, w, b, stride=2 if j == 0 else 1, padding=1))
        feats.append(x)
    return feats  # levels 1..6


def correlation(f1, f2, md=4):
    """Non-legacy cost volume [N, (2md+1)^2, H, W], zero outside the image."""
    N, C, H, W = f1.shape
    p = F.pad(f2, (md, md, md, md))
    out = []
    for dy in range(2 * md + 1):
        for dx in range(2 * md + 1):
            out.append((f1 * p[:, :, dy:dy + H, dx:dx + W]).sum(1))
    return torch.stack(out, 1)


def warp(x, flow32: np.ndarray, scale: float):
    """Warp op (warp.cc:71-134) of NCHW x by the float32 NHWC flow * scale.
    Taps and the 0.999 mask rule are evaluated in float32 exactly as the
    device evaluates them; the blend is float64."""
    N, C, H, W = x.shape
    f = flow32.astype(np.float32) * np.float32(scale)
    xs = np.arange(W, dtype=np.float32)[None, None, :] + f[..., 0]
    ys = np.arange(H, dtype=np.float32)[None, :, None] + f[..., 1]
    xL, yT = np.floor(xs), np.floor(ys)
    al, be = xs - xL, ys - yT
    xR, yB = xL + np.float32(1), yT + np.float32(1)
    mL = (xL >= 0) & (xL <= W - 1)
    mR = (xR >= 0) & (xR <= W - 1)
    mT = (yT >= 0) & (yT <= H - 1)
    mB = (yB >= 0) & (yB <= H - 1)
    one = np.float32(1)
    wTL, wTR = (one - al) * (one - be), al * (one - be)
    wBL, wBR = (one - al) * be, al * be
    z = np.float32(0)
    mask = np.zeros_like(xs)
    mask = mask + np.where(mT & mL, wTL, z)
    mask = mask + np.where(mT & mR, wTR, z)
    mask = mask + np.where(mB & mL, wBL, z)
    mask = mask + np.where(mB & mR, wBR, z)
    keep = mask >= np.float32(0.999)
    out = torch.zeros_like(x)
    for n in range(N):
        acc = torch.zeros((C, H, W), dtype=torch.float64)
        for m, ix, iy, wt in ((mT & mL, xL, yT, wTL), (mT & mR, xR, yT, wTR),
                              (mB & mL, xL, yB, wBL), (mB & mR, xR, yB, wBR)):
            sel = m[n] & keep[n]
            ixc = np.clip(ix[n], 0, W - 1).astype(np.int64)
            iyc = np.clip(iy[n], 0, H - 1).astype(np.int64)
            v = x[n][:, torch.from_numpy(iyc), torch.from_numpy(ixc)]
            acc = acc + torch.from_numpy(np.where(sel, wt[n], 0).astype(np.float64))[None] * v
        out[n] = acc
    return out


def upflow(P, lvl, flow):
    w, b = P[18 + (5 - lvl)]
    return F.conv_transpose2d(flow, w, b, stride=2, padding=1)


def decoder(P, lvl, f1, f2, upflow32=None):
    """One DecoderLight level.  upflow32: the float32 NHWC upsampled flow that
    enters the level (None at level 6).  Returns the NCHW flow and x5."""
    base = 21 + 6 * (6 - lvl)
    if upflow32 is None:
        vol = leaky(correlation(f1, f2))
        fl = torch.zeros((f1.shape[0], 2) + f1.shape[2:], dtype=torch.float64)
    else:
        fl = torch.from_numpy(upflow32.astype(np.float64)).permute(0, 3, 1, 2)
        vol = leaky(correlation(f1, warp(f2, upflow32, DBL[lvl])))
    feat = torch.cat([vol, f1, fl], 1)

    def c(j, x, act=True):
        w, b = P[base + j]
        y = F.conv2d(x, w, b, padding=1)
        return leaky(y) if act else y

    x1 = c(0, feat)
    x2 = c(1, x1)
    x3 = c(2, torch.cat([x1, x2], 1))
    x4 = c(3, torch.cat([x2, x3], 1))
    x5 = c(4, torch.cat([x3, x4], 1))
    flow = c(5, torch.cat([x4, x5], 1), act=False)
    return flow, x5


def refiner(P, x):
    for j in range(7):
        w, b = P[45 + j]
        x = F.conv2d(x, w, b, padding=REF_DIL[j], dilation=REF_DIL[j])
        if j < 6:
            x = leaky(x)
    return x


def final_flow(flow3, Hp, Wp, Hd, Wd):
    up = F.interpolate(flow3 * 20.0, scale_factor=8, mode="bilinear", align_corners=True)
    assert up.shape[2:] == (Hp, Wp)
    out = F.interpolate(up, size=(Hd, Wd), mode="bilinear", align_corners=False)
    out = out * torch.tensor([Hd / Hp, Wd / Wp], dtype=torch.float64)[None, :, None, None]
    return out.permute(0, 2, 3, 1)  # NHWC


def flow_pair(P, feats1, feats2, upflows32=None):
    """Whole decoder stack for one pair.  upflows32: optional {level: float32
    NHWC upsampled flow} from the device (teacher forcing at the warps).
    Returns dict of NHWC float64 flows: 'flow6'..'flow3', 'up5'..'up3', 'refined'."""
    res = {}
    fl = None
    for lvl in (6, 5, 4, 3):
        f1, f2 = feats1[lvl - 1], feats2[lvl - 1]
        if lvl == 6:
            fl, x5 = decoder(P, 6, f1, f2)
        else:
            up = upflow(P, lvl, fl)
            res[f"up{lvl}"] = up.permute(0, 2, 3, 1)
            up32 = upflows32[lvl] if upflows32 is not None else up.permute(0, 2, 3, 1).numpy().astype(np.float32)
            fl, x5 = decoder(P, lvl, f1, f2, up32)
        res[f"flow{lvl}"] = fl.permute(0, 2, 3, 1)
    refined = fl + refiner(P, torch.cat([x5, fl], 1))
    res["refined"] = refined.permute(0, 2, 3, 1)
    res["refined_nchw"] = refined
    return res
