"""Training data and loop for the learned flow upsampler (vsc_upsampler_*).

Frames and flows come from the seeded synthetic generator (libvsc_synth.so):
textured frames with independently moving occluder patches, so the flow has
sharp motion boundaries that a bilinear upsampler smears and that the frame's
luma can locate.  A training pair is (frame, full-resolution ground-truth flow,
half-resolution flow), the half-resolution flow being the 2x2 box average of
the full one divided by 2 (the flow the network produces at FLOWDOWNSCALE=2,
with the vectors in low-resolution pixels).  Forward, backward and Adam all
run in the library's HIP kernels; this module only feeds them.
"""
from __future__ import annotations

import importlib
import os
import sys

import numpy as np

if not __package__:  # run as a script: make the package importable from the repository root
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
_pkg = importlib.import_module(__package__ or "video-stream-consistency_amd")

WEIGHTS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "upsampler_weights.npy")


def make_pairs(n_pairs, size=256, seq_len=50, seed=0):
    """(frames [n, size, size, 4] u8, gt [n, size, size, 2] f32, lo [n, size/2, size/2, 2] f32)."""
    frames, gts = [], []
    k = 0
    while sum(len(f) for f in frames) < n_pairs:
        n = min(seq_len, n_pairs - sum(len(f) for f in frames) + 1)
        orig, _, fwd, _ = _pkg.synth_video(width=size, height=size, nframes=max(n, 2), nblobs=80, sigma_min=4.0,
                                           sigma_max=30.0, motion=8.0, npatches=4, patch_size=size // 5,
                                           seed=seed * 1000 + k)
        frames.append(orig[:-1])
        gts.append(fwd[:-1])
        k += 1
    frames = np.concatenate(frames)[:n_pairs]
    gt = np.concatenate(gts)[:n_pairs]
    lo = low_res(gt)
    return np.ascontiguousarray(frames), np.ascontiguousarray(gt), lo


def low_res(gt):
    """2x2 box average, vectors halved: [n, H, W, 2] -> [n, H/2, W/2, 2]."""
    n, H, W, _ = gt.shape
    g = gt[:, :H // 2 * 2, :W // 2 * 2].reshape(n, H // 2, 2, W // 2, 2, 2).mean(axis=(2, 4)) * 0.5
    return np.ascontiguousarray(g.astype(np.float32))


def epe(pred, gt):
    return float(np.sqrt(((np.asarray(pred, np.float64) - gt) ** 2).sum(-1)).mean())


def train(ups, frames, gt, lo, steps, batch=16, lr=1e-3, seed=0, log=None):
    """Adam steps on random batches of the pairs (device tensors)."""
    import torch
    rng = np.random.default_rng(seed)
    dev = "cuda"
    F = torch.from_numpy(frames).to(dev)
    G = torch.from_numpy(gt).to(dev)
    L = torch.from_numpy(lo).to(dev)
    n = F.shape[0]
    losses = []
    for s in range(steps):
        idx = torch.from_numpy(rng.choice(n, size=min(batch, n), replace=False)).to(dev)
        want = log is not None and (s % log == 0 or s == steps - 1)
        loss = ups.train_step(L[idx].contiguous(), F[idx].contiguous(), G[idx].contiguous(), lr=lr, want_loss=want)
        if want:
            losses.append((s, loss))
    return losses


def evaluate(ups, frames, gt, lo, scale=2.0):
    """Mean EPE of the learned and of the bilinear upsampling (gate 0)."""
    import torch
    gate0 = torch.zeros(1, dtype=torch.float32, device="cuda")
    e_l, e_b = [], []
    for i in range(frames.shape[0]):
        fr = torch.from_numpy(frames[i]).cuda()
        fl = torch.from_numpy(lo[i]).cuda()
        e_l.append(epe(ups.run(fl, fr, scale=scale).cpu().numpy(), gt[i]))
        e_b.append(epe(ups.run(fl, fr, scale=scale, gate=gate0).cpu().numpy(), gt[i]))
    return float(np.mean(e_l)), float(np.mean(e_b))


def main():
    """Train on 2000 pairs (256 x 256, batch 16, Adam lr 1e-3) and save the weights."""
    import argparse
    import time
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=2000)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--out", default=WEIGHTS_PATH)
    a = ap.parse_args()
    t0 = time.time()
    frames, gt, lo = make_pairs(a.pairs, seed=1)
    vf, vg, vl = make_pairs(64, seed=99)
    print(f"data: {frames.shape[0]} pairs in {time.time() - t0:.1f} s", flush=True)
    ups = _pkg.Upsampler(seed=0)
    t0 = time.time()
    for s0 in range(0, a.steps, 500):
        losses = train(ups, frames, gt, lo, min(500, a.steps - s0), seed=s0, log=250)
        el, eb = evaluate(ups, vf, vg, vl)
        print(f"step {s0 + min(500, a.steps - s0)}: loss {losses[-1][1]:.4f}  val EPE learned {el:.4f} "
              f"bilinear {eb:.4f}  ({time.time() - t0:.1f} s)", flush=True)
    np.save(a.out, ups.params())
    print(f"saved {a.out}")


if __name__ == "__main__":
    main()
