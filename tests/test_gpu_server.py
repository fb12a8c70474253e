"""Multi-stream server: each stream's frames are the same bits alone and in a
mix of concurrent streams of other resolutions (continuous batching over a
pool of HIP streams), equal to the single-stream pipeline; admission by the
device-memory budget; parameter changes apply from the next tick."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [(192, 128), (256, 144), (128, 96)]


def _video(vsc, W, H, n, seed):
    orig, proc, _, _ = vsc.synth_video(width=W, height=H, nframes=n, nblobs=50, motion=3.0, npatches=1,
                                       patch_size=max(8, W // 8), seed=seed)
    return orig, proc


def _run(vsc, cuda, streams, order, form="chosen", hip_streams=4, batch=True, params=None):
    """streams: {name: (W, H, orig, proc)}; order: list of names, one submission
    each (in that order), ticks after every submission; params: {name: {frame
    submitted: HyperParams}} changes.  Returns {name: {frame: out}}."""
    import torch
    sv = vsc.Server(form=form, hip_streams=hip_streams, max_streams=max(64, len(streams)), batch=batch)
    ids = {k: sv.open(v[0], v[1]) for k, v in streams.items()}
    frames = {k: torch.from_numpy(np.concatenate([v[2][None], v[3][None]])).to(cuda) for k, v in streams.items()}
    nxt = {k: 0 for k in streams}
    outs = {k: {} for k in streams}
    def collect():
        for sid, tag, frame, _ in sv.poll():
            k = [n for n, i in ids.items() if i == sid][0]
            assert tag == frame + 1
            outs[k][frame] = sv.read(sid, frame).cpu()

    for name in order:
        t = nxt[name]
        if params and t in params.get(name, {}):
            sv.set_params(ids[name], params[name][t])
        while True:
            try:
                sv.submit(ids[name], frames[name][0][t], frames[name][1][t], tag=t)
                break
            except vsc.VscError as e:           # ring slot still in use: let work finish
                assert e.code == vsc.VSC_E_AGAIN
                sv.drain()
                collect()
                sv.tick()
        nxt[name] += 1
        sv.tick()
        collect()
    for _ in range(3):
        sv.drain()
        collect()
        sv.tick()
    sv.close()
    return outs


def test_server_alone_equals_mix_equals_pipeline(vsc, cuda):
    import torch
    vids = {f"s{i}": (W, H) + _video(vsc, W, H, 7, seed=30 + i) for i, (W, H) in enumerate(SIZES)}
    alone = _run(vsc, cuda, {"s0": vids["s0"]}, ["s0"] * 7)
    rng = np.random.default_rng(3)
    order = ["s0"] * 7 + ["s1"] * 7 + ["s2"] * 7
    rng.shuffle(order)
    mix = _run(vsc, cuda, vids, order)
    assert sorted(alone["s0"]) == [1, 2, 3, 4, 5]
    for f in alone["s0"]:
        assert torch.equal(alone["s0"][f], mix["s0"][f]), f"frame {f}"
    # the server adds scheduling only: same bits as the single-stream pipeline
    W, H, orig, proc = vids["s1"]
    o = torch.from_numpy(orig).to(cuda)
    p = torch.from_numpy(proc).to(cuda)
    pipe = vsc.Pipeline(W, H, form="chosen")
    pipe.set_last_rgba(p[2])
    for t in range(1, 6):
        ref = pipe.frame((o[t - 1], o[t], o[t + 1], p[t - 1], p[t], p[t + 1]), t).cpu()
        assert torch.equal(ref, mix["s1"][t]), f"pipeline frame {t}"


def test_server_admission_by_budget(vsc, cuda):
    probe = vsc.Server(hbm_budget=1 << 40)
    need = probe.footprint(256, 144)
    probe.close()
    sv = vsc.Server(hbm_budget=int(2.5 * need))
    a = sv.open(256, 144)
    b = sv.open(256, 144)
    with pytest.raises(vsc.VscError):
        sv.open(256, 144)                       # a third does not fit
    assert sv.used_bytes() == 2 * need
    sv.close_stream(a)
    c = sv.open(256, 144)                       # admitted once memory is free
    assert c >= 0 and sv.used_bytes() == 2 * need
    sv.close()


def test_server_params_apply_from_next_tick(vsc, cuda):
    import torch
    W, H = SIZES[0]
    orig, proc = _video(vsc, W, H, 6, seed=9)
    o = torch.from_numpy(orig).to(cuda)
    p = torch.from_numpy(proc).to(cuda)
    sv = vsc.Server(form="plain")
    sid = sv.open(W, H)
    hp = vsc.HyperParams()
    hp.gamma = 0.5
    got = {}
    for t in range(6):
        if t == 4:
            sv.set_params(sid, hp)             # frames stabilized from now on use gamma 0.5
        sv.submit(sid, o[t], p[t], tag=t)
        sv.tick()
        sv.drain()
        for _, _, frame, _ in sv.poll():
            got[frame] = sv.read(sid, frame).cpu()
    sv.close()
    pipe = vsc.Pipeline(W, H, form="plain")
    pipe.set_last_rgba(p[2])
    for t in range(1, 5):
        prm = hp if t >= 3 else None
        ref = pipe.frame((o[t - 1], o[t], o[t + 1], p[t - 1], p[t], p[t + 1]), t, params=prm).cpu()
        assert torch.equal(ref, got[t]), f"frame {t}"


@pytest.mark.parametrize("batch", [True, False])
@pytest.mark.parametrize("form", ["chosen", "plain"])
def test_server_graph_replay_matches_eager(vsc, cuda, form, batch):
    """Each stream's per-frame work replayed from captured hipGraphs (6 per
    stream, one per frame phase) gives the eager launches' bits on every
    frame: several replays of each graph, eager work and synchronisations in
    between, and a parameter change that forces a re-capture."""
    import torch
    W, H = 160, 96
    n = 22
    orig, proc = _video(vsc, W, H, n, seed=17)
    o = torch.from_numpy(orig).to(cuda)
    p = torch.from_numpy(proc).to(cuda)
    hp = vsc.HyperParams()
    hp.gamma = 1.25
    res = {}
    for graphs in (False, True):
        sv = vsc.Server(form=form, use_graphs=graphs, batch=batch)
        sid = sv.open(W, H)
        got = {}
        for t in range(n):
            if t == 13:
                sv.set_params(sid, hp)
            sv.submit(sid, o[t], p[t], tag=t)
            sv.tick()
            sv.drain()
            for _, _, frame, _ in sv.poll():
                got[frame] = sv.read(sid, frame).cpu()
            torch.cuda.synchronize()
        sv.close()
        res[graphs] = got
    assert sorted(res[True]) == list(range(1, n - 1))
    for f in res[False]:
        assert torch.equal(res[False][f], res[True][f]), f"frame {f}"


def test_server_read_into_pinned_host(vsc, cuda):
    """Egress straight into pinned host memory gives the device read's bytes;
    an unpinned host destination is refused before any copy."""
    import torch
    W, H = SIZES[1]
    orig, proc = _video(vsc, W, H, 5, seed=21)
    o = torch.from_numpy(orig).to(cuda)
    p = torch.from_numpy(proc).to(cuda)
    sv = vsc.Server(form="chosen")
    sid = sv.open(W, H)
    n = 0
    for t in range(5):
        sv.submit(sid, o[t], p[t], tag=t)
        sv.tick()
        sv.drain()
        for _, _, frame, _ in sv.poll():
            dev = sv.read(sid, frame)
            host = torch.empty((H, W, 4), dtype=torch.uint8, pin_memory=True)
            sv.read(sid, frame, out=host)
            torch.cuda.synchronize()
            assert torch.equal(dev.cpu(), host), f"frame {frame}"
            n += 1
    with pytest.raises(vsc.VscError):
        sv.read(sid, 3, out=torch.empty((H, W, 4), dtype=torch.uint8))
    sv.close()
    assert n == 3


def test_server_refuses_bad_params(vsc, cuda):
    """vsc_server_set_params validates against the stream (pyramid depth of
    its frame size, num_iter >= 0); a refused change leaves the stream's
    parameters and its captured graphs as they were."""
    import torch
    W, H = SIZES[2]
    orig, proc = _video(vsc, W, H, 7, seed=12)
    o = torch.from_numpy(orig).to(cuda)
    p = torch.from_numpy(proc).to(cuda)
    outs = {}
    for refuse in (False, True):
        sv = vsc.Server(form="plain")
        sid = sv.open(W, H)
        got = {}
        for t in range(7):
            if refuse and t == 4:
                bad = vsc.HyperParams()
                bad.pyramidLevels = 9
                with pytest.raises(vsc.VscError):
                    sv.set_params(sid, bad)
                bad = vsc.HyperParams()
                bad.numIter = -3
                with pytest.raises(vsc.VscError):
                    sv.set_params(sid, bad)
            sv.submit(sid, o[t], p[t], tag=t)
            sv.tick()
            sv.drain()
            for _, _, frame, _ in sv.poll():
                got[frame] = sv.read(sid, frame).cpu()
        sv.close()
        outs[refuse] = got
    assert sorted(outs[True]) == [1, 2, 3, 4, 5]
    for f in outs[False]:
        assert torch.equal(outs[False][f], outs[True][f]), f"frame {f}"


def test_serving_mix_each_stream_equals_its_pipeline(vsc, cuda):
    """The serving mix at its real sizes: 24 streams (8 x 640x360,
    8 x 1280x720, 6 x 1920x1080, 2 x 3840x2160), each its own seeded video,
    frames submitted in a shuffled interleaving (each stream in order).  Every
    stabilized frame equals the bits of that stream run alone through
    vsc_pipeline, and the committed device memory is the sum of the streams'
    footprints."""
    import torch
    nfr = 5
    vids = {}
    for i, (w, h) in enumerate(vsc.SERVE_MIX):
        o, p = vsc.serve_stream_video(i, nfr)
        vids[f"s{i:02d}"] = (w, h, o, p)
    probe = vsc.Server(hbm_budget=1 << 40, max_streams=len(vids))
    need = sum(probe.footprint(w, h) for (w, h) in vsc.SERVE_MIX)
    probe.close()
    rng = np.random.default_rng(11)
    order = [k for k in vids for _ in range(nfr)]
    rng.shuffle(order)
    sv = vsc.Server(hbm_budget=need, max_streams=len(vids))
    ids = {k: sv.open(v[0], v[1]) for k, v in vids.items()}
    assert sv.used_bytes() == need
    with pytest.raises(vsc.VscError):
        sv.open(640, 360)                      # the budget is exactly the mix
    sv.close()
    mix = _run(vsc, cuda, vids, order)
    for k, (w, h, orig, proc) in vids.items():
        assert sorted(mix[k]) == list(range(1, nfr - 1)), k
        o = torch.from_numpy(orig).to(cuda)
        p = torch.from_numpy(proc).to(cuda)
        pipe = vsc.Pipeline(w, h, form="chosen")
        pipe.set_last_rgba(p[2])
        for t in range(1, nfr - 1):
            ref = pipe.frame((o[t - 1], o[t], o[t + 1], p[t - 1], p[t], p[t + 1]), t).cpu()
            assert torch.equal(ref, mix[k][t]), f"{k} ({w}x{h}) frame {t}"
        pipe.close()


def test_server_batched_equals_unbatched_and_fallback(vsc, cuda):
    """Batched ticks (one set of persistent launches stabilizing every stream's
    window) give the bits of per-stream ticks, in a mix of sizes (odd ones
    included) with a stream switching to 3 pyramid levels mid-stream (those
    windows run per stream) and one to 1 level and fewer sweeps (batched)."""
    import torch
    sizes = {"a": (192, 128), "b": (130, 96), "c": (256, 144), "d": (101, 75)}
    vids = {k: (W, H) + _video(vsc, W, H, 9, seed=50 + i) for i, (k, (W, H)) in enumerate(sizes.items())}
    three = vsc.HyperParams()
    three.pyramidLevels = 3
    one = vsc.HyperParams()
    one.pyramidLevels = 1
    one.numIter = 37
    params = {"a": {4: three, 7: vsc.HyperParams()}, "c": {5: one}}
    rng = np.random.default_rng(5)
    order = [k for k in vids for _ in range(9)]
    rng.shuffle(order)
    for form in ("chosen", "plain"):
        res = {b: _run(vsc, cuda, vids, order, form=form, batch=b, params=params) for b in (True, False)}
        for k in vids:
            assert sorted(res[True][k]) == list(range(1, 8)), k
            for f in res[False][k]:
                assert torch.equal(res[True][k][f], res[False][k][f]), (form, k, f)
