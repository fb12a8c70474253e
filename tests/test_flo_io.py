"""Product-side .flo reader/writer (vsc_flo_read / vsc_flo_write) against the
format (ReadFlowFile, flowIO.cpp:31-78) and the oracle's reader.  Host-only:
runs without a GPU."""
import struct

import numpy as np
import pytest


def test_flo_roundtrip_and_oracle(vsc, orc, tmp_path):
    rng = np.random.default_rng(0)
    fl = rng.standard_normal((7, 11, 2)).astype(np.float32) * 5
    p = tmp_path / "a.flo"
    vsc.write_flo(p, fl)
    raw = p.read_bytes()
    assert struct.unpack("<fii", raw[:12]) == (202021.25, 11, 7) and len(raw) == 12 + fl.size * 4
    got = vsc.read_flo(p)
    assert np.array_equal(got, fl)
    ref, code = orc.read_flo(str(p))
    assert code == 0 and np.array_equal(ref, fl)


@pytest.mark.parametrize("case", ["tag", "width", "height", "short", "long", "missing"])
def test_flo_rejects_malformed(vsc, tmp_path, case):
    fl = np.zeros((3, 4, 2), np.float32)
    head = {"tag": struct.pack("<fii", 1.0, 4, 3), "width": struct.pack("<fii", 202021.25, 0, 3),
            "height": struct.pack("<fii", 202021.25, 4, 100000)}.get(case, struct.pack("<fii", 202021.25, 4, 3))
    body = fl.tobytes()
    if case == "short":
        body = body[:-4]
    if case == "long":
        body = body + b"\0"
    p = tmp_path / "b.flo"
    if case != "missing":
        p.write_bytes(head + body)
    with pytest.raises(vsc.VscError):
        vsc.read_flo(p)
    with pytest.raises(vsc.VscError):
        vsc.write_flo(tmp_path / "c.flo", np.zeros((3, 4, 3), np.float32))


def test_flow_dir_naming_and_read(vsc, tmp_path):
    """The flow-file mode's naming (stabilizefiles.cpp:99-101,134-149): window i
    reads its forward flow i -> i+1 from frame_<i+1>.flo and its backward flow
    i -> i-1 from frame_<i>_bwd.flo, indices zero-padded to 6 digits."""
    rng = np.random.default_rng(3)
    n, H, W = 5, 6, 9
    fwd = rng.standard_normal((n, H, W, 2)).astype(np.float32)
    bwd = rng.standard_normal((n, H, W, 2)).astype(np.float32)
    for i in range(1, n - 1):
        vsc.write_flo(tmp_path / f"frame_{i + 1:06d}.flo", fwd[i])
        vsc.write_flo(tmp_path / f"frame_{i:06d}_bwd.flo", bwd[i])
    assert vsc.flow_dir_path(tmp_path, 7).endswith("/frame_000008.flo")
    assert vsc.flow_dir_path(tmp_path, 123456, backward=True).endswith("/frame_123456_bwd.flo")
    for i in range(1, n - 1):
        f, b = vsc.flow_dir_read(tmp_path, i)
        assert np.array_equal(f, fwd[i]) and np.array_equal(b, bwd[i])
    with pytest.raises(vsc.VscError):
        vsc.flow_dir_read(tmp_path, n - 1)          # frame_000005.flo does not exist
    vsc.write_flo(tmp_path / f"frame_{n:06d}.flo", np.zeros((H, W + 1, 2), np.float32))
    vsc.write_flo(tmp_path / f"frame_{n - 1:06d}_bwd.flo", np.zeros((H, W, 2), np.float32))
    with pytest.raises(vsc.VscError):
        vsc.flow_dir_read(tmp_path, n - 1)          # forward and backward sizes differ
