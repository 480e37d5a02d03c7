"""On-disk field files: the checkpoint data_RR.bin (escreve,
escreve_par.f:6-33, read back by main_par.f:763-767) and the base flow
based_RR.bin (baseflow2D/main_par_2D.f:578-595, read by main_par.f:847-850).

The checkpoint writer is pinned against the bytes the reference's own
escreve wrote (tests/golden/ref_checkpoint.json: size, SHA-256 and the
first and last bytes of its file for seeded fields; script tests/golden/
make_ref_vectors.py --checkpoint).  The base-flow layout is the same
Fortran sequential record convention, checked against bytes assembled here
by hand: a 4-byte length, the three arrays, the length again.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import trans_bl_amd as tb
from trans_bl_amd import workload as wl

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _ckpt():
    with open(os.path.join(GOLD, "ref_checkpoint.json")) as f:
        return json.load(f)


def _fields(g):
    """Seeded fields in the order the reference's writer takes them
    (ux, uy, uz, wx, wy, wz): ux = wy = dwzdx, uz = wx = wz = wx."""
    dwzdx = wl.wavepacket_rhs(g, tag="refckpt")
    wx = 0.1 * wl.wavepacket_rhs(g, tag="refckpt-wx")
    uy = wl.wavepacket_rhs(g, tag="refckpt-uy")
    return [dwzdx, uy, wx, wx, dwzdx, wx]


def test_checkpoint_writer_reproduces_reference_bytes(tmp_path):
    rec = _ckpt()
    g = wl.Grid(129, 33, 4, 4, 0.03, 1e-3, 1.03, 0.2, 0.6)
    path = str(tmp_path / "data_00.bin")
    tb.checkpoint_write(path, rec["t"], _fields(g))
    b = open(path, "rb").read()
    assert len(b) == rec["size"]
    assert b[:32].hex() == rec["head"] and b[-32:].hex() == rec["tail"]
    assert hashlib.sha256(b).hexdigest() == rec["sha256"]
    t, back = tb.checkpoint_read(path, g.nx, g.ny, g.nmodes)
    assert t == rec["t"]
    for a, e in zip(back, _fields(g)):
        assert np.array_equal(a, e)


def test_checkpoint_layout_by_hand(tmp_path):
    """Record 1: 4-byte t; record 2: the fields back to back, each
    complex*16 with x fastest, then y, then the mode."""
    rng = np.random.default_rng(3)
    nx, ny, nk = 5, 3, 2
    f = [rng.standard_normal((nk, ny, nx)) + 1j * rng.standard_normal((nk, ny, nx))
         for _ in range(7)]
    path = str(tmp_path / "d.bin")
    tb.checkpoint_write(path, -4, f)
    n = 7 * nk * ny * nx * 16
    want = (np.array([4, -4, 4, n], "<i4").tobytes()
            + b"".join(a.astype("<c16").tobytes() for a in f) + np.int32(n).tobytes())
    assert open(path, "rb").read() == want
    t, back = tb.checkpoint_read(path, nx, ny, nk, nfields=7)
    assert t == -4 and all(np.array_equal(a, e) for a, e in zip(back, f))


def test_baseflow_roundtrip_and_layout(tmp_path):
    rng = np.random.default_rng(4)
    ny, nx = 7, 11
    ux, uy, wz = (rng.standard_normal((ny, nx)) for _ in range(3))
    path = str(tmp_path / "based_00.bin")
    tb.baseflow_write(path, ux, uy, wz)
    n = 3 * ny * nx * 8
    want = (np.int32(n).tobytes() + ux.tobytes() + uy.tobytes() + wz.tobytes()
            + np.int32(n).tobytes())
    assert open(path, "rb").read() == want
    back = tb.baseflow_read(path, nx, ny)
    assert all(np.array_equal(a, e) for a, e in zip(back, (ux, uy, wz)))


def test_readers_refuse_wrong_grid_and_damaged_files(tmp_path):
    rng = np.random.default_rng(5)
    f = [rng.standard_normal((2, 3, 5)).astype(np.complex128) for _ in range(6)]
    path = str(tmp_path / "d.bin")
    tb.checkpoint_write(path, 1, f)
    with pytest.raises(tb.TblError, match="record length"):
        tb.checkpoint_read(path, 5, 4, 2)
    with pytest.raises(tb.TblError, match="record length"):
        tb.checkpoint_read(path, 5, 3, 2, nfields=7)
    b = open(path, "rb").read()
    cut = str(tmp_path / "cut.bin")
    open(cut, "wb").write(b[:-10])
    with pytest.raises(tb.TblError, match="truncated"):
        tb.checkpoint_read(cut, 5, 3, 2)
    bad = str(tmp_path / "bad.bin")
    open(bad, "wb").write(b[:-4] + np.int32(7).tobytes())
    with pytest.raises(tb.TblError, match="trailing marker"):
        tb.checkpoint_read(bad, 5, 3, 2)
    with pytest.raises(tb.TblError, match="cannot open"):
        tb.baseflow_read(str(tmp_path / "missing.bin"), 3, 3)
    with pytest.raises(tb.TblError, match="field count"):
        tb.checkpoint_write(path, 1, f[:5])
