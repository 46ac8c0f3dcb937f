"""Raw files on the CPU: engine.RawSource / RawSink / RawFile semantics
(engine/raw.go, engine/rawfile.go), which the device path must match.

- Validity from the fill value compares as floats: -0 matches a fill of 0,
  and a NaN fill matches every NaN (internal/vec/validity.go:11-19).
- A source without a fill sets every bit it is given.
- A sink with a fill writes it into every invalid cell, in the source
  raster too (fillInvalid).
- Panics and errors carry the reference's messages.
"""
import os

import numpy as np
import pytest

from strata_amd import engine
from strata_amd.engine import (EnginePanic, OpenRawFile, CreateRawFile, RawOptions, RawSink, RawSource,
                               SourceError)
from strata_amd.raster import Float32Raster, unpack_bits


def write_file(path, vals):
    np.asarray(vals, dtype="<f4").tofile(path)


def host_raster(w, h, stride=None, masked=True, off=0):
    stride = stride or w
    n = (h - 1) * stride + w
    data = np.full(n, 7.0, np.float32)
    valid = np.zeros((off + n + 63) // 64 + 1, np.uint64) if masked else None
    return Float32Raster(data, w, h, stride, valid, off)


def bits_of(r):
    out = np.zeros((r.Height, r.Width), bool)
    for y in range(r.Height):
        out[y] = unpack_bits(r.Valid, r.ValidOffset + y * r.Stride, r.Width)
    return out


def cells_of(r):
    return np.stack([r.Data[y * r.Stride:y * r.Stride + r.Width] for y in range(r.Height)])


@pytest.fixture
def tmpfile(tmp_path):
    return str(tmp_path / "f.raw")


def test_fill_zero_matches_negative_zero(tmpfile):
    W, H = 9, 4
    vals = np.arange(W * H, dtype=np.float32)
    vals[[3, 10, 20]] = [0.0, -0.0, np.nan]
    write_file(tmpfile, vals)
    f = OpenRawFile(tmpfile, os.O_RDONLY)
    src = RawSource(f, W, H, RawOptions(Fill=0.0, HasFill=True))
    assert src.Masked() and src.Size() == (W, H)
    dst = host_raster(5, 3, stride=70, off=37)
    assert src.ReadWindow(None, dst, 2, 1) is None
    want = vals.reshape(H, W)[1:4, 2:7]
    got = cells_of(dst)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    wv = ~((want == 0) & ~np.isnan(want))  # 0 and -0 invalid; NaN is a value
    assert np.array_equal(bits_of(dst), wv)
    f.Close()


def test_nan_fill_matches_every_nan(tmpfile):
    vals = np.array([1, np.nan, -np.nan, 0, np.inf, 2], np.float32)
    vals[2] = np.uint32(0xFFC00123).view(np.float32)  # a NaN with a payload and the sign bit
    write_file(tmpfile, vals)
    with open(tmpfile, "rb") as fh:
        src = RawSource(fh, 6, 1, RawOptions(Fill=float("nan"), HasFill=True))
        dst = host_raster(6, 1)
        assert src.ReadWindow(None, dst, 0, 0) is None
        assert bits_of(dst).tolist() == [[True, False, False, True, True, True]]


def test_no_fill_sets_every_bit(tmpfile):
    write_file(tmpfile, np.zeros(12, np.float32))
    f = OpenRawFile(tmpfile, os.O_RDONLY)
    src = RawSource(f, 4, 3)
    assert not src.Masked()
    dst = host_raster(4, 3)
    assert src.ReadWindow(None, dst, 0, 0) is None
    assert bits_of(dst).all()
    dst2 = host_raster(4, 3, masked=False)
    assert src.ReadWindow(None, dst2, 0, 0) is None
    f.Close()


def test_short_file_is_an_error(tmpfile):
    write_file(tmpfile, np.zeros(4 * 3 - 1, np.float32))
    f = OpenRawFile(tmpfile, os.O_RDONLY)
    err = RawSource(f, 4, 3).ReadWindow(None, host_raster(4, 3, masked=False), 0, 0)
    assert isinstance(err, SourceError) and str(err) == "engine: raw source: reading row 2: unexpected EOF"
    f.Close()


def test_sink_fills_invalid_cells(tmpfile):
    f = CreateRawFile(tmpfile, 4 * 6 * 3)
    sink = RawSink(f, 6, 3, RawOptions(Fill=-9999.0, HasFill=True))
    src = host_raster(4, 2, stride=5, off=3)
    src.Data[:] = np.arange(len(src.Data), dtype=np.float32)
    from strata_amd.engine import _put_bits
    _put_bits(src.Valid, 3, np.array([1, 0, 1, 1], bool))
    _put_bits(src.Valid, 3 + 5, np.array([0, 1, 1, 0], bool))
    assert sink.WriteWindow(None, src, 1, 1) is None
    f.Close()
    got = np.fromfile(tmpfile, dtype="<f4").reshape(3, 6)
    assert got[0].tolist() == [0.0] * 6 and got[1, 0] == 0 and got[1, 5] == 0
    assert got[1, 1:5].tolist() == [0.0, -9999.0, 2.0, 3.0]
    assert got[2, 1:5].tolist() == [-9999.0, 6.0, 7.0, -9999.0]
    # fillInvalid writes into the source raster itself.
    assert src.Data[1] == -9999.0 and src.Data[5] == -9999.0


def test_panics_and_errors(tmpfile):
    write_file(tmpfile, np.zeros(16, np.float32))
    f = OpenRawFile(tmpfile, os.O_RDWR)
    with pytest.raises(EnginePanic, match="^engine: NewRawSource: nil file$"):
        RawSource(None, 4, 4)
    with pytest.raises(EnginePanic, match="^engine: NewRawSink: dimensions must be positive, got 0×4$"):
        RawSink(f, 0, 4)
    with pytest.raises(EnginePanic, match="outside 4×4 raster"):
        RawSource(f, 4, 4).ReadWindow(None, host_raster(3, 3), 2, 0)
    with pytest.raises(EnginePanic, match="the source has a fill value and dst has no validity mask"):
        RawSource(f, 4, 4, RawOptions(HasFill=True)).ReadWindow(None, host_raster(2, 2, masked=False), 0, 0)
    with pytest.raises(EnginePanic, match="src has a validity mask and the sink has no fill value"):
        RawSink(f, 4, 4).WriteWindow(None, host_raster(2, 2), 0, 0)
    with pytest.raises(OSError, match="O_APPEND does not allow WriteAt"):
        OpenRawFile(tmpfile, os.O_WRONLY | os.O_APPEND)
    f.Close()


def test_raw_file_handles(tmpfile):
    f = CreateRawFile(tmpfile, 64, handles=3)
    assert f.Handles() == 3 and os.path.getsize(tmpfile) == 64
    assert f.WriteAt(b"abcd", 8) == 4
    assert f.ReadAt(4, 8) == b"abcd"
    f.Sync()
    f.Close()


def test_library_descriptor(tmpfile):
    """A raw source over a file with a descriptor reaches the library as a
    strata_raw, so the library reads it itself."""
    write_file(tmpfile, np.zeros(4, np.float32))
    f = OpenRawFile(tmpfile, os.O_RDONLY)
    r = engine._raw(RawSource(f, 2, 2, RawOptions(Fill=-1.0, HasFill=True)))
    assert r.fd == f.fileno() and r.has_fill == 1 and r.fill == -1.0 and r.offset == 0
    assert engine._raw(engine.MemorySource(host_raster(2, 2))) is None
    f.Close()
