"""MRC2014 I/O: round trips, header fields against the format's word layout,
other storage modes, and refusal of malformed files."""
import numpy as np
import pytest

from multigpugridder_amd import mrc


def test_volume_round_trip_and_header(tmp_path):
    vol = np.random.default_rng(1).standard_normal((12, 10, 8)).astype(np.float32)
    p = str(tmp_path / "v.mrc")
    mrc.write_mrc(p, vol, voxel_size=1.5)
    back, vs = mrc.read_mrc(p)
    assert back.shape == (12, 10, 8) and np.array_equal(back, vol) and vs == 1.5
    raw = open(p, "rb").read()
    assert len(raw) == 1024 + vol.size * 4
    words = np.frombuffer(raw[:1024], dtype="<i4")
    assert tuple(words[0:4]) == (8, 10, 12, 2)          # NX NY NZ MODE, x fastest
    assert raw[208:212] == b"MAP " and raw[212:214] == b"\x44\x44"
    h = mrc.read_header(p)
    assert (h["mapc"], h["mapr"], h["maps"]) == (1, 2, 3)
    assert abs(h["dmean"] - float(vol.mean())) < 1e-6 and abs(h["rms"] - float(vol.std())) < 1e-5
    assert h["dmin"] == float(vol.min()) and h["dmax"] == float(vol.max())
    # data order: element [z, y, x] sits at byte 1024 + 4 * (x + NX * (y + NY * z))
    z, y, x = 3, 7, 5
    off = 1024 + 4 * (x + 8 * (y + 10 * z))
    assert np.frombuffer(raw[off:off + 4], dtype="<f4")[0] == vol[z, y, x]


def test_image_stack_and_single_image(tmp_path):
    stack = np.arange(3 * 4 * 5, dtype=np.float32).reshape(3, 4, 5)
    p = str(tmp_path / "s.mrc")
    mrc.write_mrc(p, stack)
    assert np.array_equal(mrc.read_mrc(p)[0], stack)
    mrc.write_mrc(p, stack[0])
    assert mrc.read_mrc(p)[0].shape == (1, 4, 5)


@pytest.mark.parametrize("mode,dtype", [(0, np.int8), (1, np.int16), (6, np.uint16), (12, np.float16)])
def test_other_modes_read(tmp_path, mode, dtype):
    data = (np.arange(2 * 3 * 4) % 100).astype(dtype).reshape(2, 3, 4)
    p = str(tmp_path / "m.mrc")
    mrc.write_mrc(p, np.zeros((2, 3, 4), np.float32))
    raw = bytearray(open(p, "rb").read()[:1024])
    raw[12:16] = np.int32(mode).tobytes()
    open(p, "wb").write(bytes(raw) + data.astype(np.dtype(dtype).newbyteorder("<")).tobytes())
    got, _ = mrc.read_mrc(p)
    assert got.dtype == np.float32 and np.array_equal(got, data.astype(np.float32))


def test_malformed_files_are_refused(tmp_path):
    p = str(tmp_path / "bad.mrc")
    mrc.write_mrc(p, np.zeros((2, 2, 2), np.float32))
    raw = bytearray(open(p, "rb").read())
    open(p, "wb").write(bytes(raw[:-4]))                 # truncated data
    with pytest.raises(mrc.MrcError):
        mrc.read_mrc(p)
    raw2 = bytearray(raw)
    raw2[208:212] = b"XXXX"
    open(p, "wb").write(bytes(raw2))
    with pytest.raises(mrc.MrcError):
        mrc.read_mrc(p)
    raw3 = bytearray(raw)
    raw3[12:16] = np.int32(4).tobytes()                   # complex mode: not read
    open(p, "wb").write(bytes(raw3))
    with pytest.raises(mrc.MrcError):
        mrc.read_mrc(p)
    open(p, "wb").write(b"short")
    with pytest.raises(mrc.MrcError):
        mrc.read_header(p)
