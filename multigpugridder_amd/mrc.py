"""MRC (MRC2014) volumes and image stacks, the cryo-EM file format the
reference's callers read volumes and write projections in (``.mrc`` files are
what its Matlab scripts load; ``.gitignore`` of the reference lists them).

Arrays are numpy arrays indexed ``[z, y, x]`` (x fastest), which is both the
MRC data order and this package's array convention, so no transposes are
made.  Little-endian files, modes 0 (int8), 1 (int16), 2 (float32),
6 (uint16) and 12 (float16) are read; float32 (mode 2) is written.
"""
from __future__ import annotations

import os

import numpy as np

_MODES = {0: np.int8, 1: np.int16, 2: np.float32, 6: np.uint16, 12: np.float16}
_HEADER = 1024


class MrcError(ValueError):
    pass


def _header_fields(buf: bytes):
    i4 = np.frombuffer(buf, dtype="<i4", count=256)
    f4 = np.frombuffer(buf, dtype="<f4", count=256)
    return i4, f4


def read_header(path: str) -> dict:
    """Header fields of an MRC file (sizes, mode, voxel size, statistics)."""
    with open(path, "rb") as fh:
        buf = fh.read(_HEADER)
    if len(buf) < _HEADER:
        raise MrcError(f"{path}: shorter than the 1024-byte MRC header")
    i4, f4 = _header_fields(buf)
    if buf[208:212] != b"MAP ":
        raise MrcError(f"{path}: no 'MAP ' identifier at byte 208")
    if buf[212] not in (0x44, 0x41) and buf[212:214] != b"\x00\x00":
        raise MrcError(f"{path}: big-endian MRC files are not supported")
    nx, ny, nz, mode = (int(v) for v in i4[0:4])
    mx, my, mz = (int(v) for v in i4[7:10])
    cella = [float(v) for v in f4[10:13]]
    voxel = cella[0] / mx if mx > 0 else 1.0
    return {"nx": nx, "ny": ny, "nz": nz, "mode": mode, "mx": mx, "my": my, "mz": mz,
            "cella": cella, "voxel_size": voxel, "mapc": int(i4[16]), "mapr": int(i4[17]),
            "maps": int(i4[18]), "dmin": float(f4[19]), "dmax": float(f4[20]), "dmean": float(f4[21]),
            "ispg": int(i4[22]), "nsymbt": int(i4[23]), "origin": [float(v) for v in f4[49:52]],
            "rms": float(f4[54]), "nlabl": int(i4[55])}


def read_mrc(path: str) -> tuple[np.ndarray, float]:
    """(data [nz, ny, nx] as float32, voxel size in Angstrom)."""
    h = read_header(path)
    if h["mode"] not in _MODES:
        raise MrcError(f"{path}: unsupported MRC mode {h['mode']}")
    if (h["mapc"], h["mapr"], h["maps"]) not in ((1, 2, 3), (0, 0, 0)):
        raise MrcError(f"{path}: axis order {h['mapc'], h['mapr'], h['maps']} is not x, y, z")
    dt = np.dtype(_MODES[h["mode"]]).newbyteorder("<")
    count = h["nx"] * h["ny"] * h["nz"]
    offset = _HEADER + h["nsymbt"]
    need = offset + count * dt.itemsize
    if os.path.getsize(path) < need:
        raise MrcError(f"{path}: holds {os.path.getsize(path)} bytes, the header needs {need}")
    data = np.fromfile(path, dtype=dt, count=count, offset=offset)
    return data.reshape(h["nz"], h["ny"], h["nx"]).astype(np.float32), h["voxel_size"]


def write_mrc(path: str, data, voxel_size: float = 1.0, label: str = "multigpugridder-amd") -> None:
    """Writes a float32 MRC2014 file: a volume [nz, ny, nx] or an image stack
    [n, ny, nx] (2-D input is written as one image)."""
    a = np.ascontiguousarray(data, dtype=np.float32)
    if a.ndim == 2:
        a = a[None]
    if a.ndim != 3:
        raise MrcError("write_mrc takes a 2-D image or a 3-D volume / stack")
    nz, ny, nx = a.shape
    i4 = np.zeros(256, dtype="<i4")
    f4 = i4.view("<f4")
    i4[0:4] = (nx, ny, nz, 2)
    i4[7:10] = (nx, ny, nz)
    f4[10:13] = (nx * voxel_size, ny * voxel_size, nz * voxel_size)
    f4[13:16] = (90.0, 90.0, 90.0)
    i4[16:19] = (1, 2, 3)
    a64 = a.astype(np.float64)
    f4[19], f4[20], f4[21] = float(a64.min()), float(a64.max()), float(a64.mean())
    i4[22] = 1 if nz > 1 and nx == ny == nz else 0
    f4[54] = float(a64.std())
    i4[55] = 1
    buf = bytearray(i4.tobytes())
    buf[208:212] = b"MAP "
    buf[212:216] = b"\x44\x44\x00\x00"  # little-endian machine stamp
    lab = label.encode("ascii", "replace")[:80]
    buf[224:224 + len(lab)] = lab
    with open(path, "wb") as fh:
        fh.write(bytes(buf))
        fh.write(a.astype("<f4", copy=False).tobytes())
