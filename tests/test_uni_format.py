"""The .uni grid files (source/fileio.cpp:38-44, 572-599, 834-999) written
and read by libflof_hip, checked against an independent gzip/struct parser of
the documented layout: "M4T2" + UniHeader (3 dims, gridType, elementType,
bytesPerElement, char[256] info, u64 timestamp = 288 bytes) + int dimT + the
T slices; "MNT2" is the 3D variant without dimT.  No GPU needed."""
import gzip
import struct

import numpy as np
import pytest

from ofblend_amd import manta as M
from ofblend_amd.abi import FLOF_INT, FLOF_REAL, FLOF_VEC4

HDR = struct.Struct("<6i256sQ")
RNG = np.random.default_rng(3)


def parse(path):
    raw = gzip.open(path, "rb").read()
    ident = raw[:4].decode()
    dx, dy, dz, gt, et, bpe, info, ts = HDR.unpack_from(raw, 4)
    off = 4 + HDR.size
    T = 1
    if ident == "M4T2":
        (T,) = struct.unpack_from("<i", raw, off)
        off += 4
    return ident, (dx, dy, dz, T), (gt, et, bpe), raw[off:]


def write_ref(path, ident, dims, types, payload):
    dx, dy, dz, T = dims
    gt, et, bpe = types
    with gzip.open(path, "wb", compresslevel=1) as f:
        f.write(ident.encode())
        f.write(HDR.pack(dx, dy, dz, gt, et, bpe, b"mantaflow test", 0))
        if ident == "M4T2":
            f.write(struct.pack("<i", T))
        f.write(payload)


def test_header_size():
    assert HDR.size == 288


@pytest.mark.parametrize("elem,types", [(FLOF_REAL, (1, 1, 4)), (FLOF_VEC4, (8, 2, 16)),
                                        (FLOF_INT, (2, 0, 4))])
def test_library_writes_reference_layout(tmp_path, elem, types):
    shape = (3, 4, 5, 6) + ((4,) if elem == FLOF_VEC4 else ())
    a = RNG.standard_normal(shape).astype(np.float32)
    if elem == FLOF_INT:
        a = RNG.integers(-9, 9, shape).astype(np.int32)
    path = tmp_path / "g.uni"
    M.writeUni4d(path, a, elem)
    ident, dims, ty, payload = parse(path)
    assert ident == "M4T2" and dims == (6, 5, 4, 3) and ty == types
    assert np.array_equal(np.frombuffer(payload, dtype=a.dtype).reshape(shape), a)
    assert M.getUniFileSize(path) == (6, 5, 4, 3)
    assert np.array_equal(M.readUni(path, (6, 5, 4, 3), elem), a)


def test_library_reads_reference_files(tmp_path):
    # 4D Real written as the reference does
    a = RNG.standard_normal((4, 5, 6, 7)).astype(np.float32)
    p4 = tmp_path / "phi.uni"
    write_ref(p4, "M4T2", (7, 6, 5, 4), (1, 1, 4), a.tobytes())
    assert np.array_equal(M.readUni(p4, (7, 6, 5, 4), FLOF_REAL), a)
    # 3D level-set frame (gridType Real|Levelset = 17), as the frame sequences
    f = RNG.standard_normal((5, 6, 7)).astype(np.float32)
    p3 = tmp_path / "frame_0000.uni"
    write_ref(p3, "MNT2", (7, 6, 5, 1), (17, 1, 4), f.tobytes())
    assert M.getUniFileSize(p3) == (7, 6, 5, 1)
    assert np.array_equal(M.readUni(p3, (7, 6, 5, 1), FLOF_REAL)[0], f)


def test_mismatches_are_errors(tmp_path):
    a = RNG.standard_normal((2, 3, 4, 5)).astype(np.float32)
    p = tmp_path / "g.uni"
    M.writeUni4d(p, a)
    with pytest.raises(RuntimeError):
        M.readUni(p, (5, 4, 3, 3), FLOF_REAL)      # wrong T
    with pytest.raises(RuntimeError):
        M.readUni(p, (5, 4, 3, 2), FLOF_VEC4)      # wrong element type
    bad = tmp_path / "bad.uni"
    with gzip.open(bad, "wb") as f:
        f.write(b"XXXX" + bytes(400))
    with pytest.raises(RuntimeError):
        M.getUniFileSize(bad)
    with pytest.raises(RuntimeError):
        M.readUni(tmp_path / "missing.uni", (5, 4, 3, 2))
