"""PFM and CVTRaw files written and read by the library (host code, no GPU).

Pins: the byte layout the reference's PFMSaver writes (header lines, rows
bottom to top, alpha dropped, little-endian scale -1.0); CVTRaw's 16-byte
header and rows at the file's stride, read back at another stride; the
reference's own CVTRaw data file, when the reference tree is present."""
import os
import struct

import numpy as np
import pytest


def test_pfm_gray_and_rgba_bytes(tmp_path):
    from cvt_amd import formats
    rng = np.random.default_rng(0)
    g = rng.random((3, 5)).astype(np.float32)
    p = str(tmp_path / "g.pfm")
    formats.save_pfm(p, g)
    b = open(p, "rb").read()
    head = b"Pf\n5 3\n-1.0\n"
    assert b[:len(head)] == head
    assert b[len(head):] == g[::-1].astype("<f4").tobytes()
    c = rng.random((4, 6, 4)).astype(np.float32)
    formats.save_pfm(p, c)
    b = open(p, "rb").read()
    head = b"PF\n6 4\n-1.0\n"
    assert b[:len(head)] == head
    assert b[len(head):] == np.ascontiguousarray(c[::-1, :, :3]).astype("<f4").tobytes()


@pytest.mark.parametrize("shape", [(7, 9), (5, 3, 4)])
def test_cvtraw_round_trip_and_header(tmp_path, shape):
    from cvt_amd import formats
    a = np.random.default_rng(1).random(shape).astype(np.float32)
    p = str(tmp_path / "a.cvtraw")
    formats.save_cvtraw(p, a)
    b = open(p, "rb").read()
    w, h, st, fid = struct.unpack("<4I", b[:16])
    nc = 1 if a.ndim == 2 else 4
    assert (w, h, st, fid) == (shape[1], shape[0], shape[1] * 4 * nc, 4 if nc == 1 else 12)
    assert b[16:] == a.tobytes()
    back, fid2 = formats.load_cvtraw(p)
    assert fid2 == fid and np.array_equal(back, a)


def test_cvtraw_reads_a_padded_stride(tmp_path):
    from cvt_amd import formats
    a = np.arange(12, dtype=np.float32).reshape(3, 4)
    pad = np.zeros((3, 7), np.float32)
    pad[:, :4] = a
    p = str(tmp_path / "pad.cvtraw")
    with open(p, "wb") as f:
        f.write(struct.pack("<4I", 4, 3, 28, 4) + pad.tobytes())
    back, fid = formats.load_cvtraw(p)
    assert fid == 4 and np.array_equal(back, a)
    with open(p, "wb") as f:   # truncated file
        f.write(struct.pack("<4I", 4, 3, 28, 4) + pad.tobytes()[:40])
    from cvt_amd import _lib
    with pytest.raises(_lib.PmhError, match="shorter"):
        formats.load_cvtraw(p)


REF_RAW = "/root/reference/data/gbrg.cvtraw"


@pytest.mark.skipif(not os.path.exists(REF_RAW), reason="reference tree not present")
def test_reference_cvtraw_data_file():
    from cvt_amd import formats
    assert formats.cvtraw_info(REF_RAW) == (1296, 964, 1296, 19)   # BAYER_GBRG_UINT8
    img, fid = formats.load_cvtraw(REF_RAW)
    raw = open(REF_RAW, "rb").read()[16:]
    assert fid == 19 and img.shape == (964, 1296)
    assert img.tobytes() == raw
