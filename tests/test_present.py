"""Screenshot files (rt_image_write): the PNG is decoded here with Python's
own zlib and struct (chunk CRCs, zlib stream, filter bytes) and the PPM by
its header, and both must give back the exact pixels written."""
import struct
import zlib

import numpy as np
import pytest


def decode_png(data: bytes) -> np.ndarray:
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, hdr = 8, b"", None
    while pos < len(data):
        n, typ = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(typ + body) & 0xFFFFFFFF, typ
        if typ == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", body)
        elif typ == b"IDAT":
            idat += body
        elif typ == b"IEND":
            break
        pos += 12 + n
    w, h, depth, ctype, comp, filt, inter = hdr
    assert (depth, ctype, comp, filt, inter) == (8, 6, 0, 0, 0)
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 4 * w + 1)
    assert (raw[:, 0] == 0).all()  # filter type None on every row
    return raw[:, 1:].reshape(h, w, 4)


@pytest.mark.parametrize("w,h", [(1, 1), (37, 11), (300, 260)])
def test_png_and_ppm_round_trip(rtcore, tmp_path, w, h):
    img = np.random.default_rng(w * h).integers(0, 256, (h, w, 4), dtype=np.uint8)
    rtcore.image_write(str(tmp_path / "a.png"), img)
    assert np.array_equal(decode_png((tmp_path / "a.png").read_bytes()), img)
    rtcore.image_write(str(tmp_path / "a.ppm"), img)
    data = (tmp_path / "a.ppm").read_bytes()
    head = f"P6\n{w} {h}\n255\n".encode()
    assert data.startswith(head)
    assert np.array_equal(np.frombuffer(data[len(head):], np.uint8).reshape(h, w, 3), img[:, :, :3])


def test_image_write_errors(rtcore, tmp_path):
    img = np.zeros((2, 2, 4), np.uint8)
    with pytest.raises(rtcore.RtError, match="png or .ppm"):
        rtcore.image_write(str(tmp_path / "a.jpg"), img)
    with pytest.raises(rtcore.RtError, match="cannot open"):
        rtcore.image_write(str(tmp_path / "missing" / "a.png"), img)
