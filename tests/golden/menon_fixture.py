"""Reader for menon_native_base_v1.bin, the reference's frozen Menon golden
file (tests/reference/raw/menon_native_base_v1.bin in tylermcm/librawops,
SHA-256 0b753a12...feac), kept here as a data fixture.  Format (its index,
tests/reference/raw/menon_native_base_v1.index.json:2): "LRMENON1", count:u32,
then per record 10*u32 metadata (width, height, stride, pattern, phase x,
phase y, active x, y, w, h), 8*u16 levels (4 black, 4 white), sample count
u32, uint16 sensor rows, float32 RGB of the active area; little endian.
Generated upstream with colour-demosaicing 0.2.6 / colour-science 0.4.6."""
import hashlib
import os
import struct

import numpy as np

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "menon_native_base_v1.bin")
SHA256 = "0b753a122a8bbf7c8e8c609df32324061e67d291078a1802f1a7ec267b78feac"


def records():
    data = open(PATH, "rb").read()
    assert hashlib.sha256(data).hexdigest() == SHA256
    assert data[:8] == b"LRMENON1"
    off = 8
    (count,) = struct.unpack_from("<I", data, off)
    off += 4
    out = []
    for _ in range(count):
        w, h, stride, pat, px, py, ax, ay, aw, ah = struct.unpack_from("<10I", data, off)
        off += 40
        lv = struct.unpack_from("<8H", data, off)
        off += 16
        (n,) = struct.unpack_from("<I", data, off)
        off += 4
        samples = np.frombuffer(data, np.uint16, n, off).copy()
        off += 2 * n
        expected = np.frombuffer(data, np.float32, aw * ah * 3, off).reshape(ah, aw, 3).copy()
        off += 4 * aw * ah * 3
        out.append(dict(width=w, height=h, row_stride_samples=stride, pattern=pat, cfa_phase_x=px,
                        cfa_phase_y=py, active_area=(ax, ay, aw, ah), black_levels=lv[:4],
                        white_levels=lv[4:], samples=samples, expected=expected))
    assert off == len(data)
    return out
