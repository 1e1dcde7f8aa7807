"""Minimal RGBA8 PNG writer (zlib + CRC chunks); the reference writes its PNG with libattopng
(Render/src/ofApp.cpp:110-134)."""
import struct
import zlib

import numpy as np


def write_png(path, rgba, width, height):
    rgba = np.ascontiguousarray(np.asarray(rgba, np.uint8).reshape(height, width, 4))
    raw = b"".join(b"\x00" + rgba[y].tobytes() for y in range(height))

    def chunk(tag, data):
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)

    with open(path, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n")
        f.write(chunk(b"IHDR", struct.pack(">IIBBBBB", width, height, 8, 6, 0, 0, 0)))
        f.write(chunk(b"IDAT", zlib.compress(raw, 6)))
        f.write(chunk(b"IEND", b""))
