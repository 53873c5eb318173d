"""Writes camera .movie files (the container of src/video_reader.hpp:13-214)
for tests: every frame is a little-endian header followed by length_data
bytes of pixels."""
import struct

import numpy as np

MAGIC = 0x496D6554
HEADER_LEN = {1: 172, 2: 128, 3: 240}  # IIDC, Andor, Ximea


def frame_header(camera, pixelmode, w, h, data_len, t_us, version=1, magic=MAGIC):
    hl = HEADER_LEN[camera]
    b = bytearray(hl)
    struct.pack_into("<6I", b, 0, magic, version, camera, pixelmode, hl, data_len)
    if camera == 1:  # IIDC: timestamp (us) at 48, size at 64/68
        struct.pack_into("<Q", b, 48, int(t_us))
        struct.pack_into("<II", b, 64, w, h)
    elif camera == 2:  # Andor: sec/nsec at 24/32, x/y start/end/bin at 48..68
        sec, rem = divmod(int(t_us), 1_000_000)
        struct.pack_into("<QQ", b, 24, sec, rem * 1000)
        struct.pack_into("<6i", b, 48, 0, w - 1, 0, h - 1, 1, 1)
    else:  # Ximea: sec/nsec at 128/136, size at 152/156
        sec, rem = divmod(int(t_us), 1_000_000)
        struct.pack_into("<QQ", b, 128, sec, rem * 1000)
        struct.pack_into("<II", b, 152, w, h)
    return bytes(b)


def write_movie(path, frames, camera=1, pixelmode=1, fps=50.0, junk_prefix=b"", pad=0, truncate=0):
    """frames: [n, h, w] uint8 (pixelmode 1) or uint16 (2 = big endian, 3 = little endian)."""
    n, h, w = frames.shape
    if pixelmode == 1:
        data = [np.ascontiguousarray(f, dtype=np.uint8).tobytes() for f in frames]
    elif pixelmode == 2:
        data = [np.ascontiguousarray(f, dtype=">u2").tobytes() for f in frames]
    else:
        data = [np.ascontiguousarray(f, dtype="<u2").tobytes() for f in frames]
    dl = len(data[0]) + pad
    out = bytearray(junk_prefix)
    for i, d in enumerate(data):
        out += frame_header(camera, pixelmode, w, h, dl, i * 1e6 / fps)
        out += d + b"\0" * pad
    if truncate:
        out = out[:-truncate]
    with open(path, "wb") as f:
        f.write(bytes(out))
