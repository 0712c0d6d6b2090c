"""Output step after PathTracer::render: what the path-tracer app does with the image.

Rendu (src/apps/pathtracer/main.cpp:106-119) tonemaps with 1 - exp(-c) unless the output is a float
format, then saves with Image::Save::SRGB_LDR | IGNORE_ALPHA: PNG bytes min(255, max(0, 255 v)) of
v = pow(c, 1/2.2) (Image::saveAsLDR, src/engine/resources/Image.cpp:126-158), or an OpenEXR file
with half-float B, G, R channels and no compression (Image::saveAsHDR, Image.cpp:160-289). The
float -> byte conversion runs on the GPU (rdu_image_to_ldr); the file writers are host code.
"""
from __future__ import annotations

import ctypes
import struct
import zlib

import numpy as np

from . import _lib

TONEMAP = 0x10
SRGB = 0x20
IGNORE_ALPHA = 0x40

_lib.lib.rdu_image_to_ldr.restype = ctypes.c_int
_lib.lib.rdu_image_to_ldr.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t,
                                      ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]


def to_ldr(image: np.ndarray, tonemap: bool = True, srgb: bool = True, ignore_alpha: bool = True) -> np.ndarray:
    """(h, w[, c]) float image -> uint8 of the same shape, on the GPU."""
    img = np.ascontiguousarray(image, np.float32)
    h, w = img.shape[:2]
    comps = 1 if img.ndim == 2 else img.shape[2]
    out = np.zeros(img.shape, np.uint8)
    flags = (TONEMAP if tonemap else 0) | (SRGB if srgb else 0) | (IGNORE_ALPHA if ignore_alpha else 0)
    _lib.check(_lib.lib.rdu_image_to_ldr(_lib.vptr(img), w, h, comps, _lib.vptr(out), flags, None))
    return out


def to_ldr_device(src_ptr: int, width: int, height: int, components: int, dst_ptr: int, flags: int,
                  stream: int = 0) -> None:
    _lib.check(_lib.lib.rdu_image_to_ldr(ctypes.c_void_p(src_ptr), width, height, components, ctypes.c_void_p(dst_ptr),
                                         flags | _lib.RDU_DEVICE_POINTERS, ctypes.c_void_p(stream)))


def _chunk(tag: bytes, data: bytes) -> bytes:
    return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xffffffff)


def png_bytes(ldr: np.ndarray) -> bytes:
    """8-bit PNG (grey, grey+alpha, RGB or RGBA by channel count), rows top to bottom."""
    a = np.ascontiguousarray(ldr, np.uint8)
    if a.ndim == 2:
        a = a[:, :, None]
    h, w, c = a.shape
    ctype = {1: 0, 2: 4, 3: 2, 4: 6}[c]
    raw = np.concatenate([np.zeros((h, 1), np.uint8), a.reshape(h, w * c)], axis=1).tobytes()  # filter 0 per row
    return (b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, ctype, 0, 0, 0)) +
            _chunk(b"IDAT", zlib.compress(raw, 6)) + _chunk(b"IEND", b""))


def read_png(data: bytes) -> np.ndarray:
    """Reader for the files png_bytes writes (8-bit, non-interlaced, any filter)."""
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, hdr = 8, b"", None
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        tag, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        if zlib.crc32(tag + body) & 0xffffffff != struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0]:
            raise ValueError("bad PNG chunk CRC")
        if tag == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", body)
        elif tag == b"IDAT":
            idat += body
        pos += 12 + n
    w, h, depth, ctype = hdr[:4]
    c = {0: 1, 4: 2, 2: 3, 6: 4}[ctype]
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + w * c)
    out = np.zeros((h, w * c), np.int32)
    for y in range(h):
        f, line = raw[y, 0], raw[y, 1:].astype(np.int32)
        prev = out[y - 1] if y else np.zeros(w * c, np.int32)
        if f == 0:
            out[y] = line
        elif f == 2:
            out[y] = (line + prev) & 255
        else:
            for x in range(w * c):
                left = out[y, x - c] if x >= c else 0
                up, ul = prev[x], (prev[x - c] if x >= c else 0)
                if f == 1:
                    pred = left
                elif f == 3:
                    pred = (left + up) // 2
                else:
                    p = left + up - ul
                    pa, pb, pc = abs(p - left), abs(p - up), abs(p - ul)
                    pred = left if pa <= pb and pa <= pc else (up if pb <= pc else ul)
                out[y, x] = (line[x] + pred) & 255
    return out.astype(np.uint8).reshape(h, w, c).squeeze()


def _attr(name: str, typ: str, value: bytes) -> bytes:
    return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(value)) + value


def exr_bytes(image: np.ndarray, ignore_alpha: bool = True) -> bytes:
    """Scanline OpenEXR, no compression, half channels: B, G, R (plus A for 4 components), as
    Image::saveAsHDR writes them (components 1 and 2 padded to three colour channels)."""
    img = np.asarray(image, np.float32)
    if img.ndim == 2:
        img = img[:, :, None]
    h, w, comps = img.shape
    chans = [img[:, :, j] if j < comps else np.zeros((h, w), np.float32) for j in range(3)]
    names = ["B", "G", "R"]
    planes = {"R": chans[0], "G": chans[1], "B": chans[2]}
    if comps == 4:
        names = ["A"] + names
        planes["A"] = np.ones((h, w), np.float32) if ignore_alpha else img[:, :, 3]
    chlist = b"".join(n.encode() + b"\0" + struct.pack("<iBBBBii", 1, 0, 0, 0, 0, 1, 1) for n in names) + b"\0"
    box = struct.pack("<iiii", 0, 0, w - 1, h - 1)
    header = (b"\x76\x2f\x31\x01" + struct.pack("<i", 2) +
              _attr("channels", "chlist", chlist) + _attr("compression", "compression", b"\0") +
              _attr("dataWindow", "box2i", box) + _attr("displayWindow", "box2i", box) +
              _attr("lineOrder", "lineOrder", b"\0") + _attr("pixelAspectRatio", "float", struct.pack("<f", 1.0)) +
              _attr("screenWindowCenter", "v2f", struct.pack("<ff", 0.0, 0.0)) +
              _attr("screenWindowWidth", "float", struct.pack("<f", 1.0)) + b"\0")
    line_bytes = w * 2 * len(names)
    first = len(header) + 8 * h
    offsets = struct.pack(f"<{h}Q", *[first + y * (8 + line_bytes) for y in range(h)])
    rows = []
    for y in range(h):
        data = b"".join(planes[n][y].astype("<f2").tobytes() for n in names)
        rows.append(struct.pack("<ii", y, len(data)) + data)
    return header + offsets + b"".join(rows)


def read_exr(data: bytes) -> dict:
    """Reader of single-part scanline EXR files without compression and with half channels (what
    exr_bytes writes, and the layout of Rendu's precomputed tables): {channel name: (h, w) float32}."""
    if data[:4] != b"\x76\x2f\x31\x01":
        raise ValueError("not an OpenEXR file")
    pos, attrs = 8, {}
    while data[pos] != 0:
        e = data.index(b"\0", pos)
        name = data[pos:e].decode()
        e2 = data.index(b"\0", e + 1)
        size, = struct.unpack("<i", data[e2 + 1:e2 + 5])
        attrs[name] = data[e2 + 5:e2 + 5 + size]
        pos = e2 + 5 + size
    pos += 1
    if attrs.get("compression", b"\0")[:1] != b"\0":
        raise NotImplementedError("compressed EXR files are not decoded")
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    names, cl = [], attrs["channels"]
    p = 0
    while cl[p] != 0:
        e = cl.index(b"\0", p)
        names.append(cl[p:e].decode())
        if struct.unpack("<i", cl[e + 1:e + 5])[0] != 1:
            raise NotImplementedError("only half-float EXR channels are decoded")
        p = e + 1 + 16
    offsets = struct.unpack(f"<{h}Q", data[pos:pos + 8 * h])
    out = {n: np.zeros((h, w), np.float32) for n in names}
    for y, off in enumerate(offsets):
        yy, size = struct.unpack("<ii", data[off:off + 8])
        row = np.frombuffer(data[off + 8:off + 8 + size], "<f2").reshape(len(names), w)
        for k, n in enumerate(names):
            out[n][yy] = row[k]
    return out


def save(image: np.ndarray, path: str) -> None:
    """The app's save: .exr keeps floats (half); otherwise tonemap + sRGB + 8-bit PNG."""
    if path.lower().endswith(".exr"):
        blob = exr_bytes(image)
    else:
        blob = png_bytes(to_ldr(image, tonemap=True, srgb=True, ignore_alpha=True))
    with open(path, "wb") as f:
        f.write(blob)
