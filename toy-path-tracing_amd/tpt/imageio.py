"""Image output and comparison: the renderer's EXR output
(`save_output` for ".exr" with the rendering transform,
R/output_image.rs:18-39 and write_exr :103-120, an RGB 32-bit float image)
and the relMSE used to compare a render against a reference image.

write_exr writes a single-part scanline OpenEXR file with FLOAT channels
B, G, R (stored in the format's alphabetical channel order) and no
compression.  read_exr reads single-part scanline files with UINT, HALF or
FLOAT channels and NONE, RLE, ZIPS or ZIP compression (the lossless ones a
float image is written with); other compressions and tiled, deep or
multi-part files are refused with a message.

relmse(x, ref) = mean over pixels and channels of (x - r)^2 / (r^2 + 1e-2).

Command line:
    python -m tpt.imageio relmse image.exr reference.exr
    python -m tpt.imageio info image.exr
"""
from __future__ import annotations

import struct
import sys
import zlib
from pathlib import Path

import numpy as np

MAGIC = 20000630
COMPRESSIONS = {0: "NONE", 1: "RLE", 2: "ZIPS", 3: "ZIP", 4: "PIZ", 5: "PXR24", 6: "B44", 7: "B44A", 8: "DWAA",
                9: "DWAB"}
LINES_PER_CHUNK = {0: 1, 1: 1, 2: 1, 3: 16}
PIXEL_TYPES = {0: np.dtype("<u4"), 1: np.dtype("<f2"), 2: np.dtype("<f4")}


class ExrError(ValueError):
    pass


def _attr(name: str, kind: str, value: bytes) -> bytes:
    return name.encode() + b"\0" + kind.encode() + b"\0" + struct.pack("<i", len(value)) + value


def write_exr(path, pixels: np.ndarray) -> None:
    """Writes an (height, width, 3) image as a scanline OpenEXR file of
    FLOAT channels R, G, B (write_exr, R/output_image.rs:103-120)."""
    px = np.asarray(pixels, dtype=np.float32)
    if px.ndim != 3 or px.shape[2] != 3 or px.shape[0] == 0 or px.shape[1] == 0:
        raise ValueError(f"write_exr: expected a non-empty (height, width, 3) image, got {px.shape}")
    h, w, _ = px.shape
    chlist = b"".join(c.encode() + b"\0" + struct.pack("<iB3xii", 2, 0, 1, 1) for c in "BGR") + b"\0"
    window = struct.pack("<iiii", 0, 0, w - 1, h - 1)
    header = (struct.pack("<ii", MAGIC, 2)
              + _attr("channels", "chlist", chlist)
              + _attr("compression", "compression", b"\0")
              + _attr("dataWindow", "box2i", window)
              + _attr("displayWindow", "box2i", window)
              + _attr("lineOrder", "lineOrder", b"\0")
              + _attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
              + _attr("screenWindowCenter", "v2f", struct.pack("<ff", 0.0, 0.0))
              + _attr("screenWindowWidth", "float", struct.pack("<f", 1.0))
              + b"\0")
    row_bytes = 3 * 4 * w
    chunk = 8 + row_bytes
    table_end = len(header) + 8 * h
    offsets = np.arange(h, dtype=np.uint64) * np.uint64(chunk) + np.uint64(table_end)
    # one chunk per scanline: y, byte count, then the row's B, G, R values
    rows = np.empty((h, chunk), np.uint8)
    rows[:, 0:4] = np.arange(h, dtype="<i4").view(np.uint8).reshape(h, 4)
    rows[:, 4:8] = np.frombuffer(struct.pack("<i", row_bytes), np.uint8)
    planar = np.ascontiguousarray(px[:, :, ::-1].transpose(0, 2, 1)).astype("<f4")  # (h, B/G/R, w)
    rows[:, 8:] = planar.view(np.uint8).reshape(h, row_bytes)
    with open(path, "wb") as f:
        f.write(header)
        f.write(offsets.astype("<u8").tobytes())
        f.write(rows.tobytes())


def _read_header(buf: bytes):
    if len(buf) < 8:
        raise ExrError("not an OpenEXR file (too short)")
    magic, version = struct.unpack_from("<ii", buf, 0)
    if magic != MAGIC:
        raise ExrError("not an OpenEXR file (magic)")
    if version & 0xFF != 2:
        raise ExrError(f"OpenEXR version {version & 0xFF} (expected 2)")
    if version & 0x200:
        raise ExrError("tiled OpenEXR files are not read")
    if version & 0x1800:
        raise ExrError("deep or multi-part OpenEXR files are not read")
    at, attrs = 8, {}
    while True:
        end = buf.index(b"\0", at)
        name = buf[at:end].decode()
        at = end + 1
        if not name:
            break
        end = buf.index(b"\0", at)
        kind = buf[at:end].decode()
        (size,) = struct.unpack_from("<i", buf, end + 1)
        at = end + 5
        attrs[name] = (kind, buf[at:at + size])
        at += size
    for need in ("channels", "compression", "dataWindow"):
        if need not in attrs:
            raise ExrError(f"header without {need}")
    channels, raw, c = [], attrs["channels"][1], 0
    while raw[c:c + 1] != b"\0":
        end = raw.index(b"\0", c)
        name = raw[c:end].decode()
        ptype, _linear, xs, ys = struct.unpack_from("<iB3xii", raw, end + 1)
        if ptype not in PIXEL_TYPES:
            raise ExrError(f"channel {name}: pixel type {ptype}")
        if xs != 1 or ys != 1:
            raise ExrError(f"channel {name}: subsampled channels are not read")
        channels.append((name, PIXEL_TYPES[ptype]))
        c = end + 1 + 16
    return attrs, channels, at


def _undo_predictor(data: bytes, size: int) -> bytes:
    """The ZIP / RLE byte transform in reverse: differences to bytes, then the
    two half-streams interleaved back."""
    d = np.frombuffer(data, np.uint8).astype(np.int64)
    if len(d) != size:
        raise ExrError("chunk decompresses to the wrong size")
    d[1:] -= 128  # t[i] = t[i - 1] + d[i] - 128 (mod 256), t[0] = d[0]
    t = (np.cumsum(d) & 0xFF).astype(np.uint8)
    out = np.empty(size, np.uint8)
    half = (size + 1) // 2
    out[0::2] = t[:half]
    out[1::2] = t[half:]
    return out.tobytes()


def _rle_decode(data: bytes, size: int) -> bytes:
    out, i = bytearray(), 0
    while i < len(data) and len(out) < size:
        n = struct.unpack_from("b", data, i)[0]
        i += 1
        if n < 0:
            out += data[i:i - n]
            i -= n
        else:
            out += bytes([data[i]]) * (n + 1)
            i += 1
    return bytes(out)


def read_exr(path) -> np.ndarray:
    """An OpenEXR file's R, G, B channels as an (height, width, 3) float32
    image (a file of one channel Y is returned grey in all three)."""
    buf = Path(path).read_bytes()
    attrs, channels, at = _read_header(buf)
    comp = attrs["compression"][1][0]
    if comp not in LINES_PER_CHUNK:
        raise ExrError(f"compression {COMPRESSIONS.get(comp, comp)} is not read (NONE, RLE, ZIPS, ZIP are)")
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    if w <= 0 or h <= 0:
        raise ExrError("empty data window")
    lines = LINES_PER_CHUNK[comp]
    n_chunks = (h + lines - 1) // lines
    offsets = np.frombuffer(buf, "<u8", n_chunks, at)
    planes = {name: np.empty((h, w), np.float32) for name, _ in channels}
    for off in offsets:
        y, size = struct.unpack_from("<ii", buf, int(off))
        data = buf[int(off) + 8:int(off) + 8 + size]
        first = y - y0
        n_lines = min(lines, h - first)
        if first < 0 or n_lines <= 0 or len(data) != size:
            raise ExrError("chunk outside the data window or cut short")
        raw_size = n_lines * w * sum(dt.itemsize for _, dt in channels)
        if comp != 0 and size < raw_size:
            data = _undo_predictor(zlib.decompress(data) if comp in (2, 3) else _rle_decode(data, raw_size),
                                   raw_size)
        if len(data) != raw_size:
            raise ExrError("chunk of the wrong size")
        pos = 0
        for line in range(n_lines):
            for name, dt in channels:
                n = w * dt.itemsize
                planes[name][first + line] = np.frombuffer(data, dt, w, pos).astype(np.float32)
                pos += n
    if all(c in planes for c in "RGB"):
        return np.stack([planes["R"], planes["G"], planes["B"]], axis=-1)
    if "Y" in planes:
        return np.repeat(planes["Y"][:, :, None], 3, axis=2)
    raise ExrError(f"no R, G, B (or Y) channels: {sorted(planes)}")


def relmse(image: np.ndarray, reference: np.ndarray, eps: float = 1e-2) -> float:
    """Mean over pixels and channels of (x - r)^2 / (r^2 + eps), in float64."""
    x = np.asarray(image, np.float64)
    r = np.asarray(reference, np.float64)
    if x.shape != r.shape:
        raise ValueError(f"relmse: shapes differ: {x.shape} vs {r.shape}")
    return float(np.mean((x - r) ** 2 / (r * r + eps)))


def save_output(path, pixels: np.ndarray) -> None:
    """save_output (R/output_image.rs:18-63) for the rendering transform:
    ".exr" writes the rendering-space pixels; display-referred formats need
    an OCIO display transform, which this package does not carry."""
    ext = Path(path).suffix.lower().lstrip(".")
    if ext == "exr":
        write_exr(path, pixels)
    elif ext in ("png", "webp", "avif"):
        raise ValueError("non-EXR output requires an OCIO display / view transform, which is not carried here")
    else:
        raise ValueError(f"unsupported output extension `{ext}`; expected .exr")


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) == 3 and argv[0] == "relmse":
        print(f"{relmse(read_exr(argv[1]), read_exr(argv[2])):.6e}")
        return 0
    if len(argv) == 2 and argv[0] == "info":
        img = read_exr(argv[1])
        print(f"{img.shape[1]} x {img.shape[0]}, mean {img.mean(axis=(0, 1))}, max {img.max():.6g}")
        return 0
    print(__doc__.split("Command line:")[1].rstrip(), file=sys.stderr)
    return 2


if __name__ == "__main__":
    sys.exit(main())
