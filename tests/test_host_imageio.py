"""tpt.imageio: the EXR output of the renderer (write_exr, an RGB 32-bit
float scanline OpenEXR file), the EXR reader (NONE, RLE, ZIPS, ZIP; UINT,
HALF, FLOAT channels) and relMSE.

The written bytes are checked against the OpenEXR layout (magic, version,
attribute list, offset table, one chunk per scanline with the channels in
B, G, R order), and compressed files are built here by the format's
documented encoder steps (bytes split into even / odd halves, each byte
replaced by its difference to the previous + 128, then zlib or the RLE
runs) and read back."""
import struct
import zlib

import numpy as np
import pytest

from tpt import imageio


def _image(h=5, w=7, seed=0):
    rng = np.random.default_rng(seed)
    img = rng.normal(0.0, 10.0, (h, w, 3)).astype(np.float32)
    img[0, 0] = [np.inf, -0.0, 1e-40]  # non-finite, signed zero, subnormal survive
    return img


def test_round_trip_is_exact(tmp_path):
    img = _image()
    p = tmp_path / "a.exr"
    imageio.write_exr(p, img)
    back = imageio.read_exr(p)
    assert back.dtype == np.float32
    assert back.tobytes() == img.tobytes()


def test_written_bytes_follow_the_openexr_layout(tmp_path):
    img = _image(3, 4)
    p = tmp_path / "a.exr"
    imageio.write_exr(p, img)
    b = p.read_bytes()
    assert b[:4] == bytes([0x76, 0x2F, 0x31, 0x01])
    assert struct.unpack_from("<i", b, 4)[0] == 2  # version 2, single-part scanline
    names = [n.encode() for n in ("channels", "compression", "dataWindow", "displayWindow", "lineOrder",
                                  "pixelAspectRatio", "screenWindowCenter", "screenWindowWidth")]
    pos = [b.index(n + b"\0") for n in names]
    assert pos == sorted(pos)
    chl = b.index(b"chlist\0") + 7 + 4
    assert b[chl:chl + 2] == b"B\0" and struct.unpack_from("<i", b, chl + 2)[0] == 2  # FLOAT
    assert b[chl + 18:chl + 20] == b"G\0" and b[chl + 36:chl + 38] == b"R\0"
    comp = b.index(b"compression\0compression\0") + 24
    assert struct.unpack_from("<i", b, comp)[0] == 1 and b[comp + 4] == 0  # NONE
    dw = b.index(b"dataWindow\0box2i\0") + 17
    assert struct.unpack_from("<iiiii", b, dw) == (16, 0, 0, 3, 2)
    end = b.index(b"screenWindowWidth\0float\0") + 24 + 4 + 4 + 1  # name, type, size, value, terminator
    assert b[end - 1] == 0
    offsets = struct.unpack_from("<3Q", b, end)
    row = 3 * 4 * 4
    assert offsets == tuple(end + 24 + k * (8 + row) for k in range(3))
    for y, off in enumerate(offsets):
        assert struct.unpack_from("<ii", b, off) == (y, row)
        vals = np.frombuffer(b, "<f4", 12, off + 8).reshape(3, 4)
        np.testing.assert_array_equal(vals, img[y].T[::-1])  # B, G, R planes of the row


def _predict(raw: bytes) -> bytes:
    """The ZIP / RLE encoder transform: interleaved halves, then differences."""
    r = np.frombuffer(raw, np.uint8)
    t = np.concatenate([r[0::2], r[1::2]]).astype(np.int64)
    d = t.copy()
    d[1:] = (t[1:] - t[:-1] + 128) & 0xFF
    return d.astype(np.uint8).tobytes()


def _rle(data: bytes) -> bytes:
    out, i = bytearray(), 0
    while i < len(data):
        j = i
        while j + 1 < len(data) and data[j + 1] == data[i] and j - i < 126:
            j += 1
        if j - i >= 2:  # a run of j - i + 1 equal bytes
            out += struct.pack("b", j - i) + data[i:i + 1]
            i = j + 1
        else:  # literals up to the next run
            k = i
            while k < len(data) and k - i < 127 and not (k + 2 < len(data) and data[k] == data[k + 1] == data[k + 2]):
                k += 1
            k = max(k, i + 1)
            out += struct.pack("b", -(k - i)) + data[i:k]
            i = k
    return bytes(out)


def _exr(h, w, chans, comp, encode, lines):
    """A scanline EXR whose chunks are `encode(raw chunk bytes)`."""
    chlist = b"".join(n.encode() + b"\0" + struct.pack("<iB3xii", t, 0, 1, 1) for n, t, _ in chans) + b"\0"
    win = struct.pack("<iiii", 0, 0, w - 1, h - 1)
    hdr = struct.pack("<ii", 20000630, 2)
    for name, kind, val in [("channels", "chlist", chlist), ("compression", "compression", bytes([comp])),
                            ("dataWindow", "box2i", win), ("displayWindow", "box2i", win),
                            ("lineOrder", "lineOrder", b"\0"), ("pixelAspectRatio", "float", struct.pack("<f", 1)),
                            ("screenWindowCenter", "v2f", struct.pack("<ff", 0, 0)),
                            ("screenWindowWidth", "float", struct.pack("<f", 1))]:
        hdr += name.encode() + b"\0" + kind.encode() + b"\0" + struct.pack("<i", len(val)) + val
    hdr += b"\0"
    chunks = []
    for y in range(0, h, lines):
        raw = b"".join(arr[yy].tobytes() for yy in range(y, min(h, y + lines)) for _, _, arr in chans)
        enc = encode(raw)
        if comp != 0 and len(enc) >= len(raw):  # the format stores what does not compress as it is
            enc = raw
        chunks.append(struct.pack("<ii", y, len(enc)) + enc)
    at = len(hdr) + 8 * len(chunks)
    offs = []
    for c in chunks:
        offs.append(at)
        at += len(c)
    return hdr + struct.pack(f"<{len(offs)}Q", *offs) + b"".join(chunks)


def _channels(h, w, dtype, ptype, seed=1):
    rng = np.random.default_rng(seed)
    # smooth rows (compressible) with some noise
    base = np.linspace(0, 4, w, dtype=np.float32)[None, :] + rng.normal(0, 0.01, (h, w)).astype(np.float32)
    arrs = {c: (base * (k + 1)).astype(dtype) for k, c in enumerate("BGR")}
    return [(c, ptype, arrs[c]) for c in "BGR"], arrs


@pytest.mark.parametrize("comp,lines", [(3, 16), (2, 1)])
def test_zip_compressed_half_and_float_files_read(tmp_path, comp, lines):
    for dtype, ptype in ((np.float16, 1), (np.float32, 2)):
        chans, arrs = _channels(37, 29, dtype, ptype)
        p = tmp_path / f"z{comp}{ptype}.exr"
        p.write_bytes(_exr(37, 29, chans, comp, lambda raw: zlib.compress(_predict(raw)), lines))
        img = imageio.read_exr(p)
        for k, c in enumerate("RGB"):
            np.testing.assert_array_equal(img[:, :, k], arrs[c].astype(np.float32))


def test_rle_and_stored_chunks_read(tmp_path):
    # piecewise-constant rows (runs after the byte transform) and noisy ones
    # (stored: they do not compress)
    chans, arrs = _channels(9, 40, np.float32, 2)
    blocks = np.repeat(np.arange(5, dtype=np.float32), 8)[None, :] * np.arange(1, 10, dtype=np.float32)[:, None]
    for k, c in enumerate("BGR"):
        arrs[c][:5] = blocks[:5] * (k + 1)
    p = tmp_path / "rle.exr"
    data = _exr(9, 40, chans, 1, lambda raw: _rle(_predict(raw)), 1)
    assert len(data) < len(_exr(9, 40, chans, 0, lambda raw: raw, 1))  # some chunks are compressed
    p.write_bytes(data)
    img = imageio.read_exr(p)
    for k, c in enumerate("RGB"):
        np.testing.assert_array_equal(img[:, :, k], arrs[c])
    # a chunk that does not compress is stored as it is (size == raw size)
    p2 = tmp_path / "stored.exr"
    p2.write_bytes(_exr(9, 40, chans, 3, lambda raw: raw, 16))
    np.testing.assert_array_equal(imageio.read_exr(p2)[:, :, 2], arrs["B"])


def test_uint_and_luminance_channels(tmp_path):
    y = np.arange(12, dtype="<u4").reshape(3, 4)
    p = tmp_path / "y.exr"
    p.write_bytes(_exr(3, 4, [("Y", 0, y)], 0, lambda raw: raw, 1))
    img = imageio.read_exr(p)
    np.testing.assert_array_equal(img, np.repeat(y[:, :, None].astype(np.float32), 3, axis=2))


def test_unsupported_files_are_refused(tmp_path):
    chans, _ = _channels(4, 4, np.float32, 2)
    good = _exr(4, 4, chans, 0, lambda raw: raw, 1)
    cases = {
        "piz": (_exr(4, 4, chans, 4, lambda raw: raw, 32), "PIZ"),
        "tiled": (good[:4] + struct.pack("<i", 2 | 0x200) + good[8:], "tiled"),
        "multipart": (good[:4] + struct.pack("<i", 2 | 0x1000) + good[8:], "multi-part"),
        "magic": (b"\0" * 4 + good[4:], "magic"),
        "short": (good[:5], "too short"),
    }
    for name, (data, msg) in cases.items():
        p = tmp_path / f"{name}.exr"
        p.write_bytes(data)
        with pytest.raises(imageio.ExrError, match=msg):
            imageio.read_exr(p)


def test_relmse():
    r = np.zeros((2, 2, 3))
    assert imageio.relmse(r, r) == 0.0
    assert imageio.relmse(r + 0.1, r) == pytest.approx(1.0)  # 0.01 / (0 + 0.01)
    r2 = np.full((1, 1, 3), 2.0)
    assert imageio.relmse(r2 * 1.5, r2) == pytest.approx(1.0 / 4.01)
    with pytest.raises(ValueError):
        imageio.relmse(np.zeros((2, 2, 3)), np.zeros((2, 3, 3)))


def test_save_output_and_the_command_line(tmp_path, capsys):
    a, b = _image(4, 4, 1), _image(4, 4, 2)
    imageio.save_output(tmp_path / "a.EXR", a)
    imageio.save_output(tmp_path / "b.exr", b)
    with pytest.raises(ValueError, match="OCIO"):
        imageio.save_output(tmp_path / "a.png", a)
    with pytest.raises(ValueError, match="unsupported"):
        imageio.save_output(tmp_path / "a.tif", a)
    a[0, 0] = b[0, 0] = 0.0  # finite inputs for the printed value
    imageio.write_exr(tmp_path / "a.exr", a)
    imageio.write_exr(tmp_path / "b.exr", b)
    assert imageio.main(["relmse", str(tmp_path / "a.exr"), str(tmp_path / "b.exr")]) == 0
    assert float(capsys.readouterr().out) == pytest.approx(imageio.relmse(a, b), rel=1e-6)
    assert imageio.main(["bogus"]) == 2
