"""Image files for scene ingestion and output: PNG and OpenEXR readers, an
OpenEXR writer. Stand-in for the SimpleImageIO calls the reference makes
(`new RgbImage(filename)`, `Image.WriteToFile`), with numpy and zlib only.

PNG: 8-bit and 16-bit grey / grey+alpha / RGB / RGBA, non-interlaced, the five
row filters of the PNG specification; values are decoded from sRGB to linear
(LDR textures are colour data). OpenEXR: scanline files with NONE, ZIPS or ZIP
compression and HALF / FLOAT channels (PIZ and the other wavelet / lossy
codecs raise an error). Images are returned as float32 arrays of shape
(height, width, channels) with channels R, G, B (A dropped).
"""
from __future__ import annotations

import struct
import zlib

import numpy as np

# ----------------------------------------------------------------------- PNG --
_PNG_MAGIC = b"\x89PNG\r\n\x1a\n"


def _paeth_row(cur: np.ndarray, prev: np.ndarray, bpp: int) -> None:
    """In-place Paeth reconstruction of one row (uint8)."""
    for i in range(cur.shape[0]):
        left = int(cur[i - bpp]) if i >= bpp else 0
        up = int(prev[i])
        ul = int(prev[i - bpp]) if i >= bpp else 0
        p = left + up - ul
        pa, pb, pc = abs(p - left), abs(p - up), abs(p - ul)
        pred = left if (pa <= pb and pa <= pc) else (up if pb <= pc else ul)
        cur[i] = (int(cur[i]) + pred) & 0xFF


def png_unfilter(raw: bytes, width: int, height: int, bpp: int) -> np.ndarray:
    """Undoes the per-row filters; returns (height, width * bpp) uint8."""
    stride = width * bpp
    data = np.frombuffer(raw, np.uint8)
    if data.size != height * (stride + 1):
        raise ValueError("PNG image data has the wrong size")
    rows = data.reshape(height, stride + 1)
    out = np.zeros((height, stride), np.uint8)
    prev = np.zeros(stride, np.uint8)
    for y in range(height):
        ftype = rows[y, 0]
        cur = rows[y, 1:].copy()
        if ftype == 0:
            pass
        elif ftype == 1:  # Sub: a running sum over the bytes of each channel position
            c = cur.astype(np.int64)
            for r in range(bpp):
                c[r::bpp] = np.cumsum(c[r::bpp]) & 0xFF
            cur = c.astype(np.uint8)
        elif ftype == 2:  # Up
            cur = (cur.astype(np.uint16) + prev).astype(np.uint8)
        elif ftype == 3:  # Average
            c = cur.astype(np.int32)
            p = prev.astype(np.int32)
            for i in range(stride):
                left = c[i - bpp] if i >= bpp else 0
                c[i] = (c[i] + ((left + p[i]) >> 1)) & 0xFF
            cur = c.astype(np.uint8)
        elif ftype == 4:
            _paeth_row(cur, prev, bpp)
        else:
            raise ValueError(f"unknown PNG filter {ftype}")
        out[y] = cur
        prev = cur
    return out


def srgb_to_linear(v: np.ndarray) -> np.ndarray:
    """The sRGB decoding function (IEC 61966-2-1)."""
    v = np.asarray(v, np.float64)
    return np.where(v <= 0.04045, v / 12.92, ((v + 0.055) / 1.055) ** 2.4).astype(np.float32)


def read_png(path: str, linear: bool = True) -> np.ndarray:
    data = open(path, "rb").read()
    if data[:8] != _PNG_MAGIC:
        raise ValueError(f"{path}: not a PNG file")
    p = 8
    idat = []
    width = height = depth = ctype = interlace = None
    while p < len(data):
        length, kind = struct.unpack(">I4s", data[p:p + 8])
        body = data[p + 8:p + 8 + length]
        crc, = struct.unpack(">I", data[p + 8 + length:p + 12 + length])
        if zlib.crc32(kind + body) & 0xFFFFFFFF != crc:
            raise ValueError(f"{path}: corrupt {kind!r} chunk")
        if kind == b"IHDR":
            width, height, depth, ctype, _, _, interlace = struct.unpack(">IIBBBBB", body)
        elif kind == b"IDAT":
            idat.append(body)
        elif kind == b"IEND":
            break
        p += 12 + length
    if interlace:
        raise ValueError(f"{path}: interlaced PNG is not supported")
    channels = {0: 1, 2: 3, 4: 2, 6: 4}.get(ctype)
    if channels is None or depth not in (8, 16):
        raise ValueError(f"{path}: PNG colour type {ctype} / depth {depth} is not supported")
    bpp = channels * depth // 8
    rows = png_unfilter(zlib.decompress(b"".join(idat)), width, height, bpp)
    if depth == 8:
        img = rows.reshape(height, width, channels).astype(np.float32) / 255.0
    else:
        img = rows.reshape(height, width * channels, 2).astype(np.uint16)
        img = ((img[..., 0] << 8) | img[..., 1]).reshape(height, width, channels).astype(np.float32) / 65535.0
    if channels in (1, 2):
        img = np.repeat(img[..., :1], 3, axis=-1)
    img = img[..., :3]
    return srgb_to_linear(img) if linear else img.astype(np.float32)


# ------------------------------------------------------------------- OpenEXR --
_EXR_MAGIC = 20000630
_NONE, _ZIPS, _ZIP = 0, 2, 3
_HALF, _FLOAT = 1, 2


def _exr_unpredict(buf: bytes) -> bytes:
    """ZIP/ZIPS decoding after inflate: undo the byte delta predictor
    (t[i] = t[i-1] + t[i] - 128, modulo 256), then re-interleave the two
    halves (first half = even bytes, second half = odd bytes)."""
    d = np.frombuffer(buf, np.uint8).astype(np.int64)
    if d.size == 0:
        return b""
    d[1:] -= 128
    t = (np.cumsum(d) & 0xFF).astype(np.uint8)
    half = (t.size + 1) // 2
    out = np.empty(t.size, np.uint8)
    out[0::2] = t[:half]
    out[1::2] = t[half:]
    return out.tobytes()


def _exr_predict(raw: bytes) -> bytes:
    """Inverse of _exr_unpredict (used by the writer)."""
    r = np.frombuffer(raw, np.uint8)
    half = (r.size + 1) // 2
    t = np.empty(r.size, np.uint8)
    t[:half] = r[0::2]
    t[half:] = r[1::2]
    d = t.astype(np.int64)
    out = d.copy()
    out[1:] = (d[1:] - d[:-1] + 128 + 256) & 0xFF
    return out.astype(np.uint8).tobytes()


def _read_exr_header(data: bytes):
    magic, version = struct.unpack("<ii", data[:8])
    if magic != _EXR_MAGIC:
        raise ValueError("not an OpenEXR file")
    if version & 0x200:
        raise ValueError("tiled OpenEXR files are not supported")
    p = 8
    attrs = {}
    while True:
        e = data.index(b"\0", p)
        name = data[p:e].decode()
        p = e + 1
        if not name:
            break
        e = data.index(b"\0", p)
        typ = data[p:e].decode()
        p = e + 1
        size, = struct.unpack("<i", data[p:p + 4])
        p += 4
        attrs[name] = (typ, data[p:p + size])
        p += size
    return attrs, p


def read_exr_channels(path: str) -> dict:
    """Every channel of a scanline OpenEXR file (NONE / ZIPS / ZIP, HALF / FLOAT)
    as an (h, w) float32 plane, by channel name."""
    data = open(path, "rb").read()
    attrs, p = _read_exr_header(data)
    comp = attrs["compression"][1][0]
    if comp not in (_NONE, _ZIPS, _ZIP):
        raise ValueError(f"{path}: OpenEXR compression {comp} is not supported (NONE, ZIPS, ZIP are)")
    xmin, ymin, xmax, ymax = struct.unpack("<iiii", attrs["dataWindow"][1])
    width, height = xmax - xmin + 1, ymax - ymin + 1
    chl = attrs["channels"][1]
    channels = []
    q = 0
    while chl[q] != 0:
        e = chl.index(b"\0", q)
        name = chl[q:e].decode()
        q = e + 1
        ptype, = struct.unpack("<i", chl[q:q + 4])
        xs, ys = struct.unpack("<ii", chl[q + 8:q + 16])
        if (xs, ys) != (1, 1):
            raise ValueError(f"{path}: subsampled channels are not supported")
        if ptype not in (_HALF, _FLOAT):
            raise ValueError(f"{path}: channel type {ptype} is not supported")
        channels.append((name, ptype))
        q += 16
    lines_per_block = 16 if comp == _ZIP else 1
    nblocks = (height + lines_per_block - 1) // lines_per_block
    offsets = struct.unpack(f"<{nblocks}Q", data[p:p + 8 * nblocks])
    planes = {name: np.zeros((height, width), np.float32) for name, _ in channels}
    line_bytes = sum(width * (2 if t == _HALF else 4) for _, t in channels)
    for off in offsets:
        y, size = struct.unpack("<ii", data[off:off + 8])
        block = data[off + 8:off + 8 + size]
        nlines = min(lines_per_block, ymax + 1 - y)
        if comp != _NONE and size < nlines * line_bytes:
            block = _exr_unpredict(zlib.decompress(block))
        if len(block) != nlines * line_bytes:
            raise ValueError(f"{path}: block at line {y} has the wrong size")
        q = 0
        for k in range(nlines):
            row = y - ymin + k
            for name, t in channels:
                n = width * (2 if t == _HALF else 4)
                vals = np.frombuffer(block[q:q + n], np.float16 if t == _HALF else np.float32)
                planes[name][row] = vals.astype(np.float32)
                q += n
    return planes


def read_exr_layers(path: str) -> dict:
    """The layers of an OpenEXR file: "" for the main image, otherwise the layer
    name of "<layer>.<channel>" channels; RGB layers as (h, w, 3), single-channel
    (Y) layers as (h, w)."""
    planes = read_exr_channels(path)
    groups: dict = {}
    for name, plane in planes.items():
        layer, _, ch = name.rpartition(".")
        groups.setdefault(layer, {})[ch] = plane
    out = {}
    for layer, chans in groups.items():
        if all(c in chans for c in "RGB"):
            out[layer] = np.stack([chans["R"], chans["G"], chans["B"]], -1)
        elif "Y" in chans:
            out[layer] = chans["Y"]
        elif len(chans) == 1:
            out[layer] = next(iter(chans.values()))
    return out


def read_exr(path: str) -> np.ndarray:
    """The main image of an OpenEXR file as (h, w, 3) float32 (R, G, B, or Y
    repeated)."""
    layers = read_exr_layers(path)
    img = layers.get("")
    if img is None:
        raise ValueError(f"{path}: no R, G, B or Y channels")
    return img if img.ndim == 3 else np.repeat(img[..., None], 3, -1)


def read_image(path: str) -> np.ndarray:
    """Reads a PNG (sRGB decoded to linear) or OpenEXR file as (h, w, 3) float32."""
    lower = path.lower()
    if lower.endswith(".exr"):
        return read_exr(path)
    if lower.endswith(".png"):
        return read_png(path)
    raise ValueError(f"{path}: unsupported image format (PNG and OpenEXR are supported)")


def write_exr(path: str, rgb: np.ndarray, compression: str = "zip") -> None:
    """Writes an (h, w, 3) float32 image as a scanline OpenEXR file with FLOAT
    R, G, B channels, NONE or ZIP compression."""
    write_exr_layers(path, [(None, rgb)], compression)


def write_exr_layers(path: str, layers, compression: str = "zip") -> None:
    """Writes several same-sized images into one scanline OpenEXR file with FLOAT
    channels: the main image (name None or "") as R, G, B, a named RGB layer as
    <name>.R/.G/.B and a single-channel (h, w) layer as <name>.Y; channels sorted
    by name as the format requires. NONE or ZIP compression."""
    planes = {}
    shape = None
    for name, img in layers:
        img = np.asarray(img, np.float32)
        prefix = f"{name}." if name else ""
        if img.ndim == 3:
            for k, c in enumerate("RGB"):
                planes[prefix + c] = img[..., k]
        else:
            planes[prefix + "Y"] = img
        if shape is None:
            shape = img.shape[:2]
        elif img.shape[:2] != shape:
            raise ValueError("all layers of an OpenEXR file need the same size")
    h, w = shape
    names = sorted(planes, key=lambda n: n.encode())
    comp = {"none": _NONE, "zip": _ZIP}[compression]

    def attr(name, typ, payload):
        return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(payload)) + payload

    chl = b"".join(c.encode() + b"\0" + struct.pack("<iBBBBii", _FLOAT, 0, 0, 0, 0, 1, 1) for c in names) + b"\0"
    header = struct.pack("<ii", _EXR_MAGIC, 2)
    header += attr("channels", "chlist", chl)
    header += attr("compression", "compression", bytes([comp]))
    header += attr("dataWindow", "box2i", struct.pack("<iiii", 0, 0, w - 1, h - 1))
    header += attr("displayWindow", "box2i", struct.pack("<iiii", 0, 0, w - 1, h - 1))
    header += attr("lineOrder", "lineOrder", bytes([0]))
    header += attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
    header += attr("screenWindowCenter", "v2f", struct.pack("<ff", 0.0, 0.0))
    header += attr("screenWindowWidth", "float", struct.pack("<f", 1.0))
    header += b"\0"
    lines = 16 if comp == _ZIP else 1
    cols = [np.ascontiguousarray(planes[n], np.float32) for n in names]
    blocks = []
    for y in range(0, h, lines):
        raw = b"".join(c[r].tobytes() for r in range(y, min(h, y + lines)) for c in cols)
        if comp == _ZIP:
            packed = zlib.compress(_exr_predict(raw))
            if len(packed) < len(raw):
                raw = packed
        blocks.append(struct.pack("<ii", y, len(raw)) + raw)
    table_start = len(header)
    pos = table_start + 8 * len(blocks)
    offsets = []
    for b in blocks:
        offsets.append(pos)
        pos += len(b)
    with open(path, "wb") as f:
        f.write(header)
        f.write(struct.pack(f"<{len(offsets)}Q", *offsets))
        for b in blocks:
            f.write(b)
