"""Image writers for test fixtures: PNG (every colour type and bit depth, optional
Adam7 interlace, every row filter) and baseline JPEG (JFIF / Adobe / 'RGB'-id markers,
4:4:4, 4:2:2, 4:4:0 and 4:2:0 chroma, grayscale, restart intervals).

Test infrastructure only. The files they write are decoded by the product's decoder and
by the reference's stb_image (through oracle/_ref), and the two outputs must agree.
"""
from __future__ import annotations

import struct
import zlib

import numpy as np

# ------------------------------------------------------------------------------------ PNG


def _chunk(tag: bytes, data: bytes) -> bytes:
    return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)


def _filter_rows(raw: np.ndarray, bpp: int, seed: int) -> bytes:
    """raw: (h, stride) uint8 scanlines -> filtered stream, filter type chosen per row."""
    out = bytearray()
    prev = np.zeros(raw.shape[1], np.int64)
    for y in range(raw.shape[0]):
        cur = raw[y].astype(np.int64)
        ft = (y + seed) % 5
        out.append(ft)
        row = np.zeros_like(cur)
        for i in range(len(cur)):
            a = cur[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            pred = [0, a, b, (a + b) >> 1, _paeth(a, b, c)][ft]
            row[i] = (cur[i] - pred) & 0xFF
        out += bytes(row.astype(np.uint8))
        prev = cur
    return bytes(out)


def _pack_samples(samples: np.ndarray, depth: int) -> np.ndarray:
    """(h, w*ch) integer samples -> (h, stride) bytes at the given bit depth."""
    h, n = samples.shape
    if depth == 8:
        return samples.astype(np.uint8)
    if depth == 16:
        s = samples.astype(np.uint16)
        return np.stack([s >> 8, s & 0xFF], axis=-1).reshape(h, n * 2).astype(np.uint8)
    per = 8 // depth
    stride = (n * depth + 7) // 8
    out = np.zeros((h, stride), np.uint8)
    for i in range(n):
        shift = 8 - depth * (i % per + 1)
        out[:, i // per] |= (samples[:, i].astype(np.uint8) << shift).astype(np.uint8)
    return out


def write_png(path: str, samples: np.ndarray, color: int, depth: int, palette=None, interlace: bool = False,
              seed: int = 0) -> None:
    """samples: (h, w, ch) integers in [0, 2^depth) (palette indices for colour type 3)."""
    h, w, ch = samples.shape
    bpp = max(1, ch * depth // 8)
    ihdr = struct.pack(">IIBBBBB", w, h, depth, color, 0, 0, 1 if interlace else 0)
    if not interlace:
        data = _filter_rows(_pack_samples(samples.reshape(h, w * ch), depth), bpp, seed)
    else:
        data = b""
        passes = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]
        for k, (x0, y0, dx, dy) in enumerate(passes):
            sub = samples[y0::dy, x0::dx]
            if sub.shape[0] == 0 or sub.shape[1] == 0:
                continue
            sh, sw = sub.shape[:2]
            data += _filter_rows(_pack_samples(sub.reshape(sh, sw * ch), depth), bpp, seed + k)
    out = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", ihdr)
    if palette is not None:
        out += _chunk(b"PLTE", bytes(np.asarray(palette, np.uint8).reshape(-1)))
    out += _chunk(b"tEXt", b"Comment\x00fixture")  # an ancillary chunk to skip
    z = zlib.compress(data, 9)
    out += _chunk(b"IDAT", z[: len(z) // 2]) + _chunk(b"IDAT", z[len(z) // 2:])  # split IDAT
    out += _chunk(b"IEND", b"")
    with open(path, "wb") as fh:
        fh.write(out)


# ----------------------------------------------------------------------------------- JPEG

_ZIGZAG = np.array([0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48, 41, 34, 27, 20,
                    13, 6, 7, 14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52,
                    45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63])

_QL = np.array([16, 11, 10, 16, 24, 40, 51, 61, 12, 12, 14, 19, 26, 58, 60, 55, 14, 13, 16, 24, 40, 57, 69, 56, 14, 17,
                22, 29, 51, 87, 80, 62, 18, 22, 37, 56, 68, 109, 103, 77, 24, 35, 55, 64, 81, 104, 113, 92, 49, 64, 78,
                87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99])
_QC = np.array([17, 18, 24, 47, 99, 99, 99, 99, 18, 21, 26, 66, 99, 99, 99, 99, 24, 26, 56, 99, 99, 99, 99, 99, 47, 66,
                99, 99, 99, 99, 99, 99] + [99] * 32)


def _qtable(base, quality):
    s = 5000 / quality if quality < 50 else 200 - 2 * quality
    return np.clip((base * s + 50) // 100, 1, 255).astype(np.int64)


def _dct_matrix():
    m = np.zeros((8, 8))
    for k in range(8):
        for n in range(8):
            m[k, n] = (np.sqrt(1 / 8) if k == 0 else np.sqrt(2 / 8)) * np.cos((2 * n + 1) * k * np.pi / 16)
    return m


_DCT = _dct_matrix()


def _canonical(symbols_by_len):
    """{length: [symbols]} -> (bits[16], vals, {symbol: (code, length)})."""
    bits, vals, codes = [0] * 16, [], {}
    code = 0
    for ln in range(1, 17):
        for s in symbols_by_len.get(ln, []):
            codes[s] = (code, ln)
            vals.append(s)
            bits[ln - 1] += 1
            code += 1
        code <<= 1
    return bits, vals, codes


def _tables():
    """Huffman tables of mixed code lengths covering every baseline symbol."""
    dc = list(range(12))
    ac = [0x00, 0xF0] + [(r << 4) | s for r in range(16) for s in range(1, 11)]
    # a few short codes first, the rest longer (Kraft sums below 1, so no all-ones code)
    dcl = {2: dc[:1], 3: dc[1:4], 4: dc[4:8], 5: dc[8:11], 6: dc[11:]}
    acl = {2: ac[:1], 4: ac[1:7], 6: ac[7:21], 10: ac[21:]}
    for t in (dcl, acl):
        assert sum(len(v) * 2.0 ** -k for k, v in t.items()) < 1
    return _canonical(dcl), _canonical(acl)


class _BitWriter:
    def __init__(self):
        self.out = bytearray()
        self.acc = 0
        self.n = 0

    def put(self, code, ln):
        self.acc = (self.acc << ln) | code
        self.n += ln
        while self.n >= 8:
            b = (self.acc >> (self.n - 8)) & 0xFF
            self.out.append(b)
            if b == 0xFF:
                self.out.append(0)
            self.n -= 8
        self.acc &= (1 << self.n) - 1

    def flush(self):
        if self.n:
            self.put((1 << (8 - self.n)) - 1, 8 - self.n)


def _category(v):
    v = abs(int(v))
    return v.bit_length()


def _bits_of(v, s):
    return v if v >= 0 else v + (1 << s) - 1


def write_jpeg(path: str, img: np.ndarray, quality: int = 80, sampling=(2, 2), restart: int = 0,
               marker: str = "jfif") -> None:
    """img: (h, w, 3) uint8 RGB or (h, w) uint8 gray. sampling: luma (h, v) factors
    (chroma 1x1). marker: 'jfif' (YCbCr), 'adobe0' (Adobe transform 0: RGB stored),
    'rgbids' (component ids 'R','G','B': RGB stored), 'none' (no marker, YCbCr)."""
    gray = img.ndim == 2
    h, w = img.shape[:2]
    x = img.astype(np.float64)
    if gray:
        planes = [x]
        samp = [(1, 1)]
        ids = [1]
    else:
        r, g, b = x[..., 0], x[..., 1], x[..., 2]
        if marker in ("adobe0", "rgbids"):
            planes = [r, g, b]
            samp = [(1, 1)] * 3
        else:
            planes = [0.299 * r + 0.587 * g + 0.114 * b,
                      -0.168736 * r - 0.331264 * g + 0.5 * b + 128,
                      0.5 * r - 0.418688 * g - 0.081312 * b + 128]
            samp = [tuple(sampling), (1, 1), (1, 1)]
        ids = [ord("R"), ord("G"), ord("B")] if marker == "rgbids" else [1, 2, 3]
    hmax = max(s[0] for s in samp)
    vmax = max(s[1] for s in samp)
    mcux = -(-w // (8 * hmax))
    mcuy = -(-h // (8 * vmax))
    qt = [_qtable(_QL, quality), _qtable(_QC, quality)]
    tq = [0] + [1] * (len(planes) - 1) if marker not in ("adobe0", "rgbids") else [0] * len(planes)
    comps = []
    for p, (sh, sv) in zip(planes, samp):
        fx, fy = hmax // sh, vmax // sv
        cw, chh = mcux * sh * 8, mcuy * sv * 8
        pw = np.pad(p, ((0, mcuy * vmax * 8 - h), (0, mcux * hmax * 8 - w)), mode="edge")
        sub = pw.reshape(chh, fy, cw, fx).mean(axis=(1, 3))
        comps.append(sub - 128.0)
    (dcb, dcv, dcc), (acb, acv, acc) = _tables()
    bw = _BitWriter()
    pred = [0] * len(comps)
    nmcu = 0
    rst = 0
    total = mcux * mcuy
    for my in range(mcuy):
        for mx in range(mcux):
            for ci, (plane, (sh, sv)) in enumerate(zip(comps, samp)):
                q = qt[tq[ci]]
                for yy in range(sv):
                    for xx in range(sh):
                        by, bx = (my * sv + yy) * 8, (mx * sh + xx) * 8
                        blk = plane[by:by + 8, bx:bx + 8]
                        coef = _DCT @ blk @ _DCT.T
                        zz = np.round(coef.reshape(64)[_ZIGZAG] / q[_ZIGZAG]).astype(np.int64)
                        diff = int(zz[0]) - pred[ci]
                        pred[ci] = int(zz[0])
                        s = _category(diff)
                        bw.put(*dcc[s])
                        if s:
                            bw.put(_bits_of(diff, s), s)
                        run = 0
                        for k in range(1, 64):
                            v = int(zz[k])
                            if v == 0:
                                run += 1
                                continue
                            while run > 15:
                                bw.put(*acc[0xF0])
                                run -= 16
                            s = _category(v)
                            bw.put(*acc[(run << 4) | s])
                            bw.put(_bits_of(v, s), s)
                            run = 0
                        if run:
                            bw.put(*acc[0x00])
            nmcu += 1
            if restart and nmcu % restart == 0 and nmcu < total:
                bw.flush()
                bw.out += bytes([0xFF, 0xD0 + (rst & 7)])
                rst += 1
                pred = [0] * len(comps)
    bw.flush()
    seg = bytearray(b"\xff\xd8")
    if marker == "jfif":
        seg += b"\xff\xe0" + struct.pack(">H", 16) + b"JFIF\x00\x01\x01\x00\x00\x01\x00\x01\x00\x00"
    if marker == "adobe0":
        seg += b"\xff\xee" + struct.pack(">H", 14) + b"Adobe\x00" + bytes([100, 0, 0, 0, 0, 0])
    for t in range(2 if (not gray and marker not in ("adobe0", "rgbids")) else 1):
        seg += b"\xff\xdb" + struct.pack(">H", 67) + bytes([t]) + bytes(qt[t][_ZIGZAG].astype(np.uint8))
    seg += b"\xff\xc0" + struct.pack(">HBHHB", 8 + 3 * len(comps), 8, h, w, len(comps))
    for ci, (sh, sv) in enumerate(samp):
        seg += bytes([ids[ci], (sh << 4) | sv, tq[ci]])
    for cls, tid, bits, vals in ((0, 0, dcb, dcv), (1, 0, acb, acv)):
        seg += b"\xff\xc4" + struct.pack(">H", 3 + 16 + len(vals)) + bytes([(cls << 4) | tid]) + bytes(bits) + \
            bytes(vals)
    if restart:
        seg += b"\xff\xdd" + struct.pack(">HH", 4, restart)
    seg += b"\xff\xda" + struct.pack(">HB", 6 + 2 * len(comps), len(comps))
    for ci in range(len(comps)):
        seg += bytes([ids[ci], 0x00])
    seg += bytes([0, 63, 0]) + bytes(bw.out) + b"\xff\xd9"
    with open(path, "wb") as fh:
        fh.write(bytes(seg))
