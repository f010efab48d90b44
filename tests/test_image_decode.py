"""The C++ image decoders (gs::decode_image) against files written here by an
independent encoder, without a GPU.

PNG: every colour type and bit depth, tRNS keys and palette alpha, all five
row filters (forward-filtered here), Adam7 and odd sizes; the expected RGBA
follows the decoder's documented conversions.

JPEG (baseline, Huffman): the zigzag order is computed here by walking the
anti-diagonals (not typed in), Huffman tables are built from this image's
own symbol counts (canonical codes, the all-ones code reserved), the DCT is a
float64 orthonormal matrix product.  The expected image is the exact
reconstruction of the quantised coefficients (float64 IDCT, level shift,
round, clamp), upsampled with centred bilinear weights and converted with
the JFIF matrix; the decoder (float32) must agree within 2 levels.  Covered:
grayscale, 4:4:4 with restart markers, 4:2:0 with 16-bit quantisation
tables, 4:2:2, non-interleaved scans, Adobe RGB; progressive files rejected
by name.
"""
import heapq
import math
import os
import struct
import subprocess
import zlib

import numpy as np
import pytest

from helpers import ROOT

EXE = os.path.join(ROOT, "on-divice-gsplate_amd", "_build", "test_image_io")
# The same decoders built with AddressSanitizer (host code only).
ASAN_EXE = os.path.join(ROOT, "on-divice-gsplate_amd", "_build", "asan_image_io")


def _decode_with(exe, path, channels, tmp_path):
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built: run make -C on-divice-gsplate_amd tests")
    out = str(tmp_path / ("decoded_%s.bin" % os.path.basename(exe)))
    r = subprocess.run([exe, "decode", str(path), str(channels), out], capture_output=True, text=True, timeout=60)
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr
    if r.returncode != 0:
        return None, r.stderr
    raw = open(out, "rb").read()
    w, h, c = struct.unpack("<3I", raw[:12])
    return np.frombuffer(raw[12:], np.uint8).reshape(h, w, c), ""


def decode(path, channels, tmp_path, exe=None):
    """Decodes with the test binary; with exe=None also with the
    AddressSanitizer build, which must give the same result."""
    if exe is not None:
        return _decode_with(exe, path, channels, tmp_path)
    got, err = _decode_with(EXE, path, channels, tmp_path)
    got_asan, err_asan = _decode_with(ASAN_EXE, path, channels, tmp_path)
    assert (got is None) == (got_asan is None), err + err_asan
    assert got is None or np.array_equal(got, got_asan)
    return got, err


# ---- PNG -------------------------------------------------------------------------

ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]


def pack_row(samples, depth):
    """One row of samples (flat, in storage order) as bytes."""
    if depth == 16:
        return b"".join(struct.pack(">H", int(s)) for s in samples)
    if depth == 8:
        return bytes(int(s) for s in samples)
    out, acc, nbits = bytearray(), 0, 0
    for s in samples:
        acc = (acc << depth) | int(s)
        nbits += depth
        if nbits == 8:
            out.append(acc)
            acc, nbits = 0, 0
    if nbits:
        out.append(acc << (8 - nbits))
    return bytes(out)


def paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    return b if pb <= pc else c


def filter_rows(rows, bpp, first_filter):
    out, prev = bytearray(), bytes(len(rows[0])) if rows else b""
    for r, row in enumerate(rows):
        f = (first_filter + r) % 5
        enc = bytearray()
        for i, x in enumerate(row):
            left = row[i - bpp] if i >= bpp else 0
            up = prev[i]
            ul = prev[i - bpp] if i >= bpp else 0
            pred = [0, left, up, (left + up) >> 1, paeth(left, up, ul)][f]
            enc.append((x - pred) & 0xFF)
        out.append(f)
        out += enc
        prev = row
    return bytes(out)


def write_png(samples, colour, depth, interlace=False, palette=None, trns=None, first_filter=0):
    """samples: (H, W, spp) ints in storage range."""
    h, w, spp = samples.shape
    bpp = max(1, spp * depth // 8)
    data = b""
    passes = ADAM7 if interlace else [(0, 0, 1, 1)]
    for k, (x0, y0, dx, dy) in enumerate(passes):
        sub = samples[y0::dy, x0::dx]
        if sub.shape[0] == 0 or sub.shape[1] == 0:
            continue
        rows = [pack_row(sub[y].reshape(-1), depth) for y in range(sub.shape[0])]
        data += filter_rows(rows, bpp, first_filter + k)

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)

    out = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, colour, 0, 0, int(interlace)))
    if palette is not None:
        out += chunk(b"PLTE", bytes(palette.reshape(-1).astype(np.uint8)))
    if trns is not None:
        out += chunk(b"tRNS", trns)
    return out + chunk(b"IDAT", zlib.compress(data, 9)) + chunk(b"IEND", b"")


def to8(v, depth):
    v = np.asarray(v, np.int64)
    if depth == 16:
        return (v * 255 + 32767) // 65535
    if depth < 8:
        return v * (255 // ((1 << depth) - 1))
    return v


PNG_CASES = []
for colour, depths in ((0, (1, 2, 4, 8, 16)), (2, (8, 16)), (3, (1, 2, 4, 8)), (4, (8, 16)), (6, (8, 16))):
    for depth in depths:
        for interlace in (False, True):
            for alpha in ((False, True) if colour in (0, 2, 3) else (False,)):
                PNG_CASES.append((colour, depth, interlace, alpha))


@pytest.mark.parametrize("colour,depth,interlace,alpha", PNG_CASES,
                         ids=[f"c{c}-d{d}-{'adam7' if i else 'flat'}{'-trns' if a else ''}" for c, d, i, a in PNG_CASES])
def test_png_decoder(colour, depth, interlace, alpha, tmp_path):
    rng = np.random.default_rng(colour * 100 + depth * 10 + int(interlace) * 2 + int(alpha))
    w, h = (13, 7) if not interlace else (9, 17)
    spp = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[colour]
    top = (1 << depth) - 1
    palette = trns = None
    if colour == 3:
        n = min(1 << depth, 40)
        samples = rng.integers(0, n, size=(h, w, 1))
        palette = rng.integers(0, 256, size=(n, 3))
        if alpha:
            trns = bytes(rng.integers(0, 256, size=n // 2).astype(np.uint8))
    else:
        samples = rng.integers(0, top + 1, size=(h, w, spp))
        if alpha:
            key = samples[h // 2, w // 2, :spp].copy()
            samples[0, :3, :] = key  # a few more pixels carry the key
            trns = b"".join(struct.pack(">H", int(k)) for k in key)
    f = tmp_path / "img.png"
    f.write_bytes(write_png(samples, colour, depth, interlace, palette, trns, first_filter=depth % 5))
    got, err = decode(f, 4, tmp_path)
    assert got is not None, err
    want = np.zeros((h, w, 4), np.int64)
    if colour == 3:
        want[..., :3] = palette[samples[..., 0]]
        want[..., 3] = 255
        if trns is not None:
            idx = samples[..., 0]
            t = np.frombuffer(trns, np.uint8)
            want[..., 3] = np.where(idx < len(t), t[np.minimum(idx, len(t) - 1)], 255)
    else:
        g = to8(samples, depth)
        if colour in (0, 4):
            want[..., :3] = g[..., :1]
        else:
            want[..., :3] = g[..., :3]
        want[..., 3] = 255
        if colour in (4, 6):
            want[..., 3] = g[..., -1]
        if trns is not None:
            is_key = np.all(samples == key, axis=-1)
            want[..., 3] = np.where(is_key, 0, 255)
    np.testing.assert_array_equal(got, want)
    if colour == 2 and depth == 8:  # luma of RGB: BT.601 weights, rounded
        got1, _ = decode(f, 1, tmp_path)
        s = samples.astype(np.int64)
        np.testing.assert_array_equal(got1[..., 0], (299 * s[..., 0] + 587 * s[..., 1] + 114 * s[..., 2] + 500) // 1000)


# ---- JPEG ------------------------------------------------------------------------

def zigzag():
    order = []
    for s in range(15):
        cells = [(r, s - r) for r in range(max(0, s - 7), min(s, 7) + 1)]  # row increasing
        if s % 2 == 0:
            cells.reverse()
        order += [r * 8 + c for r, c in cells]
    return order


ZZ = zigzag()
A = np.array([[(math.sqrt(0.5) if u == 0 else 1.0) * 0.5 * math.cos((2 * x + 1) * u * math.pi / 16) for u in range(8)]
              for x in range(8)])


def huffman_table(freq):
    """(bits[16], values) of a canonical code for `freq` {symbol: count}; the
    all-ones code is left unused via a reserved pseudo-symbol."""
    items = [(c, i, [s]) for i, (s, c) in enumerate(sorted(freq.items()))]
    items.append((0.5, -1, [256]))  # reserved
    lengths = {s: 0 for s in list(freq) + [256]}
    heap = list(items)
    heapq.heapify(heap)
    tie = len(items)
    while len(heap) > 1:
        c1, _, s1 = heapq.heappop(heap)
        c2, _, s2 = heapq.heappop(heap)
        for s in s1 + s2:
            lengths[s] += 1
        heapq.heappush(heap, (c1 + c2, tie, s1 + s2))
        tie += 1
    assert max(lengths.values()) <= 16
    order = sorted(lengths, key=lambda s: (lengths[s], s == 256, s))
    bits = [0] * 16
    for s in order:
        bits[lengths[s] - 1] += 1
    bits[lengths[256] - 1] -= 1  # the reserved symbol is last among the longest: drop it
    values = [s for s in order if s != 256]
    return bits, values


def canonical_codes(bits, values):
    codes, code, k = {}, 0, 0
    for length in range(1, 17):
        for _ in range(bits[length - 1]):
            codes[values[k]] = (code, length)
            code += 1
            k += 1
        code <<= 1
    return codes


class BitWriter:
    def __init__(self):
        self.out, self.acc, self.n = bytearray(), 0, 0

    def put(self, value, length):
        for i in range(length - 1, -1, -1):
            self.acc = (self.acc << 1) | ((value >> i) & 1)
            self.n += 1
            if self.n == 8:
                self.out.append(self.acc)
                if self.acc == 0xFF:
                    self.out.append(0)
                self.acc, self.n = 0, 0

    def flush(self):
        while self.n:
            self.put(1, 1)


def category(v):
    return 0 if v == 0 else int(abs(v)).bit_length()


def magnitude_bits(v, s):
    return v if v >= 0 else v + (1 << s) - 1


def encode_jpeg(planes, factors, qtables, comp_q, restart=0, interleaved=True, adobe_rgb=False, q16=False):
    """planes: per-component float arrays at full resolution (levels 0..255),
    already subsampled by `factors` [(h, v)] relative to the max factors.
    Returns (bytes, decoded planes as the decoder reconstructs them)."""
    hmax = max(f[0] for f in factors)
    vmax = max(f[1] for f in factors)
    H = max(p.shape[0] * vmax // f[1] for p, f in zip(planes, factors))  # caller passes exact sizes
    W = max(p.shape[1] * hmax // f[0] for p, f in zip(planes, factors))
    mcux, mcuy = -(-W // (8 * hmax)), -(-H // (8 * vmax))
    blocks, recon = [], []
    for p, (h, v), qi in zip(planes, factors, comp_q):
        pw, ph = mcux * h * 8, mcuy * v * 8
        padded = np.pad(p, ((0, ph - p.shape[0]), (0, pw - p.shape[1])), mode="edge")
        q = np.array(qtables[qi], np.float64)  # zigzag order
        qn = np.zeros(64)
        qn[ZZ] = q
        coefs = np.zeros((ph // 8, pw // 8, 64), np.int64)
        rec = np.zeros((ph, pw))
        for by in range(ph // 8):
            for bx in range(pw // 8):
                f = padded[by * 8:(by + 1) * 8, bx * 8:(bx + 1) * 8] - 128.0
                F = A.T @ f @ A  # F[v][u]
                Q = np.rint(F.reshape(-1) / qn).astype(np.int64)
                coefs[by, bx] = Q
                rec[by * 8:(by + 1) * 8, bx * 8:(bx + 1) * 8] = A @ (Q * qn).reshape(8, 8) @ A.T + 128.0
        blocks.append(coefs)
        recon.append(np.clip(np.rint(rec), 0, 255))
    # scans: lists of component indices
    scans = [list(range(len(planes)))] if interleaved else [[c] for c in range(len(planes))]
    ncomp = len(planes)
    tclass = [0 if c == 0 else 1 for c in range(ncomp)]

    def units(scan):
        """(component, block y, block x) in coding order, grouped by MCU."""
        if len(scan) == 1:
            c = scan[0]
            h, v = factors[c]
            vw, vh = -(-W * h // hmax), -(-H * v // vmax)
            return [[(c, by, bx)] for by in range(-(-vh // 8)) for bx in range(-(-vw // 8))]
        out = []
        for my in range(mcuy):
            for mx in range(mcux):
                mcu = []
                for c in scan:
                    h, v = factors[c]
                    mcu += [(c, my * v + y, mx * h + x) for y in range(v) for x in range(h)]
                out.append(mcu)
        return out

    def symbols(scan, emit):
        pred = [0] * ncomp
        for m, mcu in enumerate(units(scan)):
            if restart and m > 0 and m % restart == 0:
                emit(("rst", (m // restart - 1) % 8))
                pred = [0] * ncomp
            for c, by, bx in mcu:
                Q = blocks[c][by, bx]
                diff = int(Q[0]) - pred[c]
                pred[c] = int(Q[0])
                s = category(diff)
                emit(("dc", tclass[c], s, diff))
                run = 0
                for k in range(1, 64):
                    v = int(Q[ZZ[k]])
                    if v == 0:
                        run += 1
                        continue
                    while run > 15:
                        emit(("ac", tclass[c], 0xF0, 0))
                        run -= 16
                    s = category(v)
                    emit(("ac", tclass[c], (run << 4) | s, v))
                    run = 0
                if run:
                    emit(("ac", tclass[c], 0x00, 0))

    freq = {("dc", 0): {}, ("dc", 1): {}, ("ac", 0): {}, ("ac", 1): {}}

    def count(e):
        if e[0] != "rst":
            d = freq[(e[0], e[1])]
            d[e[2]] = d.get(e[2], 0) + 1

    for scan in scans:
        symbols(scan, count)
    tables = {k: huffman_table(v) for k, v in freq.items() if v}
    codes = {k: canonical_codes(*t) for k, t in tables.items()}

    def seg(marker, payload):
        return bytes([0xFF, marker]) + struct.pack(">H", len(payload) + 2) + payload

    out = b"\xff\xd8"
    if adobe_rgb:
        out += seg(0xEE, b"Adobe" + bytes([0, 100, 0, 0, 0, 0, 0]))
    for i, t in enumerate(qtables):
        if q16:
            out += seg(0xDB, bytes([0x10 | i]) + b"".join(struct.pack(">H", int(x)) for x in t))
        else:
            out += seg(0xDB, bytes([i]) + bytes(int(x) for x in t))
    sof = struct.pack(">BHHB", 8, H, W, ncomp)
    for c in range(ncomp):
        sof += bytes([c + 1, (factors[c][0] << 4) | factors[c][1], comp_q[c]])
    out += seg(0xC0, sof)
    for (kind, tc), (bits, values) in tables.items():
        out += seg(0xC4, bytes([(0 if kind == "dc" else 1) << 4 | tc]) + bytes(bits) + bytes(values))
    if restart:
        out += seg(0xDD, struct.pack(">H", restart))
    for scan in scans:
        sos = bytes([len(scan)])
        for c in scan:
            sos += bytes([c + 1, (tclass[c] << 4) | tclass[c]])
        out += seg(0xDA, sos + bytes([0, 63, 0]))
        bw = BitWriter()
        data = bytearray()

        def write(e):
            nonlocal data
            if e[0] == "rst":
                bw.flush()
                data += bw.out + bytes([0xFF, 0xD0 + e[1]])
                bw.out = bytearray()
                return
            code, length = codes[(e[0], e[1])][e[2]]
            bw.put(code, length)
            s = e[2] if e[0] == "dc" else e[2] & 15
            if s:
                bw.put(magnitude_bits(e[3], s), s)

        symbols(scan, write)
        bw.flush()
        out += bytes(data + bw.out)
    out += b"\xff\xd9"
    return out, recon, (W, H, hmax, vmax)


def upsample_expected(recon, factors, geom):
    W, H, hmax, vmax = geom
    out = []
    for rec, (h, v) in zip(recon, factors):
        vw, vh = -(-W * h // hmax), -(-H * v // vmax)
        if h == hmax and v == vmax:
            out.append(rec[:H, :W])
            continue
        xs = (np.arange(W) + 0.5) * h / hmax - 0.5
        ys = (np.arange(H) + 0.5) * v / vmax - 0.5
        x0, y0 = np.floor(xs).astype(int), np.floor(ys).astype(int)
        fx, fy = xs - x0, ys - y0
        xa, xb = np.clip(x0, 0, vw - 1), np.clip(x0 + 1, 0, vw - 1)
        ya, yb = np.clip(y0, 0, vh - 1), np.clip(y0 + 1, 0, vh - 1)
        top = rec[ya][:, xa] * (1 - fx) + rec[ya][:, xb] * fx
        bot = rec[yb][:, xa] * (1 - fx) + rec[yb][:, xb] * fx
        out.append(top * (1 - fy)[:, None] + bot * fy[:, None])
    return out


def smooth_rgb(w, h, seed):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    img = np.zeros((h, w, 3))
    for c in range(3):
        a, b, p = rng.uniform(0.05, 0.4, 3)
        img[..., c] = 128 + 90 * np.sin(a * x + b * y + 6 * p) + rng.normal(0, 6, (h, w))
    return np.clip(img, 0, 255)


def subsample(p, h, v, hmax, vmax, W, H):
    fx, fy = hmax // h, vmax // v
    vw, vh = -(-W * h // hmax), -(-H * v // vmax)
    pp = np.pad(p, ((0, vh * fy - H), (0, vw * fx - W)), mode="edge")
    return pp.reshape(vh, fy, vw, fx).mean(axis=(1, 3))


QL = [2 + (k * 3) % 11 for k in range(64)]
QC = [3 + (k * 5) % 13 for k in range(64)]

JPEG_CASES = [
    # name, W, H, factors (Y, Cb, Cr) or gray, restart, interleaved, adobe, q16
    ("gray", 17, 9, [(1, 1)], 0, True, False, False),
    ("444-restart", 33, 20, [(1, 1), (1, 1), (1, 1)], 3, True, False, False),
    ("420-q16-restart", 45, 31, [(2, 2), (1, 1), (1, 1)], 5, True, False, True),
    ("422", 40, 16, [(2, 1), (1, 1), (1, 1)], 0, True, False, False),
    ("420-noninterleaved", 29, 23, [(2, 2), (1, 1), (1, 1)], 0, False, False, False),
    ("adobe-rgb", 16, 16, [(1, 1), (1, 1), (1, 1)], 2, True, True, False),
]


@pytest.mark.parametrize("name,W,H,factors,restart,interleaved,adobe,q16", JPEG_CASES, ids=[c[0] for c in JPEG_CASES])
def test_jpeg_decoder(name, W, H, factors, restart, interleaved, adobe, q16, tmp_path):
    rgb = smooth_rgb(W, H, len(name))
    hmax, vmax = max(f[0] for f in factors), max(f[1] for f in factors)
    if len(factors) == 1:
        full = [rgb[..., 0] * 0.299 + rgb[..., 1] * 0.587 + rgb[..., 2] * 0.114]
    elif adobe:
        full = [rgb[..., 0], rgb[..., 1], rgb[..., 2]]
    else:
        R, G, B = rgb[..., 0], rgb[..., 1], rgb[..., 2]
        full = [0.299 * R + 0.587 * G + 0.114 * B, -0.168736 * R - 0.331264 * G + 0.5 * B + 128,
                0.5 * R - 0.418688 * G - 0.081312 * B + 128]
    planes = [subsample(p, h, v, hmax, vmax, W, H) for p, (h, v) in zip(full, factors)]
    data, recon, geom = encode_jpeg(planes, factors, [QL, QC], [0] + [1] * (len(factors) - 1), restart, interleaved,
                                    adobe, q16)
    f = tmp_path / "img.jpg"
    f.write_bytes(data)
    got, err = decode(f, 3, tmp_path)
    assert got is not None, err
    up = upsample_expected(recon, factors, geom)
    if len(up) == 1:
        want = np.repeat(up[0][..., None], 3, axis=-1)
    elif adobe:
        want = np.stack(up, -1)
    else:
        Y, cb, cr = up[0], up[1] - 128.0, up[2] - 128.0
        want = np.stack([Y + 1.402 * cr, Y - 0.344136 * cb - 0.714136 * cr, Y + 1.772 * cb], -1)
    want = np.clip(np.floor(want + 0.5), 0, 255)
    diff = np.abs(got.astype(np.float64) - want)
    print(f"[info] jpeg {name}: {W}x{H}, {len(data)} bytes, max diff {diff.max():.0f}, mean {diff.mean():.4f}")
    assert diff.max() <= 2
    assert diff.mean() < 0.05


def test_jpeg_rejects_progressive(tmp_path):
    sof2 = bytes([0xFF, 0xC2, 0, 11, 8, 0, 8, 0, 8, 1, 1, 0x11, 0])
    f = tmp_path / "p.jpg"
    f.write_bytes(b"\xff\xd8" + sof2 + b"\xff\xd9")
    got, err = decode(f, 3, tmp_path)
    assert got is None and "progressive" in err


@pytest.mark.parametrize("exe", [EXE, ASAN_EXE], ids=["plain", "asan"])
@pytest.mark.parametrize("counts", [[255] + [0] * 15, [0, 5] + [0] * 14, [0] * 8 + [0, 0, 0, 0, 0, 0, 0, 255]],
                         ids=["len1x255", "len2x5", "len16x255"])
def test_jpeg_rejects_oversubscribed_huffman_table(counts, exe, tmp_path):
    """A DHT whose counts assign more codes of a length than that length has
    (255 one-bit codes; five two-bit codes; 255 sixteen-bit codes after none
    shorter fit, which is legal, so only the first two are bad) is rejected
    with an error, and decoding it touches no memory outside the tables."""
    total = sum(counts)
    seg = bytes([0x00]) + bytes(counts) + bytes(range(total))
    dht = b"\xff\xc4" + struct.pack(">H", 2 + len(seg)) + seg
    f = tmp_path / "h.jpg"
    f.write_bytes(b"\xff\xd8" + dht + b"\xff\xd9")
    got, err = decode(f, 3, tmp_path, exe)
    assert got is None
    assert "AddressSanitizer" not in err and "runtime error" not in err, err
    if counts[0] > 2 or counts[1] > 4:
        assert "bad Huffman table" in err, err
    else:
        assert "no image data" in err, err


def test_zigzag_table_in_the_decoder_matches_the_walk():
    """The decoder's typed table (image_io.cpp kZigzag) equals the walk."""
    src = open(os.path.join(ROOT, "on-divice-gsplate_amd", "host", "image_io.cpp")).read()
    body = src[src.index("kZigzag[64] = {") + len("kZigzag[64] = {"):]
    body = body[:body.index("}")]
    table = [int(x) for x in body.replace("\n", " ").split(",")]
    assert table == ZZ
