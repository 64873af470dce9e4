"""Writer of synthetic baseline JPEG files (ITU-T T.81 sequential DCT, Huffman coded) for the
decoder tests. The files hold seeded random coefficient blocks rather than encoded pictures, which
exercises every decoder path the same way: Huffman tables built from random symbol frequencies
(code lengths 1-16), 8- or 16-bit quantisation tables, any sampling factors, interleaved and
single-component scans, restart intervals, fill bytes and application segments.

Dequantised values are kept where natural images keep them (DC within +-1024, the sum of
dequantised AC magnitudes per block at most 600), so the inverse DCT's intermediate sums stay
within 16 bits.
"""
import heapq
import struct

import numpy as np

ZIGZAG = [0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5,
          12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6, 7, 14, 21, 28,
          35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
          58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63]

DC_SYMBOLS = list(range(12))
AC_SYMBOLS = [0x00, 0xF0] + [(r << 4) | s for r in range(16) for s in range(1, 11)]


def _code_lengths(freqs):
    """Huffman code lengths for symbol frequencies, plus a reserved least-frequent symbol that
    takes the all-ones code (T.81 K.2); None when a length exceeds 16."""
    heap = [(f, i, [i]) for i, f in enumerate(freqs + [0.5])]
    heapq.heapify(heap)
    depth = [0] * len(heap)
    uid = len(heap)
    while len(heap) > 1:
        f1, _, s1 = heapq.heappop(heap)
        f2, _, s2 = heapq.heappop(heap)
        for s in s1 + s2:
            depth[s] += 1
        heapq.heappush(heap, (f1 + f2, uid, s1 + s2))
        uid += 1
    if max(depth) > 16:
        return None
    return depth[:-1], depth[-1]


def huffman_table(symbols, rng):
    """(BITS, HUFFVAL, {symbol: (code, length)}) of a canonical table over `symbols`."""
    while True:
        freqs = list(rng.uniform(1.0, 40.0, len(symbols)) ** 2)
        got = _code_lengths(freqs)
        if got is not None:
            break
    lengths = got[0]
    # Canonical order is by length. The reserved symbol would come last (it has the longest code
    # and the highest index) and take the all-ones code, so leaving it out of BITS and HUFFVAL
    # changes no other code and keeps the all-ones code unused.
    order = sorted(range(len(symbols)), key=lambda i: (lengths[i], i))
    bits = [0] * 16
    for i in order:
        bits[lengths[i] - 1] += 1
    huffval = [symbols[i] for i in order]
    codes = {}
    code = 0
    k = 0
    for length in range(1, 17):
        for _ in range(bits[length - 1]):
            codes[huffval[k]] = (code, length)
            code += 1
            k += 1
        code <<= 1
    return bits, huffval, codes


class BitWriter:
    def __init__(self):
        self.out = bytearray()
        self.acc = 0
        self.n = 0

    def put(self, value, length):
        for i in range(length - 1, -1, -1):
            self.acc = (self.acc << 1) | ((value >> i) & 1)
            self.n += 1
            if self.n == 8:
                self.out.append(self.acc)
                if self.acc == 0xFF:
                    self.out.append(0)
                self.acc = 0
                self.n = 0

    def flush(self):
        if self.n:
            self.put((1 << (8 - self.n)) - 1, 8 - self.n)


def _category(v):
    return int(abs(int(v))).bit_length()


def _bits_of(v, size):
    return v if v >= 0 else v + (1 << size) - 1


def _segment(marker, payload, fill=0):
    return b"\xff" * fill + bytes([0xFF, marker]) + struct.pack(">H", len(payload) + 2) + payload


def encode(width, height, sampling, *, seed, restart=0, scans="interleaved", qprec=0, app="jfif",
           ids=None, fill=0, dct_process=0xC0, precision=8, sparse=8):
    """Bytes of a baseline JPEG of `width` x `height` with components sampled as `sampling`
    ([(h, v), ...], 1 or 3 entries). scans: "interleaved" (one scan), "separate" (one scan per
    component) or "split" (component 0, then the others interleaved). app: "jfif", "adobe0",
    "adobe1" or None. qprec 1 writes 16-bit quantisation tables."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nf = len(sampling)
    ids = ids or list(range(1, nf + 1))
    hmax = max(h for h, _ in sampling)
    vmax = max(v for _, v in sampling)
    mcux = -(-width // (8 * hmax))
    mcuy = -(-height // (8 * vmax))
    ntab = 1 if nf == 1 else 2
    qtabs = []
    for _ in range(ntab):
        hi = 400 if qprec else 60
        qtabs.append(rng.integers(1, hi, 64))
    dcs = [huffman_table(DC_SYMBOLS, rng) for _ in range(ntab)]
    acs = [huffman_table(AC_SYMBOLS, rng) for _ in range(ntab)]

    out = bytearray(b"\xff\xd8")
    if app == "jfif":
        out += _segment(0xE0, b"JFIF\x00\x01\x01\x00\x00\x01\x00\x01\x00\x00", fill)
    elif app in ("adobe0", "adobe1"):
        out += _segment(0xEE, b"Adobe\x00" + bytes([100, 0, 0, 0, 0, 0 if app == "adobe0" else 1]), fill)
    out += _segment(0xFE, b"synthetic", fill)
    for t, q in enumerate(qtabs):
        body = bytes([(qprec << 4) | t])
        for k in range(64):
            v = int(q[ZIGZAG[k]])
            body += struct.pack(">H", v) if qprec else bytes([v])
        out += _segment(0xDB, body, fill)
    for t in range(ntab):
        for cls, (bits, vals, _) in ((0, dcs[t]), (1, acs[t])):
            out += _segment(0xC4, bytes([(cls << 4) | t]) + bytes(bits) + bytes(vals), fill)
    if restart:
        out += _segment(0xDD, struct.pack(">H", restart), fill)
    sof = bytes([precision]) + struct.pack(">HH", height, width) + bytes([nf])
    for i, (h, v) in enumerate(sampling):
        sof += bytes([ids[i], (h << 4) | v, 0 if i == 0 else 1])
    out += _segment(dct_process, sof, fill)

    def block(comp):
        q = qtabs[0 if comp == 0 else 1]
        dc = int(round(rng.uniform(-1024, 1016) / q[0]))
        coef = np.zeros(64, np.int64)
        coef[0] = dc
        budget = 600
        for _ in range(int(rng.integers(0, sparse + 1))):
            k = int(rng.integers(1, 64))
            pos = ZIGZAG[k]
            lim = budget // int(q[pos])
            if lim < 1:
                continue
            mag = int(rng.integers(1, min(lim, 1023) + 1))
            coef[pos] = mag if rng.integers(0, 2) else -mag
            budget -= mag * int(q[pos])
        return coef

    def put_block(bw, coef, comp, pred):
        t = 0 if comp == 0 else 1
        dcodes, acodes = dcs[t][2], acs[t][2]
        diff = int(coef[0]) - pred
        s = _category(diff)
        bw.put(*dcodes[s])
        if s:
            bw.put(_bits_of(diff, s), s)
        run = 0
        zz = [int(coef[ZIGZAG[k]]) for k in range(64)]
        last = max([k for k in range(1, 64) if zz[k]] or [0])
        for k in range(1, last + 1):
            if zz[k] == 0:
                run += 1
                continue
            while run > 15:
                bw.put(*acodes[0xF0])
                run -= 16
            s = _category(zz[k])
            bw.put(*acodes[(run << 4) | s])
            bw.put(_bits_of(zz[k], s), s)
            run = 0
        if last < 63:
            bw.put(*acodes[0x00])
        return int(coef[0])

    def scan(comps):
        hdr = bytes([len(comps)])
        for c in comps:
            t = 0 if c == 0 else 1
            hdr += bytes([ids[c], (t << 4) | t])
        hdr += bytes([0, 63, 0])
        seg = _segment(0xDA, hdr, fill)
        bw = BitWriter()
        preds = {c: 0 for c in comps}
        if len(comps) == 1:
            c = comps[0]
            h, v = sampling[c]
            px = -(-width * h // hmax)
            py = -(-height * v // vmax)
            units = [[(c, None)] for _ in range(-(-px // 8) * -(-py // 8))]
        else:
            units = [[(c, None) for c in comps for _ in range(sampling[c][0] * sampling[c][1])]
                     for _ in range(mcux * mcuy)]
        rst = 0
        for u, unit in enumerate(units):
            for c, _ in unit:
                preds[c] = put_block(bw, block(c), c, preds[c])
            if restart and (u + 1) % restart == 0 and u + 1 < len(units):
                bw.flush()
                bw.out += b"\xff" * fill + bytes([0xFF, 0xD0 + rst])
                rst = (rst + 1) % 8
                preds = {c: 0 for c in comps}
        bw.flush()
        return seg + bytes(bw.out)

    if scans == "interleaved" or nf == 1:
        out += scan(list(range(nf)))
    elif scans == "separate":
        for c in range(nf):
            out += scan([c])
    elif scans == "split":
        out += scan([0])
        out += scan(list(range(1, nf)))
    else:
        raise ValueError(scans)
    out += b"\xff" * fill + b"\xff\xd9"
    return bytes(out)


# Sampling layouts of the decoder tests: every upsampling path (none, vertical 2:1, horizontal 2:1,
# 2x2, nearest for 3:1 / 4:1 / 4:2 / 1:4), chroma sampled finer than luma, and mixed factors.
LAYOUTS = {
    "gray": [(1, 1)],
    "444": [(1, 1)] * 3,
    "422": [(2, 1), (1, 1), (1, 1)],
    "420": [(2, 2), (1, 1), (1, 1)],
    "440": [(1, 2), (1, 1), (1, 1)],
    "411": [(4, 1), (1, 1), (1, 1)],
    "chroma_fine": [(1, 1), (2, 2), (2, 2)],
    "31": [(3, 1), (1, 1), (1, 1)],
    "42": [(4, 2), (1, 1), (1, 1)],
    "14": [(1, 4), (1, 1), (1, 1)],
    "mixed": [(2, 2), (2, 1), (1, 2)],
}
SIZES = [(1, 1), (1, 17), (17, 1), (8, 8), (9, 9), (33, 17), (50, 37)]


def cases():
    """(name, encode keyword arguments) of the synthetic decoder cases, in a fixed order."""
    n = 0
    for layout, sampling in LAYOUTS.items():
        for w, h in SIZES:
            for scans in ("interleaved", "separate", "split"):
                if len(sampling) == 1 and scans != "interleaved":
                    continue
                for restart in (0, 1, 3):
                    name = f"{layout}-{w}x{h}-{scans}-r{restart}"
                    yield name, dict(width=w, height=h, sampling=sampling, seed=n, restart=restart, scans=scans,
                                     qprec=n % 2, fill=n % 3)
                    n += 1
