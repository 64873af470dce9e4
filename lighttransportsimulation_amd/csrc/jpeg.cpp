// Baseline JPEG -> 8-bit RGB, for image textures.
//
// The reference reads texture files with stb_image 2.27 (vendored as extern/stb_image.h, called as
// stbi_loadf at core/ImageI.cuh:14). This decoder is written from ITU-T T.81 (sequential DCT,
// Huffman coding, 8-bit samples) and JFIF 1.02 and gives that library's bytes on every file it
// accepts. Three steps must agree with the library bit for bit, so they use its integer arithmetic:
//   - the 8-point inverse DCT: the Loeffler-Ligtenberg-Moschytz factorisation with the Independent
//     JPEG Group's 12-bit constants ("islow"), applied here as an 8x8 integer matrix holding that
//     factorisation's net coefficient for each input. The column pass keeps 2 extra bits (round,
//     >> 10); the row pass rounds, adds the 128 level shift and drops the remaining 17 bits
//     (extern/stb_image.h:2381-2474). Integer sums are exact (and wrap alike), so the matrix form
//     returns the factored form's values.
//   - chroma upsampling: a centred triangle filter (3/4 nearer sample, 1/4 farther) down rows for
//     2:1 vertical ratios, across columns for 2:1 horizontal ones, both for 2x2; the nearest sample
//     for every other ratio (extern/stb_image.h:3401-3475, 3591-3600; rows walked as at 3888-3906).
//     The horizontal-only filter weights the last output pair's first sample 3:1 towards the
//     second-to-last input (extern/stb_image.h:3441); that is kept.
//   - YCbCr -> RGB: JFIF's coefficients in 20-bit fixed point with the Cb share of G rounded down
//     to a multiple of 2^16 (extern/stb_image.h:3604-3627).
// The library's SSE2 inverse DCT holds intermediate sums in 16 bits where its scalar form (and this
// one) holds 32; the forms agree while those sums fit, which they do for blocks of natural images.
//
// Errors rather than output: progressive, lossless and arithmetic-coded files; sample precision
// other than 8 bits; four-component (CMYK / YCCK) images; tables used before they are defined; a
// component that no scan covers; a stream that ends inside a scan, holds a code that is not in its
// Huffman table or lacks a restart marker where one is due (the library fills in zero bits or
// leaves blocks undecoded there).
#include "image_io.h"

#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace lts {
namespace {

struct JpegError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
[[noreturn]] void fail(const char* what) { throw JpegError(what); }

// Row-major position within the 8x8 block of the k-th coefficient in zigzag order (T.81 Figure A.6).
constexpr uint8_t kZigzag[64] = {
     0,  1,  8, 16,  9,  2,  3, 10, 17, 24, 32, 25, 18, 11,  4,  5,
    12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13,  6,  7, 14, 21, 28,
    35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
    58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

// A constant in 12-bit fixed point, rounded as (int)(x * 4096 + 0.5): toward zero after adding one
// half, so fix12(-a) is not always -fix12(a).
int fix12(float x) { return (int)((double)(x * 4096) + 0.5); }

// The 8-point inverse DCT as m[output][input], built from the factorisation's constants.
struct Idct8 {
    int32_t m[8][8];
    Idct8() {
        const int a = fix12(0.5411961f), b = fix12(-1.847759065f), c = fix12(0.765366865f);
        const int r = fix12(1.175875602f);
        const int w1 = fix12(1.501321110f), w3 = fix12(3.072711026f), w5 = fix12(2.053119869f),
                  w7 = fix12(0.298631336f);
        const int z1 = fix12(-0.899976223f), z2 = fix12(-2.562915447f), z3 = fix12(-1.961570560f),
                  z4 = fix12(-0.390180644f);
        // outputs 0..3 from the even inputs 0, 2, 4, 6
        const int even[4][4] = {{4096, a + c, 4096, a},
                                {4096, a, -4096, a + b},
                                {4096, -a, -4096, -(a + b)},
                                {4096, -(a + c), 4096, -a}};
        // outputs 0..3 from the odd inputs 1, 3, 5, 7; outputs 7..4 take the same even part and
        // subtract this one
        const int odd[4][4] = {{w1 + r + z1 + z4, r, r + z4, r + z1},
                               {r, w3 + r + z2 + z3, r + z2, r + z3},
                               {r + z4, r + z2, w5 + r + z2 + z4, r},
                               {r + z1, r + z3, r, w7 + r + z1 + z3}};
        for (int k = 0; k < 4; k++)
            for (int j = 0; j < 4; j++) {
                m[k][2 * j] = even[k][j];
                m[k][2 * j + 1] = odd[k][j];
                m[7 - k][2 * j] = even[k][j];
                m[7 - k][2 * j + 1] = -odd[k][j];
            }
    }
    // Dequantised coefficients (row-major) -> level-shifted samples at out[row * stride + col].
    void block(const int16_t* in, uint8_t* out, size_t stride) const {
        int32_t mid[64];
        for (int col = 0; col < 8; col++)
            for (int k = 0; k < 8; k++) {
                uint32_t s = 512u;
                for (int n = 0; n < 8; n++) s += (uint32_t)m[k][n] * (uint32_t)(int32_t)in[n * 8 + col];
                mid[k * 8 + col] = (int32_t)s >> 10;
            }
        for (int row = 0; row < 8; row++)
            for (int k = 0; k < 8; k++) {
                uint32_t s = 65536u + (128u << 17);
                for (int n = 0; n < 8; n++) s += (uint32_t)m[k][n] * (uint32_t)mid[row * 8 + n];
                const int v = (int32_t)s >> 17;
                out[row * stride + k] = (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v);
            }
    }
};

// Canonical Huffman code (T.81 Annex C) with the decoding tables of F.2.2.3.
struct Huffman {
    bool defined = false;
    int count = 0;
    uint8_t symbol[256] = {};
    int32_t maxcode[17] = {};  // largest code of each length, -1 when the length has none
    int32_t mincode[17] = {};
    int32_t first[17] = {};    // index in `symbol` of the first code of each length
};

void build_huffman(Huffman& h, const uint8_t bits[16]) {
    int code = 0, k = 0;
    for (int len = 1; len <= 16; len++) {
        const int n = bits[len - 1];
        h.first[len] = k;
        h.mincode[len] = code;
        code += n;
        k += n;
        if (n && code > (1 << len)) fail("bad Huffman code lengths");
        h.maxcode[len] = n ? code - 1 : -1;
        code <<= 1;
    }
    if (k > 256) fail("bad Huffman table size");
    h.count = k;
    h.defined = true;
}

// Entropy-coded bytes of one scan: 0xFF 0x00 is a data byte 0xFF; 0xFF followed by anything else
// is a marker and ends the data.
struct BitReader {
    const std::vector<uint8_t>* b;
    size_t pos;
    uint32_t cur = 0;
    int left = 0;
    BitReader(const std::vector<uint8_t>& bytes, size_t at) : b(&bytes), pos(at) {}
    int bit() {
        if (left == 0) {
            const std::vector<uint8_t>& d = *b;
            if (pos >= d.size()) fail("JPEG data ends inside a scan");
            uint8_t v = d[pos];
            if (v == 0xFF) {
                size_t q = pos + 1;
                while (q < d.size() && d[q] == 0xFF) q++;
                if (q >= d.size() || d[q] != 0) fail("JPEG scan data ends before its last block");
                pos = q + 1;
            } else {
                pos++;
            }
            cur = v;
            left = 8;
        }
        left--;
        return (int)(cur >> left) & 1;
    }
    int bits(int n) {
        int v = 0;
        for (int i = 0; i < n; i++) v = (v << 1) | bit();
        return v;
    }
    // RECEIVE and EXTEND of T.81 F.2.2.1
    int receive_extend(int n) {
        const int v = bits(n);
        return v < (1 << (n - 1)) ? v - (1 << n) + 1 : v;
    }
    int decode(const Huffman& h) {
        int code = 0;
        for (int len = 1; len <= 16; len++) {
            code = (code << 1) | bit();
            if (code <= h.maxcode[len]) return h.symbol[h.first[len] + code - h.mincode[len]];
        }
        fail("JPEG code not in its Huffman table");
    }
};

struct Component {
    int id = 0, h = 1, v = 1, tq = 0;
    int td = 0, ta = 0;  // tables named by the current scan
    int pred = 0;        // DC prediction
    int px = 0, py = 0;  // samples that carry image data (T.81 A.1.1)
    size_t stride = 0, rows = 0;
    std::vector<uint8_t> plane;
    bool scanned = false;
};

class Decoder {
public:
    explicit Decoder(const std::vector<uint8_t>& bytes) : b(bytes) {}

    void run(int& w, int& h, std::vector<uint8_t>& rgb) {
        if (marker(false) != 0xD8) fail("not a JPEG (no SOI marker)");
        int m = marker(true);
        while (m != 0xC0 && m != 0xC1) {
            if (m == 0xC2) fail("progressive JPEG is not supported");
            if (is_sof(m)) fail("lossless or arithmetic-coded JPEG is not supported");
            segment(m);
            m = marker(true);
        }
        frame();
        m = marker(false);
        while (m != 0xD9) {
            if (m == 0xDA) {
                scan();
                m = after_scan();
                continue;
            }
            if (m == 0xDC) {  // DNL: must repeat the frame's height
                if (u16() != 4 || u16() != (int)Y) fail("bad DNL segment");
            } else {
                if (is_sof(m)) fail("second frame header in a JPEG");
                segment(m);
            }
            m = marker(false);
        }
        for (const Component& c : comp)
            if (!c.scanned) fail("JPEG component missing from its scans");
        output(rgb);
        w = (int)X;
        h = (int)Y;
    }

private:
    const std::vector<uint8_t>& b;
    size_t p = 0;
    uint16_t qt[4][64] = {};
    bool qtDefined[4] = {};
    Huffman dc[4], ac[4];
    int restart = 0;
    bool jfif = false;
    int adobeTransform = -1;
    std::vector<Component> comp;
    uint32_t X = 0, Y = 0;
    int hmax = 1, vmax = 1, mcux = 0, mcuy = 0;
    bool rgbIds = false;
    Idct8 idct;

    static bool is_sof(int m) { return m >= 0xC0 && m <= 0xCF && m != 0xC4 && m != 0xC8 && m != 0xCC; }

    int u8() {
        if (p >= b.size()) fail("truncated JPEG");
        return b[p++];
    }
    int u16() {
        const int hi = u8();
        return hi << 8 | u8();
    }
    // Next marker code: 0xFF, optional 0xFF fill bytes, the code. With skipJunk, bytes that are
    // not a marker are passed over (as between header segments); otherwise they are an error.
    int marker(bool skipJunk) {
        for (;;) {
            const int x = u8();
            if (x != 0xFF) {
                if (skipJunk) continue;
                fail("expected a JPEG marker");
            }
            int c = u8();
            while (c == 0xFF) c = u8();
            if (c != 0) return c;
            if (!skipJunk) fail("expected a JPEG marker");
        }
    }

    void segment(int m) {
        if (m == 0xDB) return define_quant();
        if (m == 0xC4) return define_huffman();
        if (m == 0xDD) {
            if (u16() != 4) fail("bad DRI segment");
            restart = u16();
            return;
        }
        if ((m >= 0xE0 && m <= 0xEF) || m == 0xFE) {
            const int L = u16();
            if (L < 2) fail("bad JPEG segment length");
            const size_t n = (size_t)L - 2;
            if (p + n > b.size()) fail("truncated JPEG");
            const uint8_t* d = b.data() + p;
            if (m == 0xE0 && n >= 5 && std::memcmp(d, "JFIF\0", 5) == 0) jfif = true;
            if (m == 0xEE && n >= 12 && std::memcmp(d, "Adobe\0", 6) == 0) adobeTransform = d[11];
            p += n;
            return;
        }
        fail("unsupported JPEG marker");
    }

    void define_quant() {
        int L = u16() - 2;
        while (L > 0) {
            const int q = u8(), prec = q >> 4, t = q & 15;
            if (prec > 1) fail("bad DQT precision");
            if (t > 3) fail("bad DQT table id");
            for (int i = 0; i < 64; i++) qt[t][kZigzag[i]] = (uint16_t)(prec ? u16() : u8());
            qtDefined[t] = true;
            L -= prec ? 129 : 65;
        }
        if (L != 0) fail("bad DQT length");
    }

    void define_huffman() {
        int L = u16() - 2;
        while (L > 0) {
            const int q = u8(), tc = q >> 4, th = q & 15;
            if (tc > 1 || th > 3) fail("bad DHT header");
            uint8_t bits[16];
            int n = 0;
            for (int i = 0; i < 16; i++) n += bits[i] = (uint8_t)u8();
            L -= 17;
            Huffman& h = tc ? ac[th] : dc[th];
            h = Huffman();
            build_huffman(h, bits);
            for (int i = 0; i < n; i++) h.symbol[i] = (uint8_t)u8();
            L -= n;
        }
        if (L != 0) fail("bad DHT length");
    }

    void frame() {
        const int L = u16();
        if (L < 11) fail("bad SOF length");
        if (u8() != 8) fail("only 8-bit JPEG samples are supported");
        Y = (uint32_t)u16();
        X = (uint32_t)u16();
        if (Y == 0) fail("JPEG without a height in its frame header");
        if (X == 0) fail("JPEG of width 0");
        if (X > kMaxImageSide || Y > kMaxImageSide || (uint64_t)X * Y > kMaxImageTexels)
            fail("JPEG dimensions out of range");
        const int nf = u8();
        if (nf == 4) fail("four-component (CMYK) JPEG is not supported");
        if (nf != 1 && nf != 3) fail("bad JPEG component count");
        if (L != 8 + 3 * nf) fail("bad SOF length");
        comp.assign(nf, Component());
        int rgbMatches = 0;
        for (int i = 0; i < nf; i++) {
            Component& c = comp[i];
            c.id = u8();
            if (nf == 3 && c.id == "RGB"[i]) rgbMatches++;
            const int s = u8();
            c.h = s >> 4;
            c.v = s & 15;
            c.tq = u8();
            if (c.h < 1 || c.h > 4) fail("bad JPEG horizontal sampling factor");
            if (c.v < 1 || c.v > 4) fail("bad JPEG vertical sampling factor");
            if (c.tq > 3) fail("bad JPEG quantisation table id");
            hmax = std::max(hmax, c.h);
            vmax = std::max(vmax, c.v);
        }
        rgbIds = rgbMatches == 3;
        for (const Component& c : comp)
            if (hmax % c.h || vmax % c.v) fail("JPEG sampling factors are not integer ratios");
        mcux = (int)((X + 8 * hmax - 1) / (8 * hmax));
        mcuy = (int)((Y + 8 * vmax - 1) / (8 * vmax));
        for (Component& c : comp) {
            c.px = (int)((X * c.h + hmax - 1) / hmax);
            c.py = (int)((Y * c.v + vmax - 1) / vmax);
            c.stride = (size_t)mcux * c.h * 8;
            c.rows = (size_t)mcuy * c.v * 8;
            c.plane.assign(c.stride * c.rows, 0);
        }
    }

    void decode_block(BitReader& br, Component& c, int16_t* blk) {
        const int t = br.decode(dc[c.td]);
        if (t > 15) fail("bad JPEG DC code");
        // the prediction wraps modulo 2^32 rather than overflowing (only hostile files get there; the
        // dequantised value keeps its low 16 bits either way)
        c.pred = (int32_t)((uint32_t)c.pred + (uint32_t)(t ? br.receive_extend(t) : 0));
        const uint16_t* q = qt[c.tq];
        std::memset(blk, 0, 64 * sizeof(int16_t));
        blk[0] = (int16_t)(uint16_t)((int64_t)c.pred * q[0]);
        for (int k = 1; k < 64;) {
            const int rs = br.decode(ac[c.ta]);
            const int run = rs >> 4, size = rs & 15;
            if (size == 0) {
                if (run != 15) break;  // end of block
                k += 16;
                continue;
            }
            k += run;
            if (k > 63) fail("JPEG coefficient run past the end of its block");
            const int z = kZigzag[k++];
            blk[z] = (int16_t)(uint16_t)(br.receive_extend(size) * (int)q[z]);
        }
    }

    void scan() {
        const int Ls = u16(), ns = u8();
        if (ns < 1 || ns > 4 || ns > (int)comp.size()) fail("bad SOS component count");
        if (Ls != 6 + 2 * ns) fail("bad SOS length");
        std::vector<int> order(ns);
        for (int i = 0; i < ns; i++) {
            const int id = u8(), q = u8();
            int k = 0;
            while (k < (int)comp.size() && comp[k].id != id) k++;
            if (k == (int)comp.size()) fail("SOS names an unknown JPEG component");
            comp[k].td = q >> 4;
            comp[k].ta = q & 15;
            if (comp[k].td > 3 || comp[k].ta > 3) fail("bad SOS table id");
            if (!dc[comp[k].td].defined || !ac[comp[k].ta].defined) fail("JPEG scan uses an undefined Huffman table");
            if (!qtDefined[comp[k].tq]) fail("JPEG scan uses an undefined quantisation table");
            order[i] = k;
        }
        const int ss = u8();
        u8();  // end of spectral selection: 63 in sequential files, not checked
        const int ahal = u8();
        if (ss != 0 || ahal != 0) fail("bad SOS spectral selection for a sequential JPEG");

        for (Component& c : comp) c.pred = 0;
        BitReader br(b, p);
        int16_t blk[64];
        const bool single = ns == 1;
        const int bw = single ? (comp[order[0]].px + 7) / 8 : mcux;
        const int bh = single ? (comp[order[0]].py + 7) / 8 : mcuy;
        const long units = (long)bw * bh;
        int todo = restart ? restart : INT_MAX;
        for (long u = 0; u < units; u++) {
            const int ux = (int)(u % bw), uy = (int)(u / bw);
            if (single) {
                Component& c = comp[order[0]];
                decode_block(br, c, blk);
                idct.block(blk, &c.plane[(size_t)uy * 8 * c.stride + (size_t)ux * 8], c.stride);
            } else {
                for (int k : order) {
                    Component& c = comp[k];
                    for (int y = 0; y < c.v; y++)
                        for (int x = 0; x < c.h; x++) {
                            decode_block(br, c, blk);
                            const size_t r = (size_t)(uy * c.v + y) * 8, col = (size_t)(ux * c.h + x) * 8;
                            idct.block(blk, &c.plane[r * c.stride + col], c.stride);
                        }
                }
            }
            if (--todo <= 0 && u + 1 < units) {
                // the rest of the current byte is padding; the next marker must be RSTn
                p = br.pos;
                const int m = next_marker_in_data();
                if (m < 0xD0 || m > 0xD7) fail("JPEG restart marker missing");
                br = BitReader(b, p);
                for (Component& c : comp) c.pred = 0;
                todo = restart;
            }
        }
        for (int k : order) comp[k].scanned = true;
        p = br.pos;
    }

    // The next marker after entropy-coded bytes (data bytes and stuffed zeros are passed over).
    int next_marker_in_data() {
        for (;;) {
            if (p >= b.size()) fail("JPEG ends without EOI");
            if (b[p++] != 0xFF) continue;
            size_t q = p;
            while (q < b.size() && b[q] == 0xFF) q++;
            if (q >= b.size()) fail("JPEG ends without EOI");
            p = q + 1;
            if (b[q] != 0) return b[q];
        }
    }

    int after_scan() {
        int m = next_marker_in_data();
        while (m >= 0xD0 && m <= 0xD7) m = next_marker_in_data();
        return m;
    }

    // One output row of component c at full resolution into `out` (at least hs * ceil(X / hs)
    // entries), or a pointer into the plane when no resampling is needed.
    const uint8_t* upsample(const Component& c, uint32_t j, uint8_t* out) const {
        const int hs = hmax / c.h, vs = vmax / c.v;
        const int w = (int)((X + hs - 1) / hs);
        const int nearRow = (int)(j / vs);
        const uint8_t* n = &c.plane[(size_t)nearRow * c.stride];
        if (hs == 1 && vs == 1) return n;
        if (vs == 2 && hs <= 2) {
            const int farRow = (j & 1) ? std::min(nearRow + 1, c.py - 1) : std::max(nearRow - 1, 0);
            const uint8_t* f = &c.plane[(size_t)farRow * c.stride];
            if (hs == 1) {
                for (int i = 0; i < w; i++) out[i] = (uint8_t)((3 * n[i] + f[i] + 2) >> 2);
                return out;
            }
            if (w == 1) {
                out[0] = out[1] = (uint8_t)((3 * n[0] + f[0] + 2) >> 2);
                return out;
            }
            int t1 = 3 * n[0] + f[0];
            out[0] = (uint8_t)((t1 + 2) >> 2);
            for (int i = 1; i < w; i++) {
                const int t0 = t1;
                t1 = 3 * n[i] + f[i];
                out[2 * i - 1] = (uint8_t)((3 * t0 + t1 + 8) >> 4);
                out[2 * i] = (uint8_t)((3 * t1 + t0 + 8) >> 4);
            }
            out[2 * w - 1] = (uint8_t)((t1 + 2) >> 2);
            return out;
        }
        if (vs == 1 && hs == 2) {
            if (w == 1) {
                out[0] = out[1] = n[0];
                return out;
            }
            out[0] = n[0];
            out[1] = (uint8_t)((3 * n[0] + n[1] + 2) >> 2);
            for (int i = 1; i < w - 1; i++) {
                out[2 * i] = (uint8_t)((3 * n[i] + n[i - 1] + 2) >> 2);
                out[2 * i + 1] = (uint8_t)((3 * n[i] + n[i + 1] + 2) >> 2);
            }
            out[2 * w - 2] = (uint8_t)((3 * n[w - 2] + n[w - 1] + 2) >> 2);
            out[2 * w - 1] = n[w - 1];
            return out;
        }
        for (int i = 0; i < w; i++)
            for (int k = 0; k < hs; k++) out[i * hs + k] = n[i];
        return out;
    }

    void output(std::vector<uint8_t>& rgb) const {
        rgb.assign((size_t)X * Y * 3, 0);
        const size_t nc = comp.size();
        std::vector<std::vector<uint8_t>> line(nc, std::vector<uint8_t>((size_t)X + 4));
        const bool asRgb = nc == 3 && (rgbIds || (adobeTransform == 0 && !jfif));
        const int kr = fix20(1.40200f), kgr = -fix20(0.71414f), kgb = -fix20(0.34414f), kb = fix20(1.77200f);
        for (uint32_t j = 0; j < Y; j++) {
            const uint8_t* row[3];
            for (size_t k = 0; k < nc; k++) row[k] = upsample(comp[k], j, line[k].data());
            uint8_t* o = &rgb[(size_t)j * X * 3];
            for (uint32_t i = 0; i < X; i++, o += 3) {
                if (nc == 1) {
                    o[0] = o[1] = o[2] = row[0][i];
                } else if (asRgb) {
                    o[0] = row[0][i];
                    o[1] = row[1][i];
                    o[2] = row[2][i];
                } else {
                    const int y = (row[0][i] << 20) + (1 << 19);
                    const int cb = row[1][i] - 128, cr = row[2][i] - 128;
                    const int r = (y + cr * kr) >> 20;
                    const int g = (y + cr * kgr + ((cb * kgb) & ~0xFFFF)) >> 20;
                    const int bl = (y + cb * kb) >> 20;
                    o[0] = clamp8(r);
                    o[1] = clamp8(g);
                    o[2] = clamp8(bl);
                }
            }
        }
    }

    // JFIF conversion coefficient in 20-bit fixed point: a 12-bit rounding shifted up by 8.
    static int fix20(float x) { return (int)(x * 4096.0f + 0.5f) << 8; }
    static uint8_t clamp8(int v) { return (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v); }
};

}  // namespace

bool jpeg_to_rgb8(const std::vector<uint8_t>& bytes, int& w, int& h, std::vector<uint8_t>& rgb, std::string& err) {
    try {
        Decoder d(bytes);
        d.run(w, h, rgb);
        return true;
    } catch (const JpegError& e) {
        err = e.what();
        return false;
    } catch (const std::bad_alloc&) {
        err = "out of memory decoding a JPEG";
        return false;
    }
}

}  // namespace lts
