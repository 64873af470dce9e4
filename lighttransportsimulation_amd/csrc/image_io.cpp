// Image input (8-bit PNG / binary PPM / baseline JPEG -> linear RGB) and output (PFM, P3 PPM).
//   input:  the reference decodes with stb_image and converts 8-bit values with
//           pow(c / 255, 2.2) (core/ImageI.cuh:11-37, extern/stb_image.h:1828)
//   output: core/ImageO.cuh:16-78
#include "image_io.h"

#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>

namespace lts {

static bool slurp(const std::string& path, std::vector<uint8_t>& buf) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    buf.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    return true;
}

static inline uint32_t rd_be32(const uint8_t* p) {
    return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | (uint32_t)p[3];
}

static inline int paeth(int a, int b, int c) {
    int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

static bool png_to_rgb8(const std::vector<uint8_t>& b, int& w, int& h, std::vector<uint8_t>& rgb, std::string& err) {
    static const uint8_t magic[8] = {0x89, 'P', 'N', 'G', 0x0d, 0x0a, 0x1a, 0x0a};
    if (b.size() < 8 || std::memcmp(b.data(), magic, 8) != 0) { err = "not a PNG file"; return false; }
    int depth = 0, color = 0, interlace = 0;
    std::vector<uint8_t> z, pal;
    for (size_t at = 8; at + 12 <= b.size();) {
        uint32_t n = rd_be32(&b[at]);
        if (at + 12 + (size_t)n > b.size()) { err = "truncated PNG chunk"; return false; }
        const uint8_t* d = &b[at + 8];
        if (!std::memcmp(&b[at + 4], "IHDR", 4)) {
            if (n != 13) { err = "bad PNG header chunk"; return false; }
            const uint32_t uw = rd_be32(d), uh = rd_be32(d + 4);
            if (uw == 0 || uh == 0 || uw > kMaxImageSide || uh > kMaxImageSide || (uint64_t)uw * uh > kMaxImageTexels) {
                err = "PNG dimensions out of range";
                return false;
            }
            w = (int)uw; h = (int)uh; depth = d[8]; color = d[9]; interlace = d[12];
        } else if (!std::memcmp(&b[at + 4], "PLTE", 4)) {
            pal.assign(d, d + n);
        } else if (!std::memcmp(&b[at + 4], "IDAT", 4)) {
            z.insert(z.end(), d, d + n);
        } else if (!std::memcmp(&b[at + 4], "IEND", 4)) {
            break;
        }
        at += 12 + (size_t)n;
    }
    int spp = color == 0 ? 1 : color == 2 ? 3 : color == 3 ? 1 : color == 4 ? 2 : color == 6 ? 4 : 0;
    if (depth != 8 || interlace || !spp || w <= 0 || h <= 0) { err = "unsupported PNG (need 8-bit, non-interlaced)"; return false; }
    const size_t row = (size_t)w * spp;
    std::vector<uint8_t> raw((row + 1) * (size_t)h);
    uLongf got = (uLongf)raw.size();
    if (uncompress(raw.data(), &got, z.data(), (uLong)z.size()) != Z_OK || got != raw.size()) { err = "PNG inflate failed"; return false; }
    std::vector<uint8_t> pix(row * (size_t)h);
    for (int y = 0; y < h; y++) {
        const uint8_t* in = &raw[(size_t)y * (row + 1)];
        uint8_t* out = &pix[(size_t)y * row];
        const uint8_t* up = y ? out - row : nullptr;
        for (size_t i = 0; i < row; i++) {
            int a = i >= (size_t)spp ? out[i - spp] : 0;
            int u = up ? up[i] : 0;
            int c = (up && i >= (size_t)spp) ? up[i - spp] : 0;
            int pred;
            switch (in[0]) {
                case 0: pred = 0; break;
                case 1: pred = a; break;
                case 2: pred = u; break;
                case 3: pred = (a + u) >> 1; break;
                case 4: pred = paeth(a, u, c); break;
                default: err = "bad PNG row filter"; return false;
            }
            out[i] = (uint8_t)(in[1 + i] + pred);
        }
    }
    rgb.resize((size_t)w * h * 3);
    for (size_t i = 0; i < (size_t)w * h; i++) {
        const uint8_t* p = &pix[i * spp];
        if (color == 3) {
            if ((size_t)p[0] * 3 + 2 >= pal.size()) { err = "PNG palette index out of range"; return false; }
            std::memcpy(&rgb[3 * i], &pal[3 * (size_t)p[0]], 3);
        } else if (spp <= 2) {
            rgb[3 * i] = rgb[3 * i + 1] = rgb[3 * i + 2] = p[0];
        } else {
            std::memcpy(&rgb[3 * i], p, 3);
        }
    }
    return true;
}

static bool ppm_to_rgb8(const std::vector<uint8_t>& b, int& w, int& h, std::vector<uint8_t>& rgb, std::string& err) {
    size_t at = 2;
    long v[3];
    for (int k = 0; k < 3; k++) {
        for (;;) {
            while (at < b.size() && std::isspace(b[at])) at++;
            if (at < b.size() && b[at] == '#') { while (at < b.size() && b[at] != '\n') at++; continue; }
            break;
        }
        if (at >= b.size() || !std::isdigit(b[at])) { err = "bad PPM header"; return false; }
        v[k] = 0;
        while (at < b.size() && std::isdigit(b[at])) {
            if (v[k] < (1L << 40)) v[k] = v[k] * 10 + (b[at] - '0');  // larger values are rejected below
            at++;
        }
    }
    at++;  // single whitespace after maxval
    if (v[0] <= 0 || v[1] <= 0 || v[0] > (long)kMaxImageSide || v[1] > (long)kMaxImageSide ||
        (uint64_t)v[0] * (uint64_t)v[1] > kMaxImageTexels) { err = "PPM dimensions out of range"; return false; }
    w = (int)v[0]; h = (int)v[1];
    if (v[2] != 255 || at + (size_t)w * h * 3 > b.size()) { err = "unsupported PPM (need P6, maxval 255)"; return false; }
    rgb.assign(b.begin() + (long)at, b.begin() + (long)(at + (size_t)w * h * 3));
    return true;
}

bool load_image_linear(const std::string& path, int& w, int& h, std::vector<float>& rgb, std::string& err) {
    std::vector<uint8_t> file, px;
    if (!slurp(path, file)) { err = "cannot open image " + path; return false; }
    bool ok;
    if (file.size() > 2 && file[0] == 'P' && file[1] == '6') ok = ppm_to_rgb8(file, w, h, px, err);
    else if (file.size() > 2 && file[0] == 0xFF && file[1] == 0xD8) ok = jpeg_to_rgb8(file, w, h, px, err);
    else ok = png_to_rgb8(file, w, h, px, err);
    if (!ok) { err = path + ": " + err; return false; }
    float lin[256];
    for (int c = 0; c < 256; c++) lin[c] = (float)(std::pow((double)(c / 255.0f), (double)2.2f) * (double)1.0f);
    rgb.resize((size_t)w * h * 3);
    // bottom row first (core/ImageI.cuh:26-35)
    size_t o = 0;
    for (int y = h - 1; y >= 0; y--)
        for (int x = 0; x < w; x++)
            for (int c = 0; c < 3; c++) rgb[o++] = lin[px[3 * ((size_t)y * w + x) + c]];
    return true;
}

bool write_pfm(const std::string& path, const float* rgb, int w, int h) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    bool ok = std::fprintf(f, "PF\n%d %d\n%f\n", w, h, -1.0f) > 0;
    for (int y = 0; ok && y < h; y++) ok = std::fwrite(rgb + (size_t)y * w * 3, sizeof(float), (size_t)w * 3, f) == (size_t)w * 3;
    std::fclose(f);
    return ok;
}

// toRGB (core/ImageO.cuh:52): int(pow(clamp(x), 1 / 2.2) * 255 + 0.5)
static int to_byte(float x) {
    float c = (x < 0) ? 0 : ((x > 1) ? 1 : x);
    return int((float)std::pow((double)c, (double)(1 / 2.2f)) * 255 + 0.5);
}

bool write_ppm(const std::string& path, const float* rgb, int w, int h) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    bool ok = std::fprintf(f, "P3\n%d %d\n%d\n", w, h, 255) > 0;
    for (size_t i = 0; ok && i < (size_t)w * h; i++)
        ok = std::fprintf(f, "%d %d %d ", to_byte(rgb[3 * i]), to_byte(rgb[3 * i + 1]), to_byte(rgb[3 * i + 2])) > 0;
    std::fclose(f);
    return ok;
}

// Shared-exponent encoding of the Radiance picture format (G. Ward, "Real Pixels", Graphics Gems
// II, 1991): e = the exponent of the largest channel as frexp gives it, each channel stored as
// floor(c * 256 / 2^e), and e + 128; values below 1e-32 (and negative ones) encode as zero.
bool write_hdr(const std::string& path, const float* rgb, int w, int h) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    bool ok = std::fprintf(f, "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %d +X %d\n", h, w) > 0;
    std::vector<unsigned char> row(4 * (size_t)w);
    for (int y = 0; ok && y < h; y++) {
        for (int x = 0; x < w; x++) {
            const float* c = rgb + 3 * ((size_t)y * w + x);
            const float r = std::fmax(c[0], 0.0f), g = std::fmax(c[1], 0.0f), b = std::fmax(c[2], 0.0f);
            const float m = std::fmax(r, std::fmax(g, b));
            unsigned char* o = &row[4 * (size_t)x];
            if (!(m >= 1e-32f)) {
                o[0] = o[1] = o[2] = o[3] = 0;
                continue;
            }
            int e;
            const float scale = (float)std::frexp(m, &e) * 256.0f / m;
            o[0] = (unsigned char)(r * scale);
            o[1] = (unsigned char)(g * scale);
            o[2] = (unsigned char)(b * scale);
            o[3] = (unsigned char)(e + 128);
        }
        ok = std::fwrite(row.data(), 1, row.size(), f) == row.size();
    }
    std::fclose(f);
    return ok;
}

// PNG (ISO/IEC 15948): signature, IHDR (8-bit RGB, no interlace), one zlib IDAT stream of rows
// with filter byte 0, IEND; chunk CRCs from zlib's crc32.
bool write_png(const std::string& path, const float* rgb, int w, int h) {
    std::vector<unsigned char> raw((size_t)h * (1 + 3 * (size_t)w));
    for (int y = 0; y < h; y++) {
        unsigned char* r = &raw[(size_t)y * (1 + 3 * (size_t)w)];
        r[0] = 0;
        for (int i = 0; i < 3 * w; i++) r[1 + i] = (unsigned char)to_byte(rgb[3 * (size_t)y * w + i]);
    }
    uLongf zlen = compressBound((uLong)raw.size());
    std::vector<unsigned char> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) != Z_OK) return false;
    z.resize(zlen);
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    auto be32 = [](unsigned char* p, uint32_t v) { p[0] = v >> 24; p[1] = v >> 16; p[2] = v >> 8; p[3] = v; };
    bool ok = true;
    auto chunk = [&](const char* type, const unsigned char* data, size_t n) {
        unsigned char head[8];
        be32(head, (uint32_t)n);
        std::memcpy(head + 4, type, 4);
        uLong crc = crc32(0L, head + 4, 4);
        if (n) crc = crc32(crc, data, (uInt)n);
        unsigned char tail[4];
        be32(tail, (uint32_t)crc);
        ok = ok && std::fwrite(head, 1, 8, f) == 8 && (n == 0 || std::fwrite(data, 1, n, f) == n) &&
             std::fwrite(tail, 1, 4, f) == 4;
    };
    static const unsigned char sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    ok = std::fwrite(sig, 1, 8, f) == 8;
    unsigned char ihdr[13];
    be32(ihdr, (uint32_t)w);
    be32(ihdr + 4, (uint32_t)h);
    ihdr[8] = 8; ihdr[9] = 2; ihdr[10] = 0; ihdr[11] = 0; ihdr[12] = 0;
    chunk("IHDR", ihdr, 13);
    chunk("IDAT", z.data(), z.size());
    chunk("IEND", nullptr, 0);
    std::fclose(f);
    return ok;
}

}  // namespace lts
