// ORACLE — test infrastructure only (see ref.h).
//
// Restatement of the reference's scene ingest:
//   app.params / .scene / .subject grammar ... core/SceneLoader.cuh:33-360, core/FileIO.cuh:234-257
//   .bin meshes ............................ core/FileIO.cuh:170-206, shapes/Triangle.cu:20-43
//   material / texture construction ........ core/SceneLoader_kernels.cu:8-50, SceneLoader.cuh:362-391
//   lights (one per emissive triangle) ..... SceneLoader_kernels.cu:52-111 (fixes: global light order)
//   images + mip pyramid .................... core/ImageI.cuh:11-37, textures/Mipmap_kernel.cu:9-73
//                                             (JPEG through the reference's stb_image, see decodeJPEG)
//   camera + Gaussian filter table ......... rendering/Camera.cu:9-43, rendering/Filter.cuh:32-99
#include "scene.h"
#include "oracle.h"
#include <dlfcn.h>
#include <zlib.h>
#include <cstdio>
#include <fstream>
#include <sstream>

namespace orc {

// ---------------------------------------------------------------- images
static bool readFile(const std::string& path, std::vector<unsigned char>& buf) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    buf.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    return true;
}

static uint32_t be32(const unsigned char* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

// 8-bit PNG (gray, RGB, palette, gray+alpha, RGBA; not interlaced) -> RGB bytes.
static bool decodePNG(const std::vector<unsigned char>& b, int& w, int& h, std::vector<unsigned char>& rgb, std::string& err) {
    static const unsigned char sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (b.size() < 8 || memcmp(b.data(), sig, 8) != 0) { err = "not a PNG"; return false; }
    size_t pos = 8;
    int depth = 0, ctype = 0, interlace = 0;
    std::vector<unsigned char> idat, plte;
    while (pos + 8 <= b.size()) {
        uint32_t len = be32(&b[pos]);
        std::string type((const char*)&b[pos + 4], 4);
        if (pos + 12 + len > b.size()) { err = "truncated PNG"; return false; }
        const unsigned char* d = &b[pos + 8];
        if (type == "IHDR") { w = (int)be32(d); h = (int)be32(d + 4); depth = d[8]; ctype = d[9]; interlace = d[12]; }
        else if (type == "PLTE") plte.assign(d, d + len);
        else if (type == "IDAT") idat.insert(idat.end(), d, d + len);
        else if (type == "IEND") break;
        pos += 12 + len;
    }
    if (depth != 8 || interlace != 0) { err = "only 8-bit non-interlaced PNG is supported"; return false; }
    int ch = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : ctype == 6 ? 4 : 0;
    if (!ch) { err = "unsupported PNG colour type"; return false; }
    size_t stride = (size_t)w * ch;
    std::vector<unsigned char> raw((stride + 1) * h);
    uLongf rawLen = raw.size();
    if (uncompress(raw.data(), &rawLen, idat.data(), idat.size()) != Z_OK || rawLen != raw.size()) { err = "PNG inflate failed"; return false; }
    std::vector<unsigned char> img(stride * h), prev(stride, 0);
    for (int y = 0; y < h; y++) {
        unsigned char ft = raw[y * (stride + 1)];
        unsigned char* cur = &img[y * stride];
        const unsigned char* src = &raw[y * (stride + 1) + 1];
        for (size_t x = 0; x < stride; x++) {
            int a = x >= (size_t)ch ? cur[x - ch] : 0, up = prev[x], c = x >= (size_t)ch ? prev[x - ch] : 0;
            int v = src[x];
            switch (ft) {
                case 0: break;
                case 1: v += a; break;
                case 2: v += up; break;
                case 3: v += (a + up) >> 1; break;
                case 4: { int p = a + up - c, pa = abs(p - a), pb = abs(p - up), pc = abs(p - c);
                          v += (pa <= pb && pa <= pc) ? a : (pb <= pc ? up : c); break; }
                default: err = "bad PNG filter"; return false;
            }
            cur[x] = (unsigned char)v;
        }
        memcpy(prev.data(), cur, stride);
    }
    rgb.resize((size_t)w * h * 3);
    for (size_t i = 0; i < (size_t)w * h; i++) {
        const unsigned char* p = &img[i * ch];
        unsigned char r, g, bb;
        if (ctype == 3) { int k = p[0]; r = plte[3 * k]; g = plte[3 * k + 1]; bb = plte[3 * k + 2]; }
        else if (ch <= 2) { r = g = bb = p[0]; }
        else { r = p[0]; g = p[1]; bb = p[2]; }
        rgb[3 * i] = r; rgb[3 * i + 1] = g; rgb[3 * i + 2] = bb;
    }
    return true;
}

static bool decodePPM(const std::vector<unsigned char>& b, int& w, int& h, std::vector<unsigned char>& rgb, std::string& err) {
    if (b.size() < 2 || b[0] != 'P' || b[1] != '6') { err = "not a binary PPM"; return false; }
    size_t pos = 2;
    int vals[3], nv = 0;
    while (nv < 3 && pos < b.size()) {
        while (pos < b.size() && isspace(b[pos])) pos++;
        if (pos < b.size() && b[pos] == '#') { while (pos < b.size() && b[pos] != '\n') pos++; continue; }
        int v = 0; bool any = false;
        while (pos < b.size() && isdigit(b[pos])) { v = v * 10 + (b[pos] - '0'); pos++; any = true; }
        if (!any) { err = "bad PPM header"; return false; }
        vals[nv++] = v;
    }
    pos++;
    w = vals[0]; h = vals[1];
    if (nv != 3 || vals[2] != 255 || pos + (size_t)w * h * 3 > b.size()) { err = "bad PPM payload"; return false; }
    rgb.assign(b.begin() + pos, b.begin() + pos + (size_t)w * h * 3);
    return true;
}

// JPEG: the reference's own decoder, stb_image 2.27 (extern/stb_image.h), compiled by this
// directory's Makefile into _ref/libstb_image.so beside _build/, and called through
// stbi_load_from_memory with 3 components requested, as stbi_loadf does (ImageI.cuh:14).
static bool decodeJPEG(const std::vector<unsigned char>& b, int& w, int& h, std::vector<unsigned char>& rgb, std::string& err) {
    typedef unsigned char* (*LoadFn)(const unsigned char*, int, int*, int*, int*, int);
    typedef void (*FreeFn)(void*);
    static LoadFn load = nullptr;
    static FreeFn release = nullptr;
    if (!load) {
        Dl_info self;
        if (!dladdr((void*)&decodeJPEG, &self) || !self.dli_fname) { err = "cannot locate the oracle library"; return false; }
        std::string dir(self.dli_fname);
        dir = dir.substr(0, dir.find_last_of('/'));
        void* lib = dlopen((dir + "/../_ref/libstb_image.so").c_str(), RTLD_NOW | RTLD_LOCAL);
        if (!lib) { err = "JPEG needs oracle/_ref/libstb_image.so (built from the reference's extern/stb_image.h)"; return false; }
        release = (FreeFn)dlsym(lib, "stbi_image_free");
        load = (LoadFn)dlsym(lib, "stbi_load_from_memory");
        if (!load || !release) { err = "stb_image symbols missing"; load = nullptr; return false; }
    }
    int comp = 0;
    unsigned char* px = load(b.data(), (int)b.size(), &w, &h, &comp, 3);
    if (!px) { err = "stb_image could not decode the JPEG"; return false; }
    rgb.assign(px, px + (size_t)w * h * 3);
    release(px);
    return true;
}

// stbi_loadf + loadImageFile_s: 8-bit -> linear float via pow(c/255, 2.2) (extern/stb_image.h:1828),
// rows flipped so that row 0 is the bottom of the image (ImageI.cuh:26-35).
static bool loadImage(const std::string& path, int& w, int& h, std::vector<Spec>& img, std::string& err) {
    std::vector<unsigned char> buf, rgb;
    if (!readFile(path, buf)) { err = "cannot open image " + path; return false; }
    bool ok = (buf.size() > 1 && buf[0] == 'P')           ? decodePPM(buf, w, h, rgb, err)
              : (buf.size() > 1 && buf[0] == 0xFF && buf[1] == 0xD8) ? decodeJPEG(buf, w, h, rgb, err)
                                                                      : decodePNG(buf, w, h, rgb, err);
    if (!ok) { err = path + ": " + err; return false; }
    float lut[256];
    for (int c = 0; c < 256; c++) lut[c] = (float)(std::pow((double)(c / 255.0f), (double)2.2f) * (double)1.0f);
    img.resize((size_t)w * h);
    int index = 0;
    for (int y = h - 1; y >= 0; y--)
        for (int x = 0; x < w; x++) {
            size_t p = (size_t)y * w + x;
            img[index++] = Spec(lut[rgb[3 * p]], lut[rgb[3 * p + 1]], lut[rgb[3 * p + 2]]);
        }
    return true;
}

bool g_textureFiltered = false;

static Spec texelAt(const Mipmap& m, int s, int t, int l, int wrap) {
    const MipLevel& L = m.lv[l];
    if (m.filtered) {
        // row-major texels; repeat wraps to [0, res), clamp clamps, black is zero outside
        if (wrap == 0) { s = ((s % L.ures) + L.ures) % L.ures; t = ((t % L.vres) + L.vres) % L.vres; }
        else if (wrap == 2) { s = std::min(std::max(s, 0), L.ures - 1); t = std::min(std::max(t, 0), L.vres - 1); }
        else if (s < 0 || s >= L.ures || t < 0 || t >= L.vres) return Spec(0.0f);
        return L.slots[(size_t)t * L.ures + s];
    }
    switch (wrap) {
        case 0: s = (int)std::fmod((float)s, (float)L.ures); t = (int)std::fmod((float)t, (float)L.vres); break;
        case 2: s = (int)clampf((float)s, 0, (float)(L.ures - 1)); t = (int)clampf((float)t, 0, (float)(L.vres - 1)); break;
        default:
            if (s < 0 || s >= L.ures || t < 0 || t >= L.vres) return Spec(0.0f);
    }
    // Repeat keeps negative indices negative (fmodf); the reference reads out of bounds there.
    if (s < 0 || t < 0 || s >= L.ures || t >= L.vres) return Spec(0.0f);
    return L.slots[L.slot(s, t)];
}

static void initLevel(MipLevel& L, int u, int v) {
    L.ures = u; L.vres = v; L.ublocks = u >> 2;
    L.slots.assign(L.slot(u - 1, v - 1) + 1, Spec(0.0f));
}

// The filtered pyramid: levels down to 1x1 (1 + floor(log2(max(w, h))) of them), level l of
// max(1, w >> l) x max(1, h >> l) texels, each the mean of its 2x2 parent block, parent indices
// clamped at odd edges, summed in the order (2s, 2t), (2s+1, 2t), (2s, 2t+1), (2s+1, 2t+1).
void buildMipmap(Mipmap& m, int w, int h, const std::vector<Spec>& img, int wrap, bool filtered) {
    m.wrap = wrap;
    m.filtered = filtered;
    if (filtered) {
        int levels = 1;
        while ((std::max(w, h) >> levels) > 0) levels++;
        m.levels = levels;
        m.lv.assign(levels, MipLevel());
        m.lv[0].ures = w; m.lv[0].vres = h;
        m.lv[0].slots = img;
        for (int l = 1; l < levels; l++) {
            const MipLevel& P = m.lv[l - 1];
            MipLevel& L = m.lv[l];
            L.ures = std::max(1, w >> l); L.vres = std::max(1, h >> l);
            L.slots.assign((size_t)L.ures * L.vres, Spec(0.0f));
            for (int t = 0; t < L.vres; t++)
                for (int s = 0; s < L.ures; s++) {
                    const int s0 = std::min(2 * s, P.ures - 1), s1 = std::min(2 * s + 1, P.ures - 1);
                    const int t0 = std::min(2 * t, P.vres - 1), t1 = std::min(2 * t + 1, P.vres - 1);
                    const Spec& a = P.slots[(size_t)t0 * P.ures + s0];
                    const Spec& b = P.slots[(size_t)t0 * P.ures + s1];
                    const Spec& c = P.slots[(size_t)t1 * P.ures + s0];
                    const Spec& d = P.slots[(size_t)t1 * P.ures + s1];
                    L.slots[(size_t)t * L.ures + s] = 0.25f * (((a + b) + c) + d);
                }
        }
        return;
    }
    // CreateMipMap (Mipmap_kernel.cu:39-73). Level l > 0 is filled only for s, t < 16 >> l
    // (Mipmap_kernel.cu:29-30); colliding writes resolve as a sequential loop would (last t wins).
    m.levels = (int)(float)std::log2((double)(float)std::max(w, h));
    m.lv.resize(m.levels);
    initLevel(m.lv[0], w, h);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) m.lv[0].slots[m.lv[0].slot(x, y)] = img[(size_t)y * w + x];
    for (int l = 1; l < m.levels; l++) {
        initLevel(m.lv[l], std::max(1, w >> l), std::max(1, h >> l));
        int lim = 16 >> l;
        for (int t = 0; t < lim; t++)
            for (int s = 0; s < lim; s++) {
                Spec a = texelAt(m, 2 * s, 2 * t, l - 1, 1), b = texelAt(m, 2 * s + 1, 2 * t, l - 1, 1);
                Spec c = texelAt(m, 2 * s, 2 * t + 1, l - 1, 1), d = texelAt(m, 2 * s + 1, 2 * t + 1, l - 1, 1);
                m.lv[l].slots[m.lv[l].slot(s, t)] = 0.25f * (a + b + c + d);
            }
    }
}

static bool createMipmap(Mipmap& m, const std::string& path, int wrap, std::string& err) {
    int w, h;
    std::vector<Spec> img;
    if (!loadImage(path, w, h, img, err)) return false;
    if (!g_textureFiltered && (w < 16 || h < 16)) { err = "images smaller than 16x16 are not supported: " + path; return false; }
    buildMipmap(m, w, h, img, wrap, g_textureFiltered);
    return true;
}

// Mipmap::triangle / lookUp (Mipmap.cuh:41-72)
static Spec mipTriangle(const Mipmap& m, int l, V2 st) {
    l = (int)clampf((float)l, 0, (float)(m.levels - 1));
    float s = st.x * m.lv[l].ures - 0.5f;
    float t = st.y * m.lv[l].vres - 0.5f;
    int s0 = (int)std::floor(s), t0 = (int)std::floor(t);
    float ds = s - s0, dt = t - t0;
    return (1 - ds) * (1 - dt) * texelAt(m, s0, t0, l, m.wrap) + (1 - ds) * dt * texelAt(m, s0, t0 + 1, l, m.wrap) +
           ds * (1 - dt) * texelAt(m, s0 + 1, t0, l, m.wrap) + ds * dt * texelAt(m, s0 + 1, t0 + 1, l, m.wrap);
}

Spec mipTexel(const Mipmap& m, int s, int t, int l) { return texelAt(m, s, t, l, m.wrap); }

// The filtered pipeline keeps this level selection and fixes the top branch, which the reference
// calls with its arguments swapped (texel(levels - 1, 0, 0): texel (levels - 1, 0) of level 0).
Spec mipLookup(const Mipmap& m, V2 st, float width) {
    float levels = (float)(m.levels - 1) + f_log2(std::fmax(width, (float)1e-8));
    if (levels < 0) return mipTriangle(m, 0, st);
    else if (levels >= m.levels - 1)
        return m.filtered ? texelAt(m, 0, 0, m.levels - 1, m.wrap) : texelAt(m, m.levels - 1, 0, 0, m.wrap);
    int iLevel = (int)std::floor(levels);
    float delta = levels - iLevel;
    return lerpS(delta, mipTriangle(m, iLevel, st), mipTriangle(m, iLevel + 1, st));
}

// ---------------------------------------------------------------- filter
void filterTable(float* out) {
    const float alpha = 1.0f, rx = 1.0f, ry = 1.0f;
    float expX = f_exp(-alpha * rx * rx), expY = f_exp(-alpha * ry * ry);
    auto gauss = [&](float d, float ev) { return std::fmax(0.0f, f_exp(-alpha * d * d) - ev); };
    for (int y = 0; y < 16; y++)
        for (int x = 0; x < 16; x++) {
            float px = (x + 0.5f) * rx / 16, py = (y + 0.5f) * ry / 16;
            out[y * 16 + x] = gauss(px, expX) * gauss(py, expY);
        }
}

// ---------------------------------------------------------------- meshes
static bool loadBin(const std::string& path, const Xform& T, Mesh& m, std::string& err) {
    std::vector<unsigned char> b;
    if (!readFile(path + ".bin", b)) { err = "missing mesh " + path + ".bin (OBJ import is not supported)"; return false; }
    if (b.size() < 16) { err = "truncated mesh " + path; return false; }
    int32_t hdr[4];
    memcpy(hdr, b.data(), 16);
    int nTri = hdr[0], nP = hdr[1], nUV = hdr[2], nN = hdr[3];
    if (nTri < 0 || nP < 0 || nUV < 0 || nN < 0) { err = "bad mesh header " + path; return false; }
    size_t need = 16 + 12ull * nP + 8ull * nUV + 12ull * nN + 36ull * nTri;
    if (b.size() < need) { err = "truncated mesh " + path; return false; }
    const unsigned char* p = b.data() + 16;
    std::vector<float> P(3 * (size_t)nP), UV(2 * (size_t)nUV), N(3 * (size_t)nN);
    memcpy(P.data(), p, 12ull * nP); p += 12ull * nP;
    memcpy(UV.data(), p, 8ull * nUV); p += 8ull * nUV;
    memcpy(N.data(), p, 12ull * nN); p += 12ull * nN;
    m.nTri = nTri;
    m.idx.resize(9 * (size_t)nTri);
    memcpy(m.idx.data(), p, 36ull * nTri);
    for (size_t i = 0; i < m.idx.size(); i++) {
        int lim = (i % 3 == 0) ? nP : (i % 3 == 1 ? nUV : nN);
        if ((i % 3 == 0 || lim > 0) && (m.idx[i] < 0 || m.idx[i] >= lim)) { err = "mesh index out of range in " + path; return false; }
    }
    m.P.resize(nP);
    for (int i = 0; i < nP; i++) m.P[i] = T.point(V3(P[3 * i], P[3 * i + 1], P[3 * i + 2]));
    m.UV.resize(nUV);
    for (int i = 0; i < nUV; i++) m.UV[i] = V2(UV[2 * i], UV[2 * i + 1]);
    m.N.resize(nN);
    for (int i = 0; i < nN; i++) m.N[i] = T.normal(V3(N[3 * i], N[3 * i + 1], N[3 * i + 2]));
    return true;
}

// ---------------------------------------------------------------- scene grammar
struct Tok {
    std::ifstream f;
    bool get(std::string& t) { return (bool)(f >> t); }
};

static bool nextTok(Tok& a, Tok& b, bool useA, std::string& t, std::string& err) {
    if (!(useA ? a.get(t) : b.get(t))) { err = "unexpected end of scene/subject file"; return false; }
    return true;
}

bool loadScene(Scene& s, const std::string& dir, const std::string& sceneName, const std::string& subjectName,
               int W, int H, std::string& err) {
    try {
        Tok sc, sb;
        bool hasSubject = !subjectName.empty();
        sc.f.open(dir + sceneName + ".scene");
        if (hasSubject) sb.f.open(dir + subjectName + ".subject");
        if (!sc.f.is_open() || (hasSubject && !sb.f.is_open())) { err = "cannot open scene or subject file"; return false; }
        std::string t;
        auto fs = [&](Tok& k) -> float { if (!k.get(t)) throw std::runtime_error("unexpected end of file"); return std::stof(t); };
        auto is = [&](Tok& k) -> int { if (!k.get(t)) throw std::runtime_error("unexpected end of file"); return std::stoi(t); };
        // camera (SceneLoader.cuh:33-57)
        float cam[11];
        for (int i = 0; i < 11; i++) cam[i] = fs(sc);
        // subject transform (SceneLoader.cuh:59-77)
        float tr[8];
        for (int i = 0; i < 8; i++) tr[i] = fs(sc);
        Xform subjT = makeXform(V3(tr[0], tr[1], tr[2]), V3(tr[3], tr[4], tr[5]), tr[6], tr[7]);
        // images (SceneLoader.cuh:79-97)
        int subImg = 0, imgCount = 0;
        if (hasSubject) { subImg = is(sb); imgCount = subImg; }
        imgCount += is(sc);
        s.mips.resize(imgCount);
        for (int i = 0; i < imgCount; i++) {
            bool useScene = i >= subImg;
            if (!nextTok(sc, sb, useScene, t, err)) return false;
            int wrap = (int)clampf((float)std::stoi(t), 0, 2);
            if (!nextTok(sc, sb, useScene, t, err)) return false;
            if (!createMipmap(s.mips[i], dir + t, wrap, err)) return false;
        }
        // meshes (SceneLoader.cuh:99-113)
        int subMesh = 0, meshCount = 0;
        if (hasSubject) { subMesh = is(sb); meshCount = subMesh; }
        meshCount += is(sc);
        std::vector<std::string> meshFiles;
        for (int m = 0; m < meshCount; m++) {
            if (!nextTok(sc, sb, m >= subMesh, t, err)) return false;
            meshFiles.push_back(dir + t);
        }
        // materials (SceneLoader.cuh:115-171)
        int subMat = 0, matCount = 0;
        if (hasSubject) { subMat = is(sb); matCount = subMat; }
        matCount += is(sc);
        s.mats.resize(matCount);
        for (int mi = 0; mi < matCount; mi++) {
            bool useScene = mi >= subMat;
            Tok& k = useScene ? sc : sb;
            if (!nextTok(sc, sb, useScene, t, err)) return false;
            Mat& M = s.mats[mi];
            M.type = t[0];
            int nTex = M.type == 'G' ? 2 : (M.type == 'M' ? 0 : 1);
            for (int ti = 0; ti < nTex; ti++) {
                if (!nextTok(sc, sb, useScene, t, err)) return false;
                Tex& X = ti == 0 ? M.t0 : M.t1;
                X.type = t[0];
                int texLen = X.type == 'i' ? 1 : (X.type == 'f' || X.type == 'w') ? 2 : X.type == 'c' ? 3 : X.type == 'm' ? 4 : 0;
                float prm[4] = {0, 0, 0, 0};
                for (int q = 0; q < texLen; q++) {
                    prm[q] = fs(k);
                    if (useScene && X.type == 'i') prm[q] = prm[q] + (float)subImg;
                }
                switch (X.type) {
                    case 'c': X.c = Spec(prm[0], prm[1], prm[2]); break;
                    case 'i': X.img = (int)prm[0]; break;
                    case 'f': case 'w': X.octaves = (int)prm[0]; X.omega = prm[1]; break;
                    case 'm': X.octaves = (int)prm[0]; X.omega = prm[1]; X.scale = prm[2]; X.variation = prm[3]; break;
                    default: err = std::string("unsupported texture type '") + X.type + "'"; return false;
                }
                if (X.type == 'i' && (X.img < 0 || X.img >= imgCount)) { err = "image index out of range"; return false; }
            }
            M.extra = fs(k);
            if (M.type != 'L' && M.type != 'G' && M.type != 'R' && M.type != 'F') {
                err = std::string("material type '") + M.type + "' has no constructor"; return false;
            }
        }
        // lights (SceneLoader.cuh:173-193)
        int areaLightMeshes = is(sc);
        std::vector<float> LEs;
        for (int a = 0; a < 3 * areaLightMeshes; a++) LEs.push_back(fs(sc));
        bool hasInf = is(sc) > 0;
        if (hasInf) for (int a = 0; a < 3; a++) LEs.push_back(fs(sc));
        if (matCount < meshCount) { err = "fewer materials than meshes"; return false; }
        if (areaLightMeshes > meshCount) { err = "more light meshes than meshes"; return false; }

        // meshes -> triangles (SceneLoader.cuh:264-282, SceneLoader_kernels.cu:52-88)
        s.meshes.resize(meshCount);
        for (int m = 0; m < meshCount; m++) {
            Xform T;  // Transform(1): identity
            if (m < subMesh) T = subjT;
            if (!loadBin(meshFiles[m], T, s.meshes[m], err)) return false;
            for (int i = 0; i < s.meshes[m].nTri; i++) {
                s.triMesh.push_back(m); s.triLocal.push_back(i); s.triLight.push_back(-1);
            }
        }
        s.nPrim = (int)s.triMesh.size();
        if (s.nPrim == 0) { err = "scene has no triangles"; return false; }
        // area lights, one per emissive triangle in global order; Le reversed (kernels.cu:76-78)
        for (int g = 0; g < s.nPrim; g++) {
            int m = s.triMesh[g];
            if (m >= meshCount - areaLightMeshes) {
                int le = meshCount - m - 1;
                Light L; L.kind = 0; L.Le = Spec(LEs[3 * le], LEs[3 * le + 1], LEs[3 * le + 2]); L.tri = g;
                const Mesh& M = s.meshes[m];
                const int* id = &M.idx[9 * (size_t)s.triLocal[g]];
                V3 p0 = M.P[id[0]], p1 = M.P[id[3]], p2 = M.P[id[6]];
                L.area = 0.5f * length(cross(p1 - p0, p2 - p0));
                s.triLight[g] = (int)s.lights.size();
                s.lights.push_back(L);
            }
        }
        if (hasInf) {
            // InfiniteLight::preprocess with the bounds of primitive 0 (kernels.cu:102, Light.cuh:136-139)
            Light L; L.kind = 1;
            L.Le = Spec(LEs[3 * areaLightMeshes], LEs[3 * areaLightMeshes + 1], LEs[3 * areaLightMeshes + 2]);
            const Mesh& M = s.meshes[s.triMesh[0]];
            const int* id = &M.idx[9 * (size_t)s.triLocal[0]];
            V3 p0 = M.P[id[0]], p1 = M.P[id[3]], p2 = M.P[id[6]];
            V3 mn(std::fmin(std::fmin(p0.x, p1.x), p2.x), std::fmin(std::fmin(p0.y, p1.y), p2.y), std::fmin(std::fmin(p0.z, p1.z), p2.z));
            V3 mx(std::fmax(std::fmax(p0.x, p1.x), p2.x), std::fmax(std::fmax(p0.y, p1.y), p2.y), std::fmax(std::fmax(p0.z, p1.z), p2.z));
            L.center = mn * 0.5f + mx * 0.5f;
            L.radius = length(mx - mn) / 2.0f;
            s.infiniteLight = (int)s.lights.size();
            s.lights.push_back(L);
        }
        // light power distribution (IntegratorHelper.cuh:25-40, Sampling.cuh:123-144)
        int nl = (int)s.lights.size();
        if (nl > 0) {
            s.lfunc.resize(nl); s.lcdf.resize(nl + 1);
            for (int l = 0; l < nl; l++) {
                const Light& L = s.lights[l];
                Spec pw = L.kind == 0 ? (L.Le * L.area) * PI_F : (PI_F * L.radius) * L.Le;
                s.lfunc[l] = pw.y();
            }
            s.lcdf[0] = 0.0f;
            for (int i = 1; i < nl + 1; i++) s.lcdf[i] = s.lcdf[i - 1] + s.lfunc[i - 1] / nl;
            s.funcInt = s.lcdf[nl];
            if (s.funcInt == 0) for (int i = 1; i < nl + 1; i++) s.lcdf[i] = (float)i / (float)nl;
            else for (int i = 1; i < nl + 1; i++) s.lcdf[i] /= s.funcInt;
        }
        // camera (Camera.cu:9-43), aperture 0 (LightTransportSimulation.cu:70)
        s.W = W; s.H = H;
        V3 lookFrom(cam[0], cam[1], cam[2]), lookAt(cam[3], cam[4], cam[5]), vUp(cam[6], cam[7], cam[8]);
        float vfov = cam[9], focusDist = cam[10], aperture = 0.0f;
        float aspect = (float)W / (float)H;
        float theta = vfov * PI_F / 180.0f;
        float hh = f_tan(theta / 2);
        float vh = 2 * hh, vw = aspect * vh;
        s.cw = normalize(lookFrom - lookAt);
        s.cu = normalize(cross(vUp, s.cw));
        s.cv = cross(s.cw, s.cu);
        s.o = lookFrom;
        s.right = s.cu * (focusDist * vw);
        s.up = s.cv * (focusDist * vh);
        s.llc = s.o - (s.right * 0.5f) - (s.up * 0.5f) - s.cw * focusDist;
        s.lensRadius = aperture / 2;
        filterTable(s.filterTable);
    } catch (const std::exception& e) {
        err = std::string("scene parse error: ") + e.what();
        return false;
    }
    // BVH over all triangles in global order (sorted-leaf mode).
    int n = s.nPrim;
    std::vector<float> soup(9 * (size_t)n);
    for (int g = 0; g < n; g++) {
        const Mesh& M = s.meshes[s.triMesh[g]];
        const int* id = &M.idx[9 * (size_t)s.triLocal[g]];
        for (int v = 0; v < 3; v++) {
            V3 p = M.P[id[3 * v]];
            soup[9 * (size_t)g + 3 * v] = p.x; soup[9 * (size_t)g + 3 * v + 1] = p.y; soup[9 * (size_t)g + 3 * v + 2] = p.z;
        }
    }
    std::vector<uint32_t> codes(n);
    std::vector<int32_t> perm(n), parent(2 * n - 1);
    s.bvhLeft.resize(n > 1 ? n - 1 : 0); s.bvhRight.resize(n > 1 ? n - 1 : 0); s.leafPrim.resize(n);
    s.ibox.resize(6 * (size_t)(n > 1 ? n - 1 : 0)); s.lbox.resize(6 * (size_t)n);
    oracle_bvh_out out;
    out.codes_sorted = codes.data(); out.perm = perm.data(); out.left = s.bvhLeft.data(); out.right = s.bvhRight.data();
    out.int_bounds = s.ibox.data(); out.leaf_prim = s.leafPrim.data(); out.leaf_bounds = s.lbox.data(); out.parent = parent.data();
    if (oracle_lbvh_build(soup.data(), n, 1, &out) != 0) { err = "BVH build failed"; return false; }
    return true;
}

} // namespace orc
