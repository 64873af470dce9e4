// Host-side scene ingest (see scene_host.h). Follows the reference's grammar and construction:
//   token grammar ................ core/SceneLoader.cuh:33-193 (camera, subject transform, images,
//                                  meshes, materials, lights; subject entries first)
//   mesh files .................... core/FileIO.cuh:170-206; world transform shapes/Triangle.cu:20-43
//                                  (quaternion transform, geometry/Transform.cuh:119-240)
//   materials / textures .......... core/SceneLoader_kernels.cu:8-50, core/SceneLoader.cuh:362-391
//   lights ........................ core/SceneLoader_kernels.cu:52-111; power distribution
//                                  intergrators/IntegratorHelper.cuh:25-40, sampling/Sampling.cuh:123-144
//   mip pyramid ................... materials/textures/Mipmap_kernel.cu:9-73, core/MemoryHelper.cuh:38-109
//   camera ........................ rendering/Camera.cu:9-43
// Fixes of undefined behaviour (documented): one light per emissive triangle in global order
// (the reference indexes lights by per-mesh triangle number), and missing materials are errors.
#include "scene_host.h"

#include <charconv>
#include <cmath>
#include <cstring>
#include <fstream>
#include <stdexcept>

#include "image_io.h"

namespace lts {

namespace {

const float kPi = 3.1415926535898f;

struct F3 { float x, y, z; };
F3 f3(float a, float b, float c) { return F3{a, b, c}; }
F3 add(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
F3 sub(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
F3 mul(F3 a, float s) { return f3(s * a.x, s * a.y, s * a.z); }
float dot3(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
F3 cross3(F3 a, F3 b) {
    double ax = a.x, ay = a.y, az = a.z, bx = b.x, by = b.y, bz = b.z;
    return f3((float)(ay * bz - az * by), (float)(az * bx - ax * bz), (float)(ax * by - ay * bx));
}
float rsqrt_quake(float f) {
    float half = f * 0.5f, y = f;
    int32_t i;
    std::memcpy(&i, &y, 4);
    i = 0x5f3759df - (i >> 1);
    std::memcpy(&y, &i, 4);
    y = y * (1.5f - (half * y * y));
    y = y * (1.5f - (half * y * y));
    return y;
}
F3 unit(F3 a) { return mul(a, rsqrt_quake(dot3(a, a))); }


struct Tokens {
    std::ifstream in;
    std::string next() {
        std::string t;
        if (!(in >> t)) throw std::runtime_error("unexpected end of file");
        return t;
    }
    float f() { return std::stof(next()); }
    int i() { return std::stoi(next()); }
};

bool read_all(const std::string& p, std::vector<char>& buf) {
    std::ifstream f(p, std::ios::binary);
    if (!f) return false;
    buf.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    return true;
}

struct MeshData {
    int nTri = 0, nP = 0, nUV = 0, nN = 0;
    std::vector<F3> P, N;
    std::vector<float> UV;
    std::vector<int32_t> idx;
};

// parseMesh (core/FileIO.cuh:212-231): <stem>.bin when it exists, else <stem>.obj, which is
// converted and written to <stem>.bin for the next load (when the directory is writable). Points
// and normals go through `transform` (the mesh's world transform, applied on the GPU).
bool read_mesh(const std::string& stem, const MeshXf& xf, const MeshTransformFn& transform, MeshData& m, std::string& err) {
    std::vector<char> b;
    if (!read_all(stem + ".bin", b)) {
        int nTri = 0;
        if (!obj_mesh_bytes(stem, b, &nTri, err)) return false;
        std::ofstream out(stem + ".bin", std::ios::binary);
        if (out) out.write(b.data(), (std::streamsize)b.size());
    }
    if (b.size() < 16) { err = "truncated mesh " + stem; return false; }
    int32_t h[4];
    std::memcpy(h, b.data(), 16);
    m.nTri = h[0]; m.nP = h[1]; m.nUV = h[2]; m.nN = h[3];
    if (m.nTri < 0 || m.nP < 0 || m.nUV < 0 || m.nN < 0) { err = "bad mesh header " + stem; return false; }
    size_t need = 16 + 12ull * m.nP + 8ull * m.nUV + 12ull * m.nN + 36ull * m.nTri;
    if (b.size() < need) { err = "truncated mesh " + stem; return false; }
    const char* at = b.data() + 16;
    std::vector<float> rawP(3ull * m.nP), rawN(3ull * m.nN);
    std::memcpy(rawP.data(), at, 12ull * m.nP); at += 12ull * m.nP;
    m.UV.resize(2ull * m.nUV);
    std::memcpy(m.UV.data(), at, 8ull * m.nUV); at += 8ull * m.nUV;
    std::memcpy(rawN.data(), at, 12ull * m.nN); at += 12ull * m.nN;
    if (!transform(xf, rawP, rawN, err)) return false;
    m.P.resize(m.nP);
    for (int i = 0; i < m.nP; i++) m.P[i] = f3(rawP[3 * i], rawP[3 * i + 1], rawP[3 * i + 2]);
    m.N.resize(m.nN);
    for (int i = 0; i < m.nN; i++) m.N[i] = f3(rawN[3 * i], rawN[3 * i + 1], rawN[3 * i + 2]);
    m.idx.resize(9ull * m.nTri);
    std::memcpy(m.idx.data(), at, 36ull * m.nTri);
    for (size_t k = 0; k < m.idx.size(); k++) {
        int which = (int)(k % 3), lim = which == 0 ? m.nP : which == 1 ? m.nUV : m.nN;
        if (which != 0 && lim == 0) continue;
        if (m.idx[k] < 0 || m.idx[k] >= lim) { err = "index out of range in mesh " + stem; return false; }
    }
    return true;
}

}  // namespace

// The token grammar of parseMeshFile: "v" takes three floats, "vt" two, "vn" three, "f" exactly
// three vertex tokens "p/uv/n" (1-based, split at '/', integers read with std::from_chars); every
// other token is skipped, so comments and statements such as "o", "s" or "usemtl" pass unless they
// contain one of the four keywords, exactly as in the reference. The uv index is 0 while no "vt"
// has been read. When the uv or normal field does not parse (e.g. "3//3" after a "vt"),
// std::from_chars leaves the integer unchanged and the reference stores the field parsed before it
// again; that is kept. Fixed instead of undefined: a missing file, a vertex token with fewer than
// three fields, a point field or a float that does not parse, and an index outside its array are
// errors (the reference writes an empty mesh, reads past the token, reads an uninitialised
// integer, throws from stof or reads out of bounds).
bool obj_mesh_bytes(const std::string& stem, std::vector<char>& bin, int* nTriOut, std::string& err) {
    std::ifstream obj(stem + ".obj");
    if (!obj) { err = "missing mesh " + stem + ".bin or " + stem + ".obj"; return false; }
    std::vector<float> P, UV, N;
    std::vector<int32_t> idx;
    int nTri = 0;
    std::string tok;
    auto num = [&](float& v) {
        std::string t;
        if (!(obj >> t)) return false;
        try {
            v = std::stof(t);
        } catch (const std::exception&) {
            return false;
        }
        return true;
    };
    while (obj >> tok) {
        if (tok == "v" || tok == "vn") {
            float a, b, c;
            if (!num(a) || !num(b) || !num(c)) { err = "bad " + tok + " in " + stem + ".obj"; return false; }
            std::vector<float>& dst = tok == "v" ? P : N;
            dst.push_back(a); dst.push_back(b); dst.push_back(c);
        } else if (tok == "vt") {
            float a, b;
            if (!num(a) || !num(b)) { err = "bad vt in " + stem + ".obj"; return false; }
            UV.push_back(a); UV.push_back(b);
        } else if (tok == "f") {
            for (int v = 0; v < 3; v++) {
                std::string t;
                if (!(obj >> t)) { err = "truncated face in " + stem + ".obj"; return false; }
                std::string_view f[3];
                size_t start = 0;
                int nf = 0;
                for (; nf < 3; nf++) {
                    const size_t end = t.find('/', start);
                    f[nf] = std::string_view(t).substr(start, end == std::string::npos ? std::string::npos : end - start);
                    if (end == std::string::npos) { nf++; break; }
                    start = end + 1;
                }
                if (nf < 3) { err = "face vertex " + t + " without p/uv/n fields in " + stem + ".obj"; return false; }
                int i = 0;
                if (std::from_chars(f[0].data(), f[0].data() + f[0].size(), i).ec != std::errc()) {
                    err = "bad face vertex " + t + " in " + stem + ".obj";
                    return false;
                }
                idx.push_back(i - 1);
                if (!UV.empty()) {
                    (void)std::from_chars(f[1].data(), f[1].data() + f[1].size(), i);  // unchanged on failure
                    idx.push_back(i - 1);
                } else {
                    idx.push_back(0);
                }
                (void)std::from_chars(f[2].data(), f[2].data() + f[2].size(), i);  // unchanged on failure
                idx.push_back(i - 1);
            }
            nTri++;
        }
    }
    const int32_t h[4] = {nTri, (int32_t)(P.size() / 3), (int32_t)(UV.size() / 2), (int32_t)(N.size() / 3)};
    for (size_t k = 0; k < idx.size(); k++) {
        const int which = (int)(k % 3), lim = which == 0 ? h[1] : which == 1 ? h[2] : h[3];
        if (which == 1 && lim == 0) continue;
        if (idx[k] < 0 || idx[k] >= lim) { err = "index out of range in " + stem + ".obj"; return false; }
    }
    bin.resize(16 + 4 * (P.size() + UV.size() + N.size() + idx.size()));
    char* at = bin.data();
    auto put = [&](const void* src, size_t bytes) {
        if (bytes) std::memcpy(at, src, bytes);
        at += bytes;
    };
    put(h, 16);
    put(P.data(), 4 * P.size());
    put(UV.data(), 4 * UV.size());
    put(N.data(), 4 * N.size());
    put(idx.data(), 4 * idx.size());
    *nTriOut = nTri;
    return true;
}

namespace {

// BlockedArray slot of (u, v) with the reference's row-block index (MemoryHelper.cuh:88-94).
inline int64_t blocked_slot(int ublocks, int u, int v) {
    return 16LL * ((int64_t)ublocks * (v >> 2) + (u >> 2)) + 4LL * (v >> 2) + (u & 3);
}

// The filtered pyramid: levels down to 1x1 (1 + floor(log2(max(w, h))) of them), level l of
// max(1, w >> l) x max(1, h >> l) row-major texels, each the mean of its 2x2 parent block (parent
// indices clamped at odd edges), summed as ((a + b) + c) + d over (2s, 2t), (2s+1, 2t), (2s, 2t+1),
// (2s+1, 2t+1).
bool build_filtered_mipmap(int w, int h, const std::vector<float>& img, int wrap, MipDesc& d, std::vector<float>& data,
                           std::string& err) {
    int levels = 1;
    while ((std::max(w, h) >> levels) > 0) levels++;
    if (levels > kMaxMipLevels) { err = "image too large"; return false; }
    d.wrap = wrap;
    d.filtered = 1;
    d.levels = levels;
    for (int l = 0; l < levels; l++) {
        MipLevelDesc& L = d.lv[l];
        L.ures = std::max(1, w >> l);
        L.vres = std::max(1, h >> l);
        L.ublocks = 0;
        L.offset = (int64_t)(data.size() / 3);
        data.resize(data.size() + 3 * (size_t)L.ures * L.vres, 0.0f);
    }
    std::memcpy(&data[3 * (size_t)d.lv[0].offset], img.data(), img.size() * sizeof(float));
    for (int l = 1; l < levels; l++) {
        const MipLevelDesc& P = d.lv[l - 1];
        const MipLevelDesc& L = d.lv[l];
        auto src = [&](int s, int t) { return &data[3 * (size_t)(P.offset + (int64_t)t * P.ures + s)]; };
        for (int t = 0; t < L.vres; t++)
            for (int s = 0; s < L.ures; s++) {
                const int s0 = std::min(2 * s, P.ures - 1), s1 = std::min(2 * s + 1, P.ures - 1);
                const int t0 = std::min(2 * t, P.vres - 1), t1 = std::min(2 * t + 1, P.vres - 1);
                float* o = &data[3 * (size_t)(L.offset + (int64_t)t * L.ures + s)];
                for (int k = 0; k < 3; k++)
                    o[k] = 0.25f * (((src(s0, t0)[k] + src(s1, t0)[k]) + src(s0, t1)[k]) + src(s1, t1)[k]);
            }
    }
    return true;
}

// Appends one mip pyramid to `data` and fills `d`: the reference's (filtered == false) or the
// filtered one.
bool build_mipmap(const std::string& path, int wrap, bool filtered, MipDesc& d, std::vector<float>& data,
                  std::string& err) {
    int w = 0, h = 0;
    std::vector<float> img;
    if (!load_image_linear(path, w, h, img, err)) return false;
    std::memset(&d, 0, sizeof(d));
    if (filtered) return build_filtered_mipmap(w, h, img, wrap, d, data, err);
    if (w < 16 || h < 16) { err = "images smaller than 16x16 are not supported: " + path; return false; }
    d.wrap = wrap;
    d.levels = (int)(float)std::log2((double)(float)(w > h ? w : h));
    if (d.levels > kMaxMipLevels) { err = "image too large: " + path; return false; }
    d.lookupLevel = (float)(d.levels - 1) + (float)std::log2((double)std::fmax(0.01f, (float)1e-8));
    for (int l = 0; l < d.levels; l++) {
        MipLevelDesc& L = d.lv[l];
        L.ures = l == 0 ? w : std::max(1, w >> l);
        L.vres = l == 0 ? h : std::max(1, h >> l);
        L.ublocks = L.ures >> 2;
        L.offset = (int64_t)(data.size() / 3);
        data.resize(data.size() + 3 * (size_t)(blocked_slot(L.ublocks, L.ures - 1, L.vres - 1) + 1), 0.0f);
    }
    auto at = [&](int l, int u, int v) { return &data[3 * (size_t)(d.lv[l].offset + blocked_slot(d.lv[l].ublocks, u, v))]; };
    for (int v = 0; v < h; v++)  // later rows overwrite the slots they share (last writer wins)
        for (int u = 0; u < w; u++) std::memcpy(at(0, u, v), &img[3 * ((size_t)v * w + u)], 12);
    for (int l = 1; l < d.levels; l++) {
        const MipLevelDesc& P = d.lv[l - 1];
        auto fetch = [&](int s, int t, float* o) {
            if (s < 0 || s >= P.ures || t < 0 || t >= P.vres) { o[0] = o[1] = o[2] = 0.0f; return; }
            std::memcpy(o, at(l - 1, s, t), 12);
        };
        int lim = 16 >> l;
        for (int t = 0; t < lim; t++)
            for (int s = 0; s < lim; s++) {
                float a[3], b[3], c[3], e[3];
                fetch(2 * s, 2 * t, a); fetch(2 * s + 1, 2 * t, b); fetch(2 * s, 2 * t + 1, c); fetch(2 * s + 1, 2 * t + 1, e);
                float* o = at(l, s, t);
                for (int k = 0; k < 3; k++) o[k] = 0.25f * (a[k] + b[k] + c[k] + e[k]);
            }
    }
    return true;
}

float luminance(float r, float g, float b) { return 0.212671f * r + 0.715160f * g + 0.072169f * b; }

float tr_alpha(float roughness) {  // TrowbridgeReitz roughnessToAlpha (Microfacet.cuh:139-144)
    roughness = std::fmax(roughness, (float)1e-3);
    float x = (float)std::log((double)roughness);
    return 1.62142f + 0.819955f * x + 0.1734f * x * x + 0.0171201f * x * x * x + 0.000640711f * x * x * x * x;
}

}  // namespace

void gaussian_filter_table(float* out) {
    const float alpha = 1.0f, rad = 1.0f;
    const float edge = (float)std::exp((double)(-alpha * rad * rad));
    for (int y = 0; y < 16; y++)
        for (int x = 0; x < 16; x++) {
            float px = (x + 0.5f) * rad / 16, py = (y + 0.5f) * rad / 16;
            float gx = std::fmax(0.0f, (float)std::exp((double)(-alpha * px * px)) - edge);
            float gy = std::fmax(0.0f, (float)std::exp((double)(-alpha * py * py)) - edge);
            out[y * 16 + x] = gx * gy;
        }
}

bool load_scene_files(const MeshTransformFn& transform, const std::string& dir, const std::string& sceneName,
                      const std::string& subjectName, int W,
                      int H, HostScene& S, std::string& err, bool filteredTextures) {
    if (W < 2 || H < 2) { err = "film must be at least 2x2"; return false; }
    Tokens sc, sb;
    const bool hasSubject = !subjectName.empty();
    sc.in.open(dir + sceneName + ".scene");
    if (hasSubject) sb.in.open(dir + subjectName + ".subject");
    if (!sc.in.is_open() || (hasSubject && !sb.in.is_open())) { err = "cannot open scene/subject file"; return false; }
    std::vector<std::string> meshFiles;
    std::vector<float> les;
    int subjectMeshes = 0, areaLightMeshes = 0;
    bool hasInfinite = false;
    float cam[11];
    MeshXf subjectXf;
    try {
        for (float& c : cam) c = sc.f();
        float tr[8];
        for (float& v : tr) v = sc.f();
        subjectXf = mesh_xf_from_tokens(tr);
        // images
        int subjectImages = hasSubject ? sb.i() : 0;
        int nImg = subjectImages + sc.i();
        for (int i = 0; i < nImg; i++) {
            Tokens& k = i < subjectImages ? sb : sc;
            int wrap = std::stoi(k.next());
            wrap = wrap < 0 ? 0 : (wrap > 2 ? 2 : wrap);
            std::string file = k.next();
            MipDesc md;
            if (!build_mipmap(dir + file, wrap, filteredTextures, md, S.mipData, err)) return false;
            S.mips.push_back(md);
        }
        // meshes
        subjectMeshes = hasSubject ? sb.i() : 0;
        int nMesh = subjectMeshes + sc.i();
        for (int m = 0; m < nMesh; m++) meshFiles.push_back(dir + (m < subjectMeshes ? sb : sc).next());
        // materials
        int subjectMats = hasSubject ? sb.i() : 0;
        int nMat = subjectMats + sc.i();
        for (int m = 0; m < nMat; m++) {
            bool fromScene = m >= subjectMats;
            Tokens& k = fromScene ? sc : sb;
            char type = k.next()[0];
            int nTex = type == 'G' ? 2 : (type == 'M' ? 0 : 1);
            MatDesc md{};
            md.tex0 = md.tex1 = -1;
            for (int t = 0; t < nTex; t++) {
                TexDesc td{};
                char tt = k.next()[0];
                int len = tt == 'i' ? 1 : (tt == 'f' || tt == 'w') ? 2 : tt == 'c' ? 3 : tt == 'm' ? 4 : 0;
                float p[4] = {0, 0, 0, 0};
                for (int q = 0; q < len; q++) {
                    p[q] = k.f();
                    if (fromScene && tt == 'i') p[q] = p[q] + (float)subjectImages;
                }
                switch (tt) {
                    case 'c': td.type = TEX_CONST; td.c[0] = p[0]; td.c[1] = p[1]; td.c[2] = p[2]; break;
                    case 'i': td.type = TEX_IMAGE; td.img = (int)p[0]; break;
                    case 'f': td.type = TEX_FBM; td.octaves = (int)p[0]; td.omega = p[1]; break;
                    case 'w': td.type = TEX_WRINKLED; td.octaves = (int)p[0]; td.omega = p[1]; break;
                    case 'm': td.type = TEX_MARBLE; td.octaves = (int)p[0]; td.omega = p[1]; td.scale = p[2]; td.variation = p[3]; break;
                    default: err = std::string("unsupported texture type '") + tt + "'"; return false;
                }
                if (td.type == TEX_IMAGE && (td.img < 0 || td.img >= (int)S.mips.size())) { err = "image index out of range"; return false; }
                (t == 0 ? md.tex0 : md.tex1) = (int)S.texs.size();
                S.texs.push_back(td);
            }
            md.extra = k.f();
            switch (type) {
                case 'L':
                    md.type = MAT_MATTE;
                    if (md.extra != 0) {  // OrenNayar constructor (Microfacet.cuh:18-25)
                        float sigma = md.extra * (kPi / 180);
                        float s2 = sigma * sigma;
                        md.onA = 1.0f - (s2 / (2.0f * (s2 + 0.33f)));
                        md.onB = 0.45f * s2 / (s2 + 0.09f);
                    }
                    break;
                case 'G': md.type = MAT_GLASS; break;
                case 'R': md.type = MAT_MIRROR; break;
                case 'F': md.type = MAT_MICROFACET; md.alpha = tr_alpha(md.extra); break;
                default: err = std::string("material type '") + type + "' has no constructor"; return false;
            }
            S.mats.push_back(md);
        }
        // lights (scene file only)
        areaLightMeshes = sc.i();
        for (int a = 0; a < 3 * areaLightMeshes; a++) les.push_back(sc.f());
        hasInfinite = sc.i() > 0;
        if (hasInfinite) for (int a = 0; a < 3; a++) les.push_back(sc.f());
    } catch (const std::exception& e) {
        err = std::string("scene parse error: ") + e.what();
        return false;
    }
    const int nMesh = (int)meshFiles.size();
    if ((int)S.mats.size() < nMesh) { err = "fewer materials than meshes"; return false; }
    if (areaLightMeshes > nMesh) { err = "more light meshes than meshes"; return false; }
    S.nMeshes = nMesh;

    // triangles in global order
    for (int m = 0; m < nMesh; m++) {
        MeshData md;
        const MeshXf identity;
        if (!read_mesh(meshFiles[m], m < subjectMeshes ? subjectXf : identity, transform, md, err)) return false;
        const bool emissive = m >= nMesh - areaLightMeshes;
        const int le = nMesh - m - 1;  // Le order reversed (SceneLoader_kernels.cu:76-78)
        for (int t = 0; t < md.nTri; t++) {
            const int32_t* id = &md.idx[9 * (size_t)t];
            F3 p[3] = {md.P[id[0]], md.P[id[3]], md.P[id[6]]};
            for (int v = 0; v < 3; v++) { S.soup.push_back(p[v].x); S.soup.push_back(p[v].y); S.soup.push_back(p[v].z); }
            ShadeTri st{};
            st.mat = m;
            st.hasN = md.nN > 0;
            if (st.hasN)
                for (int v = 0; v < 3; v++) { F3 n = md.N[id[3 * v + 2]]; st.n[3 * v] = n.x; st.n[3 * v + 1] = n.y; st.n[3 * v + 2] = n.z; }
            if (md.nUV > 0) {
                for (int v = 0; v < 3; v++) { st.uv[2 * v] = md.UV[2 * (size_t)id[3 * v + 1]]; st.uv[2 * v + 1] = md.UV[2 * (size_t)id[3 * v + 1] + 1]; }
            } else {
                const float def[6] = {0, 0, 1, 0, 1, 1};
                std::memcpy(st.uv, def, sizeof(def));
            }
            st.light = -1;
            if (emissive) {
                LightDesc L{};
                L.kind = 0;
                L.prim = (int)S.shade.size();
                L.le[0] = les[3 * le]; L.le[1] = les[3 * le + 1]; L.le[2] = les[3 * le + 2];
                L.area = 0.5f * std::sqrt(dot3(cross3(sub(p[1], p[0]), sub(p[2], p[0])), cross3(sub(p[1], p[0]), sub(p[2], p[0]))));
                st.light = (int)S.lights.size();
                S.lights.push_back(L);
            }
            S.shade.push_back(st);
        }
    }
    if (S.shade.empty()) { err = "scene has no triangles"; return false; }
    if (hasInfinite) {  // preprocess with the bounds of primitive 0 (SceneLoader_kernels.cu:102)
        LightDesc L{};
        L.kind = 1;
        L.prim = -1; L.leaf = -1;
        for (int c = 0; c < 3; c++) L.le[c] = les[3 * areaLightMeshes + c];
        const float* t = &S.soup[0];
        F3 mn = f3(std::fmin(std::fmin(t[0], t[3]), t[6]), std::fmin(std::fmin(t[1], t[4]), t[7]), std::fmin(std::fmin(t[2], t[5]), t[8]));
        F3 mx = f3(std::fmax(std::fmax(t[0], t[3]), t[6]), std::fmax(std::fmax(t[1], t[4]), t[7]), std::fmax(std::fmax(t[2], t[5]), t[8]));
        F3 c = add(mul(mn, 0.5f), mul(mx, 0.5f));
        L.center[0] = c.x; L.center[1] = c.y; L.center[2] = c.z;
        F3 dg = sub(mx, mn);
        L.radius = std::sqrt(dot3(dg, dg)) / 2.0f;
        S.infiniteLight = (int)S.lights.size();
        S.lights.push_back(L);
    }
    const int nl = (int)S.lights.size();
    if (nl > 0) {  // Distribution1D over power().y()
        S.lfunc.resize(nl);
        S.lcdf.resize(nl + 1);
        for (int l = 0; l < nl; l++) {
            const LightDesc& L = S.lights[l];
            float r, g, b;
            if (L.kind == 0) { r = (L.le[0] * L.area) * kPi; g = (L.le[1] * L.area) * kPi; b = (L.le[2] * L.area) * kPi; }
            else { float k = kPi * L.radius; r = L.le[0] * k; g = L.le[1] * k; b = L.le[2] * k; }
            S.lfunc[l] = luminance(r, g, b);
        }
        S.lcdf[0] = 0.0f;
        for (int i = 1; i <= nl; i++) S.lcdf[i] = S.lcdf[i - 1] + S.lfunc[i - 1] / nl;
        S.funcInt = S.lcdf[nl];
        for (int i = 1; i <= nl; i++) S.lcdf[i] = S.funcInt == 0 ? (float)i / (float)nl : S.lcdf[i] / S.funcInt;
    }
    // camera (aperture 0)
    S.width = W; S.height = H;
    F3 from = f3(cam[0], cam[1], cam[2]), at = f3(cam[3], cam[4], cam[5]), vup = f3(cam[6], cam[7], cam[8]);
    float vfov = cam[9], focus = cam[10];
    float aspect = (float)W / (float)H;
    float theta = vfov * kPi / 180.0f;
    float hh = (float)std::tan((double)(theta / 2));
    float vh = 2 * hh, vw = aspect * vh;
    F3 w = unit(sub(from, at));
    F3 u = unit(cross3(vup, w));
    F3 v = cross3(w, u);
    F3 right = mul(u, focus * vw), up = mul(v, focus * vh);
    F3 llc = sub(sub(sub(from, mul(right, 0.5f)), mul(up, 0.5f)), mul(w, focus));
    auto put = [](float* d, F3 a) { d[0] = a.x; d[1] = a.y; d[2] = a.z; };
    put(S.cam.o, from); put(S.cam.right, right); put(S.cam.up, up); put(S.cam.llc, llc); put(S.cam.u, u); put(S.cam.v, v);
    S.cam.lensRadius = 0.0f / 2;
    S.cam.width = W; S.cam.height = H;
    gaussian_filter_table(S.filter);
    return true;
}

}  // namespace lts
