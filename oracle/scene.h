// ORACLE — test infrastructure only (see ref.h). Scene model of the CPU restatement.
#pragma once
#include "ref.h"
#include <string>
#include <vector>

namespace orc {

struct Mesh {
    int nTri = 0;
    std::vector<V3> P, N;
    std::vector<V2> UV;
    std::vector<int> idx;  // 9 per triangle: (p, uv, n) x 3 (shapes/Triangle.cuh:19)
};

// Mipmap level with the reference's BlockedArray addressing (core/MemoryHelper.cuh:88-94).
struct MipLevel {
    int ures = 0, vres = 0, ublocks = 0;
    std::vector<Spec> slots;
    long slot(int u, int v) const {
        int bu = u >> 2, bv = v >> 2, ou = u & 3, ov = v >> 2;
        long off = 16L * ((long)ublocks * bv + bu);
        return off + 4L * ov + ou;
    }
};
struct Mipmap {
    int wrap = 1;   // 0 Repeat, 1 Black, 2 Clamp (materials/textures/Mipmap.cuh:10-12)
    int levels = 0;
    bool filtered = false;  // full box pyramid, row-major texels, width from ray differentials
    std::vector<MipLevel> lv;
};

// Texture pipeline of scenes loaded from now on: false = the reference's (bug-compatible) mipmap,
// true = the filtered one. Test infrastructure state, set through oracle_set_texture_filter.
extern bool g_textureFiltered;
// Builds a pyramid from linear RGB texels (row 0 first) in either mode.
void buildMipmap(Mipmap& m, int w, int h, const std::vector<Spec>& img, int wrap, bool filtered);

struct Tex {
    char type = 'c';
    Spec c;
    int img = 0;
    int octaves = 0;
    float omega = 0, scale = 1, variation = 0;
};
struct Mat {
    char type = 'L';
    Tex t0, t1;
    float extra = 0;
};
struct Light {
    int kind = 0;  // 0 area, 1 infinite
    Spec Le;
    int tri = -1;
    float area = 0;
    V3 center;
    float radius = 0;
};

struct Scene {
    int W = 0, H = 0;
    // camera (rendering/Camera.cu:9-43)
    V3 o, right, up, llc, cu, cv, cw;
    float lensRadius = 0;
    float filterTable[256];
    // geometry
    std::vector<Mesh> meshes;
    std::vector<int> triMesh, triLocal, triLight;
    std::vector<Mat> mats;
    std::vector<Mipmap> mips;
    std::vector<Light> lights;
    std::vector<float> lfunc, lcdf;
    float funcInt = 0;
    int infiniteLight = -1;
    // BVH (sorted-leaf mode of the reference build)
    int nPrim = 0;
    std::vector<int> bvhLeft, bvhRight, leafPrim;
    std::vector<float> ibox, lbox;
};

bool loadScene(Scene& s, const std::string& dir, const std::string& scene, const std::string& subject,
               int W, int H, std::string& err);
void filterTable(float* out256);
Spec mipLookup(const Mipmap& m, V2 st, float width);
Spec mipTexel(const Mipmap& m, int s, int t, int l);

} // namespace orc
