// Device-resident scene layout (plain structs shared by the host loader and the kernels).
//
// HBM layout of one loaded scene:
//   tris     TriLeaf[n]   48 B  vertices in BVH leaf order (leaf k = k-th primitive in Morton order)
//   nodes    BvhNode[n-1] 64 B  both child boxes + child indices (unified index space)
//   shade    ShadeTri[n]  80 B  per leaf: normals, uvs, material, light (read once per hit)
//   lights   LightDesc[L]       area lights in global triangle order, infinite light last
//   mats/texs/mips              material, texture and mip-pyramid descriptors
//   mipData  float[3*]          all pyramid levels, in the reference's blocked-array slot order
#pragma once
#include <stdint.h>

namespace lts {

constexpr int kMaxMipLevels = 16;

struct alignas(16) TriLeaf {
    float p0x, p0y, p0z, p1x;
    float p1y, p1z, p2x, p2y;
    float p2z;
    int32_t prim;  // primitive number (global triangle order of the scene)
    int32_t mat;   // material of the primitive (scenes; 0 after a standalone build)
    int32_t pad1;
};

struct alignas(16) BvhNode {
    float lmin[3], lmax[3];
    float rmin[3], rmax[3];
    int32_t left, right;  // unified index: < n-1 interior, >= n-1 leaf (k = idx - (n-1))
    int32_t pad0, pad1;
};

// Node of the 4-wide render BVH: an interior node of even depth of the binary render tree, holding
// the boxes of its children's children (a leaf child is kept as is). Child boxes are quantized to
// 8 bits per plane against the node's own box: coordinate = origin + q * 2^e per axis, with qlo
// rounded down and qhi rounded up so that the decoded box (computed exactly as the traversal does)
// contains the child's box. child >= 0 is a 4-wide node, child < 0 the leaf ~child; `count`
// children are valid. 64 bytes, four 16-byte loads.
struct alignas(64) BvhQNode {
    float ox, oy, oz;     // origin: minimum corner of the node's box
    int8_t ex, ey, ez;    // per-axis scale exponents
    uint8_t count;
    int32_t child[4];
    uint8_t qlo[3][4];    // [axis][child]
    uint8_t qhi[3][4];
    int32_t pad[2];
};
static_assert(sizeof(BvhQNode) == 64, "BvhQNode layout");

struct alignas(16) ShadeTri {
    float n[9];    // vertex normals (valid when hasN)
    float uv[6];   // vertex uvs (defaults (0,0),(1,0),(1,1) when the mesh has none)
    int32_t mat;   // material index (= mesh index)
    int32_t light; // area light index or -1
    int32_t hasN;
    int32_t pad;
};

enum TexType : int32_t { TEX_CONST = 0, TEX_IMAGE = 1, TEX_FBM = 2, TEX_WRINKLED = 3, TEX_MARBLE = 4 };
enum MatType : int32_t { MAT_MATTE = 0, MAT_GLASS = 1, MAT_MIRROR = 2, MAT_MICROFACET = 3 };

struct TexDesc {
    int32_t type;
    int32_t img;
    int32_t octaves;
    float c[3];
    float omega, scale, variation;
    int32_t pad;
};

struct MatDesc {
    int32_t type;
    int32_t tex0, tex1;
    float extra;     // roughness (matte, microfacet) or eta (glass)
    float onA, onB;  // Oren-Nayar A, B (matte with roughness != 0)
    float alpha;     // Trowbridge-Reitz alpha (microfacet)
    int32_t pad;
};

struct MipLevelDesc {
    int32_t ures, vres, ublocks;
    int64_t offset;  // first slot of the level in mipData (slots are RGB float triples)
};

struct MipDesc {
    int32_t wrap;    // 0 repeat, 1 black, 2 clamp
    int32_t levels;
    float lookupLevel;  // (levels - 1) + log2(max(width, 1e-8)) for the fixed width 0.01
    int32_t filtered;   // 1: full box pyramid, row-major texels, width from the ray differentials
    MipLevelDesc lv[kMaxMipLevels];
};

struct LightDesc {
    int32_t kind;  // 0 area, 1 infinite
    int32_t prim;  // area: primitive number
    int32_t leaf;  // area: leaf index of that primitive
    float area;
    float le[3];
    float center[3];
    float radius;
    int32_t pad;
};

struct CameraDesc {
    float o[3], right[3], up[3], llc[3], u[3], v[3];
    float lensRadius;
    int32_t width, height;
};

// Everything the integrator kernels read, by device pointer.
struct SceneView {
    const TriLeaf* tris;
    const BvhNode* nodes;
    const BvhQNode* nodes4;  // 4-wide collapse of `nodes`, root at 0 (used by the traversal)
    const ShadeTri* shade;   // per BVH leaf: shade[k] belongs to the triangle tris[k]
    const MatDesc* mats;
    const TexDesc* texs;
    const MipDesc* mips;
    const float* mipData;
    const LightDesc* lights;
    const float* lfunc;
    const float* lcdf;
    const int* perm;  // Perlin permutation, 512 entries
    float funcInt;
    int32_t nLights;
    int32_t nPrim;
    int32_t infiniteLight;
    CameraDesc cam;
    float filter[256];
};

// Traversal stack entries per lane of the 4-wide traversal. A 4-wide node leaves at most 3 entries
// on the stack, and a leaf at depth D of the binary tree lies under ceil(D / 2) 4-wide nodes, so a
// render tree is accepted only when it is at most kMaxRenderDepth deep.
constexpr int kTraceStack = 128;
constexpr int kMaxRenderDepth = 2 * (kTraceStack / 3);

}  // namespace lts
