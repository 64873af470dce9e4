// ORACLE — test infrastructure only (see ref.h).
//
// CPU restatement of the reference's Karras-2012 LBVH build, step by step:
//   primitive bounds / centroids ... acceleration/BVH.cuh:92-98, shapes/Triangle.cuh:48-53
//   global bounds ................... BVH.cuh:97 (atomicUnion, Geometry.cuh:647-654)
//   Morton codes .................... BVH.cuh:101-116, :134-137; Bound.cuh:169-175
//   stable sort ..................... BVH.cuh:140-146 (thrust::stable_sort_by_key)
//   hierarchy ....................... BVH.cuh:173-236 (delta, treeBuilding)
//   bottom-up refit ................. BVH.cuh:238-262 (addBoundsToNode)
//   flatten ......................... BVH.cuh:311-329 (setTraversalTree)
// Export layout: interior i (i < n-1) holds left/right children in the unified index
// space of the flattened tree (interior 0..n-2, leaf k at n-1+k), as BVH.cuh:313-326.
#include "oracle.h"
#include "ref.h"
#include <algorithm>
#include <numeric>
#include <vector>

namespace orc {

// Total order on floats used for the scene-bounds reduction: -0 sorts below +0.
// (The reference's fatomicMin/Max agree with it on every input without -0.0.)
static inline uint32_t orderKey(float f) {
    uint32_t u = fbits(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
static inline float minTotal(float a, float b) { return orderKey(b) < orderKey(a) ? b : a; }
static inline float maxTotal(float a, float b) { return orderKey(b) > orderKey(a) ? b : a; }

struct Box { V3 mn, mx; };

// Bounds3(p1, p2) (Bound.cuh:107-110)
static inline Box box2(V3 a, V3 b) {
    Box r;
    r.mn = V3(std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z));
    r.mx = V3(std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z));
    return r;
}
// bUnion(Bounds3, Point3) (Geometry.cuh:604-611): result goes through the Bounds3 ctor again
static inline Box unionPt(Box b, V3 p) {
    V3 a(std::fmin(b.mn.x, p.x), std::fmin(b.mn.y, p.y), std::fmin(b.mn.z, p.z));
    V3 c(std::fmax(b.mx.x, p.x), std::fmax(b.mx.y, p.y), std::fmax(b.mx.z, p.z));
    return box2(a, c);
}
// bUnion(Bounds3, Bounds3) (Geometry.cuh:631-639)
static inline Box unionBox(Box b1, Box b2) {
    V3 a(std::fmin(b1.mn.x, b2.mn.x), std::fmin(b1.mn.y, b2.mn.y), std::fmin(b1.mn.z, b2.mn.z));
    V3 c(std::fmax(b1.mx.x, b2.mx.x), std::fmax(b1.mx.y, b2.mx.y), std::fmax(b1.mx.z, b2.mx.z));
    return box2(a, c);
}

// BVH.cuh:101-109
static inline uint32_t leftShift3(uint32_t x) {
    if (x == (1u << 10)) --x;
    x = (x | (x << 16)) & 0b00000011000000000000000011111111u;
    x = (x | (x << 8)) & 0b00000011000000001111000000001111u;
    x = (x | (x << 4)) & 0b00000011000011000011000011000011u;
    x = (x | (x << 2)) & 0b00001001001001001001001001001001u;
    return x;
}

static inline int clz32(uint32_t x) { return x == 0 ? 32 : __builtin_clz(x); }

// BVH.cuh:173-190
static inline int delta(const uint32_t* codes, int n, int i, int j) {
    if (i < 0 || i >= n || j < 0 || j >= n) return -1;
    uint32_t diff = codes[i] ^ codes[j];
    int extra = 0;
    if (diff == 0) { extra = 40; diff = (uint32_t)(i ^ j); }
    return clz32(diff) + extra;
}

} // namespace orc

using namespace orc;

extern "C" uint32_t oracle_morton3(float x, float y, float z) {
    return leftShift3((uint32_t)x) | (leftShift3((uint32_t)y) << 1) | (leftShift3((uint32_t)z) << 2);
}

extern "C" int oracle_lbvh_build(const float* tri, int32_t n, int32_t mode, oracle_bvh_out* out) {
    if (n < 1 || !tri || !out) return -1;
    std::vector<Box> bounds(n);
    std::vector<V3> centroids(n);
    Box global;
    global.mn = V3(INFINITY, INFINITY, INFINITY);
    global.mx = V3(-INFINITY, -INFINITY, -INFINITY);
    for (int i = 0; i < n; i++) {
        const float* t = tri + 9 * (size_t)i;
        V3 p0(t[0], t[1], t[2]), p1(t[3], t[4], t[5]), p2(t[6], t[7], t[8]);
        Box b = unionPt(box2(p0, p1), p2);
        bounds[i] = b;
        centroids[i] = b.mn * 0.5f + b.mx * 0.5f;
        for (int a = 0; a < 3; a++) {
            global.mn[a] = minTotal(global.mn[a], b.mn[a]);
            global.mx[a] = maxTotal(global.mx[a], b.mx[a]);
        }
    }
    // Morton codes (BVH.cuh:134-137, Bound.cuh:169-175)
    std::vector<uint32_t> codes(n);
    for (int i = 0; i < n; i++) {
        V3 o = centroids[i] - global.mn;
        if (global.mx.x > global.mn.x) o.x /= global.mx.x - global.mn.x;
        if (global.mx.y > global.mn.y) o.y /= global.mx.y - global.mn.y;
        if (global.mx.z > global.mn.z) o.z /= global.mx.z - global.mn.z;
        V3 s = o * (float)(1 << 10);
        codes[i] = oracle_morton3(s.x, s.y, s.z);
    }
    // Stable sort of (code, primitive number) by code.
    std::vector<int32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return codes[a] < codes[b]; });
    std::vector<uint32_t> sorted(n);
    for (int i = 0; i < n; i++) sorted[i] = codes[perm[i]];

    // Hierarchy (BVH.cuh:192-236); children in the unified index space.
    std::vector<int32_t> left(n > 1 ? n - 1 : 0), right(n > 1 ? n - 1 : 0);
    std::vector<int32_t> parentLeaf(n, -1), parentInt(n > 1 ? n - 1 : 0, -1);
    const uint32_t* k = sorted.data();
    for (int i = 0; i < n - 1; i++) {
        int d = delta(k, n, i, i + 1) - delta(k, n, i, i - 1);
        d = d > 0 ? 1 : -1;
        int deltaMin = delta(k, n, i, i - d);
        int lMax = 128;
        while (delta(k, n, i, i + lMax * d) > deltaMin) lMax *= 4;
        int l = 0;
        for (int t = lMax / 2; t >= 1; t /= 2)
            if (delta(k, n, i, i + (l + t) * d) > deltaMin) l += t;
        int j = i + l * d;
        int deltaNode = delta(k, n, i, j);
        int s = 0;
        for (int t = (int)std::ceil(l / 2.0f); t >= 1; t = t == 1 ? 0 : (int)std::ceil(t / 2.0f))
            if (delta(k, n, i, i + (s + t) * d) > deltaNode) s += t;
        int gamma = i + s * d + std::min(d, 0);
        int lo = std::min(i, j), hi = std::max(i, j);
        if (lo == gamma) { left[i] = (n - 1) + gamma; parentLeaf[gamma] = i; }
        else { left[i] = gamma; parentInt[gamma] = i; }
        if (hi == gamma + 1) { right[i] = (n - 1) + gamma + 1; parentLeaf[gamma + 1] = i; }
        else { right[i] = gamma + 1; parentInt[gamma + 1] = i; }
    }

    // Leaf bounds: mode 0 reproduces the reference (leaf k takes primitive k's bounds,
    // BVH.cuh:242, the sorted permutation is never applied); mode 1 takes perm[k].
    std::vector<Box> leafB(n), intB(n > 1 ? n - 1 : 0);
    std::vector<int32_t> leafPrim(n);
    for (int kk = 0; kk < n; kk++) {
        leafPrim[kk] = mode == 0 ? kk : perm[kk];
        leafB[kk] = bounds[leafPrim[kk]];
    }
    // Bottom-up refit in post order (the union is order independent: min/max are exact).
    if (n > 1) {
        std::vector<int> stack; stack.push_back(0);
        std::vector<char> done(n - 1, 0);
        while (!stack.empty()) {
            int i = stack.back();
            int ch[2] = {left[i], right[i]};
            bool ready = true;
            for (int c : ch)
                if (c < n - 1 && !done[c]) { stack.push_back(c); ready = false; }
            if (!ready) continue;
            stack.pop_back();
            Box bl = left[i] < n - 1 ? intB[left[i]] : leafB[left[i] - (n - 1)];
            Box br = right[i] < n - 1 ? intB[right[i]] : leafB[right[i] - (n - 1)];
            intB[i] = unionBox(bl, br);
            done[i] = 1;
        }
    }
    // Export.
    for (int i = 0; i < n; i++) {
        out->codes_sorted[i] = sorted[i];
        out->perm[i] = perm[i];
        out->leaf_prim[i] = leafPrim[i];
        const Box& b = leafB[i];
        float* lb = out->leaf_bounds + 6 * (size_t)i;
        lb[0] = b.mn.x; lb[1] = b.mn.y; lb[2] = b.mn.z; lb[3] = b.mx.x; lb[4] = b.mx.y; lb[5] = b.mx.z;
        out->parent[n - 1 + i] = n > 1 ? parentLeaf[i] : -1;
    }
    for (int i = 0; i < n - 1; i++) {
        out->left[i] = left[i];
        out->right[i] = right[i];
        const Box& b = intB[i];
        float* ib = out->int_bounds + 6 * (size_t)i;
        ib[0] = b.mn.x; ib[1] = b.mn.y; ib[2] = b.mn.z; ib[3] = b.mx.x; ib[4] = b.mx.y; ib[5] = b.mx.z;
        out->parent[i] = parentInt[i];
    }
    out->scene_bounds[0] = global.mn.x; out->scene_bounds[1] = global.mn.y; out->scene_bounds[2] = global.mn.z;
    out->scene_bounds[3] = global.mx.x; out->scene_bounds[4] = global.mx.y; out->scene_bounds[5] = global.mx.z;
    return 0;
}
