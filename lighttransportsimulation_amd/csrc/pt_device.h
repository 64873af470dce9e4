// Path-tracer device routines shared by the megakernel and the wavefront kernels.
// Each routine restates a reference routine (tierChampion/LightTransportSimulation/src) with
// the same fp32 evaluation order; see lts_device.h for the arithmetic conventions.
#pragma once
#include "lts_device.h"
#include "lts_scene.h"

namespace lts {

struct Ray {
    V3 o, d;
    float tMax;
};

struct Hit {
    int leaf;   // leaf index of the closest triangle, -1 for a miss
    float t, u, v, invDet;
};

struct SurfacePoint {
    V3 p, pError, wo, n;
    V2 uv;
    V3 dpdu, dpdv;
    V3 sn, sdpdu;
    V3 dpdx, dpdy;
    float dudx, dvdx, dudy, dvdy;  // uv differentials (primary hits; read by filtered textures)
    int leaf;  // BVH leaf of the triangle, which indexes its ShadeTri record
};

// The scene's arrays live in device (global) memory. Kernels that read a SceneView from memory see
// its pointers as generic ones, whose flat loads also count against the LDS wait counter (so a wait
// for a traversal-stack read waits for node loads too); the traversal's loads go through global
// pointers instead.
typedef float GlobalF4 __attribute__((ext_vector_type(4)));
typedef unsigned GlobalU4 __attribute__((ext_vector_type(4)));
LTS_D GlobalF4 load_global_f4(const void* p, int i) { return ((const __attribute__((address_space(1))) GlobalF4*)p)[i]; }
LTS_D GlobalU4 load_global_u4(const void* p, int i) { return ((const __attribute__((address_space(1))) GlobalU4*)p)[i]; }

LTS_D void load_tri(const TriLeaf* __restrict__ tris, int k, V3& p0, V3& p1, V3& p2, int& prim) {
    const GlobalF4 a = load_global_f4(tris + k, 0), b = load_global_f4(tris + k, 1), c = load_global_f4(tris + k, 2);
    p0 = v3(a.x, a.y, a.z);
    p1 = v3(a.w, b.x, b.y);
    p2 = v3(b.z, b.w, c.x);
    prim = __float_as_int(c.y);
}

// The reference's triangle tests take 1.0 / det in double and compare float barycentrics with the
// double 1 -/+ SHADOW_EPS. Both have exact float forms: a double quotient of float operands rounded
// to float is the correctly rounded float quotient (53 >= 2 * 24 + 2 bits: the double rounding is
// innocuous), which the IEEE float division gives; and for a float x, (double)x > c holds exactly
// when x exceeds the largest float not above c, the two constants below.
constexpr float kOneMinusEpsF = 0x1.fff2e4p-1f;  // largest float <= 1.0 - (double)SHADOW_EPS
constexpr float kOnePlusEpsF = 0x1.00068cp+0f;   // largest float <= 1.0 + (double)SHADOW_EPS

// Triangle::intersect hit test (shapes/Triangle.cu:94-128) without its distance test: true when
// the ray meets the triangle inside the barycentric margins, at distance t of either sign.
LTS_D bool tri_mt(V3 p0, V3 p1, V3 p2, const Ray& r, float& t, float& u, float& v, float& invDetOut) {
    V3 edge1 = p1 - p0, edge2 = p2 - p0;
    V3 pVec = cross(r.d, edge2);
    float det = dot(edge1, pVec);
    if (det == 0) return false;
    float invDet = 1.0f / det;
    V3 tVec = r.o - p0;
    u = invDet * dot(tVec, pVec);
    if (u < SHADOW_EPS || u > kOneMinusEpsF) return false;
    V3 qVec = cross(tVec, edge1);
    v = invDet * dot(r.d, qVec);
    if (v < SHADOW_EPS || u + v > kOneMinusEpsF) return false;
    t = invDet * dot(edge2, qVec);
    invDetOut = invDet;
    return true;
}

// Triangle::intersect hit test (shapes/Triangle.cu:94-128): t in (0, tMax).
LTS_D bool tri_hit(V3 p0, V3 p1, V3 p2, const Ray& r, float& t, float& u, float& v, float& invDetOut) {
    return tri_mt(p0, p1, p2, r, t, u, v, invDetOut) && t > 0 && t < r.tMax;
}

// The closest-hit acceptance of leaf `leaf` against the current closest hit (distance r.tMax, leaf
// best): nearer, or exactly as near with a lower leaf index. BVH::intersect (BVH.cuh:365-398) keeps
// the first triangle it meets at the closest distance, and its depth-first, left-first walk meets
// the leaves in index order, so of several triangles at exactly that distance the lowest leaf wins;
// this rule gives the same winner in any visiting order.
LTS_D bool tri_hit_closest(V3 p0, V3 p1, V3 p2, const Ray& r, int leaf, int best, float& t, float& u, float& v,
                           float& invDetOut) {
    return tri_mt(p0, p1, p2, r, t, u, v, invDetOut) && t > 0 && (t < r.tMax || (t == r.tMax && leaf < best));
}

// Triangle::simpleIntersect (shapes/Triangle.cu:50-84)
LTS_D bool tri_occludes(V3 p0, V3 p1, V3 p2, const Ray& r) {
    V3 edge1 = p1 - p0, edge2 = p2 - p0;
    V3 pVec = cross(r.d, edge1);
    float det = dot(edge2, pVec);
    if (det == 0) return false;
    float invDet = 1.0f / det;
    V3 tVec = r.o - p0;
    float u = invDet * dot(tVec, pVec);
    if (u < SHADOW_EPS || u > kOnePlusEpsF) return false;
    V3 qVec = cross(tVec, edge2);
    float v = invDet * dot(r.d, qVec);
    if (v < SHADOW_EPS || u + v > kOnePlusEpsF) return false;
    float t = invDet * dot(edge1, qVec);
    return t > 0 && t < r.tMax;
}

// Bounds3::simpleIntersect (geometry/Bound.cuh:200-227); also returns the entry distance.
LTS_D bool box_hit(const float* mn, const float* mx, const Ray& r, V3 inv, int nx, int ny, int nz, float& tEnter) {
    const float g3 = 1 + 2 * gammaN(3);
    float tMin = ((nx ? mx[0] : mn[0]) - r.o.x) * inv.x;
    float tMax = ((nx ? mn[0] : mx[0]) - r.o.x) * inv.x;
    float tyMin = ((ny ? mx[1] : mn[1]) - r.o.y) * inv.y;
    float tyMax = ((ny ? mn[1] : mx[1]) - r.o.y) * inv.y;
    tMax *= g3;
    tyMax *= g3;
    if (tMin > tyMax || tyMin > tMax) return false;
    if (tyMin > tMin) tMin = tyMin;
    if (tyMax < tMax) tMax = tyMax;
    float tzMin = ((nz ? mx[2] : mn[2]) - r.o.z) * inv.z;
    float tzMax = ((nz ? mn[2] : mx[2]) - r.o.z) * inv.z;
    tzMax *= g3;
    if (tMin > tzMax || tzMin > tMax) return false;
    if (tzMin > tMin) tMin = tzMin;
    if (tzMax < tMax) tMax = tzMax;
    tEnter = tMin;
    return (tMin < r.tMax) && (tMax > 0);
}

constexpr int kStack = kTraceStack;
// Stack entries kept in LDS; deeper entries go to private memory. 22 entries of a 256-lane block
// (22528 B) plus the 256 B material histogram let 7 trace blocks (7 waves per SIMD) share a CU's
// 160 KB of LDS.
constexpr int kShortStack = 22;

// Traversal stack of one lane: entries [0, kShortStack) in LDS (entry e at lds[e * stride], so the
// lanes of a wave touch consecutive words), deeper entries in a private array that only deep
// traversals reach. A private 64-entry array alone is scratch memory, and its traffic dominated
// the traversal kernels.
struct TraceStack {
    int* lds;
    int stride;
};

// LDS stack of thread `tid` in a block-wide array __shared__ int[kShortStack * blockSize].
LTS_D TraceStack lds_stack(int* shared, int tid, int blockSize) { return TraceStack{shared + tid, blockSize}; }

// Child boxes of a 4-wide node: hit distance per slot (INFINITY for a miss or an empty slot). The
// quantized boxes are decoded with qdecode, as the builder did, and contain the exact child boxes;
// the slab test is monotone in the planes, so a ray that passes an exact box passes its decoded box.
// The ray's direction signs pick the near and far plane bytes once per axis for all four children,
// so only those planes are decoded. The slab test is box_hit's in closed form: its pairwise
// comparisons pass exactly when the largest near distance is at most the smallest far distance
// (rounding is monotone, so scaling the minimum equals the minimum of the scaled values), then the
// entry lies before tMax and the exit after 0. A NaN slab distance (a zero direction component with
// the origin exactly on a plane) is ignored here on every axis, where box_hit culls on x only.
// The inverse direction comes from the hardware reciprocal (at most 1 ulp, ray_inverse), so a slab
// distance (p - o) * inv carries at most about 4 roundings instead of box_hit's 3, and the far
// distances are scaled by 1 + 2 gamma(4) = (1 + 4u) / (1 - 4u) instead of 1 + 2 gamma(3): every box
// the exact arithmetic passes still passes, so the test stays conservative.
LTS_D void node4_hits(const BvhQNode* nodes4, int cur, const Ray& r, V3 inv, int nx, int ny, int nz, float (&t)[4],
                      int (&c)[4], int& count) {
    const GlobalU4 h = load_global_u4(nodes4 + cur, 0), ch = load_global_u4(nodes4 + cur, 1), lo = load_global_u4(nodes4 + cur, 2),
                   hi = load_global_u4(nodes4 + cur, 3);
    const float ox = __uint_as_float(h.x), oy = __uint_as_float(h.y), oz = __uint_as_float(h.z);
    const int ex = (int)(int8_t)(h.w & 0xffu), ey = (int)(int8_t)((h.w >> 8) & 0xffu), ez = (int)(int8_t)((h.w >> 16) & 0xffu);
    count = (int)(h.w >> 24);
    // lo = (qlo x, qlo y, qlo z, qhi x), hi = (qhi y, qhi z, pad, pad); byte k belongs to child k
    const unsigned nearX = nx ? lo.w : lo.x, farX = nx ? lo.x : lo.w;
    const unsigned nearY = ny ? hi.x : lo.y, farY = ny ? lo.y : hi.x;
    const unsigned nearZ = nz ? hi.y : lo.z, farZ = nz ? lo.z : hi.y;
    const float g4 = 1 + 2 * gammaN(4);
    const int cc[4] = {(int)ch.x, (int)ch.y, (int)ch.z, (int)ch.w};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const unsigned s = 8 * k;
        const float txMin = (qdecode(ox, ex, (nearX >> s) & 0xffu) - r.o.x) * inv.x;
        const float txMax = (qdecode(ox, ex, (farX >> s) & 0xffu) - r.o.x) * inv.x;
        const float tyMin = (qdecode(oy, ey, (nearY >> s) & 0xffu) - r.o.y) * inv.y;
        const float tyMax = (qdecode(oy, ey, (farY >> s) & 0xffu) - r.o.y) * inv.y;
        const float tzMin = (qdecode(oz, ez, (nearZ >> s) & 0xffu) - r.o.z) * inv.z;
        const float tzMax = (qdecode(oz, ez, (farZ >> s) & 0xffu) - r.o.z) * inv.z;
        const float tMin = fmaxf(fmaxf(txMin, tyMin), tzMin);
        const float tMax = fminf(fminf(txMax, tyMax), tzMax) * g4;
        const bool hit = k < count && tMin <= tMax && tMin <= r.tMax && tMax > 0;
        t[k] = hit ? tMin : INFINITY;
        c[k] = cc[k];
    }
}

// 1 / d per component by v_rcp_f32 (within 1 ulp; 1 / ±0 = ±inf), for the 4-wide node test only.
LTS_D V3 ray_inverse(V3 d) { return v3(__builtin_amdgcn_rcpf(d.x), __builtin_amdgcn_rcpf(d.y), __builtin_amdgcn_rcpf(d.z)); }

LTS_D void cswap(float& ta, int& ca, float& tb, int& cb) {
    if (tb < ta) {
        const float t = ta; ta = tb; tb = t;
        const int c = ca; ca = cb; cb = c;
    }
}

// Closest hit (BVH::intersect, acceleration/BVH.cuh:365-398) over the 4-wide collapse of the render
// tree (PLOC or the LBVH itself, over the same sorted leaves), children visited nearest first. A
// 4-wide node tests the boxes of the binary tree's grandchildren and skips the children's boxes,
// which never prunes a candidate a binary traversal would keep (a ray that passes a box passes its
// ancestors' boxes). Every triangle whose box the ray passes before the current closest distance
// is tested, so the reported hit is the reference traversal's, except for exactly equal distances.
// Render-node fetches and triangle tests of one lane's traversals (counting launches only; the
// timed kernels are instantiated without it).
struct TraceTally {
    unsigned long long nodes = 0, tris = 0;
};

LTS_D Hit trace_closest(const SceneView& sv, Ray& r, TraceStack ts, unsigned long long* nodeCount, unsigned long long* triCount,
                        TraceTally* tally = nullptr) {
    Hit h; h.leaf = -1; h.t = h.u = h.v = h.invDet = 0;
    const int n = sv.nPrim;
    const V3 inv = ray_inverse(r.d);
    int nx = inv.x < 0, ny = inv.y < 0, nz = inv.z < 0;
    unsigned long long nodes = 0, tests = 0;
    if (n == 1) {
        V3 p0, p1, p2; int prim;
        load_tri(sv.tris, 0, p0, p1, p2, prim);
        float t, u, v, id;
        tests++;
        if (tri_hit_closest(p0, p1, p2, r, 0, h.leaf, t, u, v, id)) { r.tMax = t; h.leaf = 0; h.t = t; h.u = u; h.v = v; h.invDet = id; }
    } else {
        int spill[kStack - kShortStack];
        int sp = 0;
        int cur = 0;  // the 4-wide root; entries < 0 are leaves (~leaf)
        // every node and leaf is entered at most once; the step bound and the index checks keep a
        // malformed node array from running away
        for (unsigned steps = 0; steps < 2u * (unsigned)n + 64u; steps++) {
            if (cur >= 0) {
                if (cur >= n - 1) break;
                float t[4];
                int c[4], count;
                node4_hits(sv.nodes4, cur, r, inv, nx, ny, nz, t, c, count);
                nodes++;
                cswap(t[0], c[0], t[1], c[1]);
                cswap(t[2], c[2], t[3], c[3]);
                cswap(t[0], c[0], t[2], c[2]);
                cswap(t[1], c[1], t[3], c[3]);
                cswap(t[1], c[1], t[2], c[2]);
#pragma unroll
                for (int k = 3; k >= 1; k--) {
                    if (t[k] < INFINITY && sp < kStack) {
                        if (sp < kShortStack) ts.lds[sp * ts.stride] = c[k];
                        else spill[sp - kShortStack] = c[k];
                        sp++;
                    }
                }
                if (t[0] < INFINITY) { cur = c[0]; continue; }
            } else {
                const int k = ~cur;
                if (k >= n) break;
                V3 p0, p1, p2; int prim;
                load_tri(sv.tris, k, p0, p1, p2, prim);
                float t, u, v, id;
                tests++;
                if (tri_hit_closest(p0, p1, p2, r, k, h.leaf, t, u, v, id)) {
                    r.tMax = t; h.leaf = k; h.t = t; h.u = u; h.v = v; h.invDet = id;
                }
            }
            if (sp == 0) break;
            --sp;
            cur = sp < kShortStack ? ts.lds[sp * ts.stride] : spill[sp - kShortStack];
        }
    }
    if (nodeCount) { atomicAdd(nodeCount, nodes); atomicAdd(triCount, tests); }
    if (tally) { tally->nodes += nodes; tally->tris += tests; }
    return h;
}

// The reference's leaf-box test for a shadow ray (its inverse direction, BVH.cuh:336, and the
// triangle's bounds), out of line: it runs once per occluded shadow ray.
static __device__ __attribute__((noinline)) bool leaf_box_passes(V3 p0, V3 p1, V3 p2, Ray r) {
    const V3 inv = v3(1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z);
    const float mn[3] = {fminf(fminf(p0.x, p1.x), p2.x), fminf(fminf(p0.y, p1.y), p2.y), fminf(fminf(p0.z, p1.z), p2.z)};
    const float mx[3] = {fmaxf(fmaxf(p0.x, p1.x), p2.x), fmaxf(fmaxf(p0.y, p1.y), p2.y), fmaxf(fmaxf(p0.z, p1.z), p2.z)};
    float tEnter;
    return box_hit(mn, mx, r, inv, inv.x < 0, inv.y < 0, inv.z < 0, tEnter);
}

// The triangle test of an any-hit traversal: Triangle::simpleIntersect for shadow rays (with
// kLeafBox, only when the leaf-box test passes too), or the closest-hit test (Triangle::intersect)
// when the question is whether leaf `leaf` would win the closest hit against leaf `tieLeaf` at
// distance r.tMax (nearer, or as near with a lower index).
template <bool kHitTest, bool kLeafBox>
LTS_D bool leaf_blocks(V3 p0, V3 p1, V3 p2, const Ray& r, int leaf, int tieLeaf) {
    if constexpr (kHitTest) {
        float t, u, v, id;
        return tri_hit_closest(p0, p1, p2, r, leaf, tieLeaf, t, u, v, id);
    } else if constexpr (kLeafBox) {
        return tri_occludes(p0, p1, p2, r) && leaf_box_passes(p0, p1, p2, r);
    } else {
        return tri_occludes(p0, p1, p2, r);
    }
}

// Any hit (BVH::simpleIntersect, acceleration/BVH.cuh:331-363) over the 4-wide tree: leaf children
// are tested as soon as their box is hit, interior children are stacked. Returns the first leaf
// whose test passes, or -1.
template <bool kHitTest, bool kLeafBox>
LTS_D int any_hit_leaf(int n, const BvhQNode* nodes4, const TriLeaf* tris, const Ray& r, TraceStack ts, TraceTally* tally,
                       int tieLeaf) {
    const V3 inv = ray_inverse(r.d);
    int nx = inv.x < 0, ny = inv.y < 0, nz = inv.z < 0;
    if (n == 1) {
        V3 p0, p1, p2; int prim;
        load_tri(tris, 0, p0, p1, p2, prim);
        if (tally) tally->tris++;
        return leaf_blocks<kHitTest, kLeafBox>(p0, p1, p2, r, 0, tieLeaf) ? 0 : -1;
    }
    int spill[kStack - kShortStack];
    int sp = 0;
    int cur = 0;
    for (unsigned steps = 0; steps < 2u * (unsigned)n + 64u; steps++) {
        if (cur < 0 || cur >= n - 1) return -1;
        float t[4];
        int c[4], count;
        node4_hits(nodes4, cur, r, inv, nx, ny, nz, t, c, count);
        if (tally) tally->nodes++;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (t[k] < INFINITY) {
                if (c[k] < 0) {
                    const int leaf = ~c[k];
                    if (leaf >= n) return -1;
                    V3 p0, p1, p2; int prim;
                    load_tri(tris, leaf, p0, p1, p2, prim);
                    if (tally) tally->tris++;
                    if (leaf_blocks<kHitTest, kLeafBox>(p0, p1, p2, r, leaf, tieLeaf)) return leaf;
                } else if (sp < kStack) {
                    if (sp < kShortStack) ts.lds[sp * ts.stride] = c[k];
                    else spill[sp - kShortStack] = c[k];
                    sp++;
                }
            }
        }
        if (sp == 0) return -1;
        --sp;
        cur = sp < kShortStack ? ts.lds[sp * ts.stride] : spill[sp - kShortStack];
    }
    return -1;
}

// kHitTest: the closest-hit question of an MIS ray, whether any leaf beats tieLeaf at r.tMax.
// Otherwise a shadow ray. Triangle::simpleIntersect accepts points up to SHADOW_EPS outside the
// triangle's edges, and the reference reaches a triangle only through its leaf's box test
// (BVH.cuh:343-348: the leaf node's bounds, the primitive's own bounds). The render tree's quantized
// boxes are wider than that leaf box, so a triangle the enlarged test accepts is also put to the
// leaf-box test; when that fails (a ray passing just outside a triangle's edge and its bounds, rare),
// the ray is traversed again with the leaf-box test on every triangle hit. The common path keeps the
// plain loop. (The closest-hit test accepts only points inside the triangle shrunk by SHADOW_EPS,
// which lie inside its bounds.)
template <bool kHitTest = false>
LTS_D bool trace_any(const SceneView& sv, const Ray& r, TraceStack ts, TraceTally* tally = nullptr, int tieLeaf = -1) {
    if constexpr (kHitTest) {
        return any_hit_leaf<true, false>(sv.nPrim, sv.nodes4, sv.tris, r, ts, tally, tieLeaf) >= 0;
    } else {
        const int leaf = any_hit_leaf<false, false>(sv.nPrim, sv.nodes4, sv.tris, r, ts, tally, -1);
        if (leaf < 0) return false;
        V3 p0, p1, p2; int prim;
        load_tri(sv.tris, leaf, p0, p1, p2, prim);
        if (leaf_box_passes(p0, p1, p2, r)) return true;
        return any_hit_leaf<false, true>(sv.nPrim, sv.nodes4, sv.tris, r, ts, tally, -1) >= 0;
    }
}

// Triangle::intersect interaction part + SurfaceInteraction constructor / setShaddingGeometry
// (shapes/Triangle.cu:130-200, geometry/Interaction.cuh:85-116)
LTS_D SurfacePoint make_surface(const SceneView& sv, int leaf, const Ray& r, float u, float v, float invDet) {
    V3 p0, p1, p2; int prim;
    load_tri(sv.tris, leaf, p0, p1, p2, prim);
    const ShadeTri& st = sv.shade[leaf];
    V3 edge1 = p1 - p0, edge2 = p2 - p0;
    float w = 1 - u - v;
    V2 uv0 = v2(st.uv[0], st.uv[1]), uv1 = v2(st.uv[2], st.uv[3]), uv2 = v2(st.uv[4], st.uv[5]);
    V2 duv02 = uv0 - uv2, duv12 = uv1 - uv2;
    V3 dp02 = p0 - p2, dp12 = p1 - p2;
    float determinant = duv02.x * duv12.y - duv02.y * duv12.x;
    V3 dpdu, dpdv;
    if (determinant == 0) coordinateSystem(normalize(cross(edge2, edge1)), &dpdu, &dpdv);
    else {
        dpdu = (dp02 * duv12.y - dp12 * duv02.y) * invDet;
        dpdv = (dp02 * -duv12.x + dp12 * duv02.x) * invDet;
    }
    float xs = fabsf(w * p0.x) + fabsf(u * p1.x) + fabsf(v * p2.x);
    float ys = fabsf(w * p0.y) + fabsf(u * p1.y) + fabsf(v * p2.y);
    float zs = fabsf(w * p0.z) + fabsf(u * p1.z) + fabsf(v * p2.z);
    SurfacePoint s;
    s.pError = v3(xs, ys, zs) * gammaN(7);
    s.p = p0 * w + p1 * u + p2 * v;
    s.uv = uv0 * w + uv1 * u + uv2 * v;
    s.wo = -r.d;
    s.dpdu = dpdu;
    s.n = normalize(cross(dpdu, dpdv));
    s.sn = s.n;
    s.sdpdu = dpdu;
    s.leaf = leaf;
    s.dpdx = s.dpdy = v3(0, 0, 0);
    s.dpdv = dpdv;
    s.dudx = s.dvdx = s.dudy = s.dvdy = 0.0f;
    if (st.hasN) {
        V3 n0 = v3(st.n[0], st.n[1], st.n[2]), n1 = v3(st.n[3], st.n[4], st.n[5]), n2 = v3(st.n[6], st.n[7], st.n[8]);
        V3 ns = normalize(n0 * w + n1 * u + n2 * v);
        V3 tg = s.dpdu;
        V3 bt = cross(ns, tg);
        if (lengthSq(bt) > 0.0f) { bt = normalize(bt); tg = cross(bt, ns); }
        else coordinateSystem(ns, &tg, &bt);
        s.sn = normalize(cross(tg, bt));
        s.n = faceTowards(s.n, s.sn);
        s.sdpdu = tg;
        s.n = faceTowards(s.n, s.sn);
    }
    return s;
}

LTS_D Ray spawn_ray(const SurfacePoint& it, V3 d) {
    Ray r;
    r.o = offsetRayOrigin(it.p, it.pError, it.n, d);
    r.d = d;
    r.tMax = INFINITY;
    return r;
}

// ---------------------------------------------------------------- sampling (sampling/Sampling.cuh)
LTS_D V2 uniform_disk(V2 u) {
    float r = sqrtf(u.x);
    float theta = 2 * PI_F * u.y;
    return v2(r * f_cos(theta), r * f_sin(theta));
}
LTS_D V3 cosine_hemisphere(V2 u) {
    V2 d = uniform_disk(u);
    float z = sqrtf(fmaxf(0.0f, 1 - d.x * d.x - d.y * d.y));
    return v3(d.x, d.y, z);
}
LTS_D V3 uniform_hemisphere(V2 u) {
    float z = u.x;
    float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    float phi = 2 * PI_F * u.y;
    return v3(r * f_cos(phi), r * f_sin(phi), z);
}
LTS_D float power_heuristic(float fPdf, float gPdf) {
    float f = 1 * fPdf, g = 1 * gPdf;
    return (f * f) / (f * f + g * g);
}

// ---------------------------------------------------------------- BxDFs (materials/models/*.cuh)
enum { B_REFL = 1, B_TRANS = 2, B_DIFF = 4, B_GLOSSY = 8, B_SPEC = 16, B_ALL = 31 };

LTS_D float cosT(V3 w) { return w.z; }
LTS_D float absCosT(V3 w) { return fabsf(w.z); }
LTS_D float cos2T(V3 w) { return w.z * w.z; }
LTS_D float sin2T(V3 w) { return fmaxf(1 - cos2T(w), 0.0f); }
LTS_D float sinT(V3 w) { return sqrtf(sin2T(w)); }
LTS_D float tanT(V3 w) { return sinT(w) / cosT(w); }
LTS_D float tan2T(V3 w) { return sin2T(w) / cos2T(w); }
LTS_D float cosPhi(V3 w) { float s = sinT(w); return clampf(w.x / s, -1, 1); }  // BxDF.cuh:144-147
LTS_D float sinPhi(V3 w) { float s = sinT(w); return clampf(w.y / s, -1, 1); }
LTS_D float cos2Phi(V3 w) { return cosPhi(w) * cosPhi(w); }
LTS_D float sin2Phi(V3 w) { return sinPhi(w) * sinPhi(w); }
LTS_D bool sameHemi(V3 a, V3 b) { return a.z * b.z > 0; }

// fresnelReflectionDielectric (BxDF.cuh:16-37), including its clamp(cos, 1, -1)
LTS_D float fr_dielectric(float cosThetaI, float etaI, float etaT) {
    cosThetaI = clampf(cosThetaI, 1, -1);
    bool entering = cosThetaI > 0.0f;
    if (!entering) { float tmp = etaI; etaI = etaT; etaT = tmp; cosThetaI = fabsf(cosThetaI); }
    float sinThetaI = sqrtf(fmaxf(0.0f, 1 - cosThetaI * cosThetaI));
    float sinThetaT = etaI / etaT * sinThetaI;
    if (sinThetaT >= 1) return 1;
    float cosThetaT = sqrtf(fmaxf(0.0f, 1 - sinThetaT * sinThetaT));
    float rpar = ((etaT * cosThetaI) - (etaI * cosThetaT)) / ((etaT * cosThetaI) + (etaI * cosThetaT));
    float rperp = ((etaI * cosThetaI) - (etaT * cosThetaT)) / ((etaI * cosThetaI) + (etaT * cosThetaT));
    return (rpar * rpar + rperp * rperp) / 2;
}
LTS_D bool refract_dir(V3 wi, V3 n, float eta, V3* wt) {
    float cosThetaI = dot(n, wi);
    float sin2ThetaI = fmaxf(0.0f, 1.0f - cosThetaI * cosThetaI);
    float sin2ThetaT = eta * eta * sin2ThetaI;
    if (sin2ThetaT >= 1) return false;
    float cosThetaT = sqrtf(1 - sin2ThetaT);
    *wt = -wi * eta + n * (eta * cosThetaI - cosThetaT);
    return true;
}

// Trowbridge-Reitz distribution (Microfacet.cuh:124-181)
LTS_D float tr_lambda(V3 w, float ax, float ay) {
    float absTan = fabsf(tanT(w));
    if (isinf(absTan)) return 0.0f;
    float alpha = sqrtf(cos2Phi(w) * ax * ax + sin2Phi(w) * ay * ay);
    float a2t2 = (alpha * absTan) * (alpha * absTan);
    return (-1 + sqrtf(1.0f + a2t2)) / 2;
}
LTS_D float tr_D(V3 wh, float ax, float ay) {
    float tan2 = tan2T(wh);
    if (isinf(tan2)) return 0;
    const float cos4 = cos2T(wh) * cos2T(wh);
    float e = (cos2Phi(wh) / (ax * ax) + sin2Phi(wh) / (ay * ay)) * tan2;
    return 1 / (PI_F * ax * ay * cos4 * (1 + e) * (1 + e));
}
LTS_D float tr_G1(V3 w, float ax, float ay) { return 1 / (1 + tr_lambda(w, ax, ay)); }
LTS_D float tr_G(V3 wo, V3 wi, float ax, float ay) { return 1 / (1 + tr_lambda(wo, ax, ay) + tr_lambda(wi, ax, ay)); }
LTS_D float tr_pdf(V3 wo, V3 wh, float ax, float ay) { return tr_D(wh, ax, ay) * tr_G1(wo, ax, ay) * absDot(wo, wh) / absCosT(wo); }

// TrowbridgeReitzSample11 / TrowbridgeReitzSample (Microfacet.cuh:53-121)
LTS_D void tr_sample11(float cosTheta, float U1, float U2, float* sx, float* sy) {
    if ((double)cosTheta > .9999) {
        float r = sqrtf(U1 / (1 - U1));
        float phi = (float)(6.28318530718 * (double)U2);
        *sx = r * f_cos(phi);
        *sy = r * f_sin(phi);
        return;
    }
    float sinTheta = sqrtf(fmaxf(0.0f, 1.0f - cosTheta * cosTheta));
    float tanTheta = sinTheta / cosTheta;
    float a = 1 / tanTheta;
    float G1 = 2 / (1 + sqrtf(1.f + 1.f / (a * a)));
    float A = 2 * U1 / G1 - 1;
    float tmp = 1.f / (A * A - 1.f);
    if ((double)tmp > 1e10) tmp = (float)1e10;
    float B = tanTheta;
    float D = sqrtf(fmaxf(float(B * B * tmp * tmp - (A * A - B * B) * tmp), 0.0f));
    float sx1 = B * tmp - D, sx2 = B * tmp + D;
    *sx = (A < 0 || sx2 > 1.f / tanTheta) ? sx1 : sx2;
    float S;
    if (U2 > 0.5f) { S = 1.f; U2 = 2.f * (U2 - .5f); }
    else { S = -1.f; U2 = 2.f * (.5f - U2); }
    float z = (U2 * (U2 * (U2 * 0.27385f - 0.73369f) + 0.46341f)) /
              (U2 * (U2 * (U2 * 0.093073f + 0.309420f) - 1.000000f) + 0.597999f);
    *sy = S * z * sqrtf(1.f + *sx * *sx);
}
LTS_D V3 tr_sample(V3 wi, float ax, float ay, float U1, float U2) {
    V3 ws = normalize(v3(ax * wi.x, ay * wi.y, wi.z));
    float sx, sy;
    tr_sample11(cosT(ws), U1, U2, &sx, &sy);
    float tmp = cosPhi(ws) * sx - sinPhi(ws) * sy;
    sy = sinPhi(ws) * sx + cosPhi(ws) * sy;
    sx = tmp;
    sx = ax * sx;
    sy = ay * sy;
    return normalize(v3(-sx, -sy, (float)1.));
}

// The BSDF of one hit: at most one BxDF (Material.cuh:31-130 and the microfacet mirror extension).
struct Bsdf {
    V3 ns, ng, ss, ts;
    int kind;   // -1 none, 0 Lambert, 1 OrenNayar, 2 SpecularReflection, 3 FresnelSpecular, 4 MicrofacetReflection
    int type;
    Spec R, T;
    float A, B, eta, alpha;

    LTS_D V3 toLocal(V3 v) const { return v3(dot(v, ss), dot(v, ts), dot(v, ns)); }
    LTS_D V3 toWorld(V3 v) const {
        return v3(ss.x * v.x + ts.x * v.y + ns.x * v.z, ss.y * v.x + ts.y * v.y + ns.y * v.z, ss.z * v.x + ts.z * v.y + ns.z * v.z);
    }
    LTS_D bool matches(int flags) const { return kind >= 0 && (type & flags) == type; }

    LTS_D Spec fLocal(V3 wo, V3 wi) const {
        if (kind == 0) return R * INV_PI_F;
        if (kind == 1) {
            float sinThetaO = sinT(wo), sinThetaI = sinT(wi);
            float maxCos = 0.0f;
            if ((double)sinThetaI > 1e-4 && (double)sinThetaO > 1e-4) {
                float sinPhiO = sinPhi(wo), cosPhiO = cosPhi(wo);
                float sinPhiI = sinPhi(wi), cosPhiI = cosPhi(wi);
                float dCos = cosPhiI * cosPhiO + sinPhiI * sinPhiO;
                maxCos = fmaxf(0.0f, dCos);
            }
            float sinAlpha, tanBeta;
            if (absCosT(wi) > absCosT(wo)) { sinAlpha = sinThetaO; tanBeta = sinThetaI / absCosT(wi); }
            else { sinAlpha = sinThetaI; tanBeta = sinThetaO / absCosT(wo); }
            return R * INV_PI_F * (A + B * maxCos * sinAlpha * tanBeta);
        }
        if (kind == 4) {
            float cosThetaO = absCosT(wo), cosThetaI = absCosT(wi);
            V3 wh = wi + wo;
            if (cosThetaI == 0 || cosThetaO == 0) return spec(0.0f);
            if (wh.x == 0.0f && wh.y == 0.0f && wh.z == 0.0f) return spec(0.0f);
            wh = normalize(wh);
            Spec F = spec(1.0f);
            return R * tr_D(wh, alpha, alpha) * tr_G(wo, wi, alpha, alpha) * F / (4 * cosThetaI * cosThetaO);
        }
        return spec(0.0f);
    }
    LTS_D float pdfLocal(V3 wo, V3 wi) const {
        if (kind == 0 || kind == 1) return sameHemi(wo, wi) ? absCosT(wi) * INV_PI_F : 0;
        if (kind == 4) {
            if (!sameHemi(wo, wi)) return 0;
            V3 wh = normalize(wo + wi);
            return tr_pdf(wo, wh, alpha, alpha) / (4 * dot(wo, wh));
        }
        return 0.0f;
    }
    // BSDF::f (BSDF.cuh:60-72)
    LTS_D Spec f(V3 woW, V3 wiW, int flags) const {
        V3 wi = toLocal(wiW), wo = toLocal(woW);
        bool reflect = dot(wiW, ng) * dot(woW, ng) > 0;
        Spec r = spec(0.0f);
        if (matches(flags) && ((reflect && (type & B_REFL)) || (!reflect && (type & B_TRANS)))) r = r + fLocal(wo, wi);
        return r;
    }
    // BSDF::Pdf (BSDF.cuh:129-144)
    LTS_D float pdf(V3 woW, V3 wiW, int flags) const {
        if (kind < 0) return 0.0f;
        V3 wo = toLocal(woW), wi = toLocal(wiW);
        if (wo.z == 0) return 0.0f;
        float p = 0.0f;
        int m = 0;
        if (matches(flags)) { m++; p = p + pdfLocal(wo, wi); }
        return m > 0 ? p / m : 0.0f;
    }
    // BSDF::sampleF (BSDF.cuh:74-127) with one component
    LTS_D Spec sampleF(V3 woW, V3* wiW, V2 u, float* pdfOut, int flags, int* sampledType) const {
        if (!matches(flags)) { *pdfOut = 0; return spec(0.0f); }
        V3 wo = toLocal(woW), wi;
        *pdfOut = 0;
        *sampledType = type;
        Spec f = spec(0.0f);
        switch (kind) {
            case 0: case 1:
                wi = cosine_hemisphere(u);
                if (wo.z < 0) wi.z *= -1;
                *pdfOut = pdfLocal(wo, wi);
                f = fLocal(wo, wi);
                break;
            case 2:
                wi = v3(-wo.x, -wo.y, wo.z);
                *pdfOut = 1.0f;
                f = (spec(1.0f) * R) / absCosT(wi);
                break;
            case 3: {
                float F = fr_dielectric(cosT(wo), 1.0f, eta);
                if (u.x < F) {
                    wi = v3(-wo.x, -wo.y, -wo.z);
                    *sampledType = B_SPEC | B_REFL;
                    *pdfOut = F;
                    f = F * R / absCosT(wi);
                } else {
                    bool entering = cosT(wo) > 0;
                    float etaI = entering ? 1.0f : eta, etaT = entering ? eta : 1.0f;
                    if (!refract_dir(wo, faceTowards(v3(0, 0, 1), wo), etaI / etaT, &wi)) return spec(0.0f);
                    Spec ft = T * (1 - F);
                    ft = ft * ((etaI * etaI) / (etaT * etaT));
                    *sampledType = B_SPEC | B_TRANS;
                    *pdfOut = 1.0f - F;
                    f = ft / absCosT(wi);
                }
                break;
            }
            case 4: {
                if (wo.z == 0) return spec(0.0f);
                bool flip = wo.z < 0;
                V3 wh = tr_sample(flip ? -wo : wo, alpha, alpha, u.x, u.y);
                if (flip) wh = -wh;
                if (dot(wo, wh) < 0) return spec(0.0f);
                wi = -wo + wh * (2 * dot(wo, wh));
                if (!sameHemi(wo, wi)) return spec(0.0f);
                *pdfOut = tr_pdf(wo, wh, alpha, alpha) / (4 * dot(wo, wh));
                f = fLocal(wo, wi);
                break;
            }
        }
        if (*pdfOut == 0) return spec(0.0f);
        *wiW = toWorld(wi);
        return f;
    }
};

// ---------------------------------------------------------------- textures (Procedural.cuh)
LTS_D float noise_weight(float t) { float t3 = t * t * t; float t4 = t3 * t; return 6 * t4 * t - 15 * t4 + 10 * t3; }
LTS_D float smooth_step(float mn, float mx, float val) {
    float v = clampf((val - mn) / (mx - mn), 0, 1);
    return v * v * (-2 * v + 3);
}
LTS_D float grad(const int* __restrict__ perm, int x, int y, int z, float dx, float dy, float dz) {
    int h = perm[perm[perm[x] + y] + z];
    h &= 15;
    float u = h < 8 || h == 12 || h == 13 ? dx : dy;
    float v = h < 4 || h == 12 || h == 13 ? dy : dz;
    return ((h & 1) ? -u : u) + ((h & 2) ? -v : v);
}
LTS_D float perlin(const int* __restrict__ P, V3 p) {
    int ix = (int)floorf(p.x), iy = (int)floorf(p.y), iz = (int)floorf(p.z);
    float dx = p.x - ix, dy = p.y - iy, dz = p.z - iz;
    ix &= 255; iy &= 255; iz &= 255;
    float w000 = grad(P, ix, iy, iz, dx, dy, dz);
    float w100 = grad(P, ix + 1, iy, iz, dx - 1, dy, dz);
    float w010 = grad(P, ix, iy + 1, iz, dx, dy - 1, dz);
    float w001 = grad(P, ix, iy, iz + 1, dx, dy, dz - 1);
    float w110 = grad(P, ix + 1, iy + 1, iz, dx - 1, dy - 1, dz);
    float w011 = grad(P, ix, iy + 1, iz + 1, dx, dy - 1, dz - 1);
    float w101 = grad(P, ix + 1, iy, iz + 1, dx - 1, dy, dz - 1);
    float w111 = grad(P, ix + 1, iy + 1, iz + 1, dx - 1, dy - 1, dz - 1);
    float wx = noise_weight(dx), wy = noise_weight(dy), wz = noise_weight(dz);
    float x00 = lerpf(wx, w000, w100), x10 = lerpf(wx, w010, w110);
    float x01 = lerpf(wx, w001, w101), x11 = lerpf(wx, w011, w111);
    float y0 = lerpf(wy, x00, x10), y1 = lerpf(wy, x01, x11);
    return lerpf(wz, y0, y1);
}
LTS_D float fbm(const int* P, V3 p, V3 dpdx, V3 dpdy, float omega, int maxOct) {
    float len2 = fmaxf(lengthSq(dpdx), lengthSq(dpdy));
    float n = clampf(-1 - 0.5f * f_log2(len2), 0, (float)maxOct);
    int nInt = (int)floorf(n);
    float sum = 0, lambda = 1, o = 1;
    for (int i = 0; i < nInt; i++) { sum += o * perlin(P, p * lambda); lambda *= 1.99f; o *= omega; }
    float nPartial = n - nInt;
    sum += o * smooth_step(0.3f, 0.7f, nPartial) * perlin(P, p * lambda);
    return sum;
}
LTS_D float turbulence(const int* P, V3 p, V3 dpdx, V3 dpdy, float omega, int maxOct) {
    float len2 = fmaxf(lengthSq(dpdx), lengthSq(dpdy));
    float n = clampf(-1 - 0.5f * f_log2(len2), 0, (float)maxOct);
    int nInt = (int)floorf(n);
    float sum = 0, lambda = 1, o = 1;
    for (int i = 0; i < nInt; i++) { sum += o * fabsf(perlin(P, p * lambda)); lambda *= 1.99f; o *= omega; }
    float nPartial = n - nInt;
    sum += o * lerpf(smooth_step(0.3f, 0.7f, nPartial), 0.2f, fabsf(perlin(P, p * lambda)));
    for (int i = nInt; i < maxOct; i++) { sum += o * 0.2f; o *= omega; }
    return sum;
}
LTS_D Spec marble_tex(const int* P, const TexDesc& X, V3 p, V3 dpdx, V3 dpdy) {
    // MarbleTexture colours (Procedural.cuh:278-282); the zero row stands in for the read past the
    // table at t == 1, where its weight is exactly 0
    const float C[10][3] = {{.58f, .58f, .6f}, {.58f, .58f, .6f}, {.58f, .58f, .6f}, {.5f, .5f, .5f},
                            {.6f, .59f, .58f}, {.58f, .58f, .6f}, {.58f, .58f, .6f}, {.2f, .2f, .33f},
                            {.58f, .58f, .6f}, {0, 0, 0}};
    p = p * X.scale;
    float m = p.y + X.variation * fbm(P, p, dpdx * X.scale, dpdy * X.scale, X.omega, X.octaves);
    float t = 0.5f + 0.5f * f_sin(m);
    int first = (int)floorf(t * 6.0f);
    first = first < 0 ? 0 : (first > 6 ? 6 : first);  // t is in [0, 1]; the clamp only fences NaN
    t = (t * 6.0f - first);
    Spec c0 = spec(C[first][0], C[first][1], C[first][2]);
    Spec c1 = spec(C[first + 1][0], C[first + 1][1], C[first + 1][2]);
    Spec c2 = spec(C[first + 2][0], C[first + 2][1], C[first + 2][2]);
    Spec c3 = spec(C[first + 3][0], C[first + 3][1], C[first + 3][2]);
    Spec s0 = (1.0f - t) * c0 + t * c1;
    Spec s1 = (1.0f - t) * c1 + t * c2;
    Spec s2 = (1.0f - t) * c2 + t * c3;
    s0 = (1.0f - t) * s0 + t * s1;
    s1 = (1.0f - t) * s1 + t * s2;
    return 1.5f * ((1.0f - t) * s0 + t * s1);
}

// Mipmap texel / triangle / lookUp (materials/textures/Mipmap.cuh:14-72)
LTS_D Spec mip_texel(const SceneView& sv, const MipDesc& m, int s, int t, int l) {
    const MipLevelDesc& L = m.lv[l];
    if (m.filtered) {
        // row-major texels; repeat wraps to [0, res), clamp clamps, black is zero outside
        if (m.wrap == 0) { s = ((s % L.ures) + L.ures) % L.ures; t = ((t % L.vres) + L.vres) % L.vres; }
        else if (m.wrap == 2) { s = min(max(s, 0), L.ures - 1); t = min(max(t, 0), L.vres - 1); }
        else if (s < 0 || s >= L.ures || t < 0 || t >= L.vres) return spec(0.0f);
        const float* px = sv.mipData + 3 * (L.offset + (long long)t * L.ures + s);
        return spec(px[0], px[1], px[2]);
    }
    if (m.wrap == 0) { s = (int)fmodf((float)s, (float)L.ures); t = (int)fmodf((float)t, (float)L.vres); }
    else if (m.wrap == 2) { s = (int)clampf((float)s, 0, (float)(L.ures - 1)); t = (int)clampf((float)t, 0, (float)(L.vres - 1)); }
    else if (s < 0 || s >= L.ures || t < 0 || t >= L.vres) return spec(0.0f);
    if (s < 0 || t < 0 || s >= L.ures || t >= L.vres) return spec(0.0f);
    long long slot = L.offset + 16LL * ((long long)L.ublocks * (t >> 2) + (s >> 2)) + 4LL * (t >> 2) + (s & 3);
    const float* px = sv.mipData + 3 * slot;
    return spec(px[0], px[1], px[2]);
}
LTS_D Spec mip_triangle(const SceneView& sv, const MipDesc& m, int l, V2 st) {
    l = (int)clampf((float)l, 0, (float)(m.levels - 1));
    float s = st.x * m.lv[l].ures - 0.5f;
    float t = st.y * m.lv[l].vres - 0.5f;
    int s0 = (int)floorf(s), t0 = (int)floorf(t);
    float ds = s - s0, dt = t - t0;
    return (1 - ds) * (1 - dt) * mip_texel(sv, m, s0, t0, l) + (1 - ds) * dt * mip_texel(sv, m, s0, t0 + 1, l) +
           ds * (1 - dt) * mip_texel(sv, m, s0 + 1, t0, l) + ds * dt * mip_texel(sv, m, s0 + 1, t0 + 1, l);
}
// The reference's fixed width 0.01 gives a precomputed level; the filtered pipeline takes the
// width of the lookup and reads the top level correctly (the reference swaps texel's arguments
// there: texel (levels - 1, 0) of level 0).
LTS_D Spec mip_lookup(const SceneView& sv, const MipDesc& m, V2 st, float width) {
    float levels = m.filtered ? (float)(m.levels - 1) + (float)log2((double)fmaxf(width, (float)1e-8)) : m.lookupLevel;
    if (levels < 0) return mip_triangle(sv, m, 0, st);
    if (levels >= m.levels - 1) return m.filtered ? mip_texel(sv, m, 0, 0, m.levels - 1) : mip_texel(sv, m, m.levels - 1, 0, 0);
    int iLevel = (int)floorf(levels);
    float delta = levels - iLevel;
    return lerpS(delta, mip_triangle(sv, m, iLevel, st), mip_triangle(sv, m, iLevel + 1, st));
}

// `perm` is the Perlin permutation (sv.perm or a copy of it in LDS).
LTS_D Spec eval_tex(const SceneView& sv, int ti, const SurfacePoint& s, const int* perm) {
    const TexDesc& X = sv.texs[ti];
    switch (X.type) {
        case TEX_CONST: return spec(X.c[0], X.c[1], X.c[2]);
        case TEX_IMAGE: {
            // filtered pipeline: PBRT's isotropic width 2 max(|ds/dx|, |dt/dx|, |ds/dy|, |dt/dy|) of
            // UVMapping2D (st = uv, Mapping.cuh:19-37); the reference's lookup ignores it
            const float width = 2 * fmaxf(fmaxf(fabsf(s.dudx), fabsf(s.dvdx)), fmaxf(fabsf(s.dudy), fabsf(s.dvdy)));
            return mip_lookup(sv, sv.mips[X.img], v2(1 * s.uv.x + 0, 1 * s.uv.y + 0), width);
        }
        case TEX_FBM: return spec(fbm(perm, identityMap(s.p), identityMap(s.dpdx), identityMap(s.dpdy), X.omega, X.octaves));
        case TEX_WRINKLED: return spec(turbulence(perm, identityMap(s.p), identityMap(s.dpdx), identityMap(s.dpdy), X.omega, X.octaves));
        case TEX_MARBLE: return marble_tex(perm, X, identityMap(s.p), identityMap(s.dpdx), identityMap(s.dpdy));
    }
    return spec(0.0f);
}

// Material::computeScatteringFunctions (materials/Material.cuh:31-130)
LTS_D Bsdf make_bsdf(const SceneView& sv, const SurfacePoint& s, const int* perm) {
    Bsdf b;
    b.ns = s.sn; b.ng = s.n;
    b.ss = normalize(s.sdpdu);
    b.ts = cross(b.ns, b.ss);
    b.kind = -1; b.type = 0;
    b.R = b.T = spec(0.0f);
    b.A = b.B = b.eta = b.alpha = 0;
    const MatDesc& M = sv.mats[sv.shade[s.leaf].mat];
    switch (M.type) {
        case MAT_MATTE: {
            Spec kd = eval_tex(sv, M.tex0, s, perm);
            if (!isBlack(kd)) { b.R = kd; b.type = B_REFL | B_DIFF; b.kind = M.extra == 0 ? 0 : 1; b.A = M.onA; b.B = M.onB; }
            break;
        }
        case MAT_GLASS: {
            Spec r = eval_tex(sv, M.tex0, s, perm), t = eval_tex(sv, M.tex1, s, perm);
            if (isBlack(r) && isBlack(t)) break;
            b.kind = 3; b.type = B_REFL | B_TRANS | B_SPEC; b.R = r; b.T = t; b.eta = M.extra;
            break;
        }
        case MAT_MIRROR: {
            Spec r = eval_tex(sv, M.tex0, s, perm);
            if (isBlack(r)) break;
            b.kind = 2; b.type = B_REFL | B_SPEC; b.R = r;
            break;
        }
        case MAT_MICROFACET: {
            Spec r = eval_tex(sv, M.tex0, s, perm);
            if (isBlack(r)) break;
            b.kind = 4; b.type = B_REFL | B_GLOSSY; b.R = r; b.alpha = M.alpha;
            break;
        }
    }
    return b;
}

// ---------------------------------------------------------------- lights (lights/Light.cu)
struct LightPoint { V3 p, n, pError; };

LTS_D LightPoint tri_sample(const SceneView& sv, const LightDesc& L, V2 u) {
    float t = sqrtf(u.x);
    V2 b = v2(1 - t, u.y * t);
    V3 p0, p1, p2; int prim;
    load_tri(sv.tris, L.leaf, p0, p1, p2, prim);
    const ShadeTri& st = sv.shade[L.leaf];
    LightPoint lp;
    lp.p = p0 * b.x + p1 * b.y + p2 * (1 - b.x - b.y);
    if (st.hasN)
        lp.n = normalize(v3(st.n[0], st.n[1], st.n[2]) * b.x + v3(st.n[3], st.n[4], st.n[5]) * b.y +
                         v3(st.n[6], st.n[7], st.n[8]) * (1 - b.x - b.y));
    else
        lp.n = normalize(cross(p1 - p0, p2 - p0));
    V3 pAbs = vabs(p0 * b.x) + vabs(p1 * b.y) + vabs(p2 * (1 - b.x - b.y));
    lp.pError = pAbs * gammaN(6);
    return lp;
}

// Triangle::Pdf(it, wi) (shapes/Triangle.cu:211-225)
LTS_D float tri_pdf(const SceneView& sv, const LightDesc& L, const SurfacePoint& it, V3 wi) {
    Ray r = spawn_ray(it, wi);
    V3 p0, p1, p2; int prim;
    load_tri(sv.tris, L.leaf, p0, p1, p2, prim);
    float t, u, v, id;
    if (!tri_hit(p0, p1, p2, r, t, u, v, id)) return 0;
    SurfacePoint is = make_surface(sv, L.leaf, r, u, v, id);
    return lengthSq(it.p - is.p) / (absDot(is.n, -wi) * L.area);
}

LTS_D float inf_pdf(V3 wi) {
    float sinTheta = f_sin(f_acos(clampf(wi.z, -1.0f, 1.0f)));
    return (0.5f * INV_PI_F) / (2 * PI_F * PI_F * sinTheta);
}

LTS_D Spec light_le(const LightDesc& L) { return spec(L.le[0], L.le[1], L.le[2]); }

LTS_D Spec sample_li(const SceneView& sv, const LightDesc& L, const SurfacePoint& it, V2 u, V3* wi, float* pdf, LightPoint* target) {
    if (L.kind == 0) {
        LightPoint ps = tri_sample(sv, L, u);
        *wi = normalize(ps.p - it.p);
        *pdf = tri_pdf(sv, L, it, *wi);
        *target = ps;
        return dot(ps.n, -*wi) > 0.f ? light_le(L) : spec(0.0f);
    }
    V3 w = uniform_hemisphere(u);
    V3 ref = v3(0, 1, 0), a, b;
    coordinateSystem(ref, &a, &b);
    w = a * w.x + b * w.y + ref * w.z;
    V3 pL = v3(L.center[0], L.center[1], L.center[2]) + w * L.radius;
    *wi = normalize(pL - it.p);
    *pdf = inf_pdf(*wi);
    target->p = pL; target->n = v3(0, 0, 0); target->pError = v3(0, 0, 0);
    return light_le(L);
}

LTS_D bool unoccluded(const SceneView& sv, const SurfacePoint& it, const LightPoint& p1, TraceStack ts) {
    V3 origin = offsetRayOrigin(it.p, it.pError, it.n, p1.p - it.p);
    V3 target = offsetRayOrigin(p1.p, p1.pError, p1.n, origin - p1.p);
    Ray r; r.o = origin; r.d = target - origin; r.tMax = 1 - SHADOW_EPS;
    return !trace_any(sv, r, ts);
}

// Distribution1D::sampleDiscrete with the reference's search length (Sampling.cuh:155-188)
LTS_D int sample_light(const SceneView& sv, float u, float* pdf) {
    int size = sv.nLights, off;
    if (size == 1) off = 0;
    else {
        int first = 0, len = size;
        while (len > 0) {
            int half = len >> 1, middle = first + half;
            if (sv.lcdf[middle] <= u) { first = middle + 1; len -= half + 1; }
            else len = half;
        }
        off = (int)clampf((float)(first - 1), 0, (float)(size - 2));
    }
    *pdf = sv.lfunc[off] / (sv.funcInt * size);
    return off;
}

// The light test of the MIS-BSDF ray (IntegratorHelper.cuh:103-122): whether the closest hit of r
// (spawned with tMax = infinity) is light li's triangle, seen from its emitting side, or for the
// infinite light whether r escapes. The closest-hit traversal is not needed for that. A ray that
// misses the light's triangle, or meets its back, contributes nothing whatever lies in front. A ray
// that meets its front at t reaches it unless some triangle would win the closest hit against it
// (nearer, or at exactly t with a lower leaf index, tri_hit_closest), which an any-hit traversal
// bounded at t answers; the same t, u, v and 1/det come out of the triangle test whatever tMax is.
LTS_D bool mis_reaches_light(const SceneView& sv, const LightDesc& L, int li, Ray r, TraceStack ts,
                             TraceTally* tally = nullptr) {
    if (L.kind == 0) {
        if (L.leaf < 0 || sv.shade[L.leaf].light != li) return false;
        V3 p0, p1, p2; int prim;
        load_tri(sv.tris, L.leaf, p0, p1, p2, prim);
        if (tally) tally->tris++;
        float t, u, v, id;
        if (!tri_hit(p0, p1, p2, r, t, u, v, id)) return false;
        const SurfacePoint lt = make_surface(sv, L.leaf, r, u, v, id);
        if (!(dot(lt.n, -r.d) > 0.f)) return false;
        r.tMax = t;
        return !trace_any<true>(sv, r, ts, tally, L.leaf);
    }
    return !trace_any<true>(sv, r, ts, tally, -1);
}

// estimateDirect (intergrators/IntegratorHelper.cuh:51-125)
LTS_D Spec estimate_direct(const SceneView& sv, const SurfacePoint& si, const Bsdf& bsdf, V2 uScatter, int li, V2 uLight, bool specular,
                           TraceStack ts) {
    const LightDesc& L = sv.lights[li];
    Spec Ld = spec(0.0f);
    int flags = specular ? B_ALL : (B_ALL & ~B_SPEC);
    V3 wi;
    float lightPdf = 0.0f, scatteringPdf = 0.0f;
    LightPoint vis;
    Spec Li = sample_li(sv, L, si, uLight, &wi, &lightPdf, &vis);
    if (lightPdf > 0 && !isBlack(Li)) {
        Spec f = bsdf.f(si.wo, wi, flags) * absDot(wi, si.sn);
        scatteringPdf = bsdf.pdf(si.wo, wi, flags);
        if (!isBlack(f)) {
            if (!unoccluded(sv, si, vis, ts)) Li = spec(0.0f);
            if (!isBlack(Li)) {
                float weight = power_heuristic(lightPdf, scatteringPdf);
                Ld = Ld + f * Li * weight / lightPdf;
            }
        }
    }
    int sampledType = 0;
    Spec f = bsdf.sampleF(si.wo, &wi, uScatter, &scatteringPdf, flags, &sampledType);
    f = f * absDot(wi, si.sn);
    bool sampledSpecular = (sampledType & B_SPEC) != 0;
    if (!isBlack(f) && scatteringPdf > 0) {
        float weight = 1;
        if (!sampledSpecular) {
            lightPdf = L.kind == 0 ? tri_pdf(sv, L, si, wi) : inf_pdf(wi);
            if (lightPdf == 0) return Ld;
            weight = power_heuristic(scatteringPdf, lightPdf);
        }
        const Spec Li2 = mis_reaches_light(sv, L, li, spawn_ray(si, wi), ts) ? light_le(L) : spec(0.f);
        if (!isBlack(Li2)) Ld = Ld + f * Li2 * spec(1.0f) * weight / scatteringPdf;
    }
    return Ld;
}

// Film sample position of (pixel, sample): dims 0-1, stratified per Sampler::stratifiedSample
// (sampling/Sampler.cuh:49-64) when requested.
LTS_D V2 film_sample(int px, int py, int sample, int spp, int stratified, const float d0[4]) {
    V2 film = v2(d0[0], d0[1]);
    if (stratified) {
        int strata = (int)floorf(sqrtf((float)spp));
        if (sample - 1 < strata * strata) {
            float inv = (float)1 / strata;
            film = v2(film.x * inv, film.y * inv);
            film = film + v2((sample % strata) / (float)strata, (sample / strata) / (float)strata);
        }
    }
    return v2((float)(px + film.x), (float)(py + film.y));
}

// ---------------------------------------------------------------- camera (rendering/Camera.cu:51-81)
struct CamRays {
    Ray r;
    V3 rxO, rxD, ryO, ryD;
};

LTS_D void camera_ray(const CameraDesc& c, V2 pf, V2 lens, V3* o, V3* d) {
    V3 cu = v3(c.u[0], c.u[1], c.u[2]), cv = v3(c.v[0], c.v[1], c.v[2]);
    V3 co = v3(c.o[0], c.o[1], c.o[2]);
    V3 offset = (cu * lens.x + cv * lens.y) * c.lensRadius;
    V2 pr = v2(pf.x / (c.width - 1), pf.y / (c.height - 1));
    *o = co + offset;
    *d = normalize(v3(c.llc[0], c.llc[1], c.llc[2]) + v3(c.right[0], c.right[1], c.right[2]) * pr.x +
                   v3(c.up[0], c.up[1], c.up[2]) * pr.y - co - offset);
}

// generateDifferentialRay + Ray::scaleDifferentials(rsqrt(spp)) (Camera.cu:67-81, Ray.cuh:35-40)
LTS_D CamRays camera_rays(const CameraDesc& c, V2 pFilm, V2 lens, int spp) {
    CamRays cr;
    camera_ray(c, pFilm, lens, &cr.r.o, &cr.r.d);
    camera_ray(c, pFilm + v2(1, 0), lens, &cr.rxO, &cr.rxD);
    camera_ray(c, pFilm + v2(0, 1), lens, &cr.ryO, &cr.ryD);
    cr.r.tMax = INFINITY;
    float sc = quake((float)spp);
    cr.rxO = cr.r.o + (cr.rxO - cr.r.o) * sc;
    cr.ryO = cr.r.o + (cr.ryO - cr.r.o) * sc;
    cr.rxD = cr.r.d + (cr.rxD - cr.r.d) * sc;
    cr.ryD = cr.r.d + (cr.ryD - cr.r.d) * sc;
    return cr;
}

// Camera rays of one (pixel, sample): film sample from dims 0-1, lens sample from dims 2-3.
LTS_D CamRays camera_sample(const SceneView& sv, int px, int py, int sample, int spp, int stratified, uint32_t seed) {
    Sampler smp;
    smp.seed = seed;
    smp.pixel = (uint32_t)(py * sv.cam.width + px);
    smp.sample = (uint32_t)sample;
    float d[4];
    smp.block(0, d);
    V2 pFilm = film_sample(px, py, sample, spp, stratified, d);
    V2 lens = uniform_disk(v2(d[2], d[3]));
    return camera_rays(sv.cam, pFilm, lens, spp);
}

// invDet * determinant of a hit: Triangle::intersect scales dpdu and dpdv by the ray's invDet where
// the uv parametrisation's 1 / determinant belongs (Triangle.cu:147-149), so the interaction's dpdu,
// dpdv are k times the true derivatives (k = 1 when the uv determinant is 0 and dpdu, dpdv come
// from coordinateSystem). The shading frame only uses their directions; uv differentials need k.
LTS_D float uv_derivative_scale(const SceneView& sv, int leaf, float invDet) {
    const ShadeTri& st = sv.shade[leaf];
    V2 uv0 = v2(st.uv[0], st.uv[1]), uv1 = v2(st.uv[2], st.uv[3]), uv2 = v2(st.uv[4], st.uv[5]);
    V2 duv02 = uv0 - uv2, duv12 = uv1 - uv2;
    float determinant = duv02.x * duv12.y - duv02.y * duv12.x;
    return determinant == 0 ? 1.0f : invDet * determinant;
}

// SurfaceInteraction::computeDifferentials (Interaction.cuh:118-160). The uv differentials, read
// only by the filtered texture pipeline, are those of the true uv parametrisation: solved against
// the interaction's dpdu, dpdv (k times too long, see uv_derivative_scale) and multiplied by k.
LTS_D void differentials(SurfacePoint& s, const CamRays& cr, float k) {
    float d = -dot(s.n, s.p);
    float tx = (-dot(s.n, cr.rxO) - d) / dot(s.n, cr.rxD);
    V3 px = cr.rxO + cr.rxD * tx;
    float ty = (-dot(s.n, cr.ryO) - d) / dot(s.n, cr.ryD);
    V3 py = cr.ryO + cr.ryD * ty;
    s.dpdx = px - s.p;
    s.dpdy = py - s.p;
    // du/dx, dv/dx, du/dy, dv/dy on the two axes other than the normal's dominant one
    // (Interaction.cuh:134-153, solveLinearSystem2x2 GeneralHelper.cuh:127-134)
    int d0, d1;
    if (fabsf(s.n.x) > fabsf(s.n.y) && fabsf(s.n.x) > fabsf(s.n.z)) { d0 = 1; d1 = 2; }
    else if (fabsf(s.n.y) > fabsf(s.n.z)) { d0 = 0; d1 = 2; }
    else { d0 = 0; d1 = 1; }
    auto c = [](V3 v, int k) { return k == 0 ? v.x : k == 1 ? v.y : v.z; };
    const float a00 = c(s.dpdu, d0), a01 = c(s.dpdv, d0), a10 = c(s.dpdu, d1), a11 = c(s.dpdv, d1);
    auto solve = [&](float b0, float b1, float& x0, float& x1) {
        const float det = a00 * a11 - a01 * a10;
        if (fabsf(det) < 1e-10f) { x0 = x1 = 0.0f; return; }
        x0 = (a11 * b0 - a01 * b1) / det;
        x1 = (a00 * b1 - a10 * b0) / det;
        if (isnan(x0) || isnan(x1)) x0 = x1 = 0.0f;
    };
    solve(c(px, d0) - c(s.p, d0), c(px, d1) - c(s.p, d1), s.dudx, s.dvdx);
    solve(c(py, d0) - c(s.p, d0), c(py, d1) - c(s.p, d1), s.dudy, s.dvdy);
    s.dudx *= k; s.dvdx *= k; s.dudy *= k; s.dvdy *= k;
}

}  // namespace lts
