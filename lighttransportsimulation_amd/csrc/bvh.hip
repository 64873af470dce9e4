// Karras-2012 LBVH build on gfx950, bit-exact with the reference's construction
// (acceleration/BVH.cuh:77-452, acceleration/BVH_kernels.cu:8-50):
//   k_prim_bounds  primitive AABB + centroid inputs, block-reduced scene bounds     (BVH.cuh:92-98)
//   k_morton       30-bit Morton codes of the normalised centroids                   (BVH.cuh:101-137)
//   radix sort     stable LSD sort of (code, primitive) pairs, 8-bit digits, 4 passes (BVH.cuh:140-146)
//                  count -> per-digit scan over tiles -> digit bases -> LDS-ranked scatter
//   k_karras       interior node i: direction, range end, split                     (BVH.cuh:173-236)
//   k_refit_l1/l2  leaf-to-root climb, second arriver unions its children            (BVH.cuh:238-262):
//   /global        subtrees inside 1024-leaf chunks in LDS, then 64K-leaf chunks with LDS arrival
//                  bits, then the few nodes crossing 64K boundaries with device-scope ordering
//   k_pack_*       traversal layout: both child boxes per node, triangles in leaf order
// The scene-bound reduction uses a total order on floats (-0 below +0); it equals the reference's
// float atomics (GeneralHelper.cuh:150-160) on every input without -0.0.
#include <algorithm>
#include <mutex>
#include <type_traits>

#include "lts_device.h"
#include "lts_internal.h"

namespace lts {

namespace {

constexpr int kBlock = 256;
constexpr int kItems = 16;
constexpr int kTile = kBlock * kItems;  // keys per radix tile

__device__ __forceinline__ uint32_t order_key(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float order_val(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// Triangle::worldBound = bUnion(Bounds3f(p0, p1), p2) with the Bounds3 constructor re-applied.
__device__ __forceinline__ void tri_box(const float* t, float* mn, float* mx) {
#pragma unroll
    for (int a = 0; a < 3; a++) {
        float lo = fminf(t[a], t[3 + a]), hi = fmaxf(t[a], t[3 + a]);
        float lo2 = fminf(lo, t[6 + a]), hi2 = fmaxf(hi, t[6 + a]);
        mn[a] = fminf(lo2, hi2);
        mx[a] = fmaxf(lo2, hi2);
    }
}

__global__ __launch_bounds__(kBlock) void k_prim_bounds(const float* __restrict__ soup, int n, float* __restrict__ primBox,
                                                        uint32_t* __restrict__ partial, int32_t* __restrict__ status) {
    __shared__ uint32_t red[6][kBlock / 64];
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i == 0) *status = 0;  // the build's first kernel: later kernels on the stream see the cleared flags
    uint32_t k[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
    if (i < n) {
        float t[9];
#pragma unroll
        for (int j = 0; j < 9; j++) t[j] = soup[9 * (size_t)i + j];
        float mn[3], mx[3];
        tri_box(t, mn, mx);
#pragma unroll
        for (int a = 0; a < 3; a++) {
            primBox[6 * (size_t)i + a] = mn[a];
            primBox[6 * (size_t)i + 3 + a] = mx[a];
            k[a] = order_key(mn[a]);
            k[3 + a] = order_key(mx[a]);
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            k[a] = min(k[a], (uint32_t)__shfl_xor((int)k[a], off));
            k[3 + a] = max(k[3 + a], (uint32_t)__shfl_xor((int)k[3 + a], off));
        }
    }
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        for (int a = 0; a < 6; a++) red[a][wave] = k[a];
    __syncthreads();
    if (threadIdx.x < 6) {
        int a = threadIdx.x;
        uint32_t v = red[a][0];
        for (int w = 1; w < kBlock / 64; w++) v = a < 3 ? min(v, red[a][w]) : max(v, red[a][w]);
        partial[6 * (size_t)blockIdx.x + a] = v;
    }
}

// One block reduces the per-block partial keys into the six scene-bound keys.
__global__ __launch_bounds__(1024) void k_reduce_scene_keys(const uint32_t* __restrict__ partial, int blocks,
                                                            uint32_t* __restrict__ sceneKeys) {
    __shared__ uint32_t red[6][16];
    uint32_t k[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
    for (int b = threadIdx.x; b < blocks; b += 1024) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            k[a] = min(k[a], partial[6 * (size_t)b + a]);
            k[3 + a] = max(k[3 + a], partial[6 * (size_t)b + 3 + a]);
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            k[a] = min(k[a], (uint32_t)__shfl_xor((int)k[a], off));
            k[3 + a] = max(k[3 + a], (uint32_t)__shfl_xor((int)k[3 + a], off));
        }
    }
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        for (int a = 0; a < 6; a++) red[a][wave] = k[a];
    __syncthreads();
    if (threadIdx.x < 6) {
        int a = threadIdx.x;
        uint32_t v = red[a][0];
        for (int w = 1; w < 16; w++) v = a < 3 ? min(v, red[a][w]) : max(v, red[a][w]);
        sceneKeys[a] = v;
    }
}

// leftShift3 (BVH.cuh:101-109)
__device__ __forceinline__ uint32_t left_shift3(uint32_t x) {
    if (x == (1u << 10)) --x;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

__global__ __launch_bounds__(kBlock) void k_morton(const float* __restrict__ primBox, int n, const uint32_t* __restrict__ sceneKeys,
                                                   uint32_t* __restrict__ codes, int32_t* __restrict__ ids) {
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float gmn[3], gmx[3];
#pragma unroll
    for (int a = 0; a < 3; a++) { gmn[a] = order_val(sceneKeys[a]); gmx[a] = order_val(sceneKeys[3 + a]); }
    uint32_t code = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        float c = primBox[6 * (size_t)i + a] * 0.5f + primBox[6 * (size_t)i + 3 + a] * 0.5f;  // centroid
        float o = c - gmn[a];
        if (gmx[a] > gmn[a]) o /= gmx[a] - gmn[a];                                         // Bounds3::offset
        float s = (float)(1 << 10) * o;
        code |= left_shift3((uint32_t)s) << a;
    }
    codes[i] = code;
    ids[i] = i;
}

// ---------------------------------------------------------------- radix sort
__global__ __launch_bounds__(kBlock) void k_radix_count(const uint32_t* __restrict__ keys, int n, int shift, int tiles,
                                                        uint32_t* __restrict__ counts) {
    __shared__ uint32_t hist[256];
    hist[threadIdx.x] = 0;
    __syncthreads();
    size_t base = (size_t)blockIdx.x * kTile;
#pragma unroll
    for (int j = 0; j < kItems; j++) {
        size_t idx = base + (size_t)j * kBlock + threadIdx.x;
        if (idx < (size_t)n) atomicAdd(&hist[(keys[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    counts[(size_t)threadIdx.x * tiles + blockIdx.x] = hist[threadIdx.x];
}

// Block-wide exclusive scan of one value per thread (256 threads); returns the total.
__device__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* out, uint32_t* scratch) {
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = (uint32_t)__shfl_up((int)incl, off);
        if (lane >= off) incl += y;
    }
    if (lane == 63) scratch[wave] = incl;
    __syncthreads();
    uint32_t wbase = 0, total = 0;
    for (int w = 0; w < kBlock / 64; w++) {
        if (w < wave) wbase += scratch[w];
        total += scratch[w];
    }
    *out = wbase + incl - v;
    __syncthreads();
    return total;
}

// One block per digit: exclusive scan of that digit's per-tile counts, total into digitBase.
__global__ __launch_bounds__(kBlock) void k_radix_scan_tiles(uint32_t* __restrict__ counts, int tiles, uint32_t* __restrict__ digitBase) {
    __shared__ uint32_t scratch[kBlock / 64];
    uint32_t* row = counts + (size_t)blockIdx.x * tiles;
    int per = (tiles + kBlock - 1) / kBlock;
    int beg = threadIdx.x * per, end = min(beg + per, tiles);
    uint32_t sum = 0;
    for (int t = beg; t < end; t++) sum += row[t];
    uint32_t excl;
    uint32_t total = block_exclusive_scan(sum, &excl, scratch);
    for (int t = beg; t < end; t++) { uint32_t c = row[t]; row[t] = excl; excl += c; }
    if (threadIdx.x == 0) digitBase[blockIdx.x] = total;
}

__global__ __launch_bounds__(kBlock) void k_radix_scan_digits(uint32_t* __restrict__ digitBase) {
    __shared__ uint32_t scratch[kBlock / 64];
    uint32_t v = digitBase[threadIdx.x], excl;
    block_exclusive_scan(v, &excl, scratch);
    digitBase[threadIdx.x] = excl;
}

// Stable scatter of one 4096-key tile: ranks from wave ballots, tile-local order staged in LDS,
// then written out in runs per digit.
__global__ __launch_bounds__(kBlock) void k_radix_scatter(const uint32_t* __restrict__ keysIn, const int32_t* __restrict__ valsIn,
                                                          uint32_t* __restrict__ keysOut, int32_t* __restrict__ valsOut, int n,
                                                          int shift, int tiles, const uint32_t* __restrict__ counts,
                                                          const uint32_t* __restrict__ digitBase) {
    __shared__ uint32_t sKeys[kTile];
    __shared__ int32_t sVals[kTile];
    __shared__ uint32_t waveCnt[kBlock / 64][256];
    __shared__ uint32_t runTotal[256];
    __shared__ uint32_t tileStart[256];
    __shared__ uint32_t globalBase[256];
    __shared__ uint32_t scratch[kBlock / 64];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const size_t base = (size_t)blockIdx.x * kTile;
    const int valid = (int)min((size_t)kTile, (size_t)n - base);
    for (int w = 0; w < kBlock / 64; w++) waveCnt[w][t] = 0;
    runTotal[t] = 0;
    globalBase[t] = digitBase[t] + counts[(size_t)t * tiles + blockIdx.x];
    uint32_t key[kItems];
    int32_t val[kItems];
    uint32_t rank[kItems];
#pragma unroll
    for (int j = 0; j < kItems; j++) {
        size_t idx = base + (size_t)j * kBlock + t;
        bool ok = idx < (size_t)n;
        key[j] = ok ? keysIn[idx] : 0xffffffffu;  // tail keys rank last in digit 255 and are never written
        val[j] = ok ? valsIn[idx] : 0;
    }
    __syncthreads();
    const uint64_t ltMask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int j = 0; j < kItems; j++) {
        uint32_t d = (key[j] >> shift) & 255u;
        uint64_t peers = ~0ull;
#pragma unroll
        for (int b = 0; b < 8; b++) {
            bool bit = (d >> b) & 1u;
            uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        uint32_t inWave = (uint32_t)__popcll(peers & ltMask);
        int leader = __ffsll((long long)peers) - 1;
        if (lane == leader) waveCnt[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        uint32_t pre = runTotal[d];
        for (int w = 0; w < wave; w++) pre += waveCnt[w][d];
        rank[j] = pre + inWave;
        __syncthreads();
        uint32_t add = 0;
        for (int w = 0; w < kBlock / 64; w++) { add += waveCnt[w][t]; waveCnt[w][t] = 0; }
        runTotal[t] += add;
        __syncthreads();
    }
    uint32_t start;
    block_exclusive_scan(runTotal[t], &start, scratch);
    tileStart[t] = start;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kItems; j++) {
        uint32_t d = (key[j] >> shift) & 255u;
        uint32_t pos = tileStart[d] + rank[j];
        sKeys[pos] = key[j];
        sVals[pos] = val[j];
    }
    __syncthreads();
    for (int i = t; i < valid; i += kBlock) {
        uint32_t k = sKeys[i];
        uint32_t d = (k >> shift) & 255u;
        size_t dst = (size_t)globalBase[d] + (i - tileStart[d]);
        keysOut[dst] = k;
        valsOut[dst] = sVals[i];
    }
}

// ---------------------------------------------------------------- hierarchy
__device__ __forceinline__ int delta_at(const uint32_t* __restrict__ codes, int n, int i, int j) {
    if (i < 0 || i >= n || j < 0 || j >= n) return -1;
    uint32_t x = codes[i] ^ codes[j];
    int extra = 0;
    if (x == 0) { extra = 40; x = (uint32_t)(i ^ j); }
    return (x == 0 ? 32 : __clz((int)x)) + extra;
}

__global__ __launch_bounds__(kBlock) void k_karras(const uint32_t* __restrict__ codes, int n, int32_t* __restrict__ left,
                                                   int32_t* __restrict__ right, int32_t* __restrict__ parent,
                                                   int32_t* __restrict__ rangeLo, int32_t* __restrict__ rangeHi) {
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n - 1) return;
    int d = delta_at(codes, n, i, i + 1) - delta_at(codes, n, i, i - 1);
    d = d > 0 ? 1 : -1;
    int deltaMin = delta_at(codes, n, i, i - d);
    int lMax = 128;
    while (delta_at(codes, n, i, i + lMax * d) > deltaMin) lMax *= 4;
    int l = 0;
    for (int t = lMax / 2; t >= 1; t /= 2)
        if (delta_at(codes, n, i, i + (l + t) * d) > deltaMin) l += t;
    int j = i + l * d;
    int deltaNode = delta_at(codes, n, i, j);
    int s = 0;
    for (int t = (l + 1) >> 1; t >= 1; t = (t == 1) ? 0 : ((t + 1) >> 1))  // ceil-halving
        if (delta_at(codes, n, i, i + (s + t) * d) > deltaNode) s += t;
    int gamma = i + s * d + min(d, 0);
    int lo = min(i, j), hi = max(i, j);
    int lc = (lo == gamma) ? (n - 1) + gamma : gamma;
    int rc = (hi == gamma + 1) ? (n - 1) + gamma + 1 : gamma + 1;
    left[i] = lc;
    right[i] = rc;
    rangeLo[i] = lo;
    rangeHi[i] = hi;
    parent[lc] = i;
    parent[rc] = i;
    if (i == 0) parent[0] = -1;
}

__global__ void k_fill_i32(int32_t* p, int n, int32_t v) {
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = v;
}

// bUnion(Bounds3, Bounds3) followed by the Bounds3(p1, p2) constructor
__device__ __forceinline__ void union_box(const float* a, const float* b, float* o) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
        float lo = fminf(a[k], b[k]), hi = fmaxf(a[3 + k], b[3 + k]);
        o[k] = fminf(lo, hi);
        o[3 + k] = fmaxf(lo, hi);
    }
}

constexpr int kChunk = 1024;            // sorted leaves per level-1 workgroup (subtrees refit in LDS)
constexpr int kChunk2 = 64 * kChunk;     // sorted leaves per level-2 workgroup
// level-1 roots filed per level-2 chunk: each is a child of a node containing one of the
// kChunk2 / kChunk + 1 chunk boundaries of that range, and at most 58 nodes contain a boundary
constexpr int kBucketCap = 2 * (kChunk2 / kChunk + 1) * 58;
// Build status bits (BvhBuffers::status). The capacities above follow from the depth bound of a
// Karras tree over at most 2^26 leaves (58 nodes on a root-to-leaf path); a violated bound is
// reported instead of leaving boxes unrefitted.
constexpr int kStatusBucket = 1;  // a level-2 chunk received more than kBucketCap level-1 roots
constexpr int kStatusRoots = 2;   // more level-3 roots than primitives
constexpr int kStatusLevel3 = 4;  // more level-3 roots than level-3 climbers were launched
constexpr int kStatusClimb = 8;   // a climb took more than 128 steps

__device__ __forceinline__ void load_box(const float* box, int node, float* o) {
#pragma unroll
    for (int q = 0; q < 6; q++) o[q] = box[6 * (size_t)node + q];
}
__device__ __forceinline__ void store_box(float* box, int node, const float* o) {
#pragma unroll
    for (int q = 0; q < 6; q++) box[6 * (size_t)node + q] = o[q];
}

// Level 1: one workgroup per chunk of kChunk sorted leaves. Leaf boxes, the records of the
// chunk's interior nodes and the boxes of every interior node whose leaf range lies inside the
// chunk are held in LDS; the leaf-to-root climbs meet through LDS arrival bits. A climb that
// reaches a node crossing the chunk boundary files the node below it as a level-2 root.
__global__ __launch_bounds__(kChunk) void k_refit_l1(const float* __restrict__ primBox, const int32_t* __restrict__ perm, int n,
                                                     int mode, const int32_t* __restrict__ left, const int32_t* __restrict__ right,
                                                     const int32_t* __restrict__ parent, const int32_t* __restrict__ rangeLo,
                                                     const int32_t* __restrict__ rangeHi, float* __restrict__ box,
                                                     int32_t* __restrict__ bucket, int32_t* __restrict__ bucketCount,
                                                     int32_t* __restrict__ status) {
    __shared__ float lbox[kChunk * 6];
    __shared__ float ibox[kChunk * 6];
    __shared__ int32_t sl[kChunk], sr[kChunk], sp[kChunk];
    __shared__ uint32_t arrive[kChunk / 32], local[kChunk / 32];
    const int t = threadIdx.x;
    const int cs = blockIdx.x * kChunk, ce = min(cs + kChunk, n);
    if (t < kChunk / 32) { arrive[t] = 0; local[t] = 0; }
    __syncthreads();
    const int k = cs + t;
    if (k < ce) {
        const int prim = mode == 0 ? k : perm[k];
        float b[6];
        load_box(primBox, prim, b);
#pragma unroll
        for (int q = 0; q < 6; q++) lbox[6 * t + q] = b[q];
        store_box(box, (n - 1) + k, b);
    }
    if (k < ce && k < n - 1) {  // interior node k (an interior node is an end of its own range)
        sl[t] = left[k]; sr[t] = right[k]; sp[t] = parent[k];
        if (rangeLo[k] >= cs && rangeHi[k] < ce) atomicOr(&local[t >> 5], 1u << (t & 31));
    }
    __syncthreads();
    if (k < ce && n > 1) {
        int cur = (n - 1) + k;
        int node = parent[cur];
        int step = 0;
        for (; step < 128; step++) {
            const int j = node - cs;
            const bool isLocal = node >= cs && node < ce && node < n - 1 && ((local[j >> 5] >> (j & 31)) & 1u);
            if (!isLocal) {
                if (node >= 0) {
                    int b = cs / kChunk2;
                    int slot = atomicAdd(&bucketCount[b], 1);
                    if (slot < kBucketCap) bucket[(size_t)b * kBucketCap + slot] = cur;
                    else atomicOr(status, kStatusBucket);
                }
                break;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            uint32_t old = atomicOr(&arrive[j >> 5], 1u << (j & 31));
            if (!((old >> (j & 31)) & 1u)) break;  // first arrival
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            float a[6], c[6], o[6];
            const int l = sl[j], r = sr[j];
#pragma unroll
            for (int q = 0; q < 6; q++) {
                a[q] = l >= n - 1 ? lbox[6 * (l - (n - 1) - cs) + q] : ibox[6 * (l - cs) + q];
                c[q] = r >= n - 1 ? lbox[6 * (r - (n - 1) - cs) + q] : ibox[6 * (r - cs) + q];
            }
            union_box(a, c, o);
#pragma unroll
            for (int q = 0; q < 6; q++) ibox[6 * j + q] = o[q];
            if (node == 0) break;
            cur = node;
            node = sp[j];
        }
        if (step == 128) atomicOr(status, kStatusClimb);
    }
    __syncthreads();
    if (k < ce && k < n - 1 && ((local[t >> 5] >> (t & 31)) & 1u)) {
        float o[6];
#pragma unroll
        for (int q = 0; q < 6; q++) o[q] = ibox[6 * t + q];
        store_box(box, k, o);
    }
}

// Level 2: one workgroup per chunk of kChunk2 leaves climbs from its level-1 roots through the
// nodes inside the chunk (arrival bits in LDS, boxes in global memory, workgroup-scope fences).
__global__ __launch_bounds__(1024) void k_refit_l2(int n, const int32_t* __restrict__ left, const int32_t* __restrict__ right,
                                                   const int32_t* __restrict__ parent, const int32_t* __restrict__ rangeLo,
                                                   const int32_t* __restrict__ rangeHi, float* box, const int32_t* __restrict__ bucket,
                                                   const int32_t* __restrict__ bucketCount, int32_t* __restrict__ roots,
                                                   int32_t* __restrict__ rootCount, int32_t* __restrict__ status) {
    __shared__ uint32_t arrive[kChunk2 / 32];
    const int cs = blockIdx.x * kChunk2, ce = min(cs + kChunk2, n);
    for (int w = threadIdx.x; w < kChunk2 / 32; w += 1024) arrive[w] = 0;
    __syncthreads();
    const int cnt = min(bucketCount[blockIdx.x], kBucketCap);
    for (int e = threadIdx.x; e < cnt; e += 1024) {
        int cur = bucket[(size_t)blockIdx.x * kBucketCap + e];
        int node = parent[cur];
        int step = 0;
        for (; step < 128 && node >= 0 && node < n - 1; step++) {
            if (rangeLo[node] < cs || rangeHi[node] >= ce) {
                int slot = atomicAdd(rootCount, 1);
                if (slot < n) roots[slot] = cur;
                else atomicOr(status, kStatusRoots);
                break;
            }
            const int j = node - cs;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            uint32_t old = atomicOr(&arrive[j >> 5], 1u << (j & 31));
            if (!((old >> (j & 31)) & 1u)) break;
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            float a[6], c[6], o[6];
            load_box(box, left[node], a);
            load_box(box, right[node], c);
            union_box(a, c, o);
            store_box(box, node, o);
            if (node == 0) break;
            cur = node;
            node = parent[node];
        }
        if (step == 128) atomicOr(status, kStatusClimb);
    }
}

// Level 3: nodes whose leaf range crosses a level-2 chunk boundary, climbed from the level-2 roots
// with device-scope ordering.
__global__ __launch_bounds__(kBlock) void k_refit_global(int n, const int32_t* __restrict__ left, const int32_t* __restrict__ right,
                                                         const int32_t* __restrict__ parent, float* box, int32_t* flags,
                                                         const int32_t* __restrict__ roots, const int32_t* __restrict__ rootCount,
                                                         int32_t* __restrict__ status) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    const int count = *rootCount;
    if (t == 0 && count > (int)(gridDim.x * kBlock)) atomicOr(status, kStatusLevel3);
    if (t >= count || t >= n) return;
    int node = parent[roots[t]];
    for (int step = 0;; step++) {
        if (node < 0 || node >= n - 1) return;
        if (step == 128) { atomicOr(status, kStatusClimb); return; }
        __threadfence();
        int old = __hip_atomic_fetch_add(&flags[node], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (old == 0) return;
        __threadfence();
        float a[6], c[6], o[6];
        int l = left[node], r = right[node];
#pragma unroll
        for (int q = 0; q < 6; q++) {
            a[q] = __hip_atomic_load(&box[6 * (size_t)l + q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            c[q] = __hip_atomic_load(&box[6 * (size_t)r + q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        union_box(a, c, o);
#pragma unroll
        for (int q = 0; q < 6; q++) __hip_atomic_store(&box[6 * (size_t)node + q], o[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (node == 0) return;
        node = parent[node];
    }
}

__global__ __launch_bounds__(kBlock) void k_pack_nodes(const float* __restrict__ box, const int32_t* __restrict__ left,
                                                       const int32_t* __restrict__ right, int n, BvhNode* __restrict__ nodes) {
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n - 1) return;
    BvhNode nd;
    int l = left[i], r = right[i];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        nd.lmin[a] = box[6 * (size_t)l + a]; nd.lmax[a] = box[6 * (size_t)l + 3 + a];
        nd.rmin[a] = box[6 * (size_t)r + a]; nd.rmax[a] = box[6 * (size_t)r + 3 + a];
    }
    nd.left = l; nd.right = r; nd.pad0 = nd.pad1 = 0;
    nodes[i] = nd;
}

__global__ __launch_bounds__(kBlock) void k_pack_tris(const float* __restrict__ soup, const int32_t* __restrict__ perm, int n,
                                                      TriLeaf* __restrict__ tris) {
    int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    int p = perm[k];
    const float* t = soup + 9 * (size_t)p;
    TriLeaf o;
    o.p0x = t[0]; o.p0y = t[1]; o.p0z = t[2]; o.p1x = t[3];
    o.p1y = t[4]; o.p1z = t[5]; o.p2x = t[6]; o.p2y = t[7];
    o.p2z = t[8]; o.prim = p; o.mat = 0; o.pad1 = 0;
    tris[k] = o;
}

inline int blocks_for(long n) { return (int)((n + kBlock - 1) / kBlock); }

// ---------------------------------------------------------------- 4-wide render BVH
constexpr int kScanBlock = 1024;

// flag[i] = 1 when interior node i has even depth (the root has depth 0); maxDepth (when given)
// receives the deepest level of a child, i.e. the tree depth counted in edges down to the leaves.
__global__ __launch_bounds__(kBlock) void k_even_depth(const int32_t* __restrict__ parent, int n, int32_t* __restrict__ flag,
                                                       int32_t* __restrict__ maxDepth) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    int depth = -1;
    if (i < n - 1) {
        depth = 0;
        for (int node = i; node > 0 && depth < 2 * n; node = parent[node]) depth++;
        flag[i] = (depth & 1) == 0;
    }
    if (maxDepth) {
        int d = depth + 1;
        for (int o = 32; o > 0; o >>= 1) d = max(d, __shfl_xor(d, o));
        if ((threadIdx.x & 63) == 0 && d > 0) atomicMax(maxDepth, d);
    }
}

// In-place exclusive prefix sum of v[0, m) within blocks of kScanBlock; totals[block] = block sum.
__global__ __launch_bounds__(kScanBlock) void k_scan_local(int32_t* __restrict__ v, int m, int32_t* __restrict__ totals) {
    __shared__ int32_t sh[kScanBlock / 64];
    const int i = blockIdx.x * kScanBlock + threadIdx.x;
    const int x0 = i < m ? v[i] : 0;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = x0;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wave] = x;
    __syncthreads();
    int base = 0;
    for (int w = 0; w < wave; w++) base += sh[w];
    if (i < m) v[i] = base + x - x0;
    if (threadIdx.x == kScanBlock - 1) totals[blockIdx.x] = base + x;
}

// Exclusive prefix of the block totals (one block, sequential over at most a few thousand entries).
__global__ void k_scan_totals(int32_t* __restrict__ totals, int nb) {
    if (threadIdx.x != 0) return;
    int run = 0;
    for (int b = 0; b < nb; b++) { const int t = totals[b]; totals[b] = run; run += t; }
    totals[nb] = run;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_add(int32_t* __restrict__ v, int m, const int32_t* __restrict__ totals) {
    const int i = blockIdx.x * kScanBlock + threadIdx.x;
    if (i < m) v[i] += totals[blockIdx.x];
}

// Smallest exponent e in [-126, 127] with qdecode(o, e, 255) >= hi.
__device__ int q_exponent(float o, float hi) {
    int e = -126;
    const float ext = hi - o;
    if (ext > 0) {
        int ee;
        (void)frexpf(ext / 255.0f, &ee);  // ext / 255 < 2^ee
        e = max(-126, min(127, ee));
    }
    while (e < 127 && !(qdecode(o, e, 255) >= hi)) e++;
    while (e > -126 && qdecode(o, e - 1, 255) >= hi) e--;
    return e;
}

// Largest q with qdecode(o, e, q) <= lo (q = 0 always qualifies: o is the node's minimum).
__device__ unsigned q_floor(float o, int e, float lo) {
    const float d = (lo - o) / __int_as_float((e + 127) << 23);
    int q = d > 0 ? (int)fminf(floorf(d), 255.0f) : 0;
    while (q > 0 && qdecode(o, e, q) > lo) q--;
    while (q < 255 && qdecode(o, e, q + 1) <= lo) q++;
    return (unsigned)q;
}

// Smallest q with qdecode(o, e, q) >= hi (q = 255 always qualifies by the choice of e).
__device__ unsigned q_ceil(float o, int e, float hi) {
    const float d = (hi - o) / __int_as_float((e + 127) << 23);
    int q = d > 0 ? (int)fminf(ceilf(d), 255.0f) : 0;
    while (q < 255 && qdecode(o, e, q) < hi) q++;
    while (q > 0 && qdecode(o, e, q - 1) >= hi) q--;
    return (unsigned)q;
}

// The quantized 4-wide node of `cnt` children with exact boxes mn / mx: the origin is the minimum
// corner of their union, per axis the smallest exponent that reaches the maximum corner, then each
// child's planes rounded outwards (q_floor / q_ceil).
__device__ BvhQNode quantize_node(const int (&ch)[4], const float (&mn)[4][3], const float (&mx)[4][3], int cnt) {
    BvhQNode o;
    float org[3];
    int ex[3];
    for (int a = 0; a < 3; a++) {
        float lo = mn[0][a], hi = mx[0][a];
        for (int k = 1; k < cnt; k++) { lo = fminf(lo, mn[k][a]); hi = fmaxf(hi, mx[k][a]); }
        org[a] = lo;
        ex[a] = q_exponent(lo, hi);
    }
    o.ox = org[0]; o.oy = org[1]; o.oz = org[2];
    o.ex = (int8_t)ex[0]; o.ey = (int8_t)ex[1]; o.ez = (int8_t)ex[2];
    o.count = (uint8_t)cnt;
    for (int k = 0; k < 4; k++) {
        const bool valid = k < cnt;
        o.child[k] = valid ? ch[k] : 0;
        for (int a = 0; a < 3; a++) {
            o.qlo[a][k] = valid ? (uint8_t)q_floor(org[a], ex[a], mn[k][a]) : 0;
            o.qhi[a][k] = valid ? (uint8_t)q_ceil(org[a], ex[a], mx[k][a]) : 0;
        }
    }
    o.pad[0] = o.pad[1] = 0;
    return o;
}

__global__ __launch_bounds__(kBlock) void k_build_bvh4(const BvhNode* __restrict__ nodes, const int32_t* __restrict__ flag,
                                                       const int32_t* __restrict__ index4, int n, BvhQNode* __restrict__ out) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n - 1 || !flag[i]) return;
    const BvhNode nd = nodes[i];
    int ch[4];
    float mn[4][3], mx[4][3];
    int cnt = 0;
    auto add = [&](int c, const float* lo, const float* hi) {
        ch[cnt] = c >= n - 1 ? ~(c - (n - 1)) : index4[c];
        for (int a = 0; a < 3; a++) { mn[cnt][a] = lo[a]; mx[cnt][a] = hi[a]; }
        cnt++;
    };
    for (int side = 0; side < 2; side++) {
        const int c = side ? nd.right : nd.left;
        const float* lo = side ? nd.rmin : nd.lmin;
        const float* hi = side ? nd.rmax : nd.lmax;
        if (c >= n - 1) {
            add(c, lo, hi);
        } else {
            const BvhNode g = nodes[c];  // odd depth: replaced by its two children
            add(g.left, g.lmin, g.lmax);
            add(g.right, g.rmin, g.rmax);
        }
    }
    out[index4[i]] = quantize_node(ch, mn, mx, cnt);
}

// Greedy surface-area collapse, one level of the 4-wide tree per launch (breadth first). Frontier
// entry t is binary interior node front[t]; it becomes 4-wide node levelStart + t, whose children
// start as the binary node's two children, then the interior child of largest surface area is
// replaced by its two children until there are four or only leaves are left. Leaf children are
// stored as ~leaf; the interior ones are listed in tmp[4 t + slot] (-1 elsewhere) and counted in
// cnt[t], and k_collapse_emit numbers them once cnt holds their exclusive prefix.
__global__ __launch_bounds__(kBlock) void k_collapse_count(const int32_t* __restrict__ front, int count,
                                                           const BvhNode* __restrict__ nodes, int n, int levelStart,
                                                           int32_t* __restrict__ tmp, int32_t* __restrict__ cnt,
                                                           BvhQNode* __restrict__ out) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= count) return;
    const BvhNode nd = nodes[front[t]];
    int id[4];
    float mn[4][3], mx[4][3];
    auto put = [&](int k, int c, const float* lo, const float* hi) {
        id[k] = c;
        for (int a = 0; a < 3; a++) { mn[k][a] = lo[a]; mx[k][a] = hi[a]; }
    };
    put(0, nd.left, nd.lmin, nd.lmax);
    put(1, nd.right, nd.rmin, nd.rmax);
    int c = 2;
    while (c < 4) {
        int best = -1;
        float bestArea = -1.0f;
        for (int k = 0; k < c; k++) {
            if (id[k] >= n - 1) continue;
            const float dx = mx[k][0] - mn[k][0], dy = mx[k][1] - mn[k][1], dz = mx[k][2] - mn[k][2];
            const float area = dx * dy + dy * dz + dz * dx;
            if (area > bestArea) { bestArea = area; best = k; }
        }
        if (best < 0) break;
        const BvhNode g = nodes[id[best]];
        put(best, g.left, g.lmin, g.lmax);
        put(c, g.right, g.rmin, g.rmax);
        c++;
    }
    int ch[4];
    int interior = 0;
    for (int k = 0; k < 4; k++) {
        const bool leaf = k < c && id[k] >= n - 1;
        ch[k] = leaf ? ~(id[k] - (n - 1)) : 0;
        const bool inner = k < c && !leaf;
        tmp[4 * t + k] = inner ? id[k] : -1;
        interior += inner;
    }
    out[levelStart + t] = quantize_node(ch, mn, mx, c);
    cnt[t] = interior;
}

// Numbers the interior children of the level's nodes (cnt now holds exclusive offsets): the j-th
// interior child of entry t becomes node nextStart + cnt[t] + j and frontier entry cnt[t] + j.
__global__ __launch_bounds__(kBlock) void k_collapse_emit(int count, const int32_t* __restrict__ tmp,
                                                          const int32_t* __restrict__ cnt, int levelStart, int nextStart,
                                                          BvhQNode* __restrict__ out, int32_t* __restrict__ next) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= count) return;
    int j = cnt[t];
    for (int k = 0; k < 4; k++) {
        const int c = tmp[4 * t + k];
        if (c < 0) continue;
        out[levelStart + t].child[k] = nextStart + j;
        next[j] = c;
        j++;
    }
}


// ---------------------------------------------------------------- PLOC render tree
// Parallel locally-ordered clustering (Meister & Bittner, IEEE TVCG 2018) over the Morton-ordered
// leaves of the LBVH build: every cluster finds the cluster within kPlocRadius positions whose union
// box has the smallest surface area; mutual nearest neighbours merge into a new interior node; the
// survivors are compacted in order; repeat until one cluster is left. Pairs are ordered by (area,
// lower position, higher position), a total order, so the globally smallest pair is always mutual
// and every round merges at least one pair. New nodes are numbered by a prefix sum over the merging
// pairs, counting down from n - 2, so the root is node 0 and the layout matches the Karras tree's
// (interior 0..n-2, leaf k at n-1+k over the same sorted leaves): the 4-wide collapse and the
// traversal take either tree. The result depends only on the input, not on scheduling.
constexpr int kPlocRadius = 8;
constexpr int kPlocBlock = 256;
constexpr int kFlagBlock = 1024;

__device__ __forceinline__ float union_area(const float* a, const float* b) {
    const float dx = fmaxf(a[3], b[3]) - fminf(a[0], b[0]);
    const float dy = fmaxf(a[4], b[4]) - fminf(a[1], b[1]);
    const float dz = fmaxf(a[5], b[5]) - fminf(a[2], b[2]);
    const float s = dx * dy + dy * dz + dz * dx;
    return s <= 3.0e38f ? s : INFINITY;  // NaN and overflow order as +inf, keeping the order total
}

__global__ __launch_bounds__(kBlock) void k_ploc_init(const float* __restrict__ primBox, const int32_t* __restrict__ perm, int n,
                                                      float* __restrict__ box, int32_t* __restrict__ cl, int32_t* __restrict__ parent) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k < n) {
        const int p = perm[k];
#pragma unroll
        for (int a = 0; a < 6; a++) box[6 * (size_t)(n - 1 + k) + a] = primBox[6 * (size_t)p + a];
        cl[k] = n - 1 + k;
        parent[n - 1 + k] = -1;
    }
    if (k < n - 1) parent[k] = -1;
}

__global__ __launch_bounds__(kPlocBlock) void k_ploc_nearest(const int32_t* __restrict__ cl, int m, const float* __restrict__ box,
                                                             int32_t* __restrict__ nn) {
    __shared__ float sb[(kPlocBlock + 2 * kPlocRadius) * 6];
    const int base = blockIdx.x * kPlocBlock - kPlocRadius;
    for (int e = threadIdx.x; e < kPlocBlock + 2 * kPlocRadius; e += kPlocBlock) {
        const int pos = base + e;
        if (pos >= 0 && pos < m) {
            const int c = cl[pos];
#pragma unroll
            for (int a = 0; a < 6; a++) sb[6 * e + a] = box[6 * (size_t)c + a];
        }
    }
    __syncthreads();
    const int i = blockIdx.x * kPlocBlock + threadIdx.x;
    if (i >= m) return;
    const float* bi = &sb[6 * (threadIdx.x + kPlocRadius)];
    float best = INFINITY;
    int bj = -1;
    const int lo = max(0, i - kPlocRadius), hi = min(m - 1, i + kPlocRadius);
    for (int j = lo; j <= hi; j++) {
        if (j == i) continue;
        const float a = union_area(bi, &sb[6 * (j - base)]);
        // candidates are visited in increasing j: below i the pair key (j, i) grows with j, above i
        // the key (i, j) grows with j and every key above i is larger than every key below it, so a
        // tie keeps the earlier candidate
        if (bj < 0 || a < best) { best = a; bj = j; }
    }
    nn[i] = bj;
}

// keep[i] / merge[i]: exclusive prefix within blocks of kFlagBlock; sums[2b], sums[2b+1]: block totals.
__global__ __launch_bounds__(kFlagBlock) void k_ploc_flags(const int32_t* __restrict__ nn, int m, int32_t* __restrict__ keep,
                                                           int32_t* __restrict__ merge, int32_t* __restrict__ sums) {
    __shared__ int32_t shK[kFlagBlock / 64], shM[kFlagBlock / 64];
    const int i = blockIdx.x * kFlagBlock + threadIdx.x;
    int k0 = 0, m0 = 0;
    if (i < m) {
        const int j = nn[i];
        const bool mutual = nn[j] == i;
        k0 = !mutual || i < j;
        m0 = mutual && i < j;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int xk = k0, xm = m0;
    for (int o = 1; o < 64; o <<= 1) {
        const int yk = __shfl_up(xk, o), ym = __shfl_up(xm, o);
        if (lane >= o) { xk += yk; xm += ym; }
    }
    if (lane == 63) { shK[wave] = xk; shM[wave] = xm; }
    __syncthreads();
    int bk = 0, bm = 0;
    for (int w = 0; w < wave; w++) { bk += shK[w]; bm += shM[w]; }
    if (i < m) { keep[i] = bk + xk - k0; merge[i] = bm + xm - m0; }
    if (threadIdx.x == kFlagBlock - 1) { sums[2 * blockIdx.x] = bk + xk; sums[2 * blockIdx.x + 1] = bm + xm; }
}

// Exclusive prefix of the block totals in one block; totals[0] = survivors, totals[1] = merges.
__global__ __launch_bounds__(kFlagBlock) void k_ploc_scan_sums(int32_t* __restrict__ sums, int nb, int32_t* __restrict__ totals) {
    __shared__ int32_t shK[kFlagBlock / 64], shM[kFlagBlock / 64];
    __shared__ int32_t carry[2];
    if (threadIdx.x == 0) { carry[0] = carry[1] = 0; }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int c0 = 0; c0 < nb; c0 += kFlagBlock) {
        const int b = c0 + threadIdx.x;
        const int k0 = b < nb ? sums[2 * b] : 0, m0 = b < nb ? sums[2 * b + 1] : 0;
        int xk = k0, xm = m0;
        for (int o = 1; o < 64; o <<= 1) {
            const int yk = __shfl_up(xk, o), ym = __shfl_up(xm, o);
            if (lane >= o) { xk += yk; xm += ym; }
        }
        if (lane == 63) { shK[wave] = xk; shM[wave] = xm; }
        __syncthreads();
        int bk = carry[0], bm = carry[1];
        for (int w = 0; w < wave; w++) { bk += shK[w]; bm += shM[w]; }
        if (b < nb) { sums[2 * b] = bk + xk - k0; sums[2 * b + 1] = bm + xm - m0; }
        __syncthreads();
        if (threadIdx.x == kFlagBlock - 1) { carry[0] = bk + xk; carry[1] = bm + xm; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { totals[0] = carry[0]; totals[1] = carry[1]; }
}

__global__ __launch_bounds__(kFlagBlock) void k_ploc_emit(const int32_t* __restrict__ cl, const int32_t* __restrict__ nn, int m,
                                                          const int32_t* __restrict__ keep, const int32_t* __restrict__ merge,
                                                          const int32_t* __restrict__ sums, int created, int n,
                                                          float* __restrict__ box, int32_t* __restrict__ left, int32_t* __restrict__ right,
                                                          int32_t* __restrict__ parent, int32_t* __restrict__ out) {
    const int i = blockIdx.x * kFlagBlock + threadIdx.x;
    if (i >= m) return;
    const int j = nn[i];
    const bool mutual = nn[j] == i;
    const int pos = keep[i] + sums[2 * blockIdx.x];
    if (!mutual) { out[pos] = cl[i]; return; }
    if (i > j) return;
    const int node = n - 2 - (created + merge[i] + sums[2 * blockIdx.x + 1]);
    const int a = cl[i], c = cl[j];
    left[node] = a; right[node] = c;
    parent[a] = node; parent[c] = node;
#pragma unroll
    for (int q = 0; q < 3; q++) {
        box[6 * (size_t)node + q] = fminf(box[6 * (size_t)a + q], box[6 * (size_t)c + q]);
        box[6 * (size_t)node + 3 + q] = fmaxf(box[6 * (size_t)a + 3 + q], box[6 * (size_t)c + 3 + q]);
    }
    out[pos] = node;
}
}  // namespace

template <typename T>
static hipError_t dalloc(T** p, size_t count) {
    return hipMalloc((void**)p, (count ? count : 1) * sizeof(T));
}

hipError_t bvh_reserve(BvhBuffers& b, int n) {
    if (n <= b.cap) return hipSuccess;
    bvh_release(b);
    int tiles = (n + kTile - 1) / kTile;
    hipError_t e = hipSuccess;
#define TRY(x) do { e = (x); if (e != hipSuccess) { bvh_release(b); return e; } } while (0)
    TRY(dalloc(&b.primBox, 6 * (size_t)n));
    TRY(dalloc(&b.box, 6 * (size_t)(2 * n - 1)));
    TRY(dalloc(&b.keysA, n)); TRY(dalloc(&b.keysB, n));
    TRY(dalloc(&b.valsA, n)); TRY(dalloc(&b.valsB, n));
    TRY(dalloc(&b.counts, 256 * (size_t)tiles));
    TRY(dalloc(&b.digitBase, 256));
    TRY(dalloc(&b.left, n)); TRY(dalloc(&b.right, n));
    TRY(dalloc(&b.parent, 2 * (size_t)n));
    TRY(dalloc(&b.flags, n));
    TRY(dalloc(&b.rangeLo, n)); TRY(dalloc(&b.rangeHi, n));
    TRY(dalloc(&b.roots, n));
    TRY(dalloc(&b.rootCount, (n + kChunk2 - 1) / kChunk2 + 1));
    TRY(dalloc(&b.bucket, (size_t)((n + kChunk2 - 1) / kChunk2) * kBucketCap));
    TRY(dalloc(&b.partial, 6 * (size_t)blocks_for(n)));
    TRY(dalloc(&b.sceneKeys, 8));
    TRY(dalloc(&b.status, 1));
    TRY(dalloc(&b.nodes, n));
    TRY(dalloc(&b.tris, n));
#undef TRY
    b.cap = n;
    return hipSuccess;
}

// Greedy surface-area collapse (k_collapse_count / k_collapse_emit), level by level from the root.
// Returns the number of 4-wide nodes in *nodes4, or 0 when the tree has more than kTraceStack / 3
// levels (a 4-wide node leaves at most 3 entries on the traversal stack).
static hipError_t collapse_greedy(BvhBuffers& b, hipStream_t s, int* nodes4) {
    const int n = b.n, m = n - 1;
    *nodes4 = 0;
    int32_t *fa = nullptr, *fb = nullptr, *cnt = nullptr, *tmp = nullptr;
    hipError_t e = dalloc(&fa, m);
    if (e == hipSuccess) e = dalloc(&fb, m);
    if (e == hipSuccess) e = dalloc(&cnt, m);
    if (e == hipSuccess) e = dalloc(&tmp, 4 * (size_t)m);
    const int32_t root = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(fa, &root, sizeof(root), hipMemcpyHostToDevice, s);
    int levelStart = 0, count = 1, levels = 0;
    while (e == hipSuccess && count > 0) {
        if (3 * ++levels > kTraceStack) { levelStart = 0; break; }
        const int lb = (count + kScanBlock - 1) / kScanBlock;
        k_collapse_count<<<blocks_for(count), kBlock, 0, s>>>(fa, count, b.nodes, n, levelStart, tmp, cnt, b.nodes4);
        k_scan_local<<<lb, kScanBlock, 0, s>>>(cnt, count, b.scanTotals);
        k_scan_totals<<<1, 64, 0, s>>>(b.scanTotals, lb);
        k_scan_add<<<lb, kScanBlock, 0, s>>>(cnt, count, b.scanTotals);
        k_collapse_emit<<<blocks_for(count), kBlock, 0, s>>>(count, tmp, cnt, levelStart, levelStart + count, b.nodes4, fb);
        int32_t next = 0;
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(&next, b.scanTotals + lb, sizeof(next), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        levelStart += count;
        count = next;
        std::swap(fa, fb);
    }
    void* ps[] = {fa, fb, cnt, tmp};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    if (e == hipSuccess) *nodes4 = levelStart;
    return e;
}

hipError_t bvh4_enqueue(BvhBuffers& b, hipStream_t s, int* depth, int* nodes4, bool* even) {
    const int n = b.n;
    *depth = 0;
    *nodes4 = 0;
    if (n < 2) return hipSuccess;
    const int m = n - 1;
    const int nb = (m + kScanBlock - 1) / kScanBlock;
    if (b.cap4 < n) {
        void* ps[] = {b.nodes4, b.index4, b.scanTotals};
        for (void* p : ps)
            if (p) (void)hipFree(p);
        b.nodes4 = nullptr; b.index4 = nullptr; b.scanTotals = nullptr; b.cap4 = 0;
        hipError_t e = dalloc(&b.nodes4, m);
        if (e == hipSuccess) e = dalloc(&b.index4, m);
        if (e == hipSuccess) e = dalloc(&b.scanTotals, (size_t)nb + 2);
        if (e != hipSuccess) return e;
        b.cap4 = n;
    }
    // the binary tree's depth (and its even-depth flags, for the fallback collapse)
    int32_t* depthDev = b.scanTotals + nb + 1;
    hipError_t e = hipMemsetAsync(depthDev, 0, sizeof(int32_t), s);
    if (e != hipSuccess) return e;
    k_even_depth<<<blocks_for(m), kBlock, 0, s>>>(b.parent, n, b.flags, depthDev);
    e = hipMemcpyAsync(depth, depthDev, sizeof(int32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    const int binaryDepth = *depth;
    if (!*even) {
        if ((e = collapse_greedy(b, s, nodes4)) != hipSuccess) return e;
        if (*nodes4 > 0) return hipSuccess;
        *even = true;
    }
    // fallback: every interior node of even depth becomes a 4-wide node (ceil(depth / 2) levels)
    // flags -> exclusive prefix in index4 -> 4-wide nodes
    e = hipMemcpyAsync(b.index4, b.flags, (size_t)m * sizeof(int32_t), hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return e;
    k_scan_local<<<nb, kScanBlock, 0, s>>>(b.index4, m, b.scanTotals);
    k_scan_totals<<<1, 64, 0, s>>>(b.scanTotals, nb);
    k_scan_add<<<nb, kScanBlock, 0, s>>>(b.index4, m, b.scanTotals);
    k_build_bvh4<<<blocks_for(m), kBlock, 0, s>>>(b.nodes, b.flags, b.index4, n, b.nodes4);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    int32_t total = 0;  // scanTotals[nb] = 4-wide node count
    e = hipMemcpyAsync(&total, b.scanTotals + nb, sizeof(int32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    *nodes4 = total;
    *depth = binaryDepth;
    return hipSuccess;
}


hipError_t ploc_enqueue(BvhBuffers& b, hipStream_t s, int maxDepth, int* depthOut) {
    const int n = b.n;
    *depthOut = 0;
    if (n < 2) return hipSuccess;
    // temporaries: unified boxes (interior + leaves), children, parents, two cluster lists,
    // neighbours, the two prefix arrays and the block totals
    const int nbMax = (n + kFlagBlock - 1) / kFlagBlock;
    float* box = nullptr;
    int32_t *left = nullptr, *right = nullptr, *parent = nullptr, *clA = nullptr, *clB = nullptr, *nn = nullptr, *keep = nullptr,
            *merge = nullptr, *sums = nullptr, *totals = nullptr;
    int32_t* host = nullptr;
    hipError_t e = hipSuccess;
    auto cleanup = [&]() {
        void* ps[] = {box, left, right, parent, clA, clB, nn, keep, merge, sums, totals};
        for (void* p : ps)
            if (p) (void)hipFree(p);
        if (host) (void)hipHostFree(host);
    };
#define TRY(x) do { e = (x); if (e != hipSuccess) { cleanup(); return e; } } while (0)
    TRY(dalloc(&box, 6 * (size_t)(2 * n - 1)));
    TRY(dalloc(&left, n)); TRY(dalloc(&right, n)); TRY(dalloc(&parent, 2 * (size_t)n));
    TRY(dalloc(&clA, n)); TRY(dalloc(&clB, n)); TRY(dalloc(&nn, n));
    TRY(dalloc(&keep, n)); TRY(dalloc(&merge, n));
    TRY(dalloc(&sums, 2 * (size_t)nbMax + 2)); TRY(dalloc(&totals, 4));
    TRY(hipHostMalloc((void**)&host, 4 * sizeof(int32_t), hipHostMallocDefault));
    k_ploc_init<<<blocks_for(n), kBlock, 0, s>>>(b.primBox, b.valsA, n, box, clA, parent);
    int m = n, created = 0;
    for (int round = 0; m > 1; round++) {
        if (round > n) { cleanup(); return hipErrorUnknown; }  // unreachable: every round merges
        const int nb = (m + kFlagBlock - 1) / kFlagBlock;
        k_ploc_nearest<<<(m + kPlocBlock - 1) / kPlocBlock, kPlocBlock, 0, s>>>(clA, m, box, nn);
        k_ploc_flags<<<nb, kFlagBlock, 0, s>>>(nn, m, keep, merge, sums);
        k_ploc_scan_sums<<<1, kFlagBlock, 0, s>>>(sums, nb, totals);
        k_ploc_emit<<<nb, kFlagBlock, 0, s>>>(clA, nn, m, keep, merge, sums, created, n, box, left, right, parent, clB);
        TRY(hipGetLastError());
        TRY(hipMemcpyAsync(host, totals, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        TRY(hipStreamSynchronize(s));
        if (host[1] < 1 || host[0] != m - host[1]) { cleanup(); return hipErrorUnknown; }
        m = host[0];
        created += host[1];
        std::swap(clA, clB);
    }
    if (created != n - 1) { cleanup(); return hipErrorUnknown; }
    // depth in edges down to the deepest leaf
    TRY(hipMemsetAsync(totals + 2, 0, sizeof(int32_t), s));
    k_even_depth<<<blocks_for(n - 1), kBlock, 0, s>>>(parent, n, b.flags, totals + 2);
    TRY(hipGetLastError());
    TRY(hipMemcpyAsync(host + 2, totals + 2, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    TRY(hipStreamSynchronize(s));
    *depthOut = host[2];
    if (host[2] <= maxDepth) {
        // adopt the tree: the traversal layout and the parents the 4-wide collapse climbs
        k_pack_nodes<<<blocks_for(n - 1), kBlock, 0, s>>>(box, left, right, n, b.nodes);
        TRY(hipMemcpyAsync(b.parent, parent, (size_t)(2 * n - 1) * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
        TRY(hipGetLastError());
        TRY(hipStreamSynchronize(s));
    }
#undef TRY
    cleanup();
    return hipSuccess;
}

void bvh_release(BvhBuffers& b) {
    void* ps[] = {b.primBox, b.box, b.keysA, b.keysB, b.valsA, b.valsB, b.counts, b.digitBase, b.left, b.right,
                  b.parent, b.flags, b.sceneKeys, b.nodes, b.tris, b.rangeLo, b.rangeHi, b.roots, b.rootCount, b.bucket, b.partial,
                  b.nodes4, b.index4, b.scanTotals, b.status};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    b = BvhBuffers();
}

hipError_t bvh_enqueue(BvhBuffers& b, const float* soup, int n, int mode, bool pack, hipStream_t s, BvhStageEvents* ev) {
    if (n < 1 || n > b.cap) return hipErrorInvalidValue;
    b.n = n;
    b.tiles = (n + kTile - 1) / kTile;
    if (ev) (void)hipEventRecord(ev->e[0], s);
    k_prim_bounds<<<blocks_for(n), kBlock, 0, s>>>(soup, n, b.primBox, b.partial, b.status);
    k_reduce_scene_keys<<<1, 1024, 0, s>>>(b.partial, blocks_for(n), b.sceneKeys);
    k_morton<<<blocks_for(n), kBlock, 0, s>>>(b.primBox, n, b.sceneKeys, b.keysA, b.valsA);
    if (ev) (void)hipEventRecord(ev->e[1], s);
    uint32_t* kin = b.keysA; uint32_t* kout = b.keysB;
    int32_t* vin = b.valsA; int32_t* vout = b.valsB;
    for (int pass = 0; pass < 4; pass++) {
        int shift = 8 * pass;
        k_radix_count<<<b.tiles, kBlock, 0, s>>>(kin, n, shift, b.tiles, b.counts);
        k_radix_scan_tiles<<<256, kBlock, 0, s>>>(b.counts, b.tiles, b.digitBase);
        k_radix_scan_digits<<<1, kBlock, 0, s>>>(b.digitBase);
        k_radix_scatter<<<b.tiles, kBlock, 0, s>>>(kin, vin, kout, vout, n, shift, b.tiles, b.counts, b.digitBase);
        uint32_t* tk = kin; kin = kout; kout = tk;
        int32_t* tv = vin; vin = vout; vout = tv;
    }
    // after 4 passes the sorted arrays are back in keysA / valsA
    if (ev) (void)hipEventRecord(ev->e[2], s);
    if (n > 1) {
        k_fill_i32<<<blocks_for(n - 1), kBlock, 0, s>>>(b.flags, n - 1, 0);
        k_fill_i32<<<blocks_for(2 * n - 1), kBlock, 0, s>>>(b.parent, 2 * n - 1, -1);
        k_karras<<<blocks_for(n - 1), kBlock, 0, s>>>(b.keysA, n, b.left, b.right, b.parent, b.rangeLo, b.rangeHi);
    } else {
        k_fill_i32<<<1, kBlock, 0, s>>>(b.parent, 1, -1);
    }
    const int chunks = (n + kChunk - 1) / kChunk, chunks2 = (n + kChunk2 - 1) / kChunk2;
    k_fill_i32<<<blocks_for(chunks2 + 1), kBlock, 0, s>>>(b.rootCount, chunks2 + 1, 0);  // level-3 count + buckets
    k_refit_l1<<<chunks, kChunk, 0, s>>>(b.primBox, b.valsA, n, mode, b.left, b.right, b.parent, b.rangeLo, b.rangeHi, b.box,
                                         b.bucket, b.rootCount + 1, b.status);
    if (chunks > 1)
        k_refit_l2<<<chunks2, 1024, 0, s>>>(n, b.left, b.right, b.parent, b.rangeLo, b.rangeHi, b.box, b.bucket,
                                            b.rootCount + 1, b.roots, b.rootCount, b.status);
    if (chunks2 > 1) {
        // level-3 roots: at most two per level-2-crossing node, and each of the chunks2 - 1
        // boundaries is crossed by at most 58 nodes
        const long bound = std::min<long>((long)n, 2L * chunks2 * 58);
        k_refit_global<<<blocks_for(bound), kBlock, 0, s>>>(n, b.left, b.right, b.parent, b.box, b.flags, b.roots, b.rootCount,
                                                            b.status);
    }
    if (ev) (void)hipEventRecord(ev->e[3], s);
    if (pack) {
        if (n > 1) k_pack_nodes<<<blocks_for(n - 1), kBlock, 0, s>>>(b.box, b.left, b.right, n, b.nodes);
        k_pack_tris<<<blocks_for(n), kBlock, 0, s>>>(soup, b.valsA, n, b.tris);
    }
    if (ev) (void)hipEventRecord(ev->e[4], s);
    return hipGetLastError();
}

// ---------------------------------------------------------------- build without allocation
// Every temporary array of a build of at most kStaticMaxPrims primitives, carved from one static
// device buffer (each device holds its own copy from the moment the library's code object is loaded
// until the process ends). Host code never passes `scratch` itself to a kernel: its device address
// comes from hipGetSymbolAddress, looked up once per device.
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {

std::mutex g_scratchLock;
void* g_scratchPtr[kStaticMaxDevices] = {};

size_t carve(size_t& at, size_t bytes) {
    const size_t o = at;
    at += (bytes + 255) & ~size_t(255);
    return o;
}

// The scratch layout of an n-primitive build; `base` null measures it (the returned size).
size_t static_layout(BvhBuffers& b, unsigned char* base, int n) {
    const int tiles = (n + kTile - 1) / kTile, chunks2 = (n + kChunk2 - 1) / kChunk2;
    size_t at = 0;
    auto take = [&](auto*& ptr, size_t count) {
        const size_t o = carve(at, count * sizeof(*ptr));
        ptr = base ? reinterpret_cast<std::remove_reference_t<decltype(ptr)>>(base + o) : nullptr;
    };
    take(b.primBox, 6 * (size_t)n);
    take(b.keysB, n);
    take(b.valsB, n);
    take(b.counts, 256 * (size_t)tiles);
    take(b.digitBase, 256);
    take(b.flags, n);
    take(b.rangeLo, n);
    take(b.rangeHi, n);
    take(b.roots, n);
    take(b.rootCount, chunks2 + 1);
    take(b.bucket, (size_t)chunks2 * kBucketCap);
    take(b.partial, 6 * (size_t)blocks_for(n));
    take(b.sceneKeys, 8);
    take(b.status, 1);
    return at;
}

}  // namespace

// The largest build (kStaticMaxPrims) must fit the buffer.
static_assert(kStaticScratchBytes >= 256 * 20 + 6 * 4 * (size_t)kStaticMaxPrims + 8 * 4 * (size_t)kStaticMaxPrims +
                                         256 * 4 * ((kStaticMaxPrims + kTile - 1) / kTile) + 4 * 256 +
                                         4 * ((kStaticMaxPrims + kChunk2 - 1) / kChunk2 + 1) +
                                         4 * (size_t)((kStaticMaxPrims + kChunk2 - 1) / kChunk2) * kBucketCap +
                                         4 * 6 * ((kStaticMaxPrims + kBlock - 1) / kBlock) + 4 * 8 + 4,
              "static LBVH scratch too small for kStaticMaxPrims");

hipError_t bvh_static_scratch(int device, void** ptr) {
    if (device < 0 || device >= kStaticMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratchLock);
    if (!g_scratchPtr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratchPtr[device] = p;
    }
    *ptr = g_scratchPtr[device];
    return hipSuccess;
}

hipError_t bvh_enqueue_static(void* scratchPtr, const float* soup, int n, int mode, uint32_t* codesSorted, int32_t* perm,
                              int32_t* left, int32_t* right, int32_t* parent, float* bounds, hipStream_t s,
                              int32_t** status, uint32_t** sceneKeys) {
    if (n < 1 || n > kStaticMaxPrims) return hipErrorInvalidValue;
    BvhBuffers b;
    if (static_layout(b, nullptr, n) > kStaticScratchBytes) return hipErrorInvalidValue;
    static_layout(b, static_cast<unsigned char*>(scratchPtr), n);
    b.cap = n;
    b.keysA = codesSorted;
    b.valsA = perm;
    b.left = left;
    b.right = right;
    b.parent = parent;
    b.box = bounds;
    *status = b.status;
    *sceneKeys = b.sceneKeys;
    return bvh_enqueue(b, soup, n, mode, false, s, nullptr);
}

}  // namespace lts
