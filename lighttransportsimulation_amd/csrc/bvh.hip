// Karras-2012 LBVH build on gfx950, bit-exact with the reference's construction
// (acceleration/BVH.cuh:77-452, acceleration/BVH_kernels.cu:8-50):
//   k_prim_bounds  primitive AABB + centroid inputs, block-reduced scene bounds     (BVH.cuh:92-98)
//   k_morton       30-bit Morton codes of the normalised centroids, digit counts     (BVH.cuh:101-137)
//   k_radix_*      stable LSD sort of (code, primitive) pairs, 10-bit digits, 3 passes of tile digit
//   k_scan_*       counts, a two-level scan and an LDS-ranked scatter                 (BVH.cuh:140-146)
//   k_tree_l1/up   the Karras hierarchy (BVH.cuh:173-236) and its bounds (BVH.cuh:238-262) in one
//                  bottom-up climb: 512-leaf chunks in LDS, then the chunk roots through global
//                  arrival slots and records
//   k_pack_*       traversal layout: both child boxes per node, triangles in leaf order
// The scene-bound reduction uses a total order on floats (-0 below +0); it equals the reference's
// float atomics (GeneralHelper.cuh:150-160) on every input without -0.0.
#include <algorithm>
#include <atomic>
#include <mutex>
#include <type_traits>

#include "lts_device.h"
#include "lts_internal.h"

namespace lts {

namespace {

constexpr int kBlock = 256;

// Build status bits (BvhBuffers::status, 0 = a complete build). Neither can occur on a correct
// build: a climb is bounded (the tree is at most 56 interior nodes deep for up to 2^26 leaves) and
// a scatter writes inside the output; a violation is reported instead of hanging or writing out
// of bounds.
constexpr int kStatusClimb = 8;   // a climb took more steps than its bound
constexpr int kStatusSort = 16;   // a radix scatter found inconsistent digit counts
constexpr int kStatusTree = 32;   // packing found a child index outside the tree

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
    const int i0 = blockIdx.x * kBlock + threadIdx.x;
    if (i0 == 0) *status = 0;  // the build's first kernel: later kernels on the stream see the cleared flags
    uint32_t k[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
    for (int i = i0; i < n; i += gridDim.x * kBlock) {
        float t[9];
#pragma unroll
        for (int j = 0; j < 9; j++) t[j] = soup[9 * (size_t)i + j];
        float mn[3], mx[3];
        tri_box(t, mn, mx);
#pragma unroll
        for (int a = 0; a < 3; a++) {
            primBox[6 * (size_t)i + a] = mn[a];
            primBox[6 * (size_t)i + 3 + a] = mx[a];
            k[a] = min(k[a], order_key(mn[a]));
            k[3 + a] = max(k[3 + a], order_key(mx[a]));
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

// delta (BVH.cuh:173-190): common prefix length of sorted codes i and j; equal codes compare their
// positions (offset 40 keeps them above every distinct-code value); -1 outside [0, n)
__device__ __forceinline__ int delta_codes(uint32_t ci, uint32_t cj, int n, int i, int j) {
    if (i < 0 || i >= n || j < 0 || j >= n) return -1;
    uint32_t x = ci ^ cj;
    int extra = 0;
    if (x == 0) { extra = 40; x = (uint32_t)(i ^ j); }
    return (x == 0 ? 32 : __clz((int)x)) + extra;
}
__device__ __forceinline__ int delta_at(const uint32_t* __restrict__ codes, int n, int i, int j) {
    if (i < 0 || i >= n || j < 0 || j >= n) return -1;
    return delta_codes(codes[i], codes[j], n, i, j);
}

// ---------------------------------------------------------------- radix sort
// Stable LSD sort of the 30-bit codes in 3 passes of 10-bit digits over 4096-key tiles. A pass is
// count (a row of 1024 digit counts per tile; the Morton kernel counts the first pass's digits) ->
// scan (within segments of 32 tiles, then over the segments and the digits) -> scatter. Counts are
// stored and read as whole rows, so every access is a contiguous 4 KB. The scatter ranks a tile's
// keys with per-wave digit counters: wave w owns the tile's keys w*1024 .. w*1024+1023 (16 rows of
// 64) and ranks them in row order, so the sort is stable without a barrier per row.
constexpr int kSortBits = 10;
constexpr int kBins = 1 << kSortBits;
constexpr int kSortPasses = 3;  // 30-bit codes
constexpr int kSortBlock = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortBlock * kSortItems;  // keys per tile
constexpr int kSortWaves = kSortBlock / 64;
constexpr int kDigitsPerThread = kBins / kSortBlock;
static_assert(kDigitsPerThread == 4, "a thread owns one uint4 of a count row");
constexpr int kSegTiles = 32;  // tiles per scan segment
constexpr int kScanWindow = 32;  // segment totals a thread loads at once

inline int sort_tiles(int n) { return (n + kSortTile - 1) / kSortTile; }
inline int sort_segments(int n) { return (sort_tiles(n) + kSegTiles - 1) / kSegTiles; }

// Tile of block b: the 8 XCDs (blocks are dealt to them round-robin) each take a contiguous range
// of tiles, so the runs that neighbouring tiles write for a digit meet in one XCD's L2.
__device__ __forceinline__ int xcd_tile(int b, int tiles) {
    constexpr int kXcds = 8;
    const int per = tiles / kXcds, rem = tiles % kXcds, x = b % kXcds;
    return x * per + min(x, rem) + b / kXcds;
}

// Exclusive scan of one value per thread over a block of kThreads threads; returns the total.
template <int kThreads>
__device__ uint32_t block_scan(uint32_t v, uint32_t* out, uint32_t* scratch) {
    constexpr int kWaves = kThreads / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = (uint32_t)__shfl_up((int)incl, off);
        if (lane >= off) incl += y;
    }
    if (lane == 63) scratch[wave] = incl;
    __syncthreads();
    uint32_t wbase = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kWaves; w++) {
        const uint32_t s = scratch[w];
        if (w < wave) wbase += s;
        total += s;
    }
    *out = wbase + incl - v;
    __syncthreads();
    return total;
}

// The tile's digit counts from the LDS histogram to its row counts[tile][*].
__device__ __forceinline__ void store_tile_counts(const uint32_t* hist, uint32_t* __restrict__ counts) {
    const int d = kDigitsPerThread * threadIdx.x;
    reinterpret_cast<uint4*>(counts + (size_t)blockIdx.x * kBins)[threadIdx.x] =
        make_uint4(hist[d], hist[d + 1], hist[d + 2], hist[d + 3]);
}

// Morton codes of the centroids (one sort tile per block), with the tile's first-pass digit counts.
__global__ __launch_bounds__(kSortBlock) void k_morton(const float* __restrict__ primBox, int n, const uint32_t* __restrict__ sceneKeys,
                                                       uint32_t* __restrict__ codes, uint32_t* __restrict__ counts) {
    __shared__ uint32_t hist[kBins];
#pragma unroll
    for (int u = 0; u < kDigitsPerThread; u++) hist[threadIdx.x + u * kSortBlock] = 0;
    float gmn[3], gmx[3];
#pragma unroll
    for (int a = 0; a < 3; a++) { gmn[a] = order_val(sceneKeys[a]); gmx[a] = order_val(sceneKeys[3 + a]); }
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * kSortTile;
#pragma unroll 4
    for (int j = 0; j < kSortItems; j++) {
        const size_t i = base + (size_t)j * kSortBlock + threadIdx.x;
        if (i >= (size_t)n) break;
        uint32_t code = 0;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            float c = primBox[6 * i + a] * 0.5f + primBox[6 * i + 3 + a] * 0.5f;  // centroid
            float o = c - gmn[a];
            if (gmx[a] > gmn[a]) o /= gmx[a] - gmn[a];                         // Bounds3::offset
            float s = (float)(1 << 10) * o;
            code |= left_shift3((uint32_t)s) << a;
        }
        codes[i] = code;
        atomicAdd(&hist[code & (kBins - 1)], 1u);
    }
    __syncthreads();
    store_tile_counts(hist, counts);
}

__global__ __launch_bounds__(kSortBlock) void k_radix_count(const uint32_t* __restrict__ keys, int n, int shift,
                                                            uint32_t* __restrict__ counts) {
    __shared__ uint32_t hist[kBins];
#pragma unroll
    for (int u = 0; u < kDigitsPerThread; u++) hist[threadIdx.x + u * kSortBlock] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * kSortTile;
    uint32_t k[kSortItems];
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
        const size_t i = base + (size_t)j * kSortBlock + threadIdx.x;
        k[j] = i < (size_t)n ? keys[i] : 0xffffffffu;
    }
#pragma unroll
    for (int j = 0; j < kSortItems; j++)
        if (k[j] != 0xffffffffu) atomicAdd(&hist[(k[j] >> shift) & (kBins - 1)], 1u);
    __syncthreads();
    store_tile_counts(hist, counts);
}

// Block g: exclusive prefix of each digit's counts over the tiles of segment g (in place), and the
// segment's digit totals.
__global__ __launch_bounds__(kSortBlock) void k_scan_segments(uint32_t* __restrict__ counts, int tiles, uint32_t* __restrict__ segBase) {
    const int t0 = blockIdx.x * kSegTiles, m = min(kSegTiles, tiles - t0);
    uint4* rows = reinterpret_cast<uint4*>(counts + (size_t)t0 * kBins) + threadIdx.x;
    uint4 c[kSegTiles];
#pragma unroll
    for (int i = 0; i < kSegTiles; i++)
        if (i < m) c[i] = rows[(size_t)i * (kBins / 4)];
    uint4 run = make_uint4(0, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < kSegTiles; i++) {
        if (i < m) {
            rows[(size_t)i * (kBins / 4)] = run;
            run.x += c[i].x; run.y += c[i].y; run.z += c[i].z; run.w += c[i].w;
        }
    }
    reinterpret_cast<uint4*>(segBase + (size_t)blockIdx.x * kBins)[threadIdx.x] = run;
}

// One block: thread d turns the segment totals of digit d into its base in each segment: the
// exclusive scan of the digit totals (the digit's base in the pass) plus the totals of the earlier
// segments (in place). The first kScanHeld segments (every segment up to 2^23 keys) are read once,
// all loads in flight together, and kept in registers for the second sweep; later ones are read
// again in windows.
constexpr int kScanHeld = 2 * kScanWindow;
__global__ __launch_bounds__(kBins) void k_scan_digits(uint32_t* __restrict__ segBase, int segs) {
    __shared__ uint32_t scratch[kBins / 64];
    const int d = threadIdx.x;
    uint32_t held[kScanHeld];
    uint32_t total = 0;
#pragma unroll
    for (int i = 0; i < kScanHeld; i++) held[i] = i < segs ? segBase[(size_t)i * kBins + d] : 0u;
#pragma unroll
    for (int i = 0; i < kScanHeld; i++) total += held[i];
    for (int g0 = kScanHeld; g0 < segs; g0 += kScanWindow) {
        uint32_t v[kScanWindow];
#pragma unroll
        for (int i = 0; i < kScanWindow; i++) v[i] = (g0 + i < segs) ? segBase[(size_t)(g0 + i) * kBins + d] : 0u;
#pragma unroll
        for (int i = 0; i < kScanWindow; i++) total += v[i];
    }
    uint32_t run;
    block_scan<kBins>(total, &run, scratch);
#pragma unroll
    for (int i = 0; i < kScanHeld; i++) {
        if (i < segs) segBase[(size_t)i * kBins + d] = run;
        run += held[i];
    }
    for (int g0 = kScanHeld; g0 < segs; g0 += kScanWindow) {
        uint32_t v[kScanWindow];
#pragma unroll
        for (int i = 0; i < kScanWindow; i++) v[i] = (g0 + i < segs) ? segBase[(size_t)(g0 + i) * kBins + d] : 0u;
#pragma unroll
        for (int i = 0; i < kScanWindow; i++) {
            if (g0 + i < segs) segBase[(size_t)(g0 + i) * kBins + d] = run;
            run += v[i];
        }
    }
}

// One pass's scatter of a 4096-key tile. kFirst: the values are the keys' input positions.
template <bool kFirst>
__global__ __launch_bounds__(kSortBlock) void k_radix_scatter(const uint32_t* __restrict__ keysIn, const int32_t* __restrict__ valsIn,
                                                              uint32_t* __restrict__ keysOut, int32_t* __restrict__ valsOut, int n,
                                                              int shift, int tiles, const uint32_t* __restrict__ counts,
                                                              const uint32_t* __restrict__ segBase, int32_t* __restrict__ status) {
    // The staged tile (keys, then values) overlays the per-wave digit counters: each key's position
    // is in registers before the tile is staged, so 40 KB of LDS let 3 workgroups share a CU.
    __shared__ uint32_t pool[2 * kSortTile];
    static_assert(kSortWaves * kBins <= 2 * kSortTile, "the digit counters fit the staging area");
    uint32_t (*waveCnt)[kBins] = reinterpret_cast<uint32_t (*)[kBins]>(pool);  // per-wave digit counts, then offsets
    uint32_t* sKeys = pool;
    int32_t* sVals = reinterpret_cast<int32_t*>(pool + kSortTile);
    __shared__ uint32_t tileStart[kBins];
    __shared__ uint32_t globalBase[kBins];
    __shared__ uint32_t scratch[kSortWaves];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int tile = xcd_tile(blockIdx.x, tiles);
    for (int q = t; q < kSortWaves * kBins; q += kSortBlock) (&waveCnt[0][0])[q] = 0;
    {  // the digits' bases in this tile: segment base (with the digit's base in the pass) + prefix in the segment
        const uint4 sb = reinterpret_cast<const uint4*>(segBase + (size_t)(tile / kSegTiles) * kBins)[t];
        const uint4 tb = reinterpret_cast<const uint4*>(counts + (size_t)tile * kBins)[t];
        reinterpret_cast<uint4*>(globalBase)[t] = make_uint4(sb.x + tb.x, sb.y + tb.y, sb.z + tb.z, sb.w + tb.w);
    }
    const size_t base = (size_t)tile * kSortTile;
    const int valid = (int)min((size_t)kSortTile, (size_t)n - base);
    uint32_t key[kSortItems];
    int32_t val[kSortItems];
    uint32_t rank[kSortItems];
    const size_t wbase = base + (size_t)wave * (64 * kSortItems);
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
        const size_t idx = wbase + (size_t)j * 64 + lane;
        const bool ok = idx < (size_t)n;
        key[j] = ok ? keysIn[idx] : 0xffffffffu;  // tail keys rank last in digit kBins-1 and are never written
        if (kFirst) val[j] = (int32_t)idx;
        else val[j] = ok ? valsIn[idx] : 0;
    }
    const uint64_t ltMask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t* cnt = waveCnt[wave];
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
        const uint32_t d = (key[j] >> shift) & (kBins - 1);
        uint64_t peers = ~0ull;
#pragma unroll
        for (int b = 0; b < kSortBits; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const uint32_t before = (uint32_t)__popcll(peers & ltMask);
        const uint32_t prev = cnt[d];  // every lane reads before the group's lowest lane writes
        rank[j] = prev + before;
        if (before == 0) cnt[d] = prev + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    // per-wave offsets and the tile's local digit starts (tail keys sort after every valid key)
    uint32_t sum = 0, run[kDigitsPerThread];
#pragma unroll
    for (int u = 0; u < kDigitsPerThread; u++) {
        const int d = kDigitsPerThread * t + u;
        uint32_t r = 0;
#pragma unroll
        for (int w = 0; w < kSortWaves; w++) {
            const uint32_t c = waveCnt[w][d];
            waveCnt[w][d] = r;
            r += c;
        }
        run[u] = r;
        sum += r;
    }
    uint32_t start;
    block_scan<kSortBlock>(sum, &start, scratch);
#pragma unroll
    for (int u = 0; u < kDigitsPerThread; u++) {
        tileStart[kDigitsPerThread * t + u] = start;
        start += run[u];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
        const uint32_t d = (key[j] >> shift) & (kBins - 1);
        rank[j] += tileStart[d] + waveCnt[wave][d];  // the key's position in the staged tile
    }
    __syncthreads();  // every counter is read before the staging overwrites them
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
        sKeys[rank[j]] = key[j];
        sVals[rank[j]] = val[j];
    }
    __syncthreads();
    for (int i = t; i < valid; i += kSortBlock) {
        const uint32_t k = sKeys[i];
        const uint32_t d = (k >> shift) & (kBins - 1);
        const size_t dst = (size_t)globalBase[d] + (uint32_t)(i - tileStart[d]);
        if (dst >= (size_t)n) {  // only inconsistent counts get here
            atomicOr(status, kStatusSort);
            continue;
        }
        keysOut[dst] = k;
        valsOut[dst] = sVals[i];
    }
}

// ---------------------------------------------------------------- hierarchy and bounds
// The Karras tree (BVH.cuh:173-236) is the binary radix tree of the sorted codes (with equal codes
// told apart by position, delta_at): interior node [l, r] splits where the adjacent delta inside
// it is smallest, and its parent splits at r (it is a left child) when delta(r, r+1) exceeds
// delta(l-1, l), else at l-1 (a right child). Karras numbers a left child by its last leaf and a
// right child by its first (the split g of a node has children g and g+1), the root 0. So the tree
// and its bounds (BVH.cuh:238-262: the second arrival at a node unions its children's boxes) are
// built in one bottom-up climb from the leaves: a node arriving at its parent's split records its
// outer end there; the second arrival learns the parent's range from it, unions the children and
// climbs on. The result equals the top-down construction and refit bit for bit: same nodes, same
// numbering, and every box is bUnion(left box, right box).
//   level 1: 512 sorted leaves per chunk, a persistent workgroup per resident slot; arrival slots,
//            child records and boxes in LDS. A climb stops when its parent's split lies left of
//            the chunk (one node per chunk, filed as the chunk's root) or when it arrives first.
//            Unmatched arrivals are written to the split's global slot, with a record of the
//            waiting node.
//   level 2: one workgroup per 65536 leaves climbs on from its chunks' roots; it files one root.
//   level 3: one workgroup climbs from the level-2 roots.
// Each level's climbs meet each other in LDS and read only what earlier launches wrote, so no
// climb needs a device-scope fence.
// Climbs of one workgroup exchange nodes through LDS only. A wave's LDS operations execute in the
// order it issues them, so a box written before an exchange is in LDS before the exchange is, and a
// partner reads it only after its own exchange returned; the climbs therefore order their LDS
// accesses with a compiler barrier instead of a fence (a fence would wait for every outstanding LDS
// operation at each step of a latency-bound climb).
#define LTS_LDS_ORDER() __atomic_signal_fence(__ATOMIC_SEQ_CST)

constexpr int kChunk = 512;
constexpr int kGroup2 = 128;                // level-1 chunks per level-2 group
constexpr int kChunk2 = kGroup2 * kChunk;   // leaves per level-2 group
constexpr int kGroup3 = 1024;               // level-2 groups (at most 2^26 / kChunk2)
static_assert((1 << 26) / kChunk2 <= kGroup3, "level 3 climbs every level-2 root in one workgroup");
constexpr int kUnset = INT32_MIN;
constexpr int kMaxClimb = 64;  // steps of a climb above level 1: the tree is at most 56 interior nodes deep

// v_min_f32 / v_max_f32 as fminf / fmaxf execute them, without the canonicalising v_max x, x the
// compiler puts before each operand it cannot prove canonical (IEEE mode). Canonicalising changes
// only signalling NaNs (denormals are kept), and every box value here is the result of a min or a
// max (the primitive boxes of tri_box included) or a copy of one, which is canonical.
__device__ __forceinline__ float box_min(float a, float b) {
    float r;
    asm("v_min_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ float box_max(float a, float b) {
    float r;
    asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// bUnion(Bounds3, Bounds3) followed by the Bounds3(p1, p2) constructor
__device__ __forceinline__ void union_box(const float* a, const float* b, float* o) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
        float lo = box_min(a[k], b[k]), hi = box_max(a[3 + k], b[3 + k]);
        o[k] = box_min(lo, hi);
        o[3 + k] = box_max(lo, hi);
    }
}

__device__ __forceinline__ void load_box(const float* box, int node, float* o) {
#pragma unroll
    for (int q = 0; q < 6; q++) o[q] = box[6 * (size_t)node + q];
}
__device__ __forceinline__ void store_box(float* box, int node, const float* o) {
#pragma unroll
    for (int q = 0; q < 6; q++) box[6 * (size_t)node + q] = o[q];
}

// Level 1. LDS holds the chunk's nodes under local indices: leaf cs + x at x, interior node cs + x
// at kChunk + x. The climbs advance in rounds over a dense work list: in a round every node that
// is complete and not yet parented takes one step (one thread each), and the nodes completed by
// second arrivals form the next round's list, so a lane whose climb has ended never idles in a
// wave while other lanes of the wave climb on. A work item carries the node's range and the deltas
// at both of its ends, and a first arrival leaves its outer end and the delta there in the split's
// slot, so a step reads only its item, the slot and the two child boxes. Every box a round reads
// was written before the barrier that opened the round (a node's box is written when the node is
// completed); the exchanges at the splits only decide which arrival completes the parent.
__device__ __forceinline__ int local_index(int c, int n, int cs) {
    return c >= n - 1 ? c - (n - 1) - cs : kChunk + (c - cs);
}

// A workgroup barrier that orders LDS only: global loads and stores stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

constexpr uint32_t kNone16 = 0xffffu;
static_assert(kChunk <= 512, "level-1 items pack chunk-local ends in 9 bits");

// work item: l | r << 9 | (dL + 1) << 18 | (dR + 1) << 25 (chunk-local ends, end deltas -1 .. 71)
__device__ __forceinline__ uint32_t pack_item(int l, int r, int dL, int dR) {
    return (uint32_t)l | ((uint32_t)r << 9) | ((uint32_t)(dL + 1) << 18) | ((uint32_t)(dR + 1) << 25);
}

// Level-1 inputs of one thread for a chunk, loaded a chunk ahead: the codes around leaf cs + t
// and a third of the chunk's leaf boxes (mode 0: float2 runs t, t + kChunk, t + 2 kChunk of the
// contiguous leaf section; otherwise the box of leaf cs + t).
constexpr int kTreeThreads = kChunk / 2;  // a thread per two leaves: five chunks share a CU's LDS

// Level-1 inputs of one thread for a chunk, loaded a chunk ahead: the codes around leaves cs + t and
// cs + t + kTreeThreads, and a third of the chunk's leaf boxes (mode 0: float2 entries t + j
// kTreeThreads of the contiguous leaf section; otherwise the boxes of the thread's two leaves).
struct TreeL1Fetch {
    uint32_t cPrev[2], c[2], cNext[2];
    float2 b[6];
};

__device__ __forceinline__ void tree_l1_fetch(TreeL1Fetch& f, const uint32_t* __restrict__ codes,
                                              const float* __restrict__ primBox, const int32_t* __restrict__ perm,
                                              int n, int mode, int chunk, int t) {
    // every load is unconditional, at clamped in-range addresses (lanes past the end reload the last
    // entries), so no branch joins the loaded registers and nothing waits for them until they are used
    const int cs = chunk * kChunk, m = min(kChunk, n - cs);
    int pk[2] = {0, 0};
#pragma unroll
    for (int h = 0; h < 2; h++)
        if (mode != 0) pk[h] = perm[cs + min(t + h * kTreeThreads, m - 1)];  // first: the box loads wait for them
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int k = cs + min(t + h * kTreeThreads, m - 1);
        f.c[h] = codes[k];
        f.cPrev[h] = codes[max(k - 1, 0)];
        f.cNext[h] = codes[min(k + 1, n - 1)];
    }
    const float2* src = reinterpret_cast<const float2*>(primBox) + 3 * (size_t)cs;
#pragma unroll
    for (int j = 0; j < 6; j++) {
        const float2* own = reinterpret_cast<const float2*>(primBox) + 3 * (size_t)pk[j / 3];
        f.b[j] = mode == 0 ? src[min(t + j * kTreeThreads, 3 * m - 1)] : own[j % 3];
    }
}

// Persistent: workgroup b builds chunks b, b + gridDim.x, ..., loading each chunk's inputs while
// the previous chunk climbs, so the climbs are not held up by global-load latency. The climbs are
// latency-bound, so residency is what counts: 256 threads per 512-leaf chunk (round 0 takes the
// leaves in two halves; no later round has more than 256 nodes) let five chunks share a CU's LDS,
// and at most 64 SGPRs keep the register files from allowing fewer.
__global__ __launch_bounds__(kTreeThreads) __attribute__((amdgpu_num_sgpr(64))) void k_tree_l1(
    const uint32_t* __restrict__ codes, const float* __restrict__ primBox, const int32_t* __restrict__ perm, int n, int mode,
    int32_t* __restrict__ left, int32_t* __restrict__ right, int32_t* __restrict__ parent, float* __restrict__ box,
    int32_t* __restrict__ slots, ClimbRec* __restrict__ pend, int32_t* __restrict__ chunkRoot, ClimbRec* __restrict__ chunkRec,
    int32_t* __restrict__ status) {
    __shared__ __align__(16) float nbox[2 * kChunk * 6];
    // split cs + x: -1, or a first arrival's outer end - cs | (delta at that end + 1) << 9
    __shared__ int32_t slot[kChunk];
    // interior node cs + x completed here: its split - cs, bit 9 when its left child is the leaf at
    // the split, bit 10 when its right child is the leaf after it (else the interior nodes there)
    __shared__ uint16_t ssplit[kChunk];
    __shared__ uint16_t sp[2 * kChunk];  // parent of each local node - cs (kNone16: set by a later level)
    // round r >= 1 reads list r & 1. A round's items are the nodes the round before completed: at
    // most half the leaves after round 0, and never more than the items of the round before.
    __shared__ uint32_t work[2][kTreeThreads];
    __shared__ int32_t workCount[3];  // items appended for round r: counter r % 3
    __shared__ int32_t sRoot[3];      // the chunk's root: its right end - cs (or -1) and its end deltas
    const int t = threadIdx.x, lane = t & 63;
    const int chunks = (n + kChunk - 1) / kChunk;
    TreeL1Fetch f;
    if ((int)blockIdx.x < chunks) tree_l1_fetch(f, codes, primBox, perm, n, mode, blockIdx.x, t);
    __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0): every chunk starts with its inputs in registers
    // one climb step of node [l, r] (chunk-local, end deltas a, b): true, with the completed
    // parent's item, for a second arrival; files the chunk's root
    auto step = [&](int l, int r, int a, int b, uint32_t& item) -> bool {
        if (a < 0 && b < 0) return false;  // the root of the whole tree
        const bool asLeft = b > a;
        const int p = asLeft ? r : l - 1;
        if (p < 0) {  // its parent splits left of the chunk: the chunk's root (one node)
            sRoot[0] = r;
            sRoot[1] = a;
            sRoot[2] = b;
            return false;
        }
        const int mine = asLeft ? (l | ((a + 1) << 9)) : (r | ((b + 1) << 9));
        const int other = atomicExch(&slot[p], mine);
        if (other < 0) return false;  // first arrival
        slot[p] = -1;
        const int oe = other & 511, od = (other >> 9) - 1;
        const int L = asLeft ? l : oe, R = asLeft ? oe : r;
        const int pL = asLeft ? a : od, pR = asLeft ? od : b;
        const bool leafL = L == p, leafR = R == p + 1;
        const int li = leafL ? p : kChunk + p, ri = leafR ? p + 1 : kChunk + p + 1;
        float x[6], y[6], o[6];
#pragma unroll
        for (int q = 0; q < 6; q++) {
            x[q] = nbox[6 * li + q];
            y[q] = nbox[6 * ri + q];
        }
        union_box(x, y, o);
        // Karras numbering; local 0 is the whole tree's root only in chunk 0
        const int me = (pL < 0 && pR < 0) ? 0 : (pR > pL ? R : L);
#pragma unroll
        for (int q = 0; q < 6; q++) nbox[6 * (kChunk + me) + q] = o[q];
        ssplit[me] = (uint16_t)(p | (leafL ? 512 : 0) | (leafR ? 1024 : 0));
        sp[li] = (uint16_t)me;
        sp[ri] = (uint16_t)me;
        item = pack_item(L, R, pL, pR);
        return true;
    };
    // the completed parents of the lanes with `next` join list `list`: one LDS atomic per wave
    auto append = [&](int list, int counter, bool next, uint32_t item) {
        const unsigned long long ballot = __ballot(next);
        if (ballot) {
            int base = 0;
            if (lane == 0) base = atomicAdd(&workCount[counter], (int)__popcll(ballot));
            base = __builtin_amdgcn_readfirstlane(base);
            if (next) work[list][base + (int)__popcll(ballot & ((1ull << lane) - 1ull))] = item;
        }
    };
    for (int chunk = blockIdx.x; chunk < chunks; chunk += gridDim.x) {
        const int cs = chunk * kChunk, ce = min(cs + kChunk, n);
        const int m = ce - cs;  // leaves of the chunk
        int dL[2] = {-1, -1}, dR[2] = {-1, -1};  // deltas at both ends of the thread's two leaves
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int k = cs + t + h * kTreeThreads;
            if (k < ce) {
                if (k >= 1) dL[h] = delta_codes(f.cPrev[h], f.c[h], n, k - 1, k);
                if (k + 1 < n) dR[h] = delta_codes(f.c[h], f.cNext[h], n, k, k + 1);
            }
        }
        if (t == 0) sRoot[0] = -1;
        if (t < 3) workCount[t] = 0;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            slot[t + h * kTreeThreads] = -1;
            ssplit[t + h * kTreeThreads] = kNone16;
        }
#pragma unroll
        for (int h = 0; h < 4; h++) sp[t + h * kTreeThreads] = kNone16;
#pragma unroll
        for (int j = 0; j < 6; j++) {
            if (mode == 0) {
                const int i = t + j * kTreeThreads;
                if (i < 3 * m) reinterpret_cast<float2*>(nbox)[i] = f.b[j];
            } else {
                const int x = t + (j / 3) * kTreeThreads;
                if (x < m) reinterpret_cast<float2*>(nbox)[3 * x + j % 3] = f.b[j];
            }
        }
        tree_l1_fetch(f, codes, primBox, perm, n, mode, min(chunk + (int)gridDim.x, chunks - 1), t);
        lds_barrier();
        if (n > 1) {
            // round 0: every leaf, in two halves
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int x = t + h * kTreeThreads;
                uint32_t item = 0;
                const bool next = x < m && step(x, x, dL[h], dR[h], item);
                append(1, 1, next, item);
            }
            if (t == 0) workCount[2] = 0;
            lds_barrier();
            // round r >= 1 reads list r & 1 (its count in counter r % 3) and appends to list
            // (r + 1) & 1 through counter (r + 1) % 3, while counter (r + 2) % 3, last read before the
            // barrier that opened round r, is cleared
            int count = workCount[1];
            uint32_t w = work[1][t];  // read with the count that bounds it: one LDS round trip
            for (int round = 1; count > 0; round++) {
                if (round == kChunk) {  // every round completes at least one node: never reached
                    if (t == 0) atomicOr(status, kStatusClimb);
                    break;
                }
                uint32_t item = 0;
                const bool next = t < count && step((int)(w & 511u), (int)((w >> 9) & 511u), (int)((w >> 18) & 127u) - 1,
                                                    (int)(w >> 25) - 1, item);
                append((round + 1) & 1, (round + 1) % 3, next, item);
                if (t == 0) workCount[(round + 2) % 3] = 0;
                lds_barrier();
                count = workCount[(round + 1) % 3];
                w = work[(round + 1) & 1][t];
            }
        }
        // the next chunk's inputs arrived during the climbs: wait for them here, before the stores
        // below, so that the wait at the next chunk does not also wait for those stores
        __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0)
        // boxes as contiguous runs: interior nodes cs .. (those completed above later levels overwrite),
        // leaves (n - 1) + cs .. (8-byte aligned; in mode 0 the primitive boxes were written there)
        {
            const int mi = min(ce, n - 1) - cs;
            float* dst = box + 6 * (size_t)cs;
            const float* ib = nbox + 6 * kChunk;
            const int whole = mi > 0 ? (6 * mi) / 4 : 0;
            for (int i = t; i < whole; i += kTreeThreads) reinterpret_cast<float4*>(dst)[i] = reinterpret_cast<const float4*>(ib)[i];
            for (int i = 4 * whole + t; i < 6 * mi; i += kTreeThreads) dst[i] = ib[i];
            if (mode != 0) {
                float2* ldst = reinterpret_cast<float2*>(box + 6 * ((size_t)(n - 1) + cs));
                for (int i = t; i < 3 * m; i += kTreeThreads) ldst[i] = reinterpret_cast<const float2*>(nbox)[i];
            }
        }
        for (int x = t; x < m; x += kTreeThreads) {
            const int k = cs + x;
            if (k < n - 1) {
                const uint32_t sv = ssplit[x];
                if (sv != kNone16) {
                    const int g = cs + (int)(sv & 511u);
                    left[k] = (sv & 512u) ? (n - 1) + g : g;
                    right[k] = (sv & 1024u) ? (n - 1) + g + 1 : g + 1;
                }
                const uint32_t pv = sp[kChunk + x];
                if (pv != kNone16) parent[k] = cs + (int)pv;
                const int sw = slot[x];
                const int v = sw >= 0 ? cs + (sw & 511) : -1;
                slots[k] = v;
                if (v >= 0) {  // a first arrival waits at split k: publish its box and the delta at its outer end
                    const bool leftChild = v <= k;
                    const int child = leftChild ? (v == k ? (n - 1) + k : k) : (v == k + 1 ? (n - 1) + k + 1 : k + 1);
                    ClimbRec rec;
                    const float* cb = &nbox[6 * local_index(child, n, cs)];
#pragma unroll
                    for (int q = 0; q < 6; q++) rec.box[q] = cb[q];
                    rec.a = (sw >> 9) - 1;
                    rec.b = 0;
                    pend[k] = rec;
                }
            }
            const uint32_t pv = sp[x];
            if (pv != kNone16) parent[(n - 1) + k] = cs + (int)pv;
        }
        if (t == 0) {
            const int rr = sRoot[0];
            chunkRoot[chunk] = rr >= 0 ? cs + rr : -1;
            if (rr >= 0) {  // the node [cs, cs + rr] (a right child: index cs) with its box and end deltas
                ClimbRec rec;
                const float* cb = &nbox[6 * (rr == 0 ? 0 : kChunk)];
#pragma unroll
                for (int q = 0; q < 6; q++) rec.box[q] = cb[q];
                rec.a = sRoot[1];
                rec.b = sRoot[2];
                chunkRec[chunk] = rec;
            }
            if (cs == 0) parent[0] = -1;
        }
        lds_barrier();  // the chunk's LDS is read before the next chunk overwrites it
    }
}

// Levels 2 and 3: climbs from the roots handed on by the level below. A climb carries its node's
// box and the deltas at its ends; a first arrival of an earlier level left its outer end in the
// split's global slot and its box and outer delta in the split's record, loaded together in one
// round trip, so a step needs no other load. Arrivals of the same launch meet in an LDS hash table
// keyed by the split (entries (split + 1) << 32 | climb, claimed by compare-and-swap and never
// removed); a climb's LDS record holds its node for its final, first arrival. Unmatched entries are
// published to the global slots and records for the next level.
__device__ __forceinline__ void store_rec(ClimbRec* dst, const float* b, int a, int c) {
    ClimbRec r;
#pragma unroll
    for (int q = 0; q < 6; q++) r.box[q] = b[q];
    r.a = a;
    r.b = c;
    *dst = r;
}

// Units of `span` leaves (512-leaf chunks for level 2, 128 of them per group; the groups of level 2
// for level 3, all in one group). Climb j of a group starts at the root of unit u0 + j (unit u0's
// root, whose parent splits left of the group, is the group's own root: handed on unless a climb
// completes a larger one).
template <int kClimbs, int kHashCap>
__global__ __launch_bounds__(kClimbs) void k_tree_up(int n, int units, int span, const int32_t* __restrict__ rootEnd,
                                                     const ClimbRec* __restrict__ rootRec, int32_t* __restrict__ left,
                                                     int32_t* __restrict__ right, int32_t* __restrict__ parent,
                                                     float* __restrict__ box, int32_t* __restrict__ slots,
                                                     ClimbRec* __restrict__ pend, int32_t* __restrict__ outEnd,
                                                     ClimbRec* __restrict__ outRec, int32_t* __restrict__ status) {
    __shared__ unsigned long long hash[kHashCap];
    __shared__ float recBox[kClimbs][6];
    __shared__ int32_t recEnd[kClimbs], recDelta[kClimbs];
    __shared__ int32_t sEnd;
    __shared__ ClimbRec sRec;
    const int j = threadIdx.x;
    const int u0 = blockIdx.x * kClimbs;
    const int cs = u0 * span;
    for (int i = j; i < kHashCap; i += kClimbs) hash[i] = 0;
    if (j == 0) {
        sEnd = cs > 0 ? rootEnd[u0] : -1;
        if (sEnd >= 0) sRec = rootRec[u0];
    }
    __syncthreads();
    constexpr int kHashBits = __builtin_ctz(kHashCap);
    const int u = u0 + j;
    int r = (j >= 1 && u < units) ? rootEnd[u] : -1;
    if (r >= 0) {
        int l = u * span;
        const ClimbRec start = rootRec[u];
        float o[6];
#pragma unroll
        for (int q = 0; q < 6; q++) o[q] = start.box[q];
        int dL = start.a, dR = start.b;
        bool asLeft = dR > dL;
        int p = asLeft ? r : l - 1;
        // the slot and record of the next split are loaded before the stores of the node just
        // completed: on CDNA a load's wait also waits for the stores issued before it
        int32_t g = p >= cs ? slots[p] : -1;
        ClimbRec pr = p >= cs ? pend[p] : ClimbRec{};
        for (int step = 0;; step++) {
            if (step == kMaxClimb) {
                atomicOr(status, kStatusClimb);
                break;
            }
            if (p < cs) {  // l == cs: this group's root
                sEnd = r;
                store_rec(&sRec, o, dL, dR);
                break;
            }
            float ob[6];
            int other, od;
            if (g >= 0) {  // the other child arrived in an earlier level
                other = g;
#pragma unroll
                for (int q = 0; q < 6; q++) ob[q] = pr.box[q];
                od = pr.a;
            } else {
#pragma unroll
                for (int q = 0; q < 6; q++) recBox[j][q] = o[q];
                recEnd[j] = asLeft ? l : r;
                recDelta[j] = asLeft ? dL : dR;
                LTS_LDS_ORDER();
                const unsigned long long key = ((unsigned long long)(p + 1) << 32) | (unsigned)j;
                uint32_t h = ((uint32_t)p * 0x9E3779B1u) >> (32 - kHashBits);
                int found = -1;
                bool first = false;
                for (int probe = 0; probe < kHashCap; probe++) {
                    const unsigned long long old = atomicCAS(&hash[h], 0ull, key);
                    if (old == 0ull) { first = true; break; }
                    if ((old >> 32) == (unsigned long long)(p + 1)) { found = (int)(old & 0xffffffffu); break; }
                    h = (h + 1) & (kHashCap - 1);
                }
                if (first) break;
                if (found < 0) {  // more arrivals than entries: never reached (kHashCap > kClimbs)
                    atomicOr(status, kStatusClimb);
                    break;
                }
                LTS_LDS_ORDER();
                hash[h] = ((unsigned long long)(p + 1) << 32) | 0xffffffffu;  // matched
#pragma unroll
                for (int q = 0; q < 6; q++) ob[q] = recBox[found][q];
                other = recEnd[found];
                od = recDelta[found];
            }
            const int L = asLeft ? l : other, R = asLeft ? other : r;
            const int lc = (L == p) ? (n - 1) + p : p;
            const int rc = (R == p + 1) ? (n - 1) + p + 1 : p + 1;
            float un[6];
            if (asLeft) union_box(o, ob, un);
            else union_box(ob, o, un);
#pragma unroll
            for (int q = 0; q < 6; q++) o[q] = un[q];
            if (asLeft) dR = od;
            else dL = od;
            const bool root = dL < 0 && dR < 0;
            const int me = root ? 0 : (dR > dL ? R : L);
            l = L;
            r = R;
            if (!root) {
                asLeft = dR > dL;
                p = asLeft ? r : l - 1;
                if (p >= cs) {
                    g = slots[p];
                    pr = pend[p];
                }
            }
            store_box(box, me, o);
            left[me] = lc;
            right[me] = rc;
            parent[lc] = me;
            parent[rc] = me;
            if (root) break;
        }
    }
    __syncthreads();
    // first arrivals still waiting: published for the next level
    for (int i = j; i < kHashCap; i += kClimbs) {
        const unsigned long long e = hash[i];
        const uint32_t c = (uint32_t)(e & 0xffffffffu);
        if (e != 0ull && c != 0xffffffffu) {
            const int p = (int)(e >> 32) - 1;
            slots[p] = recEnd[c];
            store_rec(&pend[p], recBox[c], recDelta[c], 0);
        }
    }
    if (j == 0 && outEnd) {
        outEnd[blockIdx.x] = sEnd;
        if (sEnd >= 0) outRec[blockIdx.x] = sRec;
    }
}

__global__ __launch_bounds__(kBlock) void k_pack_nodes(const float* __restrict__ box, const int32_t* __restrict__ left,
                                                       const int32_t* __restrict__ right, int n, BvhNode* __restrict__ nodes,
                                                       int32_t* __restrict__ status) {
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n - 1) return;
    BvhNode nd;
    int l = left[i], r = right[i];
    if (l < 1 || l > 2 * n - 2 || r < 1 || r > 2 * n - 2) {  // only a broken build gets here
        atomicOr(status, kStatusTree);
        l = r = 2 * n - 2;
    }
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

// Level-1 workgroups resident on the current device at once (the persistent grid), from the
// occupancy calculator; kept per device after the first query.
static int tree_l1_grid() {
    static std::atomic<int> cached[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    int g = cached[dev].load(std::memory_order_relaxed);
    if (g > 0) return g;
    int perCu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, k_tree_l1, kTreeThreads, 0) != hipSuccess || perCu < 1) perCu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 1;
    g = perCu * cus;
    cached[dev].store(g, std::memory_order_relaxed);
    return g;
}
// blocks of the scene-bound kernel (a grid-stride loop: few partial bounds to reduce)
inline int bound_blocks(long n) { return std::min(blocks_for(n), 2048); }

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
    hipError_t e = hipSuccess;
#define TRY(x) do { e = (x); if (e != hipSuccess) { bvh_release(b); return e; } } while (0)
    TRY(dalloc(&b.primBox, 6 * (size_t)n));
    TRY(dalloc(&b.box, 6 * (size_t)(2 * n - 1)));
    TRY(dalloc(&b.keysA, n)); TRY(dalloc(&b.keysB, n));
    TRY(dalloc(&b.valsA, n)); TRY(dalloc(&b.valsB, n));
    TRY(dalloc(&b.counts, (size_t)kBins * sort_tiles(n)));
    TRY(dalloc(&b.segBase, (size_t)kBins * sort_segments(n)));
    TRY(dalloc(&b.left, n)); TRY(dalloc(&b.right, n));
    TRY(dalloc(&b.parent, 2 * (size_t)n));
    TRY(dalloc(&b.flags, n));
    TRY(dalloc(&b.slots, n));
    TRY(dalloc(&b.chunkRoot, (n + kChunk - 1) / kChunk));
    TRY(dalloc(&b.pend, n));
    TRY(dalloc(&b.chunkRec, (n + kChunk - 1) / kChunk));
    TRY(dalloc(&b.root2, (n + kChunk2 - 1) / kChunk2));
    TRY(dalloc(&b.rec2, (n + kChunk2 - 1) / kChunk2));
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
    k_ploc_init<<<blocks_for(n), kBlock, 0, s>>>(b.lastPrimBox, b.valsA, n, box, clA, parent);
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
        k_pack_nodes<<<blocks_for(n - 1), kBlock, 0, s>>>(box, left, right, n, b.nodes, b.status);
        TRY(hipMemcpyAsync(b.parent, parent, (size_t)(2 * n - 1) * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
        TRY(hipGetLastError());
        TRY(hipStreamSynchronize(s));
    }
#undef TRY
    cleanup();
    return hipSuccess;
}

void bvh_release(BvhBuffers& b) {
    void* ps[] = {b.primBox, b.box, b.keysA, b.keysB, b.valsA, b.valsB, b.counts, b.segBase,
                  b.left, b.right, b.parent, b.flags, b.sceneKeys, b.nodes, b.tris, b.slots, b.pend, b.chunkRoot, b.chunkRec, b.root2, b.rec2, b.partial,
                  b.nodes4, b.index4, b.scanTotals, b.status};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    b = BvhBuffers();
}

hipError_t bvh_enqueue(BvhBuffers& b, const float* soup, int n, int mode, bool pack, hipStream_t s, BvhStageEvents* ev) {
    if (n < 1 || n > b.cap) return hipErrorInvalidValue;
    b.n = n;
    b.tiles = sort_tiles(n);
    if (ev) (void)hipEventRecord(ev->e[0], s);
    // primitive boxes: in mode 0 leaf k holds primitive k, so they go straight to the leaf section
    // of the node bounds
    float* pb = mode == 0 ? b.box + 6 * (size_t)(n - 1) : b.primBox;
    b.lastPrimBox = pb;
    const int boundBlocks = bound_blocks(n);
    k_prim_bounds<<<boundBlocks, kBlock, 0, s>>>(soup, n, pb, b.partial, b.status);
    k_reduce_scene_keys<<<1, 1024, 0, s>>>(b.partial, boundBlocks, b.sceneKeys);
    k_morton<<<b.tiles, kSortBlock, 0, s>>>(pb, n, b.sceneKeys, b.keysB, b.counts);
    if (ev) (void)hipEventRecord(ev->e[1], s);
    // B -> A (values: input positions), A -> B, B -> A: the sorted pairs end in keysA / valsA
    uint32_t* kin = b.keysB; uint32_t* kout = b.keysA;
    int32_t* vin = nullptr; int32_t* vout = b.valsA;
    const int segs = sort_segments(n);
    for (int pass = 0; pass < kSortPasses; pass++) {
        const int shift = kSortBits * pass;
        if (pass > 0) k_radix_count<<<b.tiles, kSortBlock, 0, s>>>(kin, n, shift, b.counts);
        k_scan_segments<<<segs, kSortBlock, 0, s>>>(b.counts, b.tiles, b.segBase);
        k_scan_digits<<<1, kBins, 0, s>>>(b.segBase, segs);
        if (pass == 0)
            k_radix_scatter<true><<<b.tiles, kSortBlock, 0, s>>>(kin, vin, kout, vout, n, shift, b.tiles, b.counts,
                                                                 b.segBase, b.status);
        else
            k_radix_scatter<false><<<b.tiles, kSortBlock, 0, s>>>(kin, vin, kout, vout, n, shift, b.tiles, b.counts,
                                                                  b.segBase, b.status);
        std::swap(kin, kout);
        vin = vout;
        vout = (vout == b.valsA) ? b.valsB : b.valsA;
    }
    if (ev) (void)hipEventRecord(ev->e[2], s);
    const int chunks = (n + kChunk - 1) / kChunk;
    k_tree_l1<<<std::min(chunks, tree_l1_grid()), kTreeThreads, 0, s>>>(b.keysA, pb, b.valsA, n, mode, b.left, b.right, b.parent, b.box, b.slots,
                                        b.pend, b.chunkRoot, b.chunkRec, b.status);
    const int chunks2 = (n + kChunk2 - 1) / kChunk2;
    if (chunks > 1)
        k_tree_up<kGroup2, 2 * kGroup2><<<chunks2, kGroup2, 0, s>>>(n, chunks, kChunk, b.chunkRoot, b.chunkRec, b.left, b.right,
                                                                   b.parent, b.box, b.slots, b.pend, b.root2, b.rec2,
                                                                   b.status);
    if (chunks2 > 1)
        k_tree_up<kGroup3, 2 * kGroup3><<<1, kGroup3, 0, s>>>(n, chunks2, kChunk2, b.root2, b.rec2, b.left, b.right, b.parent,
                                                             b.box, b.slots, b.pend, nullptr, nullptr, b.status);
    if (ev) (void)hipEventRecord(ev->e[3], s);
    if (pack) {
        if (n > 1) k_pack_nodes<<<blocks_for(n - 1), kBlock, 0, s>>>(b.box, b.left, b.right, n, b.nodes, b.status);
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
    size_t at = 0;
    auto take = [&](auto*& ptr, size_t count) {
        const size_t o = carve(at, count * sizeof(*ptr));
        ptr = base ? reinterpret_cast<std::remove_reference_t<decltype(ptr)>>(base + o) : nullptr;
    };
    take(b.primBox, 6 * (size_t)n);
    take(b.keysB, n);
    take(b.valsB, n);
    take(b.counts, (size_t)kBins * sort_tiles(n));
    take(b.segBase, (size_t)kBins * sort_segments(n));
    take(b.slots, n);
    take(b.chunkRoot, (n + kChunk - 1) / kChunk);
    take(b.pend, n);
    take(b.chunkRec, (n + kChunk - 1) / kChunk);
    take(b.root2, (n + kChunk2 - 1) / kChunk2);
    take(b.rec2, (n + kChunk2 - 1) / kChunk2);
    take(b.partial, 6 * (size_t)blocks_for(n));
    take(b.sceneKeys, 8);
    take(b.status, 1);
    return at;
}

}  // namespace

// The largest build (kStaticMaxPrims) must fit the buffer.
static_assert(kStaticScratchBytes >= 256 * 16 + (4 * (6 + 1 + 1 + 1) + sizeof(ClimbRec)) * (size_t)kStaticMaxPrims +
                                         4 * (size_t)kBins * ((kStaticMaxPrims + kSortTile - 1) / kSortTile) +
                                         4 * (size_t)kBins * ((kStaticMaxPrims + kSortTile * kSegTiles - 1) / (kSortTile * kSegTiles)) +
                                         (4 + sizeof(ClimbRec)) * ((kStaticMaxPrims + kChunk - 1) / kChunk) +
                                         (4 + sizeof(ClimbRec)) * ((kStaticMaxPrims + kChunk2 - 1) / kChunk2) +
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
