// Internal interfaces between the C ABI layer and the kernel translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>

#include "lts_scene.h"
#include "mesh_xf.h"

namespace lts {

// Process-unique number for every (re)allocation of a buffer set. Captured graphs hold raw device
// pointers, and a reallocation can return an address that was just freed, so a captured graph is
// matched to its buffers by these numbers, never by addresses.
inline uint64_t next_buffer_generation() {
    static std::atomic<uint64_t> gen{0};
    return ++gen;
}

// A node handed between the levels of the bottom-up hierarchy: its box and two deltas (a chunk
// root: the deltas at both ends; a first arrival waiting at a split: the delta at its outer end).
struct alignas(32) ClimbRec {
    float box[6];
    int32_t a, b;
};

// Device buffers of one LBVH build (sized for `cap` primitives; reused across builds).
struct BvhBuffers {
    int cap = 0;
    int n = 0;
    int tiles = 0;
    float* primBox = nullptr;     // 6n    primitive bounds (mode 1 builds)
    const float* lastPrimBox = nullptr;  // where the last build put the primitive bounds
    float* box = nullptr;         // 6(2n-1) node bounds, unified index
    uint32_t* keysA = nullptr;    // n
    uint32_t* keysB = nullptr;
    int32_t* valsA = nullptr;
    int32_t* valsB = nullptr;
    uint32_t* counts = nullptr;   // tiles x 1024: a radix pass's digit counts per tile, then prefixes in the segment
    uint32_t* segBase = nullptr;  // segments x 1024: digit totals per 16-tile segment, then the digits' bases
    int32_t* left = nullptr;      // n-1
    int32_t* right = nullptr;     // n-1
    int32_t* parent = nullptr;    // 2n-1
    int32_t* flags = nullptr;     // n-1 (4-wide collapse)
    int32_t* slots = nullptr;     // n-1 arrival slot per split of the bottom-up hierarchy
    ClimbRec* pend = nullptr;     // n-1: the node waiting at a split (written where a slot is set)
    int32_t* chunkRoot = nullptr; // per 512-leaf chunk: last leaf of the node its climbs handed on, or -1
    ClimbRec* chunkRec = nullptr; // per 512-leaf chunk: that node
    int32_t* root2 = nullptr;     // the same per 65536-leaf group (level 2 to level 3)
    ClimbRec* rec2 = nullptr;
    uint32_t* partial = nullptr;   // per-block scene-bound keys
    uint32_t* sceneKeys = nullptr;// 6 ordered keys of the scene bounds
    int32_t* status = nullptr;    // failed bounded waits of the last build (bits; 0 = none)
    BvhNode* nodes = nullptr;     // n-1 (or 1 for n == 1)
    TriLeaf* tris = nullptr;      // n
    BvhQNode* nodes4 = nullptr;   // 4-wide collapse for rendering (allocated by bvh4_enqueue)
    int32_t* index4 = nullptr;    // n-1: 4-wide node of each even-depth interior node
    int32_t* scanTotals = nullptr;
    int cap4 = 0;
    // results after the sort: sorted codes and permutation live in keysA / valsA
};

hipError_t bvh_reserve(BvhBuffers& b, int n);
void bvh_release(BvhBuffers& b);
// Collapses the packed tree of the last build (mode 1, pack) into the 4-wide render BVH: interior
// nodes of even depth become 4-wide nodes (numbered in interior order, so the root is node 0) and
// take the boxes of their children's children. Deterministic. Waits for the stream and returns
// the tree's depth (edges down to the deepest leaf) and the 4-wide node count.
// 4-wide render nodes of the binary tree in b (greedy surface-area collapse, or the even-depth
// collapse when *even is set on entry or the greedy tree is too deep; *even tells which was used).
hipError_t bvh4_enqueue(BvhBuffers& b, hipStream_t stream, int* depth, int* nodes4, bool* even);
// Render tree of a scene: PLOC over the sorted leaves of the last bvh_enqueue (which must have run
// with pack). When the PLOC tree is at most maxDepth deep it replaces b.nodes and b.parent (the
// 4-wide collapse then takes it); otherwise the Karras tree stays. *depth = the PLOC tree's depth.
hipError_t ploc_enqueue(BvhBuffers& b, hipStream_t stream, int maxDepth, int* depth);

struct BvhStageEvents {
    hipEvent_t e[5];  // start, after bounds+morton, after sort, after tree, after pack
};

// Build without allocation (lts_bvh_build_static): all temporaries in one static device buffer of
// kStaticScratchBytes per device, enough for kStaticMaxPrims primitives.
constexpr int kStaticMaxPrims = 1 << 18;
constexpr size_t kStaticScratchBytes = (size_t)24 << 20;
constexpr int kStaticMaxDevices = 64;
// The buffer's device address on `device` (the current device), looked up once per device.
hipError_t bvh_static_scratch(int device, void** ptr);
// Enqueues a build of n <= kStaticMaxPrims primitives whose results go to the caller's device arrays
// (sorted codes, permutation, children, parents, unified node bounds) and whose temporaries live
// in the static buffer; *status and *sceneKeys point at the build's status word and scene-bound
// keys inside that buffer.
hipError_t bvh_enqueue_static(void* scratch, const float* soup, int n, int mode, uint32_t* codesSorted, int32_t* perm,
                              int32_t* left, int32_t* right, int32_t* parent, float* bounds, hipStream_t s,
                              int32_t** status, uint32_t** sceneKeys);

// Enqueues the whole build on `stream`. `soup` is a device pointer to n*9 floats.
// mode 0: leaf k holds primitive k (reference); mode 1: leaf k holds perm[k].
// When pack is set, also writes nodes / tris for the traversal kernels (mode 1 layout).
hipError_t bvh_enqueue(BvhBuffers& b, const float* soup, int n, int mode, bool pack, hipStream_t stream,
                       BvhStageEvents* ev);

// Mesh world transform (mesh.hip): points and normals (3 floats each, device memory) in place.
hipError_t mesh_transform(const MeshXf& xf, float* P, int nP, float* N, int nN, hipStream_t stream);

// Render launches (render.hip).
struct RenderLaunch {
    int spp;
    int maxBounce;
    int rrStart;
    uint32_t seed;
    int stratified;
    int s0, s1;          // sample range of this pass
    float* sampleL;      // [s1 - s0][H][W][3] radiance of the pass
    float* film;         // [H][W][4] rgb + weight, accumulated across passes
    unsigned long long* counters;  // optional [2]: node visits, triangle tests
};
hipError_t launch_megakernel(const SceneView& sv, const RenderLaunch& rl, hipStream_t stream);
// Bidirectional path tracing (bdpt.hip): the same launch, sample layout and film as the megakernel.
constexpr int kBdptMaxBounce = 8;
hipError_t launch_bdpt(const SceneView& sv, const RenderLaunch& rl, hipStream_t stream);
hipError_t launch_film_gather(const SceneView& sv, const RenderLaunch& rl, hipStream_t stream);
hipError_t launch_film_clear(float* film, int npix, hipStream_t stream);
hipError_t launch_trace_rays(const SceneView& sv, const float* rays, int n, int32_t* prim, float* t, hipStream_t stream);
// tris[k].mat = shade[k].mat for the n leaves of a loaded scene (shade is in leaf order).
hipError_t launch_tri_materials(TriLeaf* tris, const ShadeTri* shade, int n, hipStream_t stream);
hipError_t launch_get_render(const SceneView& sv, const float* film, float* rgb, hipStream_t stream);
// plain copy of `bytes` (a multiple of 16) with 16-byte vector loads and stores
hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t stream);

// ---------------------------------------------------------------- wavefront tracer (wavefront.hip)
constexpr int kWfAlive = 1, kWfSpecular = 2;
// per bounce b: [0] path queue, [1] shadow rays, [2] MIS rays, [3] ended paths awaiting their
// light sample; the 8 work counters of the trace
// launch of bounce b's path rays at kWfWorkBase + kWfWorkStride * j (one 64-byte line each)
constexpr int kWfCountsPerBounce = 160;
constexpr int kWfWorkBase = 32, kWfWorkStride = 16, kWfWorkSegments = 8;
static_assert(kWfWorkBase + kWfWorkStride * kWfWorkSegments <= kWfCountsPerBounce, "work counters of a bounce");
constexpr int kWfMaxMatBins = 64;
constexpr int kWfMaxBounce = 64;

// 32 bytes, so that a record read in material order touches one 64-byte segment; the scene and the
// training slot follow from the path index, the flags share a word with the sample number.
struct alignas(32) WfPath { float beta[3]; float L[3]; int32_t pixel; uint32_t sampleFlags; };
constexpr int kWfFlagBits = 2;
constexpr int kWfMaxSpp = 1 << (31 - kWfFlagBits);
struct alignas(16) WfRay { float o[3]; float d[3]; int32_t scene; int32_t pad; };
struct alignas(16) WfHit { int32_t leaf; float u, v, invDet; };
struct alignas(16) WfNee { float nb[3]; float lpdf; float p1[3]; float p2[3]; int32_t used; int32_t pad; };
struct alignas(16) WfShadow { float o[3]; float d[3]; int32_t path; int32_t scene; float c[3]; float pad2; };
struct alignas(16) WfMis { float o[3]; float d[3]; int32_t path; int32_t light; float f[3]; float weight; float pdf; int32_t scene; int32_t pad[2]; };

// Read by the kernels from device memory, so a captured pass graph can be replayed for every pass.
struct WfParams {
    int64_t nPaths;
    int32_t nScenes, maxBounce, rrStart;
    int32_t queryBounce;   // cached integrator: paths query the cache at their first non-specular hit
                           // of this bounce or later (a value above maxBounce disables queries)
    int32_t trainStride;   // training paths: paths offset + k stride of scene 0 (path index in the pass)
    int32_t trainOffset;
    int32_t trainSlots;    // 0: no training paths
    int32_t tileRows;      // path order within a sample: bands of this many rows, walked column by column
};
struct WfSceneSlot {
    int64_t pathBase;  // first path of this scene in the pass
    int32_t width, height, spp, stratified;
    uint32_t seed;
    int32_t s0;        // first sample index of the pass
    float* sampleL;    // [pass samples][H][W][3]
    float boxMin[3];   // scene bounds, for the cache's normalised positions
    float boxInv[3];   // 1 / extent per axis
};
// One vertex of a training path: radiance gathered before the vertex's own scattering, throughput
// and reflectance there; the suffix target is (L_final - Lb) / (beta a).
// slot: the path's sample index into its scene's sample buffer ((sample - s0) * W * H + pixel)
struct alignas(16) WfTrainRec { float Lb[3]; float beta[3]; float a[3]; int32_t slot; int32_t valid; int32_t pad; };

constexpr int kMaxTileRows = 64;

struct WfTuning {
    int traceBlock = 256;         // block size of the persistent traversal kernels (64, 128, 256)
    int persistentBlocks = 2048;  // grid of the traversal kernels
    int sweepBlocks = 2048;       // grid of the generate / shade / finalize sweeps
    int sortMaterials = 1;        // counting-sort the shade queue by material
    int useGraph = 1;             // capture the pass in a hipGraph and replay it
    int tileRows = 8;             // band height of the path order (a wave of 64 paths covers tileRows x 64/tileRows pixels)
    int passSpp = 0;              // samples per pass chosen by the autotuner (0: the automatic pass)
};

struct NrcQuery;
struct NrcBuffers;

struct WfBuffers {
    uint64_t gen = 0;        // allocation generation of the pass buffers
    uint64_t cacheGen = 0;   // allocation generation of the cached-integrator buffers
    size_t cap = 0;       // paths per pass
    int maxScenes = 0;
    int matBins = 2;
    SceneView* scenes = nullptr;
    WfParams* params = nullptr;
    WfSceneSlot* slots = nullptr;
    WfPath* paths = nullptr;
    WfRay* rays = nullptr;
    WfHit* hits = nullptr;
    WfNee* nee = nullptr;
    int32_t* queue[2] = {nullptr, nullptr};
    int32_t* sorted = nullptr;
    int32_t* done = nullptr;      // paths that ended with a light sample still in flight
    int32_t* shadowQ = nullptr;   // paths with a shadow ray (their records sit at the path index)
    int32_t* misQ = nullptr;      // paths with an MIS-BSDF ray (likewise)
    int32_t* bins = nullptr;      // material bin of each path's hit, written by the extend kernel
    WfShadow* shadows = nullptr;  // per path
    WfMis* mis = nullptr;         // per path
    int32_t* counts = nullptr;
    // cached integrator (allocated by wf_reserve_cache)
    size_t cacheCap = 0;          // paths per pass covered by the query arrays
    int trainCap = 0;             // training-path slots
    float* qRaw = nullptr;        // [cacheCap][16] query records, compact in append order
    NrcQuery* qMeta = nullptr;    // [cacheCap] the queries' throughput, reflectance, radiance and sample slot
    WfTrainRec* tRec = nullptr;   // [trainCap][kWfMaxBounce]
    float* tRaw = nullptr;        // [trainCap][kWfMaxBounce][16]
    int32_t* tCount = nullptr;    // per-block counts / offsets of the batch compaction
    float* sampleBase = nullptr;  // sample buffer of the pass (path i's radiance at 3 i)
};

// counts: per bounce kWfCountsPerBounce counters, then per-bounce material histograms, then the
// pass-wide counters below.
constexpr int kWfGlobalBase = (kWfCountsPerBounce + kWfMaxMatBins) * (kWfMaxBounce + 1);
constexpr int kWfGlobalCounts = 8;  // [0] cache queries

// Optional phase timing: events recorded between the kernels of a pass launched directly (not
// captured). ev[8 kWfMaxBounce + 3] precedes the pass's first trace; bounce b uses ev[8b + 0..4]:
// before the sort, after the sort, after shading, after the trace of the next path rays with this
// bounce's shadow and MIS rays, after finalize; the cached integrator's tail uses
// ev[8 kWfMaxBounce + 0..2]: before inference, after inference, after the training steps.
constexpr int kWfPhaseEvents = 8 * kWfMaxBounce + 4;
struct WfProfile { hipEvent_t ev[kWfPhaseEvents]; };
// Traversal counters of a context: render-node fetches and triangle tests of path, shadow and MIS
// rays (the megakernel uses the first two, for its path rays).
constexpr int kTraceCounters = 6;

hipError_t wf_reserve(WfBuffers& b, size_t cap, int maxScenes);
hipError_t wf_reserve_cache(WfBuffers& b, size_t cap, int trainCap);
void wf_release(WfBuffers& b);
// Enqueues one pass: generate, then per bounce extend / (sort) / shade / shadow / mis / finalize.
// With nrc (cached integrator): queries, training records, then cache inference into the sample
// buffer, the training batch of path-suffix radiance and one Adam step.
// traceCounts (6 counters, or null): launches the counting instantiation of the traversal kernel,
// which adds render-node fetches and triangle tests of path, shadow and MIS rays.
hipError_t wf_enqueue_pass(const WfBuffers& b, const WfTuning& t, int maxBounce, const NrcBuffers* nrc, hipStream_t stream,
                           const WfProfile* prof = nullptr, unsigned long long* traceCounts = nullptr);

// ---------------------------------------------------------------- learned radiance cache (nrc.hip)
// Network: 16-float raw record -> 64-dim encoding -> 5 hidden layers of 64 (ReLU) -> 3 outputs.
// Master weights are fp32, W_l[in][out] row-major: layers 0..4 at l*4096, the output layer
// (64 x 16, columns 3..15 held at zero) at 20480.
constexpr int kNrcRaw = 16;
constexpr int kNrcParams = 5 * 4096 + 64 * 16;  // 21504
constexpr int kNrcBwdElems = 6 * 4096;          // backward fragments (output layer padded)
constexpr int kNrcKSplit = 256;                 // samples per split of the weight-gradient GEMM
constexpr int kNrcBatchAlign = 1024;            // batch capacity multiple (4 splits per GEMM block)
constexpr int kNrcMaxSteps = 64;                // Adam steps per training call
constexpr int kNrcTrainBlocks = 256;

// A query of the cached pass: the path's throughput, the reflectance at the query vertex, the
// radiance it gathered before it (L) and its sample slot in sampleL; 48 bytes.
struct NrcQuery { float beta[3]; float a[3]; float L[3]; int32_t slot; int32_t pad[2]; };
struct NrcState {
    int32_t t;        // Adam steps taken
    int32_t nTotal;   // samples collected in trainRaw / trainTarget for the next training call
    float loss;       // mean loss of the last step (before its update)
    int32_t pad;
    float lr, beta1, beta2, eps;  // Adam hyper-parameters (device-resident: captured graphs read them)
    // per step k of a training call, planned once per call: batch window and bias corrections
    int32_t stepFirst[kNrcMaxSteps];
    int32_t stepN[kNrcMaxSteps];
    float stepC1[kNrcMaxSteps];   // 1 / (1 - beta1^t_k)
    float stepC2[kNrcMaxSteps];   // 1 / (1 - beta2^t_k)
};

struct NrcBuffers {
    uint64_t gen = 0;   // allocation generation (next_buffer_generation)
    int precision = 0;  // 0 bf16, 1 fp16
    int batchCap = 0;   // training samples per step (multiple of kNrcKSplit)
    int steps = 1;      // Adam steps per cached pass
    int chunks = 1;     // batches of collected samples trainRaw / trainTarget hold
    float lr = 1e-2f, beta1 = 0.9f, beta2 = 0.99f, eps = 1e-15f;
    float* master = nullptr;      // kNrcParams
    float* m = nullptr;           // Adam first moment
    float* v = nullptr;           // Adam second moment
    float* grad = nullptr;        // gradient of the last step (kNrcParams)
    uint16_t* fwd = nullptr;      // forward MFMA fragments (kNrcParams)
    uint16_t* bwd = nullptr;      // backward MFMA fragments (kNrcBwdElems)
    uint16_t* H = nullptr;        // layer inputs [6][64][batchCap]
    uint16_t* G = nullptr;        // layer output gradients [6][64][batchCap]
    float* partial = nullptr;     // [batchCap / (4 kNrcKSplit)][kNrcParams] weight-gradient partials
    float* lossPartial = nullptr; // kNrcTrainBlocks
    NrcState* state = nullptr;
    float* trainRaw = nullptr;    // [chunks * batchCap][16]
    float* trainTarget = nullptr; // [chunks * batchCap][3]
};

hipError_t nrc_reserve(NrcBuffers& b, int batchCap, int steps, int chunks);
void nrc_release(NrcBuffers& b);
// master -> MFMA fragments (after weights are set)
hipError_t nrc_pack(const NrcBuffers& b, hipStream_t stream);
// Evaluates min(*nDev, nCap) raw records; out receives 3 floats per record.
hipError_t nrc_infer(const NrcBuffers& b, const float* raw, const int32_t* nDev, int nCap, float* out, hipStream_t stream);
// The cached pass's queries (compact records raw / meta): evaluates them and writes each query's
// sample, sampleL[3 slot + c] = L_c + beta_c * max(y_c, 0) * a_c.
hipError_t nrc_infer_queries(const NrcBuffers& b, const float* raw, const NrcQuery* meta, const int32_t* nDev, int nCap,
                             float* sampleL, hipStream_t stream);
// `steps` Adam steps over the state->nTotal samples in trainRaw / trainTarget: step k trains on batch
// (k mod ceil(nTotal / batchCap)), so steps beyond the collected batches make further passes.
hipError_t nrc_train(const NrcBuffers& b, hipStream_t stream, int steps);

}  // namespace lts
