// C ABI of liblts.so (include/lts.h): context, scene upload, LBVH build and export, rendering.
#include "lts.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "image_io.h"
#include "lts_internal.h"
#include "scene_host.h"

using namespace lts;

namespace {

// Ken Perlin's permutation (improved noise reference implementation), as tabulated by the
// reference at materials/textures/Procedural.cuh:79-107; stored twice so index + 255 stays in range.
const int kPerm256[256] = {
    151, 160, 137, 91,  90,  15,  131, 13,  201, 95,  96,  53,  194, 233, 7,   225, 140, 36,  103, 30,  69,  142,
    8,   99,  37,  240, 21,  10,  23,  190, 6,   148, 247, 120, 234, 75,  0,   26,  197, 62,  94,  252, 219, 203,
    117, 35,  11,  32,  57,  177, 33,  88,  237, 149, 56,  87,  174, 20,  125, 136, 171, 168, 68,  175, 74,  165,
    71,  134, 139, 48,  27,  166, 77,  146, 158, 231, 83,  111, 229, 122, 60,  211, 133, 230, 220, 105, 92,  41,
    55,  46,  245, 40,  244, 102, 143, 54,  65,  25,  63,  161, 1,   216, 80,  73,  209, 76,  132, 187, 208, 89,
    18,  169, 200, 196, 135, 130, 116, 188, 159, 86,  164, 100, 109, 198, 173, 186, 3,   64,  52,  217, 226, 250,
    124, 123, 5,   202, 38,  147, 118, 126, 255, 82,  85,  212, 207, 206, 59,  227, 47,  16,  58,  17,  182, 189,
    28,  42,  223, 183, 170, 213, 119, 248, 152, 2,   44,  154, 163, 70,  221, 153, 101, 155, 167, 43,  172, 9,
    129, 22,  39,  253, 19,  98,  108, 110, 79,  113, 224, 232, 178, 185, 112, 104, 218, 246, 97,  228, 251, 34,
    242, 193, 238, 210, 144, 12,  191, 179, 162, 241, 81,  51,  145, 235, 249, 14,  239, 107, 49,  192, 214, 31,
    181, 199, 106, 157, 184, 84,  204, 176, 115, 121, 50,  45,  127, 4,   150, 254, 138, 236, 205, 93,  222, 114,
    67,  29,  24,  72,  243, 141, 128, 195, 78,  66,  215, 61,  156, 180};

constexpr int kMaxPrims = 1 << 26;  // keeps the Karras depth below 58, inside the traversal stack's bound

struct DeviceScene {
    bool loaded = false;
    HostScene host;
    float* soup = nullptr;
    ShadeTri* shade = nullptr;
    MatDesc* mats = nullptr;
    TexDesc* texs = nullptr;
    MipDesc* mips = nullptr;
    float* mipData = nullptr;
    LightDesc* lights = nullptr;
    float* lfunc = nullptr;
    float* lcdf = nullptr;
    BvhBuffers bvh;
    int treeKind = 0, treeDepth = 0, nodes4 = 0;  // render tree actually used
    SceneView view{};
    float boxMin[3] = {0, 0, 0}, boxInv[3] = {0, 0, 0};  // vertex bounds, for the cache's positions
};

}  // namespace

struct lts_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    BvhBuffers build;         // standalone builds (lts_bvh_build*)
    float* buildSoup = nullptr;
    size_t buildSoupCap = 0;
    int lastMode = 0;
    int renderTree = LTS_TREE_PLOC;  // tree of the next lts_load_scene
    int textureFilter = LTS_TEXTURE_REFERENCE;  // texture pipeline of the next lts_load_scene
    DeviceScene scene;
    // progressive film: the sampler parameters of the renders accumulated in `film`, and the
    // sample the next render must start at to continue it
    int32_t filmKey[9] = {};  // width, height, spp, seed, stratified, max_bounce, rr_start, integrator, scene load
    int32_t filmNext = 0;
    int32_t sceneLoads = 0;   // lts_load_scene calls that succeeded
    float* sampleL = nullptr;
    size_t sampleCap = 0;
    uint64_t sampleGen = 0;   // allocation generation of sampleL
    float* film = nullptr;
    float* rgb = nullptr;
    size_t filmCap = 0;
    unsigned long long* counters = nullptr;
    int* perm = nullptr;  // Perlin permutation (512 ints)
    hipEvent_t ev[8] = {};
    // wavefront integrator
    WfBuffers wf;
    hipGraphExec_t wfExec = nullptr;  // captured pass, valid for the wfKey* values below
    uint64_t wfKeyGen = 0, wfKeyCacheGen = 0, wfKeyNrcGen = 0, wfKeySampleGen = 0;
    int wfKeyBounce = -1, wfKeyBins = -1;
    WfTuning wfKeyTune;
    int wfKeyPrecision = -1;
    bool wfKeyProfile = false;
    WfProfile* prof = nullptr;  // phase events, created on first use
    WfParams wfParams{};                  // host copies of the per-pass parameters (stable for async H2D)
    std::vector<WfSceneSlot> wfSlots;
    std::map<std::string, WfTuning> tuned;  // autotuner results of this context
    // learned radiance cache
    NrcBuffers nrc;
    lts_nrc_config nrcCfg{};
    bool nrcReady = false;
    float* nrcIo = nullptr;     // staging for lts_nrc_infer: raw records then outputs
    size_t nrcIoCap = 0;        // records
    int32_t* nrcCount = nullptr;  // device record count for lts_nrc_infer
    uint64_t nrcPasses = 0;       // cached passes rendered (moves the training-pixel offset)
    float* xfBuf = nullptr;       // points and normals of a mesh being transformed
    size_t xfCap = 0;             // floats
};

namespace {

// Message of the last failed call made without a context, per host thread.
thread_local std::string g_noCtxErr;

int fail(lts_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    else g_noCtxErr = msg;
    return code;
}

#define HIPCHK(ctx, expr)                                                                              \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return fail(ctx, e_ == hipErrorOutOfMemory ? LTS_E_NOMEM : LTS_E_HIP,                     \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                            \
    } while (0)

template <typename T>
hipError_t upload(T** dst, const T* src, size_t count, hipStream_t s) {
    hipError_t e = hipMalloc((void**)dst, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess || count == 0) return e;
    e = hipMemcpyAsync(*dst, src, count * sizeof(T), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

void free_scene(DeviceScene& s) {
    void* ps[] = {s.soup, s.shade, s.mats, s.texs, s.mips, s.mipData, s.lights, s.lfunc, s.lcdf};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    bvh_release(s.bvh);
    s = DeviceScene();
}

float ms_between(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, a, b);
    return ms;
}

int copy_export(lts_ctx* c, const BvhBuffers& b, int mode, lts_bvh_export* out) {
    const int n = b.n;
    hipStream_t s = c->stream;
    if (out->codes_sorted) HIPCHK(c, hipMemcpyAsync(out->codes_sorted, b.keysA, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    if (out->perm) HIPCHK(c, hipMemcpyAsync(out->perm, b.valsA, n * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    if (n > 1) {
        if (out->left) HIPCHK(c, hipMemcpyAsync(out->left, b.left, (n - 1) * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        if (out->right) HIPCHK(c, hipMemcpyAsync(out->right, b.right, (n - 1) * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        if (out->int_bounds) HIPCHK(c, hipMemcpyAsync(out->int_bounds, b.box, 6 * (size_t)(n - 1) * sizeof(float), hipMemcpyDeviceToHost, s));
    }
    if (out->leaf_bounds) HIPCHK(c, hipMemcpyAsync(out->leaf_bounds, b.box + 6 * (size_t)(n - 1), 6 * (size_t)n * sizeof(float), hipMemcpyDeviceToHost, s));
    if (out->parent) HIPCHK(c, hipMemcpyAsync(out->parent, b.parent, (2 * (size_t)n - 1) * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    uint32_t keys[6];
    HIPCHK(c, hipMemcpyAsync(keys, b.sceneKeys, sizeof(keys), hipMemcpyDeviceToHost, s));
    if (out->leaf_prim) {
        if (mode == 0) {
            for (int k = 0; k < n; k++) out->leaf_prim[k] = k;
        } else {
            HIPCHK(c, hipMemcpyAsync(out->leaf_prim, b.valsA, n * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        }
    }
    HIPCHK(c, hipStreamSynchronize(s));
    for (int a = 0; a < 6; a++) {
        uint32_t k = keys[a];
        uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
        std::memcpy(&out->scene_bounds[a], &u, 4);
    }
    return LTS_OK;
}

int run_build(lts_ctx* c, BvhBuffers& b, const float* dsoup, int n, int mode, bool pack, lts_bvh_timing* t) {
    BvhStageEvents ev;
    for (int i = 0; i < 5; i++) ev.e[i] = c->ev[i];
    HIPCHK(c, bvh_enqueue(b, dsoup, n, mode, pack, c->stream, &ev));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int32_t status = 0;
    HIPCHK(c, hipMemcpy(&status, b.status, sizeof(status), hipMemcpyDeviceToHost));
    if (status != 0)
        return fail(c, LTS_E_STATE, "LBVH build failed (status " + std::to_string(status) +
                                        "): a bounded wait of the sort or the hierarchy climb did not end");
    if (t) {
        t->bounds_ms = ms_between(ev.e[0], ev.e[1]);
        t->sort_ms = ms_between(ev.e[1], ev.e[2]);
        t->tree_ms = ms_between(ev.e[2], ev.e[4]);
        t->total_ms = ms_between(ev.e[0], ev.e[4]);
    }
    return LTS_OK;
}

uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

int nrc_ensure(lts_ctx* c);

int check_render_params(lts_ctx* c, const lts_render_params* p) {
    if (p->spp < 1 || p->spp > kWfMaxSpp || p->max_bounce < 0 || p->spp_per_pass < 0) return fail(c, LTS_E_ARG, "bad render parameters");
    if (p->stratified) {
        int s = (int)std::floor(std::sqrt((float)p->spp));
        if (s * s != p->spp) return fail(c, LTS_E_ARG, "stratified film sampling needs a square sample count");
    }
    if (p->integrator == LTS_INTEGRATOR_WAVEFRONT || p->integrator == LTS_INTEGRATOR_CACHED) {
        if (p->max_bounce > kWfMaxBounce) return fail(c, LTS_E_ARG, "wavefront max_bounce above 64");
        if (p->trace_block != 0 && p->trace_block != 64 && p->trace_block != 128 && p->trace_block != 256)
            return fail(c, LTS_E_ARG, "trace_block must be 64, 128 or 256");
        if (p->persistent_blocks < 0 || p->persistent_blocks > (1 << 20)) return fail(c, LTS_E_ARG, "bad persistent_blocks");
        if (p->tile_rows < 0 || p->tile_rows > kMaxTileRows) return fail(c, LTS_E_ARG, "tile_rows must be 0 .. 64");
    } else if (p->integrator == LTS_INTEGRATOR_BDPT) {
        if (p->max_bounce > kBdptMaxBounce) return fail(c, LTS_E_ARG, "bidirectional max_bounce above 8");
        if (p->count_traversal) return fail(c, LTS_E_ARG, "count_traversal is a megakernel option");
    } else if (p->integrator != LTS_INTEGRATOR_MEGAKERNEL) {
        return fail(c, LTS_E_ARG, "unsupported integrator");
    }
    if (p->count_traversal != 0 && p->count_traversal != 1) return fail(c, LTS_E_ARG, "count_traversal is 0 or 1");
    const int end = p->sample_end ? p->sample_end : p->spp;
    if (p->sample_begin < 0 || end <= p->sample_begin || end > p->spp)
        return fail(c, LTS_E_ARG, "sample range outside [0, spp)");
    return LTS_OK;
}

// Samples [*b0, *b1) of the render.
void sample_range(const lts_render_params* p, int* b0, int* b1) {
    *b0 = p->sample_begin;
    *b1 = p->sample_end ? p->sample_end : p->spp;
}

void film_key(const lts_ctx* c, const lts_render_params* p, int32_t* k) {
    const CameraDesc& cam = c->scene.view.cam;
    const int32_t v[9] = {cam.width, cam.height, p->spp, (int32_t)p->seed, p->stratified, p->max_bounce, p->rr_start,
                          p->integrator, c->sceneLoads};
    std::memcpy(k, v, sizeof v);
}

// A render of [b0, ...) with b0 > 0 must continue the film this context holds.
int film_continues(lts_ctx* c, const lts_render_params* p, int b0) {
    if (b0 == 0) return LTS_OK;
    int32_t k[9];
    film_key(c, p, k);
    if (c->filmNext != b0 || std::memcmp(k, c->filmKey, sizeof k) != 0)
        return fail(c, LTS_E_STATE, "progressive render does not continue the film (sample_begin or parameters differ)");
    return LTS_OK;
}

void film_advance(lts_ctx* c, const lts_render_params* p, int b1) {
    film_key(c, p, c->filmKey);
    c->filmNext = b1;
}

int ensure_film(lts_ctx* c, size_t npix) {
    if (npix <= c->filmCap) return LTS_OK;
    if (c->film) (void)hipFree(c->film);
    if (c->rgb) (void)hipFree(c->rgb);
    c->film = c->rgb = nullptr;
    c->filmCap = 0;
    HIPCHK(c, hipMalloc((void**)&c->film, npix * 4 * sizeof(float)));
    HIPCHK(c, hipMalloc((void**)&c->rgb, npix * 3 * sizeof(float)));
    c->filmCap = npix;
    return LTS_OK;
}

int ensure_samples(lts_ctx* c, size_t floats) {
    if (floats <= c->sampleCap) return LTS_OK;
    if (c->sampleL) (void)hipFree(c->sampleL);
    c->sampleL = nullptr;
    c->sampleCap = 0;
    HIPCHK(c, hipMalloc((void**)&c->sampleL, floats * sizeof(float)));
    c->sampleCap = floats;
    c->sampleGen = next_buffer_generation();
    return LTS_OK;
}

bool same_tuning(const WfTuning& a, const WfTuning& b) {
    return a.traceBlock == b.traceBlock && a.persistentBlocks == b.persistentBlocks && a.sweepBlocks == b.sweepBlocks &&
           a.sortMaterials == b.sortMaterials && a.useGraph == b.useGraph;
}

// Runs the pass whose parameters are resident in c->wf: replays the captured graph, capturing it
// first when the buffers, bounce count or tuning changed since the last capture.
int wf_run(lts_ctx* c, const WfTuning& t, int maxBounce, const NrcBuffers* nrc, const WfProfile* prof = nullptr,
           unsigned long long* traceCounts = nullptr) {
    hipStream_t s = c->stream;
    if (!t.useGraph || traceCounts) {
        HIPCHK(c, wf_enqueue_pass(c->wf, t, maxBounce, nrc, s, prof, traceCounts));
        return LTS_OK;
    }
    const uint64_t nrcGen = nrc ? nrc->gen : 0, cacheGen = nrc ? c->wf.cacheGen : 0;
    const int precision = nrc ? nrc->precision + 2 * nrc->steps : -1;  // kernels and step count of the pass
    if (!c->wfExec || c->wfKeyGen != c->wf.gen || c->wfKeyCacheGen != cacheGen || c->wfKeyNrcGen != nrcGen ||
        c->wfKeySampleGen != c->sampleGen || c->wfKeyBounce != maxBounce || c->wfKeyBins != c->wf.matBins ||
        !same_tuning(c->wfKeyTune, t) || c->wfKeyPrecision != precision || c->wfKeyProfile != (prof != nullptr)) {
        if (c->wfExec) (void)hipGraphExecDestroy(c->wfExec);
        c->wfExec = nullptr;
        hipGraph_t g = nullptr;
        HIPCHK(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        hipError_t e = wf_enqueue_pass(c->wf, t, maxBounce, nrc, s, prof);
        hipError_t e2 = hipStreamEndCapture(s, &g);
        if (e == hipSuccess) e = e2;
        if (e == hipSuccess) e = hipGraphInstantiate(&c->wfExec, g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        HIPCHK(c, e);
        c->wfKeyGen = c->wf.gen;
        c->wfKeyCacheGen = cacheGen;
        c->wfKeyNrcGen = nrcGen;
        c->wfKeySampleGen = c->sampleGen;
        c->wfKeyBounce = maxBounce;
        c->wfKeyBins = c->wf.matBins;
        c->wfKeyTune = t;
        c->wfKeyPrecision = precision;
        c->wfKeyProfile = prof != nullptr;
    }
    HIPCHK(c, hipGraphLaunch(c->wfExec, s));
    return LTS_OK;
}

// Radiance-cache settings of one pass (queries off and no training paths for plain passes).
struct PassCache {
    int queryBounce = 1 << 30;
    int trainStride = 1, trainOffset = 0, trainSlots = 0;
};

// Writes the pass parameters for samples [s0, s1) of every scene to the device.
int wf_upload_pass(lts_ctx* c, const std::vector<lts_ctx*>& scenes, const lts_render_params* p, int s0, int s1,
                   int64_t* paths, int tileRows, const PassCache& pc = PassCache()) {
    const int n = (int)scenes.size();
    c->wfSlots.resize(n);
    int64_t base = 0;
    for (int i = 0; i < n; i++) {
        const CameraDesc& cam = scenes[i]->scene.view.cam;
        WfSceneSlot& sl = c->wfSlots[i];
        sl.pathBase = base;
        sl.width = cam.width; sl.height = cam.height;
        sl.spp = p->spp; sl.stratified = p->stratified; sl.seed = p->seed; sl.s0 = s0;
        sl.sampleL = c->sampleL + 3 * base;
        for (int a = 0; a < 3; a++) {
            sl.boxMin[a] = scenes[i]->scene.boxMin[a];
            sl.boxInv[a] = scenes[i]->scene.boxInv[a];
        }
        base += (int64_t)cam.width * cam.height * (s1 - s0);
    }
    c->wfParams.nPaths = base;
    c->wfParams.nScenes = n;
    c->wfParams.maxBounce = p->max_bounce;
    c->wfParams.rrStart = p->rr_start;
    c->wfParams.queryBounce = pc.queryBounce;
    c->wfParams.trainStride = pc.trainStride;
    c->wfParams.trainOffset = pc.trainOffset;
    c->wfParams.trainSlots = pc.trainSlots;
    c->wfParams.tileRows = tileRows;
    HIPCHK(c, hipMemcpyAsync(c->wf.params, &c->wfParams, sizeof(WfParams), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->wf.slots, c->wfSlots.data(), n * sizeof(WfSceneSlot), hipMemcpyHostToDevice, c->stream));
    *paths = base;
    return LTS_OK;
}

std::string tune_key(const std::vector<lts_ctx*>& scenes, const lts_render_params* p, int pass) {
    // the automatic pass and the spp bound the pass sizes the tuner may choose
    std::ostringstream k;
    k << "gfx950";
    for (lts_ctx* x : scenes) k << ":" << x->scene.view.nPrim << "/" << x->scene.view.cam.width << "x" << x->scene.view.cam.height;
    k << ":pass" << pass << ":spp" << p->spp << ":b" << p->max_bounce << ":rr" << p->rr_start;
    return k.str();
}

bool tune_cache_read(const std::string& key, WfTuning* t) {
    const char* path = std::getenv("LTS_TUNE_CACHE");
    if (!path || !*path) return false;
    std::ifstream in(path);
    std::string line;
    bool found = false;
    while (std::getline(in, line)) {
        std::istringstream f(line);
        std::string k;
        int tb = 0, pb = 0, so = 0, rows = 8, pass = 0;
        if (!(f >> k >> tb >> pb >> so)) continue;
        f >> rows >> pass;  // absent in entries of the older format: the defaults
        if (k == key && (tb == 64 || tb == 128 || tb == 256) && pb > 0 && rows >= 1 && rows <= kMaxTileRows && pass >= 0) {
            t->traceBlock = tb; t->persistentBlocks = pb; t->sortMaterials = so != 0;
            t->tileRows = rows; t->passSpp = pass;
            found = true;  // the last entry for a key wins
        }
    }
    return found;
}

void tune_cache_append(const std::string& key, const WfTuning& t) {
    const char* path = std::getenv("LTS_TUNE_CACHE");
    if (!path || !*path) return;
    std::ofstream out(path, std::ios::app);
    out << key << " " << t.traceBlock << " " << t.persistentBlocks << " " << t.sortMaterials << " " << t.tileRows << " "
        << t.passSpp << "\n";
}

// Mean device time of one pass with tuning t (one untimed warm-up, then `reps` passes).
int time_pass(lts_ctx* c, const WfTuning& t, int maxBounce, int reps, double* ms) {
    WfTuning u = t;
    u.useGraph = 0;
    int rc = wf_run(c, u, maxBounce, nullptr);
    if (rc != LTS_OK) return rc;
    HIPCHK(c, hipEventRecord(c->ev[5], c->stream));
    for (int r = 0; r < reps; r++)
        if ((rc = wf_run(c, u, maxBounce, nullptr)) != LTS_OK) return rc;
    HIPCHK(c, hipEventRecord(c->ev[6], c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev[6]));
    *ms = ms_between(c->ev[5], c->ev[6]) / reps;
    return LTS_OK;
}

// Coordinate search over the traversal block size, the persistent grid, material sorting and the
// band height of the path order, on one pass of the render's own size (the parameters of that pass
// are uploaded here): the end of every trace launch is a tail whose share depends on the pass size,
// so smaller trial passes pick other settings than the render needs. Then the pass size itself:
// half the pass against the whole, compared by time per sample (a smaller pass trades longer tails
// for less path state in the caches).
int autotune(lts_ctx* c, const std::vector<lts_ctx*>& scenes, const lts_render_params* p, int pass, WfTuning* best) {
    const int trialSamples = std::max(1, std::min(p->spp, pass));
    int64_t paths = 0;
    WfTuning cur = *best;
    int rc = wf_upload_pass(c, scenes, p, 0, trialSamples, &paths, cur.tileRows);
    if (rc != LTS_OK) return rc;
    double bestMs = 0;
    if ((rc = time_pass(c, cur, p->max_bounce, 1, &bestMs)) != LTS_OK) return rc;
    auto consider = [&](WfTuning t) -> int {
        double ms = 0;
        int r = time_pass(c, t, p->max_bounce, 1, &ms);
        if (r != LTS_OK) return r;
        if (ms < bestMs) { bestMs = ms; cur = t; }
        return LTS_OK;
    };
    for (int tb : {64, 128, 256}) {
        if (tb == cur.traceBlock) continue;
        WfTuning t = cur; t.traceBlock = tb;
        if ((rc = consider(t)) != LTS_OK) return rc;
    }
    const WfTuning afterBlock = cur;
    for (int pb : {512, 1024, 2048, 4096, 8192}) {
        if (pb == afterBlock.persistentBlocks) continue;
        WfTuning t = afterBlock; t.persistentBlocks = pb;
        if ((rc = consider(t)) != LTS_OK) return rc;
    }
    {
        WfTuning t = cur; t.sortMaterials = !cur.sortMaterials;
        if ((rc = consider(t)) != LTS_OK) return rc;
    }
    const WfTuning afterSort = cur;
    for (int rows : {1, 4, 8, 16}) {
        if (rows == afterSort.tileRows) continue;
        WfTuning t = afterSort; t.tileRows = rows;
        if ((rc = wf_upload_pass(c, scenes, p, 0, trialSamples, &paths, rows)) != LTS_OK) return rc;
        if ((rc = consider(t)) != LTS_OK) return rc;
    }
    // The search compares single passes, so a choice can win by noise: time it against the starting
    // tuning again, two passes each, and keep the starting tuning unless the choice is faster.
    const WfTuning start = *best;
    if (cur.traceBlock != start.traceBlock || cur.persistentBlocks != start.persistentBlocks ||
        cur.sortMaterials != start.sortMaterials || cur.tileRows != start.tileRows) {
        double msStart = 0, msCur = 0;
        if ((rc = wf_upload_pass(c, scenes, p, 0, trialSamples, &paths, start.tileRows)) != LTS_OK) return rc;
        if ((rc = time_pass(c, start, p->max_bounce, 2, &msStart)) != LTS_OK) return rc;
        if ((rc = wf_upload_pass(c, scenes, p, 0, trialSamples, &paths, cur.tileRows)) != LTS_OK) return rc;
        if ((rc = time_pass(c, cur, p->max_bounce, 2, &msCur)) != LTS_OK) return rc;
        if (msCur < msStart) {
            bestMs = msCur;
        } else {
            cur = start;
            bestMs = msStart;
        }
    }
    cur.passSpp = trialSamples;
    if (trialSamples >= 2 && p->spp > trialSamples / 2) {
        // half the pass: the time of a pass of half the samples, per sample, against the whole one's
        const int half = trialSamples / 2;
        if ((rc = wf_upload_pass(c, scenes, p, 0, half, &paths, cur.tileRows)) != LTS_OK) return rc;
        double ms = 0;
        if ((rc = time_pass(c, cur, p->max_bounce, 2, &ms)) != LTS_OK) return rc;
        if (ms / half < bestMs / trialSamples) cur.passSpp = half;
    }
    *best = cur;
    return LTS_OK;
}

// The wavefront integrator over one or more scenes, executed on c's stream and buffers.
int render_wavefront(lts_ctx* c, const std::vector<lts_ctx*>& scenes, const lts_render_params* p, float* const* rgb_out,
                     float* sample_rgb, lts_stats* st) {
    hipStream_t s = c->stream;
    size_t totalPix = 0;
    int maxMats = 1;
    for (lts_ctx* x : scenes) {
        totalPix += (size_t)x->scene.view.cam.width * x->scene.view.cam.height;
        maxMats = std::max(maxMats, (int)x->scene.host.mats.size());
    }
    const bool cached = p->integrator == LTS_INTEGRATOR_CACHED;
    if (cached && scenes.size() != 1) return fail(c, LTS_E_ARG, "the cached integrator renders one scene");
    int rc = LTS_OK;
    if (cached && (rc = nrc_ensure(c)) != LTS_OK) return rc;
    int b0, b1;
    sample_range(p, &b0, &b1);
    if (b0 > 0 && scenes.size() != 1) return fail(c, LTS_E_ARG, "progressive rendering is for one scene");
    if (b0 > 0 && (rc = film_continues(c, p, b0)) != LTS_OK) return rc;
    // automatic pass: 2^27 paths (about 36 GB of path state at 268 B per path). Each pass runs its
    // bounces as separate launches whose late ones hold few paths, so fewer, larger passes keep the
    // GPU busy: 2^24 -> 2^26 paths took the 1080p 64-spp frame from 585 to 756 Msamples/s, and
    // 2^26 -> 2^27 (the whole frame in one pass) from 1392 to 1524. The cached integrator takes the
    // same passes: a 64-spp 1080p frame is one pass that queries the cache and trains it once, and
    // the cache is carried from frame to frame (its warm-up passes are the first frames after
    // lts_nrc_configure). Passes of 2^25 paths, which let one fresh frame warm up and then query,
    // cost the cached frame 112 ms against 81 ms in one pass.
    const size_t passPaths = (size_t)1 << 27;
    int pass = p->spp_per_pass > 0 ? p->spp_per_pass : (int)std::max<size_t>(1, passPaths / totalPix);
    if (sample_rgb) pass = b1 - b0;
    pass = std::min(pass, b1 - b0);
    const size_t cap = totalPix * (size_t)pass;
    if (cap > (size_t)INT32_MAX) return fail(c, LTS_E_ARG, "pass holds more than 2^31 paths");
    const bool realloc = cap > c->wf.cap || (int)scenes.size() > c->wf.maxScenes;
    HIPCHK(c, wf_reserve(c->wf, cap, (int)scenes.size()));
    if (realloc && c->wfExec) {
        (void)hipGraphExecDestroy(c->wfExec);
        c->wfExec = nullptr;
    }
    c->wf.matBins = std::min(maxMats, kWfMaxMatBins - 1) + 1;
    if ((rc = ensure_samples(c, cap * 3)) != LTS_OK) return rc;
    c->wf.sampleBase = c->sampleL;
    // training paths stay a minority of the pass (at most 1/8 of its paths) so most paths query
    const int trainPaths = cached ? (int)std::max<size_t>(1, std::min<size_t>((size_t)c->nrcCfg.train_paths, cap / 8)) : 0;
    if (cached) HIPCHK(c, wf_reserve_cache(c->wf, cap, trainPaths));
    const NrcBuffers* nrc = cached ? &c->nrc : nullptr;
    std::vector<SceneView> views;
    for (lts_ctx* x : scenes) {
        views.push_back(x->scene.view);
        if ((rc = ensure_film(x, (size_t)x->scene.view.cam.width * x->scene.view.cam.height)) != LTS_OK)
            return fail(c, rc, x->err);
    }
    HIPCHK(c, hipMemcpyAsync(c->wf.scenes, views.data(), views.size() * sizeof(SceneView), hipMemcpyHostToDevice, s));
    if (b0 == 0)
        for (lts_ctx* x : scenes)
            HIPCHK(c, launch_film_clear(x->film, x->scene.view.cam.width * x->scene.view.cam.height, s));

    WfTuning t;
    double tuneMs = 0;
    if (p->autotune && !cached) {
        const std::string key = tune_key(scenes, p, pass);
        auto it = c->tuned.find(key);
        if (it != c->tuned.end()) {
            t = it->second;
        } else if (tune_cache_read(key, &t)) {
            c->tuned[key] = t;
        } else {
            HIPCHK(c, hipEventRecord(c->ev[3], s));
            if ((rc = autotune(c, scenes, p, pass, &t)) != LTS_OK) return rc;
            HIPCHK(c, hipEventRecord(c->ev[4], s));
            HIPCHK(c, hipEventSynchronize(c->ev[4]));
            tuneMs = ms_between(c->ev[3], c->ev[4]);
            c->tuned[key] = t;
            tune_cache_append(key, t);
        }
    }
    if (p->trace_block) t.traceBlock = p->trace_block;
    if (p->persistent_blocks) t.persistentBlocks = p->persistent_blocks;
    if (p->sort_materials >= 0) t.sortMaterials = p->sort_materials != 0;
    if (p->use_graph >= 0) t.useGraph = p->use_graph != 0;
    if (p->tile_rows) t.tileRows = p->tile_rows;
    if (p->autotune && !cached && p->spp_per_pass == 0 && !sample_rgb && t.passSpp > 0) pass = std::min(pass, t.passSpp);
    if (p->profile_phases) t.useGraph = 0;  // phase events are recorded between direct launches
    // counting runs launch the counting instantiation of the traversal kernel directly
    unsigned long long* traceCounts = p->count_traversal ? c->counters : nullptr;
    if (traceCounts) {
        t.useGraph = 0;
        HIPCHK(c, hipMemsetAsync(traceCounts, 0, kTraceCounters * sizeof(unsigned long long), s));
    }

    double renderMs = 0, resolveMs = 0;
    int64_t launches = 0;
    const size_t npix0 = (size_t)views[0].cam.width * views[0].cam.height;
    double phase[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (p->profile_phases && !c->prof) {
        c->prof = new WfProfile();
        for (hipEvent_t& e : c->prof->ev) e = nullptr;
        for (hipEvent_t& e : c->prof->ev) HIPCHK(c, hipEventCreate(&e));
    }
    const WfProfile* prof = p->profile_phases ? c->prof : nullptr;
    for (int s0 = b0; s0 < b1; s0 += pass) {
        const int s1 = std::min(b1, s0 + pass);
        PassCache pc;
        if (cached) {
            // training paths: every stride-th (sample, pixel) of scene 0 in the pass, in sample-major
            // row-major order, from an offset that moves from pass to pass
            const size_t paths0 = npix0 * (size_t)(s1 - s0);
            pc.queryBounce = c->nrcPasses < (uint64_t)c->nrcCfg.warmup_passes ? (1 << 30) : c->nrcCfg.query_bounce;
            pc.trainStride = (int)((paths0 + trainPaths - 1) / trainPaths);
            uint64_t h = 0x70617373ull + (uint64_t)c->nrcPasses * 0x9e3779b97f4a7c15ull;
            pc.trainOffset = (int)(splitmix64(h) % (uint64_t)pc.trainStride);
            pc.trainSlots = (int)std::min<size_t>((size_t)trainPaths,
                                                  (paths0 - pc.trainOffset + pc.trainStride - 1) / pc.trainStride);
            c->nrcPasses++;
        }
        int64_t paths = 0;
        if ((rc = wf_upload_pass(c, scenes, p, s0, s1, &paths, t.tileRows, pc)) != LTS_OK) return rc;
        HIPCHK(c, hipEventRecord(c->ev[5], s));
        if ((rc = wf_run(c, t, p->max_bounce, nrc, prof, traceCounts)) != LTS_OK) return rc;
        HIPCHK(c, hipEventRecord(c->ev[6], s));
        for (size_t i = 0; i < scenes.size(); i++) {
            RenderLaunch rl{};
            rl.spp = p->spp; rl.maxBounce = p->max_bounce; rl.rrStart = p->rr_start; rl.seed = p->seed;
            rl.stratified = p->stratified; rl.s0 = s0; rl.s1 = s1;
            rl.sampleL = c->wfSlots[i].sampleL; rl.film = scenes[i]->film;
            HIPCHK(c, launch_film_gather(views[i], rl, s));
        }
        HIPCHK(c, hipEventRecord(c->ev[7], s));
        HIPCHK(c, hipEventSynchronize(c->ev[7]));
        renderMs += ms_between(c->ev[5], c->ev[6]);
        resolveMs += ms_between(c->ev[6], c->ev[7]);
        launches++;
        if (prof && p->max_bounce > 0) {
            // the pass's first trace, then per bounce: sort, shade, trace, finalize
            phase[LTS_PHASE_EXTEND] += ms_between(prof->ev[8 * kWfMaxBounce + 3], prof->ev[0]);
            for (int b = 0; b < p->max_bounce; b++) {
                const hipEvent_t* e = prof->ev + 8 * b;
                phase[LTS_PHASE_SORT] += ms_between(e[0], e[1]);
                phase[LTS_PHASE_SHADE] += ms_between(e[1], e[2]);
                phase[LTS_PHASE_EXTEND] += ms_between(e[2], e[3]);
                phase[LTS_PHASE_FINALIZE] += ms_between(e[3], e[4]);
            }
            if (cached) {
                const hipEvent_t* e = prof->ev + 8 * kWfMaxBounce;
                phase[LTS_PHASE_CACHE_INFER] += ms_between(e[0], e[1]);
                phase[LTS_PHASE_CACHE_TRAIN] += ms_between(e[1], e[2]);
            }
        }
        if (sample_rgb)
            HIPCHK(c, hipMemcpy(sample_rgb, c->sampleL, (size_t)paths * 3 * sizeof(float), hipMemcpyDeviceToHost));
    }
    for (size_t i = 0; i < scenes.size(); i++) {
        lts_ctx* x = scenes[i];
        HIPCHK(c, launch_get_render(views[i], x->film, x->rgb, s));
        film_advance(x, p, b1);
        if (rgb_out && rgb_out[i])
            HIPCHK(c, hipMemcpyAsync(rgb_out[i], x->rgb, (size_t)views[i].cam.width * views[i].cam.height * 3 * sizeof(float),
                                     hipMemcpyDeviceToHost, s));
    }
    HIPCHK(c, hipStreamSynchronize(s));
    if (st) {
        std::memset(st, 0, sizeof(*st));
        st->render_ms = renderMs;
        st->resolve_ms = resolveMs;
        st->launches = launches;
        st->kernel_avg_ms = launches ? renderMs / launches : 0;
        st->samples = (int64_t)totalPix * (b1 - b0);
        st->msamples_per_s = (renderMs + resolveMs) > 0 ? st->samples / ((renderMs + resolveMs) * 1e3) : 0;
        st->trace_block = t.traceBlock;
        st->persistent_blocks = t.persistentBlocks;
        st->sort_materials = t.sortMaterials;
        st->use_graph = t.useGraph;
        st->tile_rows = t.tileRows;
        st->spp_per_pass = pass;
        st->autotune_ms = tuneMs;
        for (int k = 0; k < 8; k++) st->phase_ms[k] = phase[k];
        if (prof && p->max_bounce > 0) {
            const int64_t perBounce = launches * p->max_bounce;
            st->phase_launches[LTS_PHASE_EXTEND] = perBounce + launches;
            st->phase_launches[LTS_PHASE_SORT] = t.sortMaterials ? 2 * perBounce : 0;
            st->phase_launches[LTS_PHASE_SHADE] = perBounce;
            st->phase_launches[LTS_PHASE_SHADOW] = 0;
            st->phase_launches[LTS_PHASE_MIS] = 0;
            st->phase_launches[LTS_PHASE_FINALIZE] = perBounce;
            st->phase_launches[LTS_PHASE_CACHE_INFER] = cached ? launches : 0;
            st->phase_launches[LTS_PHASE_CACHE_TRAIN] = cached ? launches * c->nrc.steps : 0;
        }
        if (traceCounts) {
            unsigned long long cnt[kTraceCounters];
            HIPCHK(c, hipMemcpy(cnt, traceCounts, sizeof cnt, hipMemcpyDeviceToHost));
            for (int k = 0; k < 3; k++) {
                st->ray_nodes[k] = (int64_t)cnt[2 * k];
                st->ray_tris[k] = (int64_t)cnt[2 * k + 1];
            }
            st->node_visits = st->ray_nodes[0];
            st->tri_tests = st->ray_tris[0];
        }
        if (cached) {
            NrcState ns;
            int32_t q = 0;
            HIPCHK(c, hipMemcpy(&ns, c->nrc.state, sizeof(ns), hipMemcpyDeviceToHost));
            HIPCHK(c, hipMemcpy(&q, c->wf.counts + kWfGlobalBase, sizeof(q), hipMemcpyDeviceToHost));
            st->nrc_loss = ns.loss;
            st->nrc_queries = q;
            st->nrc_train_samples = ns.nTotal;
            st->nrc_steps = ns.t;
        }
    }
    return LTS_OK;
}

// ---------------------------------------------------------------- learned radiance cache
lts_nrc_config nrc_defaults(const lts_nrc_config* in) {
    lts_nrc_config c{};
    if (in) c = *in;
    if (c.batch <= 0) c.batch = 65536;
    if (c.query_bounce <= 0) c.query_bounce = 2;
    if (c.train_paths <= 0) c.train_paths = 262144;
    if (c.train_steps <= 0) c.train_steps = 16;
    if (c.warmup_passes == 0) c.warmup_passes = 4;
    if (c.warmup_passes < 0) c.warmup_passes = 0;
    if (c.lr <= 0) c.lr = 1e-2f;
    if (c.beta1 <= 0) c.beta1 = 0.9f;
    if (c.beta2 <= 0) c.beta2 = 0.99f;
    if (c.eps <= 0) c.eps = 1e-15f;
    return c;
}

// Uploads fp32 master weights (padding columns forced to zero), clears Adam and repacks.
int nrc_upload_weights(lts_ctx* c, const float* w) {
    std::vector<float> m(w, w + kNrcParams);
    for (int in = 0; in < 64; in++)
        for (int out = 3; out < 16; out++) m[20480 + in * 16 + out] = 0.0f;
    hipStream_t s = c->stream;
    HIPCHK(c, hipMemcpyAsync(c->nrc.master, m.data(), kNrcParams * sizeof(float), hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemsetAsync(c->nrc.m, 0, kNrcParams * sizeof(float), s));
    HIPCHK(c, hipMemsetAsync(c->nrc.v, 0, kNrcParams * sizeof(float), s));
    HIPCHK(c, hipMemsetAsync(c->nrc.grad, 0, kNrcParams * sizeof(float), s));
    NrcState st{};
    st.lr = c->nrc.lr; st.beta1 = c->nrc.beta1; st.beta2 = c->nrc.beta2; st.eps = c->nrc.eps;
    HIPCHK(c, hipMemcpyAsync(c->nrc.state, &st, sizeof(NrcState), hipMemcpyHostToDevice, s));
    HIPCHK(c, nrc_pack(c->nrc, s));
    HIPCHK(c, hipStreamSynchronize(s));
    return LTS_OK;
}

int nrc_setup(lts_ctx* c, const lts_nrc_config* in) {
    lts_nrc_config cfg = nrc_defaults(in);
    if (cfg.precision != 0 && cfg.precision != 1) return fail(c, LTS_E_ARG, "nrc precision must be 0 (bf16) or 1 (fp16)");
    if (cfg.batch > (1 << 20)) return fail(c, LTS_E_ARG, "nrc batch above 2^20");
    if (cfg.train_paths > (1 << 22)) return fail(c, LTS_E_ARG, "nrc train_paths above 2^22");
    if (cfg.train_steps > 64) return fail(c, LTS_E_ARG, "nrc train_steps above 64");
    HIPCHK(c, hipSetDevice(c->device));
    c->nrc.precision = cfg.precision;
    c->nrc.lr = cfg.lr; c->nrc.beta1 = cfg.beta1; c->nrc.beta2 = cfg.beta2; c->nrc.eps = cfg.eps;
    // training paths yield fewer samples than paths (escaping paths yield none): hold up to
    // train_paths samples, in batches
    const int chunks = std::max(1, (cfg.train_paths + cfg.batch - 1) / cfg.batch);
    HIPCHK(c, nrc_reserve(c->nrc, cfg.batch, cfg.train_steps, chunks));
    if (!c->nrcCount) HIPCHK(c, hipMalloc((void**)&c->nrcCount, sizeof(int32_t)));
    // Xavier-uniform initialisation
    std::vector<float> w(kNrcParams, 0.0f);
    uint64_t st = 0x6e72635f696e6974ull ^ cfg.seed;
    auto uni = [&](float a) { return (float)(((splitmix64(st) >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0) * a; };
    const float hid = std::sqrt(6.0f / 128.0f), outA = std::sqrt(6.0f / 67.0f);
    for (int p = 0; p < 20480; p++) w[p] = uni(hid);
    for (int in = 0; in < 64; in++)
        for (int out = 0; out < 3; out++) w[20480 + in * 16 + out] = uni(outA);
    c->nrcCfg = cfg;
    c->nrcPasses = 0;
    int rc = nrc_upload_weights(c, w.data());
    if (rc != LTS_OK) return rc;
    c->nrcReady = true;
    return LTS_OK;
}

int nrc_ensure(lts_ctx* c) { return c->nrcReady ? LTS_OK : nrc_setup(c, nullptr); }

// Runs an entry point's body; a failed host allocation returns LTS_E_NOMEM and any other C++
// exception LTS_E_IO, with the message kept in the context.
template <class F>
int guarded(lts_ctx* c, F&& body) {
    try {
        return body();
    } catch (const std::bad_alloc&) {
        return fail(c, LTS_E_NOMEM, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(c, LTS_E_IO, std::string("unexpected error: ") + e.what());
    }
}

}  // namespace

extern "C" {

const char* lts_version(void) { return "lts-mi355x 0.1 (gfx950)"; }

int lts_create(lts_ctx** out, int dev) {
    if (!out) return LTS_E_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || dev < 0 || dev >= count) return LTS_E_HIP;
    lts_ctx* c = new lts_ctx();
    c->device = dev;
    if (hipSetDevice(dev) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return LTS_E_HIP;
    }
    for (hipEvent_t& e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) { delete c; return LTS_E_HIP; }
    int table[512];
    for (int i = 0; i < 512; i++) table[i] = kPerm256[i & 255];
    if (hipMalloc((void**)&c->perm, sizeof(table)) != hipSuccess ||
        hipMemcpy(c->perm, table, sizeof(table), hipMemcpyHostToDevice) != hipSuccess) { lts_destroy(c); return LTS_E_HIP; }
    if (hipMalloc((void**)&c->counters, kTraceCounters * sizeof(unsigned long long)) != hipSuccess) { lts_destroy(c); return LTS_E_HIP; }
    *out = c;
    return LTS_OK;
}

void lts_destroy(lts_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    free_scene(c->scene);
    bvh_release(c->build);
    void* ps[] = {c->buildSoup, c->sampleL, c->film, c->rgb, c->counters, c->perm, c->xfBuf};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    for (hipEvent_t e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->wfExec) (void)hipGraphExecDestroy(c->wfExec);
    if (c->prof) {
        for (hipEvent_t e : c->prof->ev)
            if (e) (void)hipEventDestroy(e);
        delete c->prof;
    }
    wf_release(c->wf);
    nrc_release(c->nrc);
    if (c->nrcIo) (void)hipFree(c->nrcIo);
    if (c->nrcCount) (void)hipFree(c->nrcCount);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* lts_last_error(const lts_ctx* c) { return c ? c->err.c_str() : g_noCtxErr.c_str(); }

int lts_read_app_params(const char* path, int32_t* w, int32_t* h, int32_t* spp, int32_t* bounce, int32_t* rr,
                        int32_t* fmt, char* output, char* scene, char* subject, int32_t cap) {
    if (!path || !w || !h || !spp || !bounce || !rr || !fmt || !output || !scene || !subject || cap < 2) return LTS_E_ARG;
    std::ifstream in(path);
    if (!in) return LTS_E_IO;
    std::string t[9];
    for (int i = 0; i < 9; i++)
        if (!(in >> t[i])) return LTS_E_IO;
    try {
        *w = std::stoi(t[0]); *h = std::stoi(t[1]); *spp = std::stoi(t[2]);
        *bounce = std::stoi(t[3]); *rr = std::stoi(t[4]); *fmt = std::stoi(t[5]);
    } catch (...) {
        return LTS_E_IO;
    }
    std::string outName = t[6] + (*fmt ? ".pfm" : ".hdr");  // FileIO.cuh:251
    if ((int)outName.size() >= cap || (int)t[7].size() >= cap || (int)t[8].size() >= cap) return LTS_E_ARG;
    std::strcpy(output, outName.c_str());
    std::strcpy(scene, t[7].c_str());
    std::strcpy(subject, t[8].c_str());
    return LTS_OK;
}

// Transforms nP points and nN normals (host arrays, 3 floats each) in place on c's device.
static hipError_t transform_on_device(lts_ctx* c, const MeshXf& xf, float* P, size_t nP, float* N, size_t nN) {
    const size_t need = 3 * (nP + nN);
    if (need == 0) return hipSuccess;
    if (nP > (size_t)INT32_MAX || nN > (size_t)INT32_MAX) return hipErrorInvalidValue;
    if (need > c->xfCap) {
        if (c->xfBuf) (void)hipFree(c->xfBuf);
        c->xfBuf = nullptr;
        c->xfCap = 0;
        hipError_t e = hipMalloc((void**)&c->xfBuf, need * sizeof(float));
        if (e != hipSuccess) return e;
        c->xfCap = need;
    }
    float* dP = c->xfBuf;
    float* dN = c->xfBuf + 3 * nP;
    hipError_t e = hipSuccess;
    if (nP) e = hipMemcpyAsync(dP, P, 3 * nP * sizeof(float), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && nN) e = hipMemcpyAsync(dN, N, 3 * nN * sizeof(float), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = mesh_transform(xf, dP, (int)nP, dN, (int)nN, c->stream);
    if (e == hipSuccess && nP) e = hipMemcpyAsync(P, dP, 3 * nP * sizeof(float), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && nN) e = hipMemcpyAsync(N, dN, 3 * nN * sizeof(float), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e;
}

static int load_scene_impl(lts_ctx* c, const char* dir, const char* scene, const char* subject, int32_t width, int32_t height) {
    if (!c || !dir || !scene) return fail(c, LTS_E_ARG, "null argument");
    HIPCHK(c, hipSetDevice(c->device));
    free_scene(c->scene);
    DeviceScene& S = c->scene;
    std::string err;
    std::string d(dir);
    if (!d.empty() && d.back() != '/') d += '/';
    // every mesh's points and normals go through its world transform on the GPU
    auto transform = [c](const MeshXf& xf, std::vector<float>& P, std::vector<float>& N, std::string& e) {
        const hipError_t r = transform_on_device(c, xf, P.data(), P.size() / 3, N.data(), N.size() / 3);
        if (r != hipSuccess) e = std::string("mesh transform: ") + hipGetErrorString(r);
        return r == hipSuccess;
    };
    if (!load_scene_files(transform, d, scene, subject ? subject : "", width, height, S.host, err,
                          c->textureFilter == LTS_TEXTURE_FILTERED))
        return fail(c, LTS_E_IO, err);
    const HostScene& H = S.host;
    const int n = (int)H.shade.size();
    for (int a = 0; a < 3; a++) {
        float lo = INFINITY, hi = -INFINITY;
        for (size_t i = a; i < H.soup.size(); i += 3) { lo = std::min(lo, H.soup[i]); hi = std::max(hi, H.soup[i]); }
        S.boxMin[a] = lo;
        S.boxInv[a] = hi > lo ? 1.0f / (hi - lo) : 0.0f;
    }
    if (n > kMaxPrims) return fail(c, LTS_E_ARG, "scene has more than 2^26 triangles");
    hipStream_t s = c->stream;
    HIPCHK(c, upload(&S.soup, H.soup.data(), H.soup.size(), s));
    HIPCHK(c, upload(&S.mats, H.mats.data(), H.mats.size(), s));
    HIPCHK(c, upload(&S.texs, H.texs.data(), H.texs.size(), s));
    HIPCHK(c, upload(&S.mips, H.mips.data(), H.mips.size(), s));
    HIPCHK(c, upload(&S.mipData, H.mipData.data(), H.mipData.size(), s));
    HIPCHK(c, upload(&S.lfunc, H.lfunc.data(), H.lfunc.size(), s));
    HIPCHK(c, upload(&S.lcdf, H.lcdf.data(), H.lcdf.size(), s));
    HIPCHK(c, bvh_reserve(S.bvh, n));
    int rc = run_build(c, S.bvh, S.soup, n, 1, true, nullptr);
    if (rc != LTS_OK) return rc;
    S.treeKind = LTS_TREE_KARRAS;
    if ((c->renderTree & ~LTS_TREE_COLLAPSE_EVEN) == LTS_TREE_PLOC && n > 1) {
        int depth = 0;
        HIPCHK(c, ploc_enqueue(S.bvh, s, kMaxRenderDepth, &depth));
        if (depth <= kMaxRenderDepth) S.treeKind = LTS_TREE_PLOC;
    }
    int depth = 0, nodes4 = 0;
    bool even = (c->renderTree & LTS_TREE_COLLAPSE_EVEN) != 0;
    HIPCHK(c, bvh4_enqueue(S.bvh, s, &depth, &nodes4, &even));
    if (even) S.treeKind |= LTS_TREE_COLLAPSE_EVEN;
    HIPCHK(c, hipStreamSynchronize(s));
    {
        int32_t status = 0;
        HIPCHK(c, hipMemcpy(&status, S.bvh.status, sizeof(status), hipMemcpyDeviceToHost));
        if (status != 0) return fail(c, LTS_E_STATE, "render tree build failed (status " + std::to_string(status) + ")");
    }
    // the Karras tree of at most 2^26 leaves is at most 57 deep; the check guards the stack bound
    if (depth > kMaxRenderDepth) return fail(c, LTS_E_STATE, "render tree deeper than the traversal stack allows");
    S.treeDepth = depth;
    S.nodes4 = nodes4;
    // leaf index of every area light's triangle (inverse of the Morton permutation)
    std::vector<int32_t> perm(n);
    HIPCHK(c, hipMemcpy(perm.data(), S.bvh.valsA, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    std::vector<int32_t> leafOf(n);
    for (int k = 0; k < n; k++) leafOf[perm[k]] = k;
    // shading records in leaf order, so a hit reads its record alongside its triangle
    std::vector<ShadeTri> shadeLeaf(n);
    for (int k = 0; k < n; k++) shadeLeaf[k] = H.shade[perm[k]];
    HIPCHK(c, upload(&S.shade, shadeLeaf.data(), shadeLeaf.size(), s));
    HIPCHK(c, launch_tri_materials(S.bvh.tris, S.shade, n, s));
    std::vector<LightDesc> lights = H.lights;
    for (LightDesc& L : lights)
        if (L.kind == 0) L.leaf = leafOf[L.prim];
    HIPCHK(c, upload(&S.lights, lights.data(), lights.size(), s));
    SceneView& v = S.view;
    v.tris = S.bvh.tris; v.nodes = S.bvh.nodes; v.nodes4 = S.bvh.nodes4; v.shade = S.shade; v.mats = S.mats; v.texs = S.texs;
    v.mips = S.mips; v.mipData = S.mipData; v.lights = S.lights; v.lfunc = S.lfunc; v.lcdf = S.lcdf;
    v.funcInt = H.funcInt; v.nLights = (int)H.lights.size(); v.nPrim = n; v.infiniteLight = H.infiniteLight;
    v.perm = c->perm;
    v.cam = H.cam;
    std::memcpy(v.filter, H.filter, sizeof(v.filter));
    S.loaded = true;
    c->sceneLoads++;
    c->filmNext = 0;
    return LTS_OK;
}

int lts_scene_info(lts_ctx* c, int32_t* nt, int32_t* nl, int32_t* nm) {
    if (!c) return LTS_E_ARG;
    if (!c->scene.loaded) return fail(c, LTS_E_STATE, "no scene loaded");
    if (nt) *nt = c->scene.view.nPrim;
    if (nl) *nl = c->scene.view.nLights;
    if (nm) *nm = c->scene.host.nMeshes;
    return LTS_OK;
}

int lts_set_render_tree(lts_ctx* c, int32_t kind) {
    if (!c) return LTS_E_ARG;
    const int tree = kind & ~LTS_TREE_COLLAPSE_EVEN;
    if (tree != LTS_TREE_KARRAS && tree != LTS_TREE_PLOC) return fail(c, LTS_E_ARG, "unknown render tree kind");
    c->renderTree = kind;
    return LTS_OK;
}

int lts_set_texture_filter(lts_ctx* c, int32_t mode) {
    if (!c) return LTS_E_ARG;
    if (mode != LTS_TEXTURE_REFERENCE && mode != LTS_TEXTURE_FILTERED) return fail(c, LTS_E_ARG, "unknown texture pipeline");
    c->textureFilter = mode;
    return LTS_OK;
}

static int trace_rays_impl(lts_ctx* c, const float* rays, int32_t n, int32_t* prim, float* t) {
    if (!c || !rays || !prim || !t || n < 0 || n > (1 << 26)) return fail(c, LTS_E_ARG, "bad trace_rays arguments");
    if (!c->scene.loaded) return fail(c, LTS_E_STATE, "no scene loaded");
    if (n == 0) return LTS_OK;
    HIPCHK(c, hipSetDevice(c->device));
    float* dr = nullptr;
    float* dt = nullptr;
    int32_t* dp = nullptr;
    hipError_t e = hipMalloc((void**)&dr, 6 * (size_t)n * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&dt, (size_t)n * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&dp, (size_t)n * sizeof(int32_t));
    if (e == hipSuccess) e = hipMemcpyAsync(dr, rays, 6 * (size_t)n * sizeof(float), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = launch_trace_rays(c->scene.view, dr, n, dp, dt, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(prim, dp, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(t, dt, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (dr) (void)hipFree(dr);
    if (dt) (void)hipFree(dt);
    if (dp) (void)hipFree(dp);
    HIPCHK(c, e);
    return LTS_OK;
}

int lts_trace_rays(lts_ctx* c, const float* rays, int32_t n, int32_t* prim, float* t) {
    return guarded(c, [&] { return trace_rays_impl(c, rays, n, prim, t); });
}

int lts_render_tree_info(lts_ctx* c, int32_t* kind, int32_t* depth, int32_t* nodes4) {
    if (!c) return LTS_E_ARG;
    if (!c->scene.loaded) return fail(c, LTS_E_STATE, "no scene loaded");
    if (kind) *kind = c->scene.treeKind;
    if (depth) *depth = c->scene.treeDepth;
    if (nodes4) *nodes4 = c->scene.nodes4;
    return LTS_OK;
}

static int bvh_build_impl(lts_ctx* c, const float* tri, int32_t n, int32_t mode, lts_bvh_export* out, lts_bvh_timing* t) {
    if (!c || !tri || n < 1 || n > kMaxPrims || (mode != 0 && mode != 1)) return fail(c, LTS_E_ARG, "bad build arguments");
    HIPCHK(c, hipSetDevice(c->device));
    size_t bytes = 9 * (size_t)n * sizeof(float);
    if (bytes > c->buildSoupCap) {
        if (c->buildSoup) (void)hipFree(c->buildSoup);
        c->buildSoup = nullptr;
        c->buildSoupCap = 0;
        HIPCHK(c, hipMalloc((void**)&c->buildSoup, bytes));
        c->buildSoupCap = bytes;
    }
    HIPCHK(c, hipMemcpyAsync(c->buildSoup, tri, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, bvh_reserve(c->build, n));
    int rc = run_build(c, c->build, c->buildSoup, n, mode, true, t);
    if (rc != LTS_OK) return rc;
    c->lastMode = mode;
    return out ? copy_export(c, c->build, mode, out) : LTS_OK;
}

static int bvh_build_device_impl(lts_ctx* c, const float* dtri, int32_t n, int32_t mode, int32_t repeat, lts_bvh_timing* t) {
    if (!c || n < 1 || n > kMaxPrims || (mode != 0 && mode != 1) || repeat < 1) return fail(c, LTS_E_ARG, "bad build arguments");
    if (!dtri) {
        if (!c->buildSoup || c->buildSoupCap < 9 * (size_t)n * sizeof(float)) return fail(c, LTS_E_STATE, "no uploaded soup of that size");
        dtri = c->buildSoup;
    }
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, bvh_reserve(c->build, n));
    // one build launched directly, with events between its stages
    lts_bvh_timing one{};
    int rc = run_build(c, c->build, dtri, n, mode, false, &one);
    if (rc != LTS_OK) return rc;
    if (repeat > 1) {
        // the repeats: one build captured as a graph (its kernels only) and replayed back to back
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const hipError_t ce = bvh_enqueue(c->build, dtri, n, mode, false, c->stream, nullptr);
        const hipError_t ee = hipStreamEndCapture(c->stream, &graph);
        HIPCHK(c, ce);
        HIPCHK(c, ee);
        hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e == hipSuccess) e = hipGraphLaunch(exec, c->stream);  // warm-up
        if (e == hipSuccess) e = hipEventRecord(c->ev[5], c->stream);
        for (int r = 0; e == hipSuccess && r < repeat; r++) e = hipGraphLaunch(exec, c->stream);
        if (e == hipSuccess) e = hipEventRecord(c->ev[6], c->stream);
        if (e == hipSuccess) e = hipEventSynchronize(c->ev[6]);
        if (exec) (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
        HIPCHK(c, e);
        int32_t status = 0;
        HIPCHK(c, hipMemcpy(&status, c->build.status, sizeof(status), hipMemcpyDeviceToHost));
        if (status != 0) return fail(c, LTS_E_STATE, "LBVH build failed (status " + std::to_string(status) + ")");
        one.total_ms = ms_between(c->ev[5], c->ev[6]) / repeat;
    }
    if (t) *t = one;
    c->lastMode = mode;
    return LTS_OK;
}

static int bvh_export_last_impl(lts_ctx* c, lts_bvh_export* out) {
    if (!c || !out) return LTS_E_ARG;
    if (c->build.n < 1) return fail(c, LTS_E_STATE, "no build to export");
    return copy_export(c, c->build, c->lastMode, out);
}

static int render_impl(lts_ctx* c, const lts_render_params* p, float* rgb_out, float* sample_rgb, lts_stats* st) {
    if (!c || !p) return fail(c, LTS_E_ARG, "null argument");
    if (!c->scene.loaded) return fail(c, LTS_E_STATE, "no scene loaded");
    int rc = check_render_params(c, p);
    if (rc != LTS_OK) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    if (p->integrator == LTS_INTEGRATOR_WAVEFRONT || p->integrator == LTS_INTEGRATOR_CACHED) {
        std::vector<lts_ctx*> one{c};
        return render_wavefront(c, one, p, &rgb_out, sample_rgb, st);
    }
    const SceneView& sv = c->scene.view;
    const int W = sv.cam.width, H = sv.cam.height;
    const size_t npix = (size_t)W * H;
    int b0, b1;
    sample_range(p, &b0, &b1);
    if ((rc = film_continues(c, p, b0)) != LTS_OK) return rc;
    int pass = p->spp_per_pass > 0 ? p->spp_per_pass : (int)std::max<size_t>(1, std::min<size_t>((size_t)(b1 - b0), (size_t)(1u << 28) / (npix * 12)));
    if (sample_rgb) pass = b1 - b0;  // all samples are returned: one pass
    pass = std::min(pass, b1 - b0);
    if ((rc = ensure_samples(c, npix * (size_t)pass * 3)) != LTS_OK) return rc;
    if ((rc = ensure_film(c, npix)) != LTS_OK) return rc;
    hipStream_t s = c->stream;
    if (b0 == 0) HIPCHK(c, launch_film_clear(c->film, (int)npix, s));
    RenderLaunch rl{};
    rl.spp = p->spp; rl.maxBounce = p->max_bounce; rl.rrStart = p->rr_start; rl.seed = p->seed;
    rl.stratified = p->stratified; rl.sampleL = c->sampleL; rl.film = c->film;
    rl.counters = p->count_traversal ? c->counters : nullptr;
    if (rl.counters) HIPCHK(c, hipMemsetAsync(c->counters, 0, 2 * sizeof(unsigned long long), s));
    double renderMs = 0, resolveMs = 0;
    int64_t launches = 0;
    for (int s0 = b0; s0 < b1; s0 += pass) {
        rl.s0 = s0;
        rl.s1 = std::min(b1, s0 + pass);
        HIPCHK(c, hipEventRecord(c->ev[5], s));
        HIPCHK(c, p->integrator == LTS_INTEGRATOR_BDPT ? launch_bdpt(sv, rl, s) : launch_megakernel(sv, rl, s));
        HIPCHK(c, hipEventRecord(c->ev[6], s));
        HIPCHK(c, launch_film_gather(sv, rl, s));
        HIPCHK(c, hipEventRecord(c->ev[7], s));
        HIPCHK(c, hipEventSynchronize(c->ev[7]));
        renderMs += ms_between(c->ev[5], c->ev[6]);
        resolveMs += ms_between(c->ev[6], c->ev[7]);
        launches++;
        if (sample_rgb)
            HIPCHK(c, hipMemcpy(sample_rgb, c->sampleL, npix * (size_t)(rl.s1 - rl.s0) * 3 * sizeof(float), hipMemcpyDeviceToHost));
    }
    HIPCHK(c, launch_get_render(sv, c->film, c->rgb, s));
    if (rgb_out) HIPCHK(c, hipMemcpyAsync(rgb_out, c->rgb, npix * 3 * sizeof(float), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    film_advance(c, p, b1);
    if (st) {
        std::memset(st, 0, sizeof(*st));
        st->render_ms = renderMs;
        st->resolve_ms = resolveMs;
        st->launches = launches;
        st->kernel_avg_ms = launches ? renderMs / launches : 0;
        st->samples = (int64_t)npix * (b1 - b0);
        st->msamples_per_s = (renderMs + resolveMs) > 0 ? st->samples / ((renderMs + resolveMs) * 1e3) : 0;
        if (rl.counters) {
            unsigned long long cnt[2];
            HIPCHK(c, hipMemcpy(cnt, c->counters, sizeof cnt, hipMemcpyDeviceToHost));
            st->node_visits = (int64_t)cnt[0];
            st->tri_tests = (int64_t)cnt[1];
            st->ray_nodes[0] = st->node_visits;
            st->ray_tris[0] = st->tri_tests;
        }
    }
    return LTS_OK;
}

static int render_batch_impl(lts_ctx* const* ctxs, int32_t n, const lts_render_params* p, float* const* rgb_out, lts_stats* st) {
    if (!ctxs || n < 1 || !ctxs[0] || !p) return LTS_E_ARG;
    lts_ctx* c = ctxs[0];
    if (p->integrator != LTS_INTEGRATOR_WAVEFRONT) return fail(c, LTS_E_ARG, "batched rendering uses the wavefront integrator");
    std::vector<lts_ctx*> scenes;
    for (int i = 0; i < n; i++) {
        if (!ctxs[i]) return fail(c, LTS_E_ARG, "null context in batch");
        if (ctxs[i]->device != c->device) return fail(c, LTS_E_ARG, "batch contexts on different devices");
        if (!ctxs[i]->scene.loaded) return fail(c, LTS_E_STATE, "batch context without a scene");
        for (lts_ctx* x : scenes)
            if (x == ctxs[i]) return fail(c, LTS_E_ARG, "context repeated in batch");
        scenes.push_back(ctxs[i]);
    }
    int rc = check_render_params(c, p);
    if (rc != LTS_OK) return rc;
    if (p->sample_begin != 0 || (p->sample_end != 0 && p->sample_end != p->spp))
        return fail(c, LTS_E_ARG, "batched rendering renders every sample");
    HIPCHK(c, hipSetDevice(c->device));
    for (lts_ctx* x : scenes) HIPCHK(c, hipStreamSynchronize(x->stream));
    // groups: scenes sorted by pixel count (largest first, stable), a group closes once it holds
    // batch_group_pixels pixels; large scenes fill the device alone, small ones share launches
    const size_t limit = p->batch_group_pixels < 0 ? SIZE_MAX
                         : p->batch_group_pixels == 0 ? ((size_t)1 << 20) : (size_t)p->batch_group_pixels;
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    auto pixels = [&](int i) { return (size_t)scenes[i]->scene.view.cam.width * scenes[i]->scene.view.cam.height; };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return pixels(a) > pixels(b); });
    lts_stats total{};
    size_t k = 0;
    while (k < order.size()) {
        std::vector<lts_ctx*> group;
        std::vector<float*> outs;
        size_t px = 0;
        while (k < order.size() && (group.empty() || px + pixels(order[k]) <= limit)) {
            group.push_back(scenes[order[k]]);
            outs.push_back(rgb_out ? rgb_out[order[k]] : nullptr);
            px += pixels(order[k]);
            k++;
        }
        lts_stats gs{};
        if ((rc = render_wavefront(c, group, p, outs.data(), nullptr, &gs)) != LTS_OK) return rc;
        total.render_ms += gs.render_ms;
        total.resolve_ms += gs.resolve_ms;
        total.launches += gs.launches;
        total.samples += gs.samples;
        total.autotune_ms += gs.autotune_ms;
        total.node_visits += gs.node_visits;
        total.tri_tests += gs.tri_tests;
        for (int q = 0; q < 3; q++) {
            total.ray_nodes[q] += gs.ray_nodes[q];
            total.ray_tris[q] += gs.ray_tris[q];
        }
        for (int q = 0; q < 8; q++) {
            total.phase_ms[q] += gs.phase_ms[q];
            total.phase_launches[q] += gs.phase_launches[q];
        }
        total.trace_block = gs.trace_block;
        total.persistent_blocks = gs.persistent_blocks;
        total.sort_materials = gs.sort_materials;
        total.use_graph = gs.use_graph;
        total.batch_groups++;
    }
    total.kernel_avg_ms = total.launches ? total.render_ms / total.launches : 0;
    total.msamples_per_s = (total.render_ms + total.resolve_ms) > 0 ? total.samples / ((total.render_ms + total.resolve_ms) * 1e3) : 0;
    if (st) *st = total;
    return LTS_OK;
}

static int nrc_configure_impl(lts_ctx* c, const lts_nrc_config* cfg) {
    if (!c) return LTS_E_ARG;
    return nrc_setup(c, cfg);
}

static int nrc_set_weights_impl(lts_ctx* c, const float* w) {
    if (!c || !w) return fail(c, LTS_E_ARG, "null argument");
    int rc = nrc_ensure(c);
    if (rc != LTS_OK) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    return nrc_upload_weights(c, w);
}

static int nrc_get_weights_impl(lts_ctx* c, float* w) {
    if (!c || !w) return fail(c, LTS_E_ARG, "null argument");
    int rc = nrc_ensure(c);
    if (rc != LTS_OK) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpy(w, c->nrc.master, kNrcParams * sizeof(float), hipMemcpyDeviceToHost));
    return LTS_OK;
}

static int nrc_infer_impl(lts_ctx* c, const float* raw, int32_t n, float* out) {
    if (!c || n < 0 || (n > 0 && (!raw || !out))) return fail(c, LTS_E_ARG, "bad nrc_infer arguments");
    int rc = nrc_ensure(c);
    if (rc != LTS_OK) return rc;
    if (n == 0) return LTS_OK;
    HIPCHK(c, hipSetDevice(c->device));
    if ((size_t)n > c->nrcIoCap) {
        if (c->nrcIo) (void)hipFree(c->nrcIo);
        c->nrcIo = nullptr;
        c->nrcIoCap = 0;
        HIPCHK(c, hipMalloc((void**)&c->nrcIo, (size_t)n * (kNrcRaw + 3) * sizeof(float)));
        c->nrcIoCap = n;
    }
    hipStream_t s = c->stream;
    float* dRaw = c->nrcIo;
    float* dOut = c->nrcIo + (size_t)n * kNrcRaw;
    HIPCHK(c, hipMemcpyAsync(dRaw, raw, (size_t)n * kNrcRaw * sizeof(float), hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(c->nrcCount, &n, sizeof(int32_t), hipMemcpyHostToDevice, s));
    HIPCHK(c, nrc_infer(c->nrc, dRaw, c->nrcCount, n, dOut, s));
    HIPCHK(c, hipMemcpyAsync(out, dOut, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    return LTS_OK;
}

static int nrc_train_step_impl(lts_ctx* c, const float* raw, const float* target, int32_t n, float* loss, float* grad) {
    if (!c || n < 0 || (n > 0 && (!raw || !target))) return fail(c, LTS_E_ARG, "bad nrc_train_step arguments");
    int rc = nrc_ensure(c);
    if (rc != LTS_OK) return rc;
    if (n > c->nrc.batchCap) return fail(c, LTS_E_ARG, "more samples than the configured batch");
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    if (n > 0) {
        HIPCHK(c, hipMemcpyAsync(c->nrc.trainRaw, raw, (size_t)n * kNrcRaw * sizeof(float), hipMemcpyHostToDevice, s));
        HIPCHK(c, hipMemcpyAsync(c->nrc.trainTarget, target, (size_t)n * 3 * sizeof(float), hipMemcpyHostToDevice, s));
    }
    HIPCHK(c, hipMemcpyAsync(&c->nrc.state->nTotal, &n, sizeof(int32_t), hipMemcpyHostToDevice, s));
    HIPCHK(c, nrc_train(c->nrc, s, 1));
    NrcState st;
    HIPCHK(c, hipMemcpyAsync(&st, c->nrc.state, sizeof(st), hipMemcpyDeviceToHost, s));
    if (grad) HIPCHK(c, hipMemcpyAsync(grad, c->nrc.grad, kNrcParams * sizeof(float), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    if (loss) *loss = n > 0 ? st.loss : 0.0f;
    return LTS_OK;
}

int lts_device_copy_benchmark(lts_ctx* c, uint64_t bytes, int32_t repeat, double* ms) {
    if (!c || bytes < 16 || repeat < 1 || !ms) return fail(c, LTS_E_ARG, "bad copy arguments");
    HIPCHK(c, hipSetDevice(c->device));
    void *a = nullptr, *b = nullptr;
    HIPCHK(c, hipMalloc(&a, bytes));
    hipError_t e = hipMalloc(&b, bytes);
    if (e != hipSuccess) { (void)hipFree(a); return fail(c, LTS_E_NOMEM, "copy buffers"); }
    e = launch_copy(b, a, bytes, c->stream);  // warm-up
    if (e == hipSuccess) e = hipEventRecord(c->ev[5], c->stream);
    for (int r = 0; e == hipSuccess && r < repeat; r++) e = launch_copy(b, a, bytes, c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->ev[6], c->stream);
    if (e == hipSuccess) e = hipEventSynchronize(c->ev[6]);
    if (e == hipSuccess) *ms = ms_between(c->ev[5], c->ev[6]) / repeat;
    (void)hipFree(a);
    (void)hipFree(b);
    HIPCHK(c, e);
    return LTS_OK;
}

// Film checkpoint file: "LTSFILM1", int32 width, height, next sample, spp, seed, stratified,
// max_bounce, rr_start, integrator, then width * height * 4 floats (rgb sums, filter weight).
static int film_save_impl(lts_ctx* c, const char* path) {
    if (!c || !path) return LTS_E_ARG;
    if (!c->scene.loaded || c->filmNext == 0 || !c->film) return fail(c, LTS_E_STATE, "no rendered film to save");
    HIPCHK(c, hipSetDevice(c->device));
    const size_t npix = (size_t)c->filmKey[0] * c->filmKey[1];
    std::vector<float> film(npix * 4);
    HIPCHK(c, hipMemcpy(film.data(), c->film, film.size() * sizeof(float), hipMemcpyDeviceToHost));
    std::ofstream f(path, std::ios::binary);
    const int32_t head[9] = {c->filmKey[0], c->filmKey[1], c->filmNext, c->filmKey[2], c->filmKey[3],
                             c->filmKey[4], c->filmKey[5], c->filmKey[6], c->filmKey[7]};
    f.write("LTSFILM1", 8);
    f.write(reinterpret_cast<const char*>(head), sizeof head);
    f.write(reinterpret_cast<const char*>(film.data()), (std::streamsize)(film.size() * sizeof(float)));
    if (!f) return fail(c, LTS_E_IO, std::string("cannot write ") + path);
    return LTS_OK;
}

static int film_load_impl(lts_ctx* c, const char* path) {
    if (!c || !path) return LTS_E_ARG;
    if (!c->scene.loaded) return fail(c, LTS_E_STATE, "no scene loaded");
    HIPCHK(c, hipSetDevice(c->device));
    std::ifstream f(path, std::ios::binary);
    char magic[8];
    int32_t head[9];
    if (!f.read(magic, 8) || std::memcmp(magic, "LTSFILM1", 8) != 0 || !f.read(reinterpret_cast<char*>(head), sizeof head))
        return fail(c, LTS_E_IO, std::string("not a film checkpoint: ") + path);
    const CameraDesc& cam = c->scene.view.cam;
    if (head[0] != cam.width || head[1] != cam.height) return fail(c, LTS_E_ARG, "checkpoint film size differs from the scene's");
    if (head[2] < 1 || head[2] > head[3]) return fail(c, LTS_E_IO, "bad checkpoint sample count");
    const size_t npix = (size_t)head[0] * head[1];
    std::vector<float> film(npix * 4);
    if (!f.read(reinterpret_cast<char*>(film.data()), (std::streamsize)(film.size() * sizeof(float))))
        return fail(c, LTS_E_IO, std::string("truncated film checkpoint: ") + path);
    int rc = ensure_film(c, npix);
    if (rc != LTS_OK) return rc;
    HIPCHK(c, hipMemcpy(c->film, film.data(), film.size() * sizeof(float), hipMemcpyHostToDevice));
    const int32_t key[9] = {head[0], head[1], head[3], head[4], head[5], head[6], head[7], head[8], c->sceneLoads};
    std::memcpy(c->filmKey, key, sizeof key);
    c->filmNext = head[2];
    return LTS_OK;
}

static int film_resolve_impl(lts_ctx* c, float* rgb_out) {
    if (!c || !rgb_out) return LTS_E_ARG;
    if (!c->scene.loaded || c->filmNext == 0 || !c->film) return fail(c, LTS_E_STATE, "no film to resolve");
    HIPCHK(c, hipSetDevice(c->device));
    const SceneView& sv = c->scene.view;
    const size_t npix = (size_t)sv.cam.width * sv.cam.height;
    HIPCHK(c, launch_get_render(sv, c->film, c->rgb, c->stream));
    HIPCHK(c, hipMemcpyAsync(rgb_out, c->rgb, npix * 3 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LTS_OK;
}

static int write_image_impl(const char* path, const float* rgb, int32_t w, int32_t h, int32_t format) {
    if (!path || !rgb || w < 1 || h < 1) return LTS_E_ARG;
    if (format < 0 || format > 3) return LTS_E_ARG;
    const bool ok = format == 0   ? write_ppm(path, rgb, w, h)
                    : format == 1 ? write_pfm(path, rgb, w, h)
                    : format == 2 ? write_hdr(path, rgb, w, h)
                                  : write_png(path, rgb, w, h);
    return ok ? LTS_OK : LTS_E_IO;
}

static int load_image_impl(const char* path, int32_t* w, int32_t* h, float* rgb_out, int64_t cap) {
    if (!path || !w || !h) return LTS_E_ARG;
    int iw = 0, ih = 0;
    std::vector<float> rgb;
    std::string err;
    if (!load_image_linear(path, iw, ih, rgb, err)) return fail(nullptr, LTS_E_IO, err);
    *w = iw; *h = ih;
    if (!rgb_out) return LTS_OK;
    if (cap < (int64_t)rgb.size()) return LTS_E_ARG;
    std::memcpy(rgb_out, rgb.data(), rgb.size() * sizeof(float));
    return LTS_OK;
}

static int obj_to_bin_impl(const char* stem, int32_t* nTris) {
    if (!stem) return LTS_E_ARG;
    std::vector<char> bin;
    std::string err;
    int n = 0;
    if (!obj_mesh_bytes(stem, bin, &n, err)) return fail(nullptr, LTS_E_IO, err);
    std::ofstream out(std::string(stem) + ".bin", std::ios::binary);
    if (!out || !out.write(bin.data(), (std::streamsize)bin.size()))
        return fail(nullptr, LTS_E_IO, std::string("cannot write ") + stem + ".bin");
    if (nTris) *nTris = n;
    return LTS_OK;
}

// Entry points: host allocation failures and other exceptions become error codes, never
// an exception crossing the C boundary.
int lts_load_scene(lts_ctx* c, const char* dir, const char* scene, const char* subject, int32_t width, int32_t height) {
    return guarded(c, [&] { return load_scene_impl(c, dir, scene, subject, width, height); });
}

int lts_bvh_build(lts_ctx* c, const float* tri, int32_t n, int32_t mode, lts_bvh_export* out, lts_bvh_timing* t) {
    return guarded(c, [&] { return bvh_build_impl(c, tri, n, mode, out, t); });
}

int lts_bvh_build_device(lts_ctx* c, const float* dtri, int32_t n, int32_t mode, int32_t repeat, lts_bvh_timing* t) {
    return guarded(c, [&] { return bvh_build_device_impl(c, dtri, n, mode, repeat, t); });
}

int lts_mesh_transform(lts_ctx* c, const float* subject, float* points, int32_t n_points, float* normals, int32_t n_normals) {
    return guarded(c, [&] {
        if (!c || !subject || n_points < 0 || n_normals < 0 || (n_points > 0 && !points) || (n_normals > 0 && !normals))
            return fail(c, LTS_E_ARG, "bad mesh transform arguments");
        HIPCHK(c, hipSetDevice(c->device));
        HIPCHK(c, transform_on_device(c, mesh_xf_from_tokens(subject), points, (size_t)n_points, normals, (size_t)n_normals));
        return LTS_OK;
    });
}

int lts_bvh_export_last(lts_ctx* c, lts_bvh_export* out) {
    return guarded(c, [&] { return bvh_export_last_impl(c, out); });
}

int lts_render(lts_ctx* c, const lts_render_params* p, float* rgb_out, float* sample_rgb, lts_stats* st) {
    return guarded(c, [&] { return render_impl(c, p, rgb_out, sample_rgb, st); });
}

int lts_render_batch(lts_ctx* const* ctxs, int32_t n, const lts_render_params* p, float* const* rgb_out, lts_stats* st) {
    return guarded((ctxs && n > 0) ? ctxs[0] : nullptr, [&] { return render_batch_impl(ctxs, n, p, rgb_out, st); });
}

int lts_nrc_configure(lts_ctx* c, const lts_nrc_config* cfg) {
    return guarded(c, [&] { return nrc_configure_impl(c, cfg); });
}

int lts_nrc_set_weights(lts_ctx* c, const float* w) {
    return guarded(c, [&] { return nrc_set_weights_impl(c, w); });
}

int lts_nrc_get_weights(lts_ctx* c, float* w) {
    return guarded(c, [&] { return nrc_get_weights_impl(c, w); });
}

int lts_nrc_infer(lts_ctx* c, const float* raw, int32_t n, float* out) {
    return guarded(c, [&] { return nrc_infer_impl(c, raw, n, out); });
}

int lts_nrc_train_step(lts_ctx* c, const float* raw, const float* target, int32_t n, float* loss, float* grad) {
    return guarded(c, [&] { return nrc_train_step_impl(c, raw, target, n, loss, grad); });
}

int lts_film_save(lts_ctx* c, const char* path) {
    return guarded(c, [&] { return film_save_impl(c, path); });
}

int lts_film_load(lts_ctx* c, const char* path) {
    return guarded(c, [&] { return film_load_impl(c, path); });
}

int lts_film_resolve(lts_ctx* c, float* rgb_out) {
    return guarded(c, [&] { return film_resolve_impl(c, rgb_out); });
}

int lts_write_image(const char* path, const float* rgb, int32_t w, int32_t h, int32_t format) {
    return guarded(nullptr, [&] { return write_image_impl(path, rgb, w, h, format); });
}

int lts_bvh_build_static(const float* d_tri_xyz, int32_t n, int32_t mode, uint32_t* d_codes_sorted, int32_t* d_perm,
                         int32_t* d_left, int32_t* d_right, int32_t* d_parent, float* d_bounds, float* scene_bounds,
                         void* hip_stream) {
    static_assert(LTS_BVH_STATIC_MAX_PRIMS == kStaticMaxPrims && LTS_BVH_STATIC_SCRATCH_BYTES == kStaticScratchBytes,
                  "header limits of the allocation-free build");
    if (!d_tri_xyz || n < 1 || n > kStaticMaxPrims || (mode != 0 && mode != 1) || !d_codes_sorted || !d_perm ||
        !d_parent || !d_bounds || (n > 1 && (!d_left || !d_right)))
        return LTS_E_ARG;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return LTS_E_HIP;
    void* scratch = nullptr;
    if (bvh_static_scratch(dev, &scratch) != hipSuccess) return LTS_E_HIP;
    hipStream_t s = static_cast<hipStream_t>(hip_stream);
    int32_t* dStatus = nullptr;
    uint32_t* dKeys = nullptr;
    if (bvh_enqueue_static(scratch, d_tri_xyz, n, mode, d_codes_sorted, d_perm, d_left, d_right, d_parent, d_bounds, s,
                           &dStatus, &dKeys) != hipSuccess)
        return LTS_E_HIP;
    int32_t status = 0;
    uint32_t keys[6];
    if (hipMemcpyAsync(&status, dStatus, sizeof(status), hipMemcpyDeviceToHost, s) != hipSuccess ||
        (scene_bounds && hipMemcpyAsync(keys, dKeys, sizeof(keys), hipMemcpyDeviceToHost, s) != hipSuccess) ||
        hipStreamSynchronize(s) != hipSuccess)
        return LTS_E_HIP;
    if (status != 0) return LTS_E_STATE;
    if (scene_bounds)
        for (int a = 0; a < 6; a++) {
            const uint32_t u = (keys[a] & 0x80000000u) ? (keys[a] & 0x7fffffffu) : ~keys[a];
            std::memcpy(&scene_bounds[a], &u, 4);
        }
    return LTS_OK;
}

int lts_load_image(const char* path, int32_t* w, int32_t* h, float* rgb_out, int64_t cap) {
    return guarded(nullptr, [&] { return load_image_impl(path, w, h, rgb_out, cap); });
}
int lts_obj_to_bin(const char* stem, int32_t* nTris) {
    return guarded(nullptr, [&] { return obj_to_bin_impl(stem, nTris); });
}

}  // extern "C"
