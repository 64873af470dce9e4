// Internal interfaces between the C ABI layer and the kernel translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lts_scene.h"

namespace lts {

// Device buffers of one LBVH build (sized for `cap` primitives; reused across builds).
struct BvhBuffers {
    int cap = 0;
    int n = 0;
    int tiles = 0;
    float* primBox = nullptr;     // 6n    primitive bounds
    float* box = nullptr;         // 6(2n-1) node bounds, unified index
    uint32_t* keysA = nullptr;    // n
    uint32_t* keysB = nullptr;
    int32_t* valsA = nullptr;
    int32_t* valsB = nullptr;
    uint32_t* counts = nullptr;   // 256 * tiles
    uint32_t* digitBase = nullptr;// 256
    int32_t* left = nullptr;      // n-1
    int32_t* right = nullptr;     // n-1
    int32_t* parent = nullptr;    // 2n-1
    int32_t* flags = nullptr;     // n-1
    uint32_t* sceneKeys = nullptr;// 6 ordered keys of the scene bounds
    BvhNode* nodes = nullptr;     // n-1 (or 1 for n == 1)
    TriLeaf* tris = nullptr;      // n
    // results after the sort: sorted codes and permutation live in keysA / valsA
};

hipError_t bvh_reserve(BvhBuffers& b, int n);
void bvh_release(BvhBuffers& b);

struct BvhStageEvents {
    hipEvent_t e[5];  // start, after bounds+morton, after sort, after tree, after pack
};

// Enqueues the whole build on `stream`. `soup` is a device pointer to n*9 floats.
// mode 0: leaf k holds primitive k (reference); mode 1: leaf k holds perm[k].
// When pack is set, also writes nodes / tris for the traversal kernels (mode 1 layout).
hipError_t bvh_enqueue(BvhBuffers& b, const float* soup, int n, int mode, bool pack, hipStream_t stream,
                       BvhStageEvents* ev);

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
hipError_t launch_film_gather(const SceneView& sv, const RenderLaunch& rl, hipStream_t stream);
hipError_t launch_film_clear(float* film, int npix, hipStream_t stream);
hipError_t launch_get_render(const SceneView& sv, const float* film, float* rgb, hipStream_t stream);

}  // namespace lts
