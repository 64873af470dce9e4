// Wavefront form of the plain path tracer (gfx950).
//
// The reference integrator (intergrators/PathTracingIntegrator.cuh:38-104 with estimateDirect,
// IntegratorHelper.cuh:51-125) is cut at its three ray casts:
//   extend   closest hit of every live path ray                     (persistent, wave-batched)
//   sort     optional counting sort of the hit queue by material
//   shade    emission, BSDF, light sample + MIS-BSDF sample (NEE set-up), continuation, roulette;
//            appends one shadow ray and one MIS ray per path at most
//   shadow   any hit of the shadow rays                             (persistent, wave-batched)
//   mis      closest hit of the MIS-BSDF rays, light identity test  (persistent, wave-batched)
//   finalize L += beta * ((Ld_light + Ld_bsdf) / p_pick), the reference's order of additions;
//            terminated paths write their radiance
// Each path's arithmetic is the megakernel's, so the image does not depend on queue order.
// Every count lives in device memory and is reset by kernels, so the bounce loop of a pass is
// captured once in a hipGraph and replayed.
#include "lts_internal.h"
#include "pt_device.h"

namespace lts {

namespace {

constexpr int kShadeBlock = 256;

// Pixel of the j-th path of a sample: the image is cut into bands of T rows and each band is walked
// column by column, so 64 consecutive paths (one wave) cover a T x 64/T tile and their rays stay
// close (T = 8: an 8x8 tile; T = 1: scanline order). The last band holds the remaining rows.
LTS_D int band_pixel(int j, int W, int H, int T) {
    const int band = j / (T * W);
    const int rows = min(T, H - T * band);
    const int jj = j - band * T * W;
    return (T * band + jj % rows) * W + jj / rows;
}

// Scene of path i of the pass: the last slot whose path range starts at or before i.
LTS_D int path_scene(const WfParams& P, const WfSceneSlot* slots, long i) {
    int lo = 0, hi = P.nScenes - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (slots[mid].pathBase <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Path i of the pass -> (scene, pixel, sample).
LTS_D void path_origin(const WfParams& P, const WfSceneSlot* slots, long i, int& scene, int& pixel, int& sample) {
    const int s = path_scene(P, slots, i);
    const WfSceneSlot& sl = slots[s];
    const long local = i - sl.pathBase;
    const long npix = (long)sl.width * sl.height;
    scene = s;
    sample = sl.s0 + (int)(local / npix);
    pixel = band_pixel((int)(local % npix), sl.width, sl.height, P.tileRows);
}

// Training paths of the radiance cache: in scene 0, every trainStride-th (sample, pixel) pair of
// the pass in sample-major, row-major order from trainOffset, so that the set and its slot order
// do not depend on the order in which the pass walks its pixels (tileRows); the slot, or -1.
LTS_D int train_slot(const WfParams& P, const WfSceneSlot& sl, int scene, int pixel, int sample) {
    if (P.trainSlots > 0 && scene == 0) {
        const long c = (long)(sample - sl.s0) * ((long)sl.width * sl.height) + pixel;
        if (c >= P.trainOffset) {
            const long d = c - P.trainOffset;
            if (d % P.trainStride == 0 && d / P.trainStride < P.trainSlots) return (int)(d / P.trainStride);
        }
    }
    return -1;
}

// The camera rays of the pass. The path records are not written here: at bounce 0 every field
// follows from the path index, and the first shading builds them.
__global__ __launch_bounds__(256) void k_wf_generate(const SceneView* __restrict__ scenes, const WfParams* __restrict__ prm,
                                                     const WfSceneSlot* __restrict__ slots, WfRay* __restrict__ rays,
                                                     int32_t* __restrict__ queue, int32_t* __restrict__ counts,
                                                     WfTrainRec* __restrict__ tRec) {
    const WfParams P = *prm;
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i == 0) counts[0] = (int)P.nPaths;  // bounce-0 queue holds every path of the pass
    for (long k = i; k < P.nPaths; k += (long)gridDim.x * 256) {
        int sc, pixel, sample;
        path_origin(P, slots, k, sc, pixel, sample);
        const SceneView& sv = scenes[sc];
        const WfSceneSlot& sl = slots[sc];
        const int px = pixel % sv.cam.width, py = pixel / sv.cam.width;
        CamRays cr = camera_sample(sv, px, py, sample, sl.spp, sl.stratified, sl.seed);
        const int ts = train_slot(P, sl, sc, pixel, sample);
        if (ts >= 0)
            for (int b = 0; b < P.maxBounce; b++) tRec[(size_t)ts * kWfMaxBounce + b].valid = 0;
        WfRay r;
        r.o[0] = cr.r.o.x; r.o[1] = cr.r.o.y; r.o[2] = cr.r.o.z;
        r.d[0] = cr.r.d.x; r.d[1] = cr.r.d.y; r.d[2] = cr.r.d.z;
        r.scene = sc; r.pad = 0;
        rays[k] = r;
        queue[k] = (int)k;
    }
}

// Wave-batched work fetch (lane 0 takes 64 items at a time) over kWfWorkSegments contiguous segments of segLen items (a multiple of
// 64), one counter each. Workgroups are dealt to the 8 XCDs round robin, so the waves of XCD
// blockIdx % 8 start on their own segment and keep neighbouring rays in one L2; a wave whose
// segment has run dry moves on to the next ones. Returns the item of this lane, with `j` (uniform
// over the wave) advanced past the segments found empty; j == kWfWorkSegments means no work is left.
LTS_D int wave_fetch_segmented(int32_t* work, int segLen, int& j) {
    const int lane = threadIdx.x & 63;
    const int home = blockIdx.x % kWfWorkSegments;
    while (j < kWfWorkSegments) {
        const int seg = (home + j) % kWfWorkSegments;
        int base = 0;
        if (lane == 0) base = atomicAdd(&work[kWfWorkStride * seg], 64);
        base = __shfl(base, 0);
        if (base < segLen) return seg * segLen + base + lane;
        j++;
    }
    return 0;
}

// Closest hit of path ray `pi`; with `hist` set, the hit is also filed under its material bin
// (misses in the last bin) for the counting sort: bins[path] and the block histogram lh.
LTS_D void extend_item(const SceneView* scenes, const WfRay* rays, WfHit* hits, int pi,
                       TraceStack ts, int32_t* bins, int32_t* lh, bool filing, int nBins, TraceTally* tally = nullptr) {
    const WfRay rr = rays[pi];
    Ray r;
    r.o = v3(rr.o[0], rr.o[1], rr.o[2]);
    r.d = v3(rr.d[0], rr.d[1], rr.d[2]);
    r.tMax = INFINITY;
    const SceneView& sv = scenes[rr.scene];
    Hit h = trace_closest(sv, r, ts, nullptr, nullptr, tally);
    WfHit o;
    o.leaf = h.leaf; o.u = h.u; o.v = h.v; o.invDet = h.invDet;
    hits[pi] = o;
    if (filing) {
        const int bin = h.leaf < 0 ? nBins - 1 : min(sv.tris[h.leaf].mat, nBins - 2);
        bins[pi] = bin;
        atomicAdd(&lh[bin], 1);
    }
}

// Counting sort of a queue by the material bin the extend kernel filed each hit under: one block
// turns the histogram into bin offsets; each scatter block then takes one contiguous chunk of the
// queue, counts its entries per bin, reserves its range of every bin with one atomic per bin and
// writes the chunk's entries into those ranges, ranked within each wave by ballots.
__global__ void k_wf_mat_scan(int32_t* __restrict__ hist, int nBins) {
    if (threadIdx.x == 0) {
        int run = 0;
        for (int b = 0; b < nBins; b++) { int c = hist[b]; hist[b] = run; run += c; }
    }
}

// Rank of each lane among the lanes of its wave holding the same bin (bin < 0: none), offset by
// the running count h[bin], which grows by the number of such lanes (one LDS atomic per bin).
LTS_D int wave_bin_rank(int bin, int32_t* h) {
    const int lane = threadIdx.x & 63;
    const uint64_t below = lane ? ~0ull >> (64 - lane) : 0ull;
    int slot = 0;
    uint64_t todo = __ballot(bin >= 0);
    while (todo) {
        const int leader = __ffsll((unsigned long long)todo) - 1;
        const int b = __shfl(bin, leader);
        const uint64_t peers = __ballot(bin == b);
        int base = 0;
        if (lane == leader) base = atomicAdd(&h[b], (int)__popcll(peers));
        base = __shfl(base, leader);
        if (bin == b) slot = base + (int)__popcll(peers & below);
        todo &= ~peers;
    }
    return slot;
}

__global__ __launch_bounds__(256) void k_wf_mat_scatter(const int32_t* __restrict__ bins, const int32_t* __restrict__ queue,
                                                        const int32_t* __restrict__ count, int32_t* __restrict__ offs, int nBins,
                                                        int32_t* __restrict__ sorted) {
    __shared__ int32_t h[kWfMaxMatBins], base[kWfMaxMatBins];
    const int n = *count;
    const int per = ((n + (int)gridDim.x - 1) / (int)gridDim.x + 255) / 256 * 256;
    const int begin = min((int)blockIdx.x * per, n), end = min(begin + per, n);
    if (threadIdx.x < kWfMaxMatBins) h[threadIdx.x] = 0;
    __syncthreads();
    for (int k0 = begin; k0 < end; k0 += 256) {
        const int k = k0 + threadIdx.x;
        (void)wave_bin_rank(k < end ? bins[queue[k]] : -1, h);
    }
    __syncthreads();
    if (threadIdx.x < nBins) {
        base[threadIdx.x] = h[threadIdx.x] ? atomicAdd(&offs[threadIdx.x], h[threadIdx.x]) : 0;
        h[threadIdx.x] = 0;
    }
    __syncthreads();
    for (int k0 = begin; k0 < end; k0 += 256) {
        const int k = k0 + threadIdx.x;
        const int pi = k < end ? queue[k] : -1;
        const int bin = pi >= 0 ? bins[pi] : -1;
        const int slot = wave_bin_rank(bin, h);
        if (pi >= 0) sorted[base[bin] + slot] = pi;
    }
}

LTS_D float clamp01(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

// The cache's 16-float record of a shading point (layout in include/lts.h, lts_nrc_config).
LTS_D void nrc_raw_record(const SurfacePoint& si, const Bsdf& bsdf, const WfSceneSlot& sl, float* __restrict__ o) {
    const V3 wo = si.wo, n = faceTowards(si.sn, si.wo);
    const bool glossy = bsdf.kind == 4;
    float4 r0, r1, r2, r3;
    r0.x = clamp01((si.p.x - sl.boxMin[0]) * sl.boxInv[0]);
    r0.y = clamp01((si.p.y - sl.boxMin[1]) * sl.boxInv[1]);
    r0.z = clamp01((si.p.z - sl.boxMin[2]) * sl.boxInv[2]);
    r0.w = acosf(fminf(fmaxf(wo.z, -1.0f), 1.0f)) * INV_PI_F;
    r1.x = atan2f(wo.y, wo.x) * (0.5f * INV_PI_F) + 0.5f;
    r1.y = acosf(fminf(fmaxf(n.z, -1.0f), 1.0f)) * INV_PI_F;
    r1.z = atan2f(n.y, n.x) * (0.5f * INV_PI_F) + 0.5f;
    r1.w = glossy ? fminf(bsdf.alpha, 1.0f) : 1.0f;
    r2.x = glossy ? 0.0f : bsdf.R.r;
    r2.y = glossy ? 0.0f : bsdf.R.g;
    r2.z = glossy ? 0.0f : bsdf.R.b;
    r2.w = glossy ? bsdf.R.r : 0.0f;
    r3.x = glossy ? bsdf.R.g : 0.0f;
    r3.y = glossy ? bsdf.R.b : 0.0f;
    r3.z = r3.w = 0.0f;
    float4* q = reinterpret_cast<float4*>(o);
    q[0] = r0; q[1] = r1; q[2] = r2; q[3] = r3;
}

// One shading kernel per (integrator, first bounce or not), so that each carries only its own code
// and registers: the radiance cache's records, and the camera-ray differentials of bounce 0. The
// kernel waits on a chain of dependent gathers; at 3 waves per SIMD (168 VGPRs, a few dozen spilled)
// it runs about 4 % faster than at the 2 its 209-233 registers would allow.
template <bool kCached, bool kFirst>
__global__ __launch_bounds__(kShadeBlock, 3) void k_wf_shade(const SceneView* __restrict__ scenes, const WfParams* __restrict__ prm,
                                                          const WfSceneSlot* __restrict__ slots, int bounce,
                                                          WfPath* __restrict__ paths, WfRay* __restrict__ rays,
                                                          const WfHit* __restrict__ hits, WfNee* __restrict__ nee,
                                                          const int32_t* __restrict__ queue, const int32_t* __restrict__ count,
                                                          int32_t* __restrict__ nextQueue, int32_t* __restrict__ nextCount,
                                                          WfShadow* __restrict__ shadows, int32_t* __restrict__ shadowCount,
                                                          WfMis* __restrict__ misRays, int32_t* __restrict__ misCount,
                                                          float* __restrict__ qRaw, NrcQuery* __restrict__ qMeta,
                                                          int32_t* __restrict__ qCount,
                                                          WfTrainRec* __restrict__ tRec, float* __restrict__ tRaw,
                                                          int32_t* __restrict__ doneQueue, int32_t* __restrict__ doneCount,
                                                          int32_t* __restrict__ shadowQueue, int32_t* __restrict__ misQueue) {
    const int n = *count;
    const int maxBounce = prm->maxBounce, rrStart = prm->rrStart, queryBounce = prm->queryBounce;
    constexpr bool cached = kCached;
    // the procedural textures' permutation lookups chain three deep: served from LDS (every scene
    // view points at its context's one table)
    __shared__ int perm[512];
    __shared__ int32_t appendLds[2][5 * (kShadeBlock / 64) + 5];
    // a query's raw record (4 float4) and its NrcQuery (3 float4), held by the thread that makes it
    // until the round's append gives the query its compact slot
    constexpr int kStage = 7;
    __shared__ float4 qStage[kCached ? kShadeBlock * kStage : 1];
    for (int i = threadIdx.x; i < 512; i += kShadeBlock) perm[i] = scenes[0].perm[i];
    __syncthreads();
    const int stride = gridDim.x * kShadeBlock;
    const int rounds = (n + stride - 1) / stride;  // uniform trip count: the appends below are wave-collective
    // chunk of the block within a round: workgroups are dealt to the 8 XCDs round robin, so the
    // blocks of one XCD take neighbouring chunks (neighbouring paths hit neighbouring geometry)
    const int chunk = gridDim.x % 8 == 0 ? (blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8 : blockIdx.x;
    for (int round = 0; round < rounds; round++) {
        const int k = round * stride + chunk * kShadeBlock + threadIdx.x;
        const bool active = k < n;
        const int pi = active ? queue[k] : 0;
        bool wantShadow = false, wantMis = false, wantNext = false, wantQuery = false, wantDone = false;
        WfShadow sh;
        WfMis ms;
        if (active) {
            const WfParams P = *prm;
            WfPath p;
            int scene, pixel, sample, pflags;
            if (kFirst) {
                // bounce 0: the path record follows from the path index
                p.beta[0] = p.beta[1] = p.beta[2] = 1.0f;
                p.L[0] = p.L[1] = p.L[2] = 0.0f;
                path_origin(P, slots, pi, scene, pixel, sample);
                pflags = kWfAlive;
            } else {
                p = paths[pi];
                scene = path_scene(P, slots, pi);
                pixel = p.pixel;
                sample = (int)(p.sampleFlags >> kWfFlagBits);
                pflags = (int)(p.sampleFlags & ((1u << kWfFlagBits) - 1));
            }
            const WfHit hh = hits[pi];
            const SceneView& sv = scenes[scene];
            const WfSceneSlot& sl = slots[scene];
            const int trainSlot = cached ? train_slot(P, sl, scene, pixel, sample) : -1;
            Ray r;
            CamRays cam;
            if (kFirst) {
                // the camera ray and its differentials, drawn again as the generate kernel drew
                // them (a path that escapes needs neither)
                if (hh.leaf >= 0) {
                    cam = camera_sample(sv, pixel % sv.cam.width, pixel / sv.cam.width, sample, sl.spp, sl.stratified, sl.seed);
                    r = cam.r;
                }
            } else {
                const WfRay rr = rays[pi];
                r.o = v3(rr.o[0], rr.o[1], rr.o[2]);
                r.d = v3(rr.d[0], rr.d[1], rr.d[2]);
            }
            r.tMax = INFINITY;
            Spec L = spec(p.L[0], p.L[1], p.L[2]), beta = spec(p.beta[0], p.beta[1], p.beta[2]);
            if (!kFirst) {
                // the light sample of the previous vertex, traced since: added before anything of
                // this vertex, as the reference's loop adds it before the next intersection
                const WfNee pn = nee[pi];
                if (pn.used) {
                    Spec Ld = spec(pn.p1[0], pn.p1[1], pn.p1[2]) + spec(pn.p2[0], pn.p2[1], pn.p2[2]);
                    L = L + spec(pn.nb[0], pn.nb[1], pn.nb[2]) * (Ld / pn.lpdf);
                }
            }
            const bool specularBounce = (pflags & kWfSpecular) != 0;
            WfNee ne;
            ne.used = 0;
            ne.p1[0] = ne.p1[1] = ne.p1[2] = 0.0f;
            ne.p2[0] = ne.p2[1] = ne.p2[2] = 0.0f;
            bool alive = true;
            if (kFirst || specularBounce) {
                if (hh.leaf >= 0) {
                    const int l = sv.shade[hh.leaf].light;
                    if (l >= 0) {
                        SurfacePoint e = make_surface(sv, hh.leaf, r, hh.u, hh.v, hh.invDet);
                        Spec le = dot(e.n, -r.d) > 0.f ? light_le(sv.lights[l]) : spec(0.0f);
                        L = L + beta * le;
                    } else {
                        L = L + beta * spec(0.0f);
                    }
                } else {
                    for (int l = 0; l < sv.nLights; l++) L = L + beta * (sv.lights[l].kind == 1 ? light_le(sv.lights[l]) : spec(0.0f));
                }
            }
            if (hh.leaf < 0) {
                alive = false;
            } else {
                SurfacePoint si = make_surface(sv, hh.leaf, r, hh.u, hh.v, hh.invDet);
                if (kFirst) differentials(si, cam, uv_derivative_scale(sv, hh.leaf, hh.invDet));
                Bsdf bsdf = make_bsdf(sv, si, perm);
                const int slot = (sample - sl.s0) * sl.width * sl.height + pixel;  // sample index of the path
                if (cached && (bsdf.kind == 0 || bsdf.kind == 1 || bsdf.kind == 4)) {
                    if (trainSlot >= 0) {
                        // training vertex: radiance so far and throughput, for the suffix target
                        const size_t r = (size_t)trainSlot * kWfMaxBounce + bounce;
                        nrc_raw_record(si, bsdf, sl, tRaw + 16 * r);
                        WfTrainRec tr;
                        tr.Lb[0] = L.r; tr.Lb[1] = L.g; tr.Lb[2] = L.b;
                        tr.beta[0] = beta.r; tr.beta[1] = beta.g; tr.beta[2] = beta.b;
                        tr.a[0] = bsdf.R.r; tr.a[1] = bsdf.R.g; tr.a[2] = bsdf.R.b;
                        tr.slot = slot; tr.valid = 1; tr.pad = 0;
                        tRec[r] = tr;
                    } else if (bounce >= queryBounce) {
                        // the cache stands for all radiance scattered here
                        nrc_raw_record(si, bsdf, sl, reinterpret_cast<float*>(&qStage[kStage * threadIdx.x]));
                        NrcQuery q;
                        q.beta[0] = beta.r; q.beta[1] = beta.g; q.beta[2] = beta.b;
                        q.a[0] = bsdf.R.r; q.a[1] = bsdf.R.g; q.a[2] = bsdf.R.b;
                        q.L[0] = q.L[1] = q.L[2] = 0.0f;  // set where the path ends, below
                        q.slot = slot; q.pad[0] = q.pad[1] = 0;
                        *reinterpret_cast<NrcQuery*>(&qStage[kStage * threadIdx.x + 4]) = q;
                        wantQuery = true;
                        alive = false;
                    }
                }
                Sampler smp;
                smp.seed = sl.seed; smp.pixel = (uint32_t)pixel; smp.sample = (uint32_t)sample;
                float u0[4], u1[4];
                smp.block(1 + 2 * bounce, u0);
                smp.block(2 + 2 * bounce, u1);
                if (!wantQuery && sv.nLights > 0) {
                    float lpdf;
                    const int li = sample_light(sv, u0[0], &lpdf);
                    const LightDesc& Lt = sv.lights[li];
                    ne.used = 1;
                    ne.nb[0] = beta.r; ne.nb[1] = beta.g; ne.nb[2] = beta.b;
                    ne.lpdf = lpdf;
                    const int flags = specularBounce ? B_ALL : (B_ALL & ~B_SPEC);
                    // light sample with MIS (IntegratorHelper.cuh:76-95)
                    V3 wi;
                    float lightPdf = 0.0f, scatteringPdf = 0.0f;
                    LightPoint vis;
                    Spec Li = sample_li(sv, Lt, si, v2(u0[1], u0[2]), &wi, &lightPdf, &vis);
                    if (lightPdf > 0 && !isBlack(Li)) {
                        Spec f = bsdf.f(si.wo, wi, flags) * absDot(wi, si.sn);
                        scatteringPdf = bsdf.pdf(si.wo, wi, flags);
                        if (!isBlack(f)) {
                            float weight = power_heuristic(lightPdf, scatteringPdf);
                            Spec c = f * Li * weight / lightPdf;
                            V3 origin = offsetRayOrigin(si.p, si.pError, si.n, vis.p - si.p);
                            V3 target = offsetRayOrigin(vis.p, vis.pError, vis.n, origin - vis.p);
                            V3 d = target - origin;
                            sh.o[0] = origin.x; sh.o[1] = origin.y; sh.o[2] = origin.z;
                            sh.d[0] = d.x; sh.d[1] = d.y; sh.d[2] = d.z;
                            sh.path = pi; sh.scene = scene;
                            sh.c[0] = c.r; sh.c[1] = c.g; sh.c[2] = c.b;
                            wantShadow = true;
                            shadows[pi] = sh;  // per path; the queue below lists the paths
                        }
                    }
                    // BSDF sample with MIS (IntegratorHelper.cuh:96-122)
                    int sampledType = 0;
                    Spec f2 = bsdf.sampleF(si.wo, &wi, v2(u0[3], u1[0]), &scatteringPdf, flags, &sampledType);
                    f2 = f2 * absDot(wi, si.sn);
                    const bool sampledSpecular = (sampledType & B_SPEC) != 0;
                    if (!isBlack(f2) && scatteringPdf > 0) {
                        float weight = 1;
                        bool cast = true;
                        if (!sampledSpecular) {
                            lightPdf = Lt.kind == 0 ? tri_pdf(sv, Lt, si, wi) : inf_pdf(wi);
                            if (lightPdf == 0) cast = false;
                            else weight = power_heuristic(scatteringPdf, lightPdf);
                        }
                        if (cast) {
                            Ray mr = spawn_ray(si, wi);
                            ms.o[0] = mr.o.x; ms.o[1] = mr.o.y; ms.o[2] = mr.o.z;
                            ms.d[0] = wi.x; ms.d[1] = wi.y; ms.d[2] = wi.z;
                            ms.path = pi; ms.light = li; ms.scene = scene;
                            ms.f[0] = f2.r; ms.f[1] = f2.g; ms.f[2] = f2.b;
                            ms.weight = weight; ms.pdf = scatteringPdf;
                            wantMis = true;
                            misRays[pi] = ms;
                        }
                    }
                }
                // continuation (PathTracingIntegrator.cuh:79-103)
                V3 wo = -r.d, wi;
                float pdf;
                int flags = 0;
                Spec f = wantQuery ? spec(0.0f) : bsdf.sampleF(wo, &wi, v2(u1[1], u1[2]), &pdf, B_ALL, &flags);
                if (wantQuery || isBlack(f) || pdf == 0) {
                    alive = false;
                } else {
                    pflags = (flags & B_SPEC) ? (pflags | kWfSpecular) : (pflags & ~kWfSpecular);
                    beta = beta * (f * absDot(wi, si.sn) / pdf);
                    Ray nr = spawn_ray(si, wi);
                    WfRay o;
                    o.o[0] = nr.o.x; o.o[1] = nr.o.y; o.o[2] = nr.o.z;
                    o.d[0] = wi.x; o.d[1] = wi.y; o.d[2] = wi.z;
                    o.scene = scene; o.pad = 0;
                    rays[pi] = o;
                    if (bounce >= rrStart) {
                        float q = fmaxf(0.05f, 1 - lum(beta));
                        if (u1[3] < q) alive = false;
                        else beta = beta / (1 - q);
                    }
                }
            }
            p.L[0] = L.r; p.L[1] = L.g; p.L[2] = L.b;
            p.beta[0] = beta.r; p.beta[1] = beta.g; p.beta[2] = beta.b;
            if (!alive) pflags &= ~kWfAlive;
            p.pixel = pixel;
            p.sampleFlags = ((uint32_t)sample << kWfFlagBits) | (uint32_t)pflags;
            wantNext = alive && bounce + 1 < maxBounce;
            // the records are read again only by the next shading or by finalize
            if (wantNext || ne.used) {
                paths[pi] = p;
                nee[pi] = ne;
            }
            if (!wantNext) {
                // the path ends here: its sample is final now, or once its light sample is traced
                if (ne.used) {
                    wantDone = true;
                } else if (cached && wantQuery) {
                    // the sample is written by inference: L plus the cache's estimate
                    NrcQuery* q = reinterpret_cast<NrcQuery*>(&qStage[kStage * threadIdx.x + 4]);
                    q->L[0] = L.r; q->L[1] = L.g; q->L[2] = L.b;
                } else {
                    float* o = sl.sampleL + 3 * ((size_t)(sample - sl.s0) * sl.width * sl.height + pixel);
                    o[0] = L.r; o[1] = L.g; o[2] = L.b;
                }
            }
        }
        // The round's appends, aggregated per block: each wave files its counts in LDS, one lane per
        // queue reserves the block's range with a single device atomic (one per wave on a handful of
        // shared counters serialised), and every lane writes at its offset. The buffers alternate
        // between rounds so that one barrier pair per round suffices.
        constexpr int kQueues = 5, kWaves = kShadeBlock / 64;
        const bool want[kQueues] = {wantShadow, wantMis, wantNext, wantDone, cached && wantQuery};
        int32_t* const counter[kQueues] = {shadowCount, misCount, nextCount, doneCount, qCount};
        int32_t* const lds = appendLds[round & 1];
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const uint64_t below = lane ? ~0ull >> (64 - lane) : 0ull;
        uint64_t mask[kQueues];
#pragma unroll
        for (int q = 0; q < kQueues; q++) {
            mask[q] = __ballot(want[q]);
            if (lane == 0) lds[q * kWaves + wave] = (int)__popcll(mask[q]);
        }
        __syncthreads();
        if (threadIdx.x < kQueues && (threadIdx.x < 4 || cached)) {
            const int q = threadIdx.x;
            int run = 0;
            for (int w = 0; w < kWaves; w++) { const int c = lds[q * kWaves + w]; lds[q * kWaves + w] = run; run += c; }
            lds[kQueues * kWaves + q] = run ? atomicAdd(counter[q], run) : 0;
        }
        __syncthreads();
        int32_t* const out[kQueues - 1] = {shadowQueue, misQueue, nextQueue, doneQueue};
#pragma unroll
        for (int q = 0; q < kQueues - 1; q++)
            if (want[q]) out[q][lds[kQueues * kWaves + q] + lds[q * kWaves + wave] + (int)__popcll(mask[q] & below)] = pi;
        if (cached && want[4]) {
            // queries are stored compactly in append order, so inference reads them in sequence
            const size_t at = (size_t)(lds[kQueues * kWaves + 4] + lds[4 * kWaves + wave] + (int)__popcll(mask[4] & below));
            float4* r = reinterpret_cast<float4*>(qRaw + 16 * at);
#pragma unroll
            for (int j = 0; j < 4; j++) r[j] = qStage[kStage * threadIdx.x + j];
            float4* m = reinterpret_cast<float4*>(qMeta + at);
#pragma unroll
            for (int j = 0; j < 3; j++) m[j] = qStage[kStage * threadIdx.x + 4 + j];
        }
    }
}

// Builds the training batch from this pass's training records in record order, so the batch and
// the step that follows are deterministic: (1) each block counts the valid records of its 2048,
// (2) one block turns the counts into block offsets, (3) each block writes its records at its
// offset plus an in-block prefix. A record contributes its raw encoding input and the suffix target
// (L_final - Lb) / beta / a per channel (0 where beta or a is 0).
constexpr int kCollectPerThread = 8;
constexpr int kCollectBlock = 256;
constexpr int kCollectPerBlock = kCollectPerThread * kCollectBlock;

LTS_D size_t train_record(int i, int maxBounce) { return (size_t)(i / maxBounce) * kWfMaxBounce + (i % maxBounce); }

// Block-wide exclusive prefix of v (256 threads); total receives the block sum.
LTS_D int block_prefix(int v, int* sh, int& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wave] = x;
    __syncthreads();
    int waveBase = 0;
    for (int w = 0; w < wave; w++) waveBase += sh[w];
    total = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return waveBase + x - v;
}

__global__ __launch_bounds__(kCollectBlock) void k_wf_nrc_count(const WfParams* __restrict__ prm, const WfTrainRec* __restrict__ tRec,
                                                                int32_t* __restrict__ blockCount) {
    __shared__ int sh[4];
    const int maxBounce = prm->maxBounce, nRec = prm->trainSlots * maxBounce;
    const int i0 = blockIdx.x * kCollectPerBlock + threadIdx.x * kCollectPerThread;
    int c = 0;
    for (int i = i0; i < min(i0 + kCollectPerThread, nRec); i++) c += tRec[train_record(i, maxBounce)].valid;
    int total;
    block_prefix(c, sh, total);
    if (threadIdx.x == 0) blockCount[blockIdx.x] = total;
}

__global__ __launch_bounds__(64) void k_wf_nrc_scan(int32_t* __restrict__ blockCount, int nBlocks, int cap,
                                                    NrcState* __restrict__ state) {
    if (threadIdx.x != 0) return;
    int run = 0;
    for (int b = 0; b < nBlocks; b++) { const int c = blockCount[b]; blockCount[b] = run; run += c; }
    state->nTotal = min(run, cap);
}

__global__ __launch_bounds__(kCollectBlock) void k_wf_nrc_scatter(const WfParams* __restrict__ prm, const WfTrainRec* __restrict__ tRec,
                                                                  const float* __restrict__ tRaw, const float* __restrict__ sampleBase,
                                                                  const int32_t* __restrict__ blockBase, int batchCap,
                                                                  float* __restrict__ outRaw, float* __restrict__ outTarget) {
    __shared__ int sh[4];
    const int maxBounce = prm->maxBounce, nRec = prm->trainSlots * maxBounce;
    const int i0 = blockIdx.x * kCollectPerBlock + threadIdx.x * kCollectPerThread;
    const int i1 = min(i0 + kCollectPerThread, nRec);
    int c = 0;
    for (int i = i0; i < i1; i++) c += tRec[train_record(i, maxBounce)].valid;
    int total;
    int pos = blockBase[blockIdx.x] + block_prefix(c, sh, total);
    for (int i = i0; i < i1 && pos < batchCap; i++) {
        const size_t r = train_record(i, maxBounce);
        const WfTrainRec t = tRec[r];
        if (!t.valid) continue;
        const float* Lf = sampleBase + 3 * (size_t)t.slot;
        for (int ch = 0; ch < 3; ch++) {
            const float suffix = t.beta[ch] > 0.0f ? (Lf[ch] - t.Lb[ch]) / t.beta[ch] : 0.0f;
            outTarget[3 * (size_t)pos + ch] = t.a[ch] > 0.0f ? suffix / t.a[ch] : 0.0f;
        }
        const float4* src = reinterpret_cast<const float4*>(tRaw + 16 * r);
        float4* dst = reinterpret_cast<float4*>(outRaw + 16 * (size_t)pos);
        dst[0] = src[0]; dst[1] = src[1]; dst[2] = src[2]; dst[3] = src[3];
        pos++;
    }
}

// Any hit of the shadow ray of path k: the light sample's contribution when the segment is free.
LTS_D void shadow_item(const SceneView* scenes, const WfShadow* shadows, WfNee* nee, int k,
                       TraceStack ts, TraceTally* tally = nullptr) {
    const WfShadow s = shadows[k];
    Ray r;
    r.o = v3(s.o[0], s.o[1], s.o[2]);
    r.d = v3(s.d[0], s.d[1], s.d[2]);
    r.tMax = 1 - SHADOW_EPS;
    if (!trace_any(scenes[s.scene], r, ts, tally)) {
        nee[s.path].p1[0] = s.c[0];
        nee[s.path].p1[1] = s.c[1];
        nee[s.path].p1[2] = s.c[2];
    }
}

// The MIS-BSDF ray of path k: does it reach the light it was cast for (IntegratorHelper.cuh:103-122).
LTS_D void mis_item(const SceneView* scenes, const WfMis* misRays, WfNee* nee, int k, TraceStack ts,
                    TraceTally* tally = nullptr) {
    const WfMis m = misRays[k];
    const SceneView& sv = scenes[m.scene];
    const LightDesc& Lt = sv.lights[m.light];
    Ray r;
    r.o = v3(m.o[0], m.o[1], m.o[2]);
    r.d = v3(m.d[0], m.d[1], m.d[2]);
    r.tMax = INFINITY;
    const Spec Li2 = mis_reaches_light(sv, Lt, m.light, r, ts, tally) ? light_le(Lt) : spec(0.f);
    if (!isBlack(Li2)) {
        Spec c = spec(m.f[0], m.f[1], m.f[2]) * Li2 * spec(1.0f) * m.weight / m.pdf;
        nee[m.path].p2[0] = c.r;
        nee[m.path].p2[1] = c.g;
        nee[m.path].p2[2] = c.b;
    }
}

// One persistent launch traces every ray that is ready: the path rays of the next bounce (queue
// `ext`), then the shadow rays and the MIS rays of the bounce just shaded. The three sets do not
// depend on each other, so one work counter over their concatenation balances long rays against
// short ones and leaves one tail per bounce instead of three. Counts may be null (empty set).
// The launch asks for 7 waves per SIMD (at most 72 VGPRs): enough waves to cover the load latency
// of the traversal, which then runs at the VALU issue limit; at 8 the 64-register cap spills.
// kCount: the counting instantiation, which also adds every lane's render-node fetches and
// triangle tests per ray kind (path, shadow, MIS) to traceCounts[0..5] (nodes, tests per kind).
template <int BS, bool kCount>
__global__ __launch_bounds__(BS, 7) void k_wf_trace(const SceneView* __restrict__ scenes, const WfRay* __restrict__ rays, WfHit* __restrict__ hits,
                                                 const int32_t* __restrict__ ext, const int32_t* __restrict__ extCount,
                                                 int32_t* __restrict__ bins, int32_t* __restrict__ hist, int nBins,
                                                 const WfShadow* __restrict__ shadows, const int32_t* __restrict__ shadowCount,
                                                 const WfMis* __restrict__ misRays, const int32_t* __restrict__ misCount,
                                                 const int32_t* __restrict__ shadowQueue, const int32_t* __restrict__ misQueue,
                                                 WfNee* __restrict__ nee, int32_t* __restrict__ work,
                                                 unsigned long long* __restrict__ traceCounts) {
    __shared__ int stk[kShortStack * BS];
    __shared__ int32_t lh[kWfMaxMatBins];
    const TraceStack ts = lds_stack(stk, threadIdx.x, BS);
    const bool filing = hist != nullptr;
    if (filing) {
        for (int b = threadIdx.x; b < kWfMaxMatBins; b += BS) lh[b] = 0;
        __syncthreads();
    }
    const int nE = extCount ? *extCount : 0;
    const int nS = shadowCount ? *shadowCount : 0;
    const int nM = misCount ? *misCount : 0;
    const int n = nE + nS + nM;
    const int segLen = ((n + kWfWorkSegments - 1) / kWfWorkSegments + 63) / 64 * 64;
    int j = 0;
    TraceTally tally[3];
    for (;;) {
        const int k = wave_fetch_segmented(work, segLen, j);
        if (j >= kWfWorkSegments) break;
        if (k >= n) continue;
        if (k < nE) extend_item(scenes, rays, hits, ext[k], ts, bins, lh, filing, nBins, kCount ? &tally[0] : nullptr);
        else if (k < nE + nS) shadow_item(scenes, shadows, nee, shadowQueue[k - nE], ts, kCount ? &tally[1] : nullptr);
        else mis_item(scenes, misRays, nee, misQueue[k - nE - nS], ts, kCount ? &tally[2] : nullptr);
    }
    if constexpr (kCount) {
        for (int q = 0; q < 3; q++) {
            atomicAdd(&traceCounts[2 * q], tally[q].nodes);
            atomicAdd(&traceCounts[2 * q + 1], tally[q].tris);
        }
    }
    if (filing) {
        __syncthreads();
        for (int b = threadIdx.x; b < nBins; b += BS)
            if (lh[b]) atomicAdd(&hist[b], lh[b]);
    }
}

// Paths that ended at this bounce with a light sample in flight: the traced sample completes their
// radiance (a continuing path takes its sample at the start of the next bounce's shading).
__global__ __launch_bounds__(256) void k_wf_finalize(const WfParams* __restrict__ prm, const WfSceneSlot* __restrict__ slots,
                                                     const WfPath* __restrict__ paths, const WfNee* __restrict__ nee,
                                                     const int32_t* __restrict__ done, const int32_t* __restrict__ count) {
    const int n = *count;
    const WfParams P = *prm;
    for (int k = blockIdx.x * 256 + threadIdx.x; k < n; k += gridDim.x * 256) {
        const int pi = done[k];
        const WfPath p = paths[pi];
        const WfNee ne = nee[pi];
        Spec L = spec(p.L[0], p.L[1], p.L[2]);
        Spec Ld = spec(ne.p1[0], ne.p1[1], ne.p1[2]) + spec(ne.p2[0], ne.p2[1], ne.p2[2]);
        L = L + spec(ne.nb[0], ne.nb[1], ne.nb[2]) * (Ld / ne.lpdf);
        const WfSceneSlot& sl = slots[path_scene(P, slots, pi)];
        const int sample = (int)(p.sampleFlags >> kWfFlagBits);
        float* o = sl.sampleL + 3 * ((size_t)(sample - sl.s0) * sl.width * sl.height + p.pixel);
        o[0] = L.r; o[1] = L.g; o[2] = L.b;
    }
}

// Passes with max_bounce == 0 produce black samples.
__global__ __launch_bounds__(256) void k_wf_zero(const WfParams* __restrict__ prm, const WfSceneSlot* __restrict__ slots) {
    const WfParams P = *prm;
    for (long k = (long)blockIdx.x * 256 + threadIdx.x; k < P.nPaths; k += (long)gridDim.x * 256) {
        int s = 0;
        while (s + 1 < P.nScenes && k >= slots[s + 1].pathBase) s++;
        const WfSceneSlot& sl = slots[s];
        sl.sampleL[3 * (k - sl.pathBase) + 0] = 0.0f;
        sl.sampleL[3 * (k - sl.pathBase) + 1] = 0.0f;
        sl.sampleL[3 * (k - sl.pathBase) + 2] = 0.0f;
    }
}

__global__ void k_wf_fill(int32_t* p, int n, int32_t v) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) p[i] = v;
}

}  // namespace

hipError_t wf_enqueue_pass(const WfBuffers& b, const WfTuning& t, int maxBounce, const NrcBuffers* nrc, hipStream_t s,
                           const WfProfile* prof, unsigned long long* traceCounts) {
    auto mark = [&](int i) {
        if (prof) (void)hipEventRecord(prof->ev[i], s);
    };
    const int gridAll = t.sweepBlocks;
    const int ncounts = kWfCountsPerBounce * (maxBounce + 1) + kWfMaxMatBins * (maxBounce + 1);
    k_wf_fill<<<(ncounts + 255) / 256, 256, 0, s>>>(b.counts, ncounts, 0);
    k_wf_fill<<<1, 256, 0, s>>>(b.counts + kWfGlobalBase, kWfGlobalCounts, 0);
    if (maxBounce <= 0) {
        k_wf_zero<<<gridAll, 256, 0, s>>>(b.params, b.slots);
        return hipGetLastError();
    }
    const bool cached = nrc != nullptr;
    int32_t* qCount = b.counts + kWfGlobalBase;
    k_wf_generate<<<gridAll, 256, 0, s>>>(b.scenes, b.params, b.slots, b.rays, b.queue[0], b.counts, b.tRec);
    auto histOf = [&](int bounce) {
        return t.sortMaterials ? b.counts + kWfCountsPerBounce * (maxBounce + 1) + kWfMaxMatBins * bounce : nullptr;
    };
    // traces the path rays of `bounce` (queue q, count cnt[0]) and the shadow and MIS rays of the
    // bounce before it (counts prev[1], prev[2]), with the work counters of `bounce`
    auto trace = [&](int bounce, const int32_t* extCount, const int32_t* shadowCount, const int32_t* misCount) {
        const int32_t* ext = b.queue[bounce & 1];
        int32_t* work = b.counts + kWfCountsPerBounce * bounce + kWfWorkBase;
        int32_t* hist = histOf(bounce);
        auto launch = [&](auto kernel, int bs) {
            kernel<<<t.persistentBlocks, bs, 0, s>>>(b.scenes, b.rays, b.hits, ext, extCount, b.bins, hist, b.matBins, b.shadows,
                                                     shadowCount, b.mis, misCount, b.shadowQ, b.misQ, b.nee, work, traceCounts);
        };
        const bool counting = traceCounts != nullptr;
        switch (t.traceBlock) {
            case 64: launch(counting ? k_wf_trace<64, true> : k_wf_trace<64, false>, 64); break;
            case 128: launch(counting ? k_wf_trace<128, true> : k_wf_trace<128, false>, 128); break;
            default: launch(counting ? k_wf_trace<256, true> : k_wf_trace<256, false>, 256); break;
        }
    };
    mark(8 * kWfMaxBounce + 3);
    trace(0, b.counts, nullptr, nullptr);
    for (int bounce = 0; bounce < maxBounce; bounce++) {
        const int q = bounce & 1;
        int32_t* cnt = b.counts + kWfCountsPerBounce * bounce;
        int32_t* nxt = b.counts + kWfCountsPerBounce * (bounce + 1);
        mark(8 * bounce + 0);
        const int32_t* shadeQueue = b.queue[q];
        if (t.sortMaterials) {
            int32_t* hist = histOf(bounce);
            k_wf_mat_scan<<<1, 64, 0, s>>>(hist, b.matBins);
            k_wf_mat_scatter<<<gridAll, 256, 0, s>>>(b.bins, b.queue[q], cnt, hist, b.matBins, b.sorted);
            shadeQueue = b.sorted;
        }
        mark(8 * bounce + 1);
        {
            auto shade = cached ? (bounce == 0 ? k_wf_shade<true, true> : k_wf_shade<true, false>)
                                : (bounce == 0 ? k_wf_shade<false, true> : k_wf_shade<false, false>);
            shade<<<gridAll, kShadeBlock, 0, s>>>(b.scenes, b.params, b.slots, bounce, b.paths, b.rays, b.hits, b.nee, shadeQueue,
                                                  cnt, b.queue[q ^ 1], nxt, b.shadows, cnt + 1, b.mis, cnt + 2,
                                                  cached ? b.qRaw : nullptr, b.qMeta, qCount, b.tRec, b.tRaw,
                                                  b.done, cnt + 3, b.shadowQ, b.misQ);
        }
        mark(8 * bounce + 2);
        // the next bounce's path rays (none after the last bounce) with this bounce's shadow and MIS rays
        trace(bounce + 1, nxt, cnt + 1, cnt + 2);
        mark(8 * bounce + 3);
        k_wf_finalize<<<gridAll, 256, 0, s>>>(b.params, b.slots, b.paths, b.nee, b.done, cnt + 3);
        mark(8 * bounce + 4);
    }
    if (cached) {
        mark(8 * kWfMaxBounce + 0);
        hipError_t e = nrc_infer_queries(*nrc, b.qRaw, b.qMeta, qCount, (int)b.cap, b.sampleBase, s);
        if (e != hipSuccess) return e;
        mark(8 * kWfMaxBounce + 1);
        const int nb = (b.trainCap * maxBounce + kCollectPerBlock - 1) / kCollectPerBlock;
        if (nb > 0) {
            k_wf_nrc_count<<<nb, kCollectBlock, 0, s>>>(b.params, b.tRec, b.tCount);
            const int cap = nrc->batchCap * nrc->chunks;
            k_wf_nrc_scan<<<1, 64, 0, s>>>(b.tCount, nb, cap, nrc->state);
            k_wf_nrc_scatter<<<nb, kCollectBlock, 0, s>>>(b.params, b.tRec, b.tRaw, b.sampleBase, b.tCount, cap,
                                                          nrc->trainRaw, nrc->trainTarget);
        }
        if ((e = nrc_train(*nrc, s, nrc->steps)) != hipSuccess) return e;  // batches cycle when steps > batches
        mark(8 * kWfMaxBounce + 2);
    }
    return hipGetLastError();
}

}  // namespace lts

namespace lts {

template <typename T>
static hipError_t wf_alloc(T** p, size_t count) {
    return hipMalloc((void**)p, (count ? count : 1) * sizeof(T));
}

hipError_t wf_reserve(WfBuffers& b, size_t cap, int maxScenes) {
    if (cap <= b.cap && maxScenes <= b.maxScenes) return hipSuccess;
    wf_release(b);
    hipError_t e = hipSuccess;
#define WTRY(x) do { e = (x); if (e != hipSuccess) { wf_release(b); return e; } } while (0)
    WTRY(wf_alloc(&b.scenes, maxScenes));
    WTRY(wf_alloc(&b.params, 1));
    WTRY(wf_alloc(&b.slots, maxScenes));
    WTRY(wf_alloc(&b.paths, cap));
    WTRY(wf_alloc(&b.rays, cap));
    WTRY(wf_alloc(&b.hits, cap));
    WTRY(wf_alloc(&b.nee, cap));
    WTRY(wf_alloc(&b.queue[0], cap));
    WTRY(wf_alloc(&b.queue[1], cap));
    WTRY(wf_alloc(&b.sorted, cap));
    WTRY(wf_alloc(&b.done, cap));
    WTRY(wf_alloc(&b.shadowQ, cap));
    WTRY(wf_alloc(&b.misQ, cap));
    WTRY(wf_alloc(&b.bins, cap));
    WTRY(wf_alloc(&b.shadows, cap));
    WTRY(wf_alloc(&b.mis, cap));
    WTRY(wf_alloc(&b.counts, (size_t)kWfGlobalBase + kWfGlobalCounts));
#undef WTRY
    b.cap = cap;
    b.maxScenes = maxScenes;
    b.gen = next_buffer_generation();
    return hipSuccess;
}

hipError_t wf_reserve_cache(WfBuffers& b, size_t cap, int trainCap) {
    if (cap <= b.cacheCap && trainCap <= b.trainCap) return hipSuccess;
    void* ps[] = {b.qRaw, b.qMeta, b.tRec, b.tRaw, b.tCount};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    b.qRaw = nullptr; b.qMeta = nullptr; b.tRec = nullptr; b.tRaw = nullptr; b.tCount = nullptr;
    b.cacheCap = 0; b.trainCap = 0;
    hipError_t e = hipSuccess;
#define CTRY(x) do { e = (x); if (e != hipSuccess) return e; } while (0)
    CTRY(wf_alloc(&b.qRaw, cap * kNrcRaw));
    CTRY(wf_alloc(&b.qMeta, cap));
    CTRY(wf_alloc(&b.tRec, (size_t)trainCap * kWfMaxBounce));
    CTRY(wf_alloc(&b.tRaw, (size_t)trainCap * kWfMaxBounce * kNrcRaw));
    CTRY(wf_alloc(&b.tCount, ((size_t)trainCap * kWfMaxBounce + kCollectPerBlock - 1) / kCollectPerBlock));
#undef CTRY
    b.cacheCap = cap;
    b.trainCap = trainCap;
    b.cacheGen = next_buffer_generation();
    return hipSuccess;
}

void wf_release(WfBuffers& b) {
    void* cs[] = {b.qRaw, b.qMeta, b.tRec, b.tRaw, b.tCount};
    for (void* p : cs)
        if (p) (void)hipFree(p);
    void* ps[] = {b.scenes, b.params, b.slots, b.paths, b.rays, b.hits, b.nee, b.queue[0], b.queue[1], b.sorted, b.done, b.shadowQ, b.misQ, b.bins, b.shadows, b.mis, b.counts};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    b = WfBuffers();
}

}  // namespace lts
