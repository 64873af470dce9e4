// Bidirectional path tracing for gfx950: one thread per (pixel, sample), launched and gathered onto
// the film exactly as the plain megakernel is (render.hip).
//
// The reference has a bidirectional integrator (intergrators/BidirectionalPathIntegrator.cuh:1-665)
// that its author disabled as not functional (LightTransportSimulation.cu:116-121): the light walk
// is handed `path` rather than `path + 1` and overwrites the light vertex (:396-397), the strategy
// loops stop one short of the longest subpaths (:627-628), and the walk's first reverse density is
// written through path[-1] (:296, :336). With no working output to hold parity against, this
// integrator follows the algorithm the reference sets out to implement (PBRT-v3 chapter 16: a
// camera subpath and a light subpath per sample, every connection strategy, balance-heuristic MIS
// from per-vertex forward and reverse area densities) and is pinned by convergence: where the plain
// path tracer is unbiased, the two estimate the same film.
//
// Strategies: every (s, t) with t >= 2 and s + t - 2 <= max_bounce, s counting light-subpath
// vertices and t camera-subpath ones. The light-tracing strategies (t = 1, splatted onto the film
// through the camera's importance, Camera.cu:88-185, 225-234) are left out: the camera is never
// hit, so every path they sample is also sampled by a strategy with t >= 2; the MIS weights are
// taken over the strategies used, so the estimate stays unbiased, and each sample's value belongs
// to its own film position, which keeps the deterministic film gather and reproducible images.
//
// Lights: an area light starts the light subpath (picked with the probability the scene's light
// search actually gives it, a uniform point, a cosine-weighted direction about the normal the hit
// test uses for emission) and takes the s = 1 connections. The infinite light is reached only by
// camera subpaths that escape (s = 0), the one strategy for such paths (weight 1). No Russian
// roulette: the reference's walk has none (:285-338).
#include "pt_device.h"
#include "lts_internal.h"

namespace lts {

namespace {

constexpr int kTileSide = 16;
constexpr int kCamVerts = kBdptMaxBounce + 2;
constexpr int kLightVerts = kBdptMaxBounce + 1;

// Sampler blocks (4 dimensions each): 0 film and lens, 1.. the camera walk, then the light
// subpath's start and walk, then the s = 1 connections.
constexpr uint32_t kBlockCamWalk = 1;
constexpr uint32_t kBlockLightStart = kBlockCamWalk + kCamVerts;
constexpr uint32_t kBlockLightWalk = kBlockLightStart + 2;
constexpr uint32_t kBlockConnect = kBlockLightWalk + kLightVerts;

enum : int { kVCamera = 0, kVLight = 1, kVSurface = 2, kVEscape = 3 };

struct BVertex {
    V3 p, n, sn, pError, wo;  // n: the geometric normal as the hit test faces it; sn: shading normal
    Spec beta;
    float pdfFwd, pdfRev;  // area densities of generating this vertex from either side
    int kind;
    int light;             // light index of an emitting vertex, else -1
    bool delta;            // sampled by a specular BSDF
    Bsdf bsdf;
};

LTS_D bool on_surface(const BVertex& v) { return v.kind == kVLight || v.kind == kVSurface; }

// Vertex::IsConnectible: area-light points and surfaces whose BSDF is not specular.
LTS_D bool connectible(const BVertex& v) {
    if (v.kind == kVLight) return true;
    return v.kind == kVSurface && (v.bsdf.kind == 0 || v.bsdf.kind == 1 || v.bsdf.kind == 4);
}

// A solid-angle density at v toward `next` as an area density at `next` (Vertex::ConvertDensity).
LTS_D float to_area(float pdf, const BVertex& v, const BVertex& next) {
    V3 w = next.p - v.p;
    float d2 = lengthSq(w);
    if (d2 == 0) return 0;
    float inv = 1 / d2;
    if (on_surface(next)) pdf *= absDot(next.n, w * sqrtf(inv));
    return pdf * inv;
}

// Probability that sample_light returns light j: its search never returns the last light and
// gives that light's share to the one before (Sampling.cuh:155-188).
LTS_D float pick_prob(const SceneView& sv, int j) {
    const int n = sv.nLights;
    if (n == 1) return 1.0f;
    if (j == n - 1) return 0.0f;
    const float hi = j == n - 2 ? 1.0f : sv.lcdf[j + 1];
    return hi - sv.lcdf[j];
}

// Density of starting a light subpath at area-light vertex v (Vertex::PdfLightOrigin).
LTS_D float origin_pdf(const SceneView& sv, const BVertex& v) {
    if (v.light < 0 || sv.lights[v.light].kind != 0) return 0.0f;
    return pick_prob(sv, v.light) / sv.lights[v.light].area;
}

// Emitted radiance of an emitting surface vertex toward w: area lights emit on the side of the
// normal the hit test uses (dot(n, w) > 0).
LTS_D Spec emitted(const SceneView& sv, const BVertex& v, V3 w) {
    if (v.light < 0) return spec(0.0f);
    const LightDesc& L = sv.lights[v.light];
    return L.kind == 0 && dot(v.n, w) > 0.f ? light_le(L) : spec(0.0f);
}

// PBRT-v3 CorrectShadingNormal: the adjoint BSDF of importance transport with shading normals.
LTS_D float shading_correction(const BVertex& v, V3 wo, V3 wi, bool importance) {
    if (!importance) return 1.0f;
    float num = absDot(wo, v.sn) * absDot(wi, v.n);
    float den = absDot(wo, v.n) * absDot(wi, v.sn);
    return den == 0 ? 0.0f : num / den;
}

// BSDF value of a surface vertex toward `next` (Vertex::f).
LTS_D Spec vertex_f(const BVertex& v, const BVertex& next, bool importance) {
    V3 wi = next.p - v.p;
    if (lengthSq(wi) == 0) return spec(0.0f);
    wi = normalize(wi);
    return v.bsdf.f(v.wo, wi, B_ALL) * shading_correction(v, v.wo, wi, importance);
}

// Area density at `next` of emitting vertex v sending light toward it: area lights emit
// cosine-weighted about their normal (Vertex::PdfLight).
LTS_D float emit_pdf(const BVertex& v, const BVertex& next) {
    V3 wn = next.p - v.p;
    if (lengthSq(wn) == 0) return 0.0f;
    return to_area(fmaxf(0.0f, dot(v.n, normalize(wn))) * INV_PI_F, v, next);
}

// Area density at `next` of v scattering toward it, having been reached from `prev`
// (Vertex::Pdf); an area-light vertex with no predecessor emits.
LTS_D float vertex_pdf(const BVertex& v, const BVertex* prev, const BVertex& next) {
    if (v.kind == kVLight && !prev) return emit_pdf(v, next);
    V3 wn = next.p - v.p;
    if (lengthSq(wn) == 0) return 0.0f;
    wn = normalize(wn);
    float pdf;
    if (v.kind == kVSurface && prev) {
        V3 wp = prev->p - v.p;
        if (lengthSq(wp) == 0) return 0.0f;
        pdf = v.bsdf.pdf(normalize(wp), wn, B_ALL);
    } else {
        return 0.0f;
    }
    return to_area(pdf, v, next);
}

LTS_D SurfacePoint as_point(const BVertex& v) {
    SurfacePoint s;
    s.p = v.p; s.pError = v.pError; s.n = v.n;
    return s;
}

// Mutual visibility of two surface vertices (VisibilityTester between their interactions).
LTS_D bool visible(const SceneView& sv, const BVertex& a, const BVertex& b, TraceStack ts) {
    LightPoint lp;
    lp.p = b.p; lp.n = b.n; lp.pError = b.pError;
    return unoccluded(sv, as_point(a), lp, ts);
}

// A surface vertex from a closest hit.
LTS_D void surface_vertex(const SceneView& sv, BVertex& v, const SurfacePoint& si, const int* perm) {
    v.kind = kVSurface;
    v.p = si.p; v.n = si.n; v.sn = si.sn; v.pError = si.pError; v.wo = si.wo;
    v.light = sv.shade[si.leaf].light;
    v.delta = false;
    v.bsdf = make_bsdf(sv, si, perm);
}

// Random walk from `ray` (PBRT-v3 RandomWalk): appends vertices path[k0..] until maxVerts, a miss
// or an absorbed sample. pdfDir is the solid-angle density the ray's direction was sampled with.
// Camera walks record an escape as a vertex of the infinite light.
LTS_D int walk(const SceneView& sv, Ray ray, Spec beta, float pdfDir, bool importance, BVertex* path, int k0,
               int maxVerts, const Sampler& smp, uint32_t block, TraceStack ts, const int* perm, const CamRays* cr) {
    int k = k0;
    while (k < maxVerts) {
        Hit h = trace_closest(sv, ray, ts, nullptr, nullptr);
        BVertex& v = path[k];
        BVertex& prev = path[k - 1];
        if (h.leaf < 0) {
            if (!importance) {
                v.kind = kVEscape;
                v.beta = beta;
                v.p = ray.o + ray.d;  // only its direction from prev is used
                v.n = v.sn = v.pError = v.wo = v3(0, 0, 0);
                v.pdfFwd = v.pdfRev = 0.0f;
                v.light = -1;
                v.delta = false;
                v.bsdf.kind = -1;
                k++;
            }
            break;
        }
        SurfacePoint si = make_surface(sv, h.leaf, ray, h.u, h.v, h.invDet);
        if (cr && k == 1) differentials(si, *cr, uv_derivative_scale(sv, h.leaf, h.invDet));
        surface_vertex(sv, v, si, perm);
        v.beta = beta;
        v.pdfFwd = to_area(pdfDir, prev, v);
        v.pdfRev = 0.0f;
        if (++k >= maxVerts) break;
        float u[4];
        smp.block(block + (uint32_t)k, u);
        V3 wi;
        float pdf = 0.0f;
        int type = 0;
        Spec f = v.bsdf.sampleF(v.wo, &wi, v2(u[0], u[1]), &pdf, B_ALL, &type);
        if (pdf == 0 || isBlack(f)) break;
        if (importance && v.bsdf.kind == 3 && (type & B_TRANS)) {
            // sampleF scales refracted radiance by (etaI / etaT)^2; importance is not scaled
            const bool entering = cosT(v.bsdf.toLocal(v.wo)) > 0;
            const float etaI = entering ? 1.0f : v.bsdf.eta, etaT = entering ? v.bsdf.eta : 1.0f;
            f = f * ((etaT * etaT) / (etaI * etaI));
        }
        beta = beta * (f * absDot(wi, v.sn) / pdf) * shading_correction(v, v.wo, wi, importance);
        float pdfRev = v.bsdf.pdf(wi, v.wo, B_ALL);
        if (type & B_SPEC) {
            v.delta = true;
            pdf = pdfRev = 0.0f;
        }
        prev.pdfRev = to_area(pdfRev, v, prev);
        pdfDir = pdf;
        ray = spawn_ray(si, wi);
    }
    return k;
}

// A point on area light li with its geometric normal as the hit test computes it (make_surface at
// the point's barycentrics), from the reference's triangle sampling (Triangle.cu:167-190).
LTS_D SurfacePoint light_point(const SceneView& sv, const LightDesc& L, V2 u) {
    const float t = sqrtf(u.x);
    const float b0 = 1 - t, b1 = u.y * t;
    Ray dummy;
    dummy.o = v3(0, 0, 0); dummy.d = v3(0, 0, 1); dummy.tMax = INFINITY;
    return make_surface(sv, L.leaf, dummy, b1, 1 - b0 - b1, 1.0f);
}

LTS_D int light_subpath(const SceneView& sv, BVertex* path, const Sampler& smp, int maxVerts, TraceStack ts,
                        const int* perm) {
    if (sv.nLights == 0 || maxVerts < 1) return 0;
    float u0[4], u1[4];
    smp.block(kBlockLightStart, u0);
    smp.block(kBlockLightStart + 1, u1);
    float unused;
    const int li = sample_light(sv, u0[0], &unused);
    const LightDesc& L = sv.lights[li];
    if (L.kind != 0) return 0;
    const float pick = pick_prob(sv, li);
    if (pick == 0 || L.area == 0) return 0;
    SurfacePoint lp = light_point(sv, L, v2(u0[1], u0[2]));
    V3 a, b;
    coordinateSystem(lp.n, &a, &b);
    const V3 wl = cosine_hemisphere(v2(u0[3], u1[0]));
    const V3 w = a * wl.x + b * wl.y + lp.n * wl.z;
    const float pdfDir = wl.z * INV_PI_F;
    if (pdfDir == 0) return 0;
    BVertex& v = path[0];
    v.kind = kVLight;
    v.p = lp.p; v.n = v.sn = lp.n; v.pError = lp.pError; v.wo = v3(0, 0, 0);
    v.light = li;
    v.delta = false;
    v.bsdf.kind = -1;
    v.beta = light_le(L);
    v.pdfFwd = pick / L.area;
    v.pdfRev = 0.0f;
    // Le |cos| / (pick pdfPos pdfDir), with pdfPos = 1 / area and pdfDir = cos / pi
    const Spec beta = light_le(L) * (PI_F * L.area / pick);
    Ray r;
    r.o = offsetRayOrigin(lp.p, lp.pError, lp.n, w);
    r.d = w;
    r.tMax = INFINITY;
    return walk(sv, r, beta, pdfDir, true, path, 1, maxVerts, smp, kBlockLightWalk, ts, perm, nullptr);
}

LTS_D float remap0(float f) { return f != 0 ? f : 1.0f; }

// Balance-heuristic weight of strategy (s, t) over the strategies with t' >= 2 (PBRT-v3 MISWeight,
// with the densities a connection changes computed into locals instead of written into the
// subpaths). `sampled` is the resampled light vertex of an s = 1 connection.
LTS_D float mis_weight(const SceneView& sv, const BVertex* lp, const BVertex* cp, const BVertex& sampled, int s, int t) {
    if (s + t == 2) return 1.0f;
    const BVertex& pt = cp[t - 1];
    const BVertex& ptm = cp[t - 2];
    if (pt.kind == kVEscape) return 1.0f;
    const BVertex* qs = s == 1 ? &sampled : s > 1 ? &lp[s - 1] : nullptr;
    const BVertex* qsm = s > 1 ? &lp[s - 2] : nullptr;
    // reverse densities of pt, ptm, qs, qsm under this connection
    const float ptRev = s > 0 ? vertex_pdf(*qs, qsm, pt) : origin_pdf(sv, pt);
    if (s == 0 && ptRev == 0) return 1.0f;  // no light-subpath strategy can start at this emitter
    const float ptmRev = s > 0 ? vertex_pdf(pt, qs, ptm) : emit_pdf(pt, ptm);
    const float qsRev = s > 0 ? vertex_pdf(pt, &ptm, *qs) : 0.0f;
    const float qsmRev = s > 1 ? vertex_pdf(*qs, &pt, *qsm) : 0.0f;
    float sum = 0.0f, ri = 1.0f;
    for (int i = t - 1; i > 1; i--) {
        const float rev = i == t - 1 ? ptRev : i == t - 2 ? ptmRev : cp[i].pdfRev;
        const bool d = i == t - 1 ? false : cp[i].delta;
        const bool dPrev = i - 1 == t - 1 ? false : cp[i - 1].delta;
        ri *= remap0(rev) / remap0(cp[i].pdfFwd);
        if (!d && !dPrev) sum += ri;
    }
    ri = 1.0f;
    for (int j = s - 1; j >= 0; j--) {
        const BVertex& v = j == s - 1 ? *qs : lp[j];
        const float rev = j == s - 1 ? qsRev : j == s - 2 ? qsmRev : lp[j].pdfRev;
        const bool d = j == s - 1 ? false : v.delta;
        const bool dPrev = j > 0 && (j - 1 == s - 1 ? false : lp[j - 1].delta);
        ri *= remap0(rev) / remap0(v.pdfFwd);
        if (!d && !dPrev) sum += ri;
    }
    return 1.0f / (1.0f + sum);
}

// Contribution of strategy (s, t), MIS-weighted (PBRT-v3 ConnectBDPT without t = 1).
LTS_D Spec connect(const SceneView& sv, const BVertex* lp, const BVertex* cp, int s, int t, const Sampler& smp,
                   TraceStack ts) {
    const BVertex& pt = cp[t - 1];
    Spec L = spec(0.0f);
    BVertex sampled;
    sampled.kind = -1;
    if (s == 0) {
        if (pt.kind == kVEscape) {
            for (int l = 0; l < sv.nLights; l++)
                if (sv.lights[l].kind == 1) L = L + light_le(sv.lights[l]);
            L = L * pt.beta;
        } else if (pt.kind == kVSurface) {
            const BVertex& ptm = cp[t - 2];
            V3 w = ptm.p - pt.p;
            L = emitted(sv, pt, normalize(w)) * pt.beta;
        }
    } else if (pt.kind == kVEscape || !connectible(pt)) {
        return spec(0.0f);
    } else if (s == 1) {
        float u[4];
        smp.block(kBlockConnect + (uint32_t)t, u);
        float unused;
        const int li = sample_light(sv, u[0], &unused);
        const LightDesc& Ld = sv.lights[li];
        const float pick = pick_prob(sv, li);
        if (Ld.kind != 0 || pick == 0 || Ld.area == 0) return spec(0.0f);
        SurfacePoint lpnt = light_point(sv, Ld, v2(u[1], u[2]));
        sampled.kind = kVLight;
        sampled.p = lpnt.p; sampled.n = sampled.sn = lpnt.n; sampled.pError = lpnt.pError;
        sampled.wo = v3(0, 0, 0);
        sampled.light = li;
        sampled.delta = false;
        sampled.bsdf.kind = -1;
        sampled.pdfFwd = pick / Ld.area;
        sampled.pdfRev = 0.0f;
        V3 d = sampled.p - pt.p;
        const float d2 = lengthSq(d);
        if (d2 == 0) return spec(0.0f);
        d = d * (1 / sqrtf(d2));
        const Spec Le = emitted(sv, sampled, -d);
        if (isBlack(Le)) return spec(0.0f);
        const float G = absDot(pt.sn, d) * absDot(sampled.n, d) / d2;
        L = pt.beta * vertex_f(pt, sampled, false) * Le * (G / sampled.pdfFwd);
        if (!isBlack(L) && !visible(sv, pt, sampled, ts)) L = spec(0.0f);
    } else {
        const BVertex& qs = lp[s - 1];
        if (!connectible(qs)) return spec(0.0f);
        V3 d = qs.p - pt.p;
        const float d2 = lengthSq(d);
        if (d2 == 0) return spec(0.0f);
        d = d * (1 / sqrtf(d2));
        L = qs.beta * vertex_f(qs, pt, true) * vertex_f(pt, qs, false) * pt.beta;
        if (isBlack(L)) return L;
        const float G = absDot(pt.sn, d) * absDot(qs.sn, d) / d2;
        L = L * G;
        if (!visible(sv, pt, qs, ts)) L = spec(0.0f);
    }
    if (isBlack(L)) return L;
    return L * mis_weight(sv, lp, cp, sampled, s, t);
}

LTS_D Spec bdpt_radiance(const SceneView& sv, int px, int py, int sample, const RenderLaunch& rl, TraceStack ts,
                         const int* perm) {
    Sampler smp;
    smp.seed = rl.seed;
    smp.pixel = (uint32_t)(py * sv.cam.width + px);
    smp.sample = (uint32_t)sample;
    const int maxBounce = rl.maxBounce;
    BVertex cp[kCamVerts], lp[kLightVerts];
    CamRays cr = camera_sample(sv, px, py, sample, rl.spp, rl.stratified, rl.seed);
    BVertex& c0 = cp[0];
    c0.kind = kVCamera;
    c0.p = cr.r.o;
    c0.n = c0.sn = c0.pError = c0.wo = v3(0, 0, 0);
    c0.beta = spec(1.0f);
    c0.pdfFwd = c0.pdfRev = 0.0f;
    c0.light = -1;
    c0.delta = false;
    c0.bsdf.kind = -1;
    // pdfDir of the camera ray only feeds the t = 1 strategies, which are not used
    const int nCam = walk(sv, cr.r, spec(1.0f), 0.0f, false, cp, 1, maxBounce + 2, smp, kBlockCamWalk, ts, perm, &cr);
    const int nLight = light_subpath(sv, lp, smp, maxBounce + 1, ts, perm);
    Spec L = spec(0.0f);
    for (int t = 2; t <= nCam; t++)
        for (int s = 0; s <= nLight; s++) {
            if (s + t - 2 > maxBounce) break;
            L = L + connect(sv, lp, cp, s, t, smp, ts);
        }
    return L;
}

// Four waves per SIMD (128 VGPRs): 32.4 ms per 4-spp frame of the 1M-triangle displaced sphere at
// 1920x1080, against 38.0 ms at the three waves its natural 165 VGPRs allow and 34.3 ms at five.
__global__ __launch_bounds__(256, 4) void k_bdpt(SceneView sv, RenderLaunch rl) {
    __shared__ int stk[kShortStack * 256];
    __shared__ int perm[512];
    for (int i = threadIdx.x; i < 512; i += 256) perm[i] = sv.perm[i];
    __syncthreads();
    const int W = sv.cam.width, H = sv.cam.height;
    const int tilesX = (W + kTileSide - 1) / kTileSide;
    const int tile = blockIdx.x;
    const int px = (tile % tilesX) * kTileSide + (threadIdx.x & (kTileSide - 1));
    const int py = (tile / tilesX) * kTileSide + (threadIdx.x / kTileSide);
    const int sample = rl.s0 + blockIdx.y;
    if (px >= W || py >= H || sample >= rl.s1) return;
    Spec L = bdpt_radiance(sv, px, py, sample, rl, lds_stack(stk, threadIdx.x, 256), perm);
    float* o = rl.sampleL + 3 * ((size_t)blockIdx.y * W * H + (size_t)py * W + px);
    o[0] = L.r; o[1] = L.g; o[2] = L.b;
}

}  // namespace

hipError_t launch_bdpt(const SceneView& sv, const RenderLaunch& rl, hipStream_t stream) {
    if (rl.maxBounce < 0 || rl.maxBounce > kBdptMaxBounce) return hipErrorInvalidValue;
    const int W = sv.cam.width, H = sv.cam.height;
    dim3 grid(((W + kTileSide - 1) / kTileSide) * ((H + kTileSide - 1) / kTileSide), rl.s1 - rl.s0);
    k_bdpt<<<grid, 256, 0, stream>>>(sv, rl);
    return hipGetLastError();
}

}  // namespace lts
