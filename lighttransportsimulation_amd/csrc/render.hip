// Plain path tracer for gfx950: one thread per (pixel, sample), the reference integrator
// (intergrators/PathTracingIntegrator.cuh:26-109) per thread, followed by a deterministic film
// gather (rendering/Camera.cu:192-218 without atomics, then getRender, Camera.cu:256-289).
//
// Launch shape: blocks of 256 threads cover a 16x16 pixel tile for one sample index, so each
// 64-wide wavefront traces 4 x 16 neighbouring pixels; grid.y walks the samples of the pass.
#include "pt_device.h"
#include "lts_internal.h"

namespace lts {

namespace {

constexpr int kTileSide = 16;

// One camera sample of PathTracingIntegrator::evaluate.
LTS_D Spec radiance(const SceneView& sv, int px, int py, int sample, const RenderLaunch& rl, unsigned long long* cnt,
                     TraceStack ts, const int* perm) {
    Sampler smp;
    smp.seed = rl.seed;
    smp.pixel = (uint32_t)(py * sv.cam.width + px);
    smp.sample = (uint32_t)sample;
    CamRays cr = camera_sample(sv, px, py, sample, rl.spp, rl.stratified, rl.seed);
    Ray r = cr.r;
    Spec L = spec(0.0f), beta = spec(1.0f);
    bool specularBounce = false;
    for (int bounce = 0; bounce < rl.maxBounce; bounce++) {
        Hit h = trace_closest(sv, r, ts, cnt, cnt ? cnt + 1 : nullptr);
        if (bounce == 0 || specularBounce) {
            if (h.leaf >= 0) {
                int l = sv.shade[h.leaf].light;
                // the geometric normal decides emission; it needs the full interaction
                if (l >= 0) {
                    SurfacePoint e = make_surface(sv, h.leaf, r, h.u, h.v, h.invDet);
                    Spec le = dot(e.n, -r.d) > 0.f ? light_le(sv.lights[l]) : spec(0.0f);
                    L = L + beta * le;
                } else {
                    L = L + beta * spec(0.0f);
                }
            } else {
                for (int l = 0; l < sv.nLights; l++) L = L + beta * (sv.lights[l].kind == 1 ? light_le(sv.lights[l]) : spec(0.0f));
            }
        }
        if (h.leaf < 0) break;
        SurfacePoint si = make_surface(sv, h.leaf, r, h.u, h.v, h.invDet);
        if (bounce == 0) differentials(si, cr, uv_derivative_scale(sv, h.leaf, h.invDet));
        Bsdf bsdf = make_bsdf(sv, si, perm);
        float u0[4], u1[4];
        smp.block(1 + 2 * bounce, u0);
        smp.block(2 + 2 * bounce, u1);
        if (sv.nLights > 0) {
            float lpdf;
            int ln = sample_light(sv, u0[0], &lpdf);
            Spec ld = estimate_direct(sv, si, bsdf, v2(u0[3], u1[0]), ln, v2(u0[1], u0[2]), specularBounce, ts);
            L = L + beta * (ld / lpdf);
        }
        V3 wo = -r.d, wi;
        float pdf;
        int flags = 0;
        Spec f = bsdf.sampleF(wo, &wi, v2(u1[1], u1[2]), &pdf, B_ALL, &flags);
        if (isBlack(f) || pdf == 0) break;
        specularBounce = (flags & B_SPEC) != 0;
        beta = beta * (f * absDot(wi, si.sn) / pdf);
        r = spawn_ray(si, wi);
        if (bounce >= rl.rrStart) {
            float q = fmaxf(0.05f, 1 - lum(beta));
            if (u1[3] < q) break;
            beta = beta / (1 - q);
        }
    }
    return L;
}

__global__ __launch_bounds__(256) void k_tri_materials(TriLeaf* __restrict__ tris, const ShadeTri* __restrict__ shade, int n) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k < n) tris[k].mat = shade[k].mat;
}

// Latency-bound: five waves per SIMD (the compiler spills the rest of the path state to scratch)
// measured 696 Msamples/s on the 1080p workload against 679 at four, 647 at three and 470 at the two
// waves its natural register count allows.
__global__ __launch_bounds__(256, 5) void k_megakernel(SceneView sv, RenderLaunch rl) {
    __shared__ int stk[kShortStack * 256];
    __shared__ int perm[512];  // the procedural textures' permutation lookups chain three deep
    for (int i = threadIdx.x; i < 512; i += 256) perm[i] = sv.perm[i];
    __syncthreads();
    const int W = sv.cam.width, H = sv.cam.height;
    const int tilesX = (W + kTileSide - 1) / kTileSide;
    // Workgroups are dealt to the 8 XCDs round robin in dispatch order. Tiles stay dealt round robin
    // (every XCD gets every eighth tile, which balances sky against geometry), but an XCD runs all
    // samples of a tile back to back, so the blocks it holds at a time cover a few tiles.
    const int dispatch = blockIdx.y * gridDim.x + blockIdx.x;
    int tile = blockIdx.x, sy = blockIdx.y;
    if (gridDim.x % 8 == 0) {
        const int i = dispatch / 8;  // the XCD's i-th block
        tile = dispatch % 8 + 8 * (i / gridDim.y);
        sy = i % gridDim.y;
    }
    const int px = (tile % tilesX) * kTileSide + (threadIdx.x & (kTileSide - 1));
    const int py = (tile / tilesX) * kTileSide + (threadIdx.x / kTileSide);
    const int sample = rl.s0 + sy;
    if (px >= W || py >= H || sample >= rl.s1) return;
    Spec L = radiance(sv, px, py, sample, rl, rl.counters, lds_stack(stk, threadIdx.x, 256), perm);
    float* o = rl.sampleL + 3 * ((size_t)sy * W * H + (size_t)py * W + px);
    o[0] = L.r; o[1] = L.g; o[2] = L.b;
}

// Pixel (x, y) gathers, for each sample in order and each owner pixel (ox, oy) with ox in
// x-2..x+1 and oy in y-2..y+1 (row major), the owner's filtered contribution when its addSample
// footprint [ceil(p - 1), floor(p + 1) + 1) covers (x, y). A film position p = owner + f with f in
// (0, 1) covers owner..owner+1; f = 1 adds owner+2; and p rounding down to the owner itself (f
// below half an ulp of the owner coordinate, possible from 1024 pixels on) adds owner-1, which is
// why the window reaches one owner past the pixel. A 16x16 block stages, per sample, the film
// positions and radiance of the 19x19 owners its pixels read in LDS, so each owner's film position
// is drawn once per block instead of once per covered pixel; the sums keep their order.
__global__ __launch_bounds__(256) void k_film_gather(SceneView sv, RenderLaunch rl) {
    constexpr int T = 16, R = T + 3;
    __shared__ float own[R * R][5];  // film position x, y and radiance r, g, b of each owner
    const int W = sv.cam.width, H = sv.cam.height;
    const int x0 = blockIdx.x * T, y0 = blockIdx.y * T;
    const int x = x0 + (threadIdx.x & 15), y = y0 + (threadIdx.x >> 4);
    const bool inside = x < W && y < H;
    float4 acc = inside ? reinterpret_cast<float4*>(rl.film)[(size_t)y * W + x] : make_float4(0.f, 0.f, 0.f, 0.f);
    Spec rgb = spec(acc.x, acc.y, acc.z);
    float wsum = acc.w;
    for (int s = rl.s0; s < rl.s1; s++) {
        __syncthreads();  // the previous sample's owners have been read
        for (int e = threadIdx.x; e < R * R; e += 256) {
            const int ox = x0 - 2 + e % R, oy = y0 - 2 + e / R;
            if (ox < 0 || oy < 0 || ox >= W || oy >= H) continue;
            Sampler smp;
            smp.seed = rl.seed; smp.pixel = (uint32_t)(oy * W + ox); smp.sample = (uint32_t)s;
            float d[4];
            smp.block(0, d);
            const V2 pf = film_sample(ox, oy, s, rl.spp, rl.stratified, d);
            const float* l = rl.sampleL + 3 * ((size_t)(s - rl.s0) * W * H + (size_t)oy * W + ox);
            own[e][0] = pf.x; own[e][1] = pf.y;
            own[e][2] = l[0]; own[e][3] = l[1]; own[e][4] = l[2];
        }
        __syncthreads();
        if (!inside) continue;
        for (int oy = y - 2; oy <= y + 1; oy++) {
            if (oy < 0 || oy >= H) continue;
            for (int ox = x - 2; ox <= x + 1; ox++) {
                if (ox < 0 || ox >= W) continue;
                const float* o = own[(oy - y0 + 2) * R + (ox - x0 + 2)];
                const V2 pf = v2(o[0], o[1]);
                int p0x = max((int)ceilf(pf.x - 1.0f), 0), p0y = max((int)ceilf(pf.y - 1.0f), 0);
                int p1x = min((int)floorf(pf.x + 1.0f) + 1, W), p1y = min((int)floorf(pf.y + 1.0f) + 1, H);
                if (x < p0x || x >= p1x || y < p0y || y >= p1y) continue;
                float dx = fabsf((x - pf.x - 0.5f) * 1.0f * 16);
                float dy = fabsf((y - pf.y - 0.5f) * 1.0f * 16);
                int ix = (int)fminf((float)(int)floorf(dx), 15.0f);
                int iy = (int)fminf((float)(int)floorf(dy), 15.0f);
                float w = sv.filter[iy * 16 + ix];
                rgb = rgb + spec(o[2], o[3], o[4]) * w;
                wsum += w;
            }
        }
    }
    if (inside) reinterpret_cast<float4*>(rl.film)[(size_t)y * W + x] = make_float4(rgb.r, rgb.g, rgb.b, wsum);
}

__global__ void k_film_clear(float4* film, int npix) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < npix) film[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ void k_get_render(const float4* film, float* rgb, int W, int H) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= W * H) return;
    int j = i / W, c = i % W;
    float4 p = film[(size_t)(H - 1 - j) * W + c];
    float inv = 1 / p.w;
    rgb[3 * (size_t)i] = fmaxf(0.0f, p.x * inv);
    rgb[3 * (size_t)i + 1] = fmaxf(0.0f, p.y * inv);
    rgb[3 * (size_t)i + 2] = fmaxf(0.0f, p.z * inv);
}

}  // namespace

hipError_t launch_megakernel(const SceneView& sv, const RenderLaunch& rl, hipStream_t stream) {
    const int W = sv.cam.width, H = sv.cam.height;
    dim3 grid(((W + kTileSide - 1) / kTileSide) * ((H + kTileSide - 1) / kTileSide), rl.s1 - rl.s0);
    k_megakernel<<<grid, 256, 0, stream>>>(sv, rl);
    return hipGetLastError();
}

hipError_t launch_film_gather(const SceneView& sv, const RenderLaunch& rl, hipStream_t stream) {
    const int W = sv.cam.width, H = sv.cam.height;
    dim3 grid((W + 15) / 16, (H + 15) / 16);
    k_film_gather<<<grid, 256, 0, stream>>>(sv, rl);
    return hipGetLastError();
}

hipError_t launch_tri_materials(TriLeaf* tris, const ShadeTri* shade, int n, hipStream_t s) {
    if (n < 1) return hipSuccess;
    k_tri_materials<<<(n + 255) / 256, 256, 0, s>>>(tris, shade, n);
    return hipGetLastError();
}

// Closest hits of caller-given rays: the renderer's own traversal (trace_closest), one thread per
// ray. rays[6 i ..] = origin, direction; a miss gives prim -1.
__global__ __launch_bounds__(256) void k_trace_rays(SceneView sv, const float* __restrict__ rays, int n,
                                                    int32_t* __restrict__ prim, float* __restrict__ tOut) {
    __shared__ int stk[kShortStack * 256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    Ray r;
    r.o = v3(rays[6 * (size_t)i], rays[6 * (size_t)i + 1], rays[6 * (size_t)i + 2]);
    r.d = v3(rays[6 * (size_t)i + 3], rays[6 * (size_t)i + 4], rays[6 * (size_t)i + 5]);
    r.tMax = INFINITY;
    const Hit h = trace_closest(sv, r, lds_stack(stk, threadIdx.x, 256), nullptr, nullptr);
    int p = -1;
    if (h.leaf >= 0) {
        V3 p0, p1, p2;
        load_tri(sv.tris, h.leaf, p0, p1, p2, p);
    }
    prim[i] = p;
    tOut[i] = h.leaf >= 0 ? h.t : INFINITY;
}

hipError_t launch_trace_rays(const SceneView& sv, const float* rays, int n, int32_t* prim, float* t, hipStream_t stream) {
    if (n < 1) return hipSuccess;
    k_trace_rays<<<(n + 255) / 256, 256, 0, stream>>>(sv, rays, n, prim, t);
    return hipGetLastError();
}

hipError_t launch_film_clear(float* film, int npix, hipStream_t stream) {
    k_film_clear<<<(npix + 255) / 256, 256, 0, stream>>>(reinterpret_cast<float4*>(film), npix);
    return hipGetLastError();
}

hipError_t launch_get_render(const SceneView& sv, const float* film, float* rgb, hipStream_t stream) {
    const int W = sv.cam.width, H = sv.cam.height;
    k_get_render<<<(W * H + 255) / 256, 256, 0, stream>>>(reinterpret_cast<const float4*>(film), rgb, W, H);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_copy(float4* __restrict__ dst, const float4* __restrict__ src, size_t n4) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}

hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t stream) {
    size_t n4 = bytes / 16;
    k_copy<<<2048, 256, 0, stream>>>(reinterpret_cast<float4*>(dst), reinterpret_cast<const float4*>(src), n4);
    return hipGetLastError();
}



}  // namespace lts
