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

struct CamRays {
    Ray r;
    V3 rxO, rxD, ryO, ryD;
};

// Camera::generateRay / generateDifferentialRay + Ray::scaleDifferentials (Camera.cu:51-81, Ray.cuh:35-40)
LTS_D void camera_ray(const CameraDesc& c, V2 pf, V2 lens, V3* o, V3* d) {
    V3 cu = v3(c.u[0], c.u[1], c.u[2]), cv = v3(c.v[0], c.v[1], c.v[2]);
    V3 co = v3(c.o[0], c.o[1], c.o[2]);
    V3 offset = (cu * lens.x + cv * lens.y) * c.lensRadius;
    V2 pr = v2(pf.x / (c.width - 1), pf.y / (c.height - 1));
    *o = co + offset;
    *d = normalize(v3(c.llc[0], c.llc[1], c.llc[2]) + v3(c.right[0], c.right[1], c.right[2]) * pr.x +
                   v3(c.up[0], c.up[1], c.up[2]) * pr.y - co - offset);
}

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

// SurfaceInteraction::computeDifferentials, dpdx/dpdy part (Interaction.cuh:118-133)
LTS_D void differentials(SurfacePoint& s, const CamRays& cr) {
    float d = -dot(s.n, s.p);
    float tx = (-dot(s.n, cr.rxO) - d) / dot(s.n, cr.rxD);
    V3 px = cr.rxO + cr.rxD * tx;
    float ty = (-dot(s.n, cr.ryO) - d) / dot(s.n, cr.ryD);
    V3 py = cr.ryO + cr.ryD * ty;
    s.dpdx = px - s.p;
    s.dpdy = py - s.p;
}

// One camera sample of PathTracingIntegrator::evaluate.
LTS_D Spec radiance(const SceneView& sv, int px, int py, int sample, const RenderLaunch& rl, unsigned long long* cnt) {
    Sampler smp;
    smp.seed = rl.seed;
    smp.pixel = (uint32_t)(py * sv.cam.width + px);
    smp.sample = (uint32_t)sample;
    float d[4];
    smp.block(0, d);
    V2 pFilm = film_sample(px, py, sample, rl.spp, rl.stratified, d);
    V2 lens = uniform_disk(v2(d[2], d[3]));
    CamRays cr = camera_rays(sv.cam, pFilm, lens, rl.spp);
    Ray r = cr.r;
    Spec L = spec(0.0f), beta = spec(1.0f);
    bool specularBounce = false;
    for (int bounce = 0; bounce < rl.maxBounce; bounce++) {
        Hit h = trace_closest(sv, r, cnt, cnt ? cnt + 1 : nullptr);
        if (bounce == 0 || specularBounce) {
            if (h.leaf >= 0) {
                int prim = sv.tris[h.leaf].prim;
                int l = sv.shade[prim].light;
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
        if (bounce == 0) differentials(si, cr);
        Bsdf bsdf = make_bsdf(sv, si);
        float u0[4], u1[4];
        smp.block(1 + 2 * bounce, u0);
        smp.block(2 + 2 * bounce, u1);
        if (sv.nLights > 0) {
            float lpdf;
            int ln = sample_light(sv, u0[0], &lpdf);
            Spec ld = estimate_direct(sv, si, bsdf, v2(u0[3], u1[0]), ln, v2(u0[1], u0[2]), specularBounce);
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

__global__ __launch_bounds__(256) void k_megakernel(SceneView sv, RenderLaunch rl) {
    const int W = sv.cam.width, H = sv.cam.height;
    const int tilesX = (W + kTileSide - 1) / kTileSide;
    const int px = (blockIdx.x % tilesX) * kTileSide + (threadIdx.x & (kTileSide - 1));
    const int py = (blockIdx.x / tilesX) * kTileSide + (threadIdx.x / kTileSide);
    const int sample = rl.s0 + blockIdx.y;
    if (px >= W || py >= H || sample >= rl.s1) return;
    Spec L = radiance(sv, px, py, sample, rl, rl.counters);
    float* o = rl.sampleL + 3 * ((size_t)blockIdx.y * W * H + (size_t)py * W + px);
    o[0] = L.r; o[1] = L.g; o[2] = L.b;
}

// Pixel (x, y) gathers, for each sample in order and each owner pixel of the 3x3 window above
// and left of it (row major), the owner's filtered contribution when its addSample footprint
// covers (x, y).
__global__ __launch_bounds__(256) void k_film_gather(SceneView sv, RenderLaunch rl) {
    const int W = sv.cam.width, H = sv.cam.height;
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= W || y >= H) return;
    float4 acc = reinterpret_cast<float4*>(rl.film)[(size_t)y * W + x];
    Spec rgb = spec(acc.x, acc.y, acc.z);
    float wsum = acc.w;
    for (int s = rl.s0; s < rl.s1; s++) {
        for (int oy = y - 2; oy <= y; oy++) {
            if (oy < 0) continue;
            for (int ox = x - 2; ox <= x; ox++) {
                if (ox < 0) continue;
                Sampler smp;
                smp.seed = rl.seed; smp.pixel = (uint32_t)(oy * W + ox); smp.sample = (uint32_t)s;
                float d[4];
                smp.block(0, d);
                V2 pf = film_sample(ox, oy, s, rl.spp, rl.stratified, d);
                int p0x = max((int)ceilf(pf.x - 1.0f), 0), p0y = max((int)ceilf(pf.y - 1.0f), 0);
                int p1x = min((int)floorf(pf.x + 1.0f) + 1, W), p1y = min((int)floorf(pf.y + 1.0f) + 1, H);
                if (x < p0x || x >= p1x || y < p0y || y >= p1y) continue;
                float dx = fabsf((x - pf.x - 0.5f) * 1.0f * 16);
                float dy = fabsf((y - pf.y - 0.5f) * 1.0f * 16);
                int ix = (int)fminf((float)(int)floorf(dx), 15.0f);
                int iy = (int)fminf((float)(int)floorf(dy), 15.0f);
                float w = sv.filter[iy * 16 + ix];
                const float* l = rl.sampleL + 3 * ((size_t)(s - rl.s0) * W * H + (size_t)oy * W + ox);
                rgb = rgb + spec(l[0], l[1], l[2]) * w;
                wsum += w;
            }
        }
    }
    reinterpret_cast<float4*>(rl.film)[(size_t)y * W + x] = make_float4(rgb.r, rgb.g, rgb.b, wsum);
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

hipError_t launch_film_clear(float* film, int npix, hipStream_t stream) {
    k_film_clear<<<(npix + 255) / 256, 256, 0, stream>>>(reinterpret_cast<float4*>(film), npix);
    return hipGetLastError();
}

hipError_t launch_get_render(const SceneView& sv, const float* film, float* rgb, hipStream_t stream) {
    const int W = sv.cam.width, H = sv.cam.height;
    k_get_render<<<(W * H + 255) / 256, 256, 0, stream>>>(reinterpret_cast<const float4*>(film), rgb, W, H);
    return hipGetLastError();
}

hipError_t upload_perlin_table(const int* table512) {
    return hipMemcpyToSymbol(HIP_SYMBOL(c_perm), table512, 512 * sizeof(int));
}

}  // namespace lts
