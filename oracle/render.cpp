// ORACLE — test infrastructure only (see ref.h).
//
// CPU restatement of the reference's path tracer, one function per reference routine:
//   PathTracingIntegrator::evaluate .......... intergrators/PathTracingIntegrator.cuh:26-109
//   uniformSampleOneLight / estimateDirect ... intergrators/IntegratorHelper.cuh:51-149
//   BVH traversal (closest / any hit) ........ acceleration/BVH.cuh:331-398, geometry/Bound.cuh:200-227
//   Triangle intersect / sample / Pdf ........ shapes/Triangle.cu:50-259
//   SurfaceInteraction + differentials ....... geometry/Interaction.cuh:85-160
//   BSDF / BxDFs ............................. materials/BSDF.cuh, materials/models/*.cuh
//   textures ................................. materials/textures/{Texture,Procedural,ImageTexture}.cuh
//   lights ................................... lights/Light.cuh, lights/Light.cu
//   camera + film ............................ rendering/Camera.cu:51-81, :192-289
// Deliberate, documented deviations: the sampler (philox.h) and a deterministic film
// (per-sample radiance, then a fixed-order filter gather instead of float atomics and
// the racy weight add at Camera.cu:215).
#include "oracle.h"
#include "philox.h"
#include "scene.h"
#include <atomic>
#include <chrono>
#include <thread>

namespace orc {

struct Ray {
    V3 o, d;
    float tMax = INFINITY;
    bool hasDiff = false;
    V3 rxO, ryO, rxD, ryD;
};

struct SI {
    V3 p, pError, wo, n;
    V2 uv;
    V3 dpdu, dpdv;
    V3 sn, sdpdu;
    V3 dpdx, dpdy;
    float dudx = 0, dvdx = 0, dudy = 0, dvdy = 0;
    float uvScale = 1;  // invDet * uv determinant: dpdu, dpdv are uvScale times the true derivatives
    int tri = -1;
};

struct Counters { int64_t nodes = 0, tris = 0, rays = 0; };

static inline void triVerts(const Scene& s, int g, V3& p0, V3& p1, V3& p2, const int** idOut) {
    const Mesh& M = s.meshes[s.triMesh[g]];
    const int* id = &M.idx[9 * (size_t)s.triLocal[g]];
    p0 = M.P[id[0]]; p1 = M.P[id[3]]; p2 = M.P[id[6]];
    if (idOut) *idOut = id;
}

// Triangle::intersect, hit test part (Triangle.cu:94-128)
static bool triHit(const Scene& s, int g, const Ray& r, float& tOut, float& uOut, float& vOut, float& invDetOut) {
    V3 p0, p1, p2;
    triVerts(s, g, p0, p1, p2, nullptr);
    V3 edge1 = p1 - p0, edge2 = p2 - p0;
    V3 pVec = cross(r.d, edge2);
    float det = dot(edge1, pVec);
    if (det == 0) return false;
    float invDet = (float)(1.0 / (double)det);
    V3 tVec = r.o - p0;
    float u = invDet * dot(tVec, pVec);
    if (u < SHADOW_EPS || (double)u > 1.0 - (double)SHADOW_EPS) return false;
    V3 qVec = cross(tVec, edge1);
    float v = invDet * dot(r.d, qVec);
    if (v < SHADOW_EPS || (double)(u + v) > 1.0 - (double)SHADOW_EPS) return false;
    float t = invDet * dot(edge2, qVec);
    if (t > 0 && t < r.tMax) { tOut = t; uOut = u; vOut = v; invDetOut = invDet; return true; }
    return false;
}

// Triangle::simpleIntersect (Triangle.cu:50-84)
static bool triOccludes(const Scene& s, int g, const Ray& r) {
    V3 p0, p1, p2;
    triVerts(s, g, p0, p1, p2, nullptr);
    V3 edge1 = p1 - p0, edge2 = p2 - p0;
    V3 pVec = cross(r.d, edge1);
    float det = dot(edge2, pVec);
    if (det == 0) return false;
    float invDet = (float)(1.0 / (double)det);
    V3 tVec = r.o - p0;
    float u = invDet * dot(tVec, pVec);
    if (u < SHADOW_EPS || (double)u > 1.0 + (double)SHADOW_EPS) return false;
    V3 qVec = cross(tVec, edge2);
    float v = invDet * dot(r.d, qVec);
    if (v < SHADOW_EPS || (double)(u + v) > 1.0 + (double)SHADOW_EPS) return false;
    float t = invDet * dot(edge1, qVec);
    return t > 0 && t < r.tMax;
}

// Triangle::intersect, interaction part (Triangle.cu:130-200) and the SurfaceInteraction
// constructor / setShaddingGeometry (Interaction.cuh:85-116)
static SI triInteraction(const Scene& s, int g, const Ray& r, float u, float v, float invDet) {
    const Mesh& M = s.meshes[s.triMesh[g]];
    V3 p0, p1, p2;
    const int* id;
    triVerts(s, g, p0, p1, p2, &id);
    V3 edge1 = p1 - p0, edge2 = p2 - p0;
    float w = 1 - u - v;
    V2 uvs[3];
    if (!M.UV.empty()) { uvs[0] = M.UV[id[1]]; uvs[1] = M.UV[id[4]]; uvs[2] = M.UV[id[7]]; }
    else { uvs[0] = V2(0, 0); uvs[1] = V2(1, 0); uvs[2] = V2(1, 1); }
    V2 duv02 = uvs[0] - uvs[2], duv12 = uvs[1] - uvs[2];
    V3 dp02 = p0 - p2, dp12 = p1 - p2;
    float determinant = duv02.x * duv12.y - duv02.y * duv12.x;
    V3 dpdu, dpdv;
    if (determinant == 0) coordinateSystem(normalize(cross(edge2, edge1)), &dpdu, &dpdv);
    else {
        dpdu = (dp02 * duv12.y - dp12 * duv02.y) * invDet;
        dpdv = (dp02 * -duv12.x + dp12 * duv02.x) * invDet;
    }
    float xs = std::fabs(w * p0.x) + std::fabs(u * p1.x) + std::fabs(v * p2.x);
    float ys = std::fabs(w * p0.y) + std::fabs(u * p1.y) + std::fabs(v * p2.y);
    float zs = std::fabs(w * p0.z) + std::fabs(u * p1.z) + std::fabs(v * p2.z);
    SI si;
    // Triangle.cu:147-149 scales by the ray's invDet instead of 1 / determinant (Q11); the uv
    // differentials undo it (computeDifferentials)
    si.uvScale = determinant == 0 ? 1.0f : invDet * determinant;
    si.pError = V3(xs, ys, zs) * gammaN(7);
    si.p = p0 * w + p1 * u + p2 * v;
    si.uv = uvs[0] * w + uvs[1] * u + uvs[2] * v;
    si.wo = -r.d;
    si.dpdu = dpdu; si.dpdv = dpdv;
    si.n = normalize(cross(dpdu, dpdv));
    si.sn = si.n;
    si.sdpdu = dpdu;
    si.tri = g;
    if (!M.N.empty()) {
        V3 n0 = M.N[id[2]], n1 = M.N[id[5]], n2 = M.N[id[8]];
        V3 ns = normalize(n0 * w + n1 * u + n2 * v);
        V3 st = si.dpdu;
        V3 sb = cross(ns, st);
        if (lengthSq(sb) > 0.0f) { sb = normalize(sb); st = cross(sb, ns); }
        else coordinateSystem(ns, &st, &sb);
        si.sn = normalize(cross(st, sb));
        si.n = faceTowards(si.n, si.sn);
        si.sdpdu = st;
        si.n = faceTowards(si.n, si.sn);
    }
    return si;
}

// Bounds3::simpleIntersect (Bound.cuh:200-227)
static inline bool boxHit(const float* b, const Ray& r, V3 inv, const int neg[3]) {
    const float g3 = 1 + 2 * gammaN(3);
    float tMin = (b[3 * neg[0] + 0] - r.o.x) * inv.x;
    float tMax = (b[3 * (1 - neg[0]) + 0] - r.o.x) * inv.x;
    float tyMin = (b[3 * neg[1] + 1] - r.o.y) * inv.y;
    float tyMax = (b[3 * (1 - neg[1]) + 1] - r.o.y) * inv.y;
    tMax *= g3; tyMax *= g3;
    if (tMin > tyMax || tyMin > tMax) return false;
    if (tyMin > tMin) tMin = tyMin;
    if (tyMax < tMax) tMax = tyMax;
    float tzMin = (b[3 * neg[2] + 2] - r.o.z) * inv.z;
    float tzMax = (b[3 * (1 - neg[2]) + 2] - r.o.z) * inv.z;
    tzMax *= g3;
    if (tMin > tzMax || tzMin > tMax) return false;
    if (tzMin > tMin) tMin = tzMin;
    if (tzMax < tMax) tMax = tzMax;
    return (tMin < r.tMax) && (tMax > 0);
}

// BVH::intersect / simpleIntersect (BVH.cuh:331-398): DFS, right child pushed first.
static bool traverse(const Scene& s, Ray& r, bool anyHit, SI* si, Counters& c) {
    int n = s.nPrim;
    V3 inv((float)1 / r.d.x, (float)1 / r.d.y, (float)1 / r.d.z);
    int neg[3] = {inv.x < 0, inv.y < 0, inv.z < 0};
    int stack[128];
    int sp = 0;
    stack[sp++] = -1;
    int cur = 0;  // unified index: interior 0..n-2, leaf k = n-1+k
    bool hit = false;
    int hitTri = -1; float hu = 0, hv = 0, hInv = 0;
    c.rays++;
    do {
        bool leaf = cur >= n - 1;
        const float* b = leaf ? &s.lbox[6 * (size_t)(cur - (n - 1))] : &s.ibox[6 * (size_t)cur];
        c.nodes++;
        if (boxHit(b, r, inv, neg)) {
            if (leaf) {
                int g = s.leafPrim[cur - (n - 1)];
                c.tris++;
                if (anyHit) { if (triOccludes(s, g, r)) return true; }
                else {
                    float t, u, v, id;
                    if (triHit(s, g, r, t, u, v, id)) { r.tMax = t; hit = true; hitTri = g; hu = u; hv = v; hInv = id; }
                }
            } else {
                stack[sp++] = s.bvhRight[cur];
                stack[sp++] = s.bvhLeft[cur];
            }
        }
        cur = stack[--sp];
    } while (cur != -1);
    if (hit && si) *si = triInteraction(s, hitTri, r, hu, hv, hInv);
    return hit;
}

static Ray spawnRay(const SI& it, V3 d) {
    Ray r; r.o = offsetRayOrigin(it.p, it.pError, it.n, d); r.d = d; r.tMax = INFINITY; return r;
}

// ----------------------------------------------------------------- sampling helpers (Sampling.cuh)
static inline V2 uniformSampleDisk(V2 u) {
    float r = std::sqrt(u.x);
    float theta = 2 * PI_F * u.y;
    return V2(r * f_cos(theta), r * f_sin(theta));
}
static inline V3 cosineSampleHemisphere(V2 u) {
    V2 d = uniformSampleDisk(u);
    float z = std::sqrt(std::fmax(0.0f, 1 - d.x * d.x - d.y * d.y));
    return V3(d.x, d.y, z);
}
static inline V3 uniformSampleHemisphere(V2 u) {
    float z = u.x;
    float r = std::sqrt(std::fmax(0.0f, 1.0f - z * z));
    float phi = 2 * PI_F * u.y;
    return V3(r * f_cos(phi), r * f_sin(phi), z);
}
static inline float powerHeuristic(float fPdf, float gPdf) {
    float f = 1 * fPdf, g = 1 * gPdf;
    return (f * f) / (f * f + g * g);
}

// ----------------------------------------------------------------- BxDFs (BxDF.cuh, models/*.cuh)
enum { REFL = 1, TRANS = 2, DIFF = 4, GLOSSY = 8, SPEC = 16, ALL = 31 };

static inline float cosT(V3 w) { return w.z; }
static inline float absCosT(V3 w) { return std::fabs(w.z); }
static inline float cos2T(V3 w) { return w.z * w.z; }
static inline float sin2T(V3 w) { return std::fmax(1 - cos2T(w), 0.0f); }
static inline float sinT(V3 w) { return std::sqrt(sin2T(w)); }
static inline float tanT(V3 w) { return sinT(w) / cosT(w); }
static inline float tan2T(V3 w) { return sin2T(w) / cos2T(w); }
// BxDF.cuh:144-154: the `sinTheta == 0` test compares a function address, so it never fires.
static inline float cosPhi(V3 w) { float s = sinT(w); return clampf(w.x / s, -1, 1); }
static inline float sinPhi(V3 w) { float s = sinT(w); return clampf(w.y / s, -1, 1); }
static inline float cos2Phi(V3 w) { return cosPhi(w) * cosPhi(w); }
static inline float sin2Phi(V3 w) { return sinPhi(w) * sinPhi(w); }
static inline bool sameHemi(V3 a, V3 b) { return a.z * b.z > 0; }

// fresnelReflectionDielectric (BxDF.cuh:16-37) with its clamp(cos, 1, -1)
static float frDielectric(float cosThetaI, float etaI, float etaT) {
    cosThetaI = clampf(cosThetaI, 1, -1);
    bool entering = cosThetaI > 0.0f;
    if (!entering) { float tmp = etaI; etaI = etaT; etaT = tmp; cosThetaI = std::fabs(cosThetaI); }
    float sinThetaI = std::sqrt(std::fmax(0.0f, 1 - cosThetaI * cosThetaI));
    float sinThetaT = etaI / etaT * sinThetaI;
    if (sinThetaT >= 1) return 1;
    float cosThetaT = std::sqrt(std::fmax(0.0f, 1 - sinThetaT * sinThetaT));
    float rpar = ((etaT * cosThetaI) - (etaI * cosThetaT)) / ((etaT * cosThetaI) + (etaI * cosThetaT));
    float rperp = ((etaI * cosThetaI) - (etaT * cosThetaT)) / ((etaI * cosThetaI) + (etaT * cosThetaT));
    return (rpar * rpar + rperp * rperp) / 2;
}

static bool refract(V3 wi, V3 n, float eta, V3* wt) {
    float cosThetaI = dot(n, wi);
    float sin2ThetaI = std::fmax(0.0f, 1.0f - cosThetaI * cosThetaI);
    float sin2ThetaT = eta * eta * sin2ThetaI;
    if (sin2ThetaT >= 1) return false;
    float cosThetaT = std::sqrt(1 - sin2ThetaT);
    *wt = -wi * eta + n * (eta * cosThetaI - cosThetaT);
    return true;
}

struct TR {
    float ax, ay;
    float lambda(V3 w) const {
        float absTan = std::fabs(tanT(w));
        if (std::isinf(absTan)) return 0.0f;
        float alpha = std::sqrt(cos2Phi(w) * ax * ax + sin2Phi(w) * ay * ay);
        float a2t2 = (alpha * absTan) * (alpha * absTan);
        return (-1 + std::sqrt(1.0f + a2t2)) / 2;
    }
    float D(V3 wh) const {
        float tan2 = tan2T(wh);
        if (std::isinf(tan2)) return 0;
        const float cos4 = cos2T(wh) * cos2T(wh);
        float e = (cos2Phi(wh) / (ax * ax) + sin2Phi(wh) / (ay * ay)) * tan2;
        return 1 / (PI_F * ax * ay * cos4 * (1 + e) * (1 + e));
    }
    float G1(V3 w) const { return 1 / (1 + lambda(w)); }
    float G(V3 wo, V3 wi) const { return 1 / (1 + lambda(wo) + lambda(wi)); }
    float pdf(V3 wo, V3 wh) const { return D(wh) * G1(wo) * absDot(wo, wh) / absCosT(wo); }
};

// Microfacet.cuh:53-121
static void trSample11(float cosTheta, float U1, float U2, float* sx, float* sy) {
    if ((double)cosTheta > .9999) {
        float r = std::sqrt(U1 / (1 - U1));
        float phi = (float)(6.28318530718 * (double)U2);
        *sx = r * f_cos(phi);
        *sy = r * f_sin(phi);
        return;
    }
    float sinTheta = std::sqrt(std::fmax(0.0f, 1.0f - cosTheta * cosTheta));
    float tanTheta = sinTheta / cosTheta;
    float a = 1 / tanTheta;
    float G1 = 2 / (1 + std::sqrt(1.f + 1.f / (a * a)));
    float A = 2 * U1 / G1 - 1;
    float tmp = 1.f / (A * A - 1.f);
    if ((double)tmp > 1e10) tmp = (float)1e10;
    float B = tanTheta;
    float D = std::sqrt(std::fmax(float(B * B * tmp * tmp - (A * A - B * B) * tmp), 0.0f));
    float sx1 = B * tmp - D, sx2 = B * tmp + D;
    *sx = (A < 0 || sx2 > 1.f / tanTheta) ? sx1 : sx2;
    float S;
    if (U2 > 0.5f) { S = 1.f; U2 = 2.f * (U2 - .5f); }
    else { S = -1.f; U2 = 2.f * (.5f - U2); }
    float z = (U2 * (U2 * (U2 * 0.27385f - 0.73369f) + 0.46341f)) /
              (U2 * (U2 * (U2 * 0.093073f + 0.309420f) - 1.000000f) + 0.597999f);
    *sy = S * z * std::sqrt(1.f + *sx * *sx);
}
static V3 trSample(V3 wi, float ax, float ay, float U1, float U2) {
    V3 ws = normalize(V3(ax * wi.x, ay * wi.y, wi.z));
    float sx, sy;
    trSample11(cosT(ws), U1, U2, &sx, &sy);
    float tmp = cosPhi(ws) * sx - sinPhi(ws) * sy;
    sy = sinPhi(ws) * sx + cosPhi(ws) * sy;
    sx = tmp;
    sx = ax * sx; sy = ay * sy;
    return normalize(V3(-sx, -sy, (float)1.));
}
static float roughnessToAlpha(float roughness) {
    roughness = std::fmax(roughness, (float)1e-3);
    float x = f_log(roughness);
    return 1.62142f + 0.819955f * x + 0.1734f * x * x + 0.0171201f * x * x * x + 0.000640711f * x * x * x * x;
}

// One BxDF per material in this scene model (Material.cuh:31-130 + the microfacet mirror extension).
struct BxDF {
    int kind = -1;  // 0 Lambert, 1 OrenNayar, 2 SpecularReflection, 3 FresnelSpecular, 4 MicrofacetReflection
    int type = 0;
    Spec R, T;
    float A = 0, B = 0, etaA = 1, etaB = 1;
    TR dist{1, 1};

    Spec f(V3 wo, V3 wi) const {
        switch (kind) {
            case 0: return R * INV_PI_F;
            case 1: {
                float sinThetaO = sinT(wo), sinThetaI = sinT(wi);
                float maxCos = 0.0f;
                if ((double)sinThetaI > 1e-4 && (double)sinThetaO > 1e-4) {
                    float sinPhiO = sinPhi(wo), cosPhiO = cosPhi(wo);
                    float sinPhiI = sinPhi(wi), cosPhiI = cosPhi(wi);
                    float dCos = cosPhiI * cosPhiO + sinPhiI * sinPhiO;
                    maxCos = std::fmax(0.0f, dCos);
                }
                float sinAlpha, tanBeta;
                if (absCosT(wi) > absCosT(wo)) { sinAlpha = sinThetaO; tanBeta = sinThetaI / absCosT(wi); }
                else { sinAlpha = sinThetaI; tanBeta = sinThetaO / absCosT(wo); }
                return R * INV_PI_F * (A + B * maxCos * sinAlpha * tanBeta);
            }
            case 4: {
                float cosThetaO = absCosT(wo), cosThetaI = absCosT(wi);
                V3 wh = wi + wo;
                if (cosThetaI == 0 || cosThetaO == 0) return Spec(0.0f);
                if (wh.x == 0.0f && wh.y == 0.0f && wh.z == 0.0f) return Spec(0.0f);
                wh = normalize(wh);
                Spec F(1.0f);  // FresnelNoOp
                return R * dist.D(wh) * dist.G(wo, wi) * F / (4 * cosThetaI * cosThetaO);
            }
            default: return Spec(0.0f);
        }
    }
    float pdf(V3 wo, V3 wi) const {
        switch (kind) {
            case 0: case 1: return sameHemi(wo, wi) ? absCosT(wi) * INV_PI_F : 0;
            case 4: {
                if (!sameHemi(wo, wi)) return 0;
                V3 wh = normalize(wo + wi);
                return dist.pdf(wo, wh) / (4 * dot(wo, wh));
            }
            default: return 0.0f;
        }
    }
    // returns f; *pdf left untouched on the early-out paths (BSDF::sampleF zeroes it first)
    Spec sampleF(V3 wo, V3* wi, V2 u, float* pdf, int* sampledType) const {
        switch (kind) {
            case 0: case 1:
                *wi = cosineSampleHemisphere(u);
                if (wo.z < 0) wi->z *= -1;
                *pdf = this->pdf(wo, *wi);
                return f(wo, *wi);
            case 2:
                *wi = V3(-wo.x, -wo.y, wo.z);
                *pdf = 1.0f;
                return (Spec(1.0f) * R) / absCosT(*wi);
            case 3: {
                float F = frDielectric(cosT(wo), etaA, etaB);
                if (u.x < F) {
                    *wi = V3(-wo.x, -wo.y, -wo.z);
                    *sampledType = SPEC | REFL;
                    *pdf = F;
                    return F * R / absCosT(*wi);
                }
                bool entering = cosT(wo) > 0;
                float etaI = entering ? etaA : etaB, etaT = entering ? etaB : etaA;
                if (!refract(wo, faceTowards(V3(0, 0, 1), wo), etaI / etaT, wi)) return Spec(0.0f);
                Spec ft = T * (1 - F);
                ft = ft * ((etaI * etaI) / (etaT * etaT));
                *sampledType = SPEC | TRANS;
                *pdf = 1.0f - F;
                return ft / absCosT(*wi);
            }
            case 4: {
                if (wo.z == 0) return Spec(0.0f);
                bool flip = wo.z < 0;
                V3 wh = trSample(flip ? -wo : wo, dist.ax, dist.ay, u.x, u.y);
                if (flip) wh = -wh;
                if (dot(wo, wh) < 0) return Spec(0.0f);
                *wi = -wo + wh * (2 * dot(wo, wh));
                if (!sameHemi(wo, *wi)) return Spec(0.0f);
                *pdf = dist.pdf(wo, wh) / (4 * dot(wo, wh));
                return f(wo, *wi);
            }
        }
        return Spec(0.0f);
    }
};

struct BSDF {
    V3 ns, ng, ss, ts;
    int nB = 0;
    BxDF b;
    V3 toLocal(V3 v) const { return V3(dot(v, ss), dot(v, ts), dot(v, ns)); }
    V3 toWorld(V3 v) const {
        return V3(ss.x * v.x + ts.x * v.y + ns.x * v.z, ss.y * v.x + ts.y * v.y + ns.y * v.z,
                  ss.z * v.x + ts.z * v.y + ns.z * v.z);
    }
    bool matches(int flags) const { return nB > 0 && (b.type & flags) == b.type; }
    // BSDF.cuh:60-72
    Spec f(V3 woW, V3 wiW, int flags) const {
        V3 wi = toLocal(wiW), wo = toLocal(woW);
        bool reflect = dot(wiW, ng) * dot(woW, ng) > 0;
        Spec r(0.0f);
        if (matches(flags) && ((reflect && (b.type & REFL)) || (!reflect && (b.type & TRANS)))) r += b.f(wo, wi);
        return r;
    }
    // BSDF.cuh:129-144
    float pdf(V3 woW, V3 wiW, int flags) const {
        if (nB == 0) return 0.0f;
        V3 wo = toLocal(woW), wi = toLocal(wiW);
        if (wo.z == 0) return 0.0f;
        float p = 0.0f; int m = 0;
        if (matches(flags)) { m++; p = p + b.pdf(wo, wi); }
        return m > 0 ? p / m : 0.0f;
    }
    // BSDF.cuh:74-127 with at most one component
    Spec sampleF(V3 woW, V3* wiW, V2 u, float* pdf, int flags, int* sampledType) const {
        if (!matches(flags)) { *pdf = 0; return Spec(0.0f); }
        int comp = (int)std::fmin((float)(int)std::floor(u.x * 1), (float)(1 - 1));
        V2 rem(u.x * 1 - comp, u.y);
        V3 wi, wo = toLocal(woW);
        *pdf = 0;
        if (sampledType) *sampledType = b.type;
        Spec f = b.sampleF(wo, &wi, rem, pdf, sampledType);
        if (*pdf == 0) return Spec(0.0f);
        *wiW = toWorld(wi);
        return f;
    }
};

// ----------------------------------------------------------------- textures (Procedural.cuh)
static const int PERM[512] = {
#define P256 151,160,137,91,90,15,131,13,201,95,96,53,194,233,7,225,140,36,103,30,69,142,8,99,37,240,21,10,23, \
    190,6,148,247,120,234,75,0,26,197,62,94,252,219,203,117,35,11,32,57,177,33,88,237,149,56,87,174,20,125,136,171,168, \
    68,175,74,165,71,134,139,48,27,166,77,146,158,231,83,111,229,122,60,211,133,230,220,105,92,41,55,46,245,40,244, \
    102,143,54,65,25,63,161,1,216,80,73,209,76,132,187,208,89,18,169,200,196,135,130,116,188,159,86,164,100,109,198, \
    173,186,3,64,52,217,226,250,124,123,5,202,38,147,118,126,255,82,85,212,207,206,59,227,47,16,58,17,182,189,28,42, \
    223,183,170,213,119,248,152,2,44,154,163,70,221,153,101,155,167,43,172,9,129,22,39,253,19,98,108,110,79,113,224, \
    232,178,185,112,104,218,246,97,228,251,34,242,193,238,210,144,12,191,179,162,241,81,51,145,235,249,14,239,107,49, \
    192,214,31,181,199,106,157,184,84,204,176,115,121,50,45,127,4,150,254,138,236,205,93,222,114,67,29,24,72,243,141, \
    128,195,78,66,215,61,156,180
    P256, P256
#undef P256
};
static inline float noiseWeight(float t) { float t3 = t * t * t; float t4 = t3 * t; return 6 * t4 * t - 15 * t4 + 10 * t3; }
static inline float smoothStep(float mn, float mx, float val) {
    float v = clampf((val - mn) / (mx - mn), 0, 1);
    return v * v * (-2 * v + 3);
}
static inline float grad(int x, int y, int z, float dx, float dy, float dz) {
    int h = PERM[PERM[PERM[x] + y] + z];
    h &= 15;
    float u = h < 8 || h == 12 || h == 13 ? dx : dy;
    float v = h < 4 || h == 12 || h == 13 ? dy : dz;
    return ((h & 1) ? -u : u) + ((h & 2) ? -v : v);
}
static float perlin(float x, float y, float z) {
    int ix = (int)std::floor(x), iy = (int)std::floor(y), iz = (int)std::floor(z);
    float dx = x - ix, dy = y - iy, dz = z - iz;
    ix &= 255; iy &= 255; iz &= 255;
    float w000 = grad(ix, iy, iz, dx, dy, dz);
    float w100 = grad(ix + 1, iy, iz, dx - 1, dy, dz);
    float w010 = grad(ix, iy + 1, iz, dx, dy - 1, dz);
    float w001 = grad(ix, iy, iz + 1, dx, dy, dz - 1);
    float w110 = grad(ix + 1, iy + 1, iz, dx - 1, dy - 1, dz);
    float w011 = grad(ix, iy + 1, iz + 1, dx, dy - 1, dz - 1);
    float w101 = grad(ix + 1, iy, iz + 1, dx - 1, dy, dz - 1);
    float w111 = grad(ix + 1, iy + 1, iz + 1, dx - 1, dy - 1, dz - 1);
    float wx = noiseWeight(dx), wy = noiseWeight(dy), wz = noiseWeight(dz);
    float x00 = lerpf(wx, w000, w100), x10 = lerpf(wx, w010, w110);
    float x01 = lerpf(wx, w001, w101), x11 = lerpf(wx, w011, w111);
    float y0 = lerpf(wy, x00, x10), y1 = lerpf(wy, x01, x11);
    return lerpf(wz, y0, y1);
}
static inline float perlinP(V3 p) { return perlin(p.x, p.y, p.z); }
static float fbm(V3 p, V3 dpdx, V3 dpdy, float omega, int maxOct) {
    float len2 = std::fmax(lengthSq(dpdx), lengthSq(dpdy));
    float n = clampf(-1 - 0.5f * f_log2(len2), 0, (float)maxOct);
    int nInt = (int)std::floor(n);
    float sum = 0, lambda = 1, o = 1;
    for (int i = 0; i < nInt; i++) { sum += o * perlinP(p * lambda); lambda *= 1.99f; o *= omega; }
    float nPartial = n - nInt;
    sum += o * smoothStep(0.3f, 0.7f, nPartial) * perlinP(p * lambda);
    return sum;
}
static float turbulence(V3 p, V3 dpdx, V3 dpdy, float omega, int maxOct) {
    float len2 = std::fmax(lengthSq(dpdx), lengthSq(dpdy));
    float n = clampf(-1 - 0.5f * f_log2(len2), 0, (float)maxOct);
    int nInt = (int)std::floor(n);
    float sum = 0, lambda = 1, o = 1;
    for (int i = 0; i < nInt; i++) { sum += o * std::fabs(perlinP(p * lambda)); lambda *= 1.99f; o *= omega; }
    float nPartial = n - nInt;
    sum += o * lerpf(smoothStep(0.3f, 0.7f, nPartial), 0.2f, std::fabs(perlinP(p * lambda)));
    for (int i = nInt; i < maxOct; i++) { sum += o * 0.2f; o *= omega; }
    return sum;
}
// MarbleTexture (Procedural.cuh:267-311); a tenth, zero row stands in for the read one
// past the table when t == 1 (its weight is then exactly 0).
static const float MARBLE[10][3] = {{.58f, .58f, .6f}, {.58f, .58f, .6f}, {.58f, .58f, .6f}, {.5f, .5f, .5f},
                                    {.6f, .59f, .58f}, {.58f, .58f, .6f}, {.58f, .58f, .6f}, {.2f, .2f, .33f},
                                    {.58f, .58f, .6f}, {0, 0, 0}};
static Spec marble(const Tex& X, V3 p, V3 dpdx, V3 dpdy) {
    p = p * X.scale;
    float m = p.y + X.variation * fbm(p, dpdx * X.scale, dpdy * X.scale, X.omega, X.octaves);
    float t = 0.5f + 0.5f * f_sin(m);
    const float NSEG = 6.0f;
    int first = (int)std::floor(t * NSEG);
    t = (t * NSEG - first);
    Spec c0(MARBLE[first][0], MARBLE[first][1], MARBLE[first][2]);
    Spec c1(MARBLE[first + 1][0], MARBLE[first + 1][1], MARBLE[first + 1][2]);
    Spec c2(MARBLE[first + 2][0], MARBLE[first + 2][1], MARBLE[first + 2][2]);
    Spec c3(MARBLE[first + 3][0], MARBLE[first + 3][1], MARBLE[first + 3][2]);
    Spec s0 = (1.0f - t) * c0 + t * c1;
    Spec s1 = (1.0f - t) * c1 + t * c2;
    Spec s2 = (1.0f - t) * c2 + t * c3;
    s0 = (1.0f - t) * s0 + t * s1;
    s1 = (1.0f - t) * s1 + t * s2;
    return 1.5f * ((1.0f - t) * s0 + t * s1);
}

static Spec evalTex(const Scene& s, const Tex& X, const SI& si) {
    switch (X.type) {
        case 'c': return X.c;
        case 'i': {
            // ImageTexture::evaluate uses the fixed width 0.01 (ImageTexture.cuh:8, 26); the filtered
            // pipeline takes PBRT's isotropic width 2 max(|ds/dx|, |dt/dx|, |ds/dy|, |dt/dy|) of
            // UVMapping2D (st = uv, Mapping.cuh:19-37)
            const Mipmap& m = s.mips[X.img];
            const float width = m.filtered ? 2 * std::fmax(std::fmax(std::fabs(si.dudx), std::fabs(si.dvdx)),
                                                           std::fmax(std::fabs(si.dudy), std::fabs(si.dvdy)))
                                           : 0.01f;
            return mipLookup(m, V2(1 * si.uv.x + 0, 1 * si.uv.y + 0), width);
        }
        case 'f': return Spec(fbm(identityMap(si.p), identityMap(si.dpdx), identityMap(si.dpdy), X.omega, X.octaves));
        case 'w': return Spec(turbulence(identityMap(si.p), identityMap(si.dpdx), identityMap(si.dpdy), X.omega, X.octaves));
        case 'm': return marble(X, identityMap(si.p), identityMap(si.dpdx), identityMap(si.dpdy));
    }
    return Spec(0.0f);
}

// SurfaceInteraction::computeDifferentials (Interaction.cuh:118-160)
static void computeDifferentials(SI& si, const Ray& r) {
    if (!r.hasDiff) { si.dpdx = si.dpdy = V3(0, 0, 0); return; }
    float d = -dot(si.n, si.p);
    float tx = (-dot(si.n, r.rxO) - d) / dot(si.n, r.rxD);
    V3 px = r.rxO + r.rxD * tx;
    float ty = (-dot(si.n, r.ryO) - d) / dot(si.n, r.ryD);
    V3 py = r.ryO + r.ryD * ty;
    si.dpdx = px - si.p;
    si.dpdy = py - si.p;
    // du/dx, dv/dx, du/dy, dv/dy on the two axes other than the normal's dominant one
    // (Interaction.cuh:134-153, solveLinearSystem2x2 GeneralHelper.cuh:127-134); only the filtered
    // texture pipeline reads them, as the derivatives of the true uv parametrisation: solved against
    // the interaction's dpdu, dpdv (uvScale times too long) and multiplied by uvScale
    int d0, d1;
    if (std::fabs(si.n.x) > std::fabs(si.n.y) && std::fabs(si.n.x) > std::fabs(si.n.z)) { d0 = 1; d1 = 2; }
    else if (std::fabs(si.n.y) > std::fabs(si.n.z)) { d0 = 0; d1 = 2; }
    else { d0 = 0; d1 = 1; }
    auto c = [](V3 v, int k) { return k == 0 ? v.x : k == 1 ? v.y : v.z; };
    const float a00 = c(si.dpdu, d0), a01 = c(si.dpdv, d0), a10 = c(si.dpdu, d1), a11 = c(si.dpdv, d1);
    auto solve = [&](float b0, float b1, float& x0, float& x1) {
        const float det = a00 * a11 - a01 * a10;
        if (std::fabs(det) < 1e-10f) { x0 = x1 = 0; return; }
        x0 = (a11 * b0 - a01 * b1) / det;
        x1 = (a00 * b1 - a10 * b0) / det;
        if (std::isnan(x0) || std::isnan(x1)) x0 = x1 = 0;
    };
    solve(c(px, d0) - c(si.p, d0), c(px, d1) - c(si.p, d1), si.dudx, si.dvdx);
    solve(c(py, d0) - c(si.p, d0), c(py, d1) - c(si.p, d1), si.dudy, si.dvdy);
    si.dudx *= si.uvScale; si.dvdx *= si.uvScale; si.dudy *= si.uvScale; si.dvdy *= si.uvScale;
}

// Material::computeScatteringFunctions (Material.cuh:31-130)
static BSDF makeBSDF(const Scene& s, const SI& si) {
    BSDF b;
    b.ns = si.sn; b.ng = si.n;
    b.ss = normalize(si.sdpdu);
    b.ts = cross(b.ns, b.ss);
    const Mat& M = s.mats[s.triMesh[si.tri]];
    switch (M.type) {
        case 'L': {
            Spec kd = evalTex(s, M.t0, si);
            if (!kd.isBlack()) {
                b.nB = 1;
                b.b.R = kd;
                b.b.type = REFL | DIFF;
                if (M.extra == 0) b.b.kind = 0;
                else {
                    b.b.kind = 1;
                    float sigma = M.extra * (PI_F / 180);
                    float sigma2 = sigma * sigma;
                    b.b.A = 1.0f - (sigma2 / (2.0f * (sigma2 + 0.33f)));
                    b.b.B = 0.45f * sigma2 / (sigma2 + 0.09f);
                }
            }
            break;
        }
        case 'G': {
            Spec r = evalTex(s, M.t0, si), t = evalTex(s, M.t1, si);
            if (r.isBlack() && t.isBlack()) break;
            b.nB = 1; b.b.kind = 3; b.b.type = REFL | TRANS | SPEC; b.b.R = r; b.b.T = t; b.b.etaA = 1.0f; b.b.etaB = M.extra;
            break;
        }
        case 'R': {
            Spec r = evalTex(s, M.t0, si);
            if (r.isBlack()) break;
            b.nB = 1; b.b.kind = 2; b.b.type = REFL | SPEC; b.b.R = r;
            break;
        }
        case 'F': {
            Spec r = evalTex(s, M.t0, si);
            if (r.isBlack()) break;
            float a = roughnessToAlpha(M.extra);
            b.nB = 1; b.b.kind = 4; b.b.type = REFL | GLOSSY; b.b.R = r; b.b.dist = TR{a, a};
            break;
        }
    }
    return b;
}

// ----------------------------------------------------------------- lights
struct LightSample { V3 p, n, pError; };

// Triangle::sample (Triangle.cu:232-259)
static LightSample triSample(const Scene& s, int g, V2 u) {
    float t = std::sqrt(u.x);
    V2 b(1 - t, u.y * t);
    const Mesh& M = s.meshes[s.triMesh[g]];
    V3 p0, p1, p2; const int* id;
    triVerts(s, g, p0, p1, p2, &id);
    LightSample ls;
    ls.p = p0 * b.x + p1 * b.y + p2 * (1 - b.x - b.y);
    if (!M.N.empty())
        ls.n = normalize(M.N[id[2]] * b.x + M.N[id[5]] * b.y + M.N[id[8]] * (1 - b.x - b.y));
    else
        ls.n = normalize(cross(p1 - p0, p2 - p0));
    V3 pAbs = vabs(p0 * b.x) + vabs(p1 * b.y) + vabs(p2 * (1 - b.x - b.y));
    ls.pError = pAbs * gammaN(6);
    return ls;
}

// Triangle::Pdf(it, wi) (Triangle.cu:211-225)
static float triPdf(const Scene& s, int g, const SI& it, V3 wi, const Light& L) {
    Ray r = spawnRay(it, wi);
    float t, u, v, invDet;
    if (!triHit(s, g, r, t, u, v, invDet)) return 0;
    SI isect = triInteraction(s, g, r, u, v, invDet);
    return lengthSq(it.p - isect.p) / (absDot(isect.n, -wi) * L.area);
}

static float infPdf(V3 wi) {
    V3 w = wi;
    float sinTheta = f_sin(f_acos(clampf(w.z, -1.0f, 1.0f)));
    return (0.5f * INV_PI_F) / (2 * PI_F * PI_F * sinTheta);
}

// Light::sampleLi (Light.cu:28-37, :68-90). Returns Li; fills wi, pdf and the visibility target.
static Spec sampleLi(const Scene& s, const Light& L, const SI& it, V2 u, V3* wi, float* pdf, LightSample* target) {
    if (L.kind == 0) {
        LightSample ps = triSample(s, L.tri, u);
        *wi = normalize(ps.p - it.p);
        *pdf = triPdf(s, L.tri, it, *wi, L);
        *target = ps;
        return dot(ps.n, -*wi) > 0.f ? L.Le : Spec(0.0f);
    }
    V3 w = uniformSampleHemisphere(u);
    V3 ref(0, 1, 0), v1, v2;
    coordinateSystem(ref, &v1, &v2);
    w = v1 * w.x + v2 * w.y + ref * w.z;
    V3 pL = L.center + w * L.radius;
    *wi = normalize(pL - it.p);
    *pdf = (0.5f * INV_PI_F) / (2 * PI_F * PI_F * f_sin(f_acos(clampf(wi->z, -1.0f, 1.0f))));
    target->p = pL; target->n = V3(0, 0, 0); target->pError = V3(0, 0, 0);
    return L.Le;
}

// Interaction::spawnRayTo(Interaction) (Interaction.cuh:47-51) + VisibilityTester::unoccluded
static bool unoccluded(const Scene& s, const SI& it, const LightSample& p1, Counters& c) {
    V3 origin = offsetRayOrigin(it.p, it.pError, it.n, p1.p - it.p);
    V3 target = offsetRayOrigin(p1.p, p1.pError, p1.n, origin - p1.p);
    Ray r; r.o = origin; r.d = target - origin; r.tMax = 1 - SHADOW_EPS;
    return !traverse(s, r, true, nullptr, c);
}

// Distribution1D::findInterval / sampleDiscrete (Sampling.cuh:155-188); the search runs
// over size (not size + 1) entries, so the last entry is never returned when size > 1.
static int sampleLight(const Scene& s, float u, float* pdf) {
    int size = (int)s.lfunc.size();
    int off;
    if (size == 1) off = 0;
    else {
        int first = 0, len = size;
        while (len > 0) {
            int half = len >> 1, middle = first + half;
            if (s.lcdf[middle] <= u) { first = middle + 1; len -= half + 1; }
            else len = half;
        }
        off = (int)clampf((float)(first - 1), 0, (float)(size - 2));
    }
    *pdf = s.lfunc[off] / (s.funcInt * size);
    return off;
}

// estimateDirect (IntegratorHelper.cuh:51-125)
static Spec estimateDirect(const Scene& s, const SI& si, const BSDF& bsdf, V2 uScatter, int lightIdx, V2 uLight,
                           bool specular, Counters& c) {
    const Light& L = s.lights[lightIdx];
    Spec Ld(0.0f);
    int flags = specular ? ALL : (ALL & ~SPEC);
    V3 wi;
    float lightPdf = 0.0f, scatteringPdf = 0.0f;
    LightSample vis;
    Spec Li = sampleLi(s, L, si, uLight, &wi, &lightPdf, &vis);
    if (lightPdf > 0 && !Li.isBlack()) {
        Spec f = bsdf.f(si.wo, wi, flags) * absDot(wi, si.sn);
        scatteringPdf = bsdf.pdf(si.wo, wi, flags);
        if (!f.isBlack()) {
            if (!unoccluded(s, si, vis, c)) Li = Spec(0.0f);
            if (!Li.isBlack()) {
                float weight = powerHeuristic(lightPdf, scatteringPdf);
                Ld += f * Li * weight / lightPdf;
            }
        }
    }
    {
        int sampledType = 0;
        Spec f = bsdf.sampleF(si.wo, &wi, uScatter, &scatteringPdf, flags, &sampledType);
        f = f * absDot(wi, si.sn);
        bool sampledSpecular = (sampledType & SPEC) != 0;
        if (!f.isBlack() && scatteringPdf > 0) {
            float weight = 1;
            if (!sampledSpecular) {
                lightPdf = L.kind == 0 ? triPdf(s, L.tri, si, wi, L) : infPdf(wi);
                if (lightPdf == 0) return Ld;
                weight = powerHeuristic(scatteringPdf, lightPdf);
            }
            Ray r = spawnRay(si, wi);
            SI lightIt;
            bool found = traverse(s, r, false, &lightIt, c);
            Spec Li2(0.f);
            if (found) {
                if (s.triLight[lightIt.tri] == lightIdx) Li2 = dot(lightIt.n, -wi) > 0.f ? L.Le : Spec(0.0f);
            } else if (L.kind == 1) Li2 = L.Le;
            if (!Li2.isBlack()) Ld += f * Li2 * Spec(1.0f) * weight / scatteringPdf;
        }
    }
    return Ld;
}

// Camera::generateRay / generateDifferentialRay (Camera.cu:51-81) at a film position and lens
// sample, the differentials scaled by scaleDifferentials(1 / sqrt(spp)) (PathTracingIntegrator.cuh:46).
static Ray cameraRay(const Scene& s, V2 pFilm, V2 lens, int spp) {
    auto gen = [&](V2 pf, V3* o, V3* d) {
        V3 offset = (s.cu * lens.x + s.cv * lens.y) * s.lensRadius;
        V2 pr(pf.x / (s.W - 1), pf.y / (s.H - 1));
        *o = s.o + offset;
        *d = normalize(s.llc + s.right * pr.x + s.up * pr.y - s.o - offset);
    };
    Ray r;
    gen(pFilm, &r.o, &r.d);
    gen(pFilm + V2(1, 0), &r.rxO, &r.rxD);
    gen(pFilm + V2(0, 1), &r.ryO, &r.ryD);
    r.hasDiff = true;
    float sc = quake((float)spp);
    r.rxO = r.o + (r.rxO - r.o) * sc;
    r.ryO = r.o + (r.ryO - r.o) * sc;
    r.rxD = r.d + (r.rxD - r.d) * sc;
    r.ryD = r.d + (r.ryD - r.d) * sc;
    return r;
}
// PathTracingIntegrator::evaluate, one camera sample (PathTracingIntegrator.cuh:38-104)
static Spec Li(const Scene& s, int px, int py, int sample, const oracle_render_params& prm, V2* pFilmOut, Counters& c) {
    Sampler smp;
    smp.seed = prm.seed; smp.pixel = (uint32_t)(py * s.W + px); smp.sample = (uint32_t)sample;
    V2 film(smp.get(0), smp.get(1));
    if (prm.stratified) {
        // Sampler::stratifiedSample (Sampler.cuh:49-64)
        int strata = (int)std::floor(std::sqrt((float)prm.spp));
        if (sample - 1 < strata * strata) {
            float inv = (float)1 / strata;
            film = V2(film.x * inv, film.y * inv);
            film = film + V2((sample % strata) / (float)strata, (sample / strata) / (float)strata);
        }
    }
    V2 lens = uniformSampleDisk(V2(smp.get(2), smp.get(3)));
    V2 pFilm((float)(px + film.x), (float)(py + film.y));
    *pFilmOut = pFilm;
    // Camera::generateRay / generateDifferentialRay (Camera.cu:51-81)
    Ray r = cameraRay(s, pFilm, lens, prm.spp);

    Spec L(0.0f), beta(1.0f);
    bool specularBounce = false;
    int nLights = (int)s.lights.size();
    for (int bounce = 0; bounce < prm.max_bounce; bounce++) {
        SI si;
        bool hit = traverse(s, r, false, &si, c);
        if (bounce == 0 || specularBounce) {
            if (hit) {
                int l = s.triLight[si.tri];
                Spec le = (l >= 0 && dot(si.n, -r.d) > 0.f) ? s.lights[l].Le : Spec(0.0f);
                L += beta * le;
            } else {
                for (int l = 0; l < nLights; l++)
                    L += beta * (s.lights[l].kind == 1 ? s.lights[l].Le : Spec(0.0f));
            }
        }
        if (!hit) break;
        computeDifferentials(si, r);
        BSDF bsdf = makeBSDF(s, si);
        int base = 4 + 8 * bounce;
        if (nLights > 0) {
            float lpdf;
            int ln = sampleLight(s, smp.get(base + 0), &lpdf);
            V2 uL(smp.get(base + 1), smp.get(base + 2));
            V2 uS(smp.get(base + 3), smp.get(base + 4));
            L += beta * (estimateDirect(s, si, bsdf, uS, ln, uL, specularBounce, c) / lpdf);
        }
        V3 wo = -r.d, wi;
        float pdf;
        int flags = 0;
        Spec f = bsdf.sampleF(wo, &wi, V2(smp.get(base + 5), smp.get(base + 6)), &pdf, ALL, &flags);
        if (f.isBlack() || pdf == 0) break;
        specularBounce = (flags & SPEC) != 0;
        beta = beta * (f * absDot(wi, si.sn) / pdf);
        r = spawnRay(si, wi);
        if (bounce >= prm.rr_start) {
            float q = std::fmax(0.05f, 1 - beta.y());
            if (smp.get(base + 7) < q) break;
            beta = beta / (1 - q);
        }
    }
    return L;
}

// Film gather: pixel (x, y) accumulates, for each sample s and each owner pixel (ox, oy) in
// x-2..x+1, y-2..y+1 (row-major), the owner's sample if its footprint covers (x, y)
// (Camera::addSample, Camera.cu:192-218). The footprint [ceil(p - 1), floor(p + 1) + 1) of a film
// position p = owner + f spans owner..owner+1 for f in (0, 1), adds owner+2 when f is exactly 1,
// and adds owner-1 when owner + f rounds to the owner itself (f below half an ulp, from 1024 on).
struct FilmPix { Spec rgb; float w = 0; };

static inline bool footprintWeight(int W, int H, const float* table, V2 pFilm, int x, int y, float* w) {
    int p0x = (int)std::ceil(pFilm.x - 1.0f), p0y = (int)std::ceil(pFilm.y - 1.0f);
    int p1x = (int)std::floor(pFilm.x + 1.0f) + 1, p1y = (int)std::floor(pFilm.y + 1.0f) + 1;
    p0x = std::max(p0x, 0); p0y = std::max(p0y, 0);
    p1x = std::min(p1x, W); p1y = std::min(p1y, H);
    if (x < p0x || x >= p1x || y < p0y || y >= p1y) return false;
    float dx = std::fabs((x - pFilm.x - 0.5f) * 1.0f * 16);
    float dy = std::fabs((y - pFilm.y - 0.5f) * 1.0f * 16);
    int ix = (int)std::fmin((float)(int)std::floor(dx), (float)(16 - 1));
    int iy = (int)std::fmin((float)(int)std::floor(dy), (float)(16 - 1));
    *w = table[iy * 16 + ix];
    return true;
}

// The gather over a film of W x H pixels: pf and L hold each sample's film position and radiance,
// [sample][y][x]; film receives the rgb and weight sums.
static void gatherFilm(int W, int H, int spp, const V2* pf, const float* L, const float* table, FilmPix* film) {
    const size_t np = (size_t)W * H;
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            FilmPix& fp = film[(size_t)y * W + x];
            for (int smp = 0; smp < spp; smp++)
                for (int oy = y - 2; oy <= y + 1; oy++)
                    for (int ox = x - 2; ox <= x + 1; ox++) {
                        if (ox < 0 || oy < 0 || ox >= W || oy >= H) continue;
                        size_t oi = (size_t)smp * np + (size_t)oy * W + ox;
                        float w;
                        if (!footprintWeight(W, H, table, pf[oi], x, y, &w)) continue;
                        const float* l = &L[3 * oi];
                        fp.rgb += Spec(l[0], l[1], l[2]) * w;
                        fp.w += w;
                    }
        }
}

} // namespace orc

using namespace orc;

static std::string g_err;

extern "C" const char* oracle_last_error(void) { return g_err.c_str(); }

struct oracle_scene { Scene s; };

extern "C" oracle_scene* oracle_load_scene(const char* dir, const char* scene, const char* subject, int32_t w, int32_t h) {
    oracle_scene* o = new oracle_scene();
    std::string err;
    if (w < 2 || h < 2 || !loadScene(o->s, dir ? dir : "", scene ? scene : "", subject ? subject : "", w, h, err)) {
        g_err = err.empty() ? "invalid film size" : err;
        delete o;
        return nullptr;
    }
    return o;
}
extern "C" void oracle_free_scene(oracle_scene* s) { delete s; }
extern "C" int32_t oracle_scene_info(const oracle_scene* s, int32_t* nt, int32_t* nl, int32_t* nm) {
    if (!s) return -1;
    if (nt) *nt = s->s.nPrim;
    if (nl) *nl = (int32_t)s->s.lights.size();
    if (nm) *nm = (int32_t)s->s.meshes.size();
    return 0;
}

extern "C" void oracle_filter_table(float* out) { filterTable(out); }

extern "C" void oracle_film_gather(int32_t w, int32_t h, int32_t spp, const float* pfilm, const float* L, float* film) {
    float table[256];
    filterTable(table);
    std::vector<FilmPix> f((size_t)w * h);
    gatherFilm(w, h, spp, reinterpret_cast<const V2*>(pfilm), L, table, f.data());
    for (size_t i = 0; i < f.size(); i++) {
        film[4 * i] = f[i].rgb.r; film[4 * i + 1] = f[i].rgb.g; film[4 * i + 2] = f[i].rgb.b; film[4 * i + 3] = f[i].w;
    }
}

extern "C" void oracle_set_texture_filter(int32_t filtered) { g_textureFiltered = filtered != 0; }

// Pyramid of a linear RGB image (row 0 first) and one lookup at (s, t) with the given width.
extern "C" void oracle_mip_lookup(const float* rgb, int32_t w, int32_t h, int32_t wrap, int32_t filtered, float s, float t,
                                  float width, float* out) {
    std::vector<Spec> img((size_t)w * h);
    for (size_t i = 0; i < img.size(); i++) img[i] = Spec(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]);
    Mipmap m;
    buildMipmap(m, w, h, img, wrap, filtered != 0);
    const Spec v = mipLookup(m, V2(s, t), width);
    out[0] = v.r; out[1] = v.g; out[2] = v.b;
}

// The camera ray at film position (px, py) with a centred lens sample: out = o, d, rxO, rxD, ryO,
// ryD (18 floats).
extern "C" void oracle_camera_ray(oracle_scene* sc, float px, float py, int32_t spp, float* out) {
    Ray r = cameraRay(sc->s, V2(px, py), V2(0, 0), spp);
    const V3 v[6] = {r.o, r.d, r.rxO, r.rxD, r.ryO, r.ryD};
    for (int i = 0; i < 6; i++) { out[3 * i] = v[i].x; out[3 * i + 1] = v[i].y; out[3 * i + 2] = v[i].z; }
}
// The primary hit of that ray with its differentials: out = p, n, uv, dpdx, dpdy, du/dx, dv/dx,
// du/dy, dv/dy (18 floats). Returns the hit triangle or -1.
extern "C" int32_t oracle_primary_hit(oracle_scene* sc, float px, float py, int32_t spp, float* out) {
    Ray r = cameraRay(sc->s, V2(px, py), V2(0, 0), spp);
    Counters c;
    SI si;
    if (!traverse(sc->s, r, false, &si, c)) return -1;
    computeDifferentials(si, r);
    const float v[18] = {si.p.x, si.p.y, si.p.z, si.n.x, si.n.y, si.n.z, si.uv.x, si.uv.y, si.dpdx.x, si.dpdx.y,
                         si.dpdx.z, si.dpdy.x, si.dpdy.y, si.dpdy.z, si.dudx, si.dvdx, si.dudy, si.dvdy};
    std::memcpy(out, v, sizeof v);
    return si.tri;
}
// InfiniteLight::sampleLi from a point p (Light.cu:68-90): out = wi, pdf; returns -1 without an
// infinite light.
extern "C" int32_t oracle_infinite_sample(oracle_scene* sc, const float p[3], const float u[2], float* out) {
    const Scene& s = sc->s;
    if (s.infiniteLight < 0) return -1;
    SI it; it.p = V3(p[0], p[1], p[2]);
    V3 wi; float pdf; LightSample tgt;
    sampleLi(s, s.lights[s.infiniteLight], it, V2(u[0], u[1]), &wi, &pdf, &tgt);
    out[0] = wi.x; out[1] = wi.y; out[2] = wi.z; out[3] = pdf;
    return 0;
}
// InfiniteLight::PdfLi (Light.cuh:154-158)
extern "C" float oracle_infinite_pdf(const float wi[3]) { return infPdf(V3(wi[0], wi[1], wi[2])); }
// Light l: out = kind, power distribution entry, centre (3), radius (infinite light), area (area
// light), triangle.
extern "C" int32_t oracle_light_info(oracle_scene* sc, int32_t l, float* out) {
    const Scene& s = sc->s;
    if (l < 0 || l >= (int)s.lights.size()) return -1;
    const Light& L = s.lights[l];
    const float v[8] = {(float)L.kind, s.lfunc[l], L.center.x, L.center.y, L.center.z, L.radius, L.area, (float)L.tri};
    std::memcpy(out, v, sizeof v);
    return 0;
}
// Distribution1D::sampleDiscrete over the light powers (Sampling.cuh:155-188).
extern "C" int32_t oracle_sample_light(oracle_scene* sc, float u, float* pdf) { return sampleLight(sc->s, u, pdf); }

extern "C" void oracle_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    U4 c; for (int i = 0; i < 4; i++) c.v[i] = ctr[i];
    U4 r = philox4x32_10(c, key[0], key[1]);
    for (int i = 0; i < 4; i++) out[i] = r.v[i];
}
extern "C" float oracle_sample(uint32_t seed, uint32_t pixel, uint32_t sample, int32_t dim) {
    Sampler s; s.seed = seed; s.pixel = pixel; s.sample = sample;
    return s.get(dim);
}

extern "C" int32_t oracle_intersect(oracle_scene* sc, const float o[3], const float d[3], float tmax, float* t_out) {
    Ray r; r.o = V3(o[0], o[1], o[2]); r.d = V3(d[0], d[1], d[2]); r.tMax = tmax;
    Counters c; SI si;
    if (!traverse(sc->s, r, false, &si, c)) return -1;
    if (t_out) *t_out = r.tMax;
    return si.tri;
}
extern "C" int32_t oracle_occluded(oracle_scene* sc, const float o[3], const float d[3], float tmax) {
    Ray r; r.o = V3(o[0], o[1], o[2]); r.d = V3(d[0], d[1], d[2]); r.tMax = tmax;
    Counters c;
    return traverse(sc->s, r, true, nullptr, c) ? 1 : 0;
}

extern "C" int oracle_render(oracle_scene* sc, const oracle_render_params* p, float* rgb_out, float* sample_rgb,
                             oracle_render_stats* st) {
    if (!sc || !p || p->spp < 1 || p->max_bounce < 0) { g_err = "bad arguments"; return -1; }
    const Scene& s = sc->s;
    int W = s.W, H = s.H, spp = p->spp;
    int r0 = std::max(0, p->row_begin), r1 = p->row_end > 0 ? std::min(H, p->row_end) : H;
    size_t np = (size_t)W * H;
    // per-sample radiance is kept only when the film gather or the caller needs it
    std::vector<float> Ls((sample_rgb || !rgb_out) ? 0 : np * spp * 3);
    float* L = sample_rgb ? sample_rgb : (rgb_out ? Ls.data() : nullptr);
    int nth = p->threads > 0 ? p->threads : (int)std::max(1u, std::thread::hardware_concurrency());
    std::atomic<int> nextTile(0);
    const int TILE = 16;
    int tilesX = (W + TILE - 1) / TILE, tilesY = (r1 - r0 + TILE - 1) / TILE;
    std::vector<Counters> cnt(nth);
    auto t0 = std::chrono::steady_clock::now();
    auto work = [&](int tid) {
        for (;;) {
            int t = nextTile.fetch_add(1);
            if (t >= tilesX * tilesY) break;
            int tx = t % tilesX, ty = t / tilesX;
            for (int y = r0 + ty * TILE; y < std::min(r1, r0 + (ty + 1) * TILE); y++)
                for (int x = tx * TILE; x < std::min(W, (tx + 1) * TILE); x++)
                    for (int smp = 0; smp < spp; smp++) {
                        V2 pf;
                        Spec v = Li(s, x, y, smp, *p, &pf, cnt[tid]);
                        if (L) {
                            float* o = &L[3 * ((size_t)smp * np + (size_t)y * W + x)];
                            o[0] = v.r; o[1] = v.g; o[2] = v.b;
                        }
                    }
        }
    };
    std::vector<std::thread> th;
    for (int i = 0; i < nth; i++) th.emplace_back(work, i);
    for (auto& t : th) t.join();
    auto t1 = std::chrono::steady_clock::now();
    if (rgb_out) {
        // deterministic film gather + getRender (Camera.cu:256-289)
        std::vector<FilmPix> film(np);
        std::vector<V2> pf((size_t)spp * np);
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++)
                for (int smp = 0; smp < spp; smp++) {
                    Sampler sm; sm.seed = p->seed; sm.pixel = (uint32_t)(y * W + x); sm.sample = (uint32_t)smp;
                    V2 film2(sm.get(0), sm.get(1));
                    if (p->stratified) {
                        int strata = (int)std::floor(std::sqrt((float)spp));
                        if (smp - 1 < strata * strata) {
                            float inv = (float)1 / strata;
                            film2 = V2(film2.x * inv, film2.y * inv);
                            film2 = film2 + V2((smp % strata) / (float)strata, (smp / strata) / (float)strata);
                        }
                    }
                    pf[(size_t)smp * np + (size_t)y * W + x] = V2((float)(x + film2.x), (float)(y + film2.y));
                }
        gatherFilm(W, H, spp, pf.data(), L, s.filterTable, film.data());
        for (int j = 0; j < H; j++)
            for (int i = 0; i < W; i++) {
                const FilmPix& fp = film[(size_t)(H - 1 - j) * W + i];
                float inv = 1 / fp.w;
                float* o = &rgb_out[3 * ((size_t)j * W + i)];
                o[0] = std::fmax(0.0f, fp.rgb.r * inv);
                o[1] = std::fmax(0.0f, fp.rgb.g * inv);
                o[2] = std::fmax(0.0f, fp.rgb.b * inv);
            }
    }
    if (st) {
        st->seconds = std::chrono::duration<double>(t1 - t0).count();
        st->samples = (int64_t)W * (r1 - r0) * spp;
        st->node_visits = st->tri_tests = st->rays = 0;
        for (auto& c : cnt) { st->node_visits += c.nodes; st->tri_tests += c.tris; st->rays += c.rays; }
        st->threads = nth;
    }
    return 0;
}

// ---------------------------------------------------------------- component entry points
extern "C" float oracle_perlin(float x, float y, float z) { return perlin(x, y, z); }
extern "C" float oracle_fbm(const float p[3], const float dx[3], const float dy[3], float omega, int32_t oct) {
    return fbm(V3(p[0], p[1], p[2]), V3(dx[0], dx[1], dx[2]), V3(dy[0], dy[1], dy[2]), omega, oct);
}
extern "C" float oracle_turbulence(const float p[3], const float dx[3], const float dy[3], float omega, int32_t oct) {
    return turbulence(V3(p[0], p[1], p[2]), V3(dx[0], dx[1], dx[2]), V3(dy[0], dy[1], dy[2]), omega, oct);
}
extern "C" void oracle_marble(const float p[3], int32_t oct, float omega, float scale, float variation, float rgb[3]) {
    Tex X; X.type = 'm'; X.octaves = oct; X.omega = omega; X.scale = scale; X.variation = variation;
    Spec s = marble(X, V3(p[0], p[1], p[2]), V3(0, 0, 0), V3(0, 0, 0));
    rgb[0] = s.r; rgb[1] = s.g; rgb[2] = s.b;
}
// The subject transform of a .scene file (SceneLoader.cuh:59-77, Transform.cuh:119-240): out = the
// transformed point (3), then the transformed, renormalised normal (3).
extern "C" void oracle_transform(const float t[3], const float axis[3], float angle, float scale, const float p[3],
                                 const float n[3], float* out) {
    Xform x = makeXform(V3(t[0], t[1], t[2]), V3(axis[0], axis[1], axis[2]), angle, scale);
    V3 a = x.point(V3(p[0], p[1], p[2])), b = x.normal(V3(n[0], n[1], n[2]));
    out[0] = a.x; out[1] = a.y; out[2] = a.z; out[3] = b.x; out[4] = b.y; out[5] = b.z;
}
extern "C" float oracle_fr_dielectric(float c, float ei, float et) { return frDielectric(c, ei, et); }
extern "C" float oracle_roughness_to_alpha(float r) { return roughnessToAlpha(r); }
extern "C" void oracle_bxdf(int32_t kind, float param, int32_t mode, const float wo_[3], const float wi_[3], const float u_[2], float* out) {
    BxDF b;
    b.kind = kind; b.R = Spec(1.0f); b.T = Spec(1.0f);
    switch (kind) {
        case 0: b.type = REFL | DIFF; break;
        case 1: {
            b.type = REFL | DIFF;
            float sigma = param * (PI_F / 180), s2 = sigma * sigma;
            b.A = 1.0f - (s2 / (2.0f * (s2 + 0.33f)));
            b.B = 0.45f * s2 / (s2 + 0.09f);
            break;
        }
        case 2: b.type = REFL | SPEC; break;
        case 3: b.type = REFL | TRANS | SPEC; b.etaA = 1.0f; b.etaB = param; break;
        case 4: { b.type = REFL | GLOSSY; float a = roughnessToAlpha(param); b.dist = TR{a, a}; break; }
    }
    V3 wo(wo_[0], wo_[1], wo_[2]), wi(wi_[0], wi_[1], wi_[2]);
    if (mode == 0) { out[0] = b.f(wo, wi).r; return; }
    if (mode == 1) { out[0] = b.pdf(wo, wi); return; }
    V3 w; float pdf = 0; int st = b.type;
    Spec f = b.sampleF(wo, &w, V2(u_[0], u_[1]), &pdf, &st);
    out[0] = w.x; out[1] = w.y; out[2] = w.z; out[3] = pdf; out[4] = f.r; out[5] = (float)st;
}
