// Device-side arithmetic of the path tracer (HIP, gfx950).
//
// Every expression keeps the reference's fp32 evaluation order (the library is compiled with
// -ffp-contract=off, IEEE divide/sqrt and denormals kept) so that a render is a pure function of
// the scene and the sampler and the CPU restatement can check it bit for bit. Reference routines
// are cited next to each function (paths inside tierChampion/LightTransportSimulation/src).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define LTS_D __device__ __forceinline__

namespace lts {

constexpr float PI_F = 3.1415926535898f;              // core/GeneralHelper.cuh:8-10
constexpr float INV_PI_F = 0.318309886183790671538f;  // core/GeneralHelper.cuh:12-14
constexpr float SHADOW_EPS = 0.0001f;                 // core/ErrorHelper.cuh:40

// Transcendentals: evaluated in fp64 and rounded once, a last-bit-stable definition of the fp32
// function shared by every code path.
LTS_D float f_sin(float x) { return (float)sin((double)x); }
LTS_D float f_cos(float x) { return (float)cos((double)x); }
LTS_D float f_acos(float x) { return (float)acos((double)x); }
LTS_D float f_log2(float x) { return (float)log2((double)x); }
LTS_D float f_log(float x) { return (float)log((double)x); }

LTS_D uint32_t fbits(float f) { return __float_as_uint(f); }
LTS_D float bitsf(uint32_t u) { return __uint_as_float(u); }

// core/ErrorHelper.cuh:77-85
LTS_D float gammaN(int n) {
    const float me = 0.5f * 1.19209290e-07f;
    return ((float)n * me) / (1 - (float)n * me);
}
// core/ErrorHelper.cuh:55-75
LTS_D float nextFloatUp(float f) {
    if (isinf(f) && f > 0.0f) return f;
    if (f == -0.0f) f = 0.0f;
    uint32_t ui = fbits(f);
    if (f >= 0) ui++; else ui--;
    return bitsf(ui);
}
LTS_D float nextFloatDown(float f) {
    if (isinf(f) && f > 0.0f) return f;
    if (f == -0.0f) f = 0.0f;
    uint32_t ui = fbits(f);
    if (f >= 0) ui--; else ui++;
    return bitsf(ui);
}
LTS_D float clampf(float f, float mn, float mx) { return (f < mn) ? mn : ((f > mx) ? mx : f); }
LTS_D float lerpf(float t, float a, float b) { return (1 - t) * a + t * b; }

// geometry/Geometry.cuh:15-29 (Quake III rsqrt with 32-bit integer punning)
LTS_D float quake(float f) {
    float x2 = f * 0.5f, y = f;
    int32_t i = __float_as_int(y);
    i = 0x5f3759df - (i >> 1);
    y = __int_as_float(i);
    y = y * (1.5f - (x2 * y * y));
    y = y * (1.5f - (x2 * y * y));
    return y;
}

struct V2 { float x, y; };
struct V3 { float x, y, z; };
LTS_D V2 v2(float a, float b) { V2 r; r.x = a; r.y = b; return r; }
LTS_D V3 v3(float a, float b, float c) { V3 r; r.x = a; r.y = b; r.z = c; return r; }
LTS_D float comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
LTS_D V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
LTS_D V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
LTS_D V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
LTS_D V3 operator*(V3 a, float f) { return v3(f * a.x, f * a.y, f * a.z); }  // Vector.cuh:141-143
LTS_D V2 operator+(V2 a, V2 b) { return v2(a.x + b.x, a.y + b.y); }
LTS_D V2 operator-(V2 a, V2 b) { return v2(a.x - b.x, a.y - b.y); }
LTS_D V2 operator*(V2 a, float f) { return v2(f * a.x, f * a.y); }
LTS_D float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
LTS_D float absDot(V3 a, V3 b) { return fabsf(dot(a, b)); }
LTS_D float lengthSq(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
LTS_D float length(V3 a) { return sqrtf(lengthSq(a)); }
LTS_D V3 vabs(V3 a) { return v3(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }
// geometry/Geometry.cuh:77-83 (fp64 cross product)
LTS_D V3 cross(V3 a, V3 b) {
    double ax = a.x, ay = a.y, az = a.z, bx = b.x, by = b.y, bz = b.z;
    return v3((float)((ay * bz) - (az * by)), (float)((az * bx) - (ax * bz)), (float)((ax * by) - (ay * bx)));
}
LTS_D V3 normalize(V3 a) { return a * quake(lengthSq(a)); }
LTS_D V3 faceTowards(V3 n, V3 v) { return (dot(n, v) < 0.0f) ? -n : n; }
// geometry/Geometry.cuh:251-259
LTS_D void coordinateSystem(V3 a, V3* b, V3* c) {
    if (fabsf(a.x) > fabsf(a.y)) *b = v3(-a.z, 0, a.x) * quake(a.x * a.x + a.z * a.z);
    else *b = v3(0, a.z, -a.y) * quake(a.y * a.y + a.z * a.z);
    *c = cross(a, *b);
}
// geometry/Geometry.cuh:782-796
LTS_D V3 offsetRayOrigin(V3 p, V3 pError, V3 n, V3 w) {
    float d = dot(vabs(n), pError);
    V3 off = n * d;
    if (dot(w, n) < 0) off = -off;
    V3 po = p + off;
    if (off.x > 0) po.x = nextFloatUp(po.x); else if (off.x < 0) po.x = nextFloatDown(po.x);
    if (off.y > 0) po.y = nextFloatUp(po.y); else if (off.y < 0) po.y = nextFloatDown(po.y);
    if (off.z > 0) po.z = nextFloatUp(po.z); else if (off.z < 0) po.z = nextFloatDown(po.z);
    return po;
}
// identity Transform(): -0 becomes +0, everything else passes through
LTS_D V3 identityMap(V3 v) { return v3(v.x * 1.0f + 0.0f, v.y * 1.0f + 0.0f, v.z * 1.0f + 0.0f); }

// rendering/Spectrum.cuh
struct Spec { float r, g, b; };
LTS_D Spec spec(float v) { Spec s; s.r = s.g = s.b = v; return s; }
LTS_D Spec spec(float a, float b, float c) { Spec s; s.r = a; s.g = b; s.b = c; return s; }
LTS_D bool isBlack(Spec a) { return a.r == 0.0f && a.g == 0.0f && a.b == 0.0f; }
LTS_D float lum(Spec a) { return 0.212671f * a.r + 0.715160f * a.g + 0.072169f * a.b; }
LTS_D Spec operator+(Spec a, Spec b) { return spec(a.r + b.r, a.g + b.g, a.b + b.b); }
LTS_D Spec operator*(Spec a, Spec b) { return spec(a.r * b.r, a.g * b.g, a.b * b.b); }
LTS_D Spec operator*(Spec a, float v) { return spec(a.r * v, a.g * v, a.b * v); }
LTS_D Spec operator*(float v, Spec a) { return a * v; }
LTS_D Spec operator/(Spec a, float v) { return spec(a.r / v, a.g / v, a.b / v); }
LTS_D Spec lerpS(float t, Spec a, Spec b) { return (1 - t) * a + t * b; }

// ------------------------------------------------------------------ counter-based sampler
// Philox-4x32-10 (Salmon et al., SC'11) keyed by (seed, pixel), counter (sample, dim / 4, 0, 0).
struct U4 { uint32_t a, b, c, d; };
LTS_D U4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    U4 o; o.a = c0; o.b = c1; o.c = c2; o.d = c3;
    return o;
}
LTS_D float u01(uint32_t x) { return (float)((x >> 8) + 1u) * (1.0f / 16777216.0f); }

struct Sampler {
    uint32_t seed, pixel, sample;
    // Draw block b holds dims 4b..4b+3.
    LTS_D void block(uint32_t b, float out[4]) const {
        U4 r = philox(sample, b, 0u, 0u, seed, pixel);
        out[0] = u01(r.a); out[1] = u01(r.b); out[2] = u01(r.c); out[3] = u01(r.d);
    }
};

// Decoded plane of a quantized render-node box: origin + q * 2^e. The product is exact (q has 8
// bits and e is at least -126), so the only rounding is the addition, and the fused form is one
// instruction; the builder evaluates this same expression when it picks q, which keeps decoded
// boxes conservative.
LTS_D float qdecode(float origin, int e, unsigned q) { return __builtin_fmaf((float)q, __int_as_float((e + 127) << 23), origin); }

}  // namespace lts
