// ORACLE — test infrastructure only. Nothing in the product links, imports or
// executes this code; tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg use it as the checker and as the host-core baseline.
//
// CPU restatement of the reference's fp32 vector / spectrum arithmetic, written so
// that every expression keeps the reference's evaluation order (compile with
// -ffp-contract=off: the reference's results are defined per operation here, and
// the GPU path reproduces the same sequence).
//
// Citations are into tierChampion/LightTransportSimulation (src/...).
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <cfloat>

namespace orc {

// GeneralHelper.cuh:8-18: the reference defines M_PI / M_1_PI as float literals
// when the platform does not; the Windows/nvcc build takes these definitions.
static const float PI_F = 3.1415926535898f;
static const float INV_PI_F = 0.318309886183790671538f;
// ErrorHelper.cuh:40
static const float SHADOW_EPS = 0.0001f;

// Transcendentals: the reference calls CUDA's fp32 sinf/cosf/... whose ULP behaviour
// is not pinnable here. Both the oracle and the GPU path evaluate them in fp64 and
// round once to fp32, so they agree to the last bit.
inline float f_sin(float x) { return (float)std::sin((double)x); }
inline float f_cos(float x) { return (float)std::cos((double)x); }
inline float f_acos(float x) { return (float)std::acos((double)x); }
inline float f_log2(float x) { return (float)std::log2((double)x); }
inline float f_log(float x) { return (float)std::log((double)x); }
inline float f_tan(float x) { return (float)std::tan((double)x); }
inline float f_exp(float x) { return (float)std::exp((double)x); }

inline uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float bitsf(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

// ErrorHelper.cuh:77-85
inline float machineEpsilon() { return 0.5f * FLT_EPSILON; }
inline float gammaN(int n) { return ((float)n * machineEpsilon()) / (1 - (float)n * machineEpsilon()); }

// ErrorHelper.cuh:55-75
inline float nextFloatUp(float f) {
    if (std::isinf(f) && f > 0.0f) return f;
    if (f == -0.0f) f = 0.0f;
    uint32_t ui = fbits(f);
    if (f >= 0) ui++; else ui--;
    return bitsf(ui);
}
inline float nextFloatDown(float f) {
    if (std::isinf(f) && f > 0.0f) return f;
    if (f == -0.0f) f = 0.0f;
    uint32_t ui = fbits(f);
    if (f >= 0) ui--; else ui++;
    return bitsf(ui);
}

// GeneralHelper.cuh:58-63 (float clamp, also used with int arguments)
inline float clampf(float f, float mn, float mx) { return (f < mn) ? mn : ((f > mx) ? mx : f); }
inline float lerpf(float t, float v1, float v2) { return (1 - t) * v1 + t * v2; }

// Geometry.cuh:15-29 — Quake III rsqrt; `long` is 32-bit on the reference's MSVC build.
inline float quake(float f) {
    int32_t i; float x2, y;
    const float threeHalfs = 1.5f;
    x2 = f * 0.5f; y = f;
    std::memcpy(&i, &y, 4);
    i = 0x5f3759df - (i >> 1);
    std::memcpy(&y, &i, 4);
    y = y * (threeHalfs - (x2 * y * y));
    y = y * (threeHalfs - (x2 * y * y));
    return y;
}

struct V2 { float x = 0, y = 0; V2() {} V2(float a, float b) : x(a), y(b) {} };
struct V3 {
    float x = 0, y = 0, z = 0;
    V3() {}
    V3(float a, float b, float c) : x(a), y(b), z(c) {}
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 operator+(V3 a, V3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator-(V3 a) { return V3(-a.x, -a.y, -a.z); }
// Vector.cuh:141-143: operator*(U f) -> (f*x, f*y, f*z)
inline V3 operator*(V3 a, float f) { return V3(f * a.x, f * a.y, f * a.z); }
// Vector.cuh:146-150: division by a scalar multiplies by the reciprocal
inline V3 divs(V3 a, float f) { float inv = (float)1 / f; return V3(a.x * inv, a.y * inv, a.z * inv); }
inline V2 operator+(V2 a, V2 b) { return V2(a.x + b.x, a.y + b.y); }
inline V2 operator-(V2 a, V2 b) { return V2(a.x - b.x, a.y - b.y); }
inline V2 operator*(V2 a, float f) { return V2(f * a.x, f * a.y); }

inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float absDot(V3 a, V3 b) { return std::fabs(dot(a, b)); }
inline float lengthSq(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline float length(V3 a) { return std::sqrt(lengthSq(a)); }
inline V3 vabs(V3 a) { return V3(std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)); }
// Geometry.cuh:77-83 — cross product evaluated in fp64
inline V3 cross(V3 a, V3 b) {
    double ax = a.x, ay = a.y, az = a.z, bx = b.x, by = b.y, bz = b.z;
    return V3((float)((ay * bz) - (az * by)), (float)((az * bx) - (ax * bz)), (float)((ax * by) - (ay * bx)));
}
inline V3 normalize(V3 a) { return a * quake(lengthSq(a)); }
// Geometry.cuh:538-560
inline V3 faceTowards(V3 n, V3 v) { return (dot(n, v) < 0.0f) ? -n : n; }
// Geometry.cuh:251-259
inline void coordinateSystem(V3 v1, V3* v2, V3* v3) {
    if (std::fabs(v1.x) > std::fabs(v1.y))
        *v2 = V3(-v1.z, 0, v1.x) * quake(v1.x * v1.x + v1.z * v1.z);
    else
        *v2 = V3(0, v1.z, -v1.y) * quake(v1.y * v1.y + v1.z * v1.z);
    *v3 = cross(v1, *v2);
}
// Geometry.cuh:782-796
inline V3 offsetRayOrigin(V3 p, V3 pError, V3 n, V3 w) {
    float d = dot(vabs(n), pError);
    V3 offset = n * d;
    if (dot(w, n) < 0) offset = -offset;
    V3 po = p + offset;
    for (int i = 0; i < 3; i++) {
        if (offset[i] > 0) po[i] = nextFloatUp(po[i]);
        else if (offset[i] < 0) po[i] = nextFloatDown(po[i]);
    }
    return po;
}

// Spectrum.cuh:11-260 (RGB triple; division by a scalar is a true division)
struct Spec {
    float r = 0, g = 0, b = 0;
    Spec() {}
    explicit Spec(float v) : r(v), g(v), b(v) {}
    Spec(float a, float bb, float c) : r(a), g(bb), b(c) {}
    bool isBlack() const { return r == 0.0f && g == 0.0f && b == 0.0f; }
    // Spectrum.cuh:229-232
    float y() const { return 0.212671f * r + 0.715160f * g + 0.072169f * b; }
};
inline Spec operator+(Spec a, Spec b) { return Spec(a.r + b.r, a.g + b.g, a.b + b.b); }
inline Spec operator-(Spec a, Spec b) { return Spec(a.r - b.r, a.g - b.g, a.b - b.b); }
inline Spec operator*(Spec a, Spec b) { return Spec(a.r * b.r, a.g * b.g, a.b * b.b); }
inline Spec operator*(Spec a, float v) { return Spec(a.r * v, a.g * v, a.b * v); }
inline Spec operator*(float v, Spec a) { return a * v; }
inline Spec operator/(Spec a, float v) { return Spec(a.r / v, a.g / v, a.b / v); }
inline Spec& operator+=(Spec& a, Spec b) { a = a + b; return a; }
inline Spec lerpS(float t, Spec a, Spec b) { return (1 - t) * a + t * b; }

// Transform.cuh:131-240 — quaternion transform (translation, rotation, scale).
struct Quat { float w = 1, x = 0, y = 0, z = 0; Quat() {} Quat(float a, float b, float c, float d) : w(a), x(b), y(c), z(d) {} };
inline Quat qmul(Quat q1, Quat q2) {
    float rw = q1.w * q2.w - q1.x * q2.x - q1.y * q2.y - q1.z * q2.z;
    float rx = q1.w * q2.x + q1.x * q2.w + q1.y * q2.z - q1.z * q2.y;
    float ry = q1.w * q2.y - q1.x * q2.z + q1.y * q2.w + q1.z * q2.x;
    float rz = q1.w * q2.z + q1.x * q2.y - q1.y * q2.x + q1.z * q2.w;
    return Quat(rw, rx, ry, rz);
}
inline Quat qconj(Quat q) { return Quat(q.w, -q.x, -q.y, -q.z); }
struct Xform {
    V3 t; Quat q; float scale = 1;
    V3 point(V3 v) const {
        Quat p(0.0f, v.x, v.y, v.z);
        Quat r = qmul(qmul(q, p), qconj(q));
        return V3(r.x, r.y, r.z) * scale + t;
    }
    V3 normal(V3 n) const {
        Quat p(0.0f, n.x, n.y, n.z);
        Quat r = qmul(qmul(qconj(q), p), q);
        return normalize(V3(r.x, r.y, r.z));
    }
};
// Transform.cuh:119-124 + :162-166 (fmodf of the angle, host sin/cos)
inline Xform makeXform(V3 t, V3 axis, float theta, float scl) {
    Xform x; x.t = t; x.scale = scl;
    float th = std::fmod(theta, 2 * PI_F);
    float s = f_sin(th / 2);
    x.q = Quat(f_cos(th / 2), axis.x * s, axis.y * s, axis.z * s);
    return x;
}
// The identity Transform() applied to a vector or point returns it unchanged except
// that a negative zero becomes +0 (x*1 + 0).
inline V3 identityMap(V3 v) { return V3(v.x * 1.0f + 0.0f, v.y * 1.0f + 0.0f, v.z * 1.0f + 0.0f); }

} // namespace orc
