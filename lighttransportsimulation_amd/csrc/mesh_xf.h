// World transform of a mesh (shapes/Triangle.cu:26-41 with geometry/Transform.cuh:119-240):
// translation, unnormalised quaternion rotation and uniform scale. Points map to s (r p r*) + t;
// normals to normalize(r* n r), the inverse rotation (Transform::operator()(Normal3) conjugates
// first), normalised with the reference's Quake inverse square root. The same expressions serve
// the host (LTS_XF_FN inline) and the device (the kernel defines LTS_XF_FN as a device function
// before including this header), so both round every operation identically (no contraction).
#pragma once
#include <stdint.h>
#include <string.h>

#include <cmath>

#ifndef LTS_XF_FN
#define LTS_XF_FN inline
#endif

namespace lts {

struct MeshXf {
    float t[3] = {0.0f, 0.0f, 0.0f};
    float r[4] = {1.0f, 0.0f, 0.0f, 0.0f};  // w, x, y, z
    float s = 1.0f;
};

struct XfQ { float w, x, y, z; };

LTS_XF_FN XfQ xf_qmul(XfQ a, XfQ b) {
    return XfQ{a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z, a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
               a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x, a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w};
}

LTS_XF_FN float xf_rsqrt_quake(float f) {
    float half = f * 0.5f, y = f;
    int32_t i;
    memcpy(&i, &y, 4);
    i = 0x5f3759df - (i >> 1);
    memcpy(&y, &i, 4);
    y = y * (1.5f - (half * y * y));
    y = y * (1.5f - (half * y * y));
    return y;
}

LTS_XF_FN void xf_point(const MeshXf& xf, const float* p, float* out) {
    const XfQ r{xf.r[0], xf.r[1], xf.r[2], xf.r[3]};
    const XfQ o = xf_qmul(xf_qmul(r, XfQ{0.0f, p[0], p[1], p[2]}), XfQ{r.w, -r.x, -r.y, -r.z});
    out[0] = xf.s * o.x + xf.t[0];
    out[1] = xf.s * o.y + xf.t[1];
    out[2] = xf.s * o.z + xf.t[2];
}

LTS_XF_FN void xf_normal(const MeshXf& xf, const float* n, float* out) {
    const XfQ r{xf.r[0], xf.r[1], xf.r[2], xf.r[3]};
    const XfQ o = xf_qmul(xf_qmul(XfQ{r.w, -r.x, -r.y, -r.z}, XfQ{0.0f, n[0], n[1], n[2]}), r);
    const float k = xf_rsqrt_quake(o.x * o.x + o.y * o.y + o.z * o.z);
    out[0] = k * o.x;
    out[1] = k * o.y;
    out[2] = k * o.z;
}

// The subject transform of a scene file (core/SceneLoader.cuh:33-193): translation (3), rotation
// axis (3), angle in radians (taken modulo 2 pi), uniform scale. Host only.
inline MeshXf mesh_xf_from_tokens(const float tr[8]) {
    const float kPi = 3.1415926535898f;
    const float th = std::fmod(tr[6], 2 * kPi);
    const float sn = (float)std::sin((double)(th / 2));
    MeshXf xf;
    xf.t[0] = tr[0]; xf.t[1] = tr[1]; xf.t[2] = tr[2];
    xf.r[0] = (float)std::cos((double)(th / 2));
    xf.r[1] = tr[3] * sn; xf.r[2] = tr[4] * sn; xf.r[3] = tr[5] * sn;
    xf.s = tr[7];
    return xf;
}

}  // namespace lts
