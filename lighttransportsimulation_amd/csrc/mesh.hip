// Mesh world transform on the GPU (shapes/Triangle.cu:26-41, the TriangleMesh constructor's loops
// over points and normals), with the expressions of mesh_xf.h compiled as device code.
#include <hip/hip_runtime.h>

#define LTS_XF_FN __device__ __forceinline__
#include "mesh_xf.h"
#include "lts_internal.h"

namespace lts {

namespace {

constexpr int kXfBlock = 256;

// Thread i transforms point i and normal i in place (either may be past its array's end).
__global__ __launch_bounds__(kXfBlock) void k_mesh_transform(MeshXf xf, float* __restrict__ P, int nP, float* __restrict__ N,
                                                             int nN) {
    const int i = blockIdx.x * kXfBlock + threadIdx.x;
    if (i < nP) {
        const float p[3] = {P[3 * (size_t)i], P[3 * (size_t)i + 1], P[3 * (size_t)i + 2]};
        float o[3];
        xf_point(xf, p, o);
        P[3 * (size_t)i] = o[0]; P[3 * (size_t)i + 1] = o[1]; P[3 * (size_t)i + 2] = o[2];
    }
    if (i < nN) {
        const float n[3] = {N[3 * (size_t)i], N[3 * (size_t)i + 1], N[3 * (size_t)i + 2]};
        float o[3];
        xf_normal(xf, n, o);
        N[3 * (size_t)i] = o[0]; N[3 * (size_t)i + 1] = o[1]; N[3 * (size_t)i + 2] = o[2];
    }
}

}  // namespace

hipError_t mesh_transform(const MeshXf& xf, float* P, int nP, float* N, int nN, hipStream_t s) {
    const int m = nP > nN ? nP : nN;
    if (m <= 0) return hipSuccess;
    k_mesh_transform<<<(m + kXfBlock - 1) / kXfBlock, kXfBlock, 0, s>>>(xf, P, nP, N, nN);
    return hipGetLastError();
}

}  // namespace lts
