// Learned radiance cache network on CDNA4 matrix cores (v_mfma_f32_16x16x32_{bf16,f16}).
//
// Orientation. Every layer is evaluated as Y^T = W^T X^T: the weights are the A operand (16 output
// neurons x 32 inputs per fragment) and 16 samples are the B operand's columns. The accumulator of
// a 16x16 tile holds, in lane l, neurons 4(l/16)+i of sample l%16, which is already the sample
// column the next layer's B operand needs; only the order of the K (input-neuron) index differs.
// Since an MFMA sums over K, any permutation of K applied to both operands gives the same product,
// so the weight fragments are stored with K permuted to match the accumulator registers:
//   K slot (ks, g, j) of lane group g = l/16  <->  neuron 16 (2 ks + j/4) + 4 g + (j % 4).
// The input layer's K order is free; its slot (ks, g, j) holds encoding feature 4 (8 ks + j) + g so
// that every unrolled slot computes one kind of feature in all lanes.
// Activations therefore pass from layer to layer in registers (fp32 accumulate, ReLU, round to
// bf16/fp16), with no LDS round trip. Weight fragments sit in LDS (43 KB forward, 49 KB backward).
//
// Training step (batch B): one fused kernel runs forward, relative-L2 loss and the backward pass of
// the activations for 16-sample tiles, writing each layer's input H_l and output gradient G_l
// (rounded as used); a split-K MFMA GEMM forms dW_l = H_l G_l^T per 1024-sample split; one kernel
// sums the splits in a fixed order and applies Adam, writing the fp32 master weights and both
// fragment copies. Every reduction has a fixed order, so a step is deterministic.
#include "lts_device.h"
#include "lts_internal.h"

namespace lts {

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

template <typename T> struct Vec8;
template <> struct Vec8<__bf16> { typedef bf16x8 type; };
template <> struct Vec8<_Float16> { typedef f16x8 type; };

LTS_D f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
LTS_D f32x4 mfma(f16x8 a, f16x8 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }

constexpr int kFwdVec = kNrcParams / 8;    // 16-byte fragment rows of the forward weights
constexpr int kBwdVec = kNrcBwdElems / 8;
constexpr int kTrainBlock = 512;

// Neuron held by K slot (ks, g, j) of a hidden layer's input.
LTS_D int slot_neuron(int ks, int g, int j) { return 16 * (2 * ks + (j >> 2)) + 4 * g + (j & 3); }

// Encoding feature held by K slot (ks, g, j) of the input layer. The input layer's K order is free
// (only its weight fragments must agree), so slot (ks, j) holds features 4 (8 ks + j) + g of the four
// lane groups: four features of one kind, the kind fixed by (ks, j) once the loops are unrolled.
LTS_D int slot_feature(int ks, int g, int j) { return 4 * (8 * ks + j) + g; }

LTS_D int w_index(int l, int in, int out) { return l < 5 ? l * 4096 + in * 64 + out : 20480 + in * 16 + out; }

// One-blob bin value: triweight kernel (1 - d^2/9)^3 of the distance d to the bin centre in
// units of the bin width 1/4 (support 3 bins).
LTS_D float one_blob(float s, int bin) {
    const float d = (s - ((float)bin + 0.5f) * 0.25f) * 4.0f;
    const float u = fmaxf(1.0f - d * d * (1.0f / 9.0f), 0.0f);
    return u * u * u;
}

// Encoding feature idx (0..63) of one raw record: 36 triangle waves of frequency 2^(f-1) on the
// normalised position, one-blob (4 bins) of the direction and normal angles and of roughness,
// diffuse and specular reflectance, and two constant ones.
// Features 4m + g (g = lane group) of block m; m is a compile-time constant in the unrolled callers,
// so the kind of feature is resolved without branches.
LTS_D float nrc_feature_block(const float* __restrict__ raw, int m, int g) {
    if (m < 9) {
        // triangle waves: coordinate d = 4m / 12 is the same for the block, frequency (4m % 12) + g
        const int d = (4 * m) / 12, f = (4 * m) % 12 + g;
        const float x = ldexpf(raw[d], f - 1);
        const float fr = x - floorf(x);
        return fabsf(2.0f * fr - 1.0f);
    }
    if (m < 13) return one_blob(raw[3 + (m - 9)], g);
    if (m == 13) return one_blob(raw[7], g);
    if (m == 14) return raw[8 + g];
    return g < 2 ? raw[12 + g] : 1.0f;
}

// The fields of a raw record one lane reads (raw[0..7], raw[8 + g], raw[12 + g]), so the next tile's
// can be loaded while the current one is evaluated.
struct RawLane { float4 a, b; float c, d; };

LTS_D RawLane load_raw_lane(const float* __restrict__ raw, int g) {
    RawLane r;
    const float4* q = reinterpret_cast<const float4*>(raw);
    r.a = q[0];
    r.b = q[1];
    r.c = raw[8 + g];
    r.d = g < 2 ? raw[12 + g] : 1.0f;
    return r;
}

// nrc_feature_block on the lane's fields (the same operations on the same values).
LTS_D float nrc_feature_regs(const RawLane& r, int m, int g) {
    if (m < 9) {
        const int d = (4 * m) / 12, f = (4 * m) % 12 + g;
        const float x = ldexpf(d == 0 ? r.a.x : d == 1 ? r.a.y : r.a.z, f - 1);
        const float fr = x - floorf(x);
        return fabsf(2.0f * fr - 1.0f);
    }
    if (m < 13) {
        const int k = m - 9;
        return one_blob(k == 0 ? r.a.w : k == 1 ? r.b.x : k == 2 ? r.b.y : r.b.z, g);
    }
    if (m == 13) return one_blob(r.b.w, g);
    if (m == 14) return r.c;
    return r.d;
}

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef int i32x4 __attribute__((ext_vector_type(4)));

// The next layer's B fragment from two accumulator tiles: ReLU as a signed integer max with 0 on
// the float bits (negative values and -0 have the sign bit set and become +0; for every input but
// NaN this is max(x, 0)), then rounding to T in pairs. One integer op per value and one packed
// conversion per two, where a float max canonicalises its input first (two ops per value) and
// single conversions need a permute to pack.
template <typename T>
LTS_D typename Vec8<T>::type relu_fragment(const f32x4& lo, const f32x4& hi) {
    typedef T t2 __attribute__((ext_vector_type(2)));
    typedef T t4 __attribute__((ext_vector_type(4)));
    const f32x4 a = __builtin_bit_cast(f32x4, __builtin_elementwise_max(__builtin_bit_cast(i32x4, lo), (i32x4)0));
    const f32x4 b = __builtin_bit_cast(f32x4, __builtin_elementwise_max(__builtin_bit_cast(i32x4, hi), (i32x4)0));
    const t2 p0 = __builtin_convertvector((f32x2{a[0], a[1]}), t2);
    const t2 p1 = __builtin_convertvector((f32x2{a[2], a[3]}), t2);
    const t2 p2 = __builtin_convertvector((f32x2{b[0], b[1]}), t2);
    const t2 p3 = __builtin_convertvector((f32x2{b[2], b[3]}), t2);
    const t4 q0 = __builtin_shufflevector(p0, p1, 0, 1, 2, 3), q1 = __builtin_shufflevector(p2, p3, 0, 1, 2, 3);
    return __builtin_shufflevector(q0, q1, 0, 1, 2, 3, 4, 5, 6, 7);
}

template <typename T>
LTS_D typename Vec8<T>::type lds_frag(const uint4* w, int row) {
    const uint4 q = w[row];
    return __builtin_bit_cast(typename Vec8<T>::type, q);
}

// Copies `count` 16-byte rows from global to LDS with the whole block.
LTS_D void stage(uint4* dst, const uint16_t* src, int count) {
    const uint4* s = reinterpret_cast<const uint4*>(src);
    for (int i = threadIdx.x; i < count; i += blockDim.x) dst[i] = s[i];
}

// Forward pass of one 16-sample tile; h[l] receives the B-operand fragments of layer l's input.
// With Reload, weight fragments are read from LDS layer by layer (a compiler barrier keeps the
// loads from being hoisted out of the tile loop and held in registers across it).
template <typename T, bool Reload>
LTS_D f32x4 forward_tile(const uint4* wf, const float* __restrict__ raw, bool valid, int lane,
                         typename Vec8<T>::type (&h)[6][2]) {
    const int g = lane >> 4;
#pragma unroll
    for (int ks = 0; ks < 2; ks++)
#pragma unroll
        for (int j = 0; j < 8; j++) h[0][ks][j] = valid ? (T)nrc_feature_block(raw, 8 * ks + j, g) : (T)0.0f;
#pragma unroll
    for (int l = 0; l < 5; l++) {
        if (Reload) asm volatile("" ::: "memory");
        f32x4 acc[4];
#pragma unroll
        for (int mt = 0; mt < 4; mt++) {
            acc[mt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < 2; ks++) acc[mt] = mfma(lds_frag<T>(wf, l * 512 + (mt * 2 + ks) * 64 + lane), h[l][ks], acc[mt]);
        }
#pragma unroll
        for (int ks = 0; ks < 2; ks++) h[l + 1][ks] = relu_fragment<T>(acc[2 * ks], acc[2 * ks + 1]);
    }
    if (Reload) asm volatile("" ::: "memory");
    f32x4 y = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < 2; ks++) y = mfma(lds_frag<T>(wf, 2560 + ks * 64 + lane), h[5][ks], y);
    return y;  // lane group 0 holds outputs 0..3 of its sample
}

// Forward pass of NT 16-sample tiles at once (independent MFMA chains interleaved); only the current
// layer's activations are kept.
template <typename T, int NT>
LTS_D void forward_tiles(const uint4* wf, const RawLane (&raw)[NT], const bool (&valid)[NT], int lane, f32x4 (&y)[NT]) {
    const int g = lane >> 4;
    typename Vec8<T>::type h[NT][2];
#pragma unroll
    for (int u = 0; u < NT; u++)
#pragma unroll
        for (int ks = 0; ks < 2; ks++)
#pragma unroll
            for (int j = 0; j < 8; j++) h[u][ks][j] = valid[u] ? (T)nrc_feature_regs(raw[u], 8 * ks + j, g) : (T)0.0f;
#pragma unroll
    for (int l = 0; l < 5; l++) {
        // a compiler barrier per layer: only this layer's weight fragments are live in registers
        asm volatile("" ::: "memory");
        f32x4 acc[NT][4];
#pragma unroll
        for (int mt = 0; mt < 4; mt++)
#pragma unroll
            for (int u = 0; u < NT; u++) acc[u][mt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 2; ks++)
#pragma unroll
            for (int mt = 0; mt < 4; mt++) {
                const typename Vec8<T>::type w = lds_frag<T>(wf, l * 512 + (mt * 2 + ks) * 64 + lane);
#pragma unroll
                for (int u = 0; u < NT; u++) acc[u][mt] = mfma(w, h[u][ks], acc[u][mt]);
            }
#pragma unroll
        for (int u = 0; u < NT; u++)
#pragma unroll
            for (int ks = 0; ks < 2; ks++) h[u][ks] = relu_fragment<T>(acc[u][2 * ks], acc[u][2 * ks + 1]);
    }
#pragma unroll
    for (int u = 0; u < NT; u++) y[u] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < 2; ks++) {
        const typename Vec8<T>::type w = lds_frag<T>(wf, 2560 + ks * 64 + lane);
#pragma unroll
        for (int u = 0; u < NT; u++) y[u] = mfma(w, h[u][ks], y[u]);
    }
}

constexpr int kInferTiles = 2;   // 16-sample tiles per wave iteration (each weight fragment read serves both)
constexpr int kInferBlock = 512;  // 8 waves share one LDS copy of the weights

// Inference over records 0 .. n-1 in sequence. Each wave takes a contiguous run of tiles, so its
// loads walk memory in order. Without meta, output c of record s goes to out[3 s + c]; with meta
// (the cached pass's queries), sampleL[3 slot + c] = L_c + beta_c * max(y_c, 0) * a_c.
template <typename T>
__global__ __launch_bounds__(kInferBlock) void k_nrc_infer(const uint16_t* __restrict__ fwd, const float* __restrict__ raw,
                                                   const int32_t* __restrict__ nDev, int nCap, float* __restrict__ out,
                                                   const NrcQuery* __restrict__ meta, float* __restrict__ sampleL) {
    __shared__ uint4 wf[kFwdVec];
    stage(wf, fwd, kFwdVec);
    __syncthreads();
    const int n = min(*nDev, nCap);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int groups = (n + 16 * kInferTiles - 1) / (16 * kInferTiles);
    const int waves = gridDim.x * (kInferBlock / 64), w = blockIdx.x * (kInferBlock / 64) + wave;
    const int per = (groups + waves - 1) / waves;
    const int g0 = w * per, g1 = min(groups, g0 + per);
    const int g = lane >> 4;
    // tile t's record fields are loaded during tile t - 1
    RawLane next[kInferTiles];
    auto fetch = [&](int t) {
#pragma unroll
        for (int u = 0; u < kInferTiles; u++) {
            const int su = (t * kInferTiles + u) * 16 + (lane & 15);
            next[u] = load_raw_lane(raw + (size_t)(su < n ? su : 0) * kNrcRaw, g);
        }
    };
    if (g0 < g1) fetch(g0);
    for (int t = g0; t < g1; t++) {
        int s[kInferTiles];
        bool valid[kInferTiles];
        RawLane cur[kInferTiles];
#pragma unroll
        for (int u = 0; u < kInferTiles; u++) {
            s[u] = (t * kInferTiles + u) * 16 + (lane & 15);
            valid[u] = s[u] < n;
            cur[u] = next[u];
        }
        if (t + 1 < g1) fetch(t + 1);
        f32x4 y[kInferTiles];
        forward_tiles<T, kInferTiles>(wf, cur, valid, lane, y);
#pragma unroll
        for (int u = 0; u < kInferTiles; u++) {
            if (lane >= 16 || !valid[u]) continue;
            if (meta) {
                const NrcQuery q = meta[s[u]];
                float* o = sampleL + 3 * (size_t)q.slot;
                o[0] = q.L[0] + q.beta[0] * fmaxf(y[u][0], 0.0f) * q.a[0];
                o[1] = q.L[1] + q.beta[1] * fmaxf(y[u][1], 0.0f) * q.a[1];
                o[2] = q.L[2] + q.beta[2] * fmaxf(y[u][2], 0.0f) * q.a[2];
            } else {
                out[3 * (size_t)s[u] + 0] = y[u][0];
                out[3 * (size_t)s[u] + 1] = y[u][1];
                out[3 * (size_t)s[u] + 2] = y[u][2];
            }
        }
    }
}

// Forward + loss + backward of the activations. H / G hold, per layer (stride 64 ld), blocks of 32
// samples laid out [neuron][32] (element (l, neuron, s) at l 64 ld + (s/32) 2048 + 32 neuron + s%32),
// so the GEMM's 8-sample operand rows are 16 contiguous bytes. Samples in [n, roundup(n, 32)) are
// written as zeros so the weight-gradient GEMM can run over whole 32-sample steps.
template <typename T>
__global__ __launch_bounds__(kTrainBlock) void k_nrc_fwdbwd(const uint16_t* __restrict__ fwd, const uint16_t* __restrict__ bwd,
                                                            const float* __restrict__ raw, const float* __restrict__ target,
                                                            const NrcState* __restrict__ state, int step, int nCap,
                                                            T* __restrict__ H, T* __restrict__ G, int ld,
                                                            float* __restrict__ lossPartial) {
    __shared__ uint4 wf[kFwdVec];
    __shared__ uint4 wb[kBwdVec];
    __shared__ float red[kTrainBlock];
    stage(wf, fwd, kFwdVec);
    stage(wb, bwd, kBwdVec);
    __syncthreads();
    const int n = min(state->stepN[step], nCap);
    const int npad = (n + 31) & ~31;
    raw += (size_t)state->stepFirst[step] * kNrcRaw;
    target += (size_t)state->stepFirst[step] * 3;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4;
    const int wavesPerBlock = kTrainBlock / 64;
    float lossAcc = 0.0f;
    for (int t = blockIdx.x * wavesPerBlock + wave; t < (npad >> 4); t += gridDim.x * wavesPerBlock) {
        const int s = t * 16 + (lane & 15);
        const bool valid = s < n;
        typename Vec8<T>::type h[6][2];
        const f32x4 y = forward_tile<T, true>(wf, raw + (size_t)(valid ? s : 0) * kNrcRaw, valid, lane, h);
        // relative L2 loss (y - t)^2 / (y^2 + 0.01) with the denominator held constant
        float dy[4] = {0.f, 0.f, 0.f, 0.f};
        if (g == 0 && valid) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const float diff = y[c] - target[3 * (size_t)s + c];
                const float den = y[c] * y[c] + 0.01f;
                lossAcc += diff * diff / den;
                dy[c] = 2.0f * diff / den;
            }
        }
        typename Vec8<T>::type gb[2];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            gb[0][j] = j < 4 ? (T)dy[j & 3] : (T)0.0f;
            gb[1][j] = (T)0.0f;
        }
        // layer inputs H_0..H_5 and the output gradient G_5 (rows 4g + j); neuron 4g + c lives at
        // 32 (4g + c) within the sample's block, so only c varies between the stores
        const size_t layer = (size_t)64 * ld;
        T* Hs = H + (size_t)(s >> 5) * 2048 + (s & 31) + 128 * g;
        T* Gs = G + (size_t)(s >> 5) * 2048 + (s & 31) + 128 * g;
        T* H0 = H + (size_t)(s >> 5) * 2048 + (s & 31) + 32 * g;  // input layer: feature 4 (8 ks + j) + g
#pragma unroll
        for (int ks = 0; ks < 2; ks++)
#pragma unroll
            for (int j = 0; j < 8; j++) H0[32 * 4 * (8 * ks + j)] = h[0][ks][j];
#pragma unroll
        for (int l = 1; l < 6; l++)
#pragma unroll
            for (int ks = 0; ks < 2; ks++)
#pragma unroll
                for (int j = 0; j < 8; j++) Hs[l * layer + 32 * (slot_neuron(ks, 0, j))] = h[l][ks][j];
#pragma unroll
        for (int j = 0; j < 4; j++) Gs[5 * layer + 32 * j] = gb[0][j];
        // back-propagate: g_l = (W_l g_{l+1}) * [h_l > 0], stored as G_{l-1}
#pragma unroll
        for (int l = 5; l >= 1; l--) {
            asm volatile("" ::: "memory");
            f32x4 acc[4];
#pragma unroll
            for (int it = 0; it < 4; it++) {
                acc[it] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int ks = 0; ks < (l == 5 ? 1 : 2); ks++)
                    acc[it] = mfma(lds_frag<T>(wb, l * 512 + (it * 2 + ks) * 64 + lane), gb[ks], acc[it]);
            }
#pragma unroll
            for (int ks = 0; ks < 2; ks++)
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const float v = acc[2 * ks + (j >> 2)][j & 3];
                    gb[ks][j] = (float)h[l][ks][j] > 0.0f ? (T)v : (T)0.0f;
                    Gs[(l - 1) * layer + 32 * slot_neuron(ks, 0, j)] = gb[ks][j];
                }
        }
    }
    red[threadIdx.x] = lossAcc;
    __syncthreads();
    for (int w = kTrainBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) lossPartial[blockIdx.x] = red[0];
}

// dW_l = H_l G_l^T: one wave per (layer, 256-sample split) computes the whole layer (16 tiles of
// 16 x 16; the output layer 4), loading 4 A and 4 B fragments per 32-sample step for 16 MFMAs.
// Block: 4 waves = 4 consecutive splits of layer blockIdx.x, summed in LDS in wave order into one
// partial per block (1024 samples).
template <typename T>
__global__ __launch_bounds__(256) void k_nrc_wgrad(const T* __restrict__ H, const T* __restrict__ G, int ld,
                                                   const NrcState* __restrict__ state, int step, int nCap,
                                                   float* __restrict__ partial) {
    using V = typename Vec8<T>::type;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4, col = lane & 15;
    const int l = blockIdx.x;
    const int split = blockIdx.y * 4 + wave;
    const int nOt = l < 5 ? 4 : 1;
    const int n = min(state->stepN[step], nCap);
    const int npad = (n + 31) & ~31;
    const int k0 = split * kNrcKSplit, k1 = min(npad, k0 + kNrcKSplit);
    const T* a = H + (size_t)l * 64 * ld + (size_t)(k0 >> 5) * 2048 + 32 * col + 8 * g;
    const T* b = G + (size_t)l * 64 * ld + (size_t)(k0 >> 5) * 2048 + 32 * col + 8 * g;
    f32x4 acc[4][4];
#pragma unroll
    for (int it = 0; it < 4; it++)
#pragma unroll
        for (int ot = 0; ot < 4; ot++) acc[it][ot] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k = k0; k < k1; k += 32, a += 2048, b += 2048) {
        V A[4], B[4];
#pragma unroll
        for (int it = 0; it < 4; it++) A[it] = *reinterpret_cast<const V*>(a + 512 * it);
#pragma unroll
        for (int ot = 0; ot < 4; ot++)
            if (ot < nOt) B[ot] = *reinterpret_cast<const V*>(b + 512 * ot);
#pragma unroll
        for (int it = 0; it < 4; it++)
#pragma unroll
            for (int ot = 0; ot < 4; ot++)
                if (ot < nOt) acc[it][ot] = mfma(A[it], B[ot], acc[it][ot]);
    }
    // per wave: element (it, ot, i, lane) of the layer's 64 x 64 (or 64 x 16) block
    __shared__ float red[4][4096];
#pragma unroll
    for (int it = 0; it < 4; it++)
#pragma unroll
        for (int ot = 0; ot < 4; ot++)
#pragma unroll
            for (int i = 0; i < 4; i++) red[wave][((it * 4 + ot) * 4 + i) * 64 + lane] = acc[it][ot][i];
    __syncthreads();
    float* p = partial + (size_t)blockIdx.y * kNrcParams;
    for (int e = threadIdx.x; e < 4096; e += 256) {
        const int ln = e & 63, i = (e >> 6) & 3, ot = (e >> 8) & 3, it = e >> 10;
        if (ot >= nOt) continue;
        const float v = ((red[0][e] + red[1][e]) + red[2][e]) + red[3][e];
        p[w_index(l, 16 * it + 4 * (ln >> 4) + i, 16 * ot + (ln & 15))] = v;
    }
}

// Plans the steps of one training call: batch window of step k (cycling over the collected batches)
// and Adam's bias corrections for step t + k + 1 (nothing advances when no samples were collected).
__global__ void k_nrc_plan(NrcState* __restrict__ state, int steps, int batchCap) {
    const int k = threadIdx.x;
    if (k >= steps) return;
    const int total = state->nTotal;
    const int chunks = max(1, (total + batchCap - 1) / batchCap);
    const int c = k % chunks;
    state->stepFirst[k] = c * batchCap;
    state->stepN[k] = max(0, min(batchCap, total - c * batchCap));
    const double t = (double)state->t + k + 1;
    state->stepC1[k] = (float)(1.0 / (1.0 - pow((double)state->beta1, t)));
    state->stepC2[k] = (float)(1.0 / (1.0 - pow((double)state->beta2, t)));
}

// After the steps of a training call: advance the step counter and record the last step's loss.
__global__ void k_nrc_finish(NrcState* __restrict__ state, int steps, const float* __restrict__ lossPartial, int nPartial) {
    if (threadIdx.x != 0) return;
    const int n = state->stepN[steps - 1];
    if (n <= 0) return;
    state->t += steps;
    float s = 0.0f;
    for (int i = 0; i < nPartial; i++) s += lossPartial[i];
    state->loss = s / (float)n;
}

LTS_D int fwd_pos(int l, int in, int out) {
    // A operand of W^T: row = out, K slot of `in` (input layer: slot_feature, else slot_neuron)
    int ks, j, g;
    if (l == 0) {
        const int m = in >> 2;
        g = in & 3; ks = m >> 3; j = m & 7;
    } else {
        const int q = in >> 4, r = in & 15;
        ks = q >> 1; j = ((q & 1) << 2) | (r & 3); g = r >> 2;
    }
    const int lane = 16 * g + (out & 15);
    return l * 4096 + (((out >> 4) * 2 + ks) * 64 + lane) * 8 + j;
}

LTS_D int bwd_pos(int l, int in, int out) {
    // A operand of W: row = in, K slot of `out`
    const int q = out >> 4, r = out & 15;
    const int ks = q >> 1, j = ((q & 1) << 2) | (r & 3), g = r >> 2;
    const int lane = 16 * g + (in & 15);
    return l * 4096 + (((in >> 4) * 2 + ks) * 64 + lane) * 8 + j;
}

LTS_D void decode_param(int p, int& l, int& in, int& out) {
    if (p < 20480) { l = p >> 12; in = (p >> 6) & 63; out = p & 63; }
    else { l = 5; in = (p - 20480) >> 4; out = (p - 20480) & 15; }
}

template <typename T>
__global__ __launch_bounds__(256) void k_nrc_adam(const float* __restrict__ partial, int splits, const NrcState* __restrict__ state,
                                                  int step, int nCap, float* __restrict__ master, float* __restrict__ m,
                                                  float* __restrict__ v, float* __restrict__ gradOut, T* __restrict__ fwd,
                                                  T* __restrict__ bwd) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= kNrcParams) return;
    const int n = min(state->stepN[step], nCap);
    if (n <= 0) return;
    const float lr = state->lr, beta1 = state->beta1, beta2 = state->beta2, eps = state->eps;
    int l, in, out;
    decode_param(p, l, in, out);
    if (l == 5 && out >= 3) return;  // padding columns stay zero
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;  // fixed-order reduction with 4 independent chains
    int s = 0;
    for (; s + 4 <= splits; s += 4) {
        s0 += partial[(size_t)(s + 0) * kNrcParams + p];
        s1 += partial[(size_t)(s + 1) * kNrcParams + p];
        s2 += partial[(size_t)(s + 2) * kNrcParams + p];
        s3 += partial[(size_t)(s + 3) * kNrcParams + p];
    }
    for (; s < splits; s++) s0 += partial[(size_t)s * kNrcParams + p];
    const float gr = ((s0 + s1) + (s2 + s3)) / (float)n;
    gradOut[p] = gr;
    const float mm = beta1 * m[p] + (1.0f - beta1) * gr;
    const float vv = beta2 * v[p] + (1.0f - beta2) * gr * gr;
    m[p] = mm;
    v[p] = vv;
    const float w = master[p] - lr * (mm * state->stepC1[step]) / (sqrtf(vv * state->stepC2[step]) + eps);
    master[p] = w;
    fwd[fwd_pos(l, in, out)] = (T)w;
    bwd[bwd_pos(l, in, out)] = (T)w;
}

// Writes every fragment element from the master weights (slots outside the network get zero).
template <typename T>
__global__ __launch_bounds__(256) void k_nrc_pack(const float* __restrict__ master, T* __restrict__ fwd, T* __restrict__ bwd) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e < kNrcParams) {
        const int l = e >> 12, rem = e & 4095;
        const int j = rem & 7, lane = (rem >> 3) & 63, ks = (rem >> 9) & 1, mt = rem >> 10;
        const int out = 16 * mt + (lane & 15);
        const int in = l == 0 ? slot_feature(ks, lane >> 4, j) : slot_neuron(ks, lane >> 4, j);
        fwd[e] = (T)master[w_index(l, in, out)];
    }
    if (e < kNrcBwdElems) {
        const int l = e >> 12, rem = e & 4095;
        const int j = rem & 7, lane = (rem >> 3) & 63, ks = (rem >> 9) & 1, it = rem >> 10;
        const int in = 16 * it + (lane & 15), out = slot_neuron(ks, lane >> 4, j);
        bwd[e] = (l < 5 || out < 16) ? (T)master[w_index(l, in, out)] : (T)0.0f;
    }
}

template <typename T>
hipError_t infer_t(const NrcBuffers& b, const float* raw, const int32_t* nDev, int nCap, float* out, const NrcQuery* meta,
                   float* sampleL, hipStream_t s) {
    // two blocks per CU are resident (114 VGPRs: 4 waves per SIMD): a grid of that size stages the
    // weights once per block and loops over the tiles, with no second round of blocks
    constexpr int kWaves = kInferBlock / 64;
    const int groups = (nCap + 16 * kInferTiles - 1) / (16 * kInferTiles);
    const int blocks = std::max(1, std::min((groups + kWaves - 1) / kWaves, 2 * 256));
    k_nrc_infer<T><<<blocks, kInferBlock, 0, s>>>(b.fwd, raw, nDev, nCap, out, meta, sampleL);
    return hipGetLastError();
}

template <typename T>
hipError_t train_t(const NrcBuffers& b, hipStream_t s, int steps) {
    T* H = reinterpret_cast<T*>(b.H);
    T* G = reinterpret_cast<T*>(b.G);
    const int tiles = b.batchCap / 16;
    const int blocks = std::min(kNrcTrainBlocks, (tiles + kTrainBlock / 64 - 1) / (kTrainBlock / 64));
    const int splits = b.batchCap / kNrcKSplit;  // 256-sample GEMM splits; 4 per block and partial
    k_nrc_plan<<<1, 64, 0, s>>>(b.state, steps, b.batchCap);
    for (int k = 0; k < steps; k++) {
        k_nrc_fwdbwd<T><<<blocks, kTrainBlock, 0, s>>>(b.fwd, b.bwd, b.trainRaw, b.trainTarget, b.state, k, b.batchCap, H, G,
                                                         b.batchCap, b.lossPartial);
        k_nrc_wgrad<T><<<dim3(6, splits / 4), 256, 0, s>>>(H, G, b.batchCap, b.state, k, b.batchCap, b.partial);
        k_nrc_adam<T><<<(kNrcParams + 255) / 256, 256, 0, s>>>(b.partial, splits / 4, b.state, k, b.batchCap, b.master, b.m, b.v,
                                                               b.grad, reinterpret_cast<T*>(b.fwd), reinterpret_cast<T*>(b.bwd));
    }
    k_nrc_finish<<<1, 64, 0, s>>>(b.state, steps, b.lossPartial, blocks);
    return hipGetLastError();
}

}  // namespace

hipError_t nrc_pack(const NrcBuffers& b, hipStream_t s) {
    const int blocks = (kNrcBwdElems + 255) / 256;
    if (b.precision == 1)
        k_nrc_pack<_Float16><<<blocks, 256, 0, s>>>(b.master, reinterpret_cast<_Float16*>(b.fwd), reinterpret_cast<_Float16*>(b.bwd));
    else
        k_nrc_pack<__bf16><<<blocks, 256, 0, s>>>(b.master, reinterpret_cast<__bf16*>(b.fwd), reinterpret_cast<__bf16*>(b.bwd));
    return hipGetLastError();
}

hipError_t nrc_infer(const NrcBuffers& b, const float* raw, const int32_t* nDev, int nCap, float* out, hipStream_t s) {
    if (nCap <= 0) return hipSuccess;
    return b.precision == 1 ? infer_t<_Float16>(b, raw, nDev, nCap, out, nullptr, nullptr, s)
                            : infer_t<__bf16>(b, raw, nDev, nCap, out, nullptr, nullptr, s);
}

hipError_t nrc_infer_queries(const NrcBuffers& b, const float* raw, const NrcQuery* meta, const int32_t* nDev, int nCap,
                             float* sampleL, hipStream_t s) {
    if (nCap <= 0) return hipSuccess;
    return b.precision == 1 ? infer_t<_Float16>(b, raw, nDev, nCap, nullptr, meta, sampleL, s)
                            : infer_t<__bf16>(b, raw, nDev, nCap, nullptr, meta, sampleL, s);
}

hipError_t nrc_train(const NrcBuffers& b, hipStream_t s, int steps) {
    if (steps < 1 || steps > kNrcMaxSteps) return hipErrorInvalidValue;
    return b.precision == 1 ? train_t<_Float16>(b, s, steps) : train_t<__bf16>(b, s, steps);
}

template <typename X>
static hipError_t nrc_alloc(X** p, size_t count) {
    return hipMalloc((void**)p, (count ? count : 1) * sizeof(X));
}

hipError_t nrc_reserve(NrcBuffers& b, int batchCap, int steps, int chunks) {
    const int cap = std::max(kNrcBatchAlign, (batchCap + kNrcBatchAlign - 1) / kNrcBatchAlign * kNrcBatchAlign);
    steps = std::max(1, steps);
    chunks = std::max(1, chunks);
    if (b.master && cap == b.batchCap && chunks == b.chunks) {
        b.steps = steps;
        return hipSuccess;
    }
    const int precision = b.precision;
    const float lr = b.lr, b1 = b.beta1, b2 = b.beta2, eps = b.eps;
    nrc_release(b);
    b.precision = precision;
    b.lr = lr; b.beta1 = b1; b.beta2 = b2; b.eps = eps;
    hipError_t e = hipSuccess;
#define NTRY(x) do { e = (x); if (e != hipSuccess) { nrc_release(b); return e; } } while (0)
    NTRY(nrc_alloc(&b.master, kNrcParams));
    NTRY(nrc_alloc(&b.m, kNrcParams));
    NTRY(nrc_alloc(&b.v, kNrcParams));
    NTRY(nrc_alloc(&b.grad, kNrcParams));
    NTRY(nrc_alloc(&b.fwd, kNrcParams));
    NTRY(nrc_alloc(&b.bwd, kNrcBwdElems));
    NTRY(nrc_alloc(&b.H, (size_t)6 * 64 * cap));
    NTRY(nrc_alloc(&b.G, (size_t)6 * 64 * cap));
    NTRY(nrc_alloc(&b.partial, (size_t)(cap / (4 * kNrcKSplit)) * kNrcParams));
    NTRY(nrc_alloc(&b.lossPartial, kNrcTrainBlocks));
    NTRY(nrc_alloc(&b.state, 1));
    NTRY(nrc_alloc(&b.trainRaw, (size_t)chunks * cap * kNrcRaw));
    NTRY(nrc_alloc(&b.trainTarget, (size_t)chunks * cap * 3));
#undef NTRY
    b.batchCap = cap;
    b.steps = steps;
    b.chunks = chunks;
    b.gen = next_buffer_generation();
    return hipSuccess;
}

void nrc_release(NrcBuffers& b) {
    void* ps[] = {b.master, b.m, b.v, b.grad, b.fwd, b.bwd, b.H, b.G, b.partial, b.lossPartial, b.state, b.trainRaw, b.trainTarget};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    const int precision = b.precision;
    const float lr = b.lr, b1 = b.beta1, b2 = b.beta2, eps = b.eps;
    b = NrcBuffers();
    b.precision = precision;
    b.lr = lr; b.beta1 = b1; b.beta2 = b2; b.eps = eps;
}

}  // namespace lts
