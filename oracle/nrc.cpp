// ORACLE — test infrastructure only. CPU restatement of the learned radiance cache's network:
// input encoding, forward pass, relative-L2 loss and its gradient, in fp64 with an optional
// emulation of the GPU's bf16 / fp16 rounding points. The reference has no such component;
// the definition restated here is the product's (64-dim encoding of a 16-float raw record,
// 64 -> 64 x5 ReLU -> 3, no biases, Adam), checked by tests against finite differences and
// closed forms.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle.h"

namespace {

constexpr int kIn = 64, kWidth = 64, kHidden = 5, kOutPad = 16, kOut = 3;
constexpr int kParams = kHidden * kIn * kWidth + kWidth * kOutPad;

// Round-to-nearest-even of an fp32 value to bf16 / fp16, returned as fp32.
float to_bf16(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return x;
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}

float to_fp16(float x) {
    if (!std::isfinite(x)) return x;
    float a = std::fabs(x);
    if (a >= 65520.0f) return std::copysign(INFINITY, x);  // beyond the last halfway point below 2^16
    int e;
    std::frexp(a, &e);                                     // a = m 2^e, m in [0.5, 1)
    float q = std::ldexp(1.0f, std::max(e - 11, -24));    // spacing of fp16 values around a
    return std::copysign(std::nearbyint(a / q) * q, x);    // ties to even (default rounding mode)
}

double round_to(double v, int mode) {
    if (mode == 0) return v;
    float f = (float)v;
    return mode == 1 ? to_bf16(f) : to_fp16(f);
}

int w_index(int l, int in, int out) { return l < kHidden ? l * kIn * kWidth + in * kWidth + out : kHidden * kIn * kWidth + in * kOutPad + out; }

float one_blob(float s, int bin) {
    // triweight kernel (1 - d^2/9)^3, d = distance to the bin centre (i + 0.5) / 4 in bin widths
    const float d = (s - ((float)bin + 0.5f) * 0.25f) * 4.0f;
    const float u = std::fmax(1.0f - d * d * (1.0f / 9.0f), 0.0f);
    return u * u * u;
}

}  // namespace

extern "C" {

void oracle_nrc_encode(const float raw[16], float out[64]) {
    for (int idx = 0; idx < 64; idx++) {
        float v;
        if (idx < 36) {
            // triangle wave of frequency 2^(f-1) on the normalised coordinate
            int d = idx / 12, f = idx % 12;
            float x = std::ldexp(raw[d], f - 1);
            float fr = x - std::floor(x);
            v = std::fabs(2.0f * fr - 1.0f);
        } else if (idx < 52) {
            v = one_blob(raw[3 + (idx - 36) / 4], idx % 4);
        } else if (idx < 56) {
            v = one_blob(raw[7], idx - 52);
        } else if (idx < 62) {
            v = raw[8 + idx - 56];
        } else {
            v = 1.0f;
        }
        out[idx] = v;
    }
}

// Forward pass of n samples; y receives n*3 outputs. rounding: 0 none, 1 bf16, 2 fp16.
void oracle_nrc_forward(const float* params, const float* raw, int32_t n, int32_t rounding, double* y) {
    std::vector<double> w(kParams);
    for (int p = 0; p < kParams; p++) w[p] = round_to(params[p], rounding);
    double h[kWidth], z[kWidth];
    float enc[kIn];
    for (int i = 0; i < n; i++) {
        oracle_nrc_encode(raw + 16 * i, enc);
        for (int k = 0; k < kIn; k++) h[k] = round_to(enc[k], rounding);
        for (int l = 0; l < kHidden; l++) {
            for (int o = 0; o < kWidth; o++) {
                double s = 0;
                for (int k = 0; k < kWidth; k++) s += w[w_index(l, k, o)] * h[k];
                z[o] = s;
            }
            for (int o = 0; o < kWidth; o++) h[o] = round_to(z[o] > 0 ? z[o] : 0.0, rounding);
        }
        for (int c = 0; c < kOut; c++) {
            double s = 0;
            for (int k = 0; k < kWidth; k++) s += w[w_index(kHidden, k, c)] * h[k];
            y[3 * i + c] = s;
        }
    }
}

// Loss (1/n) sum_i sum_c (y_ic - t_ic)^2 / (d_ic + 0.01) with d = y^2 held constant (stop-gradient)
// or d = denom_ref when given; grad (kParams entries, padding columns of the output layer stay 0)
// receives dLoss/dW. With rounding, layer inputs and back-propagated gradients are rounded where
// the GPU kernels round them, and the output-layer gradient is formed from y rounded to fp32.
double oracle_nrc_grad(const float* params, const float* raw, const float* target, int32_t n, int32_t rounding,
                       const double* denom_ref, double* grad, double* y_out) {
    std::vector<double> w(kParams);
    for (int p = 0; p < kParams; p++) w[p] = round_to(params[p], rounding);
    if (grad) std::memset(grad, 0, sizeof(double) * kParams);
    std::vector<double> acts((kHidden + 1) * kWidth);
    double z[kWidth], g[kWidth], gn[kWidth], dy[kOut];
    float enc[kIn];
    double loss = 0;
    for (int i = 0; i < n; i++) {
        oracle_nrc_encode(raw + 16 * i, enc);
        double* h0 = &acts[0];
        for (int k = 0; k < kIn; k++) h0[k] = round_to(enc[k], rounding);
        for (int l = 0; l < kHidden; l++) {
            const double* hi = &acts[l * kWidth];
            double* ho = &acts[(l + 1) * kWidth];
            for (int o = 0; o < kWidth; o++) {
                double s = 0;
                for (int k = 0; k < kWidth; k++) s += w[w_index(l, k, o)] * hi[k];
                z[o] = s;
            }
            for (int o = 0; o < kWidth; o++) ho[o] = round_to(z[o] > 0 ? z[o] : 0.0, rounding);
        }
        const double* h5 = &acts[kHidden * kWidth];
        for (int c = 0; c < kOut; c++) {
            double s = 0;
            for (int k = 0; k < kWidth; k++) s += w[w_index(kHidden, k, c)] * h5[k];
            double yv = rounding ? (double)(float)s : s;
            if (y_out) y_out[3 * i + c] = yv;
            double d = denom_ref ? denom_ref[3 * i + c] : yv * yv;
            double diff = yv - target[3 * i + c];
            loss += diff * diff / (d + 0.01);
            double gy = 2.0 * diff / (d + 0.01);
            dy[c] = rounding ? round_to((double)(float)gy, rounding) : gy;
        }
        if (!grad) continue;
        // output layer
        for (int k = 0; k < kWidth; k++)
            for (int c = 0; c < kOut; c++) grad[w_index(kHidden, k, c)] += h5[k] * dy[c];
        for (int k = 0; k < kWidth; k++) {
            double s = 0;
            for (int c = 0; c < kOut; c++) s += w[w_index(kHidden, k, c)] * dy[c];
            g[k] = round_to(h5[k] > 0 ? s : 0.0, rounding);
        }
        // hidden layers, last to first: g holds dLoss/dz of layer l's output
        for (int l = kHidden - 1; l >= 0; l--) {
            const double* hi = &acts[l * kWidth];
            for (int k = 0; k < kWidth; k++)
                for (int o = 0; o < kWidth; o++) grad[w_index(l, k, o)] += hi[k] * g[o];
            if (l == 0) break;
            for (int k = 0; k < kWidth; k++) {
                double s = 0;
                for (int o = 0; o < kWidth; o++) s += w[w_index(l, k, o)] * g[o];
                gn[k] = round_to(hi[k] > 0 ? s : 0.0, rounding);
            }
            std::memcpy(g, gn, sizeof(g));
        }
    }
    if (grad && n > 0)
        for (int p = 0; p < kParams; p++) grad[p] /= n;
    return n > 0 ? loss / n : 0.0;
}

int32_t oracle_nrc_param_count(void) { return kParams; }

}  // extern "C"
