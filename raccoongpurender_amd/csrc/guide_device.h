// Path-guiding network: shared constants, the vMF-mixture head and its sampling.
//
// Network: 96 encoded inputs (12-band sin/cos of the position normalised to the scene box,
// normal, outgoing direction, zero padding) -> 3 ReLU layers of width 64 -> 64 outputs, of
// which 40 parameterise 8 von Mises-Fisher lobes (unnormalised mean direction, log
// concentration, mixture logit). Weights are kept in fp32 and mirrored in bfloat16 for the
// matrix cores (row-major [out][in] for forward passes, transposed for back-propagation).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rgpu {

constexpr int kGIn = 96, kGHid = 64, kGOut = 64, kGLobes = 8, kGBands = 12;
constexpr int kGoW0 = 0, kGoB0 = kGoW0 + kGHid * kGIn, kGoW1 = kGoB0 + kGHid, kGoB1 = kGoW1 + kGHid * kGHid,
              kGoW2 = kGoB1 + kGHid, kGoB2 = kGoW2 + kGHid * kGHid, kGoW3 = kGoB2 + kGHid,
              kGoB3 = kGoW3 + kGOut * kGHid, kGParams = kGoB3 + kGOut;
constexpr float kGLogKappaMin = -3.0f, kGLogKappaMax = 7.5f;
constexpr int kGLobeFloats = 5;                  // mu.xyz, kappa, log weight
constexpr int kGSlotFloats = kGLobes * kGLobeFloats;  // 40 floats of lobe parameters per slot

// One training sample: position (normalised to the scene box), normal, outgoing and
// incident directions, and the weight F/p of the direction (luminance of the incident
// radiance times BSDF cosine over the sampling pdf).
struct GuideSample {
    float p[3], n[3], wo[3], wi[3], weight, pad[3];
};
static_assert(sizeof(GuideSample) == 64, "guide sample is 64 bytes");

// Per-vertex record kept until the path ends (p, n, wo, wi, throughput before the vertex,
// radiance after the vertex's own emission).
struct GuideRecord {
    float p[3], n[3], wo[3], wi[3], T[3], L[3], pad[2];
};
static_assert(sizeof(GuideRecord) == 80, "guide record is 80 bytes");
constexpr int kGRecords = 4;  // guided vertices recorded per path

struct DevGuide {
    float* params = nullptr;          // fp32 master weights
    uint16_t* w_bf = nullptr;         // bf16 weights [out][in] at the parameter offsets
    uint16_t* wt_bf = nullptr;        // bf16 transposed weights [in][out]
    float* grad = nullptr;
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    float* lobes = nullptr;           // per slot kGSlotFloats
    GuideRecord* records = nullptr;   // per slot kGRecords
    uint32_t* rec_count = nullptr;    // per slot
    GuideSample* samples = nullptr;   // training buffer
    uint32_t* sample_count = nullptr; // appended samples (may exceed capacity; extra are dropped)
    uint32_t capacity = 0;
    float lo[3] = {0, 0, 0}, inv_ext[3] = {1, 1, 1};  // scene box normalisation
    float select_p = 0.5f;            // one-sample MIS selection probability of the guide
    uint32_t enabled = 0;             // guided sampling on
    uint32_t train = 0;               // record training samples
};

// ---------------------------------------------------------------- vMF mixture (device)
__device__ __forceinline__ float guide_log_vmf(float kappa, float c) {
    return (float)(log((double)kappa) - log(2.0 * 3.14159265358979323846) - log1p(-exp(-2.0 * (double)kappa))) +
           kappa * (c - 1.0f);
}

// q(w) of the lobes stored for one slot
__device__ __forceinline__ float guide_pdf(const float* lobes, float wx, float wy, float wz) {
    float a[kGLobes], mx = -3.0e38f;
    for (int k = 0; k < kGLobes; ++k) {
        const float* L = lobes + kGLobeFloats * k;
        float c = L[0] * wx + L[1] * wy + L[2] * wz;
        a[k] = L[4] + guide_log_vmf(L[3], c);
        mx = fmaxf(mx, a[k]);
    }
    float s = 0.0f;
    for (int k = 0; k < kGLobes; ++k) s += expf(a[k] - mx);
    return expf(mx) * s;
}

// Sample a direction: lobe by weight (u0), then Jakob's inversion of the vMF (u1, u2).
__device__ __forceinline__ void guide_sample(const float* lobes, float u0, float u1, float u2, float* w) {
    float c = 0.0f;
    int k = kGLobes - 1;
    for (int j = 0; j < kGLobes; ++j) {
        c += expf(lobes[kGLobeFloats * j + 4]);
        if (u0 < c) {
            k = j;
            break;
        }
    }
    const float* L = lobes + kGLobeFloats * k;
    const float kap = L[3];
    float ct = 1.0f + (float)(log((double)u1 + (1.0 - (double)u1) * exp(-2.0 * (double)kap)) / (double)kap);
    ct = fminf(fmaxf(ct, -1.0f), 1.0f);
    float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    float ph = 2.0f * 3.14159265f * u2;
    float cp = (float)cos((double)ph), sp = (float)sin((double)ph);
    const float z0 = L[0], z1 = L[1], z2 = L[2];
    const float sign = copysignf(1.0f, z2);
    const float a = -1.0f / (sign + z2);
    const float b = z0 * z1 * a;
    const float x0 = 1.0f + sign * z0 * z0 * a, x1 = sign * b, x2 = -sign * z0;
    const float y0 = b, y1 = sign + z1 * z1 * a, y2 = -z1;
    w[0] = x0 * st * cp + y0 * st * sp + z0 * ct;
    w[1] = x1 * st * cp + y1 * st * sp + z1 * ct;
    w[2] = x2 * st * cp + y2 * st * sp + z2 * ct;
}

}  // namespace rgpu
