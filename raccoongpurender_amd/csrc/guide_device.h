// Path-guiding network: shared constants, the vMF-mixture head and its sampling.
//
// Network: 96 encoded inputs (12-band sin/cos of the position normalised to the scene box,
// normal, outgoing direction, zero padding) -> 3 ReLU layers of width 64 -> 64 outputs, of
// which 40 parameterise 8 von Mises-Fisher lobes (unnormalised mean direction, log
// concentration, mixture logit) and output 40 is the logit of the learned selection
// probability (used when the guide is configured to learn it). Weights are kept in fp32 and
// mirrored in bfloat16 (row-major [out][in]) for the matrix cores.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rgpu {

constexpr int kGIn = 96, kGHid = 64, kGOut = 64, kGLobes = 8, kGBands = 12;
constexpr int kGoW0 = 0, kGoB0 = kGoW0 + kGHid * kGIn, kGoW1 = kGoB0 + kGHid, kGoB1 = kGoW1 + kGHid * kGHid,
              kGoW2 = kGoB1 + kGHid, kGoB2 = kGoW2 + kGHid * kGHid, kGoW3 = kGoB2 + kGHid,
              kGoB3 = kGoW3 + kGOut * kGHid, kGParams = kGoB3 + kGOut;
constexpr int kGTrainGroups = 128;  // partial gradient sums of one training step (tile splits)
constexpr int kGActBytesPerRow = 544 * 2;  // layer inputs and output deltas of one training row
constexpr float kGLogKappaMin = -3.0f, kGLogKappaMax = 7.5f;
// Stored per lobe: mu.xyz (unit), kappa, mixture weight, log(weight * C(kappa)) with
// C(kappa) = kappa / (2 pi (1 - exp(-2 kappa))) the vMF normalisation, so that evaluating
// the mixture needs one exp per lobe.
constexpr int kGLobeFloats = 6;
constexpr int kGSlotFloats = kGLobes * kGLobeFloats;  // 48 floats of lobe parameters per slot
// Learned selection probability of the guide in the one-sample mixture (Mueller et al. 2019,
// "Neural Importance Sampling", learned selection probabilities): c = sigmoid(output 40),
// clamped to [kGSelMin, kGSelMax] so that neither strategy is ever switched off.
constexpr int kGSelOut = 5 * kGLobes;
constexpr float kGSelMin = 0.05f, kGSelMax = 0.95f;
__device__ __forceinline__ float guide_selection(float z) {
    const float c = 1.0f / (1.0f + __expf(-z));
    return fminf(fmaxf(c, kGSelMin), kGSelMax);
}

// One training sample: position (normalised to the scene box), normal, outgoing and
// incident directions, the weight F/p of the direction (luminance of the incident radiance
// times BSDF cosine over the sampling pdf), the scene class code of a batch job
// (0 for a single render, 1..kGClasses for the job's scene; an exact small float) and the
// density of the other strategies (the BSDF/envmap mixture) at wi as the vertex had it, which
// the learned selection probability is trained on (0: not known, no selection gradient).
struct GuideSample {
    float p[3], n[3], wo[3], wi[3], weight, scene_class, other_pdf, pad;
};
static_assert(sizeof(GuideSample) == 64, "guide sample is 64 bytes");

// Per-vertex record kept until the path ends: the sample's geometry (p, n, wo, wi), and in
// separate arrays the throughput before the vertex and the radiance after its own emission.
// Record j of slot s sits at j * nslots + s, so that a wave touches consecutive records.
struct GuideVertex {
    float p[3], n[3], wo[3], wi[3];
};
static_assert(sizeof(GuideVertex) == 48, "guide vertex is 48 bytes");
constexpr int kGRecords = 4;  // guided vertices recorded per path
constexpr int kGClasses = 8;  // distinct scenes one network tells apart (one-hot inputs 78..85)

// Where a vertex's guide inputs come from: the box that normalises its position and the
// class code of its scene (0 = none: a single render; c = 1..kGClasses: one-hot input 77 + c).
struct GuideFrame {
    float lo[3], inv_ext[3];
    uint32_t cls;
};

struct DevGuide {
    float* params = nullptr;          // fp32 master weights
    uint16_t* w_bf = nullptr;         // bf16 weights [out][in] at the parameter offsets
    uint16_t* wt_bf = nullptr;        // bf16 transposed weights [in][out]
    uint16_t* act = nullptr;          // training buffer of layer inputs and deltas (bf16)
    uint32_t act_rows = 0;            // rows it holds (the minibatch)
    float* grad = nullptr;            // summed gradient (test hook only)
    float* partials = nullptr;        // kGTrainGroups per-workgroup gradient sums
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    // per guided vertex of the bounce, in queue order (the lambertian queue, then the Ward
    // queue), as twelve planes of nslots float4 (plane k: floats 4k..4k+3 of every vertex), so
    // that inference writes and the shading kernels read them contiguously (guide_load_lobes)
    float* lobes = nullptr;           // kGSlotFloats per vertex
    float* sel = nullptr;             // learned selection probability (learn_select)
    GuideVertex* rec_geo = nullptr;   // kGRecords * nslots
    float4* rec_T = nullptr;          // kGRecords * nslots (xyz; w: the other strategies' pdf at wi)
    float4* rec_L = nullptr;          // kGRecords * nslots (xyz used)
    uint32_t* rec_count = nullptr;    // per slot
    float4* emit_w = nullptr;         // per slot: weights of the samples a finished path emits
    uint32_t nslots = 0;
    uint32_t infer_blocks = 0;        // persistent inference grid
    GuideSample* samples = nullptr;   // training buffer
    uint32_t* sample_count = nullptr; // appended samples (may exceed capacity; extra are dropped)
    uint32_t capacity = 0;
    float lo[3] = {0, 0, 0}, inv_ext[3] = {1, 1, 1};  // scene box normalisation
    float select_p = 0.5f;            // one-sample MIS selection probability of the guide
    uint32_t learn_select = 0;        // per-vertex selection probability from output kGSelOut
    uint32_t max_depth = 0;           // guided vertices per path (0 = all)
    float weight_cap = 0.0f;          // training weights are clamped to it (0 = not clamped)
    float log_kappa_max = kGLogKappaMax;  // upper clamp of a lobe's log concentration
    uint32_t enabled = 0;             // guided sampling on
    uint32_t train = 0;               // record training samples
};

// ---------------------------------------------------------------- vMF mixture (device)
// log C(kappa) for kappa in [e^-3, e^7.5]
__device__ __forceinline__ float guide_log_norm(float kappa) {
    return logf(kappa) - 1.83787706641f - log1pf(-expf(-2.0f * kappa));
}

// q(w) of the lobes stored for one slot. kappa (c - 1) is formed as -kappa |w - mu|^2 / 2,
// which keeps its precision for sharp lobes where c is close to 1.
__device__ __forceinline__ float guide_pdf(const float* lobes, float wx, float wy, float wz) {
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < kGLobes; ++k) {
        const float* L = lobes + kGLobeFloats * k;
        const float dx = wx - L[0], dy = wy - L[1], dz = wz - L[2];
        s += expf(L[5] - 0.5f * L[3] * (dx * dx + dy * dy + dz * dz));
    }
    return s;
}

// The 48 lobe floats of guided vertex `at` from the twelve float4 planes of DevGuide::lobes.
__device__ __forceinline__ void guide_load_lobes(const float* lobes, uint32_t nslots, size_t at, float* L) {
    const float4* q = reinterpret_cast<const float4*>(lobes);
#pragma unroll
    for (int k = 0; k < kGSlotFloats / 4; ++k) {
        const float4 v = q[(size_t)k * nslots + at];
        L[4 * k] = v.x;
        L[4 * k + 1] = v.y;
        L[4 * k + 2] = v.z;
        L[4 * k + 3] = v.w;
    }
}

// Sample a direction: lobe by weight (u0), then Jakob's inversion of the vMF (u1, u2).
// The selected lobe is gathered with compile-time indices so that a lobe table held in
// registers is never indexed dynamically.
__device__ __forceinline__ void guide_sample(const float* lobes, float u0, float u1, float u2, float* w) {
    constexpr int kLast = kGLobeFloats * (kGLobes - 1);
    float z0 = lobes[kLast], z1 = lobes[kLast + 1], z2 = lobes[kLast + 2], kap = lobes[kLast + 3];
    float c = 0.0f;
    bool found = false;
#pragma unroll
    for (int j = 0; j < kGLobes - 1; ++j) {
        c += lobes[kGLobeFloats * j + 4];
        if (!found && u0 < c) {
            z0 = lobes[kGLobeFloats * j];
            z1 = lobes[kGLobeFloats * j + 1];
            z2 = lobes[kGLobeFloats * j + 2];
            kap = lobes[kGLobeFloats * j + 3];
            found = true;
        }
    }
    float ct = 1.0f + logf(u1 + (1.0f - u1) * expf(-2.0f * kap)) / kap;
    ct = fminf(fmaxf(ct, -1.0f), 1.0f);
    const float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    const float cp = cospif(2.0f * u2), sp = sinpif(2.0f * u2);
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
