// Path-guiding network kernels on gfx950 matrix cores (bf16 operands, fp32 accumulation).
//
// Rows (path vertices or training samples) go through the network in tiles of 32, one
// workgroup of two wave64s per tile, each wave owning 16 rows. Activations live in LDS as
// bf16; the bf16 weights (37 KB, row-major [out][in], plus a transposed copy for
// back-propagation) are read from global memory, where they stay cache-resident, so that a
// tile needs only 29-38 KB of LDS and four or five workgroups share a CU. Products run
// 16x16x32 at a time on v_mfma_f32_16x16x32_bf16:
//   A operand: lane l holds 8 consecutive k of row (l mod 16),
//   B operand: lane l holds the same 8 k of column (l mod 16),
//   result:    lane l holds rows 4*(l/16)..+3 of column (l mod 16).
// Training is split in two: k_guide_fwdbwd runs forward, loss gradient and the input deltas
// of every tile and writes the layer inputs and output deltas to a bf16 buffer (column-major
// per tile); k_guide_dw then forms the weight gradients dW = dZ^T In with the rows as the
// reduction dimension, one partial sum per split of the tiles, which k_adam adds up.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "guide_device.h"
#include "rgpu_wavefront.h"

namespace rgpu {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kTile = 32;                  // rows per workgroup
constexpr int kPX = kGIn + 8;              // LDS pitch of the input tile (u16), 16-byte aligned rows
constexpr int kPH = kGHid + 8;             // LDS pitch of hidden tiles
constexpr int kPF = kGOut + 4;             // LDS pitch of the fp32 output tile

// float -> bf16, round to nearest even (v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
// two floats -> packed bf16 pair, a in the low half
__device__ __forceinline__ uint32_t f2bf2(float a, float b) {
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    const bf16x2 v = {(__bf16)a, (__bf16)b};
    return __builtin_bit_cast(uint32_t, v);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ bf16x8 ld8(const uint16_t* p) {
    uint4 v = *reinterpret_cast<const uint4*>(p);
    return __builtin_bit_cast(bf16x8, v);
}
// 8 bf16 values at stride `step` (column reads of a row-major LDS tile)
__device__ __forceinline__ bf16x8 ld8s(const uint16_t* p, int step) {
    uint16_t t[8];
    for (int j = 0; j < 8; ++j) t[j] = p[j * step];
    uint4 v = make_uint4(t[0] | ((uint32_t)t[1] << 16), t[2] | ((uint32_t)t[3] << 16), t[4] | ((uint32_t)t[5] << 16),
                         t[6] | ((uint32_t)t[7] << 16));
    return __builtin_bit_cast(bf16x8, v);
}
__device__ __forceinline__ f32x4 mma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Forward layer for this wave's 16 rows: out = act(X W^T + b), X: LDS [rows][K] pitch px.
// The product is formed transposed (A: weights, B: rows), so that each lane ends with four
// consecutive outputs of one row and writes them with one 8-byte (bf16) or 16-byte (fp32)
// store.
// W: rows of K weights at pitch wp (global memory or LDS). NT: 16-column output tiles formed
// (the last layer's outputs 48..63 feed nothing; inference skips them).
template <int K, int NT = 4>
__device__ void fwd_layer(const uint16_t* X, int px, const uint16_t* W, int wp, const float* bias, uint16_t* H,
                          float* F, int r0, int lane, bool relu) {
    const int row = r0 + (lane & 15);
    bf16x8 x[K / 32];  // this wave's rows, read once for all output tiles
#pragma unroll
    for (int kb = 0; kb < K / 32; ++kb) x[kb] = ld8(X + row * px + kb * 32 + (lane >> 4) * 8);
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kb = 0; kb < K / 32; ++kb) {
            const bf16x8 w = ld8(W + (nt * 16 + (lane & 15)) * wp + kb * 32 + (lane >> 4) * 8);
            acc = mma(w, x[kb], acc);
        }
        const int col = nt * 16 + (lane >> 4) * 4;
        const float4 bb = *reinterpret_cast<const float4*>(bias + col);
        float v0 = acc[0] + bb.x, v1 = acc[1] + bb.y, v2 = acc[2] + bb.z, v3 = acc[3] + bb.w;
        if (F) {
            *reinterpret_cast<float4*>(F + row * kPF + col) = make_float4(v0, v1, v2, v3);
        } else {
            if (relu) {
                v0 = fmaxf(v0, 0.0f);
                v1 = fmaxf(v1, 0.0f);
                v2 = fmaxf(v2, 0.0f);
                v3 = fmaxf(v3, 0.0f);
            }
            *reinterpret_cast<uint2*>(H + row * kPH + col) = make_uint2(f2bf2(v0, v1), f2bf2(v2, v3));
        }
    }
}

// sin(pi x) and cos(pi x) for x in [0, 2]: x = r + q/2 with q = rint(2x) (exact, |r| <= 1/4),
// Taylor series of sin(pi r) to r^9 and cos(pi r) to r^10 (truncation below 3e-9), then the
// quadrant q mod 4.
__device__ __forceinline__ void sincospi_02(float x, float* s, float* c) {
    const float q = rintf(2.0f * x);
    const float r = x - 0.5f * q;
    const float z = r * r;
    float ps = __builtin_fmaf(z, 0.0821458866f, -0.599264529f);  // pi^9/9!, -pi^7/7!
    ps = __builtin_fmaf(z, ps, 2.55016404f);                       // pi^5/5!
    ps = __builtin_fmaf(z, ps, -5.16771278f);                      // -pi^3/3!
    ps = __builtin_fmaf(z, ps, 3.14159265f);                       // pi
    const float sr = r * ps;
    float pc = __builtin_fmaf(z, -0.0258068914f, 0.235330630f);  // -pi^10/10!, pi^8/8!
    pc = __builtin_fmaf(z, pc, -1.33526277f);                      // -pi^6/6!
    pc = __builtin_fmaf(z, pc, 4.05871213f);                       // pi^4/4!
    pc = __builtin_fmaf(z, pc, -4.93480220f);                      // -pi^2/2!
    const float cr = __builtin_fmaf(z, pc, 1.0f);
    const int qi = (int)q;
    const float a = (qi & 1) ? cr : sr;
    const float b = (qi & 1) ? sr : cr;
    *s = (qi & 2) ? -a : a;
    *c = ((qi + 1) & 2) ? -b : b;
}

// Encoded input features of one row (written as bf16 into X): for band k and axis a,
// features 6k + 2a and 6k + 2a + 1 are sin and cos of 2^k pi p_a. 2^k p_a is exact in fp32 and
// so is its reduction modulo 2, which leaves sincospi_02 an argument in [0, 2). Then the normal,
// the outgoing direction and the scene class one-hot (class code c = 1..kGClasses sets
// feature 77 + c). Features 2j and 2j + 1 form pair j; each of the row's four lanes writes nine
// sin/cos pairs and three plain pairs, one 4-byte store per pair.
__device__ void encode_row(uint16_t* X, const float* p01, const float* n, const float* wo, uint32_t cls,
                           int part) {
    uint32_t* X2 = reinterpret_cast<uint32_t*>(X);
    // The 36 sin/cos pairs form 12 chains (axis a, bands k0..k0+2, k0 = 0, 3, 6, 9); chain
    // c = part + 4 m (m = 0..2) belongs to this lane, with axis c mod 3 = (part + m) mod 3 and
    // k0 = 3 (c / 3). The first band of a chain is evaluated, the next two follow by the double
    // angle formulas sin 2t = 2 sin t cos t, cos 2t = (cos t - sin t)(cos t + sin t), which add
    // a few float ulps, far below bf16 resolution.
    const int rot = part == 3 ? 0 : part;
    float v0 = rot == 0 ? p01[0] : rot == 1 ? p01[1] : p01[2];
    float v1 = rot == 0 ? p01[1] : rot == 1 ? p01[2] : p01[0];
    float v2 = rot == 0 ? p01[2] : rot == 1 ? p01[0] : p01[1];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        const int c = part + 4 * m;
        const int k0 = c - c % 3;  // 3 (c / 3)
        const int a = (part + m) % 3;
        float x = ldexpf(v0, k0);
        x = x - 2.0f * floorf(0.5f * x);
        float sv, cv;
        sincospi_02(x, &sv, &cv);
        int j = k0 * 3 + a;  // pair of band k0, axis a
        X2[j] = f2bf2(sv, cv);
#pragma unroll
        for (int d = 1; d < 3; ++d) {
            const float s2 = (sv + sv) * cv;
            cv = (cv - sv) * (cv + sv);
            sv = s2;
            j += 3;
            X2[j] = f2bf2(sv, cv);
        }
        const float t = v0;  // next chain: next axis
        v0 = v1;
        v1 = v2;
        v2 = t;
    }
    // pairs 36..47: features 72..95. Pair 36 + part: (n0, n1), (n2, wo0), (wo1, wo2) or the
    // class features 78, 79; pairs 40 + part and 44 + part: class features 80 + 2 part ...
    float lo, hi;
    if (part == 0) {
        lo = n[0];
        hi = n[1];
    } else if (part == 1) {
        lo = n[2];
        hi = wo[0];
    } else if (part == 2) {
        lo = wo[1];
        hi = wo[2];
    } else {
        lo = cls == 1u ? 1.0f : 0.0f;
        hi = cls == 2u ? 1.0f : 0.0f;
    }
    X2[36 + part] = f2bf2(lo, hi);
#pragma unroll
    for (int i = 10; i < 12; ++i) {
        const int j = part + 4 * i;
        const uint32_t c = (uint32_t)(2 * j - 77);  // class code of the pair's first feature
        X2[j] = f2bf2(c == cls ? 1.0f : 0.0f, c + 1u == cls ? 1.0f : 0.0f);
    }
}

// Class code of a stored sample; anything but 1..kGClasses reads as none.
__device__ __forceinline__ uint32_t sample_class(float c) {
    return (c >= 1.0f && c <= (float)kGClasses) ? (uint32_t)c : 0u;
}

// Largest component below which a lobe's raw mean has no direction: every hidden unit of a row
// can be zero (dead ReLUs), leaving the mean outputs at their biases, which may be zero. Such a
// lobe (and one with a non-finite mean) takes the mean +z and passes no gradient to its mean
// outputs. The mean is scaled by its largest component before it is squared, so that large
// outputs do not overflow.
constexpr float kMinMean = 1e-15f;

// Unit mean of raw outputs o[0..2] into mu; returns the length of o (0: no direction). The
// head's reciprocals, square roots, exponentials and logarithms use the hardware's
// single-instruction forms (v_rcp_f32, v_rsq_f32, v_exp_f32, v_log_f32: about one ulp), since
// the lobes only need to be consistent between sampling and evaluation, which both read the
// stored records.
__device__ __forceinline__ float lobe_mean(const float* o, float* mu) {
    const float mx = fmaxf(fmaxf(fabsf(o[0]), fabsf(o[1])), fabsf(o[2]));
    if (!(mx > kMinMean) || !(mx <= 3.0e38f)) {
        mu[0] = 0.0f;
        mu[1] = 0.0f;
        mu[2] = 1.0f;
        return 0.0f;
    }
    const float inv = __builtin_amdgcn_rcpf(mx);
    const float q0 = o[0] * inv, q1 = o[1] * inv, q2 = o[2] * inv;
    const float len2 = q0 * q0 + q1 * q1 + q2 * q2;  // in [1, 3]
    const float il = __builtin_amdgcn_rsqf(len2);
    mu[0] = q0 * il;
    mu[1] = q1 * il;
    mu[2] = q2 * il;
    return (len2 * il) * mx;
}

// Lobe k of one row from its raw outputs o[0..4], the row's log-sum-exp shift (max + log of the
// sum) and the lobe's softmax weight w. log C(kappa) = log kappa - log 2 pi - log(1 - e^-2kappa),
// with log kappa the clamped raw output itself.
__device__ __forceinline__ void lobe_params(const float* o, float shift, float w, float lkmax, float* L) {
    const float lk = fminf(fmaxf(o[3], kGLogKappaMin), lkmax);
    const float kap = __expf(lk);
    lobe_mean(o, L);
    L[3] = kap;
    L[4] = w;
    L[5] = (o[4] - shift) + ((lk - 1.83787706641f) - __logf(1.0f - __expf(-2.0f * kap)));
}

// The mixture head runs on four lanes per row (lanes 4r..4r+3 of a workgroup hold row r),
// two lobes per lane; softmax and log-sum-exp combine over the four lanes.
__device__ __forceinline__ float quad_max(float v) {
    v = fmaxf(v, __shfl_xor(v, 1));
    return fmaxf(v, __shfl_xor(v, 2));
}
__device__ __forceinline__ float quad_sum(float v) {
    v += __shfl_xor(v, 1);
    return v + __shfl_xor(v, 2);
}

// Raw head outputs of one row -> this lane's two lobes (mu, kappa, weight, log(weight C(kappa))).
__device__ void head_lobes_part(const float* out, int part, float lkmax, float* L) {
    const float* o0 = out + 5 * (2 * part);
    const float* o1 = o0 + 5;
    const float mx = quad_max(fmaxf(o0[4], o1[4]));
    const float e0 = __expf(o0[4] - mx), e1 = __expf(o1[4] - mx);
    const float sum = quad_sum(e0 + e1);  // >= 1: the largest term is e^0
    const float shift = mx + __logf(sum), is = __builtin_amdgcn_rcpf(sum);
    lobe_params(o0, shift, e0 * is, lkmax, L);
    lobe_params(o1, shift, e1 * is, lkmax, L + kGLobeFloats);
}

// d(loss) / d(raw outputs) for this lane's two lobes: d[j] is the derivative by output
// 10 part + j; *dsel (every lane) the derivative by the selection logit, output kGSelOut (the 23
// other outputs have none). The loss is -weight * log q(wi), and with a learned selection
// probability c = guide_selection(out[kGSelOut]) and a known density p_o > 0 of the other
// strategies at wi also -weight * log(c q(wi) + (1 - c) p_o), the log density of the whole
// one-sample mixture (Mueller et al. 2019): its derivative scales the lobes' gradient by
// 1 + c q / p_mix and gives the logit -weight (q - p_o) c (1 - c) / p_mix inside the clamp.
__device__ void head_grad_part(const float* out, int part, const float* wi, float weight, float lkmax, float other_pdf,
                               bool learn_select, float* d, float* dsel) {
    float L[2 * kGLobeFloats], a[2], cm1[2];
    head_lobes_part(out, part, lkmax, L);
    for (int q = 0; q < 2; ++q) {
        const float* Lk = L + kGLobeFloats * q;
        const float dx = wi[0] - Lk[0], dy = wi[1] - Lk[1], dz = wi[2] - Lk[2];
        cm1[q] = -0.5f * (dx * dx + dy * dy + dz * dz);  // mu . wi - 1
        a[q] = Lk[5] + Lk[3] * cm1[q];
    }
    const float mx = quad_max(fmaxf(a[0], a[1]));
    const float lse = mx + logf(quad_sum(expf(a[0] - mx) + expf(a[1] - mx)));
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const float* Lk = L + kGLobeFloats * q;
        const float* o = out + 5 * (2 * part + q);
        const float r = expf(a[q] - lse);  // responsibility of the lobe for wi
        d[5 * q + 4] = weight * (Lk[4] - r);
        const float kap = Lk[3];
        const float e = expf(-2.0f * kap);
        const float dlogf = 1.0f / kap - 2.0f * e / (1.0f - e) + cm1[q];
        d[5 * q + 3] = (o[3] > kGLogKappaMin && o[3] < lkmax) ? -weight * r * dlogf * kap : 0.0f;
        float mu[3];
        const float m = lobe_mean(o, mu);
        const float c = cm1[q] + 1.0f;
        for (int j = 0; j < 3; ++j) d[5 * q + j] = m > 0.0f ? -weight * r * kap * (wi[j] - c * Lk[j]) / m : 0.0f;
    }
    *dsel = 0.0f;
    if (learn_select && other_pdf > 0.0f) {
        const float z = out[kGSelOut];
        const float sg = 1.0f / (1.0f + __expf(-z));
        const float c = fminf(fmaxf(sg, kGSelMin), kGSelMax);
        const float qv = __expf(lse);
        const float pm = c * qv + (1.0f - c) * other_pdf;
        const float scale = 1.0f + c * qv / pm;
        for (int j = 0; j < 10; ++j) d[j] *= scale;
        *dsel = (sg > kGSelMin && sg < kGSelMax) ? -weight * (qv - other_pdf) * c * (1.0f - c) / pm : 0.0f;
    }
}

// Shared forward pass of a tile: X filled by the caller; leaves H1..H3 (bf16) and F (fp32 outputs).
struct alignas(16) TileLds {
    uint16_t X[kTile * kPX];
    uint16_t H1[kTile * kPH];
    uint16_t H2[kTile * kPH];
    uint16_t H3[kTile * kPH];
    float F[kTile * kPF];
};

__device__ void tile_forward(TileLds& t, const DevGuide& g, int r0, int lane) {
    fwd_layer<kGIn>(t.X, kPX, g.w_bf + kGoW0, kGIn, g.params + kGoB0, t.H1, nullptr, r0, lane, true);
    fwd_layer<kGHid>(t.H1, kPH, g.w_bf + kGoW1, kGHid, g.params + kGoB1, t.H2, nullptr, r0, lane, true);
    fwd_layer<kGHid>(t.H2, kPH, g.w_bf + kGoW2, kGHid, g.params + kGoB2, t.H3, nullptr, r0, lane, true);
    fwd_layer<kGHid>(t.H3, kPH, g.w_bf + kGoW3, kGHid, g.params + kGoB3, nullptr, t.F, r0, lane, false);
}

// Inference runs one workgroup of kInferWaves waves per CU. The workgroup copies the bf16
// weights and the biases into LDS once (rows padded to the activation pitches), so that every
// layer reads both MFMA operands from LDS; each wave then works through 16-row blocks of the
// queues on its own: two private activation buffers (A: input, H2 and the fp32 outputs; B: H1
// and H3) and no workgroup barrier inside the loop.
constexpr int kInferWaves = 16;
constexpr int kInferWaveA = (16 * kPX * 2 > 16 * kPF * 4 ? 16 * kPX * 2 : 16 * kPF * 4) / 2;
struct alignas(16) InferWaveLds {
    uint16_t A[kInferWaveA];
    uint16_t B[16 * kPH];
};
struct alignas(16) InferLds {
    uint16_t W0[kGHid * kPX];
    uint16_t W1[kGHid * kPH];
    uint16_t W2[kGHid * kPH];
    uint16_t W3[kGOut * kPH];
    float bias[4][kGHid];
    InferWaveLds wave[kInferWaves];
};
static_assert(sizeof(InferLds) <= 160 * 1024, "inference LDS exceeds a CU's 160 KB");

// Orders this wave's LDS accesses across lanes: a wave's LDS operations execute in program
// order, so the fences only stop the compiler from moving accesses across this point.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Copies `rows` rows of `cols` bf16 weights (cols a multiple of 8) to LDS at pitch `pitch`.
__device__ __forceinline__ void stage_weights(const uint16_t* src, int rows, int cols, uint16_t* dst, int pitch) {
    const int chunks = cols / 8;
    for (int i = threadIdx.x; i < rows * chunks; i += blockDim.x) {
        const int r = i / chunks, c = i - r * chunks;
        *reinterpret_cast<uint4*>(dst + r * pitch + 8 * c) = *reinterpret_cast<const uint4*>(src + r * cols + 8 * c);
    }
}

// The inference loop of one wave over its 16-row blocks of the lambertian and ward queues (the
// workgroup has staged the weights). kJobs: the vertex's guide frame comes from its job's
// table entry, else from the guide (a single render). Two-stage software pipeline, unrolled
// twice so that the two vertex register sets alternate without copies: while block b is
// encoded and run through the network, the vertex records of block b + stride and the queue
// entries of block b + 2 stride are in flight. Every load is issued unconditionally (positions
// past the queue end read its last entry; those rows run but are not stored), so that no
// branch makes the compiler wait for the loads in flight.
template <bool kJobs>
__device__ void infer_blocks(InferLds& t, const DevState& s, const WFParams& p, const DevGuide& g,
                             const DevJobs& jobs) {
    const uint32_t c1 = s.counts[kMatLambertian], c4 = s.counts[kMatWard];
    const uint32_t rows = c1 + c4;
    const uint32_t nblocks = (rows + 15) / 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane >> 2, part = lane & 3;  // encode and head: 4 lanes per row
    uint16_t* A = t.wave[wave].A;
    uint16_t* B = t.wave[wave].B;
    float* F = reinterpret_cast<float*>(A);
    const uint32_t stride = gridDim.x * kInferWaves;
    uint32_t blk = blockIdx.x * kInferWaves + wave;
    if (blk >= nblocks) return;
    auto slot_at = [&](uint32_t b) -> uint32_t {
        const uint32_t i = min(b * 16 + row, rows - 1);
        return s.items[i < c1 ? (size_t)kMatLambertian * p.n + i : (size_t)kMatWard * p.n + (i - c1)];
    };
    struct Vtx {
        TraceRay r;
        ExtRec e;
        uint32_t job;
    };
    auto fetch = [&](uint32_t slot, Vtx& v) {
        v.r = s.ray[slot];
        v.e = s.ext[slot];
        if (kJobs) v.job = jobs.job_of_block[slot / kJobBlock];
    };
    auto process = [&](uint32_t b, uint32_t slot, const Vtx& v) {
        {
            const GuideFrame fr = kJobs ? jobs.jobs[v.job].gframe
                                        : GuideFrame{{g.lo[0], g.lo[1], g.lo[2]},
                                                     {g.inv_ext[0], g.inv_ext[1], g.inv_ext[2]}, 0u};
            float p01[3], n[3], wo[3];
            for (int a = 0; a < 3; ++a) {
                float P = v.r.ro[a] + v.r.rd[a] * v.e.tmin;
                p01[a] = (P - fr.lo[a]) * fr.inv_ext[a];
                wo[a] = -v.r.rd[a];
            }
            // the normal flipped toward wo
            const float sg = v.e.ngx * wo[0] + v.e.ngy * wo[1] + v.e.ngz * wo[2] < 0.0f ? -1.0f : 1.0f;
            n[0] = sg * v.e.ngx;
            n[1] = sg * v.e.ngy;
            n[2] = sg * v.e.ngz;
            encode_row(A + row * kPX, p01, n, wo, fr.cls, part);
        }
        wave_lds_sync();
        fwd_layer<kGIn>(A, kPX, t.W0, kPX, t.bias[0], B, nullptr, 0, lane, true);
        wave_lds_sync();
        fwd_layer<kGHid>(B, kPH, t.W1, kPH, t.bias[1], A, nullptr, 0, lane, true);
        wave_lds_sync();
        fwd_layer<kGHid>(A, kPH, t.W2, kPH, t.bias[2], B, nullptr, 0, lane, true);
        wave_lds_sync();
        fwd_layer<kGHid, (5 * kGLobes + 15) / 16>(B, kPH, t.W3, kPH, t.bias[3], nullptr, F, 0, lane, false);
        wave_lds_sync();
        {
            float L[2 * kGLobeFloats];
            // all four lanes of a row take part
            head_lobes_part(F + row * kPF, part, g.log_kappa_max, L);
            const uint32_t at = b * 16 + row;  // the vertex's position in the two queues
            if (at < rows) {
                // this lane's 12 floats are planes 3 part .. 3 part + 2 of the vertex
                float4* dst = reinterpret_cast<float4*>(g.lobes) + (size_t)(3 * part) * g.nslots + at;
                dst[0] = make_float4(L[0], L[1], L[2], L[3]);
                dst[g.nslots] = make_float4(L[4], L[5], L[6], L[7]);
                dst[2 * (size_t)g.nslots] = make_float4(L[8], L[9], L[10], L[11]);
                if (g.learn_select && part == 0) g.sel[at] = guide_selection(F[row * kPF + kGSelOut]);
            }
        }
        wave_lds_sync();  // the next block's encode overwrites the outputs
    };
    Vtx va, vb;
    uint32_t sa = slot_at(blk);
    fetch(sa, va);
    uint32_t sb = slot_at(blk + stride);
    while (true) {
        fetch(sb, vb);
        const uint32_t sa2 = slot_at(blk + 2 * stride);
        process(blk, sa, va);
        blk += stride;
        if (blk >= nblocks) break;
        fetch(sa2, va);
        const uint32_t sb2 = slot_at(blk + 2 * stride);
        process(blk, sb, vb);
        blk += stride;
        if (blk >= nblocks) break;
        sa = sa2;
        sb = sb2;
    }
}

// Inference over the lambertian and ward queues: lobe parameters for every guided vertex.
__global__ void __launch_bounds__(64 * kInferWaves) k_guide_infer(DevState s, WFParams p, DevGuide g, DevJobs jobs) {
    __shared__ InferLds t;
    stage_weights(g.w_bf + kGoW0, kGHid, kGIn, t.W0, kPX);
    stage_weights(g.w_bf + kGoW1, kGHid, kGHid, t.W1, kPH);
    stage_weights(g.w_bf + kGoW2, kGHid, kGHid, t.W2, kPH);
    stage_weights(g.w_bf + kGoW3, kGOut, kGHid, t.W3, kPH);
    if (threadIdx.x < 4 * kGHid) {
        const int l = threadIdx.x / kGHid, c = threadIdx.x % kGHid;
        const int off = l == 0 ? kGoB0 : l == 1 ? kGoB1 : l == 2 ? kGoB2 : kGoB3;
        t.bias[l][c] = g.params[off + c];
    }
    __syncthreads();
    if (jobs.njobs) infer_blocks<true>(t, s, p, g, jobs);
    else infer_blocks<false>(t, s, p, g, jobs);
}

// Back-propagation through one layer for this wave's 16 rows:
// dIn = (dZ W) masked by In > 0, written as bf16 into dIn. WT: [in][64] transposed weights. The
// product is formed transposed (A: WT rows, B: dZ rows), so that each lane holds four
// consecutive columns of one row: In is read and dIn written 8 bytes at a time.
template <int K>
__device__ void bwd_dx(const uint16_t* dZ, const uint16_t* WT, const uint16_t* In, int pin, uint16_t* dIn, int r0,
                       int lane) {
    const int row = r0 + (lane & 15);
    bf16x8 z[2];  // this wave's rows of dZ, read once for all output tiles
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) z[kb] = ld8(dZ + row * kPH + kb * 32 + (lane >> 4) * 8);
#pragma unroll
    for (int it = 0; it < K / 16; ++it) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) {
            const bf16x8 w = ld8(WT + (it * 16 + (lane & 15)) * kGHid + kb * 32 + (lane >> 4) * 8);
            acc = mma(w, z[kb], acc);
        }
        const int col = it * 16 + (lane >> 4) * 4;
        const uint2 in = *reinterpret_cast<const uint2*>(In + row * pin + col);
        const float x0 = __uint_as_float(in.x << 16), x1 = __uint_as_float(in.x & 0xFFFF0000u);
        const float x2 = __uint_as_float(in.y << 16), x3 = __uint_as_float(in.y & 0xFFFF0000u);
        *reinterpret_cast<uint2*>(dIn + row * kPH + col) =
            make_uint2(f2bf2(x0 > 0.0f ? acc[0] : 0.0f, x1 > 0.0f ? acc[1] : 0.0f),
                       f2bf2(x2 > 0.0f ? acc[2] : 0.0f, x3 > 0.0f ? acc[3] : 0.0f));
    }
}

// Per-tile block of the training buffer: 32 rows of each matrix below, column-major
// (element (r, c) of a matrix at its column offset * 32 + c * 32 + r).
constexpr int kAcX = 0, kAcH1 = kAcX + kGIn, kAcH2 = kAcH1 + kGHid, kAcH3 = kAcH2 + kGHid;
constexpr int kAcD0 = kAcH3 + kGHid, kAcD1 = kAcD0 + kGHid, kAcD2 = kAcD1 + kGHid, kAcD3 = kAcD2 + kGHid;
constexpr int kAcCols = kAcD3 + kGOut;  // 544 columns, 34,816 bytes per tile

__device__ __forceinline__ void store_cm(const uint16_t* lds, int pitch, int cols, uint16_t* dst) {
    // each thread gathers 8 rows of one column from LDS and stores them as one 16-byte word;
    // four consecutive threads write a column's 64 bytes
    for (int idx = threadIdx.x; idx < cols * (kTile / 8); idx += blockDim.x) {
        const int c = idx / (kTile / 8), r0 = (idx % (kTile / 8)) * 8;
        const uint16_t* src = lds + r0 * pitch + c;
        uint32_t w[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) w[j] = (uint32_t)src[(2 * j) * pitch] | ((uint32_t)src[(2 * j + 1) * pitch] << 16);
        *reinterpret_cast<uint4*>(dst + c * kTile + r0) = make_uint4(w[0], w[1], w[2], w[3]);
    }
}

// Forward, loss gradient and input deltas of one tile of training rows; layer inputs and
// output deltas go to the tile's block of g.act.
__global__ void __launch_bounds__(128) k_guide_fwdbwd(DevGuide g, uint32_t first, uint32_t rows, float inv_batch) {
    __shared__ TileLds t;
    __shared__ uint16_t dA[kTile * kPH];
    __shared__ uint16_t dB[kTile * kPH];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r0 = wave * 16;
    const uint32_t base = blockIdx.x * kTile;
    if (base >= rows) return;
    uint16_t* blk = g.act + (size_t)blockIdx.x * kAcCols * kTile;
    {
        const int row = threadIdx.x >> 2, part = threadIdx.x & 3;
        float p01[3] = {0, 0, 0}, n[3] = {0, 0, 0}, wo[3] = {0, 0, 0};
        uint32_t cls = 0;
        if (base + row < rows) {
            const GuideSample& sm = g.samples[first + base + row];
            for (int a = 0; a < 3; ++a) {
                p01[a] = sm.p[a];
                n[a] = sm.n[a];
                wo[a] = sm.wo[a];
            }
            cls = sample_class(sm.scene_class);
        }
        encode_row(t.X + row * kPX, p01, n, wo, cls, part);
    }
    __syncthreads();
    tile_forward(t, g, r0, lane);
    __syncthreads();
    {
        const int row = threadIdx.x >> 2, part = threadIdx.x & 3;
        const bool live = base + row < rows;
        float wi[3] = {0.0f, 0.0f, 1.0f}, wgt = 0.0f, po = 0.0f;  // padding rows contribute nothing
        if (live) {
            const GuideSample& sm = g.samples[first + base + row];
            wi[0] = sm.wi[0];
            wi[1] = sm.wi[1];
            wi[2] = sm.wi[2];
            wgt = (g.weight_cap > 0.0f ? fminf(sm.weight, g.weight_cap) : sm.weight) * inv_batch;
            po = sm.other_pdf;
        }
        float d[10], dsel;
        head_grad_part(t.F + row * kPF, part, wi, wgt, g.log_kappa_max, po, g.learn_select != 0, d, &dsel);
        uint16_t* drow = dA + row * kPH;
#pragma unroll
        for (int i = 0; i < 10; ++i) drow[10 * part + i] = f2bf(d[i]);
#pragma unroll
        for (int i = 0; i < 6; ++i) drow[5 * kGLobes + 6 * part + i] = part == 0 && i == 0 ? f2bf(dsel) : (uint16_t)0;
    }
    __syncthreads();
    store_cm(t.X, kPX, kGIn, blk + kAcX * kTile);
    store_cm(t.H1, kPH, kGHid, blk + kAcH1 * kTile);
    store_cm(t.H2, kPH, kGHid, blk + kAcH2 * kTile);
    store_cm(t.H3, kPH, kGHid, blk + kAcH3 * kTile);
    store_cm(dA, kPH, kGOut, blk + kAcD3 * kTile);
    bwd_dx<kGHid>(dA, g.wt_bf + kGoW3, t.H3, kPH, dB, r0, lane);  // dZ2 = (dZ3 W3) * (H3 > 0)
    __syncthreads();
    store_cm(dB, kPH, kGHid, blk + kAcD2 * kTile);
    bwd_dx<kGHid>(dB, g.wt_bf + kGoW2, t.H2, kPH, dA, r0, lane);
    __syncthreads();
    store_cm(dA, kPH, kGHid, blk + kAcD1 * kTile);
    bwd_dx<kGHid>(dA, g.wt_bf + kGoW1, t.H1, kPH, dB, r0, lane);
    __syncthreads();
    store_cm(dB, kPH, kGHid, blk + kAcD0 * kTile);
}

// Weight and bias gradients of one layer over a split of the tiles:
// dW[o][i] = sum_r dZ[r][o] In[r][i] with the 32 rows of a tile as the MFMA's k, written as
// this split's partial sum. blockIdx.x = split * 4 + layer; four waves share the output tiles.
__global__ void __launch_bounds__(256) k_guide_dw(DevGuide g, uint32_t ntiles, uint32_t splits) {
    const int layer = blockIdx.x & 3, split = blockIdx.x >> 2;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int K = layer == 0 ? kGIn : kGHid;
    const int in_col = layer == 0 ? kAcX : layer == 1 ? kAcH1 : layer == 2 ? kAcH2 : kAcH3;
    const int dz_col = layer == 0 ? kAcD0 : layer == 1 ? kAcD1 : layer == 2 ? kAcD2 : kAcD3;
    const int w_off = layer == 0 ? kGoW0 : layer == 1 ? kGoW1 : layer == 2 ? kGoW2 : kGoW3;
    const int b_off = layer == 0 ? kGoB0 : layer == 1 ? kGoB1 : layer == 2 ? kGoB2 : kGoB3;
    const uint32_t t0 = (uint32_t)((uint64_t)ntiles * split / splits);
    const uint32_t t1 = (uint32_t)((uint64_t)ntiles * (split + 1) / splits);
    const int ntile_out = 4 * (K / 16);  // 16x16 output tiles of this layer: 16 or 24
    __shared__ float bpart[4][kGHid];
    f32x4 acc[6];
    for (int j = 0; j < 6; ++j) acc[j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    float bsum = 0.0f;  // column (threadIdx & 63) of dZ, rows 8 (threadIdx / 64) .. + 7 of each tile
    // unrolled so that the loads of several tiles are in flight together
#pragma unroll 4
    for (uint32_t tt = t0; tt < t1; ++tt) {
        const uint16_t* blk = g.act + (size_t)tt * kAcCols * kTile;
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            const int tix = wave + 4 * j;
            if (tix < ntile_out) {
                const int ot = tix / (K / 16), it = tix % (K / 16);
                const bf16x8 a = ld8(blk + (dz_col + ot * 16 + (lane & 15)) * kTile + (lane >> 4) * 8);
                const bf16x8 b = ld8(blk + (in_col + it * 16 + (lane & 15)) * kTile + (lane >> 4) * 8);
                acc[j] = mma(a, b, acc[j]);
            }
        }
        const uint4 v = *reinterpret_cast<const uint4*>(blk + (dz_col + lane) * kTile + wave * 8);
        bsum += (bf2f((uint16_t)v.x) + bf2f((uint16_t)(v.x >> 16))) + (bf2f((uint16_t)v.y) + bf2f((uint16_t)(v.y >> 16))) +
                (bf2f((uint16_t)v.z) + bf2f((uint16_t)(v.z >> 16))) + (bf2f((uint16_t)v.w) + bf2f((uint16_t)(v.w >> 16)));
    }
    bpart[wave][lane] = bsum;
    __syncthreads();
    float* dst = g.partials + (size_t)split * kGParams;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        const int tix = wave + 4 * j;
        if (tix < ntile_out) {
            const int ot = tix / (K / 16), it = tix % (K / 16);
            const int i = it * 16 + (lane & 15);
            for (int r = 0; r < 4; ++r) {
                const int o = ot * 16 + (lane >> 4) * 4 + r;
                dst[w_off + o * K + i] = acc[j][r];
            }
        }
    }
    if (threadIdx.x < kGHid)
        dst[b_off + threadIdx.x] = (bpart[0][threadIdx.x] + bpart[1][threadIdx.x]) +
                                   (bpart[2][threadIdx.x] + bpart[3][threadIdx.x]);
}

// Forward only over given encoded-input samples (tests): raw head outputs per row.
__global__ void __launch_bounds__(128) k_guide_forward(DevGuide g, const GuideSample* in, uint32_t rows,
                                                       float* out) {
    __shared__ TileLds t;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t base = blockIdx.x * kTile;
    if (base >= rows) return;
    {
        const int row = threadIdx.x >> 2, part = threadIdx.x & 3;
        float p01[3] = {0, 0, 0}, n[3] = {0, 0, 0}, wo[3] = {0, 0, 0};
        uint32_t cls = 0;
        if (base + row < rows) {
            for (int a = 0; a < 3; ++a) {
                p01[a] = in[base + row].p[a];
                n[a] = in[base + row].n[a];
                wo[a] = in[base + row].wo[a];
            }
            cls = sample_class(in[base + row].scene_class);
        }
        encode_row(t.X + row * kPX, p01, n, wo, cls, part);
    }
    __syncthreads();
    tile_forward(t, g, wave * 16, lane);
    __syncthreads();
    for (int i = threadIdx.x; i < kTile * kGOut; i += blockDim.x) {
        const int row = i / kGOut, col = i % kGOut;
        if (base + row < rows) out[(size_t)(base + row) * kGOut + col] = t.F[row * kPF + col];
    }
}

__device__ __forceinline__ bool is_weight(int i, int* layer_off, int* nin, int* nout) {
    if (i < kGoB0) { *layer_off = kGoW0; *nin = kGIn; *nout = kGHid; return true; }
    if (i >= kGoW1 && i < kGoB1) { *layer_off = kGoW1; *nin = kGHid; *nout = kGHid; return true; }
    if (i >= kGoW2 && i < kGoB2) { *layer_off = kGoW2; *nin = kGHid; *nout = kGHid; return true; }
    if (i >= kGoW3 && i < kGoB3) { *layer_off = kGoW3; *nin = kGHid; *nout = kGOut; return true; }
    return false;
}

// Sum of the `groups` partial gradients, then Adam (Kingma & Ba 2015) on the fp32 master
// weights and a refresh of the bf16 copies. lr == 0: the summed gradient is stored in g.grad
// instead (groups == 0 only refreshes the copies). A workgroup covers 64 parameters; its four
// waves sum interleaved quarters of the partials.
__global__ void __launch_bounds__(256) k_adam(DevGuide g, uint32_t groups, float lr, float b1, float b2, float eps,
                                              float c1, float c2) {
    __shared__ float part[4][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + lane;
    float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (i < kGParams) {
        uint32_t w = (uint32_t)wave;
        for (; w + 28 < groups; w += 32) {
#pragma unroll
            for (int u = 0; u < 8; ++u) acc[u] += g.partials[(size_t)(w + 4 * u) * kGParams + i];
        }
        for (; w < groups; w += 4) acc[0] += g.partials[(size_t)w * kGParams + i];
    }
    part[wave][lane] = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    __syncthreads();
    if (wave != 0 || i >= kGParams) return;
    const float gr = (part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]);
    float w = g.params[i];
    if (lr != 0.0f) {
        float m = b1 * g.adam_m[i] + (1.0f - b1) * gr;
        float v = b2 * g.adam_v[i] + (1.0f - b2) * gr * gr;
        g.adam_m[i] = m;
        g.adam_v[i] = v;
        const float mh = m / c1, vh = v / c2;
        w -= lr * mh / (sqrtf(vh) + eps);
        g.params[i] = w;
    } else if (groups != 0) {
        g.grad[i] = gr;
    }
    int off, nin, nout;
    if (is_weight(i, &off, &nin, &nout)) {
        const int o = (i - off) / nin, in = (i - off) % nin;
        const uint16_t b = f2bf(w);
        g.w_bf[i] = b;
        g.wt_bf[off + in * nout + o] = b;
    }
}

}  // namespace

hipError_t guide_infer(const DevState& st, const WFParams& p, const DevGuide& g, const DevJobs& jobs,
                       hipStream_t s) {
    hipLaunchKernelGGL(k_guide_infer, dim3(g.infer_blocks ? g.infer_blocks : 256u), dim3(64 * kInferWaves), 0, s, st,
                       p, g, jobs);
    return hipGetLastError();
}

// Forward/backward of `rows` samples from `first`, then the weight gradients; returns the
// number of partial sums written.
static uint32_t launch_gradients(const DevGuide& g, uint32_t first, uint32_t rows, hipStream_t s) {
    const uint32_t ntiles = (rows + kTile - 1) / kTile;
    const uint32_t splits = std::min<uint32_t>(kGTrainGroups, ntiles);
    hipLaunchKernelGGL(k_guide_fwdbwd, dim3(ntiles), dim3(128), 0, s, g, first, rows, 1.0f / (float)rows);
    hipLaunchKernelGGL(k_guide_dw, dim3(4 * splits), dim3(256), 0, s, g, ntiles, splits);
    return splits;
}

hipError_t guide_train_batch(const DevGuide& g, uint32_t first, uint32_t rows, float lr, uint32_t step,
                             hipStream_t s) {
    if (rows == 0) return hipSuccess;
    if (rows > g.act_rows) return hipErrorInvalidValue;
    const uint32_t groups = launch_gradients(g, first, rows, s);
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3((kGParams + 63) / 64), dim3(256), 0, s, g, groups, lr, b1, b2, 1e-8f, c1, c2);
    return hipGetLastError();
}

hipError_t guide_refresh(const DevGuide& g, hipStream_t s) {
    hipLaunchKernelGGL(k_adam, dim3((kGParams + 63) / 64), dim3(256), 0, s, g, 0u, 0.0f, 0.9f, 0.999f, 1e-8f,
                       1.0f, 1.0f);
    return hipGetLastError();
}

hipError_t guide_forward(const DevGuide& g, const GuideSample* in, uint32_t rows, float* out, hipStream_t s) {
    if (rows == 0) return hipSuccess;
    hipLaunchKernelGGL(k_guide_forward, dim3((rows + kTile - 1) / kTile), dim3(128), 0, s, g, in, rows, out);
    return hipGetLastError();
}

hipError_t guide_gradients(const DevGuide& g, uint32_t rows, hipStream_t s) {
    if (rows == 0) return hipSuccess;
    if (rows > g.act_rows) return hipErrorInvalidValue;
    const uint32_t groups = launch_gradients(g, 0u, rows, s);
    hipLaunchKernelGGL(k_adam, dim3((kGParams + 63) / 64), dim3(256), 0, s, g, groups, 0.0f, 0.9f, 0.999f, 1e-8f,
                       1.0f, 1.0f);
    return hipGetLastError();
}

}  // namespace rgpu
