// Path-guiding network kernels on gfx950 matrix cores (bf16 operands, fp32 accumulation).
//
// A workgroup of two wave64s processes a tile of 32 rows (one row = one path vertex or one
// training sample); each wave owns 16 rows. Every layer is a product of the tile's bf16
// activations held in LDS with the bf16 weights read from global memory (they stay in L1/L2:
// 37 KB per copy), computed 16x16x32 at a time by v_mfma_f32_16x16x32_bf16:
//   A operand: lane l holds 8 consecutive k of row (l mod 16),
//   B operand: lane l holds the same 8 k of column (l mod 16),
//   result:    lane l holds rows 4*(l/16)..+3 of column (l mod 16).
// Back-propagation computes input deltas with the transposed weight copy and weight
// gradients over the tile's 32 rows, accumulated across workgroups with float atomics.
#include <hip/hip_runtime.h>
#include <stdint.h>

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

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
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
template <int K>
__device__ void fwd_layer(const uint16_t* X, int px, const uint16_t* W, const float* bias, uint16_t* H,
                          float* F, int r0, int lane, bool relu) {
    for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int kb = 0; kb < K / 32; ++kb) {
            bf16x8 a = ld8(X + (r0 + (lane & 15)) * px + kb * 32 + (lane >> 4) * 8);
            bf16x8 b = ld8(W + (nt * 16 + (lane & 15)) * K + kb * 32 + (lane >> 4) * 8);
            acc = mma(a, b, acc);
        }
        const int col = nt * 16 + (lane & 15);
        const float bb = bias[col];
        for (int r = 0; r < 4; ++r) {
            const int row = r0 + (lane >> 4) * 4 + r;
            float v = acc[r] + bb;
            if (F) F[row * kPF + col] = v;
            else H[row * kPH + col] = f2bf(relu ? fmaxf(v, 0.0f) : v);
        }
    }
}

// Encoded input features of one row (written as bf16 into X).
__device__ void encode_row(uint16_t* X, const float* p01, const float* n, const float* wo, int part) {
    // part 0..3: each of four lanes writes 24 of the 96 features
    for (int f = part * 24; f < part * 24 + 24; ++f) {
        float v;
        if (f < 72) {
            int k = f / 6, a = (f % 6) / 2;
            double arg = (double)(1u << k) * 3.14159265358979323846 * (double)p01[a];
            v = (f & 1) ? (float)cos(arg) : (float)sin(arg);
        } else if (f < 75) {
            v = n[f - 72];
        } else if (f < 78) {
            v = wo[f - 75];
        } else {
            v = 0.0f;
        }
        X[f] = f2bf(v);
    }
}

// Raw head outputs of one row -> 8 lobes (mu normalised, kappa, log weight).
__device__ void head_lobes(const float* out, float* lobes) {
    float mx = -3.0e38f;
    for (int k = 0; k < kGLobes; ++k) mx = fmaxf(mx, out[5 * k + 4]);
    float z = 0.0f;
    for (int k = 0; k < kGLobes; ++k) z += expf(out[5 * k + 4] - mx);
    const float lz = logf(z);
    for (int k = 0; k < kGLobes; ++k) {
        const float* o = out + 5 * k;
        float m = sqrtf(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]) + 1e-12f;
        lobes[5 * k + 0] = o[0] / m;
        lobes[5 * k + 1] = o[1] / m;
        lobes[5 * k + 2] = o[2] / m;
        lobes[5 * k + 3] = expf(fminf(fmaxf(o[3], kGLogKappaMin), kGLogKappaMax));
        lobes[5 * k + 4] = o[4] - mx - lz;
    }
}

// d(-weight * log q(wi)) / d(raw outputs); zero for the 24 unused outputs.
__device__ void head_grad(const float* out, const float* wi, float weight, float* d) {
    float L[kGSlotFloats];
    head_lobes(out, L);
    float a[kGLobes], mx = -3.0e38f;
    float c[kGLobes];
    for (int k = 0; k < kGLobes; ++k) {
        c[k] = L[5 * k] * wi[0] + L[5 * k + 1] * wi[1] + L[5 * k + 2] * wi[2];
        a[k] = L[5 * k + 4] + guide_log_vmf(L[5 * k + 3], c[k]);
        mx = fmaxf(mx, a[k]);
    }
    float s = 0.0f;
    for (int k = 0; k < kGLobes; ++k) s += expf(a[k] - mx);
    const float lse = mx + logf(s);
    for (int i = 0; i < kGOut; ++i) d[i] = 0.0f;
    for (int k = 0; k < kGLobes; ++k) {
        const float r = expf(a[k] - lse);
        const float pi_k = expf(L[5 * k + 4]);
        d[5 * k + 4] = weight * (pi_k - r);
        const float kap = L[5 * k + 3];
        const float e = expf(-2.0f * kap);
        const float dlogf = 1.0f / kap - 2.0f * e / (1.0f - e) + (c[k] - 1.0f);
        const float raw = out[5 * k + 3];
        d[5 * k + 3] = (raw > kGLogKappaMin && raw < kGLogKappaMax) ? -weight * r * dlogf * kap : 0.0f;
        const float* o = out + 5 * k;
        const float m = sqrtf(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]) + 1e-12f;
        for (int q = 0; q < 3; ++q) d[5 * k + q] = -weight * r * kap * (wi[q] - c[k] * L[5 * k + q]) / m;
    }
}

// Shared forward pass of a tile: X filled by the caller; leaves H1..H3 (bf16) and F (fp32 outputs).
struct TileLds {
    uint16_t X[kTile * kPX];
    uint16_t H1[kTile * kPH];
    uint16_t H2[kTile * kPH];
    uint16_t H3[kTile * kPH];
    float F[kTile * kPF];
};

__device__ void tile_forward(TileLds& t, const DevGuide& g, int r0, int lane) {
    fwd_layer<kGIn>(t.X, kPX, g.w_bf + kGoW0, g.params + kGoB0, t.H1, nullptr, r0, lane, true);
    fwd_layer<kGHid>(t.H1, kPH, g.w_bf + kGoW1, g.params + kGoB1, t.H2, nullptr, r0, lane, true);
    fwd_layer<kGHid>(t.H2, kPH, g.w_bf + kGoW2, g.params + kGoB2, t.H3, nullptr, r0, lane, true);
    fwd_layer<kGHid>(t.H3, kPH, g.w_bf + kGoW3, g.params + kGoB3, nullptr, t.F, r0, lane, false);
}

// Inference over the lambertian and ward queues: lobe parameters for every guided vertex.
__global__ void __launch_bounds__(128) k_guide_infer(DevState s, WFParams p, DevGuide g) {
    __shared__ TileLds t;
    const uint32_t c1 = s.counts[kMatLambertian], c4 = s.counts[kMatWard];
    const uint32_t rows = c1 + c4;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r0 = wave * 16;
    for (uint32_t base = blockIdx.x * kTile; base < rows; base += gridDim.x * kTile) {
        // encode: 4 threads per row
        {
            const int row = threadIdx.x >> 2, part = threadIdx.x & 3;
            const uint32_t i = base + row;
            float p01[3] = {0, 0, 0}, n[3] = {0, 0, 0}, wo[3] = {0, 0, 0};
            if (i < rows) {
                const uint32_t slot = i < c1 ? s.items[(size_t)kMatLambertian * p.n + i]
                                             : s.items[(size_t)kMatWard * p.n + (i - c1)];
                const TraceRay r = s.ray[slot];
                const ExtRec e = s.ext[slot];
                for (int a = 0; a < 3; ++a) {
                    float P = r.ro[a] + r.rd[a] * e.tmin;
                    p01[a] = (P - g.lo[a]) * g.inv_ext[a];
                    wo[a] = -r.rd[a];
                }
                n[0] = e.ngx;
                n[1] = e.ngy;
                n[2] = e.ngz;
                if (n[0] * wo[0] + n[1] * wo[1] + n[2] * wo[2] < 0.0f) {
                    n[0] = -n[0];
                    n[1] = -n[1];
                    n[2] = -n[2];
                }
            }
            encode_row(t.X + row * kPX, p01, n, wo, part);
        }
        __syncthreads();
        tile_forward(t, g, r0, lane);
        __syncthreads();
        if (threadIdx.x < kTile) {
            const uint32_t i = base + threadIdx.x;
            if (i < rows) {
                const uint32_t slot = i < c1 ? s.items[(size_t)kMatLambertian * p.n + i]
                                             : s.items[(size_t)kMatWard * p.n + (i - c1)];
                float L[kGSlotFloats];
                head_lobes(t.F + threadIdx.x * kPF, L);
                float* dst = g.lobes + (size_t)slot * kGSlotFloats;
                for (int q = 0; q < kGSlotFloats; ++q) dst[q] = L[q];
            }
        }
        __syncthreads();
    }
}

// Back-propagation through one layer for this wave's 16 rows:
// dIn = (dZ W) masked by In > 0, written as bf16 into dIn (skip with dIn == nullptr).
template <int K>
__device__ void bwd_dx(const uint16_t* dZ, const uint16_t* WT, const uint16_t* In, int pin, uint16_t* dIn, int r0,
                       int lane) {
    for (int it = 0; it < K / 16; ++it) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int kb = 0; kb < 2; ++kb) {
            bf16x8 a = ld8(dZ + (r0 + (lane & 15)) * kPH + kb * 32 + (lane >> 4) * 8);
            bf16x8 b = ld8(WT + (it * 16 + (lane & 15)) * kGHid + kb * 32 + (lane >> 4) * 8);
            acc = mma(a, b, acc);
        }
        const int col = it * 16 + (lane & 15);
        for (int r = 0; r < 4; ++r) {
            const int row = r0 + (lane >> 4) * 4 + r;
            const float x = bf2f(In[row * pin + col]);
            dIn[row * kPH + col] = f2bf(x > 0.0f ? acc[r] : 0.0f);
        }
    }
}

// Weight and bias gradients of one layer over the tile's 32 rows: dW[o][i] += sum_r dZ[r][o] In[r][i].
template <int K>
__device__ void bwd_dw(const uint16_t* dZ, const uint16_t* In, int pin, float* gW, float* gB, int wave, int lane) {
    const int ntiles = 4 * (K / 16);
    for (int tix = wave; tix < ntiles; tix += 2) {
        const int ot = tix / (K / 16), it = tix % (K / 16);
        bf16x8 a = ld8s(dZ + ((lane >> 4) * 8) * kPH + ot * 16 + (lane & 15), kPH);
        bf16x8 b = ld8s(In + ((lane >> 4) * 8) * pin + it * 16 + (lane & 15), pin);
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = mma(a, b, acc);
        const int i = it * 16 + (lane & 15);
        for (int r = 0; r < 4; ++r) {
            const int o = ot * 16 + (lane >> 4) * 4 + r;
            unsafeAtomicAdd(gW + o * K + i, acc[r]);
        }
    }
    const int tid = wave * 64 + lane;
    if (tid < kGHid) {
        float sb = 0.0f;
        for (int r = 0; r < kTile; ++r) sb += bf2f(dZ[r * kPH + tid]);
        unsafeAtomicAdd(gB + tid, sb);
    }
}

// Forward + loss gradient + back-propagation for `rows` training samples starting at `first`.
__global__ void __launch_bounds__(128) k_guide_train(DevGuide g, uint32_t first, uint32_t rows, float inv_batch) {
    __shared__ TileLds t;
    __shared__ uint16_t dA[kTile * kPH];
    __shared__ uint16_t dB[kTile * kPH];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r0 = wave * 16;
    const uint32_t base = blockIdx.x * kTile;
    if (base >= rows) return;
    {
        const int row = threadIdx.x >> 2, part = threadIdx.x & 3;
        float p01[3] = {0, 0, 0}, n[3] = {0, 0, 0}, wo[3] = {0, 0, 0};
        if (base + row < rows) {
            const GuideSample& sm = g.samples[first + base + row];
            for (int a = 0; a < 3; ++a) {
                p01[a] = sm.p[a];
                n[a] = sm.n[a];
                wo[a] = sm.wo[a];
            }
        }
        encode_row(t.X + row * kPX, p01, n, wo, part);
    }
    __syncthreads();
    tile_forward(t, g, r0, lane);
    __syncthreads();
    if (threadIdx.x < kTile) {
        const int row = threadIdx.x;
        float d[kGOut];
        if (base + row < rows) {
            const GuideSample& sm = g.samples[first + base + row];
            head_grad(t.F + row * kPF, sm.wi, sm.weight * inv_batch, d);
        } else {
            for (int i = 0; i < kGOut; ++i) d[i] = 0.0f;
        }
        for (int i = 0; i < kGOut; ++i) dA[row * kPH + i] = f2bf(d[i]);
    }
    __syncthreads();
    // layer 3: dW3, then dZ2 = (dY3 W3) * (H3 > 0)
    bwd_dw<kGHid>(dA, t.H3, kPH, g.grad + kGoW3, g.grad + kGoB3, wave, lane);
    bwd_dx<kGHid>(dA, g.wt_bf + kGoW3, t.H3, kPH, dB, r0, lane);
    __syncthreads();
    bwd_dw<kGHid>(dB, t.H2, kPH, g.grad + kGoW2, g.grad + kGoB2, wave, lane);
    bwd_dx<kGHid>(dB, g.wt_bf + kGoW2, t.H2, kPH, dA, r0, lane);
    __syncthreads();
    bwd_dw<kGHid>(dA, t.H1, kPH, g.grad + kGoW1, g.grad + kGoB1, wave, lane);
    bwd_dx<kGHid>(dA, g.wt_bf + kGoW1, t.H1, kPH, dB, r0, lane);
    __syncthreads();
    bwd_dw<kGIn>(dB, t.X, kPX, g.grad + kGoW0, g.grad + kGoB0, wave, lane);
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
        if (base + row < rows)
            for (int a = 0; a < 3; ++a) {
                p01[a] = in[base + row].p[a];
                n[a] = in[base + row].n[a];
                wo[a] = in[base + row].wo[a];
            }
        encode_row(t.X + row * kPX, p01, n, wo, part);
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

// Adam (Kingma & Ba 2015) on the fp32 master weights, then refresh the bf16 copies and clear
// the gradient. lr == 0 only refreshes the copies.
__global__ void __launch_bounds__(256) k_adam(DevGuide g, float lr, float b1, float b2, float eps, float c1,
                                              float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kGParams) return;
    float w = g.params[i];
    if (lr != 0.0f) {
        const float gr = g.grad[i];
        float m = b1 * g.adam_m[i] + (1.0f - b1) * gr;
        float v = b2 * g.adam_v[i] + (1.0f - b2) * gr * gr;
        g.adam_m[i] = m;
        g.adam_v[i] = v;
        const float mh = m / c1, vh = v / c2;
        w -= lr * mh / (sqrtf(vh) + eps);
        g.params[i] = w;
    }
    g.grad[i] = 0.0f;
    int off, nin, nout;
    if (is_weight(i, &off, &nin, &nout)) {
        const int o = (i - off) / nin, in = (i - off) % nin;
        const uint16_t b = f2bf(w);
        g.w_bf[i] = b;
        g.wt_bf[off + in * nout + o] = b;
    }
}

}  // namespace

hipError_t guide_infer(const DevState& st, const WFParams& p, const DevGuide& g, int blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_guide_infer, dim3(blocks > 0 ? blocks : 2048), dim3(128), 0, s, st, p, g);
    return hipGetLastError();
}

hipError_t guide_train_batch(const DevGuide& g, uint32_t first, uint32_t rows, float lr, uint32_t step,
                             hipStream_t s) {
    if (rows == 0) return hipSuccess;
    hipLaunchKernelGGL(k_guide_train, dim3((rows + kTile - 1) / kTile), dim3(128), 0, s, g, first, rows,
                       1.0f / (float)rows);
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3((kGParams + 255) / 256), dim3(256), 0, s, g, lr, b1, b2, 1e-8f, c1, c2);
    return hipGetLastError();
}

hipError_t guide_refresh(const DevGuide& g, hipStream_t s) {
    hipLaunchKernelGGL(k_adam, dim3((kGParams + 255) / 256), dim3(256), 0, s, g, 0.0f, 0.9f, 0.999f, 1e-8f, 1.0f,
                       1.0f);
    return hipGetLastError();
}

hipError_t guide_forward(const DevGuide& g, const GuideSample* in, uint32_t rows, float* out, hipStream_t s) {
    if (rows == 0) return hipSuccess;
    hipLaunchKernelGGL(k_guide_forward, dim3((rows + kTile - 1) / kTile), dim3(128), 0, s, g, in, rows, out);
    return hipGetLastError();
}

hipError_t guide_gradients(const DevGuide& g, uint32_t rows, hipStream_t s) {
    if (rows == 0) return hipSuccess;
    hipLaunchKernelGGL(k_guide_train, dim3((rows + kTile - 1) / kTile), dim3(128), 0, s, g, 0u, rows,
                       1.0f / (float)rows);
    return hipGetLastError();
}

}  // namespace rgpu
