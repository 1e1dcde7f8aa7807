// Wavefront path-tracing stages on gfx950.
//
// One bounce per step, with the reference's stage order and per-slot semantics
// (common/wavefront_path_tracing.hpp:591-939, kernels/*.cl), reorganised for MI355X:
//  - the strategy selection, envmap / BSDF sampling, pdf evaluation and shading that the
//    reference runs as separate full-N launches are fused into one queue-driven kernel per
//    material, which keeps the incident sample in registers. Every slot still draws its
//    random numbers in the reference order: strategy (1), then the envmap sample or the
//    BSDF sample, then the dielectric choice (1); a slot belongs to exactly one queue, so
//    fusing per material preserves each slot's stream;
//  - volume free-flight sampling (homogeneous_volume_through) is fused into the
//    extension ray cast, which runs over every slot;
//  - queues are built by wave ballots and a block prefix sum into per-block counts, a
//    scan over blocks and a scatter: a stable, slot-ordered compaction, so the pixel that
//    new_path assigns to a slot (next_pixel_index + queue position) is deterministic;
//  - queue kernels run a device-sized grid that strides over the queue length read from
//    device memory, instead of N threads that mostly exit.
// Transcendentals are evaluated in double and rounded once to float, and the library is
// built with -ffp-contract=off, so results agree with the CPU oracle to the last bit
// except where accumulation order differs (pixel sums use float atomics).
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <type_traits>
#include <stdint.h>

#include "guide_device.h"
#include "rgpu_wavefront.h"
#include "stage_math.h"
#include "trace_device.h"

namespace rgpu {
namespace {

const float kPiF = 3.14159265358979323846f;
const double kPiD = 3.14159265358979323846;
const float kFltMax = 3.402823466e+38f;

// ---------------------------------------------------------------- math
// The reference's float builtins, evaluated in double by the stage_math.h functions (defined
// results, independent of any libm) and rounded once.
__device__ __forceinline__ float fsin(float x) { return (float)sm_sin((double)x); }
__device__ __forceinline__ float fcos(float x) { return (float)sm_cos((double)x); }
__device__ __forceinline__ float flog(float x) { return (float)sm_log((double)x); }
__device__ __forceinline__ float fexp(float x) { return (float)sm_exp((double)x); }
__device__ __forceinline__ float fatan(float x) { return (float)sm_atan((double)x); }
__device__ __forceinline__ float fatan2(float y, float x) { return (float)sm_atan2((double)y, (double)x); }
__device__ __forceinline__ float fpow(float x, float y) { return (float)sm_pow((double)x, (double)y); }
// "x * M_PI * 2.0f" with OpenCL's double M_PI (kernels/lambertian.cl:11, new_path.cl:63)
__device__ __forceinline__ float two_pi_times(float u) { return (float)((double)u * kPiD * 2.0); }

__device__ __forceinline__ F3 add3(F3 a, F3 b) { return F3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ F3 mul3(F3 a, float s) { return F3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ F3 mul3v(F3 a, F3 b) { return F3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ F3 neg3(F3 a) { return F3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ F3 normalize3(F3 v) {
    float inv = 1.0f / sqrtf(dot3(v, v));
    return mul3(v, inv);
}
__device__ __forceinline__ F3 load3(const float* p) { return F3{p[0], p[1], p[2]}; }
__device__ __forceinline__ bool finite3(F3 v) { return isfinite(v.x) && isfinite(v.y) && isfinite(v.z); }

// ---------------------------------------------------------------- RNG (kernels/peseudo_random.cl)
__device__ __forceinline__ uint32_t rotl(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }
__device__ __forceinline__ uint32_t xo_next(uint4& s) {
    const uint32_t r = rotl(s.x * 5u, 7) * 9u;
    const uint32_t t = s.y << 9;
    s.z ^= s.x;
    s.w ^= s.y;
    s.y ^= s.z;
    s.x ^= s.w;
    s.z ^= t;
    s.w = rotl(s.w, 11);
    return r;
}
__device__ __forceinline__ float xo_uniform(uint4& s) {
    return __uint_as_float((xo_next(s) >> 9) | 0x3f800000u) - 1.0f;
}
__device__ __forceinline__ unsigned long long xo_uint62(uint4& s) {
    unsigned long long a = xo_next(s) >> 1;
    unsigned long long b = xo_next(s) >> 1;
    return (a << 31) | b;
}
__device__ __forceinline__ unsigned long long splitmix64(unsigned long long* x) {
    unsigned long long z = (*x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// ---------------------------------------------------------------- envmap (kernels/envmap_sampling.cl)
__device__ __forceinline__ void cartesian_to_polar(F3 rd, float* theta, float* phi) {
    float z = rd.y, x = rd.z, y = rd.x;
    *theta = fatan2(sqrtf(x * x + y * y), z);
    *phi = fatan2(y, x);
    if (!isfinite(*phi)) *phi = 0.0f;
}

// read_imagef with normalised coordinates and CLK_ADDRESS_REPEAT, nearest or bilinear
// (OpenCL 1.2 section 8.2 addressing/filtering formulas).
__device__ __forceinline__ F3 texel(const DevEnvmap& e, int i, int j) {
    float4 t = e.texels[(size_t)j * e.width + i];
    return F3{t.x, t.y, t.z};
}
__device__ F3 envmap_value(const DevEnvmap& e, F3 rd, bool filter) {
    float theta, phi;
    cartesian_to_polar(rd, &theta, &phi);
    const float pi = kPiF;
    float u = 1.0f - phi / (2.0f * pi);
    float v = theta / pi;
    const int W = (int)e.width, H = (int)e.height;
    float uu = (u - floorf(u)) * (float)W;
    float vv = (v - floorf(v)) * (float)H;
    // The final clamps change nothing for a finite direction; they keep a non-finite one in bounds.
    if (!filter) {
        int i = (int)floorf(uu);
        int j = (int)floorf(vv);
        if (i > W - 1) i -= W;
        if (j > H - 1) j -= H;
        return texel(e, min(max(i, 0), W - 1), min(max(j, 0), H - 1));
    }
    float us = uu - 0.5f, vs = vv - 0.5f;
    int i0 = (int)floorf(us), j0 = (int)floorf(vs);
    int i1 = i0 + 1, j1 = j0 + 1;
    float a = us - floorf(us), b = vs - floorf(vs);
    if (i0 < 0) i0 += W;
    if (j0 < 0) j0 += H;
    if (i1 > W - 1) i1 -= W;
    if (j1 > H - 1) j1 -= H;
    i0 = min(max(i0, 0), W - 1);
    i1 = min(max(i1, 0), W - 1);
    j0 = min(max(j0, 0), H - 1);
    j1 = min(max(j1, 0), H - 1);
    F3 t00 = texel(e, i0, j0), t10 = texel(e, i1, j0), t01 = texel(e, i0, j1), t11 = texel(e, i1, j1);
    float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b), w01 = (1.0f - a) * b, w11 = a * b;
    return F3{w00 * t00.x + w10 * t10.x + w01 * t01.x + w11 * t11.x,
              w00 * t00.y + w10 * t10.y + w01 * t01.y + w11 * t11.y,
              w00 * t00.z + w10 * t10.z + w01 * t01.z + w11 * t11.z};
}

__device__ __forceinline__ int fract_int(int x, int m) {
    int r = x % m;
    return r < 0 ? r + m : r;
}
__device__ __forceinline__ int envmap_index(const DevEnvmap& e, F3 wi) {
    float theta, phi;
    cartesian_to_polar(wi, &theta, &phi);
    const float pi = kPiF;
    float u = 1.0f - phi / (2.0f * pi);
    float v = theta / pi;
    int ix = (int)floorf(u * (float)e.width);
    int iy = (int)floorf(v * (float)e.height);
    ix = fract_int(ix, (int)e.width);
    iy = min(max(iy, 0), (int)e.height - 1);
    return iy * (int)e.width + ix;
}

__device__ __forceinline__ uint32_t alias_sample(uint4& st, const AliasBucket* table, uint32_t count) {
    uint32_t idx = (uint32_t)(xo_uint62(st) % (unsigned long long)count);
    AliasBucket b = table[idx];
    if (b.alias < 0) return idx;
    float u = xo_uniform(st);
    if (u < b.height) return idx;
    return (uint32_t)b.alias;
}

__device__ __forceinline__ F3 sample_from_fragment(const EnvFragment& f, uint4& st) {
    float y = f.beg_y + (f.end_y - f.beg_y) * xo_uniform(st);
    float phi = f.beg_phi + (f.end_phi - f.beg_phi) * xo_uniform(st);
    F3 p = F3{fsin(phi), y, fcos(phi)};
    float r_xz = sqrtf(fmaxf(1.0f - p.y * p.y, 0.0f));
    p.x *= r_xz;
    p.z *= r_xz;
    return p;
}

// The six-axis mixture n.x^2 p(+-X) + n.y^2 p(+-Y) + n.z^2 p(+-Z) of one texel, summed in the
// reference's order (envmap_sampling.cl:117-121, 152-155); the sign of each normal component
// picks the table (0 < n.c: the + table). Two 16-byte loads of the texel's record.
__device__ __forceinline__ float six_axis_mix(const DevEnvmap& e, uint32_t idx, F3 n, F3 ap) {
    const float4* r = reinterpret_cast<const float4*>(e.rec + idx);
    const float4 lo = r[0], hi = r[1];  // pdf[0..3], pdf[4..7]
    const float px = 0.0f < n.x ? lo.y : lo.z;
    const float py = 0.0f < n.y ? lo.w : hi.x;
    const float pz = 0.0f < n.z ? hi.y : hi.z;
    float p = 0.0f;
    p += ap.x * px;
    p += ap.y * py;
    p += ap.z * pz;
    return p;
}

// sample_envmap_6axis + sample_from_fragment (envmap_sampling.cl:126-157, 169-217)
__device__ F3 sample_env_6axis(const DevEnvmap& e, F3 n, uint4& st, float* pdf) {
    const uint32_t count = e.width * e.height;
    const uint32_t tx = 0.0f < n.x ? 1u : 2u, ty = 0.0f < n.y ? 3u : 4u, tz = 0.0f < n.z ? 5u : 6u;
    float u = xo_uniform(st);
    F3 ap = mul3v(n, n);
    uint32_t sel = (u < ap.x) ? tx : ((u < ap.x + ap.y) ? ty : tz);
    uint32_t idx = alias_sample(st, e.alias + (size_t)sel * count, count);
    *pdf = six_axis_mix(e, idx, n, ap);
    return sample_from_fragment(e.rec[idx].frag, st);
}
// envmap_pdf_sixAxis (envmap_sampling.cl:100-124)
__device__ float pdf_env_6axis(const DevEnvmap& e, F3 wi, F3 n) {
    return six_axis_mix(e, (uint32_t)envmap_index(e, wi), n, mul3v(n, n));
}
// sample_envmap_stage (envmap_sampling.cl:219-248): the uniform-weight table
__device__ F3 sample_env_uniform(const DevEnvmap& e, uint4& st, float* pdf) {
    const uint32_t count = e.width * e.height;
    uint32_t idx = alias_sample(st, e.alias, count);
    *pdf = e.rec[idx].pdf[0];
    return sample_from_fragment(e.rec[idx].frag, st);
}
__device__ __forceinline__ float pdf_env_uniform(const DevEnvmap& e, F3 wi) { return e.pdf0[envmap_index(e, wi)]; }

// sample_envmap_6axis_stage / sample_envmap_stage (kernels/envmap_sampling.cl:169-248) over a
// flat batch: one sample per RNG state (advanced in place); given normals (already facing the
// outgoing side), from the six-axis tables mixed by n^2, else from the uniform-weight table.
// out: direction and pdf per sample.
__global__ void __launch_bounds__(256) k_envmap_sample(DevEnvmap e, uint4* states, const float* normals, uint32_t n,
                                                       float4* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint4 st = states[i];
    float pdf;
    F3 wi;
    if (normals) {
        const F3 nn = F3{normals[3 * (size_t)i], normals[3 * (size_t)i + 1], normals[3 * (size_t)i + 2]};
        wi = sample_env_6axis(e, nn, st, &pdf);
    } else {
        wi = sample_env_uniform(e, st, &pdf);
    }
    states[i] = st;
    out[i] = make_float4(wi.x, wi.y, wi.z, pdf);
}

// evaluate_envmap_6axis_pdf_stage / evaluate_envmap_pdf_stage (kernels/envmap_sampling.cl:250-283)
// for given directions.
__global__ void __launch_bounds__(256) k_envmap_pdf(DevEnvmap e, const float* dirs, const float* normals, uint32_t n,
                                                    float* pdfs) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const F3 wi = F3{dirs[3 * (size_t)i], dirs[3 * (size_t)i + 1], dirs[3 * (size_t)i + 2]};
    if (normals) {
        const F3 nn = F3{normals[3 * (size_t)i], normals[3 * (size_t)i + 1], normals[3 * (size_t)i + 2]};
        pdfs[i] = pdf_env_6axis(e, wi, nn);
    } else {
        pdfs[i] = pdf_env_uniform(e, wi);
    }
}

// ---------------------------------------------------------------- BSDF helpers
// Duff et al. 2017 orthonormal basis, z up (kernels/lambertian.cl:24-32)
__device__ __forceinline__ void onb(F3 z, F3* x, F3* y) {
    const float sign = copysignf(1.0f, z.z);
    const float a = -1.0f / (sign + z.z);
    const float b = z.x * z.y * a;
    *x = F3{1.0f + sign * z.x * z.x * a, sign * b, -sign * z.x};
    *y = F3{b, sign + z.y * z.y * a, -z.y};
}
__device__ __forceinline__ F3 local_to_world(F3 x, F3 y, F3 z, F3 l) {
    return add3(add3(mul3(x, l.x), mul3(y, l.y)), mul3(z, l.z));
}
__device__ __forceinline__ F3 reflect3(F3 d, F3 n) { return sub3(d, mul3(n, 2.0f * dot3(d, n))); }
__device__ __forceinline__ F3 flip_toward(F3 n, F3 wo) { return dot3(n, wo) < 0.0f ? neg3(n) : n; }
__device__ __forceinline__ float sqr(float x) { return x * x; }

// Ward sampling / pdf / BRDF. alpha_x == alpha_y uses the reference's isotropic formulas
// (kernels/ward.cl:40-74, 155-176) verbatim; otherwise the anisotropic generalisation of
// the same (Walter 2005 half-vector sampling, Geisler-Moroder normalisation) in the ONB
// of the shading normal.
__device__ F3 sample_ward(const Ward& w, F3 wo, F3 z, float u0, float u1) {
    F3 x, y;
    float thetaH, phiH;
    const float pi = kPiF;
    if (w.alpha_x == w.alpha_y) {
        phiH = u0 * 2.0f * pi;
        thetaH = fatan(w.alpha_x * sqrtf(-flog(u1)));
    } else {
        float tp = 2.0f * pi * u0;
        phiH = fatan2(w.alpha_y * fsin(tp), w.alpha_x * fcos(tp));
        float c = fcos(phiH), s = fsin(phiH);
        float denom = c * c / (w.alpha_x * w.alpha_x) + s * s / (w.alpha_y * w.alpha_y);
        thetaH = fatan(sqrtf(-flog(u1) / denom));
    }
    float st = fsin(thetaH);
    F3 hl = F3{st * fcos(phiH), st * fsin(phiH), fcos(thetaH)};
    onb(z, &x, &y);
    F3 h = local_to_world(x, y, z, hl);
    return reflect3(neg3(wo), h);
}
__device__ float ward_exponent(const Ward& w, F3 h, F3 z) {
    // tan^2(theta_h) weighted by the two roughnesses
    if (w.alpha_x == w.alpha_y) {
        float c2 = sqr(dot3(h, z));
        float tan2 = (1.0f - c2) / c2;
        return -tan2 / sqr(w.alpha_x);
    }
    F3 x, y;
    onb(z, &x, &y);
    float hx = dot3(h, x) / w.alpha_x, hy = dot3(h, y) / w.alpha_y, hz = dot3(h, z);
    return -(hx * hx + hy * hy) / (hz * hz);
}
__device__ float pdf_ward(const Ward& w, F3 wo, F3 wi, F3 z) {
    if ((double)(dot3(wo, z) * dot3(wi, z)) < 0.0) return 0.0f;
    const float pi = kPiF;
    F3 h = normalize3(add3(wi, wo));
    float hz = dot3(h, z);
    float axy = w.alpha_x == w.alpha_y ? sqr(w.alpha_x) : w.alpha_x * w.alpha_y;
    float k0 = 1.0f / (4.0f * pi * axy * dot3(h, wi) * (hz * hz * hz));
    float k1 = fexp(ward_exponent(w, h, z));
    return k0 * k1;
}

// ---------------------------------------------------------------- per-slot records
__device__ __forceinline__ void load_ray_rec(const TraceRay* r, uint32_t s, F3* ro, float* tmax, F3* rd, uint32_t* pix) {
    const float4* q = reinterpret_cast<const float4*>(r + s);
    float4 a = q[0], b = q[1];
    *ro = F3{a.x, a.y, a.z};
    *tmax = a.w;
    *rd = F3{b.x, b.y, b.z};
    *pix = __float_as_uint(b.w);
}
__device__ __forceinline__ void store_ray_rec(TraceRay* r, uint32_t s, F3 ro, float tmax, F3 rd, uint32_t pix) {
    float4* q = reinterpret_cast<float4*>(r + s);
    q[0] = make_float4(ro.x, ro.y, ro.z, tmax);
    q[1] = make_float4(rd.x, rd.y, rd.z, __uint_as_float(pix));
}
__device__ __forceinline__ uint32_t ray_pixel(const TraceRay* r, uint32_t s) {
    return reinterpret_cast<const uint32_t*>(r + s)[7];
}
__device__ __forceinline__ void store_shade(ShadeRec* sh, uint32_t s, F3 Le, F3 T) {
    float4* q = reinterpret_cast<float4*>(sh + s);
    q[0] = make_float4(Le.x, Le.y, Le.z, T.x);
    q[1] = make_float4(T.y, T.z, 0.0f, 0.0f);
}

// Strategy selection (kernels/mixture_density.cl:25-67) for a surface hit; returns true for BSDF.
__device__ __forceinline__ bool select_bxdf(uint4& st, float sel_bxdf) {
    float u = xo_uniform(st);
    return u <= sel_bxdf;
}

// ---------------------------------------------------------------- where a slot's inputs come from
// OneJob: a single render, its scene, materials, envmap and camera passed as kernel arguments
// (the kernels compile to straight uses of those). ManyJobs: a batch, the slot's job found
// through its 256-slot block and its inputs read from the job table.
struct OneJob {
    TraceScene sc;
    DevMaterials m;
    DevEnvmap env;
    Camera cam;
    static constexpr bool kBatch = false;
    __device__ __forceinline__ uint32_t job(uint32_t) const { return 0; }
    __device__ __forceinline__ TraceScene scene(uint32_t) const { return sc; }
    __device__ __forceinline__ DevMaterials mats(uint32_t) const { return m; }
    __device__ __forceinline__ DevEnvmap envmap(uint32_t) const { return env; }
    __device__ __forceinline__ const Camera& camera(uint32_t) const { return cam; }
    __device__ __forceinline__ float* accum(uint32_t, const DevState& s) const { return s.accum; }
    __device__ __forceinline__ GuideFrame guide_frame(uint32_t, const DevGuide& g) const {
        return GuideFrame{{g.lo[0], g.lo[1], g.lo[2]}, {g.inv_ext[0], g.inv_ext[1], g.inv_ext[2]}, 0u};
    }
    DevJobs table() const { return DevJobs{}; }
};

struct ManyJobs {
    DevJobs d;
    static constexpr bool kBatch = true;
    __device__ __forceinline__ uint32_t job(uint32_t slot) const { return d.job_of_block[slot / kJobBlock]; }
    __device__ __forceinline__ TraceScene scene(uint32_t j) const { return d.jobs[j].sc; }
    __device__ __forceinline__ DevMaterials mats(uint32_t j) const { return d.jobs[j].m; }
    __device__ __forceinline__ DevEnvmap envmap(uint32_t j) const { return d.jobs[j].env; }
    __device__ __forceinline__ const Camera& camera(uint32_t j) const { return d.jobs[j].cam; }
    __device__ __forceinline__ float* accum(uint32_t j, const DevState&) const { return d.jobs[j].accum; }
    __device__ __forceinline__ GuideFrame guide_frame(uint32_t j, const DevGuide&) const { return d.jobs[j].gframe; }
    DevJobs table() const { return d; }
};

// ---------------------------------------------------------------- kernels
// Start state of every slot. A single render seeds slot i with seed_offset + i; a batch
// seeds the k-th slot of job j with that job's seed + k, as a single render of that job.
template <class J>
__global__ void __launch_bounds__(256) k_init(DevState s, WFParams p, J jobs, unsigned long long seed_offset,
                                              unsigned long long next_pixel) {
    uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid == 0) {
        *s.next_pixel = next_pixel;  // a batch keeps its cursors in the job table (this stays 0)
        for (int q = 0; q < kNumQueues; ++q) s.counts[q] = q == 0 ? p.n : 0u;
    }
    // every slot starts in the new-path queue: its per-block exclusive offsets
    if (gid < s.nblocks) s.block_counts[gid] = gid * 256u;
    if (gid >= p.n) return;
    // random_initialize (kernels/peseudo_random.cl:48-64)
    unsigned long long x;
    if constexpr (J::kBatch) {
        const JobDesc& jd = jobs.d.jobs[jobs.job(gid)];
        x = jd.seed + (gid - jd.first_block * kJobBlock);
    } else {
        x = seed_offset + gid;
    }
    unsigned long long r0 = splitmix64(&x);
    unsigned long long r1 = splitmix64(&x);
    uint4 st = make_uint4((uint32_t)r0, (uint32_t)(r0 >> 32), (uint32_t)r1, (uint32_t)(r1 >> 32));
    if ((st.x | st.y | st.z | st.w) == 0) st.x = 1;
    s.rng[gid] = st;
    // initialize_all_as_new_path (kernels/new_path.cl:9-15)
    s.items[gid] = gid;
    // defined start state (the reference leaves these buffers uninitialised)
    ExtRec e = {-1, -1, kFltMax, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    s.ext[gid] = e;
    VolList v = {};
    s.vol[gid] = v;
    PathRec pr = {};
    s.path[gid] = pr;
    ShadeRec sh = {};
    s.shade[gid] = sh;
}

__global__ void __launch_bounds__(256) k_clear_accum(DevState s, uint32_t npixels, uint32_t accum64) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npixels) return;
    if (accum64) {
        s.accum64[3 * (size_t)i + 0] = 0.0;
        s.accum64[3 * (size_t)i + 1] = 0.0;
        s.accum64[3 * (size_t)i + 2] = 0.0;
        s.accum_count[i] = 0;
    } else {
        reinterpret_cast<float4*>(s.accum)[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

// new_path (kernels/new_path.cl:17-83) over the new-path queue, plus the strategy
// selection and envmap draws the reference makes for a new path whose extension result
// still carries a volume id from its previous life (kernels/mixture_density.cl:26-55,
// kernels/envmap_sampling.cl:219-248).
// In a batch the sample index is the job's cursor plus the item's rank among the job's new
// paths (its queue position less the job's first position, the scanned offset of its first
// block); a job past its sample budget parks the slot (pixel kDeadPixel, in no queue).
template <class J>
__global__ void __launch_bounds__(256) k_new_path(DevState s, WFParams p, J jobs, DevGuide g) {
    const uint32_t count = s.counts[0];
    const unsigned long long base0 = J::kBatch ? 0ull : *s.next_pixel;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const uint32_t slot = s.items[i];
        const uint32_t jb = jobs.job(slot);
        const Camera& cam = jobs.camera(jb);
        const DevEnvmap env = jobs.envmap(jb);
        unsigned long long sample = base0 + i;
        if constexpr (J::kBatch) {
            const JobDesc& jd = jobs.d.jobs[jb];
            sample = jd.next_pixel + (i - s.block_counts[jd.first_block]);
            if (jd.budget != 0 && sample >= jd.budget) {
                store_ray_rec(s.ray, slot, F3{0.0f, 0.0f, 0.0f}, -1.0f, F3{0.0f, 0.0f, 1.0f}, kDeadPixel);
                continue;
            }
        }
        const unsigned long long P = (unsigned long long)cam.width * cam.height;
        uint4 st = s.rng[slot];
        uint32_t pixel = (uint32_t)(sample % P);
        uint32_t x = pixel % cam.width, y = pixel / cam.width;
        F3 o = load3(cam.o), r = load3(cam.r), b = load3(cam.b);
        float jx = (float)x + xo_uniform(st);
        float jy = (float)y + xo_uniform(st);
        F3 on_plane = add3(add3(o, mul3(r, jx)), mul3(b, jy));
        float u0 = xo_uniform(st);
        float u1 = xo_uniform(st);
        float rr = sqrtf(u0);
        float theta = two_pi_times(u1);
        float lensx = cam.lens_radius * rr * fcos(theta);
        float lensy = cam.lens_radius * rr * fsin(theta);
        F3 lens = add3(add3(load3(cam.eye), mul3(load3(cam.up), lensx)), mul3(load3(cam.right), lensy));
        F3 rd = normalize3(sub3(on_plane, lens));
        store_ray_rec(s.ray, slot, lens, kFltMax, rd, pixel);
        PathRec pr;
        pr.T[0] = pr.T[1] = pr.T[2] = 1.0f;
        pr.L[0] = pr.L[1] = pr.L[2] = 0.0f;
        pr.logic_i = 0;
        pr.pad = 0;
        s.path[slot] = pr;
        s.vol[slot].count = 0;
        if (g.train) g.rec_count[slot] = 0;
        store_shade(s.shade, slot, F3{0.0f, 0.0f, 0.0f}, F3{1.0f, 1.0f, 1.0f});
        ExtRec e = s.ext[slot];
        e.hit_surface = -1;
        e.tmin = kFltMax;
        e.ngx = e.ngy = e.ngz = 0.0f;
        s.ext[slot] = e;
        if (p.quirks && e.hit_volume >= 0) {
            bool bxdf = select_bxdf(st, p.bxdf_p);
            if (!bxdf) {
                float pdf;
                (void)sample_env_uniform(env, st, &pdf);
            }
        }
        s.rng[slot] = st;
    }
}

__device__ __forceinline__ void spawn(DevState& s, uint32_t slot, F3 ro, F3 rd, float t, F3 wi, F3 ng_offset,
                                      bool offset_normal, uint32_t pixel) {
    // ro + rd * tmin + wi * 1e-5 (+ Ng * 1e-5)  (kernels/lambertian.cl:117-119)
    F3 o = add3(add3(ro, mul3(rd, t)), mul3(wi, 1.0e-5f));
    if (offset_normal) o = add3(o, mul3(ng_offset, 1.0e-5f));
    store_ray_rec(s.ray, slot, o, kFltMax, wi, pixel);
}

// Guided one-sample MIS: the guide is chosen with probability g.select_p (or the vertex's
// learned selection probability), otherwise the reference's BSDF / envmap choice is made with
// the same uniform rescaled.
enum { kStratBxdf = 0, kStratEnv = 1, kStratGuide = 2 };
// logic_i counts the path's vertices from the camera (1 at the first hit).
__device__ __forceinline__ bool guided_vertex(const DevGuide& g, uint32_t logic_i) {
    return g.max_depth == 0 || logic_i <= g.max_depth;
}
__device__ __forceinline__ int select_strategy(uint4& st, float bxdf_p, bool guided, float guide_p) {
    float u = xo_uniform(st);
    if (guided) {
        if (u < guide_p) return kStratGuide;
        u = (u - guide_p) / (1.0f - guide_p);
    }
    return u <= bxdf_p ? kStratBxdf : kStratEnv;
}

// Remember a guided vertex until its path ends (throughput before it, radiance after its
// emission, and the density of the other strategies at wi, for the learned selection).
__device__ __forceinline__ void record_vertex(const DevGuide& g, const GuideFrame& fr, uint32_t slot, F3 hit, F3 n,
                                              F3 wo, F3 wi, const PathRec& pr, F3 Le, float other_pdf) {
    const uint32_t c = g.rec_count[slot];
    if (c >= (uint32_t)kGRecords) return;
    GuideVertex r;
    r.p[0] = (hit.x - fr.lo[0]) * fr.inv_ext[0];
    r.p[1] = (hit.y - fr.lo[1]) * fr.inv_ext[1];
    r.p[2] = (hit.z - fr.lo[2]) * fr.inv_ext[2];
    r.n[0] = n.x; r.n[1] = n.y; r.n[2] = n.z;
    r.wo[0] = wo.x; r.wo[1] = wo.y; r.wo[2] = wo.z;
    r.wi[0] = wi.x; r.wi[1] = wi.y; r.wi[2] = wi.z;
    const size_t at = (size_t)c * g.nslots + slot;
    g.rec_geo[at] = r;
    g.rec_T[at] = make_float4(pr.T[0], pr.T[1], pr.T[2], other_pdf);
    g.rec_L[at] = make_float4(pr.L[0] + pr.T[0] * Le.x, pr.L[1] + pr.T[1] * Le.y, pr.L[2] + pr.T[2] * Le.z, 0.0f);
    g.rec_count[slot] = c + 1;
}

// Lambertian: strategy selection + envmap 6-axis sample / cosine sample + pdfs +
// lambertian_stage (mixture_density.cl, envmap_sampling.cl:169-217,250-266, lambertian.cl:34-123).
template <bool kGuided, class J>
__global__ void __launch_bounds__(256) k_shade_lambertian(DevState s, WFParams p, J jobs, DevGuide g) {
    const uint32_t count = s.counts[kMatLambertian];
    const uint32_t* items = s.items + (size_t)kMatLambertian * p.n;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const uint32_t slot = items[i];
        const uint32_t jb = jobs.job(slot);
        const DevMaterials m = jobs.mats(jb);
        const DevEnvmap env = jobs.envmap(jb);
        uint4 st = s.rng[slot];
        ExtRec e = s.ext[slot];
        F3 ro, rd;
        float tmax;
        uint32_t pixel;
        load_ray_rec(s.ray, slot, &ro, &tmax, &rd, &pixel);
        const Lambertian L = m.lambertian[m.mat[e.hit_surface].index];
        const F3 Ng = F3{e.ngx, e.ngy, e.ngz};
        const F3 wo = neg3(rd);
        const bool guide_here = kGuided && guided_vertex(g, s.path[slot].logic_i);
        // the guide's records of this vertex: at its position in the lambertian queue
        const float gp = guide_here ? (g.learn_select ? g.sel[i] : g.select_p) : 0.0f;
        const float sel_b = (1.0f - gp) * p.bxdf_p, sel_e = (1.0f - gp) * (1.0f - p.bxdf_p);
        const int strat = guide_here ? select_strategy(st, p.bxdf_p, true, gp)
                                     : (select_bxdf(st, sel_b) ? kStratBxdf : kStratEnv);
        const bool bxdf = strat == kStratBxdf;
        F3 wi;
        float pdf_e = 0.0f, pdf_b;
        if (strat == kStratEnv) wi = sample_env_6axis(env, flip_toward(Ng, wo), st, &pdf_e);
        if (kGuided && strat == kStratGuide) {
            float u0 = xo_uniform(st), u1 = xo_uniform(st), u2 = xo_uniform(st);
            float w[3], L[kGSlotFloats];
            guide_load_lobes(g.lobes, g.nslots, i, L);
            guide_sample(L, u0, u1, u2, w);
            wi = F3{w[0], w[1], w[2]};
        }
        // The reference samples and evaluates the cosine lobe and the envmap pdf with the
        // unflipped normal (kernels/lambertian.cl:59-64, envmap_sampling.cl:263-265); the
        // consistent-normals option flips it toward wo like the shading does. A guided vertex
        // always flips: at a back face the reference's envmap pdf of a direction the envmap
        // did not sample comes from the opposite tables, which only the guide's samples (not
        // the cosine lobe's, which point into the surface) would expose as bias.
        const F3 Ns = (p.consistent || guide_here) ? flip_toward(Ng, wo) : Ng;
        float cosTheta;
        if (bxdf) {
            float u0 = xo_uniform(st);
            float u1 = xo_uniform(st);
            float r = sqrtf(u0);
            float th = two_pi_times(u1);
            F3 l = F3{r * fcos(th), r * fsin(th), sqrtf(fmaxf(1.0f - u0, 0.0f))};
            F3 x, y;
            onb(Ns, &x, &y);
            wi = local_to_world(x, y, Ns, l);
            cosTheta = l.z;
        } else {
            cosTheta = dot3(Ns, wi);
        }
        pdf_b = fmaxf(cosTheta, 0.0f) / kPiF;
        if (strat != kStratEnv) pdf_e = pdf_env_6axis(env, wi, Ns);
        const F3 Ngf = flip_toward(Ng, wo);
        float pdf = 0.0f;
        pdf += sel_b * pdf_b;
        pdf += sel_e * pdf_e;
        if (guide_here) {
            float L[kGSlotFloats];
            guide_load_lobes(g.lobes, g.nslots, i, L);
            pdf += gp * guide_pdf(L, wi.x, wi.y, wi.z);
        }
        float c = dot3(wi, Ngf);
        F3 T = F3{0.0f, 0.0f, 0.0f};
        if (0.0 < (double)c) {
            T.x = L.R[0] / kPiF * c / pdf;
            T.y = L.R[1] / kPiF * c / pdf;
            T.z = L.R[2] / kPiF * c / pdf;
        }
        store_shade(s.shade, slot, load3(L.Le), T);
        if (guide_here && g.train)
            record_vertex(g, jobs.guide_frame(jb, g), slot, add3(ro, mul3(rd, e.tmin)), Ngf, wo, wi, s.path[slot],
                          load3(L.Le), p.bxdf_p * pdf_b + (1.0f - p.bxdf_p) * pdf_e);
        spawn(s, slot, ro, rd, e.tmin, wi, Ngf, true, pixel);
        s.rng[slot] = st;
    }
}

// Ward (mixture_density.cl, envmap_sampling.cl, ward.cl:76-186)
template <bool kGuided, class J>
__global__ void __launch_bounds__(256) k_shade_ward(DevState s, WFParams p, J jobs, DevGuide g) {
    const uint32_t count = s.counts[kMatWard];
    const uint32_t* items = s.items + (size_t)kMatWard * p.n;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const uint32_t slot = items[i];
        const uint32_t jb = jobs.job(slot);
        const DevMaterials m = jobs.mats(jb);
        const DevEnvmap env = jobs.envmap(jb);
        uint4 st = s.rng[slot];
        ExtRec e = s.ext[slot];
        F3 ro, rd;
        float tmax;
        uint32_t pixel;
        load_ray_rec(s.ray, slot, &ro, &tmax, &rd, &pixel);
        const Ward w = m.ward[m.mat[e.hit_surface].index];
        const F3 Ng = F3{e.ngx, e.ngy, e.ngz};
        const F3 wo = neg3(rd);
        const F3 Ngf = flip_toward(Ng, wo);
        const bool guide_here = kGuided && guided_vertex(g, s.path[slot].logic_i);
        // the guide's records of this vertex: after the lambertian queue's, at its position
        const size_t at = (size_t)s.counts[kMatLambertian] + i;
        const float gp = guide_here ? (g.learn_select ? g.sel[at] : g.select_p) : 0.0f;
        const float sel_b = (1.0f - gp) * p.bxdf_p, sel_e = (1.0f - gp) * (1.0f - p.bxdf_p);
        const int strat = guide_here ? select_strategy(st, p.bxdf_p, true, gp)
                                     : (select_bxdf(st, sel_b) ? kStratBxdf : kStratEnv);
        const bool bxdf = strat == kStratBxdf;
        F3 wi;
        float pdf_e = 0.0f;
        if (strat == kStratEnv) wi = sample_env_6axis(env, Ngf, st, &pdf_e);
        if (kGuided && strat == kStratGuide) {
            float u0 = xo_uniform(st), u1 = xo_uniform(st), u2 = xo_uniform(st);
            float d[3], L[kGSlotFloats];
            guide_load_lobes(g.lobes, g.nslots, at, L);
            guide_sample(L, u0, u1, u2, d);
            wi = F3{d[0], d[1], d[2]};
        }
        if (bxdf) {
            float u0 = xo_uniform(st);
            float u1 = xo_uniform(st);
            wi = sample_ward(w, wo, Ngf, u0, u1);
        }
        float pdf_b = pdf_ward(w, wo, wi, Ngf);
        if (strat != kStratEnv) pdf_e = pdf_env_6axis(env, wi, (p.consistent || guide_here) ? Ngf : Ng);
        float pdf = 0.0f;
        pdf += sel_b * pdf_b;
        pdf += sel_e * pdf_e;
        if (guide_here) {
            float L[kGSlotFloats];
            guide_load_lobes(g.lobes, g.nslots, at, L);
            pdf += gp * guide_pdf(L, wi.x, wi.y, wi.z);
        }
        float c = dot3(wi, Ngf);
        F3 T = F3{0.0f, 0.0f, 0.0f};
        if (0.0f < c) {
            const float pi = kPiF;
            F3 lv = add3(wi, wo);
            F3 h = normalize3(lv);
            float mu = dot3(h, wi);
            float k = fpow(1.0f - mu, 1.0f / w.falloff);
            F3 rho = F3{w.reflectance[0] + (w.edgetint[0] - w.reflectance[0]) * k,
                        w.reflectance[1] + (w.edgetint[1] - w.reflectance[1]) * k,
                        w.reflectance[2] + (w.edgetint[2] - w.reflectance[2]) * k};
            float axy = w.alpha_x == w.alpha_y ? sqr(w.alpha_x) : w.alpha_x * w.alpha_y;
            F3 k0 = F3{rho.x / (pi * axy), rho.y / (pi * axy), rho.z / (pi * axy)};
            float k1 = fexp(ward_exponent(w, h, Ngf));
            float k2 = dot3(lv, lv) / sqr(sqr(dot3(lv, Ngf)));
            T.x = k0.x * k1 * k2 * c / pdf;
            T.y = k0.y * k1 * k2 * c / pdf;
            T.z = k0.z * k1 * k2 * c / pdf;
        }
        store_shade(s.shade, slot, F3{0.0f, 0.0f, 0.0f}, T);
        if (guide_here && g.train)
            record_vertex(g, jobs.guide_frame(jb, g), slot, add3(ro, mul3(rd, e.tmin)), Ngf, wo, wi, s.path[slot],
                          F3{0.0f, 0.0f, 0.0f}, p.bxdf_p * pdf_b + (1.0f - p.bxdf_p) * pdf_e);
        spawn(s, slot, ro, rd, e.tmin, wi, Ngf, true, pixel);
        s.rng[slot] = st;
    }
}

// Mirror (kernels/delta_materials.cl:48-77)
__global__ void __launch_bounds__(256) k_shade_specular(DevState s, WFParams p) {
    const uint32_t count = s.counts[kMatSpecular];
    const uint32_t* items = s.items + (size_t)kMatSpecular * p.n;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const uint32_t slot = items[i];
        uint4 st = s.rng[slot];
        ExtRec e = s.ext[slot];
        F3 ro, rd;
        float tmax;
        uint32_t pixel;
        load_ray_rec(s.ray, slot, &ro, &tmax, &rd, &pixel);
        (void)select_bxdf(st, 1.0f);  // strategy draw: selection (1, 0)
        const F3 Ngf = flip_toward(F3{e.ngx, e.ngy, e.ngz}, neg3(rd));
        F3 wi = reflect3(rd, Ngf);
        store_shade(s.shade, slot, F3{0.0f, 0.0f, 0.0f}, F3{1.0f, 1.0f, 1.0f});
        spawn(s, slot, ro, rd, e.tmin, wi, Ngf, false, pixel);
        s.rng[slot] = st;
    }
}

// Dielectric (kernels/delta_materials.cl:79-134)
template <class J>
__global__ void __launch_bounds__(256) k_shade_dielectric(DevState s, WFParams p, J jobs) {
    const uint32_t count = s.counts[kMatDielectric];
    const uint32_t* items = s.items + (size_t)kMatDielectric * p.n;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const uint32_t slot = items[i];
        const DevMaterials m = jobs.mats(jobs.job(slot));
        uint4 st = s.rng[slot];
        ExtRec e = s.ext[slot];
        F3 ro, rd;
        float tmax;
        uint32_t pixel;
        load_ray_rec(s.ray, slot, &ro, &tmax, &rd, &pixel);
        (void)select_bxdf(st, 1.0f);
        const Dielectric d = m.dielectric[m.mat[e.hit_surface].index];
        F3 Ng = F3{e.ngx, e.ngy, e.ngz};
        const F3 wo = neg3(rd);
        const bool backside = dot3(Ng, wo) < 0.0f;
        float eta_i = 1.0f, eta_t = d.ior;
        if (backside) {
            float t = eta_i;
            eta_i = eta_t;
            eta_t = t;
            Ng = neg3(Ng);
        }
        // fresnel_dielectrics(|cos|, eta_t, eta_i)
        float cth = fabsf(dot3(Ng, wo));
        float g = sqrtf(eta_t * eta_t / sqr(eta_i) - 1.0f + sqr(cth));
        float fa = 0.5f * sqr(g - cth) / sqr(g + cth);
        float fb = 1.0f + sqr(cth * (g + cth) - 1.0f) / sqr(cth * (g - cth) + 1.0f);
        float f = fa * fb;
        float u = xo_uniform(st);
        float comp = 1.0f;
        F3 wi;
        if (u < f) {
            wi = reflect3(rd, Ng);
        } else {
            float eta = eta_i / eta_t;
            float NoD = dot3(Ng, rd);
            float k = 1.0f - eta * eta * (1.0f - NoD * NoD);
            if (k <= 0.0f) {
                wi = reflect3(rd, Ng);
            } else {
                wi = sub3(mul3(rd, eta), mul3(Ng, eta * NoD + sqrtf(k)));
                comp = sqr(eta_t / eta_i);
            }
        }
        store_shade(s.shade, slot, F3{0.0f, 0.0f, 0.0f}, F3{comp, comp, comp});
        spawn(s, slot, ro, rd, e.tmin, wi, Ng, false, pixel);
        s.rng[slot] = st;
    }
}

__device__ __forceinline__ uint32_t volume_hash(const Volume& v) {
    return __float_as_uint(v.C) ^ __float_as_uint(v.R[0]) ^ __float_as_uint(v.R[1]) ^ __float_as_uint(v.R[2]);
}

// Medium boundary (kernels/homogeneous_volume.cl:19-64, InVolumeList ops types.cl:108-132)
template <class J>
__global__ void __launch_bounds__(256) k_shade_volume_surface(DevState s, WFParams p, J jobs) {
    const uint32_t count = s.counts[kMatVolume];
    const uint32_t* items = s.items + (size_t)kMatVolume * p.n;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const uint32_t slot = items[i];
        const DevMaterials m = jobs.mats(jobs.job(slot));
        uint4 st = s.rng[slot];
        ExtRec e = s.ext[slot];
        F3 ro, rd;
        float tmax;
        uint32_t pixel;
        load_ray_rec(s.ray, slot, &ro, &tmax, &rd, &pixel);
        (void)select_bxdf(st, 1.0f);
        const bool backside = dot3(F3{e.ngx, e.ngy, e.ngz}, neg3(rd)) < 0.0f;
        store_shade(s.shade, slot, F3{0.0f, 0.0f, 0.0f}, F3{1.0f, 1.0f, 1.0f});
        VolList v = s.vol[slot];
        if (backside) {
            uint32_t hsh = volume_hash(m.volume[m.mat[e.hit_surface].index]);
            bool move = false;
            for (int k = 0; k < v.count; ++k) {
                if (move) v.volumes[k - 1] = v.volumes[k];
                else if (volume_hash(m.volume[m.mat[v.volumes[k]].index]) == hsh) move = true;
            }
            if (move) v.count--;
        } else if (v.count < 7) {
            v.volumes[v.count] = e.hit_surface;
            v.count++;
        }
        s.vol[slot] = v;
        F3 o = add3(ro, mul3(rd, e.tmin + 1.0e-5f));
        store_ray_rec(s.ray, slot, o, kFltMax, rd, pixel);
        s.rng[slot] = st;
    }
}

// Scattering inside a medium (mixture_density.cl:47-51, envmap_sampling.cl:219-248,268-281,
// homogeneous_volume.cl:66-129)
template <class J>
__global__ void __launch_bounds__(256) k_shade_volume_inside(DevState s, WFParams p, J jobs) {
    const uint32_t count = s.counts[kMatVolumeInside];
    const uint32_t* items = s.items + (size_t)kMatVolumeInside * p.n;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const uint32_t slot = items[i];
        const uint32_t jb = jobs.job(slot);
        const DevMaterials m = jobs.mats(jb);
        const DevEnvmap env = jobs.envmap(jb);
        uint4 st = s.rng[slot];
        ExtRec e = s.ext[slot];
        F3 ro, rd;
        float tmax;
        uint32_t pixel;
        load_ray_rec(s.ray, slot, &ro, &tmax, &rd, &pixel);
        const float sel_b = p.bxdf_p, sel_e = 1.0f - p.bxdf_p;
        const bool bxdf = select_bxdf(st, sel_b);
        F3 wi;
        float pdf_e = 0.0f;
        if (!bxdf) wi = sample_env_uniform(env, st, &pdf_e);
        if (bxdf) {
            float u0 = xo_uniform(st);
            float u1 = xo_uniform(st);
            float phi = two_pi_times(u0);
            float z = -1.0f + (1.0f - -1.0f) * u1;
            float r = sqrtf(fmaxf(1.0f - z * z, 0.0f));
            wi = F3{r * fcos(phi), r * fsin(phi), z};
        }
        const float pdf_b = (float)(1.0 / (kPiD * 4.0));
        if (bxdf) pdf_e = pdf_env_uniform(env, wi);
        float pdf = 0.0f;
        pdf += sel_b * pdf_b;
        pdf += sel_e * pdf_e;
        const Volume v = m.volume[m.mat[e.hit_volume].index];
        const float four_pi = (float)(4.0f * kPiD);
        F3 T = F3{v.R[0] / four_pi / pdf, v.R[1] / four_pi / pdf, v.R[2] / four_pi / pdf};
        store_shade(s.shade, slot, F3{0.0f, 0.0f, 0.0f}, T);
        store_ray_rec(s.ray, slot, add3(ro, mul3(rd, e.tmin)), kFltMax, wi, pixel);
        s.rng[slot] = st;
    }
}

// finalize_new_path (kernels/new_path.cl:85-88): advance the sample cursor(s) by the new paths
// started this bounce (a job's share: the span of its run in queue 0). Thread gid of the grid.
template <class J>
__device__ __forceinline__ void advance_cursors(const DevState& s, const J& jobs, uint32_t gid) {
    if constexpr (!J::kBatch) {
        if (gid == 0) *s.next_pixel += s.counts[0];
    } else if (gid < jobs.d.njobs) {
        JobDesc& jd = jobs.d.jobs[gid];
        const uint32_t lo = s.block_counts[jd.first_block];
        const uint32_t hi = gid + 1 < jobs.d.njobs ? s.block_counts[jobs.d.jobs[gid + 1].first_block] : s.counts[0];
        jd.next_pixel += hi - lo;
    }
}

// homogeneous_volume_through: the medium the ray is in (drawn from the slot's list) and its
// free-flight distance, which becomes the extension ray's far clip.
__device__ __forceinline__ void volume_through(const DevState& s, const DevMaterials& m, uint32_t slot,
                                               int32_t* hit_volume, float* tmin) {
    uint4 st = s.rng[slot];
    VolList v = s.vol[slot];
    *hit_volume = -1;
    if (0 < v.count) {
        unsigned long long u = xo_uint62(st);
        // the drawn entry by selects: an indexed read of the register copy would go through
        // scratch memory
        const uint32_t k = (uint32_t)(u % (unsigned long long)v.count);
        int32_t h = v.volumes[0];
#pragma unroll
        for (uint32_t j = 1; j < 7; ++j) h = k == j ? v.volumes[j] : h;
        *hit_volume = h;
    }
    *tmin = kFltMax;
    if (0 <= *hit_volume) {
        const Volume vol = m.volume[m.mat[*hit_volume].index];
        float u = xo_uniform(st);
        *tmin = -1.0f / vol.C * flog(u);
    }
    s.rng[slot] = st;
}

// Extension result of a slot: the surface hit (with its geometric normal) or, failing that,
// the medium event at the free-flight distance.
__device__ __forceinline__ void store_extension(const DevState& s, const TraceScene& sc, uint32_t slot, F3 ro, F3 rd,
                                                int32_t prim, float t, int32_t hit_volume, float tvol) {
    ExtRec e = s.ext[slot];
    e.hit_surface = -1;
    e.hit_volume = hit_volume;
    e.tmin = tvol;
    if (prim >= 0) {
        F3 n;
        if (prim < sc.sphere_begin) {
            const uint32_t* ix = sc.indices + 3 * (size_t)prim;
            float4 a = sc.points[ix[0]], b = sc.points[ix[1]], c = sc.points[ix[2]];
            F3 v0 = F3{a.x, a.y, a.z}, v1 = F3{b.x, b.y, b.z}, v2 = F3{c.x, c.y, c.z};
            n = cross3(sub3(v2, v1), sub3(v1, v0));  // clockwise winding
        } else {
            float4 sp = sc.spheres[prim - sc.sphere_begin];
            n = sub3(add3(ro, mul3(rd, t)), F3{sp.x, sp.y, sp.z});
        }
        n = normalize3(n);
        e.hit_surface = prim;
        e.hit_volume = -1;
        e.tmin = t;
        e.ngx = n.x;
        e.ngy = n.y;
        e.ngz = n.z;
    }
    s.ext[slot] = e;
}

// homogeneous_volume_through (kernels/homogeneous_volume_through.cl:7-45) fused with
// extension_ray_cast (kernels/extension_ray_cast_stackless.cl:202-246) over every slot, one
// thread per slot.
template <int kVariant, class J>
__global__ void __launch_bounds__(256) k_trace_step(DevState s, WFParams p, J jobs, uint32_t top) {
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    if constexpr (kVariant == kVariantStackless) {
        // the top `top` records of the block's scene in LDS (a job owns whole 256-slot blocks,
        // so every slot of a workgroup traces the same scene)
        const uint32_t first = blockIdx.x * blockDim.x;
        const TraceScene sb = jobs.scene(jobs.job(first < p.n ? first : p.n - 1));
        const uint32_t t = top < sb.nnodes ? top : sb.nnodes;
        const float4* g = reinterpret_cast<const float4*>(sb.sl_nodes);
        for (uint32_t j = threadIdx.x; j < 4 * t; j += blockDim.x) sl_treelet[j] = g[j];
        __syncthreads();
        top = t;
    }
    advance_cursors(s, jobs, slot);
    if (slot >= p.n) return;
    if (J::kBatch && ray_pixel(s.ray, slot) == kDeadPixel) return;
    const uint32_t jb = jobs.job(slot);
    const TraceScene sc = jobs.scene(jb);
    int32_t hit_volume;
    float tmin;
    volume_through(s, jobs.mats(jb), slot, &hit_volume, &tmin);
    F3 ro, rd;
    float tmax;
    uint32_t pixel;
    load_ray_rec(s.ray, slot, &ro, &tmax, &rd, &pixel);
    float t = tmin, u = 0.0f, vv = 0.0f;
    int32_t prim = -1;
    if (sc.nnodes != 0) prim = trace_batched<kVariant, true>(sc, ro, rd, &t, &u, &vv, kPerRayLeafBatch, top);
    store_extension(s, sc, slot, ro, rd, prim, t, hit_volume, tmin);
}

// The same stage for a single render as three launches: the medium draw and far clip of every
// slot written as a flat ray (the medium id in the spare word), the library's persistent
// ray-cast kernel over those rays (leaf batching included), and the extension records from the
// hits. Each slot's traversal and results are those of k_trace_step.
template <class J>
__global__ void __launch_bounds__(256) k_trace_setup(DevState s, WFParams p, J jobs) {
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    advance_cursors(s, jobs, slot);
    if (slot >= p.n) return;
    int32_t hit_volume;
    float tvol;
    volume_through(s, jobs.mats(0), slot, &hit_volume, &tvol);
    F3 ro, rd;
    float tmax;
    uint32_t pixel;
    load_ray_rec(s.ray, slot, &ro, &tmax, &rd, &pixel);
    float4* x = reinterpret_cast<float4*>(s.xray + slot);
    x[0] = make_float4(ro.x, ro.y, ro.z, tvol);
    x[1] = make_float4(rd.x, rd.y, rd.z, __int_as_float(hit_volume));
}

template <class J>
__global__ void __launch_bounds__(256) k_trace_finish(DevState s, WFParams p, J jobs) {
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= p.n) return;
    const float4* x = reinterpret_cast<const float4*>(s.xray + slot);
    const float4 a = x[0], b = x[1];
    const TraceHit h = s.xhit[slot];
    store_extension(s, jobs.scene(0), slot, F3{a.x, a.y, a.z}, F3{b.x, b.y, b.z}, h.prim, h.t,
                    __float_as_int(b.w), a.w);
}

// logic (kernels/logic.cl:117-280): path update, Russian roulette, envmap on a miss,
// accumulation of finished paths, and the per-block queue counts of a stable compaction.
template <class J>
__global__ void __launch_bounds__(256) k_logic(DevState s, WFParams p, J jobs, DevGuide g) {
    __shared__ uint32_t wcount[kNumQueues][4];
    __shared__ uint32_t wsamples[4];
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t q = 0xFFu;
    uint32_t rec_mask = 0;  // training samples this slot emits (recorded vertices of a finished path)
    float rec_w[kGRecords];
    bool live = slot < p.n;
    if (J::kBatch && live && ray_pixel(s.ray, slot) == kDeadPixel) {
        live = false;
        s.qsel[slot] = 0xFFu;  // parked: in no queue
    }
    if (live) {
        const uint32_t jb = jobs.job(slot);
        const DevMaterials m = jobs.mats(jb);
        const DevEnvmap env = jobs.envmap(jb);
        PathRec pr = s.path[slot];
        const uint4 sh0 = reinterpret_cast<const uint4*>(s.shade + slot)[0];
        const uint4 sh1 = reinterpret_cast<const uint4*>(s.shade + slot)[1];
        const F3 Le = F3{__uint_as_float(sh0.x), __uint_as_float(sh0.y), __uint_as_float(sh0.z)};
        const F3 Ts = F3{__uint_as_float(sh0.w), __uint_as_float(sh1.x), __uint_as_float(sh1.y)};
        const ExtRec e = s.ext[slot];
        uint4 st = s.rng[slot];
        const uint32_t logic_i = pr.logic_i;
        F3 L = add3(load3(pr.L), mul3v(load3(pr.T), Le));
        F3 T = mul3v(load3(pr.T), Ts);
        bool newPath = false, evalEnv = false;
        if (e.hit_surface < 0 && e.hit_volume < 0) newPath = evalEnv = true;
        float lumT = 0.2126f * T.x + 0.7152f * T.y + 0.0722f * T.z;
        float cp = 1.0f;
        if (p.rr_start < logic_i) cp = fminf(lumT, 1.0f);
        float u = xo_uniform(st);
        if (u < cp) {
            T.x = T.x / cp;
            T.y = T.y / cp;
            T.z = T.z / cp;
        } else {
            newPath = true;
        }
        if (lumT < 1.0e-5f) newPath = true;
        if (p.max_depth != 0 && logic_i >= p.max_depth) newPath = true;
        F3 ro, rd;
        float tmax;
        uint32_t pixel;
        load_ray_rec(s.ray, slot, &ro, &tmax, &rd, &pixel);
        if (evalEnv) {
            F3 em = envmap_value(env, rd, logic_i == 0);
            L = add3(L, mul3v(T, em));
        }
        const bool accept = newPath && finite3(L) && L.x < p.clamp && L.y < p.clamp && L.z < p.clamp;
        if (g.train && newPath) {
            // a finished path turns its recorded guided vertices into training samples:
            // weight = luminance((L_final - L_after) / T_before) = F / p of the sampled direction
            const uint32_t nrec = accept ? g.rec_count[slot] : 0u;
#pragma unroll
            for (uint32_t j = 0; j < (uint32_t)kGRecords; ++j) {
                if (j >= nrec) break;
                const float4 rT = g.rec_T[(size_t)j * g.nslots + slot];
                const float4 rL = g.rec_L[(size_t)j * g.nslots + slot];
                const float Tr[3] = {rT.x, rT.y, rT.z}, Lr[3] = {rL.x, rL.y, rL.z};
                float c[3];
                const float Lf[3] = {L.x, L.y, L.z};
                for (int k = 0; k < 3; ++k) c[k] = Tr[k] > 0.0f ? (Lf[k] - Lr[k]) / Tr[k] : 0.0f;
                const float wgt = 0.2126f * c[0] + 0.7152f * c[1] + 0.0722f * c[2];
                if (wgt > 0.0f && isfinite(wgt)) {
                    rec_w[j] = wgt;
                    rec_mask |= 1u << j;
                }
            }
            g.rec_count[slot] = 0;
        }
        if (accept) {
            if (p.accum64) {
                unsafeAtomicAdd(s.accum64 + 3 * (size_t)pixel + 0, (double)L.x);
                unsafeAtomicAdd(s.accum64 + 3 * (size_t)pixel + 1, (double)L.y);
                unsafeAtomicAdd(s.accum64 + 3 * (size_t)pixel + 2, (double)L.z);
                atomicAdd(s.accum_count + pixel, 1u);
            } else {
                float* a = jobs.accum(jb, s) + 4 * (size_t)pixel;
                unsafeAtomicAdd(a + 0, L.x);
                unsafeAtomicAdd(a + 1, L.y);
                unsafeAtomicAdd(a + 2, L.z);
                atomicAdd(reinterpret_cast<uint32_t*>(a + 3), 1u);
            }
        }
        pr.L[0] = L.x;
        pr.L[1] = L.y;
        pr.L[2] = L.z;
        pr.T[0] = T.x;
        pr.T[1] = T.y;
        pr.T[2] = T.z;
        pr.logic_i = logic_i + 1;
        s.path[slot] = pr;
        s.rng[slot] = st;
        if (newPath) q = 0;
        else if (0 <= e.hit_surface) q = (uint32_t)m.mat[e.hit_surface].type;
        else q = kMatVolumeInside;
        s.qsel[slot] = (uint8_t)q;
    }
    if (g.train) {
        // training samples leave through the queue scan: the per-block count goes to row
        // kNumQueues of block_counts, the slot's sample mask to qsel bits 3..6 and the
        // weights to emit_w; k_scatter_queues writes the samples in slot order
        if (rec_mask != 0) {
            g.emit_w[slot] = make_float4(rec_w[0], rec_w[1], rec_w[2], rec_w[3]);
            s.qsel[slot] = (uint8_t)(q | (rec_mask << 3));
        }
        const uint32_t cnt = (uint32_t)__popc(rec_mask);
        uint32_t total = 0;
        for (int b = 0; b < 3; ++b) total += (uint32_t)__popcll(__ballot((cnt >> b) & 1u)) << b;
        if (lane == 0) wsamples[wave] = total;
    }
    for (int k = 0; k < kNumQueues; ++k) {
        unsigned long long b = __ballot(q == (uint32_t)k);
        if (lane == 0) wcount[k][wave] = (uint32_t)__popcll(b);
    }
    __syncthreads();
    if (threadIdx.x < kNumQueues) {
        const int k = threadIdx.x;
        s.block_counts[(size_t)k * s.nblocks + blockIdx.x] = wcount[k][0] + wcount[k][1] + wcount[k][2] + wcount[k][3];
    }
    if (g.train && threadIdx.x == kNumQueues)
        s.block_counts[(size_t)kNumQueues * s.nblocks + blockIdx.x] =
            wsamples[0] + wsamples[1] + wsamples[2] + wsamples[3];
}

// Exclusive scan of the per-block counts of one queue (one workgroup per queue). Workgroup
// kNumQueues scans the training-sample counts on top of the samples already buffered and
// advances the buffer cursor (saturating at the buffer's capacity).
__global__ void __launch_bounds__(1024) k_scan_queues(DevState s, uint32_t* sample_count, uint32_t sample_cap) {
    // One workgroup per queue (and one for the training samples). Each of its 16 waves owns a
    // contiguous run of the queue's per-block counts and reads it 64 at a time, coalesced: the
    // runs' totals first, scanned across the waves in LDS, then the exclusive prefixes written
    // in place, each 64-count piece scanned across the wave.
    __shared__ uint32_t wtot[16];
    const int q = blockIdx.x;
    if (q == kNumQueues && !sample_count) return;
    uint32_t* c = s.block_counts + (size_t)q * s.nblocks;
    const uint32_t nb = s.nblocks;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t run = (nb + 15) / 16;  // counts per wave
    const uint32_t beg = min(wave * run, nb), end = min(beg + run, nb);
    uint32_t sum = 0;
    for (uint32_t i = beg + lane; i < end; i += 64) sum += c[i];
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
    if (lane == 0) wtot[wave] = sum;
    __syncthreads();
    const uint32_t base = q == kNumQueues ? *sample_count : 0u;
    uint32_t carry = base;  // exclusive prefix of this wave's run
    for (uint32_t w = 0; w < wave; ++w) carry += wtot[w];
    for (uint32_t i0 = beg; i0 < end; i0 += 64) {
        const uint32_t i = i0 + lane;
        const uint32_t v = i < end ? c[i] : 0u;
        uint32_t incl = v;  // inclusive scan of the piece across the wave
        for (uint32_t off = 1; off < 64; off <<= 1) {
            const uint32_t o = __shfl_up(incl, off, 64);
            if (lane >= off) incl += o;
        }
        if (i < end) c[i] = carry + incl - v;
        carry += __shfl(incl, 63, 64);
    }
    if (threadIdx.x == 1023) {
        uint32_t total = 0;
        for (int w = 0; w < 16; ++w) total += wtot[w];
        if (q == kNumQueues) *sample_count = min(base + total, sample_cap);
        else s.counts[q] = total;
    }
}

// Scatter slots into their queues in slot order; with training on, also write the finished
// paths' training samples (slot order) into the guide's buffer.
__global__ void __launch_bounds__(256) k_scatter_queues(DevState s, WFParams p, DevGuide g, DevJobs jobs) {
    __shared__ uint32_t wpre[kNumQueues][4];
    __shared__ uint32_t wsam[4];
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t qs = slot < p.n ? (uint32_t)s.qsel[slot] : 0xFFu;
    const uint32_t q = qs == 0xFFu ? 0xFFu : (qs & 7u);
    const uint32_t mask = (qs == 0xFFu || !g.train) ? 0u : (qs >> 3);
    const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint32_t rank = 0;
    for (int k = 0; k < kNumQueues; ++k) {
        unsigned long long b = __ballot(q == (uint32_t)k);
        if (lane == 0) wpre[k][wave] = (uint32_t)__popcll(b);
        if (q == (uint32_t)k) rank = (uint32_t)__popcll(b & lt);
    }
    uint32_t srank = 0;
    if (g.train) {
        const uint32_t cnt = (uint32_t)__popc(mask);
        uint32_t total = 0;
        for (int b = 0; b < 3; ++b) {
            const unsigned long long bb = __ballot((cnt >> b) & 1u);
            srank += (uint32_t)__popcll(bb & lt) << b;
            total += (uint32_t)__popcll(bb) << b;
        }
        if (lane == 0) wsam[wave] = total;
    }
    __syncthreads();
    if (q < (uint32_t)kNumQueues) {
        uint32_t off = s.block_counts[(size_t)q * s.nblocks + blockIdx.x];
        for (uint32_t w = 0; w < wave; ++w) off += wpre[q][w];
        s.items[(size_t)q * p.n + off + rank] = slot;
    }
    if (mask != 0) {
        uint32_t idx = s.block_counts[(size_t)kNumQueues * s.nblocks + blockIdx.x] + srank;
        for (uint32_t w = 0; w < wave; ++w) idx += wsam[w];
        const float4 wts = g.emit_w[slot];
        const float wj[kGRecords] = {wts.x, wts.y, wts.z, wts.w};
        const float cls = jobs.njobs ? (float)jobs.jobs[jobs.job_of_block[slot / kJobBlock]].gframe.cls : 0.0f;
#pragma unroll
        for (int j = 0; j < kGRecords; ++j) {
            if (!((mask >> j) & 1u)) continue;
            if (idx < g.capacity) {
                // 48-byte vertex -> 64-byte sample (p n wo wi | weight), 16-byte moves
                const float4* r = reinterpret_cast<const float4*>(g.rec_geo + (size_t)j * g.nslots + slot);
                float4* o = reinterpret_cast<float4*>(g.samples + idx);
                o[0] = r[0];
                o[1] = r[1];
                o[2] = r[2];
                o[3] = make_float4(wj[j], cls, g.rec_T[(size_t)j * g.nslots + slot].w, 0.0f);
            }
            ++idx;
        }
    }
}

// Mean radiance -> RGBA8: accumlation_to_intermediate (half-float means, accumlation.cl:6-25)
// followed by tonemap (pow(x, 1/2.2) * 256, accumlation.cl:68-79). Counts stay 32-bit.
__global__ void __launch_bounds__(256) k_tonemap(DevState s, uint32_t npixels, uint32_t accum64, uchar4* out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npixels) return;
    float r, g, b, cnt;
    if (accum64) {
        cnt = (float)s.accum_count[i];
        r = (float)s.accum64[3 * (size_t)i + 0];
        g = (float)s.accum64[3 * (size_t)i + 1];
        b = (float)s.accum64[3 * (size_t)i + 2];
    } else {
        float4 a = reinterpret_cast<const float4*>(s.accum)[i];
        r = a.x;
        g = a.y;
        b = a.z;
        cnt = (float)__float_as_uint(a.w);
    }
    float div = 1.0f / cnt;
    if (!isfinite(div)) div = 0.0f;
    float m[3] = {(float)(__half)(r * div), (float)(__half)(g * div), (float)(__half)(b * div)};
    unsigned char c[3];
    for (int k = 0; k < 3; ++k) {
        int v = (int)(fpow(m[k], 1.0f / 2.2f) * 256.0f);
        c[k] = (unsigned char)min(max(v, 0), 255);
    }
    out[i] = make_uchar4(c[0], c[1], c[2], 255);
}

// accumlation_to_intermediate (kernels/accumlation.cl:6-25) without its mutex: per pixel the
// half-float mean and the sample count, widened to 32 bits.
// held (optional): the kernel does nothing unless *held is set (copy_if_locked, kernels/mutex.cl).
__global__ void __launch_bounds__(256) k_to_intermediate(DevState s, uint32_t npixels, uint32_t accum64,
                                                         RGB16* __restrict__ out, const int* held) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npixels) return;
    if (held && *held == 0) return;
    float r, g, b;
    uint32_t cnt;
    if (accum64) {
        cnt = s.accum_count[i];
        r = (float)s.accum64[3 * (size_t)i + 0];
        g = (float)s.accum64[3 * (size_t)i + 1];
        b = (float)s.accum64[3 * (size_t)i + 2];
    } else {
        const float4 a = reinterpret_cast<const float4*>(s.accum)[i];
        r = a.x;
        g = a.y;
        b = a.z;
        cnt = __float_as_uint(a.w);
    }
    float div = 1.0f / (float)cnt;
    if (!isfinite(div)) div = 0.0f;
    RGB16 o;
    o.r = __half_as_ushort(__float2half_rn(r * div));
    o.g = __half_as_ushort(__float2half_rn(g * div));
    o.b = __half_as_ushort(__float2half_rn(b * div));
    o.pad = 0;
    o.count = cnt;
    out[i] = o;
}

// merge_intermediate (kernels/accumlation.cl:27-50): count-weighted mean of two intermediates
// into a, counts 32-bit.
__global__ void __launch_bounds__(256) k_merge_intermediate(RGB16* __restrict__ a, const RGB16* __restrict__ b,
                                                            uint32_t npixels) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npixels) return;
    RGB16 x = a[i];
    const RGB16 y = b[i];
    const uint32_t sab = x.count + y.count;
    if (sab == 0) {
        a[i] = RGB16{0, 0, 0, 0, 0u};
        return;
    }
    const float one_over = 1.0f / (float)sab, sa = (float)x.count, sb = (float)y.count;
    const uint16_t* xa = &x.r;
    const uint16_t* ya = &y.r;
    uint16_t m[3];
    for (int k = 0; k < 3; ++k) {
        const float v = (__half2float(__ushort_as_half(xa[k])) * sa + __half2float(__ushort_as_half(ya[k])) * sb) * one_over;
        m[k] = __half_as_ushort(__float2half_rn(v));
    }
    a[i] = RGB16{m[0], m[1], m[2], 0, sab};
}

// tonemap (kernels/accumlation.cl:68-79) of an intermediate: pow(x, 1/2.2) * 256, clamped.
__global__ void __launch_bounds__(256) k_tonemap_intermediate(const RGB16* __restrict__ in, uint32_t npixels,
                                                              uchar4* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npixels) return;
    const RGB16 x = in[i];
    const float m[3] = {__half2float(__ushort_as_half(x.r)), __half2float(__ushort_as_half(x.g)),
                        __half2float(__ushort_as_half(x.b))};
    unsigned char c[3];
    for (int k = 0; k < 3; ++k) {
        int v = (int)(fpow(m[k], 1.0f / 2.2f) * 256.0f);
        c[k] = (unsigned char)min(max(v, 0), 255);
    }
    out[i] = make_uchar4(c[0], c[1], c[2], 255);
}

// stat (kernels/stat.cl): total number of accumulated samples.
__global__ void __launch_bounds__(256) k_stat(DevState s, uint32_t npixels, uint32_t accum64,
                                              unsigned long long* total) {
    __shared__ unsigned long long part[4];
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long v = 0;
    if (i < npixels) v = accum64 ? s.accum_count[i] : __float_as_uint(s.accum[4 * (size_t)i + 3]);
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(total, part[0] + part[1] + part[2] + part[3]);
}

}  // namespace

hipError_t envmap_sample(const DevEnvmap& e, uint32_t* states, const float* normals, uint32_t n, float* out,
                         hipStream_t s) {
    if (n == 0) return hipSuccess;
    const dim3 grid((uint32_t)(((uint64_t)n + 255) / 256));
    hipLaunchKernelGGL(k_envmap_sample, grid, dim3(256), 0, s, e, reinterpret_cast<uint4*>(states),
                       normals, n, reinterpret_cast<float4*>(out));
    return hipGetLastError();
}

hipError_t envmap_pdf(const DevEnvmap& e, const float* dirs, const float* normals, uint32_t n, float* pdfs,
                      hipStream_t s) {
    if (n == 0) return hipSuccess;
    const dim3 grid((uint32_t)(((uint64_t)n + 255) / 256));
    hipLaunchKernelGGL(k_envmap_pdf, grid, dim3(256), 0, s, e, dirs, normals, n, pdfs);
    return hipGetLastError();
}

hipError_t wf_tonemap(const DevState& st, uint32_t npixels, bool accum64, void* rgba8, hipStream_t s) {
    dim3 grid((npixels + 255) / 256);
    hipLaunchKernelGGL(k_tonemap, grid, dim3(256), 0, s, st, npixels, accum64 ? 1u : 0u, (uchar4*)rgba8);
    return hipGetLastError();
}

hipError_t wf_intermediate(const DevState& st, uint32_t npixels, bool accum64, RGB16* out, hipStream_t s) {
    if (npixels == 0) return hipSuccess;
    hipLaunchKernelGGL(k_to_intermediate, dim3((npixels + 255) / 256), dim3(256), 0, s, st, npixels, accum64 ? 1u : 0u,
                       out, (const int*)nullptr);
    return hipGetLastError();
}

// ---------------------------------------------------------------- preview handoff
// The weak device mutex of kernels/mutex.cl: word 1 = free. A side takes it with one
// compare-and-swap (1 -> 0) and records in its own flag whether it did; it never waits, so the
// step stream and the preview stream never block each other. Work guarded by the flag runs in
// later kernels of the same stream (kernel boundaries order the data), and the holder frees the
// word in a final kernel.
__global__ void k_try_lock(int* mutex, int* held) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *held = atomicCAS(mutex, 1, 0) == 1 ? 1 : 0;
}
__global__ void k_unlock(int* mutex, int* held) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && *held != 0) {
        atomicExch(mutex, 1);
        *held = 0;
    }
}
__global__ void __launch_bounds__(256) k_copy_rgb16_if(const RGB16* __restrict__ src, RGB16* __restrict__ dst,
                                                      uint32_t npixels, const int* held) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npixels || *held == 0) return;
    dst[i] = src[i];
}
__global__ void k_copy_flag(const int* src, int* dst) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *dst = *src;
}
__global__ void k_fill_i32(int* p, int v, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

hipError_t handoff_init(int* words, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, words, 0, 4u);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, words, 1, 1u);  // word 0: the mutex, free
    return hipGetLastError();
}

hipError_t handoff_publish(const DevState& st, uint32_t npixels, bool accum64, RGB16* shared, int* words,
                           hipStream_t s) {
    if (npixels == 0) return hipSuccess;
    hipLaunchKernelGGL(k_try_lock, dim3(1), dim3(64), 0, s, words, words + 1);
    hipLaunchKernelGGL(k_to_intermediate, dim3((npixels + 255) / 256), dim3(256), 0, s, st, npixels,
                       accum64 ? 1u : 0u, shared, (const int*)(words + 1));
    hipLaunchKernelGGL(k_unlock, dim3(1), dim3(64), 0, s, words, words + 1);
    return hipGetLastError();
}

hipError_t handoff_take(const RGB16* shared, RGB16* dst, uint32_t npixels, int* words, int* took, hipStream_t s) {
    if (npixels == 0) return hipSuccess;
    hipLaunchKernelGGL(k_try_lock, dim3(1), dim3(64), 0, s, words, words + 2);
    hipLaunchKernelGGL(k_copy_rgb16_if, dim3((npixels + 255) / 256), dim3(256), 0, s, shared, dst, npixels,
                       (const int*)(words + 2));
    if (took) hipLaunchKernelGGL(k_copy_flag, dim3(1), dim3(64), 0, s, (const int*)(words + 2), took);
    hipLaunchKernelGGL(k_unlock, dim3(1), dim3(64), 0, s, words, words + 2);
    return hipGetLastError();
}

hipError_t merge_intermediate(RGB16* a, const RGB16* b, uint32_t npixels, hipStream_t s) {
    if (npixels == 0) return hipSuccess;
    hipLaunchKernelGGL(k_merge_intermediate, dim3((npixels + 255) / 256), dim3(256), 0, s, a, b, npixels);
    return hipGetLastError();
}

hipError_t tonemap_intermediate(const RGB16* in, uint32_t npixels, void* rgba8, hipStream_t s) {
    if (npixels == 0) return hipSuccess;
    hipLaunchKernelGGL(k_tonemap_intermediate, dim3((npixels + 255) / 256), dim3(256), 0, s, in, npixels,
                       (uchar4*)rgba8);
    return hipGetLastError();
}

hipError_t wf_stat(const DevState& st, uint32_t npixels, bool accum64, unsigned long long* total, hipStream_t s) {
    dim3 grid((npixels + 255) / 256);
    hipLaunchKernelGGL(k_stat, grid, dim3(256), 0, s, st, npixels, accum64 ? 1u : 0u, total);
    return hipGetLastError();
}

hipError_t wf_initialize(const DevState& st, const WFParams& p, uint64_t seed_offset, uint64_t next_pixel,
                         uint32_t npixels, hipStream_t s) {
    (void)npixels;
    dim3 grid((p.n + 255) / 256);
    hipLaunchKernelGGL(k_init<OneJob>, grid, dim3(256), 0, s, st, p, OneJob{}, (unsigned long long)seed_offset,
                       (unsigned long long)next_pixel);
    return hipGetLastError();
}

hipError_t wf_initialize_jobs(const DevState& st, const WFParams& p, const DevJobs& jobs, hipStream_t s) {
    dim3 grid((p.n + 255) / 256);
    hipLaunchKernelGGL(k_init<ManyJobs>, grid, dim3(256), 0, s, st, p, ManyJobs{jobs}, 0ull, 0ull);
    return hipGetLastError();
}

hipError_t wf_clear_accum(const DevState& st, uint32_t npixels, bool accum64, hipStream_t s) {
    dim3 grid((npixels + 255) / 256);
    hipLaunchKernelGGL(k_clear_accum, grid, dim3(256), 0, s, st, npixels, accum64 ? 1u : 0u);
    return hipGetLastError();
}


namespace {
template <class J>
hipError_t step_impl(const DevState& st, const WFParams& p, const J& jobs, const DevGuide& g, int variant,
                     int persistent_blocks, int trace_block, hipStream_t s, hipEvent_t* marks) {
    const dim3 qgrid(persistent_blocks > 0 ? persistent_blocks : 1024), blk(256);
    const dim3 full((p.n + 255) / 256);
    const uint32_t tb = (trace_block == 64 || trace_block == 128) ? (uint32_t)trace_block : 256u;
    const dim3 tgrid((p.n + tb - 1) / tb), tblk(tb);
    // optional stage marks: marks[k] is recorded before stage k (kStageMarks - 1 stages)
    auto mark = [&](int k) {
        if (marks) (void)hipEventRecord(marks[k], s);
    };
    mark(0);
    hipLaunchKernelGGL(k_new_path<J>, qgrid, blk, 0, s, st, p, jobs, g);
    mark(1);
    if (g.enabled) {
        hipError_t e = guide_infer(st, p, g, jobs.table(), s);
        if (e != hipSuccess) return e;
        mark(2);
        hipLaunchKernelGGL((k_shade_lambertian<true, J>), qgrid, blk, 0, s, st, p, jobs, g);
        mark(3);
        hipLaunchKernelGGL((k_shade_ward<true, J>), qgrid, blk, 0, s, st, p, jobs, g);
    } else {
        mark(2);
        hipLaunchKernelGGL((k_shade_lambertian<false, J>), qgrid, blk, 0, s, st, p, jobs, g);
        mark(3);
        hipLaunchKernelGGL((k_shade_ward<false, J>), qgrid, blk, 0, s, st, p, jobs, g);
    }
    mark(4);
    hipLaunchKernelGGL(k_shade_specular, qgrid, blk, 0, s, st, p);
    hipLaunchKernelGGL(k_shade_dielectric<J>, qgrid, blk, 0, s, st, p, jobs);
    mark(5);
    hipLaunchKernelGGL(k_shade_volume_surface<J>, qgrid, blk, 0, s, st, p, jobs);
    hipLaunchKernelGGL(k_shade_volume_inside<J>, qgrid, blk, 0, s, st, p, jobs);
    mark(6);
    // trace_block 0: a single render casts through the persistent ray-cast kernel; a batch (several
    // scenes) keeps one thread per slot (tgrid is 256 wide then)
    if (!J::kBatch && trace_block == 0) {
        if constexpr (!J::kBatch) {
            hipLaunchKernelGGL(k_trace_setup<J>, full, blk, 0, s, st, p, jobs);
            hipError_t e = launch_trace(jobs.sc, variant, st.xray, st.xhit, p.n, 0, s, nullptr, nullptr);
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL(k_trace_finish<J>, full, blk, 0, s, st, p, jobs);
        }
    } else if (variant == kVariantThreaded) {
        hipLaunchKernelGGL((k_trace_step<kVariantThreaded, J>), tgrid, tblk, 0, s, st, p, jobs, 0u);
    } else if (variant == kVariantSingleThreaded) {
        hipLaunchKernelGGL((k_trace_step<kVariantSingleThreaded, J>), tgrid, tblk, 0, s, st, p, jobs, 0u);
    } else {
        // the top of the tree in LDS: 80 records per wave, so that 8 waves per SIMD fit the 160 KB
        const uint32_t top = kDefaultTreelet * tb / 256u;
        hipLaunchKernelGGL((k_trace_step<kVariantStackless, J>), tgrid, tblk, (size_t)top * 64u, s, st, p, jobs,
                           top);
    }
    mark(7);
    hipLaunchKernelGGL(k_logic<J>, full, blk, 0, s, st, p, jobs, g);
    mark(8);
    hipLaunchKernelGGL(k_scan_queues, dim3(kNumQueues + 1), dim3(1024), 0, s, st,
                       g.train ? g.sample_count : (uint32_t*)nullptr, g.capacity);
    hipLaunchKernelGGL(k_scatter_queues, full, blk, 0, s, st, p, g, jobs.table());
    mark(9);
    return hipGetLastError();
}
}  // namespace

hipError_t wf_step(const DevState& st, const WFParams& p, const TraceScene& sc, const DevMaterials& m,
                   const DevEnvmap& env, const Camera& cam, const DevGuide& g, int variant, int persistent_blocks,
                   int trace_block, hipStream_t s, const DevJobs* jobs, hipEvent_t* marks) {
    if (jobs) return step_impl(st, p, ManyJobs{*jobs}, g, variant, persistent_blocks, trace_block, s, marks);
    return step_impl(st, p, OneJob{sc, m, env, cam}, g, variant, persistent_blocks, trace_block, s, marks);
}

}  // namespace rgpu
