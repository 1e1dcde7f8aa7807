// ORACLE — test infrastructure only. CPU restatement of the reference's wavefront step
// (common/wavefront_path_tracing.hpp:591-939) kernel by kernel, over the reference's AoS
// slot records (kernels/types.cl), with stable (slot-ordered) queues. Used as the checker
// by tests/ and __graft_entry__.smoke(); the product never links it.
//
// Floating point: -ffp-contract=off, IEEE denormals; sin/cos/log/exp/atan/atan2/pow are
// evaluated in double and rounded once to float; "x * M_PI * 2.0f" is evaluated in double
// as OpenCL does with its double M_PI on fp64 devices.
//
// Pinning: queue counts and RNG streams follow from the restated control flow; the
// materials are pinned by closed-form furnace tests (tests/test_oracle_wavefront.py); the
// envmap tables by the alias-table and pdf invariants.
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "oracle_math.h"

extern "C" int orc_stackless_intersect(const void* nodes, const uint32_t* prim_ids, const float* points,
                                       uint32_t stride, const uint32_t* indices, int32_t sphere_begin,
                                       const float* spheres, uint32_t nrays, const void* rays, void* hits,
                                       uint64_t* counters, int nthreads);
extern "C" int orc_mtbvh_intersect(const void* nodes, const int32_t* links, uint32_t nn, const uint32_t* prim_ids,
                                   const float* points, uint32_t stride, const uint32_t* indices, int32_t sphere_begin,
                                   const float* spheres, uint32_t nrays, const void* rays, void* hits,
                                   uint64_t* counters, int nthreads);
extern "C" int orc_tbvh_intersect(const void* nodes, const int32_t* links, uint32_t nn, const uint32_t* prim_ids,
                                  const float* points, uint32_t stride, const uint32_t* indices, int32_t sphere_begin,
                                  const float* spheres, uint32_t nrays, const void* rays, void* hits,
                                  uint64_t* counters, int nthreads);
extern "C" void orc_hit_normals(uint32_t nhits, const void* rays, const void* hits, const float* points,
                                uint32_t stride, const uint32_t* indices, int32_t sphere_begin, const float* spheres,
                                float* out_ng);

namespace {

const double kPiD = 3.14159265358979323846;
const float kPiF = (float)kPiD;

// ------------------------------------------------------------------ float3
struct f3 {
    float x, y, z;
};
inline f3 F(float x, float y, float z) { return f3{x, y, z}; }
inline f3 operator+(f3 a, f3 b) { return F(a.x + b.x, a.y + b.y, a.z + b.z); }
inline f3 operator-(f3 a, f3 b) { return F(a.x - b.x, a.y - b.y, a.z - b.z); }
inline f3 operator-(f3 a) { return F(-a.x, -a.y, -a.z); }
inline f3 operator*(f3 a, f3 b) { return F(a.x * b.x, a.y * b.y, a.z * b.z); }
inline f3 operator*(f3 a, float s) { return F(a.x * s, a.y * s, a.z * s); }
inline f3 operator*(float s, f3 a) { return F(s * a.x, s * a.y, s * a.z); }
inline f3 operator/(f3 a, float s) { return F(a.x / s, a.y / s, a.z / s); }
inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline f3 cross(f3 a, f3 b) { return F(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline f3 normalize(f3 v) {
    float inv = 1.0f / std::sqrt(dot(v, v));
    return v * inv;
}
inline float fmn(float a, float b) { return std::fmin(a, b); }
inline float fmx(float a, float b) { return std::fmax(a, b); }
// The reference's float builtins: the oracle_math.h functions in double, rounded once.
inline float sinf_(float x) { return (float)orcm::sin((double)x); }
inline float cosf_(float x) { return (float)orcm::cos((double)x); }
inline float logf_(float x) { return (float)orcm::log((double)x); }
inline float expf_(float x) { return (float)orcm::exp((double)x); }
inline float atanf_(float x) { return (float)orcm::atan((double)x); }
inline float atan2f_(float y, float x) { return (float)orcm::atan2((double)y, (double)x); }
inline float powf_(float x, float y) { return (float)orcm::pow((double)x, (double)y); }
inline float mix(float a, float b, float t) { return a + (b - a) * t; }
inline float sqr(float x) { return x * x; }

// ------------------------------------------------------------------ RNG (kernels/peseudo_random.cl)
struct Rng {
    uint32_t s[4];
};
inline uint32_t rotl(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }
inline uint32_t next(Rng& r) {
    uint32_t* s = r.s;
    const uint32_t res = rotl(s[0] * 5u, 7) * 9u;
    const uint32_t t = s[1] << 9;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 11);
    return res;
}
inline float uniform(Rng& r) {
    uint32_t bits = (next(r) >> 9) | 0x3f800000u;
    float f;
    memcpy(&f, &bits, 4);
    return f - 1.0f;
}
inline uint64_t uniform_integer(Rng& r) {
    uint64_t a = next(r) >> 1;
    uint64_t b = next(r) >> 1;
    return (a << 31) | b;
}
inline uint64_t splitmix64(uint64_t* x) {
    uint64_t z = (*x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// ------------------------------------------------------------------ reference records
enum { kLambertian = 1, kSpecular = 2, kDierectric = 3, kWard = 4, kVolume = 5, kVolumeInside = 6 };
enum { kBxdf = 0, kEnv = 1, kNone = 2 };

struct WavefrontPath {
    f3 T, L, ro, rd;
    uint32_t logic_i, pixel_index;
};
struct InVolumeList {
    int32_t volumes[7];
    int32_t count;
};
struct ExtensionResult {
    int32_t hit_surface_material;
    int32_t hit_volume_material;
    float tmin;
    f3 Ng;
};
struct ShadingResult {
    f3 Le, T;
};
struct IncidentSample {
    f3 wi;
    float selection_probs[2];
    float pdfs[2];
    uint32_t strategy;
};
struct Material {
    int32_t type, index;
};
struct Lambertian {
    float Le[3], R[3];
    int32_t back;
};
struct Ward {
    float ax, ay, refl[3], edge[3], falloff;
};
struct Volume {
    float C, R[3];
};
struct Camera {
    float eye[3], fwd[3], up[3], right[3], o[3], r[3], b[3];
    uint32_t w, h;
    float lens;
};
struct AliasB {
    float height;
    int32_t alias;
};
struct Frag {
    float beg_y, end_y, beg_phi, end_phi;
};
struct Ray {
    float ro[3];
    float tmax;
    float rd[3];
    uint32_t pad;
};
struct Hit {
    int32_t prim;
    float t, u, v;
};

inline f3 v3(const float* p) { return F(p[0], p[1], p[2]); }

// ------------------------------------------------------------------ envmap tables (restated)
struct Kahan {
    double sum = 0.0, c = 0.0;
    void add(double x) {
        double y = x - c;
        double t = sum + y;
        c = (t - sum) - y;
        sum = t;
    }
    double get() const { return sum - c; }
};

// common/share/alias_method.hpp:47-114
void alias_prepare(const std::vector<double>& w, std::vector<double>* probs, std::vector<AliasB>* out) {
    int N = (int)w.size();
    Kahan s;
    for (double v : w) s.add(v);
    double inv = 1.0 / s.get();
    probs->resize(N);
    std::vector<Kahan> h(N);
    std::vector<int> alias(N, -1);
    for (int i = 0; i < N; ++i) {
        (*probs)[i] = w[i] * inv;
        h[i].sum = (*probs)[i] * N;
        h[i].c = 0.0;
    }
    std::vector<int> lower, upper;
    for (int i = 0; i < N; ++i) {
        if (h[i].get() < 1.0) lower.push_back(i);
        else upper.push_back(i);
    }
    for (;;) {
        if (lower.empty() || upper.empty()) break;
        int li = lower.back();
        lower.pop_back();
        int ui = upper.back();
        upper.pop_back();
        double mov = 1.0 - h[li].get();
        h[ui].add(-mov);
        alias[li] = ui;
        if (h[ui].get() < 1.0) lower.push_back(ui);
        else upper.push_back(ui);
    }
    out->resize(N);
    for (int i = 0; i < N; ++i) (*out)[i] = AliasB{(float)h[i].get(), alias[i]};
}

struct Envmap {
    uint32_t w = 0, h = 0;
    std::vector<float> rgba;
    std::vector<Frag> frags;
    std::vector<float> pdf[7];
    std::vector<AliasB> alias[7];
};

// common/share/envmap.hpp:151-235 (ImageEnvmap) and 291-312 (six-axis weights); a table
// whose weights sum to zero falls back to solid-angle x direction weights, then solid angle.
void build_envmap(const float* rgb, uint32_t ch, uint32_t W, uint32_t H, Envmap* e) {
    e->w = W;
    e->h = H;
    size_t P = (size_t)W * H;
    e->rgba.assign(4 * P, 0.0f);
    for (size_t i = 0; i < P; ++i)
        for (int k = 0; k < 3; ++k) e->rgba[4 * i + k] = rgb[i * ch + k];
    double phi_step = 2.0 * kPiD / W, theta_step = kPiD / H;
    e->frags.resize(P);
    std::vector<double> sr(H), th(H);
    for (uint32_t y = 0; y < H; ++y) {
        double lt = theta_step * y, ut = theta_step * (y + 1);
        double by = std::cos(ut), ey = std::cos(lt);
        sr[y] = (std::cos(lt) - std::cos(ut)) * (2.0 * kPiD) / W;
        th[y] = (std::acos(by) + std::acos(ey)) * 0.5;
        for (uint32_t x = 0; x < W; ++x) {
            double up = 2.0 * kPiD - x * phi_step, lp = 2.0 * kPiD - (x + 1) * phi_step;
            e->frags[(size_t)y * W + x] = Frag{(float)by, (float)ey, (float)lp, (float)up};
        }
    }
    auto dirw = [](int t, double theta, double phi) {
        double s = std::sin(theta);
        double d[3] = {s * std::sin(phi), std::cos(theta), s * std::cos(phi)};
        if (t == 0) return 1.0;
        int a = (t - 1) / 2;
        return std::max(((t - 1) & 1) ? -d[a] : d[a], 0.0);
    };
    for (int t = 0; t < 7; ++t) {
        std::vector<double> w(P);
        double tot = 0.0;
        for (uint32_t y = 0; y < H; ++y)
            for (uint32_t x = 0; x < W; ++x) {
                size_t i = (size_t)y * W + x;
                double phi = ((2.0 * kPiD - (x + 1) * phi_step) + (2.0 * kPiD - x * phi_step)) * 0.5;
                const float* c = &e->rgba[4 * i];
                float Y = 0.2126f * c[0] + 0.7152f * c[1] + 0.0722f * c[2];
                w[i] = (double)Y * sr[y] * dirw(t, th[y], phi);
                tot += w[i];
            }
        if (!(tot > 0.0)) {
            tot = 0.0;
            for (uint32_t y = 0; y < H; ++y)
                for (uint32_t x = 0; x < W; ++x) {
                    size_t i = (size_t)y * W + x;
                    w[i] = sr[y] * dirw(t, th[y], 2.0 * kPiD - (x + 0.5) * phi_step);
                    tot += w[i];
                }
            if (!(tot > 0.0))
                for (uint32_t y = 0; y < H; ++y)
                    for (uint32_t x = 0; x < W; ++x) w[(size_t)y * W + x] = sr[y];
        }
        std::vector<double> probs;
        alias_prepare(w, &probs, &e->alias[t]);
        e->pdf[t].resize(P);
        for (uint32_t y = 0; y < H; ++y)
            for (uint32_t x = 0; x < W; ++x) {
                size_t i = (size_t)y * W + x;
                e->pdf[t][i] = (float)(probs[i] * (1.0 / sr[y]));
            }
    }
}

// kernels/envmap_sampling.cl:25-34
void cartesian_to_polar(f3 rd, float* theta, float* phi) {
    float z = rd.y, x = rd.z, y = rd.x;
    *theta = atan2f_(std::sqrt(x * x + y * y), z);
    *phi = atan2f_(y, x);
    if (!std::isfinite(*phi)) *phi = 0.0f;
}
// read_imagef, normalised coordinates, REPEAT addressing (OpenCL 1.2 section 8.2)
f3 tex(const Envmap& e, int i, int j) {
    const float* c = &e.rgba[4 * ((size_t)j * e.w + i)];
    return F(c[0], c[1], c[2]);
}
f3 get_envmap_value(const Envmap& e, f3 rd, bool use_filter) {
    float theta, phi;
    cartesian_to_polar(rd, &theta, &phi);
    const float pi = kPiF;
    float u = 1.0f - phi / (2.0f * pi);
    float v = theta / pi;
    int W = (int)e.w, H = (int)e.h;
    float uu = (u - std::floor(u)) * (float)W, vv = (v - std::floor(v)) * (float)H;
    // The clamps only matter for a non-finite direction (an out-of-range index otherwise).
    auto cl = [](int v, int hi) { return std::min(std::max(v, 0), hi); };
    if (!use_filter) {
        int i = std::isfinite(uu) ? (int)std::floor(uu) : 0, j = std::isfinite(vv) ? (int)std::floor(vv) : 0;
        if (i > W - 1) i = i - W;
        if (j > H - 1) j = j - H;
        return tex(e, cl(i, W - 1), cl(j, H - 1));
    }
    float us = uu - 0.5f, vs = vv - 0.5f;
    int i0 = std::isfinite(us) ? (int)std::floor(us) : 0, j0 = std::isfinite(vs) ? (int)std::floor(vs) : 0;
    int i1 = i0 + 1, j1 = j0 + 1;
    float a = us - std::floor(us), b = vs - std::floor(vs);
    if (i0 < 0) i0 += W;
    if (j0 < 0) j0 += H;
    if (i1 > W - 1) i1 -= W;
    if (j1 > H - 1) j1 -= H;
    i0 = cl(i0, W - 1);
    i1 = cl(i1, W - 1);
    j0 = cl(j0, H - 1);
    j1 = cl(j1, H - 1);
    f3 t00 = tex(e, i0, j0), t10 = tex(e, i1, j0), t01 = tex(e, i0, j1), t11 = tex(e, i1, j1);
    float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b), w01 = (1.0f - a) * b, w11 = a * b;
    return F(w00 * t00.x + w10 * t10.x + w01 * t01.x + w11 * t11.x, w00 * t00.y + w10 * t10.y + w01 * t01.y + w11 * t11.y,
             w00 * t00.z + w10 * t10.z + w01 * t01.z + w11 * t11.z);
}
// kernels/envmap_sampling.cl:54-66
uint32_t alias_method(Rng& st, const AliasB* b, uint32_t n) {
    uint32_t idx = (uint32_t)(uniform_integer(st) % (uint64_t)n);
    int alias = b[idx].alias;
    if (alias < 0) return idx;
    float height = b[idx].height;
    float u = uniform(st);
    if (u < height) return idx;
    return (uint32_t)alias;
}
// kernels/envmap_sampling.cl:68-81
int fract_int(int x, int m) {
    int r = x % m;
    return r < 0 ? r + m : r;
}
int envmap_index(const Envmap& e, f3 wi) {
    float theta, phi;
    cartesian_to_polar(wi, &theta, &phi);
    const float pi = kPiF;
    float u = 1.0f - phi / (2.0f * pi);
    float v = theta / pi;
    int ix = std::isfinite(u) ? (int)std::floor(u * (float)e.w) : 0;
    int iy = std::isfinite(v) ? (int)std::floor(v * (float)e.h) : 0;
    ix = fract_int(ix, (int)e.w);
    iy = std::min(std::max(iy, 0), (int)e.h - 1);
    return iy * (int)e.w + ix;
}
float envmap_pdf_sixAxis(const Envmap& e, f3 wi, f3 n) {
    int xt = 0.0f < n.x ? 1 : 2, yt = 0.0f < n.y ? 3 : 4, zt = 0.0f < n.z ? 5 : 6;
    f3 ap = n * n;
    int index = envmap_index(e, wi);
    float pdf = 0.0f;
    pdf += ap.x * e.pdf[xt][index];
    pdf += ap.y * e.pdf[yt][index];
    pdf += ap.z * e.pdf[zt][index];
    return pdf;
}
f3 sample_from_fragment(const Frag& f, Rng& st) {
    float y = mix(f.beg_y, f.end_y, uniform(st));
    float phi = mix(f.beg_phi, f.end_phi, uniform(st));
    f3 p = F(sinf_(phi), y, cosf_(phi));
    float r_xz = std::sqrt(fmx(1.0f - p.y * p.y, 0.0f));
    p.x *= r_xz;
    p.z *= r_xz;
    return p;
}

// ------------------------------------------------------------------ BSDF helpers
void get_orthonormal_basis(f3 z, f3* x, f3* y) {
    const float sign = std::copysign(1.0f, z.z);
    const float a = -1.0f / (sign + z.z);
    const float b = z.x * z.y * a;
    *x = F(1.0f + sign * z.x * z.x * a, sign * b, -sign * z.x);
    *y = F(b, sign + z.y * z.y * a, -z.y);
}
f3 reflect(f3 d, f3 n) { return d - 2.0f * dot(d, n) * n; }
float cubic(float x) { return x * x * x; }

// Anisotropic Ward helpers (isotropic case = kernels/ward.cl formulas).
float ward_exp(const Ward& w, f3 h, f3 z) {
    if (w.ax == w.ay) {
        float c2 = sqr(dot(h, z));
        float t2 = (1.0f - c2) / c2;
        return -t2 / sqr(w.ax);
    }
    f3 x, y;
    get_orthonormal_basis(z, &x, &y);
    float hx = dot(h, x) / w.ax, hy = dot(h, y) / w.ay, hz = dot(h, z);
    return -(hx * hx + hy * hy) / (hz * hz);
}
f3 sample_ward(const Ward& w, f3 wo, f3 zaxis, float u0, float u1) {
    const float pi = kPiF;
    float phiH, thetaH;
    if (w.ax == w.ay) {
        phiH = u0 * 2.0f * pi;
        thetaH = atanf_(w.ax * std::sqrt(-logf_(u1)));
    } else {
        float tp = 2.0f * pi * u0;
        phiH = atan2f_(w.ay * sinf_(tp), w.ax * cosf_(tp));
        float c = cosf_(phiH), s = sinf_(phiH);
        float den = c * c / (w.ax * w.ax) + s * s / (w.ay * w.ay);
        thetaH = atanf_(std::sqrt(-logf_(u1) / den));
    }
    float sinTheta = sinf_(thetaH);
    f3 hl = F(sinTheta * cosf_(phiH), sinTheta * sinf_(phiH), cosf_(thetaH));
    f3 x, y;
    get_orthonormal_basis(zaxis, &x, &y);
    f3 h = x * hl.x + y * hl.y + zaxis * hl.z;
    return reflect(-wo, h);
}
float pdf_ward(const Ward& w, f3 wo, f3 wi, f3 z) {
    if ((double)(dot(wo, z) * dot(wi, z)) < 0.0) return 0.0f;
    const float pi = kPiF;
    f3 h = normalize(wi + wo);
    float a2 = w.ax == w.ay ? sqr(w.ax) : w.ax * w.ay;
    float k0 = 1.0f / (4.0f * pi * a2 * dot(h, wi) * cubic(dot(h, z)));
    float k1 = expf_(ward_exp(w, h, z));
    return k0 * k1;
}
// Ward BRDF value f(wo, wi) of ward_stage (kernels/ward.cl:155-176): edge-tinted rho over
// pi*alpha^2 times the exponential and |wi + wo|^2 / ((wi + wo).n)^4 (Geisler-Moroder form).
f3 ward_f(const Ward& w, f3 wo, f3 wi, f3 Ng) {
    f3 lv = wi + wo;
    f3 h = normalize(lv);
    float pi = kPiF;
    float a2 = w.ax == w.ay ? sqr(w.ax) : w.ax * w.ay;
    float mu = dot(h, wi);
    float t = powf_(1.0f - mu, 1.0f / w.falloff);
    f3 rho = F(mix(w.refl[0], w.edge[0], t), mix(w.refl[1], w.edge[1], t), mix(w.refl[2], w.edge[2], t));
    f3 k0 = rho / (pi * a2);
    float k1 = expf_(ward_exp(w, h, Ng));
    float k2 = dot(lv, lv) / sqr(sqr(dot(lv, Ng)));
    return F(k0.x * k1 * k2, k0.y * k1 * k2, k0.z * k1 * k2);
}

// ------------------------------------------------------------------ the lane
struct Cfg {
    uint32_t n, rr_start, max_depth, quirks, variant, nthreads;
    float clamp, bxdf_p;
    uint64_t seed;
    uint32_t consistent = 0;  // flip Ng toward wo for the cosine lobe and the 6-axis envmap pdf too
};

struct Lane {
    Cfg cfg;
    // scene
    const float* points;
    uint32_t stride;
    const uint32_t* indices;
    int32_t sphere_begin;
    const float* spheres;
    const void* sl_nodes;
    const uint32_t* sl_prims;
    const void* mt_nodes;
    const int32_t* mt_links;
    uint32_t mt_nn;
    const uint32_t* mt_prims;
    std::vector<Material> mats;
    std::vector<Lambertian> lamb;
    std::vector<Ward> ward;
    std::vector<float> ior;
    std::vector<Volume> vols;
    Envmap env;
    Camera cam;
    // slot state
    std::vector<WavefrontPath> path;
    std::vector<InVolumeList> vol;
    std::vector<ExtensionResult> ext;
    std::vector<ShadingResult> shade;
    std::vector<IncidentSample> isample;
    std::vector<Rng> rng;
    std::vector<uint32_t> q[7];          // new_path + 6 material queues
    std::vector<uint32_t> sq[4];         // sample_env, eval_env_pdf, sample_env_6axis, eval_env_6axis_pdf
    uint64_t next_pixel = 0;
    std::vector<double> accum;           // 3 per pixel
    std::vector<uint32_t> accum_count;
};

// The primary ray of new_path (kernels/new_path.cl:44-71): the pixel's image-plane point
// jittered by two draws, the origin on the thin lens by two more (radius sqrt(u0), angle
// 2 pi u1).
void camera_ray(const Camera& cam, uint32_t pixel, Rng& st, f3* ro, f3* rd) {
    uint32_t x = pixel % cam.w, y = pixel / cam.w;
    float jx = (float)x + uniform(st);
    float jy = (float)y + uniform(st);
    f3 on_plane = v3(cam.o) + v3(cam.r) * jx + v3(cam.b) * jy;
    float u0 = uniform(st);
    float u1 = uniform(st);
    float dof = cam.lens;
    float r = std::sqrt(u0);
    float theta = (float)((double)u1 * kPiD * 2.0);
    float lensx = dof * r * cosf_(theta);
    float lensy = dof * r * sinf_(theta);
    f3 lens = v3(cam.eye) + v3(cam.up) * lensx + v3(cam.right) * lensy;
    *ro = lens;
    *rd = normalize(on_plane - lens);
}

// homogeneous_volume_through for one slot (kernels/homogeneous_volume_through.cl:24-40): a
// uniform choice among the media the path is inside (a 62-bit integer draw modulo their count),
// then a free-flight distance -ln(u) / C of that medium; no medium: no choice, far clip FLT_MAX.
void free_flight(const std::vector<Material>& mats, const std::vector<Volume>& vols, const InVolumeList& v,
                 Rng& st, int* hv_out, float* tmin_out) {
    int hv = -1;
    if (0 < v.count) {
        uint64_t u = uniform_integer(st);
        hv = v.volumes[u % (uint64_t)v.count];
    }
    float tmin = FLT_MAX;
    if (0 <= hv) {
        const Volume& V = vols[mats[hv].index];
        float u = uniform(st);
        tmin = -1.0f / V.C * logf_(u);
    }
    *hv_out = hv;
    *tmin_out = tmin;
}

// sample_envmap_6axis_stage for one sample (kernels/envmap_sampling.cl:169-217): a table by the
// squared components of n, a texel by its alias table, a direction in the texel's fragment; the
// pdf is the n^2-weighted mixture of the three tables at that texel.
f3 sample_envmap_6axis(const Envmap& env, f3 n, Rng& st, float* pdf_out) {
    const uint32_t count_env = env.w * env.h;
    int xt = 0.0f < n.x ? 1 : 2, yt = 0.0f < n.y ? 3 : 4, zt = 0.0f < n.z ? 5 : 6;
    float u = uniform(st);
    f3 ap = n * n;
    int sel = u < ap.x ? xt : (u < ap.x + ap.y ? yt : zt);
    uint32_t frag = alias_method(st, env.alias[sel].data(), count_env);
    float pdf = 0.0f;
    pdf += ap.x * env.pdf[xt][frag];
    pdf += ap.y * env.pdf[yt][frag];
    pdf += ap.z * env.pdf[zt][frag];
    *pdf_out = pdf;
    return sample_from_fragment(env.frags[frag], st);
}

// sample_envmap_stage for one sample (kernels/envmap_sampling.cl:219-248): the uniform-weight table.
f3 sample_envmap_uniform(const Envmap& env, Rng& st, float* pdf_out) {
    uint32_t frag = alias_method(st, env.alias[0].data(), env.w * env.h);
    *pdf_out = env.pdf[0][frag];
    return sample_from_fragment(env.frags[frag], st);
}

void strategy_selection(Lane& L) {
    for (int k = 0; k < 4; ++k) L.sq[k].clear();
    for (uint32_t item = 0; item < L.cfg.n; ++item) {
        const ExtensionResult& e = L.ext[item];
        bool surface_hit = 0 <= e.hit_surface_material;
        bool volume_hit = 0 <= e.hit_volume_material;
        bool any_hit = surface_hit || volume_hit;
        IncidentSample s = {};
        if (any_hit) {
            if (surface_hit) {
                int t = L.mats[e.hit_surface_material].type;
                if (t == kLambertian || t == kWard) {
                    s.selection_probs[kBxdf] = L.cfg.bxdf_p;
                    s.selection_probs[kEnv] = 1.0f - L.cfg.bxdf_p;
                } else {
                    s.selection_probs[kBxdf] = 1.0f;
                    s.selection_probs[kEnv] = 0.0f;
                }
            } else if (volume_hit) {
                s.selection_probs[kBxdf] = L.cfg.bxdf_p;
                s.selection_probs[kEnv] = 1.0f - L.cfg.bxdf_p;
            }
            float u = uniform(L.rng[item]);
            s.strategy = u <= s.selection_probs[kBxdf] ? kBxdf : kEnv;
        } else {
            s.strategy = kNone;
        }
        L.isample[item] = s;
        const int idx_sample = volume_hit ? 0 : 2, idx_eval = volume_hit ? 1 : 3;
        if (any_hit && s.selection_probs[kEnv] != 0.0f) L.sq[s.strategy == kEnv ? idx_sample : idx_eval].push_back(item);
    }
}

float incident_pdf(const IncidentSample& s) {
    float pdf = 0.0f;
    for (int k = 0; k < 2; ++k) pdf += s.selection_probs[k] * s.pdfs[k];
    return pdf;
}

void step(Lane& L) {
    const Cfg& c = L.cfg;
    const uint32_t P = L.cam.w * L.cam.h;
    // new_path (kernels/new_path.cl:17-83)
    for (uint32_t gid = 0; gid < L.q[0].size(); ++gid) {
        uint32_t pi = L.q[0][gid];
        Rng st = L.rng[pi];
        WavefrontPath& p = L.path[pi];
        p.T = F(1, 1, 1);
        p.L = F(0, 0, 0);
        p.logic_i = 0;
        uint64_t gpi = L.next_pixel + gid;
        uint32_t pixel = (uint32_t)(gpi % (uint64_t)P);
        p.pixel_index = pixel;
        camera_ray(L.cam, pixel, st, &p.ro, &p.rd);
        L.vol[pi].count = 0;
        L.rng[pi] = st;
        L.shade[pi].Le = F(0, 0, 0);
        L.shade[pi].T = F(1, 1, 1);
        L.ext[pi].hit_surface_material = -1;
        L.ext[pi].tmin = FLT_MAX;
        L.ext[pi].Ng = F(0, 0, 0);
    }
    // finalize_new_path
    L.next_pixel += L.q[0].size();
    L.q[0].clear();

    strategy_selection(L);

    // sample_envmap_6axis_stage (envmap_sampling.cl:169-217)
    for (uint32_t index : L.sq[2]) {
        Rng st = L.rng[index];
        f3 wo = -L.path[index].rd;
        f3 n = L.ext[index].Ng;
        if (dot(n, wo) < 0.0f) n = -n;
        float pdf;
        f3 wi = sample_envmap_6axis(L.env, n, st, &pdf);
        L.isample[index].wi = wi;
        L.isample[index].pdfs[kEnv] = pdf;
        L.rng[index] = st;
    }
    // sample_envmap_stage (envmap_sampling.cl:219-248)
    for (uint32_t index : L.sq[0]) {
        Rng st = L.rng[index];
        float pdf;
        f3 wi = sample_envmap_uniform(L.env, st, &pdf);
        L.isample[index].wi = wi;
        L.isample[index].pdfs[kEnv] = pdf;
        L.rng[index] = st;
    }
    // sample_or_eval_lambertian_stage (lambertian.cl:34-71): Ng is used unflipped there
    for (uint32_t item : L.q[kLambertian]) {
        f3 Ng = L.ext[item].Ng;
        if (L.cfg.consistent && dot(Ng, -L.path[item].rd) < 0.0f) Ng = -Ng;
        float cosTheta;
        if (L.isample[item].strategy == kBxdf) {
            Rng st = L.rng[item];
            float a = uniform(st);
            float b = uniform(st);
            L.rng[item] = st;
            float r = std::sqrt(a);
            float th = (float)((double)b * kPiD * 2.0);
            f3 l = F(r * cosf_(th), r * sinf_(th), std::sqrt(fmx(1.0f - a, 0.0f)));
            f3 x, y;
            get_orthonormal_basis(Ng, &x, &y);
            L.isample[item].wi = x * l.x + y * l.y + Ng * l.z;
            cosTheta = l.z;
        } else {
            cosTheta = dot(Ng, L.isample[item].wi);
        }
        L.isample[item].pdfs[kBxdf] = fmx(cosTheta, 0.0f) / kPiF;
    }
    // sample_or_eval_ward_stage (ward.cl:76-117)
    for (uint32_t item : L.q[kWard]) {
        const Ward& w = L.ward[L.mats[L.ext[item].hit_surface_material].index];
        f3 wo = -L.path[item].rd;
        f3 Ng = L.ext[item].Ng;
        if (dot(Ng, wo) < 0.0f) Ng = -Ng;
        f3 wi;
        if (L.isample[item].strategy == kBxdf) {
            Rng st = L.rng[item];
            float u0 = uniform(st);
            float u1 = uniform(st);
            L.rng[item] = st;
            wi = sample_ward(w, wo, Ng, u0, u1);
        } else {
            wi = L.isample[item].wi;
        }
        L.isample[item].wi = wi;
        L.isample[item].pdfs[kBxdf] = pdf_ward(w, wo, wi, Ng);
    }
    // sample_or_eval_homogeneous_volume_inside_stage (homogeneous_volume.cl:66-88)
    for (uint32_t item : L.q[kVolumeInside]) {
        if (L.isample[item].strategy == kBxdf) {
            Rng st = L.rng[item];
            float u0 = uniform(st);
            float u1 = uniform(st);
            L.rng[item] = st;
            float phi = (float)((double)u0 * kPiD * 2.0);
            float z = mix(-1.0f, 1.0f, u1);
            float r = std::sqrt(fmx(1.0f - z * z, 0.0f));
            L.isample[item].wi = F(r * cosf_(phi), r * sinf_(phi), z);
        }
        L.isample[item].pdfs[kBxdf] = (float)(1.0 / (kPiD * 4.0));
    }
    // evaluate_envmap_6axis_pdf_stage / evaluate_envmap_pdf_stage (envmap_sampling.cl:250-281)
    for (uint32_t index : L.sq[3]) {
        f3 n = L.ext[index].Ng;  // unflipped in envmap_sampling.cl:263-265
        if (L.cfg.consistent && dot(n, -L.path[index].rd) < 0.0f) n = -n;
        L.isample[index].pdfs[kEnv] = envmap_pdf_sixAxis(L.env, L.isample[index].wi, n);
    }
    for (uint32_t index : L.sq[1]) L.isample[index].pdfs[kEnv] = L.env.pdf[0][envmap_index(L.env, L.isample[index].wi)];
    // lambertian_stage (lambertian.cl:73-123)
    for (uint32_t item : L.q[kLambertian]) {
        const Lambertian& m = L.lamb[L.mats[L.ext[item].hit_surface_material].index];
        f3 Ng = L.ext[item].Ng;
        f3 ro = L.path[item].ro, rd = L.path[item].rd, wo = -rd;
        if (dot(Ng, wo) < 0.0f) Ng = -Ng;
        f3 wi = L.isample[item].wi;
        float pdf = incident_pdf(L.isample[item]);
        float cosTheta = dot(wi, Ng);
        f3 T = F(0, 0, 0);
        if (0.0 < cosTheta) T = v3(m.R) / kPiF * cosTheta / pdf;
        L.shade[item].Le = v3(m.Le);
        L.shade[item].T = T;
        float tmin = L.ext[item].tmin;
        L.path[item].ro = ro + rd * tmin + wi * 1.0e-5f + Ng * 1.0e-5f;
        L.path[item].rd = wi;
    }
    // ward_stage (ward.cl:119-186)
    for (uint32_t item : L.q[kWard]) {
        const Ward& w = L.ward[L.mats[L.ext[item].hit_surface_material].index];
        float tmin = L.ext[item].tmin;
        f3 Ng = L.ext[item].Ng;
        f3 ro = L.path[item].ro, rd = L.path[item].rd, wo = -rd;
        if (dot(Ng, wo) < 0.0f) Ng = -Ng;
        f3 wi = L.isample[item].wi;
        float pdf = incident_pdf(L.isample[item]);
        float cosTheta = dot(wi, Ng);
        f3 T = F(0, 0, 0);
        if (0.0f < cosTheta) {
            f3 f = ward_f(w, wo, wi, Ng);
            T = F(f.x * cosTheta / pdf, f.y * cosTheta / pdf, f.z * cosTheta / pdf);
        }
        L.shade[item].Le = F(0, 0, 0);
        L.shade[item].T = T;
        L.path[item].ro = ro + rd * tmin + wi * 1.0e-5f + Ng * 1.0e-5f;
        L.path[item].rd = wi;
    }
    // homogeneous_volume_stage (homogeneous_volume.cl:19-64)
    auto vhash = [&](int mat) {
        const Volume& v = L.vols[L.mats[mat].index];
        uint32_t a, b, cc, d;
        memcpy(&a, &v.C, 4);
        memcpy(&b, &v.R[0], 4);
        memcpy(&cc, &v.R[1], 4);
        memcpy(&d, &v.R[2], 4);
        return a ^ b ^ cc ^ d;
    };
    for (uint32_t item : L.q[kVolume]) {
        int hsm = L.ext[item].hit_surface_material;
        float tmin = L.ext[item].tmin;
        f3 Ng = L.ext[item].Ng, ro = L.path[item].ro, rd = L.path[item].rd, wo = -rd;
        bool backside = dot(Ng, wo) < 0.0f;
        L.shade[item].Le = F(0, 0, 0);
        L.shade[item].T = F(1, 1, 1);
        InVolumeList v = L.vol[item];
        if (backside) {
            uint32_t h = vhash(hsm);
            bool move = false;
            for (int i = 0; i < v.count; ++i) {
                if (move) v.volumes[i - 1] = v.volumes[i];
                else if (vhash(v.volumes[i]) == h) move = true;
            }
            if (move) v.count--;
        } else if (v.count < 7) {
            v.volumes[v.count] = hsm;
            v.count++;
        }
        L.vol[item] = v;
        L.path[item].ro = ro + rd * (tmin + 1.0e-5f);
    }
    // homogeneous_volume_inside_stage (homogeneous_volume.cl:90-129)
    for (uint32_t item : L.q[kVolumeInside]) {
        const Volume& V = L.vols[L.mats[L.ext[item].hit_volume_material].index];
        float tmin = L.ext[item].tmin;
        f3 ro = L.path[item].ro, rd = L.path[item].rd;
        f3 wi = L.isample[item].wi;
        float pdf = incident_pdf(L.isample[item]);
        L.shade[item].Le = F(0, 0, 0);
        L.shade[item].T = v3(V.R) / (float)(4.0f * kPiD) / pdf;
        L.path[item].ro = ro + rd * tmin;
        L.path[item].rd = wi;
    }
    // delta_materials (delta_materials.cl:34-135): specular then dielectric per gid
    size_t nd = std::max(L.q[kSpecular].size(), L.q[kDierectric].size());
    for (size_t gid = 0; gid < nd; ++gid) {
        if (gid < L.q[kSpecular].size()) {
            uint32_t pi = L.q[kSpecular][gid];
            float tmin = L.ext[pi].tmin;
            f3 Ng = L.ext[pi].Ng, ro = L.path[pi].ro, rd = L.path[pi].rd, wo = -rd;
            if (dot(Ng, wo) < 0.0f) Ng = -Ng;
            f3 wi = reflect(rd, Ng);
            L.shade[pi].Le = F(0, 0, 0);
            L.shade[pi].T = F(1, 1, 1);
            L.path[pi].ro = ro + rd * tmin + wi * 1.0e-5f;
            L.path[pi].rd = wi;
        }
        if (gid < L.q[kDierectric].size()) {
            uint32_t pi = L.q[kDierectric][gid];
            float tmin = L.ext[pi].tmin;
            f3 Ng = L.ext[pi].Ng, ro = L.path[pi].ro, rd = L.path[pi].rd, wo = -rd;
            bool backside = dot(Ng, wo) < 0.0f;
            float eta_i = 1.0f, eta_t = L.ior[L.mats[L.ext[pi].hit_surface_material].index];
            if (backside) std::swap(eta_i, eta_t);
            if (backside) Ng = -Ng;
            float cc = std::fabs(dot(Ng, wo));
            float g = std::sqrt(eta_t * eta_t / sqr(eta_i) - 1.0f + sqr(cc));
            float fa = 0.5f * sqr(g - cc) / sqr(g + cc);
            float fb = 1.0f + sqr(cc * (g + cc) - 1.0f) / sqr(cc * (g - cc) + 1.0f);
            float f = fa * fb;
            Rng st = L.rng[pi];
            float u = uniform(st);
            L.rng[pi] = st;
            float comp = 1.0f;
            f3 wi;
            if (u < f) {
                wi = reflect(rd, Ng);
            } else {
                float eta = eta_i / eta_t;
                float NoD = dot(Ng, rd);
                float k = 1.0f - eta * eta * (1.0f - NoD * NoD);
                if (k <= 0.0f) {
                    wi = reflect(rd, Ng);
                } else {
                    wi = eta * rd - (eta * NoD + std::sqrt(k)) * Ng;
                    comp = sqr(eta_t / eta_i);
                }
            }
            L.shade[pi].Le = F(0, 0, 0);
            L.shade[pi].T = F(comp, comp, comp);
            L.path[pi].ro = ro + rd * tmin + wi * 1.0e-5f;
            L.path[pi].rd = wi;
        }
    }
    // homogeneous_volume_through (homogeneous_volume_through.cl:7-45)
    for (uint32_t item = 0; item < c.n; ++item) {
        Rng st = L.rng[item];
        L.ext[item].hit_surface_material = -1;
        int hv;
        float tmin;
        free_flight(L.mats, L.vols, L.vol[item], st, &hv, &tmin);
        L.ext[item].hit_volume_material = hv;
        L.ext[item].tmin = tmin;
        L.rng[item] = st;
    }
    // extension_ray_cast (extension_ray_cast_stackless.cl:202-246), all slots as one batch
    std::vector<Ray> rays(c.n);
    std::vector<Hit> hits(c.n);
    for (uint32_t i = 0; i < c.n; ++i) {
        const WavefrontPath& p = L.path[i];
        rays[i] = Ray{{p.ro.x, p.ro.y, p.ro.z}, L.ext[i].tmin, {p.rd.x, p.rd.y, p.rd.z}, 0};
    }
    if (L.sl_nodes || L.mt_nodes) {
        if (c.variant == 1)
            orc_mtbvh_intersect(L.mt_nodes, L.mt_links, L.mt_nn, L.mt_prims, L.points, L.stride, L.indices,
                                L.sphere_begin, L.spheres, c.n, rays.data(), hits.data(), nullptr, (int)c.nthreads);
        else if (c.variant == 2)
            orc_tbvh_intersect(L.mt_nodes, L.mt_links, L.mt_nn, L.mt_prims, L.points, L.stride, L.indices,
                               L.sphere_begin, L.spheres, c.n, rays.data(), hits.data(), nullptr, (int)c.nthreads);
        else
            orc_stackless_intersect(L.sl_nodes, L.sl_prims, L.points, L.stride, L.indices, L.sphere_begin, L.spheres,
                                    c.n, rays.data(), hits.data(), nullptr, (int)c.nthreads);
        std::vector<float> ng(3 * (size_t)c.n);
        orc_hit_normals(c.n, rays.data(), hits.data(), L.points, L.stride, L.indices, L.sphere_begin, L.spheres,
                        ng.data());
        for (uint32_t i = 0; i < c.n; ++i) {
            if (hits[i].prim < 0) continue;
            L.ext[i].hit_surface_material = hits[i].prim;
            L.ext[i].tmin = hits[i].t;
            L.ext[i].Ng = F(ng[3 * i], ng[3 * i + 1], ng[3 * i + 2]);
            L.ext[i].hit_volume_material = -1;
        }
    }
    for (int k = 1; k < 7; ++k) L.q[k].clear();
    // logic (logic.cl:117-280), stable queue order = slot order
    for (uint32_t gid = 0; gid < c.n; ++gid) {
        WavefrontPath& p = L.path[gid];
        uint32_t logic_i = p.logic_i++;
        f3 Lr = p.L + p.T * L.shade[gid].Le;
        f3 T = p.T * L.shade[gid].T;
        int hs = L.ext[gid].hit_surface_material, hv = L.ext[gid].hit_volume_material;
        bool evalEnv, newPath;
        if (hs < 0 && hv < 0) newPath = evalEnv = true;
        else newPath = evalEnv = false;
        float lum = 0.2126f * T.x + 0.7152f * T.y + 0.0722f * T.z;
        float cp = 1.0f;
        if (c.rr_start < logic_i) cp = fmn(lum, 1.0f);
        float u = uniform(L.rng[gid]);
        if (u < cp) T = T / cp;
        else newPath = true;
        if (lum < 1.0e-5f) newPath = true;
        if (c.max_depth != 0 && logic_i >= c.max_depth) newPath = true;
        if (evalEnv) Lr = Lr + T * get_envmap_value(L.env, p.rd, logic_i == 0);
        if (newPath) {
            bool ok = std::isfinite(Lr.x) && std::isfinite(Lr.y) && std::isfinite(Lr.z) && Lr.x < c.clamp &&
                      Lr.y < c.clamp && Lr.z < c.clamp;
            if (ok) {
                uint32_t px = p.pixel_index;
                L.accum[3 * (size_t)px + 0] += Lr.x;
                L.accum[3 * (size_t)px + 1] += Lr.y;
                L.accum[3 * (size_t)px + 2] += Lr.z;
                L.accum_count[px] += 1;
            }
        }
        p.L = Lr;
        p.T = T;
        int qi;
        if (newPath) qi = 0;
        else if (0 <= hs) qi = L.mats[hs].type;
        else qi = kVolumeInside;
        L.q[qi].push_back(gid);
    }
}

}  // namespace

extern "C" {

// Reference-order CPU lane. Arrays passed in must outlive the handle.
struct orc_wf_desc {
    uint32_t n, rr_start, max_depth, quirks, variant, nthreads;
    float clamp, bxdf_p;
    uint64_t seed;
    const float* points;
    uint32_t stride;
    const uint32_t* indices;
    int32_t sphere_begin;
    const float* spheres;
    const void* sl_nodes;
    const uint32_t* sl_prims;
    const void* mt_nodes;
    const int32_t* mt_links;
    uint32_t mt_nn;
    const uint32_t* mt_prims;
    uint32_t nprims;
    const int32_t* mat_types;
    const int32_t* mat_indices;
    const float* lambertians;  // 7 floats each: Le, R, back (as float)
    uint32_t nl;
    const float* wards;  // 9 floats each: ax, ay, refl, edge, falloff
    uint32_t nw;
    const float* iors;
    uint32_t nd;
    const float* volumes;  // 4 floats each: C, R
    uint32_t nv;
    const float* env_rgb;
    uint32_t env_channels, env_w, env_h;
    const float* camera;  // 21 floats: eye, forward, up, right, o, r, b
    uint32_t width, height;
    float lens_radius;
    uint32_t consistent;
};

void* orc_wf_create(const orc_wf_desc* d) {
    Lane* L = new Lane();
    L->cfg = Cfg{d->n, d->rr_start ? d->rr_start : 25u, d->max_depth, d->quirks, d->variant,
                 d->nthreads ? d->nthreads : 1u, d->clamp > 0 ? d->clamp : 1000.0f, d->bxdf_p > 0 ? d->bxdf_p : 0.5f,
                 d->seed};
    L->cfg.consistent = d->consistent;
    L->points = d->points;
    L->stride = d->stride;
    L->indices = d->indices;
    L->sphere_begin = d->sphere_begin;
    L->spheres = d->spheres;
    L->sl_nodes = d->sl_nodes;
    L->sl_prims = d->sl_prims;
    L->mt_nodes = d->mt_nodes;
    L->mt_links = d->mt_links;
    L->mt_nn = d->mt_nn;
    L->mt_prims = d->mt_prims;
    L->mats.resize(d->nprims);
    for (uint32_t i = 0; i < d->nprims; ++i) L->mats[i] = Material{d->mat_types[i], d->mat_indices[i]};
    L->lamb.resize(d->nl);
    for (uint32_t i = 0; i < d->nl; ++i) {
        const float* s = d->lambertians + 7 * (size_t)i;
        for (int k = 0; k < 3; ++k) {
            L->lamb[i].Le[k] = s[k];
            L->lamb[i].R[k] = s[3 + k];
        }
        L->lamb[i].back = (int32_t)s[6];
    }
    L->ward.resize(d->nw);
    for (uint32_t i = 0; i < d->nw; ++i) {
        const float* s = d->wards + 9 * (size_t)i;
        Ward& w = L->ward[i];
        w.ax = s[0];
        w.ay = s[1];
        for (int k = 0; k < 3; ++k) {
            w.refl[k] = s[2 + k];
            w.edge[k] = s[5 + k];
        }
        w.falloff = s[8];
    }
    L->ior.assign(d->iors, d->iors + d->nd);
    L->vols.resize(d->nv);
    for (uint32_t i = 0; i < d->nv; ++i) {
        const float* s = d->volumes + 4 * (size_t)i;
        L->vols[i].C = s[0];
        for (int k = 0; k < 3; ++k) L->vols[i].R[k] = s[1 + k];
    }
    build_envmap(d->env_rgb, d->env_channels, d->env_w, d->env_h, &L->env);
    Camera& c = L->cam;
    float* dst[7] = {c.eye, c.fwd, c.up, c.right, c.o, c.r, c.b};
    for (int a = 0; a < 7; ++a)
        for (int k = 0; k < 3; ++k) dst[a][k] = d->camera[3 * a + k];
    c.w = d->width;
    c.h = d->height;
    c.lens = d->lens_radius;
    const uint32_t n = d->n;
    L->path.assign(n, WavefrontPath{});
    L->vol.assign(n, InVolumeList{});
    L->ext.assign(n, ExtensionResult{-1, -1, FLT_MAX, F(0, 0, 0)});
    L->shade.assign(n, ShadingResult{});
    L->isample.assign(n, IncidentSample{});
    L->rng.assign(n, Rng{});
    L->accum.assign(3 * (size_t)c.w * c.h, 0.0);
    L->accum_count.assign((size_t)c.w * c.h, 0);
    return L;
}

void orc_wf_destroy(void* h) { delete (Lane*)h; }

// WavefrontLane::initialize (common/wavefront_path_tracing.hpp:532-570)
void orc_wf_initialize(void* h, uint32_t lane_index, uint32_t num_lanes) {
    Lane* L = (Lane*)h;
    const uint32_t n = L->cfg.n, P = L->cam.w * L->cam.h;
    L->next_pixel = (uint64_t)lane_index * (P / num_lanes);
    uint64_t seed_offset = L->cfg.seed + (uint64_t)lane_index * n;
    for (uint32_t gid = 0; gid < n; ++gid) {
        uint64_t x = seed_offset + gid;
        uint64_t r0 = splitmix64(&x), r1 = splitmix64(&x);
        Rng& s = L->rng[gid];
        s.s[0] = (uint32_t)r0;
        s.s[1] = (uint32_t)(r0 >> 32);
        s.s[2] = (uint32_t)r1;
        s.s[3] = (uint32_t)(r1 >> 32);
        if ((s.s[0] | s.s[1] | s.s[2] | s.s[3]) == 0) s.s[0] = 1;
    }
    for (int k = 0; k < 7; ++k) L->q[k].clear();
    for (uint32_t gid = 0; gid < n; ++gid) L->q[0].push_back(gid);
    std::fill(L->accum.begin(), L->accum.end(), 0.0);
    std::fill(L->accum_count.begin(), L->accum_count.end(), 0u);
    for (uint32_t i = 0; i < n; ++i) {
        L->ext[i] = ExtensionResult{-1, -1, FLT_MAX, F(0, 0, 0)};
        L->vol[i] = InVolumeList{};
    }
}

void orc_wf_step(void* h, uint32_t steps) {
    Lane* L = (Lane*)h;
    for (uint32_t s = 0; s < steps; ++s) {
        if (!L->cfg.quirks) {
            // without the quirk, a new path does no strategy draw for a stale volume id
            for (uint32_t pi : L->q[0]) L->ext[pi].hit_volume_material = -1;
        }
        step(*L);
    }
}

void orc_wf_queue_counts(void* h, uint32_t* counts7, uint64_t* next_pixel) {
    Lane* L = (Lane*)h;
    for (int k = 0; k < 7; ++k) counts7[k] = (uint32_t)L->q[k].size();
    if (next_pixel) *next_pixel = L->next_pixel;
}

void orc_wf_read_accumulation(void* h, double* rgb, uint32_t* counts) {
    Lane* L = (Lane*)h;
    size_t P = (size_t)L->cam.w * L->cam.h;
    if (rgb) memcpy(rgb, L->accum.data(), 3 * P * 8);
    if (counts) memcpy(counts, L->accum_count.data(), P * 4);
}

// Same record layouts as rgpu_wavefront_read_state.
void orc_wf_read_state(void* h, uint32_t* rng, float* rays, float* paths, void* ext, float* shade, int32_t* vol,
                       uint32_t* items) {
    Lane* L = (Lane*)h;
    const uint32_t n = L->cfg.n;
    for (uint32_t i = 0; i < n; ++i) {
        if (rng) memcpy(rng + 4 * (size_t)i, L->rng[i].s, 16);
        const WavefrontPath& p = L->path[i];
        if (rays) {
            float* r = rays + 8 * (size_t)i;
            r[0] = p.ro.x;
            r[1] = p.ro.y;
            r[2] = p.ro.z;
            r[3] = 0.0f;
            r[4] = p.rd.x;
            r[5] = p.rd.y;
            r[6] = p.rd.z;
            memcpy(&r[7], &p.pixel_index, 4);
        }
        if (paths) {
            float* q = paths + 8 * (size_t)i;
            q[0] = p.T.x;
            q[1] = p.T.y;
            q[2] = p.T.z;
            memcpy(&q[3], &p.logic_i, 4);
            q[4] = p.L.x;
            q[5] = p.L.y;
            q[6] = p.L.z;
            q[7] = 0.0f;
        }
        if (ext) {
            const ExtensionResult& e = L->ext[i];
            int32_t* q = (int32_t*)ext + 8 * (size_t)i;
            q[0] = e.hit_surface_material;
            q[1] = e.hit_volume_material;
            memcpy(&q[2], &e.tmin, 4);
            memcpy(&q[3], &e.Ng.x, 4);
            memcpy(&q[4], &e.Ng.y, 4);
            memcpy(&q[5], &e.Ng.z, 4);
            q[6] = q[7] = 0;
        }
        if (shade) {
            const ShadingResult& s = L->shade[i];
            float* q = shade + 8 * (size_t)i;
            q[0] = s.Le.x;
            q[1] = s.Le.y;
            q[2] = s.Le.z;
            q[3] = s.T.x;
            q[4] = s.T.y;
            q[5] = s.T.z;
            q[6] = q[7] = 0.0f;
        }
        if (vol) {
            memcpy(vol + 8 * (size_t)i, L->vol[i].volumes, 28);
            vol[8 * (size_t)i + 7] = L->vol[i].count;
        }
    }
    if (items)
        for (int k = 0; k < 7; ++k)
            for (size_t j = 0; j < L->q[k].size(); ++j) items[(size_t)k * n + j] = L->q[k][j];
}

// Envmap tables of the restated ImageEnvmap, for comparison with the product's.
void* orc_envmap_build(const float* rgb, uint32_t channels, uint32_t w, uint32_t h) {
    Envmap* e = new Envmap();
    build_envmap(rgb, channels, w, h, e);
    return e;
}
void orc_envmap_export(void* h, int table, float* pdf, float* heights, int32_t* aliases, float* frags) {
    Envmap* e = (Envmap*)h;
    size_t P = (size_t)e->w * e->h;
    for (size_t i = 0; i < P; ++i) {
        if (pdf) pdf[i] = e->pdf[table][i];
        if (heights) heights[i] = e->alias[table][i].height;
        if (aliases) aliases[i] = e->alias[table][i].alias;
        if (frags) memcpy(frags + 4 * i, &e->frags[i], 16);
    }
}
void orc_envmap_destroy(void* h) { delete (Envmap*)h; }
// Envmap radiance lookups and samples for direct tests.
void orc_envmap_lookup(void* h, uint32_t n, const float* dirs, int filter, float* out) {
    Envmap* e = (Envmap*)h;
    for (uint32_t i = 0; i < n; ++i) {
        f3 r = get_envmap_value(*e, F(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]), filter != 0);
        out[3 * i] = r.x;
        out[3 * i + 1] = r.y;
        out[3 * i + 2] = r.z;
    }
}
// Draw n samples from table `table` (0 uniform; 1..6 by the normal's six-axis mixture when
// normal != NULL): directions and pdfs, consuming `state`.
void orc_envmap_sample(void* h, uint32_t* state, uint32_t n, const float* normal, float* dirs, float* pdfs) {
    Envmap* e = (Envmap*)h;
    Rng st;
    memcpy(st.s, state, 16);
    for (uint32_t i = 0; i < n; ++i) {
        float pdf;
        f3 wi = normal ? sample_envmap_6axis(*e, F(normal[0], normal[1], normal[2]), st, &pdf)
                       : sample_envmap_uniform(*e, st, &pdf);
        dirs[3 * i] = wi.x;
        dirs[3 * i + 1] = wi.y;
        dirs[3 * i + 2] = wi.z;
        pdfs[i] = pdf;
    }
    memcpy(state, st.s, 16);
}
// One sample per state (4 words each, advanced in place), the normal of sample i at
// normals[3i] (NULL = uniform table): sample_envmap_6axis_stage / sample_envmap_stage
// (kernels/envmap_sampling.cl:169-248) over a flat batch.
void orc_envmap_sample_each(void* h, uint32_t* states, uint32_t n, const float* normals, float* dirs, float* pdfs) {
    for (uint32_t i = 0; i < n; ++i)
        orc_envmap_sample(h, states + 4 * (size_t)i, 1, normals ? normals + 3 * (size_t)i : nullptr,
                          dirs + 3 * (size_t)i, pdfs + i);
}
void orc_envmap_pdf(void* h, uint32_t n, const float* dirs, const float* normal, float* pdfs) {
    Envmap* e = (Envmap*)h;
    for (uint32_t i = 0; i < n; ++i) {
        f3 wi = F(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]);
        pdfs[i] = normal ? envmap_pdf_sixAxis(*e, wi, F(normal[0], normal[1], normal[2]))
                         : e->pdf[0][envmap_index(*e, wi)];
    }
}
// pdf of direction i under the normal of row i (NULL = uniform table).
void orc_envmap_pdf_each(void* h, uint32_t n, const float* dirs, const float* normals, float* pdfs) {
    for (uint32_t i = 0; i < n; ++i)
        orc_envmap_pdf(h, 1, dirs + 3 * (size_t)i, normals ? normals + 3 * (size_t)i : nullptr, pdfs + i);
}

// ---------------------------------------------------------------- single-function test hooks
// The lane's own stage functions over flat batches, so that tests can pin them against closed
// forms and statistics: Ward (w: 9 floats ax, ay, refl, edge, falloff) sampling, pdf and BRDF.
void orc_ward_sample(const float* w9, uint32_t n, const float* wo, const float* nrm, const float* u, float* wi) {
    Ward w{};
    w.ax = w9[0];
    w.ay = w9[1];
    for (int k = 0; k < 3; ++k) {
        w.refl[k] = w9[2 + k];
        w.edge[k] = w9[5 + k];
    }
    w.falloff = w9[8];
    for (uint32_t i = 0; i < n; ++i) {
        f3 r = sample_ward(w, v3(wo + 3 * (size_t)i), v3(nrm + 3 * (size_t)i), u[2 * (size_t)i], u[2 * (size_t)i + 1]);
        wi[3 * (size_t)i] = r.x;
        wi[3 * (size_t)i + 1] = r.y;
        wi[3 * (size_t)i + 2] = r.z;
    }
}
// pdf[i] (solid angle) and f[3i..] (BRDF) of wi given wo about the normal.
void orc_ward_eval(const float* w9, uint32_t n, const float* wo, const float* wi, const float* nrm, float* pdf,
                   float* f) {
    Ward w{};
    w.ax = w9[0];
    w.ay = w9[1];
    for (int k = 0; k < 3; ++k) {
        w.refl[k] = w9[2 + k];
        w.edge[k] = w9[5 + k];
    }
    w.falloff = w9[8];
    for (uint32_t i = 0; i < n; ++i) {
        f3 a = v3(wo + 3 * (size_t)i), b = v3(wi + 3 * (size_t)i), z = v3(nrm + 3 * (size_t)i);
        if (pdf) pdf[i] = pdf_ward(w, a, b, z);
        if (f) {
            f3 r = ward_f(w, a, b, z);
            f[3 * (size_t)i] = r.x;
            f[3 * (size_t)i + 1] = r.y;
            f[3 * (size_t)i + 2] = r.z;
        }
    }
}
// new_path's primary rays: camera (21 floats eye, forward, up, right, o, r, b), image size and
// lens radius; one ray per pixel index, each drawing from its own state (4 words, advanced).
void orc_camera_rays(const float* camera, uint32_t width, uint32_t height, float lens, uint32_t n,
                     const uint32_t* pixels, uint32_t* states, float* ro, float* rd) {
    Camera c{};
    float* dst[7] = {c.eye, c.fwd, c.up, c.right, c.o, c.r, c.b};
    for (int a = 0; a < 7; ++a)
        for (int k = 0; k < 3; ++k) dst[a][k] = camera[3 * a + k];
    c.w = width;
    c.h = height;
    c.lens = lens;
    for (uint32_t i = 0; i < n; ++i) {
        Rng st;
        memcpy(st.s, states + 4 * (size_t)i, 16);
        f3 o, d;
        camera_ray(c, pixels[i], st, &o, &d);
        memcpy(states + 4 * (size_t)i, st.s, 16);
        memcpy(ro + 3 * (size_t)i, &o, 12);
        memcpy(rd + 3 * (size_t)i, &d, 12);
    }
}
// homogeneous_volume_through per row: the medium list (7 ids + count; ids index extinctions C),
// a state (advanced); out: the chosen medium (-1: none) and the free-flight distance.
void orc_free_flight(const float* C, uint32_t nvol, uint32_t n, const int32_t* lists8, uint32_t* states,
                     int32_t* hv, float* tmin) {
    std::vector<Material> mats(nvol);
    std::vector<Volume> vols(nvol);
    for (uint32_t i = 0; i < nvol; ++i) {
        mats[i] = Material{kVolume, (int32_t)i};
        vols[i].C = C[i];
    }
    for (uint32_t i = 0; i < n; ++i) {
        InVolumeList v{};
        for (int k = 0; k < 7; ++k) v.volumes[k] = lists8[8 * (size_t)i + k];
        v.count = lists8[8 * (size_t)i + 7];
        Rng st;
        memcpy(st.s, states + 4 * (size_t)i, 16);
        free_flight(mats, vols, v, st, hv + i, tmin + i);
        memcpy(states + 4 * (size_t)i, st.s, 16);
    }
}
// The double-precision functions of oracle_math.h over a batch (fn: 0 exp, 1 log, 2 sin, 3 cos,
// 4 atan, 5 atan2(x, y), 6 pow(x, y)), for checks against a reference libm.
void orc_math_f64(int fn, uint32_t n, const double* x, const double* y, double* out) {
    for (uint32_t i = 0; i < n; ++i) {
        const double a = x[i], b = y ? y[i] : 0.0;
        double r;
        switch (fn) {
            case 0: r = orcm::exp(a); break;
            case 1: r = orcm::log(a); break;
            case 2: r = orcm::sin(a); break;
            case 3: r = orcm::cos(a); break;
            case 4: r = orcm::atan(a); break;
            case 5: r = orcm::atan2(a, b); break;
            default: r = orcm::pow(a, b); break;
        }
        out[i] = r;
    }
}
// (float)pow((double)x[i], (double)y) as the tonemap evaluates it.
void orc_pow_f32(uint32_t n, const float* x, float y, float* out) {
    for (uint32_t i = 0; i < n; ++i) out[i] = powf_(x[i], y);
}

// Texel index of directions (envmap_index, kernels/envmap_sampling.cl:100-114).
void orc_envmap_index(void* h, uint32_t n, const float* dirs, int32_t* out) {
    Envmap* e = (Envmap*)h;
    for (uint32_t i = 0; i < n; ++i) out[i] = envmap_index(*e, v3(dirs + 3 * (size_t)i));
}

}  // extern "C"
