// ORACLE — test infrastructure only. FP64 CPU model of the path-guiding network: input
// encoding, 3 hidden ReLU layers of width 64, an 8-lobe von Mises-Fisher mixture head,
// the guiding loss and its gradient, and Adam. No reference counterpart exists (the
// reference has no path guiding); the model is pinned by finite differences, the vMF
// normalisation integral and Adam's closed form (tests/test_oracle_mlp.py).
//
// With `emulate` set, values are rounded to bfloat16 (round to nearest even) exactly where
// the GPU kernels round them before a matrix-core product (encoded inputs, weights,
// hidden activations, back-propagated deltas); sums stay in double.
#include <stdint.h>
#include <string.h>

#include <cmath>
#include <vector>

namespace {

const int kIn = 96, kHid = 64, kOut = 64, kLobes = 8, kBands = 12;
const int oW0 = 0, oB0 = oW0 + kHid * kIn, oW1 = oB0 + kHid, oB1 = oW1 + kHid * kHid, oW2 = oB1 + kHid,
          oB2 = oW2 + kHid * kHid, oW3 = oB2 + kHid, oB3 = oW3 + kOut * kHid, kParams = oB3 + kOut;
const double kPi = 3.14159265358979323846;
const double kLogKappaMin = -3.0, kLogKappaMax = 7.5;
// learned selection probability: c = clamp(sigmoid(out[40]), 0.05, 0.95)
const int kSelOut = 5 * kLobes;
const double kSelMin = 0.05, kSelMax = 0.95;

double bf16(double x) {
    float f = (float)x;
    uint32_t u;
    memcpy(&u, &f, 4);
    uint32_t lsb = (u >> 16) & 1u;
    u = (u + 0x7FFFu + lsb) & 0xFFFF0000u;
    memcpy(&f, &u, 4);
    return (double)f;
}

struct Net {
    const double* w;
    bool emulate;
    double R(double x) const { return emulate ? bf16(x) : x; }
};

// features: 12-band sin/cos of the normalised position, then normal and outgoing direction,
// then the scene class of a batch job one-hot in features 78..85 (class code c = 1..8 sets
// feature 77 + c; c = 0, a single render, leaves them zero)
void encode(const double* p01, const double* n, const double* wo, uint32_t cls, double* x) {
    for (int k = 0; k < kBands; ++k)
        for (int a = 0; a < 3; ++a) {
            double arg = (double)(1u << k) * kPi * p01[a];
            x[6 * k + 2 * a] = std::sin(arg);
            x[6 * k + 2 * a + 1] = std::cos(arg);
        }
    for (int a = 0; a < 3; ++a) {
        x[72 + a] = n[a];
        x[75 + a] = wo[a];
    }
    for (int i = 78; i < kIn; ++i) x[i] = 0.0;
    if (cls >= 1 && cls <= 8) x[77 + cls] = 1.0;
}

struct Fwd {
    double x[kIn], h1[kHid], h2[kHid], h3[kHid], out[kOut];
};

void layer(const Net& net, int oW, int oB, int nin, int nout, const double* in, double* out, bool relu) {
    for (int o = 0; o < nout; ++o) {
        double s = 0.0;
        for (int i = 0; i < nin; ++i) s += net.R(net.w[oW + o * nin + i]) * in[i];
        s += net.w[oB + o];
        out[o] = relu ? (s > 0.0 ? s : 0.0) : s;
    }
}

void forward(const Net& net, const double* x_in, Fwd* f) {
    for (int i = 0; i < kIn; ++i) f->x[i] = net.R(x_in[i]);
    double t[kHid];
    layer(net, oW0, oB0, kIn, kHid, f->x, t, true);
    for (int i = 0; i < kHid; ++i) f->h1[i] = net.R(t[i]);
    layer(net, oW1, oB1, kHid, kHid, f->h1, t, true);
    for (int i = 0; i < kHid; ++i) f->h2[i] = net.R(t[i]);
    layer(net, oW2, oB2, kHid, kHid, f->h2, t, true);
    for (int i = 0; i < kHid; ++i) f->h3[i] = net.R(t[i]);
    layer(net, oW3, oB3, kHid, kOut, f->h3, f->out, false);
}

struct Lobe {
    double mu[3], mnorm, kappa, raw_kappa, logw;
};

// Head: lobe k uses outputs 5k..5k+4 = (mean direction (unnormalised), log kappa, logit).
void lobes(const double* out, Lobe* L) {
    double mx = -1e300;
    for (int k = 0; k < kLobes; ++k) mx = std::fmax(mx, out[5 * k + 4]);
    double z = 0.0;
    for (int k = 0; k < kLobes; ++k) z += std::exp(out[5 * k + 4] - mx);
    for (int k = 0; k < kLobes; ++k) {
        const double* o = out + 5 * k;
        // a mean whose largest component is below 1e-15 has no direction (a row whose hidden
        // units are all zero leaves the outputs at their biases): the lobe takes +z and its mean
        // no gradient; so does a non-finite one
        const double mxo = std::fmax(std::fmax(std::fabs(o[0]), std::fabs(o[1])), std::fabs(o[2]));
        if (mxo > 1e-15 && mxo <= 3.0e38) {
            const double m = std::sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
            for (int a = 0; a < 3; ++a) L[k].mu[a] = o[a] / m;
            L[k].mnorm = m;
        } else {
            L[k].mu[0] = 0.0;
            L[k].mu[1] = 0.0;
            L[k].mu[2] = 1.0;
            L[k].mnorm = 0.0;
        }
        L[k].raw_kappa = o[3];
        L[k].kappa = std::exp(std::fmin(std::fmax(o[3], kLogKappaMin), kLogKappaMax));
        L[k].logw = o[4] - mx - std::log(z);
    }
}

// log of the vMF density kappa / (2 pi (1 - exp(-2 kappa))) * exp(kappa (mu.w - 1))
double log_vmf(const Lobe& l, const double* w) {
    double c = l.mu[0] * w[0] + l.mu[1] * w[1] + l.mu[2] * w[2];
    return std::log(l.kappa) - std::log(2.0 * kPi) - std::log1p(-std::exp(-2.0 * l.kappa)) + l.kappa * (c - 1.0);
}

double log_mixture(const Lobe* L, const double* w, double* resp) {
    double a[kLobes], mx = -1e300;
    for (int k = 0; k < kLobes; ++k) {
        a[k] = L[k].logw + log_vmf(L[k], w);
        mx = std::fmax(mx, a[k]);
    }
    double s = 0.0;
    for (int k = 0; k < kLobes; ++k) s += std::exp(a[k] - mx);
    double lse = mx + std::log(s);
    if (resp)
        for (int k = 0; k < kLobes; ++k) resp[k] = std::exp(a[k] - lse);
    return lse;
}

// d(-weight * log q(w)) / d(out)
void head_grad(const double* out, const double* w, double weight, double* dout) {
    Lobe L[kLobes];
    lobes(out, L);
    double r[kLobes];
    log_mixture(L, w, r);
    for (int i = 0; i < kOut; ++i) dout[i] = 0.0;
    for (int k = 0; k < kLobes; ++k) {
        double pi_k = std::exp(L[k].logw);
        dout[5 * k + 4] = weight * (pi_k - r[k]);
        double kap = L[k].kappa;
        double c = L[k].mu[0] * w[0] + L[k].mu[1] * w[1] + L[k].mu[2] * w[2];
        double e = std::exp(-2.0 * kap);
        double dlogf_dk = 1.0 / kap - 2.0 * e / (1.0 - e) + (c - 1.0);
        bool inside = L[k].raw_kappa > kLogKappaMin && L[k].raw_kappa < kLogKappaMax;
        dout[5 * k + 3] = inside ? -weight * r[k] * dlogf_dk * kap : 0.0;
        for (int a = 0; a < 3; ++a)
            dout[5 * k + a] = L[k].mnorm > 0.0 ? -weight * r[k] * kap * (w[a] - c * L[k].mu[a]) / L[k].mnorm : 0.0;
    }
}

void backprop_layer(const Net& net, int oW, int oB, int nin, int nout, const double* in, const double* dz,
                    double* g, double* din) {
    for (int o = 0; o < nout; ++o) {
        g[oB + o] += dz[o];
        for (int i = 0; i < nin; ++i) g[oW + o * nin + i] += dz[o] * in[i];
    }
    if (din)
        for (int i = 0; i < nin; ++i) {
            double s = 0.0;
            for (int o = 0; o < nout; ++o) s += dz[o] * net.R(net.w[oW + o * nin + i]);
            din[i] = s;
        }
}

}  // namespace

extern "C" {

int orc_mlp_param_count() { return kParams; }

// cls: per-row class codes (0 = none), or null for none
void orc_mlp_encode(uint32_t n, const double* p01, const double* nrm, const double* wo, const uint32_t* cls,
                    double* x) {
    for (uint32_t i = 0; i < n; ++i)
        encode(p01 + 3 * i, nrm + 3 * i, wo + 3 * i, cls ? cls[i] : 0u, x + (size_t)kIn * i);
}

// out: n x 64 raw head outputs
void orc_mlp_forward(const double* params, uint32_t n, const double* x, double* out, int emulate) {
    Net net{params, emulate != 0};
    for (uint32_t i = 0; i < n; ++i) {
        Fwd f;
        forward(net, x + (size_t)kIn * i, &f);
        memcpy(out + (size_t)kOut * i, f.out, sizeof(f.out));
    }
}

// Guide density of directions w (n x 3) given raw head outputs (n x 64).
void orc_mlp_pdf(uint32_t n, const double* out, const double* w, double* pdf) {
    for (uint32_t i = 0; i < n; ++i) {
        Lobe L[kLobes];
        lobes(out + (size_t)kOut * i, L);
        pdf[i] = std::exp(log_mixture(L, w + 3 * i, nullptr));
    }
}

// Mean over the batch of -weight * log q(w), and its gradient with respect to every parameter.
// With other_pdf (may be null), rows whose density of the other strategies p_o is positive add
// -weight * log(c q(w) + (1 - c) p_o), the log density of the one-sample mixture with the
// learned selection probability c = clamp(sigmoid(out[40]), 0.05, 0.95).
double orc_mlp_loss_grad_sel(const double* params, uint32_t n, const double* x, const double* w, const double* weight,
                             const double* other_pdf, double* grad, int emulate) {
    Net net{params, emulate != 0};
    for (int i = 0; i < kParams; ++i) grad[i] = 0.0;
    double loss = 0.0;
    for (uint32_t b = 0; b < n; ++b) {
        Fwd f;
        forward(net, x + (size_t)kIn * b, &f);
        Lobe L[kLobes];
        lobes(f.out, L);
        const double lq = log_mixture(L, w + 3 * b, nullptr);
        loss += -weight[b] * lq;
        double d3[kOut], dh[kHid], dz[kHid];
        head_grad(f.out, w + 3 * b, weight[b] / n, d3);
        const double po = other_pdf ? other_pdf[b] : 0.0;
        if (po > 0.0) {
            const double sg = 1.0 / (1.0 + std::exp(-f.out[kSelOut]));
            const double c = std::fmin(std::fmax(sg, kSelMin), kSelMax);
            const double q = std::exp(lq);
            const double pm = c * q + (1.0 - c) * po;
            loss += -weight[b] * std::log(pm);
            // d(-w log pm)/d(lobe outputs) = (c q / pm) d(-w log q)/d(lobe outputs)
            const double scale = 1.0 + c * q / pm;
            for (int i = 0; i < kSelOut; ++i) d3[i] *= scale;
            d3[kSelOut] = (sg > kSelMin && sg < kSelMax) ? -weight[b] / n * (q - po) * c * (1.0 - c) / pm : 0.0;
        }
        for (int i = 0; i < kOut; ++i) d3[i] = net.R(d3[i]);
        backprop_layer(net, oW3, oB3, kHid, kOut, f.h3, d3, grad, dh);
        for (int i = 0; i < kHid; ++i) dz[i] = net.R(f.h3[i] > 0.0 ? dh[i] : 0.0);
        backprop_layer(net, oW2, oB2, kHid, kHid, f.h2, dz, grad, dh);
        for (int i = 0; i < kHid; ++i) dz[i] = net.R(f.h2[i] > 0.0 ? dh[i] : 0.0);
        backprop_layer(net, oW1, oB1, kHid, kHid, f.h1, dz, grad, dh);
        for (int i = 0; i < kHid; ++i) dz[i] = net.R(f.h1[i] > 0.0 ? dh[i] : 0.0);
        backprop_layer(net, oW0, oB0, kIn, kHid, f.x, dz, grad, nullptr);
    }
    return loss / n;
}

double orc_mlp_loss_grad(const double* params, uint32_t n, const double* x, const double* w, const double* weight,
                         double* grad, int emulate) {
    return orc_mlp_loss_grad_sel(params, n, x, w, weight, nullptr, grad, emulate);
}

// Adam (Kingma & Ba 2015), bias-corrected, step index t >= 1.
void orc_adam(uint32_t n, double* w, double* m, double* v, const double* g, double lr, double b1, double b2,
              double eps, uint32_t t) {
    for (uint32_t i = 0; i < n; ++i) {
        m[i] = b1 * m[i] + (1.0 - b1) * g[i];
        v[i] = b2 * v[i] + (1.0 - b2) * g[i] * g[i];
        double mh = m[i] / (1.0 - std::pow(b1, t));
        double vh = v[i] / (1.0 - std::pow(b2, t));
        w[i] -= lr * mh / (std::sqrt(vh) + eps);
    }
}

// vMF mixture sampling (Jakob 2012 inversion for one lobe): u = 3 uniforms per sample.
void orc_mlp_sample(const double* out, uint32_t n, const double* u, double* w) {
    Lobe L[kLobes];
    lobes(out, L);
    for (uint32_t s = 0; s < n; ++s) {
        double c = 0.0;
        int k = kLobes - 1;
        for (int j = 0; j < kLobes; ++j) {
            c += std::exp(L[j].logw);
            if (u[3 * s] < c) {
                k = j;
                break;
            }
        }
        double kap = L[k].kappa;
        double ct = 1.0 + std::log(u[3 * s + 1] + (1.0 - u[3 * s + 1]) * std::exp(-2.0 * kap)) / kap;
        double st = std::sqrt(std::fmax(0.0, 1.0 - ct * ct));
        double ph = 2.0 * kPi * u[3 * s + 2];
        const double* z = L[k].mu;
        double sign = std::copysign(1.0, z[2]);
        double a = -1.0 / (sign + z[2]);
        double b = z[0] * z[1] * a;
        double x[3] = {1.0 + sign * z[0] * z[0] * a, sign * b, -sign * z[0]};
        double y[3] = {b, sign + z[1] * z[1] * a, -z[1]};
        for (int i = 0; i < 3; ++i) w[3 * s + i] = x[i] * st * std::cos(ph) + y[i] * st * std::sin(ph) + z[i] * ct;
    }
}

double orc_bf16(double x) { return bf16(x); }

}  // extern "C"
