// Environment-map importance-sampling tables, built on the host in double precision.
//
// Follows the reference's ImageEnvmap (common/share/envmap.hpp:151-235): texel x covers
// phi in [2pi-(x+1)dphi, 2pi-x dphi] (clockwise map, envmap.hpp:42-49), texel row y covers
// theta in [y dtheta, (y+1) dtheta]; the selection weight of a texel is its luminance times
// its solid angle times a direction weight evaluated at the texel centre; pdf = p / sr.
// The six-axis tables weight by max(+-direction component, 0) (envmap.hpp:291-312).
// The alias table is Vose's method with Kahan-compensated heights and LIFO small/large
// worklists (common/share/alias_method.hpp:45-114).
//
// One deliberate extension: a table whose weights sum to zero (a black map, or a six-axis
// table whose hemisphere is black) would divide by zero in the reference; here it falls
// back to weights proportional to solid angle times the direction weight (then to solid
// angle alone), and the table is flagged degenerate. Any valid pdf keeps the one-sample
// MIS estimator unbiased, and the radiance it samples is zero anyway.
#include "envmap_build.h"

#include <cmath>
#include <stdexcept>

namespace rgpu {
namespace {

const double kPi = 3.14159265358979323846;

struct Kahan {
    double sum = 0.0, c = 0.0;
    void add(double x) {
        double y = x - c;
        double t = sum + y;
        c = (t - sum) - y;
        sum = t;
    }
    double value() const { return sum - c; }
    void set(double x) {
        sum = x;
        c = 0.0;
    }
};

void prepare_alias(const std::vector<double>& weights, std::vector<double>* probs, std::vector<AliasBucket>* out) {
    const int N = (int)weights.size();
    Kahan wsum;
    for (double w : weights) wsum.add(w);
    double one_over = 1.0 / wsum.value();
    probs->assign(N, 0.0);
    std::vector<Kahan> height(N);
    std::vector<int> alias(N, -1);
    for (int i = 0; i < N; ++i) {
        (*probs)[i] = weights[i] * one_over;
        height[i].set((*probs)[i] * N);
    }
    std::vector<int> lower, upper;
    lower.reserve(N);
    upper.reserve(N);
    for (int i = 0; i < N; ++i) (height[i].value() < 1.0 ? lower : upper).push_back(i);
    while (!lower.empty() && !upper.empty()) {
        int lo = lower.back();
        lower.pop_back();
        int up = upper.back();
        upper.pop_back();
        double mov = 1.0 - height[lo].value();
        height[up].add(-mov);
        alias[lo] = up;
        (height[up].value() < 1.0 ? lower : upper).push_back(up);
    }
    out->resize(N);
    for (int i = 0; i < N; ++i) {
        (*out)[i].height = (float)height[i].value();
        (*out)[i].alias = alias[i];
    }
}

}  // namespace

void build_envmap_tables(const float* rgb, uint32_t channels, uint32_t width, uint32_t height, EnvmapTables* out) {
    if (!rgb || width == 0 || height == 0 || (channels != 3 && channels != 4))
        throw std::invalid_argument("envmap needs a width x height RGB(A) image");
    const uint32_t W = width, H = height;
    const size_t P = (size_t)W * H;
    out->width = W;
    out->height = H;
    out->texels.assign(P * 4, 0.0f);
    for (size_t i = 0; i < P; ++i)
        for (int k = 0; k < 3; ++k) out->texels[4 * i + k] = rgb[i * channels + k];

    const double phi_step = 2.0 * kPi / W;
    const double theta_step = kPi / H;
    out->fragments.resize(P);
    std::vector<double> row_sr(H), row_theta(H);
    for (uint32_t y = 0; y < H; ++y) {
        double beg_theta = theta_step * y, end_theta = theta_step * (y + 1);
        double beg_y = std::cos(end_theta), end_y = std::cos(beg_theta);
        // solid_angle_sliced_sphere (envmap.hpp:73-78), shared by the row's W texels
        row_sr[y] = (std::cos(beg_theta) - std::cos(end_theta)) * (2.0 * kPi) / W;
        row_theta[y] = (std::acos(beg_y) + std::acos(end_y)) * 0.5;
        for (uint32_t x = 0; x < W; ++x) {
            double upper_phi = 2.0 * kPi - x * phi_step;
            double lower_phi = 2.0 * kPi - (x + 1) * phi_step;
            EnvFragment& f = out->fragments[(size_t)y * W + x];
            f.beg_y = (float)beg_y;
            f.end_y = (float)end_y;
            f.beg_phi = (float)lower_phi;
            f.end_phi = (float)upper_phi;
        }
    }
    // Direction of a texel centre: polar_to_cartesian returns (y, z, x) = (sin t sin p, cos t, sin t cos p).
    auto dir_weight = [&](int table, double theta, double phi) {
        double st = std::sin(theta);
        double d[3] = {st * std::sin(phi), std::cos(theta), st * std::cos(phi)};
        if (table == 0) return 1.0;
        int axis = (table - 1) / 2;
        bool neg = ((table - 1) & 1) != 0;
        return std::max(neg ? -d[axis] : d[axis], 0.0);
    };
    for (int table = 0; table < 7; ++table) {
        std::vector<double> w(P);
        double total = 0.0;
        for (uint32_t y = 0; y < H; ++y) {
            for (uint32_t x = 0; x < W; ++x) {
                size_t i = (size_t)y * W + x;
                double upper_phi = 2.0 * kPi - x * phi_step;
                double lower_phi = 2.0 * kPi - (x + 1) * phi_step;
                // the fragment stores float phi limits; the reference averages its double ones
                double phi = (lower_phi + upper_phi) * 0.5;
                const float* t = &out->texels[4 * i];
                float Y = 0.2126f * t[0] + 0.7152f * t[1] + 0.0722f * t[2];
                w[i] = (double)Y * row_sr[y] * dir_weight(table, row_theta[y], phi);
                total += w[i];
            }
        }
        if (!(total > 0.0)) {
            out->degenerate[table] = true;
            total = 0.0;
            for (uint32_t y = 0; y < H; ++y)
                for (uint32_t x = 0; x < W; ++x) {
                    size_t i = (size_t)y * W + x;
                    double phi = (2.0 * kPi - (x + 0.5) * phi_step);
                    w[i] = row_sr[y] * dir_weight(table, row_theta[y], phi);
                    total += w[i];
                }
            if (!(total > 0.0))
                for (uint32_t y = 0; y < H; ++y)
                    for (uint32_t x = 0; x < W; ++x) w[(size_t)y * W + x] = row_sr[y];
        }
        std::vector<double> probs;
        prepare_alias(w, &probs, &out->alias[table]);
        out->pdf[table].resize(P);
        for (uint32_t y = 0; y < H; ++y)
            for (uint32_t x = 0; x < W; ++x) {
                size_t i = (size_t)y * W + x;
                out->pdf[table][i] = (float)(probs[i] * (1.0 / row_sr[y]));
            }
    }
}

}  // namespace rgpu
