// ORACLE (test infrastructure) — the double-precision exp, log, sin, cos, atan, atan2 and pow the
// restated stages use. The kernels of the HIP library evaluate these functions with the same
// textbook methods in the same operation order, from IEEE +, -, *, / (correctly rounded; both
// sides compile with -ffp-contract=off), floor, frexp and ldexp, so both sides get identical
// doubles and identical floats after one rounding, whatever either platform's libm would have
// returned. The reference kernels call OpenCL float builtins here (kernels/ward.cl:48-71,
// kernels/new_path.cl:61-65, kernels/homogeneous_volume_through.cl:37, kernels/lambertian.cl:11,
// kernels/envmap_sampling.cl:11-22,142-160, kernels/accumlation.cl:68-79), whose last-place
// behaviour is unspecified; these functions are accurate to a few double ulps.
//
// Methods: exp by Cody-Waite reduction k ln2 + r (ln2 as a 32-bit head plus tail) and the
// Taylor series to r^13; log by x = m 2^e, m in [sqrt(1/2), sqrt(2)), 2 atanh((m-1)/(m+1)) to
// the 23rd power; sin and cos by Cody-Waite reduction with pi/2 in three 33-bit parts and Taylor
// series on |r| <= pi/4; atan by atan(a) = pi/2 - atan(1/a) and atan(a) = pi/6 +
// atan((a sqrt3 - 1)/(a + sqrt3)), then the Taylor series on |a| <= tan(pi/12); atan2 from atan
// by quadrant; pow = exp(y log |x|) with the IEEE rules for zero, negative and infinite bases.
#pragma once
#include <cmath>
#include <limits>

namespace orcm {

inline double inf() { return std::numeric_limits<double>::infinity(); }
inline double qnan() { return std::numeric_limits<double>::quiet_NaN(); }

inline double exp(double x) {
    if (x != x) return x;
    if (x > 709.78) return inf();
    if (x < -745.2) return 0.0;
    const double kd = std::floor(x * 1.4426950408889634 + 0.5);
    const double r = (x - kd * 0.6931471803691238) - kd * 1.9082149292705877e-10;
    // Taylor coefficients 1/n!, n = 13 .. 0, by Horner's rule
    static const double c[14] = {1.6059043836821613e-10, 2.08767569878681e-09, 2.505210838544172e-08,
                                 2.755731922398589e-07,  2.7557319223985893e-06, 2.48015873015873e-05,
                                 0.0001984126984126984,  0.001388888888888889,   0.008333333333333333,
                                 0.041666666666666664,   0.16666666666666666,    0.5,
                                 1.0,                    1.0};
    double p = c[0];
    for (int i = 1; i < 14; ++i) p = p * r + c[i];
    return std::ldexp(p, (int)kd);
}

inline double log(double x) {
    if (x != x) return x;
    if (x < 0.0) return qnan();
    if (x == 0.0) return -inf();
    if (x == inf()) return x;
    int e;
    double m = std::frexp(x, &e);
    if (m < 0.7071067811865476) {
        m = m * 2.0;
        e = e - 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double z = s * s;
    // 1/(2k+1), k = 11 .. 0
    static const double c[12] = {0.043478260869565216, 0.047619047619047616, 0.05263157894736842,
                                 0.058823529411764705, 0.06666666666666667,  0.07692307692307693,
                                 0.09090909090909091,  0.1111111111111111,   0.14285714285714285,
                                 0.2,                  0.3333333333333333,   1.0};
    double p = c[0];
    for (int i = 1; i < 12; ++i) p = p * z + c[i];
    const double lm = (2.0 * s) * p;
    const double ed = (double)e;
    return ed * 0.6931471803691238 + (ed * 1.9082149292705877e-10 + lm);
}

inline double sin_poly(double r) {
    const double z = r * r;
    // (-1)^k / (2k+1)!, k = 8 .. 1
    static const double c[8] = {2.8114572543455206e-15, -7.647163731819816e-13, 1.6059043836821613e-10,
                                -2.505210838544172e-08, 2.7557319223985893e-06, -0.0001984126984126984,
                                0.008333333333333333,   -0.16666666666666666};
    double p = c[0];
    for (int i = 1; i < 8; ++i) p = p * z + c[i];
    return r + r * (z * p);
}

inline double cos_poly(double r) {
    const double z = r * r;
    // (-1)^k / (2k)!, k = 9 .. 1
    static const double c[9] = {-1.5619206968586225e-16, 4.779477332387385e-14, -1.1470745597729725e-11,
                                2.08767569878681e-09,    -2.755731922398589e-07, 2.48015873015873e-05,
                                -0.001388888888888889,   0.041666666666666664,   -0.5};
    double p = c[0];
    for (int i = 1; i < 9; ++i) p = p * z + c[i];
    return 1.0 + z * p;
}

inline double reduce_pio2(double x, int* q) {
    const double n = std::floor(x * 0.6366197723675814 + 0.5);
    *q = (int)(n - 4.0 * std::floor(n * 0.25));
    return ((x - n * 1.5707963267341256) - n * 6.077100506303966e-11) - n * 2.0222662487959506e-21;
}

inline double sin(double x) {
    if (!(std::fabs(x) < inf())) return x - x;
    if (x == 0.0) return x;
    int q;
    const double r = reduce_pio2(x, &q);
    switch (q) {
        case 0: return sin_poly(r);
        case 1: return cos_poly(r);
        case 2: return -sin_poly(r);
        default: return -cos_poly(r);
    }
}

inline double cos(double x) {
    if (!(std::fabs(x) < inf())) return x - x;
    int q;
    const double r = reduce_pio2(x, &q);
    switch (q) {
        case 0: return cos_poly(r);
        case 1: return -sin_poly(r);
        case 2: return -cos_poly(r);
        default: return sin_poly(r);
    }
}

inline double atan(double x) {
    if (x != x || x == 0.0) return x;
    double a = std::fabs(x);
    const bool inv = a > 1.0;
    if (inv) a = 1.0 / a;
    double base = 0.0;
    if (a > 0.2679491924311228) {
        a = (a * 1.7320508075688772 - 1.0) / (a + 1.7320508075688772);
        base = 0.5235987755982988;
    }
    const double z = a * a;
    // (-1)^k / (2k+1), k = 14 .. 1
    static const double c[14] = {0.034482758620689655, -0.037037037037037035, 0.04,
                                 -0.043478260869565216, 0.047619047619047616, -0.05263157894736842,
                                 0.058823529411764705, -0.06666666666666667,  0.07692307692307693,
                                 -0.09090909090909091, 0.1111111111111111,    -0.14285714285714285,
                                 0.2,                  -0.3333333333333333};
    double p = c[0];
    for (int i = 1; i < 14; ++i) p = p * z + c[i];
    double t = base + (a + a * (z * p));
    if (inv) t = 1.5707963267948966 - t;
    return x < 0.0 ? -t : t;
}

inline double atan2(double y, double x) {
    const double pi = 3.141592653589793;
    if (x != x || y != y) return x + y;
    if (y == 0.0) return (x > 0.0 || (x == 0.0 && !std::signbit(x))) ? y : std::copysign(pi, y);
    if (x == 0.0) return std::copysign(1.5707963267948966, y);
    if (std::fabs(x) == inf()) {
        if (std::fabs(y) == inf()) return std::copysign(x > 0.0 ? 0.7853981633974483 : 2.356194490192345, y);
        return x > 0.0 ? std::copysign(0.0, y) : std::copysign(pi, y);
    }
    if (std::fabs(y) == inf()) return std::copysign(1.5707963267948966, y);
    const double t = atan(std::fabs(y / x));
    return std::copysign(x > 0.0 ? t : pi - t, y);
}

// IEEE rules: a negative base takes integer exponents only (sign by parity), a zero base keeps
// its sign under odd exponents.
inline double pow(double x, double y) {
    if (y == 0.0 || x == 1.0) return 1.0;
    if (x != x || y != y) return x + y;
    const bool integral = std::floor(y) == y;
    const bool odd = integral && std::floor(y * 0.5) != y * 0.5;
    if (x == 0.0) {
        if (y > 0.0) return odd ? x : 0.0;
        return odd ? std::copysign(inf(), x) : inf();
    }
    double a = x;
    if (x < 0.0) {
        if (!integral) return qnan();
        a = -x;
    }
    double r;
    if (a == inf()) r = y > 0.0 ? inf() : 0.0;
    else if (a == 1.0) r = 1.0;  // (-1)^(+-inf)
    else r = exp(y * log(a));
    return (x < 0.0 && odd) ? -r : r;
}

}  // namespace orcm
