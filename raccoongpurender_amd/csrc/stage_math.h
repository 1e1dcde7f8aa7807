// Double-precision exp, log, sin, cos, atan, atan2 and pow for the wavefront stages, built only
// from IEEE-754 +, -, *, / (correctly rounded; the library is compiled with -ffp-contract=off),
// comparisons, floor, frexp and ldexp, all of which give one defined result. The float values the
// stages use are these results rounded once, so they do not depend on any math library's
// last-place rounding, and a restatement that performs the same operations in the same order
// reproduces them bit for bit.
//
// Methods (each a textbook one):
// - exp: Cody & Waite argument reduction x = k ln2 + r with ln2 split into a 32-bit head and a
//   tail, |r| <= ln2/2, Taylor series of e^r to r^13 (truncation < 5e-18 relative);
// - log: x = m 2^e with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s), s = (m-1)/(m+1),
//   |s| <= 0.172, odd series to s^23;
// - sin, cos: reduction by pi/2 with pi/2 split into three 33-bit parts (Cody & Waite), |r| <=
//   pi/4, Taylor series to r^17 (sin) and r^18 (cos), quadrant by n mod 4;
// - atan: atan(a) = pi/2 - atan(1/a) for a > 1, then atan(a) = pi/6 + atan((a sqrt3 - 1) /
//   (a + sqrt3)) for a > tan(pi/12), Taylor series to a^29 on |a| <= tan(pi/12);
// - atan2: quadrants of atan(|y/x|) with the IEEE special cases;
// - pow: exp(y log |x|), with the IEEE rules for zero, negative and infinite bases.
// Accuracy: a few units in the last place of double, far below float resolution.
#pragma once
#include <hip/hip_runtime.h>

namespace rgpu {

__device__ __forceinline__ double sm_exp(double x) {
    if (x != x) return x;
    if (x > 709.78) return __longlong_as_double(0x7ff0000000000000ll);
    if (x < -745.2) return 0.0;
    const double kd = floor(x * 1.4426950408889634 + 0.5);
    const double r = (x - kd * 0.6931471803691238) - kd * 1.9082149292705877e-10;
    double p = 1.6059043836821613e-10;
    p = p * r + 2.08767569878681e-09;
    p = p * r + 2.505210838544172e-08;
    p = p * r + 2.755731922398589e-07;
    p = p * r + 2.7557319223985893e-06;
    p = p * r + 2.48015873015873e-05;
    p = p * r + 0.0001984126984126984;
    p = p * r + 0.001388888888888889;
    p = p * r + 0.008333333333333333;
    p = p * r + 0.041666666666666664;
    p = p * r + 0.16666666666666666;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    return ldexp(p, (int)kd);
}

__device__ __forceinline__ double sm_log(double x) {
    if (x != x) return x;
    if (x < 0.0) return __longlong_as_double(0x7ff8000000000000ll);
    if (x == 0.0) return __longlong_as_double(0xfff0000000000000ll);
    if (x == __longlong_as_double(0x7ff0000000000000ll)) return x;
    int e;
    double m = frexp(x, &e);
    if (m < 0.7071067811865476) {
        m = m * 2.0;
        e = e - 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double z = s * s;
    double p = 0.043478260869565216;
    p = p * z + 0.047619047619047616;
    p = p * z + 0.05263157894736842;
    p = p * z + 0.058823529411764705;
    p = p * z + 0.06666666666666667;
    p = p * z + 0.07692307692307693;
    p = p * z + 0.09090909090909091;
    p = p * z + 0.1111111111111111;
    p = p * z + 0.14285714285714285;
    p = p * z + 0.2;
    p = p * z + 0.3333333333333333;
    p = p * z + 1.0;
    const double lm = (2.0 * s) * p;
    const double ed = (double)e;
    return ed * 0.6931471803691238 + (ed * 1.9082149292705877e-10 + lm);
}

__device__ __forceinline__ double sm_sin_poly(double r) {
    const double z = r * r;
    double p = 2.8114572543455206e-15;
    p = p * z + -7.647163731819816e-13;
    p = p * z + 1.6059043836821613e-10;
    p = p * z + -2.505210838544172e-08;
    p = p * z + 2.7557319223985893e-06;
    p = p * z + -0.0001984126984126984;
    p = p * z + 0.008333333333333333;
    p = p * z + -0.16666666666666666;
    return r + r * (z * p);
}

__device__ __forceinline__ double sm_cos_poly(double r) {
    const double z = r * r;
    double p = -1.5619206968586225e-16;
    p = p * z + 4.779477332387385e-14;
    p = p * z + -1.1470745597729725e-11;
    p = p * z + 2.08767569878681e-09;
    p = p * z + -2.755731922398589e-07;
    p = p * z + 2.48015873015873e-05;
    p = p * z + -0.001388888888888889;
    p = p * z + 0.041666666666666664;
    p = p * z + -0.5;
    return 1.0 + z * p;
}

// x = n pi/2 + r, |r| <= pi/4; returns r and n mod 4. |x| must be finite.
__device__ __forceinline__ double sm_reduce_pio2(double x, int* q) {
    const double n = floor(x * 0.6366197723675814 + 0.5);
    *q = (int)(n - 4.0 * floor(n * 0.25));
    return ((x - n * 1.5707963267341256) - n * 6.077100506303966e-11) - n * 2.0222662487959506e-21;
}

__device__ __forceinline__ double sm_sin(double x) {
    if (!(fabs(x) < __longlong_as_double(0x7ff0000000000000ll))) return x - x;  // NaN for inf and NaN
    if (x == 0.0) return x;  // keeps the sign of zero
    int q;
    const double r = sm_reduce_pio2(x, &q);
    if (q == 0) return sm_sin_poly(r);
    if (q == 1) return sm_cos_poly(r);
    if (q == 2) return -sm_sin_poly(r);
    return -sm_cos_poly(r);
}

__device__ __forceinline__ double sm_cos(double x) {
    if (!(fabs(x) < __longlong_as_double(0x7ff0000000000000ll))) return x - x;
    int q;
    const double r = sm_reduce_pio2(x, &q);
    if (q == 0) return sm_cos_poly(r);
    if (q == 1) return -sm_sin_poly(r);
    if (q == 2) return -sm_cos_poly(r);
    return sm_sin_poly(r);
}

__device__ __forceinline__ double sm_atan(double x) {
    if (x != x || x == 0.0) return x;
    double a = fabs(x);
    const bool inv = a > 1.0;
    if (inv) a = 1.0 / a;
    double base = 0.0;
    if (a > 0.2679491924311228) {
        a = (a * 1.7320508075688772 - 1.0) / (a + 1.7320508075688772);
        base = 0.5235987755982988;
    }
    const double z = a * a;
    double p = 0.034482758620689655;
    p = p * z + -0.037037037037037035;
    p = p * z + 0.04;
    p = p * z + -0.043478260869565216;
    p = p * z + 0.047619047619047616;
    p = p * z + -0.05263157894736842;
    p = p * z + 0.058823529411764705;
    p = p * z + -0.06666666666666667;
    p = p * z + 0.07692307692307693;
    p = p * z + -0.09090909090909091;
    p = p * z + 0.1111111111111111;
    p = p * z + -0.14285714285714285;
    p = p * z + 0.2;
    p = p * z + -0.3333333333333333;
    double t = base + (a + a * (z * p));
    if (inv) t = 1.5707963267948966 - t;
    return x < 0.0 ? -t : t;
}

__device__ __forceinline__ double sm_atan2(double y, double x) {
    const double inf = __longlong_as_double(0x7ff0000000000000ll);
    const double pi = 3.141592653589793;
    if (x != x || y != y) return x + y;
    if (y == 0.0) return (x > 0.0 || (x == 0.0 && !signbit(x))) ? y : copysign(pi, y);
    if (x == 0.0) return copysign(1.5707963267948966, y);
    if (fabs(x) == inf) {
        if (fabs(y) == inf) return copysign(x > 0.0 ? 0.7853981633974483 : 2.356194490192345, y);
        return x > 0.0 ? copysign(0.0, y) : copysign(pi, y);
    }
    if (fabs(y) == inf) return copysign(1.5707963267948966, y);
    const double t = sm_atan(fabs(y / x));
    return copysign(x > 0.0 ? t : pi - t, y);
}

// x^y by the IEEE rules: a negative base takes integer exponents only (its sign by the
// exponent's parity), a zero base keeps its sign under odd exponents. The stages raise values
// such as 1 - h.wi that can round just below zero.
__device__ __forceinline__ double sm_pow(double x, double y) {
    const double inf = __longlong_as_double(0x7ff0000000000000ll);
    if (y == 0.0 || x == 1.0) return 1.0;
    if (x != x || y != y) return x + y;
    const bool integral = floor(y) == y;
    const bool odd = integral && floor(y * 0.5) != y * 0.5;
    if (x == 0.0) {
        if (y > 0.0) return odd ? x : 0.0;
        return odd ? copysign(inf, x) : inf;
    }
    double a = x;
    if (x < 0.0) {
        if (!integral) return __longlong_as_double(0x7ff8000000000000ll);
        a = -x;
    }
    double r;
    if (a == inf) r = y > 0.0 ? inf : 0.0;
    else if (a == 1.0) r = 1.0;  // (-1)^(+-inf)
    else r = sm_exp(y * sm_log(a));
    return (x < 0.0 && odd) ? -r : r;
}

}  // namespace rgpu
