"""The oracle's double-precision exp, log, sin, cos, atan, atan2 and pow (oracle/oracle_math.h),
which the HIP kernels evaluate with the same operations in the same order. Pinned against the
platform libm (glibc, accurate to within an ulp): the double results within a few ulps, the
float results (what the stages use) equal to libm's rounded result almost always and never more
than one float ulp away, and the IEEE special values."""
import math

import numpy as np
import pytest

import oracle as O


def _ulps(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.abs(a - b) / np.spacing(np.abs(b))


def _float_agreement(ours, ref):
    fo, fr = ours.astype(np.float32), ref.astype(np.float32)
    same = (fo == fr) | (np.isnan(fo) & np.isnan(fr))
    diff = np.abs(fo.view(np.int32).astype(np.int64) - fr.view(np.int32).astype(np.int64))
    return same.mean(), diff[~same].max(initial=0)


RNG = np.random.default_rng(2024)


def _args(name):
    if name == "exp":
        return np.concatenate([RNG.uniform(-745.0, 709.0, 200000), RNG.uniform(-20.0, 20.0, 200000),
                               RNG.uniform(-1e-6, 1e-6, 20000)])
    if name == "log":
        return np.concatenate([10.0 ** RNG.uniform(-300, 300, 200000), RNG.uniform(0.5, 2.0, 200000),
                               1.0 + RNG.uniform(-1e-6, 1e-6, 20000), np.array([5e-324, 2.2e-308, 1e-310])])
    if name in ("sin", "cos"):
        return np.concatenate([RNG.uniform(0.0, 2.0 * math.pi, 200000), RNG.uniform(-50.0, 50.0, 200000),
                               RNG.uniform(-1e-5, 1e-5, 20000)])
    if name == "atan":
        return np.concatenate([RNG.uniform(-2.0, 2.0, 200000), 10.0 ** RNG.uniform(-20, 20, 200000)
                               * RNG.choice([-1.0, 1.0], 200000)])
    raise KeyError(name)


@pytest.mark.parametrize("name", ["exp", "log", "sin", "cos", "atan"])
def test_unary_functions_against_libm(name):
    x = _args(name)
    ours = O.math_f64(name, x)
    ref = getattr(np, {"atan": "arctan"}.get(name, name))(x)
    u = _ulps(ours, ref)
    if name in ("sin", "cos"):
        # near a zero of sin or cos the reduction's absolute error (~1e-16 |x|) dominates
        ok = (u <= 4) | (np.abs(ours - ref) <= 4e-17 * np.maximum(1.0, np.abs(x)))
    else:
        ok = u <= 4
    print(f"{name}: max ulps {u[np.isfinite(u)].max():.2f}")
    assert ok.all(), (name, x[~ok][:5], ours[~ok][:5], ref[~ok][:5])
    # what the stages see: float arguments, float results
    xf = x.astype(np.float32).astype(np.float64)
    xf = xf[np.isfinite(getattr(np, {"atan": "arctan"}.get(name, name))(xf))]
    same, worst = _float_agreement(O.math_f64(name, xf), getattr(np, {"atan": "arctan"}.get(name, name))(xf))
    print(f"{name}: float results equal libm's for {same:.7f} of arguments, worst {worst} float ulp")
    assert same > 0.99999 and worst <= 1


def test_atan2_quadrants_and_special_values():
    y = RNG.normal(size=200000) * 10.0 ** RNG.uniform(-5, 5, 200000)
    x = RNG.normal(size=200000) * 10.0 ** RNG.uniform(-5, 5, 200000)
    ours, ref = O.math_f64("atan2", y, x), np.arctan2(y, x)
    assert (_ulps(ours, ref) <= 4).all()
    inf, nan = np.inf, np.nan
    special = [(0.0, 1.0), (-0.0, 1.0), (0.0, -1.0), (-0.0, -1.0), (0.0, 0.0), (-0.0, 0.0), (0.0, -0.0),
               (-0.0, -0.0), (1.0, 0.0), (-1.0, -0.0), (inf, inf), (-inf, inf), (inf, -inf), (-inf, -inf),
               (1.0, inf), (-1.0, inf), (1.0, -inf), (-1.0, -inf), (inf, 1.0), (-inf, -1.0), (nan, 1.0), (1.0, nan)]
    ys, xs = np.array([s[0] for s in special]), np.array([s[1] for s in special])
    got, want = O.math_f64("atan2", ys, xs), np.arctan2(ys, xs)
    for a, b, s in zip(got, want, special):
        assert (math.isnan(a) and math.isnan(b)) or (a == b and math.copysign(1, a) == math.copysign(1, b)), (s, a, b)


def test_pow_against_libm():
    x = np.concatenate([RNG.uniform(0.0, 1.0, 200000), 10.0 ** RNG.uniform(-30, 30, 200000)])
    y = np.concatenate([RNG.uniform(0.0, 8.0, 200000), RNG.uniform(-5.0, 5.0, 200000)])
    ours, ref = O.math_f64("pow", x, y), np.power(x, y)
    rel = np.abs(ours - ref) / np.abs(ref)
    bound = 8e-16 * (1.0 + np.abs(y * np.log(x)))
    ok = (rel <= bound) | (ref == 0) & (ours == 0)
    assert ok.all(), (x[~ok][:4], y[~ok][:4])
    # the shapes the stages raise: (1 - mu)^(1/falloff) and half-float means ^ (1/2.2)
    m = np.arange(1, 31744, dtype=np.uint16).view(np.float16).astype(np.float32)
    e = np.float32(1.0) / np.float32(2.2)
    same, worst = _float_agreement(O.pow_f32(m, e).astype(np.float64),
                                   np.power(m.astype(np.float64), np.float64(e)))
    assert same > 0.9999 and worst <= 1
    xs = RNG.uniform(0.0, 1.0, 100000).astype(np.float32)
    same, worst = _float_agreement(O.pow_f32(xs, 2.0).astype(np.float64), xs.astype(np.float64) ** 2)
    assert same > 0.9999 and worst <= 1


def test_special_values():
    inf, nan = np.inf, np.nan
    e = O.math_f64("exp", np.array([0.0, -0.0, inf, -inf, nan, 710.0, -746.0, 709.78, 1.0]))
    assert e[0] == 1 and e[1] == 1 and e[2] == inf and e[3] == 0 and np.isnan(e[4]) and e[5] == inf and e[6] == 0
    assert np.isfinite(e[7]) and abs(e[8] - math.e) <= 2 * np.spacing(math.e)
    lg = O.math_f64("log", np.array([1.0, 0.0, -0.0, -1.0, inf, nan, 2.0, 5e-324]))
    assert lg[0] == 0 and lg[1] == -inf and lg[2] == -inf and np.isnan(lg[3]) and lg[4] == inf and np.isnan(lg[5])
    assert abs(lg[6] - math.log(2)) <= np.spacing(math.log(2)) and abs(lg[7] - math.log(5e-324)) < 1e-12
    s = O.math_f64("sin", np.array([0.0, -0.0, inf, nan]))
    assert s[0] == 0 and math.copysign(1, s[1]) == -1 and np.isnan(s[2]) and np.isnan(s[3])
    c = O.math_f64("cos", np.array([0.0, inf]))
    assert c[0] == 1 and np.isnan(c[1])
    a = O.math_f64("atan", np.array([inf, -inf, 0.0, -0.0, 1.0]))
    assert a[0] == math.pi / 2 and a[1] == -math.pi / 2 and a[2] == 0 and math.copysign(1, a[3]) == -1
    assert abs(a[4] - math.pi / 4) <= np.spacing(math.pi / 4)
    bases = np.array([0.0, 1.0, 2.0, inf, 0.0, -1.0, 2.0, -0.0, -0.0, -0.0, -1e-8, -1e-8, -2.0, -inf, -inf, -1.0,
                      0.5, 3.0, nan])
    exps = np.array([2.0, nan, 0.0, 0.5, -1.0, 0.5, 10.0, 3.0, 2.0, -3.0, 2.0, 1.0, 3.0, 3.0, -2.0, inf,
                     inf, -inf, 0.0])
    got = O.math_f64("pow", bases, exps)
    want = np.power(bases, exps)
    for b, e, a, w in zip(bases, exps, got, want):
        if np.isnan(w):
            assert np.isnan(a), (b, e, a)
        else:
            assert math.copysign(1, a) == math.copysign(1, w), (b, e, a, w)
            # exp(y log|x|): the product's rounding grows with |y log|x|| (as in the random test)
            assert a == w or abs(a - w) <= 8e-16 * (1.0 + abs(e * math.log(abs(b)))) * abs(w), (b, e, a, w)
