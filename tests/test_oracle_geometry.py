"""Pins the oracle's slab, triangle and sphere tests with closed-form cases.

The AABB cases restate the reference's lost aabb_cases.abc fixture from its generator
(UnitTest/bin/data/aabb_cases.hiplc:2796-2805 directions, 1802-1813 ground truth) and
apply the checks of UnitTest/src/unit_test.cpp:187-229.
"""
import numpy as np

import oracle as O

P0 = np.array([-0.5, -1.0, -1.5], np.float32)
P1 = np.array([0.5, 1.0, 1.5], np.float32)
DIRS = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [0, 1, 1], [1, 0, 1]], np.float64)
DIRS /= np.linalg.norm(DIRS, axis=1, keepdims=True)


def _first_surface_hit(ro, rd):
    """Distance along rd to the first box face in front of ro (float64), or -1."""
    t0 = np.full(3, -np.inf)
    t1 = np.full(3, np.inf)
    for k in range(3):
        if rd[k] != 0:
            a = (P0[k] - ro[k]) / rd[k]
            b = (P1[k] - ro[k]) / rd[k]
            t0[k], t1[k] = min(a, b), max(a, b)
        elif not (P0[k] <= ro[k] <= P1[k]):
            return -1.0, False
    tn, tf = t0.max(), t1.min()
    if tn > tf or tf <= 0:
        return -1.0, False
    if tn > 0:
        return tn, False  # enters through a face: outside
    return tf, True  # starts inside: exits through a face


def test_aabb_cases_like_reference_unit_test():
    rng = np.random.default_rng(5)
    n = 0
    for i in range(3000):
        ro = rng.uniform(-3, 3, 3)
        rd = DIRS[rng.integers(0, 6)]
        # keep the cases away from the faces so float32 rounding cannot flip the truth
        if np.min(np.abs(np.concatenate([ro - P0, ro - P1]))) < 1e-3:
            continue
        tmin, inside = _first_surface_hit(ro, rd)
        if tmin > 0 and abs(tmin) < 1e-3:
            continue
        ro32, rd32 = ro.astype(np.float32), rd.astype(np.float32)
        hit, h0, h1 = O.slabs_with_hit(P0, P1, ro32, rd32, np.finfo(np.float32).max)
        assert hit == (tmin > 0), (ro, rd, tmin)
        if hit:
            n += 1
            half, _, _ = O.slabs_with_hit(P0, P1, ro32, rd32, np.float32(tmin * 0.5))
            assert half == inside, (ro, rd, tmin, inside)
            # the entry distance is clamped at 0 and the exit distance at the far clip
            assert h0 >= 0.0 and (inside or abs(h0 - tmin) < 1e-4)
    assert n > 200


def test_slab_nan_cases_origin_on_face_plane():
    # Origin on the +x face with a direction parallel to it: (0.5-0.5)*inf = NaN is
    # replaced by -t0 (kernels/slab.cl:15), so the ray counts as inside that slab.
    hit, _, _ = O.slabs_with_hit(P0, P1, [0.5, 0.0, 0.0], [0.0, 1.0, 0.0], 1e30)
    assert hit
    hit, _, _ = O.slabs_with_hit(P0, P1, [0.5, 5.0, 0.0], [0.0, 1.0, 0.0], 1e30)
    assert not hit  # behind
    hit, _, _ = O.slabs_with_hit(P0, P1, [0.5, -5.0, 0.0], [0.0, 1.0, 0.0], 1e30)
    assert hit
    hit, h0, h1 = O.slabs_with_hit(P0, P1, [0.5, -5.0, 0.0], [0.0, 1.0, 0.0], 3.0)
    assert not hit  # the box starts at t = 4
    hit, _, _ = O.slabs_with_hit(P0, P1, [0.50001, -5.0, 0.0], [0.0, 1.0, 0.0], 1e30)
    assert not hit  # just outside the face plane


def test_triangle_closed_forms():
    v0, v1, v2 = [0, 0, 0], [1, 0, 0], [0, 1, 0]
    hit, tuv = O.intersect_triangle([0.25, 0.25, 2.0], [0, 0, -1], v0, v1, v2, 1e30)
    assert hit and tuv[0] == 2.0 and tuv[1] == 0.25 and tuv[2] == 0.25
    # vertex (u = v = 0) and edge (u + v = 1) are inside
    hit, tuv = O.intersect_triangle([0, 0, 1.0], [0, 0, -1], v0, v1, v2, 1e30)
    assert hit and tuv[1] == 0.0 and tuv[2] == 0.0
    hit, tuv = O.intersect_triangle([0.5, 0.5, 1.0], [0, 0, -1], v0, v1, v2, 1e30)
    assert hit and tuv[1] + tuv[2] == 1.0
    # outside, parallel, behind
    assert not O.intersect_triangle([0.6, 0.6, 1.0], [0, 0, -1], v0, v1, v2, 1e30)[0]
    assert not O.intersect_triangle([0.2, 0.2, 1.0], [1, 0, 0], v0, v1, v2, 1e30)[0]
    assert not O.intersect_triangle([0.2, 0.2, -1.0], [0, 0, -1], v0, v1, v2, 1e30)[0]
    # t == tmin is accepted (a later primitive wins an exact tie), t > tmin is not
    assert O.intersect_triangle([0.2, 0.2, 1.0], [0, 0, -1], v0, v1, v2, 1.0)[0]
    assert not O.intersect_triangle([0.2, 0.2, 1.0], [0, 0, -1], v0, v1, v2, np.nextafter(np.float32(1.0), np.float32(0)))[0]
    # t == 0 is accepted
    assert O.intersect_triangle([0.2, 0.2, 0.0], [0, 0, -1], v0, v1, v2, 1e30)[0]
    # |det| < 1e-8 is rejected: a sliver whose det is ~1e-9
    assert not O.intersect_triangle([0.0, 0.0, 1.0], [0, 0, -1], [0, 0, 0], [1e-4, 0, 0], [0, 1e-5, 0], 1e30)[0]


def test_triangle_agrees_with_float64_plane_intersection():
    rng = np.random.default_rng(9)
    for _ in range(500):
        tri = rng.normal(size=(3, 3))
        ro = rng.normal(size=3) * 3
        w = rng.dirichlet([1, 1, 1])
        target = w @ tri
        rd = target - ro
        rd /= np.linalg.norm(rd)
        hit, tuv = O.intersect_triangle(ro, rd, tri[0], tri[1], tri[2], 1e30)
        if min(w) < 1e-3:
            continue
        e1, e2 = tri[1] - tri[0], tri[2] - tri[0]
        if abs(np.dot(np.cross(e1, e2) / np.linalg.norm(np.cross(e1, e2)), rd)) < 1e-2:
            continue
        assert hit
        t_exact = np.linalg.norm(target - ro)
        assert abs(tuv[0] - t_exact) < 1e-4 * max(1.0, t_exact)
        assert abs(tuv[1] - w[1]) < 1e-3 and abs(tuv[2] - w[2]) < 1e-3
