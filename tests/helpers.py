"""Shared fixtures for the traversal tests: meshes, ray sets, comparisons."""
import numpy as np

from raccoongpurender_amd import scenes
from raccoongpurender_amd._lib import RAY_DTYPE

FLT_MAX = np.float32(np.finfo(np.float32).max)


def triangle_soup(n, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1, 1, (n, 1, 3))
    v = (c + rng.normal(scale=0.08, size=(n, 3, 3))).reshape(-1, 3).astype(np.float32)
    f = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    return v, f


def grid_plane(n):
    """n x n quads on z = 0: many coplanar triangles with shared edges (tie-prone)."""
    xs = np.linspace(-1, 1, n + 1, dtype=np.float32)
    X, Y = np.meshgrid(xs, xs)
    v = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size, np.float32)], 1)
    idx = []
    for j in range(n):
        for i in range(n):
            a = j * (n + 1) + i
            idx += [[a, a + 1, a + n + 2], [a, a + n + 2, a + n + 1]]
    return v, np.array(idx, np.uint32)


def degenerate_rays(points, n, seed):
    """Axis-aligned and diagonal rays starting on vertices (exact box-plane origins)."""
    rng = np.random.default_rng(seed)
    dirs = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
                     [1, 1, 0], [0, 1, 1], [1, 0, 1]], np.float64)
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    o = points[rng.integers(0, len(points), n)].astype(np.float64)
    d = dirs[rng.integers(0, len(dirs), n)]
    o = o - d * rng.choice([0.0, 0.5, 2.0], n)[:, None]
    return scenes.make_rays(o, d)


def with_tmax(rays, tmax):
    r = rays.copy()
    r["tmax"] = tmax
    return r


def _bits_differ(x, y):
    # bitwise, except that two NaNs agree: IEEE 754 leaves the sign and payload of a NaN that an
    # operation creates to the implementation (x86 and the GPU pick different ones)
    return (x.view(np.uint32) != y.view(np.uint32)) & ~(np.isnan(x) & np.isnan(y))


def compare_hits(a, b):
    """Counts of disagreements between two hit arrays (bitwise on t, u, v; NaN equals NaN)."""
    prim = int((a["prim"] != b["prim"]).sum())
    t = int(_bits_differ(a["t"], b["t"]).sum())
    hit = a["prim"] >= 0
    uv = int((_bits_differ(a["u"], b["u"]) | _bits_differ(a["v"], b["v"]))[hit].sum())
    return dict(prim=prim, t=t, uv=uv)


def tie_excepted_mismatch(a, ref, rel=1e-6):
    """Rays whose primitive differs from `ref` although t agrees within rel (ties), and those that truly differ."""
    diff = a["prim"] != ref["prim"]
    ta, tr = a["t"].astype(np.float64), ref["t"].astype(np.float64)
    tie = diff & (np.abs(ta - tr) <= rel * np.maximum(np.abs(tr), 1e-30))
    return int(tie.sum()), int((diff & ~tie).sum())
