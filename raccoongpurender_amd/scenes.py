"""Procedural scenes and ray sets for the benchmark and test workloads.

All generators are deterministic functions of their seed (numpy PCG64 streams and an
integer lattice hash), computed in float64 and rounded once to float32, so the same
arrays reach the GPU kernels and the CPU oracle.
"""
import math

import numpy as np

from ._lib import RAY_DTYPE

FLT_MAX = np.float32(np.finfo(np.float32).max)


# ------------------------------------------------------------------ meshes
def icosphere(level):
    """Unit icosphere: 20 * 4**level triangles, vertices on the unit sphere."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    for _ in range(level):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]], axis=0)
        e.sort(axis=1)
        key = e[:, 0] * (len(v) + 1) + e[:, 1]
        uniq, inv = np.unique(key, return_inverse=True)
        a = uniq // (len(v) + 1)
        b = uniq % (len(v) + 1)
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = len(v)
        v = np.concatenate([v, mid], axis=0)
        m = inv.reshape(3, -1).T + base  # midpoints of edges (0,1), (1,2), (2,0)
        m01, m12, m20 = m[:, 0], m[:, 1], m[:, 2]
        f = np.concatenate([
            np.stack([f[:, 0], m01, m20], 1),
            np.stack([f[:, 1], m12, m01], 1),
            np.stack([f[:, 2], m20, m12], 1),
            np.stack([m01, m12, m20], 1),
        ], axis=0)
    return v, f


def _lattice_value(ix, iy, iz, seed):
    """Hash of integer lattice points to [-1, 1) (splitmix64 finaliser on a packed key)."""
    with np.errstate(over="ignore"):
        k = (ix.astype(np.uint64) * np.uint64(0x9E3779B1) ^ iy.astype(np.uint64) * np.uint64(0x85EBCA77)
             ^ iz.astype(np.uint64) * np.uint64(0xC2B2AE3D) ^ np.uint64(seed) * np.uint64(0x27D4EB2F165667C5))
        k = k + np.uint64(0x9E3779B97F4A7C15)
        k = (k ^ (k >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        k = (k ^ (k >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        k = k ^ (k >> np.uint64(31))
    return (k >> np.uint64(11)).astype(np.float64) / float(1 << 53) * 2.0 - 1.0


def value_noise3(p, seed):
    """Trilinear value noise with smoothstep weights, in [-1, 1]."""
    fl = np.floor(p)
    f = p - fl
    i = fl.astype(np.int64)
    w = f * f * (3.0 - 2.0 * f)
    out = np.zeros(len(p))
    for dx in (0, 1):
        wx = w[:, 0] if dx else 1.0 - w[:, 0]
        for dy in (0, 1):
            wy = w[:, 1] if dy else 1.0 - w[:, 1]
            for dz in (0, 1):
                wz = w[:, 2] if dz else 1.0 - w[:, 2]
                out += wx * wy * wz * _lattice_value(i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz, seed)
    return out


def fbm(p, octaves=4, seed=7, base_frequency=2.0):
    """Fractal sum of value noise, octave o at frequency base*2**o and weight 2**-o, normalised to [-1, 1]."""
    total = np.zeros(len(p))
    norm = 0.0
    for o in range(octaves):
        a = 0.5 ** o
        total += a * value_noise3(p * (base_frequency * 2.0 ** o), seed + 1013 * o)
        norm += a
    return total / norm


def displaced_icosphere(level=8, amplitude=0.15, seed=7, octaves=4):
    """Icosphere radially displaced by fBm value noise: r = 1 + amplitude * fbm(p)."""
    v, f = icosphere(level)
    r = 1.0 + amplitude * fbm(v, octaves=octaves, seed=seed)
    v = v * r[:, None]
    return v.astype(np.float32), f.astype(np.uint32)


# ------------------------------------------------------------------ rays
def make_rays(ro, rd, tmax=None):
    n = len(ro)
    rays = np.zeros(n, dtype=RAY_DTYPE)
    rays["ro"] = ro.astype(np.float32)
    rays["rd"] = rd.astype(np.float32)
    rays["tmax"] = FLT_MAX if tmax is None else tmax
    return rays


def look_at_camera(eye, target, up, fov_deg, width, height):
    eye = np.asarray(eye, np.float64)
    fwd = np.asarray(target, np.float64) - eye
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, np.asarray(up, np.float64))
    right /= np.linalg.norm(right)
    upv = np.cross(right, fwd)
    h = math.tan(math.radians(fov_deg) * 0.5)
    w = h * width / height
    return dict(eye=eye, forward=fwd, right=right, up=upv, half_w=w, half_h=h, width=width, height=height)


def primary_rays(width, height, eye=(0.0, 0.0, 3.0), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov_deg=45.0):
    """Coherent pinhole rays through pixel centres, row-major from the top-left pixel."""
    cam = look_at_camera(eye, target, up, fov_deg, width, height)
    xs = (np.arange(width) + 0.5) / width * 2.0 - 1.0
    ys = 1.0 - (np.arange(height) + 0.5) / height * 2.0
    X, Y = np.meshgrid(xs, ys)
    d = (cam["forward"][None, :] + X.reshape(-1, 1) * cam["half_w"] * cam["right"][None, :]
         + Y.reshape(-1, 1) * cam["half_h"] * cam["up"][None, :])
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = np.broadcast_to(cam["eye"], d.shape)
    return make_rays(o, d)


def incoherent_rays(n, lo, hi, seed=11):
    """Origins uniform in the box [lo, hi], directions uniform on the unit sphere."""
    rng = np.random.default_rng(seed)
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    o = lo + (hi - lo) * rng.random((n, 3))
    z = 2.0 * rng.random(n) - 1.0
    phi = 2.0 * math.pi * rng.random(n)
    s = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    d = np.stack([s * np.cos(phi), s * np.sin(phi), z], axis=1)
    return make_rays(o, d)


def raycast_benchmark_scene(level=8, n_rays=None, width=2048, height=2048, seed_mesh=7, seed_rays=11):
    """The ray-cast microbenchmark: displaced icosphere (20*4**level triangles),
    `width*height` coherent primary rays and as many incoherent rays."""
    v, f = displaced_icosphere(level=level, seed=seed_mesh)
    n = width * height if n_rays is None else n_rays
    lo, hi = v.min(axis=0), v.max(axis=0)
    coherent = primary_rays(width, height)
    incoherent = incoherent_rays(n, lo, hi, seed=seed_rays)
    return dict(points=v, indices=f, coherent=coherent, incoherent=incoherent, aabb=(lo, hi))


def load_obj(source, clockwise=True):
    """Triangle mesh of a Wavefront OBJ file (a path, or the text itself when it contains a
    newline): (points float32 (n, 3), indices uint32 (m, 3)). Reads `v` and `f` records; a face
    may give `i`, `i/t`, `i//n` or `i/t/n` per corner, negative indices count back from the
    last vertex, and polygons are split into a fan (0, k, k+1). Other records (vt, vn, o, g, s,
    usemtl, mtllib, comments) are skipped. OBJ faces are counter-clockwise seen from the front;
    with `clockwise` the corners are reversed to the clockwise order of addPolymesh
    (common/scene_manager.hpp:339-370). Raises ValueError naming the line of a bad record."""
    if "\n" in source:
        lines = source.splitlines()
    else:
        with open(source, "r") as fh:
            lines = fh.read().splitlines()
    pts, tris = [], []
    for ln, line in enumerate(lines, 1):
        rec = line.split("#", 1)[0].split()
        if not rec:
            continue
        if rec[0] == "v":
            if len(rec) < 4:
                raise ValueError(f"line {ln}: a vertex needs three coordinates")
            try:
                pts.append((float(rec[1]), float(rec[2]), float(rec[3])))
            except ValueError:
                raise ValueError(f"line {ln}: bad vertex coordinate") from None
        elif rec[0] == "f":
            if len(rec) < 4:
                raise ValueError(f"line {ln}: a face needs three corners")
            corner = []
            for c in rec[1:]:
                try:
                    i = int(c.split("/", 1)[0])
                except ValueError:
                    raise ValueError(f"line {ln}: bad face corner {c!r}") from None
                j = i - 1 if i > 0 else len(pts) + i
                if i == 0 or not 0 <= j < len(pts):
                    raise ValueError(f"line {ln}: vertex index {i} out of range")
                corner.append(j)
            for k in range(1, len(corner) - 1):
                t = (corner[0], corner[k], corner[k + 1])
                tris.append((t[0], t[2], t[1]) if clockwise else t)
    return np.asarray(pts, np.float32).reshape(-1, 3), np.asarray(tris, np.uint32).reshape(-1, 3)
