"""Procedural render scenes: triangles, per-primitive materials, envmap, camera.

Triangle winding follows the reference's clockwise convention: the geometric normal is
normalize(cross(v2 - v1, v1 - v0)) (kernels/extension_ray_cast_stackless.cl:235), so the
meshes here are wound so that this normal points out of closed objects (which the
medium boundaries rely on to tell entering from leaving).
"""
import math

import numpy as np

from . import scenes
from ._lib import MAT_DIELECTRIC, MAT_LAMBERTIAN, MAT_SPECULAR, MAT_VOLUME, MAT_WARD, MaterialSet


class MeshBuilder:
    """Concatenates meshes and records one material per triangle."""

    def __init__(self):
        self.v, self.f, self.types, self.idx = [], [], [], []
        self.nv = 0

    def add(self, verts, faces, mat_type, mat_index):
        verts = np.asarray(verts, np.float64)
        faces = np.asarray(faces, np.int64)
        self.v.append(verts)
        self.f.append(faces + self.nv)
        self.nv += len(verts)
        self.types.append(np.full(len(faces), mat_type, np.int32))
        self.idx.append(np.full(len(faces), mat_index, np.int32))

    def build(self):
        v = np.concatenate(self.v).astype(np.float32)
        f = np.concatenate(self.f).astype(np.uint32)
        return v, f, np.concatenate(self.types), np.concatenate(self.idx)


def quad(p0, p1, p2, p3):
    """Quad p0..p3 counter-clockwise seen from the side it faces; emitted clockwise so the
    reference's normal convention points toward that side."""
    return np.array([p0, p1, p2, p3], np.float64), np.array([[0, 2, 1], [0, 3, 2]])


def box(lo, hi):
    """Closed box, 12 triangles, normals outward."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    faces = [
        quad((x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1)),  # +z
        quad((x1, y0, z0), (x0, y0, z0), (x0, y1, z0), (x1, y1, z0)),  # -z
        quad((x1, y0, z1), (x1, y0, z0), (x1, y1, z0), (x1, y1, z1)),  # +x
        quad((x0, y0, z0), (x0, y0, z1), (x0, y1, z1), (x0, y1, z0)),  # -x
        quad((x0, y1, z1), (x1, y1, z1), (x1, y1, z0), (x0, y1, z0)),  # +y
        quad((x0, y0, z0), (x1, y0, z0), (x1, y0, z1), (x0, y0, z1)),  # -y
    ]
    v = np.concatenate([q[0] for q in faces])
    f = np.concatenate([q[1] + 4 * i for i, q in enumerate(faces)])
    return v, f


def uv_sphere(center, radius, segments=64, rings=32):
    """UV sphere: `segments` longitudes, `rings` latitude bands, triangle-fan caps
    (segments * (2 * rings - 2) triangles), normals outward."""
    v = [(0.0, 1.0, 0.0)]
    for i in range(1, rings):
        th = math.pi * i / rings
        for j in range(segments):
            ph = 2.0 * math.pi * j / segments
            v.append((math.sin(th) * math.cos(ph), math.cos(th), math.sin(th) * math.sin(ph)))
    v.append((0.0, -1.0, 0.0))
    v = np.array(v) * radius + np.asarray(center, np.float64)
    south = len(v) - 1

    def ring(i, j):
        return 1 + (i - 1) * segments + (j % segments)

    f = []
    for j in range(segments):
        f.append((0, ring(1, j), ring(1, j + 1)))
    for i in range(1, rings - 1):
        for j in range(segments):
            a, b, c, d = ring(i, j), ring(i, j + 1), ring(i + 1, j + 1), ring(i + 1, j)
            f += [(a, d, c), (a, c, b)]
    for j in range(segments):
        f.append((south, ring(rings - 1, j + 1), ring(rings - 1, j)))
    return v, np.array(f)


def sky_envmap(width, height, sun_dir=(0.35, 0.55, -0.75), sun_peak=5.0e4, sun_width_deg=0.5,
               zenith=(0.25, 0.45, 0.95), horizon=(0.9, 0.9, 1.0), ground=(0.12, 0.1, 0.08)):
    """Sky gradient plus a Gaussian sun lobe (peak radiance sun_peak, standard deviation
    sun_width_deg). Row 0 is the zenith; texel (x, y) covers the directions of the
    reference's envmap parameterisation (common/share/envmap.hpp:42-49)."""
    ys = (np.arange(height) + 0.5) / height * math.pi
    xs = 2.0 * math.pi - (np.arange(width) + 0.5) / width * 2.0 * math.pi
    th, ph = np.meshgrid(ys, xs, indexing="ij")
    d = np.stack([np.sin(th) * np.sin(ph), np.cos(th), np.sin(th) * np.cos(ph)], -1)
    up = d[..., 1]
    t = np.clip(up, 0.0, 1.0) ** 0.5
    sky = (1 - t)[..., None] * np.asarray(horizon) + t[..., None] * np.asarray(zenith)
    sky = np.where((up < 0)[..., None], np.asarray(ground), sky)
    s = np.asarray(sun_dir, np.float64)
    s /= np.linalg.norm(s)
    ang = np.degrees(np.arccos(np.clip(d @ s, -1.0, 1.0)))
    sun = sun_peak * np.exp(-0.5 * (ang / sun_width_deg) ** 2)
    img = sky + sun[..., None] * np.array([1.0, 0.95, 0.85])
    return img.astype(np.float32)


def constant_envmap(value, width=16, height=8):
    return np.full((height, width, 3), value, np.float32)


# ------------------------------------------------------------------ scenes
def cornell_box(width=256, height=256, segments=64, rings=32, lens_radius=0.0045):
    """The plain path-tracer scene: five walls of two quads each (20 triangles), a 0.25 x 0.25
    area light emitting (17, 12, 4) and a UV sphere of albedo 0.8; black environment."""
    mb = MeshBuilder()
    lam = [((0, 0, 0), (0.75, 0.75, 0.75)),       # 0 white
           ((0, 0, 0), (0.63, 0.06, 0.04)),       # 1 red
           ((0, 0, 0), (0.14, 0.45, 0.09)),       # 2 green
           ((17.0, 12.0, 4.0), (0.78, 0.78, 0.78)),  # 3 light
           ((0, 0, 0), (0.8, 0.8, 0.8))]          # 4 sphere

    def wall(p0, p1, p2, p3, mat):
        # two quads per wall, split between p0-p1 and p3-p2
        m01 = tuple((np.add(p0, p1) / 2).tolist())
        m32 = tuple((np.add(p3, p2) / 2).tolist())
        for q in (quad(p0, m01, m32, p3), quad(m01, p1, p2, m32)):
            mb.add(q[0], q[1], MAT_LAMBERTIAN, mat)

    wall((-1, -1, 1), (1, -1, 1), (1, -1, -1), (-1, -1, -1), 0)    # floor, faces +y
    wall((-1, 1, -1), (1, 1, -1), (1, 1, 1), (-1, 1, 1), 0)        # ceiling, faces -y
    wall((-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1), 0)    # back, faces +z
    wall((-1, -1, 1), (-1, -1, -1), (-1, 1, -1), (-1, 1, 1), 1)    # left (red), faces +x
    wall((1, -1, -1), (1, -1, 1), (1, 1, 1), (1, 1, -1), 2)        # right (green), faces -x
    h = 0.125
    lq = quad((-h, 0.999, -h), (h, 0.999, -h), (h, 0.999, h), (-h, 0.999, h))
    mb.add(lq[0], lq[1], MAT_LAMBERTIAN, 3)
    sv, sf = uv_sphere((0.25, -0.55, -0.15), 0.45, segments, rings)
    mb.add(sv, sf, MAT_LAMBERTIAN, 4)
    v, f, types, idx = mb.build()
    cam = dict(eye=(0.0, 0.0, 3.6), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov=40.0,
               width=width, height=height, lens_radius=lens_radius)
    return dict(points=v, indices=f, materials=MaterialSet(types, idx, lambertians=lam),
                env=constant_envmap(0.0, 8, 4), camera=cam)


def mesh_scene(points, indices, width=320, height=240, albedo=0.75, env_size=(256, 128)):
    """A loaded mesh (clockwise triangles, e.g. scenes.load_obj) on a lambertian ground under a
    sky with a sun, the camera framing the mesh's bounding box from its front-left."""
    p = np.asarray(points, np.float64)
    lo, hi = p.min(axis=0), p.max(axis=0)
    c, r = 0.5 * (lo + hi), 0.5 * float(np.linalg.norm(hi - lo))
    mb = MeshBuilder()
    lam = [((0, 0, 0), (albedo, albedo, albedo)), ((0, 0, 0), (0.5, 0.5, 0.5))]
    mb.add(p, np.asarray(indices, np.int64), MAT_LAMBERTIAN, 0)
    e = 4.0 * r
    y = lo[1] - 1e-3 * r
    g = quad((c[0] - e, y, c[2] + e), (c[0] + e, y, c[2] + e), (c[0] + e, y, c[2] - e), (c[0] - e, y, c[2] - e))
    mb.add(g[0], g[1], MAT_LAMBERTIAN, 1)
    v, f, types, idx = mb.build()
    env = sky_envmap(env_size[0], env_size[1], sun_peak=200.0, sun_width_deg=4.0)
    eye = c + r * np.array([-1.2, 0.8, 2.4])
    cam = dict(eye=tuple(eye.tolist()), target=tuple(c.tolist()), up=(0.0, 1.0, 0.0), fov=40.0, width=width,
               height=height, lens_radius=0.0)
    return dict(points=v, indices=f, materials=MaterialSet(types, idx, lambertians=lam), env=env, camera=cam)


def material_test_scene(width=64, height=48, env_size=(64, 32), sun_width_deg=6.0, volume=True):
    """Every material type: a lambertian ground, icospheres of lambertian, isotropic Ward,
    anisotropic Ward, mirror and dielectric, and a homogeneous-medium box, under a sky with
    a sun."""
    mb = MeshBuilder()
    lam = [((0, 0, 0), (0.6, 0.6, 0.55)), ((0, 0, 0), (0.7, 0.2, 0.2))]
    wards = [(0.15, 0.15, (0.8, 0.7, 0.4), (1.0, 1.0, 1.0), 0.5), (0.08, 0.3, (0.4, 0.6, 0.9), (1, 1, 1), 0.5)]
    g = quad((-4, 0, 4), (4, 0, 4), (4, 0, -4), (-4, 0, -4))
    mb.add(g[0], g[1], MAT_LAMBERTIAN, 0)
    iv, iff = scenes.icosphere(2)
    spheres = [((-2.2, 0.6, 0), MAT_LAMBERTIAN, 1), ((-1.1, 0.6, 0.4), MAT_WARD, 0), ((0.0, 0.6, 0), MAT_WARD, 1),
               ((1.1, 0.6, 0.4), MAT_SPECULAR, 0), ((2.2, 0.6, 0), MAT_DIELECTRIC, 0)]
    for c, t, k in spheres:
        # icosphere faces are counter-clockwise seen from outside; swap to the clockwise convention
        mb.add(iv * 0.55 + np.asarray(c), iff[:, [0, 2, 1]], t, k)
    vols = []
    if volume:
        bv, bf = box((-0.5, 0.0, 1.0), (0.5, 1.0, 2.0))
        mb.add(bv, bf, MAT_VOLUME, 0)
        vols = [(0.9, (0.8 / 0.9, 0.8 / 0.9, 0.8 / 0.9))]
    v, f, types, idx = mb.build()
    ms = MaterialSet(types, idx, lambertians=lam, wards=wards, dielectrics=[1.5], volumes=vols)
    env = sky_envmap(env_size[0], env_size[1], sun_peak=200.0, sun_width_deg=sun_width_deg)
    cam = dict(eye=(0.0, 2.2, 6.0), target=(0.0, 0.5, 0.0), up=(0.0, 1.0, 0.0), fov=45.0, width=width,
               height=height, lens_radius=0.0045)
    return dict(points=v, indices=f, materials=ms, env=env, camera=cam)


def furnace_scene(kind, width=16, height=16, value=0.5, fill=False, flip=False):
    """A single convex object under a constant environment (closed-form furnace tests).

    fill=True narrows the camera so that every primary ray hits the object: every path then
    has the same number of bounces, and a lambertian sphere of reflectance R reads R * value
    in every pixel. flip=True reverses the sphere's winding so that the camera sees back
    faces only (geometric normals point inward)."""
    mb = MeshBuilder()
    iv, iff = scenes.icosphere(3)
    if not flip:
        iff = iff[:, [0, 2, 1]]
    lam, wards, diel, vols = [], [], [], []
    if kind == "lambertian":
        lam = [((0, 0, 0), (0.5, 0.5, 0.5))]
        mb.add(iv, iff, MAT_LAMBERTIAN, 0)
    elif kind == "ward":
        wards = [(0.12, 0.3, (0.7, 0.6, 0.5), (1.0, 1.0, 1.0), 0.5)]
        mb.add(iv, iff, MAT_WARD, 0)
    elif kind == "mirror":
        mb.add(iv, iff, MAT_SPECULAR, 0)
    elif kind == "dielectric":
        diel = [1.5]
        mb.add(iv, iff, MAT_DIELECTRIC, 0)
    elif kind == "volume":
        vols = [(0.9, (1.0, 1.0, 1.0))]
        bv, bf = box((-0.8, -0.8, -0.8), (0.8, 0.8, 0.8))
        mb.add(bv, bf, MAT_VOLUME, 0)
    else:
        raise ValueError(kind)
    v, f, types, idx = mb.build()
    ms = MaterialSet(types, idx, lambertians=lam, wards=wards, dielectrics=diel, volumes=vols)
    # fill: 10 degree field of view from distance 3 stays inside the sphere's 19 degree silhouette
    cam = dict(eye=(0.0, 0.0, 3.0 if fill else 4.5), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0),
               fov=10.0 if fill else 30.0, width=width, height=height, lens_radius=0.0)
    return dict(points=v, indices=f, materials=ms, env=constant_envmap(value), camera=cam)
